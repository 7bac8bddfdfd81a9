// Internal definitions shared by the host engine and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tnr.h"

namespace tnr {

constexpr int MAXN = 8;      // largest Z_N order supported
constexpr int MAXLEG = 8;    // largest tensor rank (HOTRG_3D intermediates have 8)
constexpr int LMAX = 96;     // largest Rayleigh-Ritz block / direct-solve size

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TNR_HIP(x)                                                            \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess)                                                 \
            throw ::tnr::Error(TNR_ERR_HIP, std::string(#x) + ": " +          \
                                                hipGetErrorString(e_));       \
    } while (0)

#define TNR_REQUIRE(cond, code, msg)                                          \
    do {                                                                      \
        if (!(cond)) throw ::tnr::Error((code), (msg));                       \
    } while (0)

// ---------------------------------------------------------------- device pool
// Caching allocator: freed buffers go back to a size-bucketed free list and
// are reused, so the coarse-graining loop never calls hipMalloc after warm-up.
class Pool {
  public:
    ~Pool();
    void* get(size_t bytes);
    void put(void* p, size_t bytes);
    size_t in_use() const { return in_use_; }
    size_t cached() const { return cached_; }
    void trim();
    static size_t bucket(size_t bytes);

  private:
    std::multimap<size_t, void*> free_;
    size_t in_use_ = 0, cached_ = 0;
};

// RAII device buffer of doubles (or raw bytes) from a Pool.
class Buf {
  public:
    Buf() = default;
    Buf(Pool* pool, size_t bytes) : pool_(pool), bytes_(Pool::bucket(bytes)) {
        p_ = bytes_ ? pool->get(bytes_) : nullptr;
    }
    ~Buf() { release(); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept { *this = std::move(o); }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) {
            release();
            pool_ = o.pool_; p_ = o.p_; bytes_ = o.bytes_;
            o.p_ = nullptr; o.bytes_ = 0;
        }
        return *this;
    }
    double* d() const { return static_cast<double*>(p_); }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    size_t bytes() const { return bytes_; }
    void release() {
        if (p_ && pool_) pool_->put(p_, bytes_);
        p_ = nullptr; bytes_ = 0;
    }

  private:
    Pool* pool_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// Pinned host ring + device ring for small asynchronous uploads (kernel
// descriptors, weight vectors).  A chunk is reused only after the stream has
// passed the work that read it (the ring synchronises when it wraps).
class Staging {
  public:
    explicit Staging(size_t bytes = size_t(4) << 20);
    ~Staging();
    // copy `bytes` from host memory to a device chunk; returns device pointer
    void* put(hipStream_t st, const void* src, size_t bytes);
    template <class T> T* put(hipStream_t st, const std::vector<T>& v) {
        return static_cast<T*>(put(st, v.data(), v.size() * sizeof(T)));
    }

  private:
    void alloc(size_t bytes);
    void release();
    char* host_ = nullptr;
    char* dev_ = nullptr;
    size_t cap_ = 0, head_ = 0;
};

// Execution context: one HIP stream, one pool, staging ring.
struct Ctx {
    hipStream_t stream = nullptr;
    Pool pool;
    std::unique_ptr<Staging> stage;
    int device = 0;
    int precision = TNR_PREC_FP64;
    Ctx();
    ~Ctx();
    void sync();
    // an idle context gives its stream back (streams share the device's few
    // hardware queues) and makes a new one before it is used again
    void drop_stream();
    void ensure_stream();
    // device -> host copy that waits for the stream (through a pinned
    // bounce buffer, so the DMA engine writes host memory directly)
    void fetch(void* dst, const void* src, size_t bytes);

  private:
    void* bounce_ = nullptr;
    size_t bounce_cap_ = 0;
};

void ensure_device();   // throws TNR_ERR_NODEVICE if no gfx950 GPU

// Called when a device allocation fails even after the failing pool dropped
// its own cache: frees what idle contexts and idle batch engines hold cached.
// Set once by the C API layer.
extern void (*g_release_idle_memory)();
void set_last_error(const std::string& msg);   // message for tnr_last_error()

}  // namespace tnr
