// Device pool, staging ring and execution context.
#include <algorithm>
#include <cstring>

#include "common.h"

namespace tnr {

void (*g_release_idle_memory)() = nullptr;

// ------------------------------------------------------------------- Pool
size_t Pool::bucket(size_t bytes) {
    if (bytes == 0) return 0;
    // 256-byte granularity below 1 MiB, then ~1/8 geometric steps
    if (bytes <= (size_t(1) << 20)) return (bytes + 255) & ~size_t(255);
    size_t p = size_t(1) << 20;
    while (p < bytes) p <<= 1;
    size_t step = p >> 3;
    return ((bytes + step - 1) / step) * step;
}

void* Pool::get(size_t bytes) {
    auto it = free_.find(bytes);
    if (it != free_.end()) {
        void* p = it->second;
        free_.erase(it);
        cached_ -= bytes;
        in_use_ += bytes;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess && g_release_idle_memory) {
            (void)hipGetLastError();
            g_release_idle_memory();
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(TNR_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    in_use_ += bytes;
    return p;
}

void Pool::put(void* p, size_t bytes) {
    free_.emplace(bytes, p);
    in_use_ -= bytes;
    cached_ += bytes;
}

void Pool::trim() {
    (void)hipDeviceSynchronize();
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

Pool::~Pool() {
    for (auto& kv : free_) (void)hipFree(kv.second);
}

// ---------------------------------------------------------------- Staging
void Staging::alloc(size_t bytes) {
    cap_ = bytes;
    head_ = 0;
    TNR_HIP(hipHostMalloc((void**)&host_, cap_, hipHostMallocDefault));
    TNR_HIP(hipMalloc((void**)&dev_, cap_));
}

void Staging::release() {
    if (host_) (void)hipHostFree(host_);
    if (dev_) (void)hipFree(dev_);
    host_ = dev_ = nullptr;
}

Staging::Staging(size_t bytes) { alloc(bytes); }

Staging::~Staging() { release(); }

void* Staging::put(hipStream_t st, const void* src, size_t bytes) {
    size_t need = (bytes + 255) & ~size_t(255);
    if (need == 0) need = 256;
    if (need > cap_) {
        // grow: everything queued so far may still read the old ring
        TNR_HIP(hipStreamSynchronize(st));
        release();
        size_t c = cap_ * 2;
        while (c < need) c *= 2;
        alloc(c);
    }
    if (head_ + need > cap_) {
        TNR_HIP(hipStreamSynchronize(st));
        head_ = 0;
    }
    char* h = host_ + head_;
    char* d = dev_ + head_;
    head_ += need;
    if (bytes) {
        std::memcpy(h, src, bytes);
        TNR_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    }
    return d;
}

// -------------------------------------------------------------------- Ctx
void ensure_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        throw Error(TNR_ERR_NODEVICE, "no HIP device visible");
    }
    hipDeviceProp_t prop;
    TNR_HIP(hipGetDeviceProperties(&prop, 0));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw Error(TNR_ERR_NODEVICE, std::string("device 0 is ") + prop.gcnArchName +
                                          ", this library is built for gfx950 only");
}

Ctx::Ctx() {
    ensure_device();
    TNR_HIP(hipGetDevice(&device));
    TNR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    stage.reset(new Staging());
}

Ctx::~Ctx() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        stage.reset();
        (void)hipStreamDestroy(stream);
    }
    if (bounce_) (void)hipHostFree(bounce_);
}

void Ctx::sync() { TNR_HIP(hipStreamSynchronize(stream)); }

void Ctx::drop_stream() {
    if (!stream) return;
    TNR_HIP(hipStreamSynchronize(stream));
    TNR_HIP(hipStreamDestroy(stream));
    stream = nullptr;
}

void Ctx::ensure_stream() {
    if (!stream) TNR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

void Ctx::fetch(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (bytes > bounce_cap_) {
        TNR_HIP(hipStreamSynchronize(stream));
        if (bounce_) TNR_HIP(hipHostFree(bounce_));
        bounce_ = nullptr;
        bounce_cap_ = 0;
        const size_t cap = std::max<size_t>(bytes, 1 << 16);
        TNR_HIP(hipHostMalloc(&bounce_, cap, hipHostMallocDefault));
        bounce_cap_ = cap;
    }
    TNR_HIP(hipMemcpyAsync(bounce_, src, bytes, hipMemcpyDeviceToHost, stream));
    TNR_HIP(hipStreamSynchronize(stream));
    std::memcpy(dst, bounce_, bytes);
}

}  // namespace tnr
