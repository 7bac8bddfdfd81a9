// tnr_kernel_jacobi_noalloc: the grouped Jacobi eigensolver of
// tnr_kernel_jacobi on caller-owned device buffers, allocating nothing.
//
// Its only working memory (the solver descriptors and the per-problem sweep
// counts) lives in `scratch`, a file-scope device array of NOALLOC_SCRATCH
// bytes that every GPU the library runs on holds until the process ends.
// The descriptors are written on the device by k_noalloc_descs, so no host
// staging buffer, device allocation or descriptor copy is involved.  All
// calls on one GPU share its copy of the array: calls that overlap on one
// GPU (from several host threads or streams) race.
#include <mutex>

#include "kernels.h"
#include "solver.h"

namespace tnr {
namespace {

constexpr size_t NOALLOC_SCRATCH = size_t(1) << 20;   // 1 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH];

// descriptor + sweep count of each problem
constexpr size_t kPerProblem = sizeof(JacobiDesc) + sizeof(int);
constexpr int kMaxProblems = (int)(NOALLOC_SCRATCH / kPerProblem);

// The device address of `scratch` on each GPU (the array's name is a host
// stand-in, not a device address); looked up once per device, with that
// device current.
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

void* scratch_ptr(int dev) {
    TNR_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, TNR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        TNR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TNR_REQUIRE(p != nullptr, TNR_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Problem i: G, W at i * n * n, lam at i * n; sweep counts start at 0 (the
// array keeps the previous call's contents).
__global__ __launch_bounds__(256) void k_noalloc_descs(JacobiDesc* descs, int* sweeps, int count,
                                                       int n, const double* G, double* W,
                                                       double* lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t nn = (int64_t)n * n;
    JacobiDesc d{G + nn * i, W + nn * i, lam + (int64_t)n * i, sweeps + i, n, n, n, nullptr,
                 nullptr, 0};
    descs[i] = d;
    sweeps[i] = 0;
}

}  // namespace
}  // namespace tnr

using namespace tnr;

extern "C" int tnr_kernel_jacobi_noalloc(int count, int n, const double* G, double* W,
                                         double* lam, int* sweeps, void* stream) {
    try {
        TNR_REQUIRE(count > 0 && n > 0 && n <= LMAX && G && W && lam, TNR_ERR_ARG,
                    "bad Jacobi arguments");
        TNR_REQUIRE(count <= kMaxProblems, TNR_ERR_ARG,
                    "count = " + std::to_string(count) + " exceeds the " +
                        std::to_string(kMaxProblems) +
                        " problems the 1 MiB scratch array holds");
        ensure_device();
        int dev = 0;
        TNR_HIP(hipGetDevice(&dev));
        unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));
        JacobiDesc* descs = reinterpret_cast<JacobiDesc*>(base);
        int* sw = reinterpret_cast<int*>(base + (size_t)count * sizeof(JacobiDesc));
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_noalloc_descs, dim3((count + 255) / 256), dim3(256), 0, st, descs, sw,
                           count, n, G, W, lam);
        TNR_HIP(hipGetLastError());
        const SolverCfg& cfg = solver_cfg();
        launch_jacobi(st, descs, count, n, cfg.jacobi_tol, cfg.jacobi_sweeps);
        if (sweeps)
            TNR_HIP(hipMemcpyAsync(sweeps, sw, (size_t)count * sizeof(int), hipMemcpyDeviceToHost,
                                   st));
        TNR_HIP(hipStreamSynchronize(st));
        return TNR_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return TNR_ERR_INTERNAL;
    }
}
