// ba_noalloc.hip — the dense camera-system solve for callers that cannot allocate device memory.
//
// ba_hip_cholesky_solve allocates its workspace (a copy of the matrix, the factor's scratch and a
// status flag) on every call.  ba_hip_cholesky_solve_noalloc runs the same kernels with that
// workspace in `scratch`, a file-scope device array of fixed size: the caller passes only the
// inputs and the output, and the call makes no hipMalloc, hipFree or other allocating call.
// Each device holds its own copy of the array for the life of the process; its address on a
// device is resolved once, with that device current, when a handle is created on it
// (noalloc_prepare_device), and kept per device.  Every call on one device shares that device's
// array, so calls that overlap on one device race; the array keeps the last call's contents, and
// the flags the kernels read are cleared by the kernels themselves at the start of each call.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "../../include/ba_hip.h"
#include "ba_handle.hpp"
#include "ba_kernels.hpp"

namespace ba {
namespace {

constexpr size_t NOALLOC_N = BA_NOALLOC_MAX_DIMENSION;
constexpr size_t TILE = 64;
// the factor's workspace for NOALLOC_N (cholesky_work_doubles): 16 scalars, six n-vectors and one
// 64 x 64 inverse per diagonal tile
constexpr size_t WORK_DOUBLES = 16 + 6 * NOALLOC_N + ((NOALLOC_N + TILE - 1) / TILE) * TILE * TILE;
constexpr size_t MATRIX_DOUBLES = NOALLOC_N * NOALLOC_N;  // FP64 copy, or the FP32 factor in its first half
constexpr size_t FLAG_DOUBLES = 2;                        // two int flags
constexpr size_t SCRATCH_DOUBLES = MATRIX_DOUBLES + WORK_DOUBLES + FLAG_DOUBLES;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
void* g_scratch[MAX_DEVICES] = {};
bool g_tried[MAX_DEVICES] = {};

}  // namespace

// Resolve this device's copy of the scratch array; the caller has made `device` current.
void noalloc_prepare_device(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> guard(g_lock);
    if (g_tried[device]) return;
    g_tried[device] = true;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch[device] = ptr;
}

static double* scratch_of(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> guard(g_lock);
    return static_cast<double*>(g_scratch[device]);
}

}  // namespace ba

extern "C" int ba_hip_cholesky_solve_noalloc(ba_hip_handle* h, size_t n, const double* a, const double* b, double* x,
                                              int32_t precision) {
    using namespace ba;
    if (!h || !h->pinned) return BA_ERR_INVALID_CONFIG;
    if (precision != BA_PRECISION_FP64 && precision != BA_PRECISION_MIXED) return BA_ERR_INVALID_CONFIG;
    if (n > NOALLOC_N) return BA_ERR_RESOURCE_EXHAUSTED;
    if (n == 0) return BA_OK;
    if (!a || !b || !x) return BA_ERR_INVALID_CONFIG;
    if (hipSetDevice(h->device) != hipSuccess) return BA_ERR_INTERNAL;
    double* base = scratch_of(h->device);  // the device address, never the host-side name
    if (!base) return BA_ERR_INTERNAL;
    // wait for the caller's work on the inputs, as the other device-buffer entry points do
    if (hipDeviceSynchronize() != hipSuccess) return BA_ERR_INTERNAL;
    if (cholesky_work_doubles((uint32_t)n) > WORK_DOUBLES) return BA_ERR_INTERNAL;  // the layout above is stale
    hipStream_t s = h->stream;
    double* work = base + MATRIX_DOUBLES;
    int* fail = reinterpret_cast<int*>(work + WORK_DOUBLES);
    int* done = fail + 1;
    const uint32_t m = (uint32_t)n;
    double* out = h->pinned;  // 4 doubles of pinned host memory from ba_hip_create
    if (precision == BA_PRECISION_MIXED) {  // FP32 factor in the scratch, FP64 refinement against `a`
        float* F = reinterpret_cast<float*>(base);
        launch_cholesky_solve_mixed(s, a, F, m, m, b, x, work, fail, done, 8);
        if (hipMemcpyAsync(out, work + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(out + 3, fail, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return BA_ERR_INTERNAL;
        int f = 0;
        __builtin_memcpy(&f, out + 3, sizeof(int));
        const double rel = __builtin_sqrt(out[1] / out[2]);
        return (f || !(rel <= 1e-10)) ? BA_ERR_NUMERICAL : BA_OK;
    }
    if (hipMemcpyAsync(base, a, n * n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return BA_ERR_INTERNAL;
    launch_cholesky_solve(s, base, m, m, b, x, work, fail);
    if (hipMemcpyAsync(out, fail, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return BA_ERR_INTERNAL;
    int f = 0;
    __builtin_memcpy(&f, out, sizeof(int));
    return f ? BA_ERR_NUMERICAL : BA_OK;
}
