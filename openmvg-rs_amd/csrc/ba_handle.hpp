// ba_handle.hpp — the opaque handle behind ba_hip_handle*: one device and one non-blocking stream.
#pragma once

#include <hip/hip_runtime.h>

struct ba_hip_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int schur_product = 0;       // ba_schur_product requested for iterative solves
    int last_schur_product = 0;  // the form the last iterative solve used
    double* pinned = nullptr;    // 4 doubles of pinned host memory (status readback of the
                                 // allocation-free entry point)
};

namespace ba {
// Resolve the calling thread's current device's copy of the allocation-free entry point's
// scratch array (once per device; ba_noalloc.hip).
void noalloc_prepare_device(int device);
}  // namespace ba
