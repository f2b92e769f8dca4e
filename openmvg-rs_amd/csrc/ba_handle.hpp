// ba_handle.hpp — the opaque handle behind ba_hip_handle*: one device and one non-blocking stream.
#pragma once

#include <hip/hip_runtime.h>

struct ba_hip_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int schur_product = 0;       // ba_schur_product requested for iterative solves
    int last_schur_product = 0;  // the form the last iterative solve used
};
