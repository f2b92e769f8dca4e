// ba_policy.hpp — host decisions of the LM loop that need no device (checked on the host alone).
#pragma once

#include <cstdint>

#include "../../include/ba_hip.h"

namespace ba {

// The PCG product of the next LM iteration, chosen before its assembly.  With the explicit
// reduced system set up, EXPLICIT takes it always, AUTO when the PCG is expected to take at least
// EXPLICIT_MIN_PCG iterations (the iteration count of the last PCG solve: 0 before the first
// solve and after one whose preconditioner failed), IMPLICIT never.  The explicit
// assembly costs about as much as ten matrix-free products save (2.1 ms against 0.21 ms per PCG
// iteration at the 2,000-camera scene), so solves that need only a few PCG iterations stay
// matrix-free.
constexpr uint32_t EXPLICIT_MIN_PCG = 10;
inline bool use_explicit_product(bool explicit_set_up, int32_t mode, uint32_t expected_pcg) {
    if (!explicit_set_up || mode == BA_SCHUR_IMPLICIT) return false;
    return mode == BA_SCHUR_EXPLICIT || expected_pcg >= EXPLICIT_MIN_PCG;
}

}  // namespace ba
