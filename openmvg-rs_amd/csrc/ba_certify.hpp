// ba_certify.hpp — host side of the post-loop certification of a native self-calibrating solve:
// owned problems and solutions for the re-solves, the weak-point verdict, the intrinsic
// marginals' verdict and the report bookkeeping.
//
// Follows crates/openmvg-ba/src:
//   native/mod.rs:252-350           solve_with_quarantine, needs_resolve
//   native/self_calibration.rs:223-303  unconstrained_intrinsics, the held re-solve
//   problem.rs:650-701              BundleProblem::without_points
//   solver.rs:527-538, 891-935      SolveReport::continued_from, Solution::lifted
//   solver.rs:1030-1101             Solution::canonicalize_roundoff_non_descent
//   solver.rs:1149-1219             Solution::weak_points
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/ba_hip.h"
#include "ba_host.hpp"

namespace ba {

// Weak-point re-solves and their iteration cap (quarantine.rs:22-27).
constexpr uint32_t MAX_RESOLVES = 4;
constexpr uint32_t RESOLVE_ITERATIONS = 8;
// Largest accepted marginal deviation of an intrinsic's effect at the image corner
// (native/self_calibration.rs:215-221).
constexpr double MAX_INTRINSIC_CORNER_DISPLACEMENT_STD = 0.01;

// A problem that owns its arrays; `p` views them.
struct OwnedProblem {
    std::vector<double> rot, cen, xyz, pix;
    std::vector<uint8_t> var_rot, var_trans, point_fixed;
    std::vector<ba_intrinsic> intr;
    std::vector<uint32_t> obs_pose, obs_intr, obs_point;
    ba_problem p{};
    void bind();
};

// A solution in the caller's frame that owns its arrays, with its report.
struct HostSolution {
    std::vector<double> rot, cen, focal, k1, k2, xyz, cov;
    std::vector<uint8_t> quarantined, unconstrained;
    ba_solve_report rep{};
    void alloc(const ba_problem* p);
    ba_solution view();
    void copy_to(ba_solution* out) const;  // the caller's buffers (optional ones may be NULL)
};

// BundleProblem::without_points: the problem without the points in `drop` (a mask over p's
// points) and their observations, its state taken from `state` (NULL: p's own); kept[i] = the
// p point of new point i.
void without_points(const ba_problem* p, const HostSolution* state, const std::vector<uint8_t>& drop,
                    OwnedProblem& out, std::vector<uint32_t>& kept);

// Solution::weak_points: free points whose 3x3 information at the solution is missing, non-finite
// or below 1e-5 of its largest eigenvalue, and points the solution's geometry makes degenerate.
std::vector<uint8_t> weak_points(const ba_problem* p, const HostSolution& s);

// Whether a solve failed its certificate in a way held-out points can repair.
bool needs_resolve(const ba_solve_report& r);

// Groups whose marginal deviation at the image corner exceeds the bound (or is NaN); every free
// group when `cov` is NULL (the marginals could not be formed).  L: the lowering the marginals
// were formed on, its group parameters at the final state.
std::vector<uint32_t> unconstrained_groups(const ba_problem* p, const Lowered& L, const double* final_gp,
                                           const double* cov);

// SolveReport::continued_from: `later` continues `first` (initial metrics from `first`, work
// counts added).
void continue_report(ba_solve_report& later, const ba_solve_report& first);

// Symmetric 3x3 eigenvalues by cyclic Jacobi rotations (row-major input), ascending.
void symmetric_eigenvalues3(const double m[9], double out[3]);

}  // namespace ba
