/*
 * ba_oracle.h — CPU oracle of the openmvg-ba self-calibrating Levenberg–Marquardt path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain sequential FP64 restatement of the reference
 * algorithm (openmvg-rs, crates/openmvg-ba), written to check the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the
 * checker or as the timed CPU baseline; the product never links or calls it.
 *
 * The problem, solution and report structs are the product's boundary types
 * (include/ba_hip.h) so that one set of caller buffers feeds both implementations.
 *
 * Parity pins (what fixes each function independently of this file; the test that does it):
 *   radial model      closed form; central differences and a dual-number oracle
 *                     (self_calibration/model.rs:396-467, 598-634): test_oracle_pins.py
 *   cost evaluation   golden initial cost 863.6448936532983 of tests/self_calibration.rs:135:
 *                     test_host.py
 *   robust losses     derivative identities (numerics/robust.rs:50-131): test_oracle_pins.py
 *   point block       known inverse Q diag(1e6,1,1e-3) Q^T (point_block.rs:237-266)
 *   Nielsen/run state tables of controller.rs:93-115 and run_state.rs:187-350
 *   stages            continuation.rs:81-111
 *   index maps        topology example (workspace/topology.rs:611-629), product vs oracle
 *                     lowering bit for bit: test_host.py
 *   Jacobians         central differences of the residuals along the reference's retraction
 *                     coordinates, the FixedBaseline S^2 chart included:
 *                     test_oracle_assembly.py
 *   Schur assembly    dense normal equations of the full Jacobian in numpy with the reference's
 *                     damping clamps and numpy's point-block inverses (S, g^, gradient,
 *                     diagonal to 1e-10): test_oracle_assembly.py
 *   camera solve      LAPACK (numpy/scipy) Cholesky: test_oracle_pins.py
 *   PCG               at rtol 1e-10 equals numpy's solve of the dense-normal-equation S on scenes
 *                     scaled by 1e-6, 1, 1e6: test_oracle_assembly.py
 *   candidate rule    a scene that holds a point and rejects steps: test_candidate_scene.py
 *                     (its branch decisions are otherwise checked only against the product)
 *   LM driver         truth recovery on the noise-free fixture of
 *                     tests/sparse_self_calibration/main.rs:48-135 (final rmse < 1e-6)
 *   Auto policy       bao_plan: the reference's rule and estimate formulas, and the pattern
 *                     count against the product's independent count: test_plan.py
 *   certificate       bao_certify: FULL_RANK exactly when numpy's smallest eigenvalue of the
 *                     dense-normal-equation S is positive (and RANK_DEFICIENT for a camera seen
 *                     through one point); marginals equal the blocks of numpy's inverse of that
 *                     S to 1e-8: test_certify.py
 *   weak points       bao_weak_points against a numpy restatement with central-difference
 *                     projection Jacobians and numpy eigenvalues: test_certify.py
 *   certified solve   flows of native/mod.rs:252-350 and native/self_calibration.rs:223-303
 *                     (held groups, re-solves, quarantine) on four scenes: test_certify.py;
 *                     the round-off canonicalisation (solver.rs:1030-1101) is not triggered by
 *                     any test scene (parity unpinned)
 */
#ifndef OPENMVG_BA_ORACLE_H
#define OPENMVG_BA_ORACLE_H

#include "../include/ba_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full solve: NativeBundleAdjuster::solve on the self-calibrating pixel path. */
int bao_solve(const ba_problem* problem, const ba_solve_config* config, ba_solution* solution,
              ba_solve_report* report, ba_iteration_trace* trace, char* message,
              size_t message_len);

/* LinearSolver::Auto's camera-step choice, same contract as ba_hip_plan. */
int bao_plan(const ba_problem* problem, const ba_solve_config* config, ba_execution_plan* plan,
             char* message, size_t message_len);

/* The certificate of a native state at zero damping: the rank verdict (enum ba_observability)
 * of the undamped reduced system and the weight-normalised marginal covariance of every free
 * intrinsic group (9 per group, NaN for fixed groups); BA_ERR_NUMERICAL when the marginals cannot
 * be formed (sparse/self_calibration/mod.rs:110-209, direct.rs:288-334). */
int bao_certify(const ba_device_problem* problem, int32_t* observability, double* covariance);

/* Solution::weak_points (solver.rs:1149-1219) of a solution of `problem`: weak[i] = 1 for each
 * weak point. */
int bao_weak_points(const ba_problem* problem, const ba_solution* solution, uint8_t* weak);

/* Lowering and index maps, same contract as ba_hip_prepare. */
int bao_prepare(const ba_problem* problem, ba_native_sizes* sizes, ba_native_problem* native,
                char* message, size_t message_len);

/* Kernel (a) on host buffers (ba_device_problem used as a host view). */
int bao_linearize(const ba_device_problem* problem, double* residual, double* point_jacobian,
                  double* camera_jacobian, double* cost, uint8_t* valid);

/* Kernel (b): materialised Schur system (schur/assembly.rs:832-1069) at damping lambda.
 * schur is D*D row-major, both triangles written.  Optional outputs may be NULL:
 * point_inverse (6 per point, packed V^-1), inverse_gradient (3 per point),
 * inverse_cross (3 per point parameter, R = V^-1 W rows in cross-offset order). */
int bao_assemble(const ba_device_problem* problem, double lambda, double* schur, double* gradient,
                 double* full_gradient, double* diagonal, double* point_inverse,
                 double* inverse_gradient, double* inverse_cross, ba_schur_scalars* scalars);

/* Kernel (c): retained block-Jacobi PCG (self_calibration_retained.rs:512-579, pcg.rs:23-119)
 * on the assembly at damping lambda. */
int bao_pcg(const ba_device_problem* problem, double lambda, double relative_tolerance,
            uint32_t max_iterations, double* step, uint32_t* iterations,
            double* relative_residual, int32_t* accepted);

/* Kernel (d): dense Cholesky solve of dense/linear_solver.rs:9-61; returns 1 on success. */
int bao_cholesky_solve(size_t n, const double* a, const double* b, double* x);

/* Scalar pieces, for the golden-vector tests. */
int bao_radial_linearization(const double camera[3], const double observation[2],
                             const double point[3], double residual[2], double jacobian[6],
                             double intrinsic_jacobian[6]);
int bao_radial_projection(const double camera[3], const double observation[2],
                          const double point[3], double residual[2], double jacobian[6]);
int bao_radial_bearing(const double camera[3], const double measurement[2], double bearing[3]);
double bao_first_positive_radial_turn(const double camera[3]); /* NaN when none */
double bao_loss_cost(int32_t loss, double scale, double residual);
double bao_loss_sqrt_weight(int32_t loss, double scale, double residual);
void bao_exp_so3(const double vector[3], double rotation[9]);
void bao_s2_tangent_basis(const double t[3], double basis[6]);
void bao_retract_s2(const double t[3], const double basis[6], const double d[2], double out[3]);
int bao_point_block_factor(const double matrix[9], double damping, double inverse[6]);
double bao_point_block_rcond(const double matrix[9]);
/* TrackGeometry (point_geometry.rs:29-101): returns bit0 = degenerate, bit1 = collapsed. */
int bao_track_geometry(size_t count, const double* offsets, double scene_length);
/* NielsenController::accept on (value, nu); returns accepted and updates the pair. */
int bao_nielsen_accept(double* value, double* nu, double actual, double predicted, double* gain);
void bao_nielsen_reject(double* value, double* nu);
/* stages() of continuation.rs:26-74: writes 3 (loss, iterations) pairs. */
void bao_stages(int32_t loss, uint32_t max_iterations, int32_t loss_out[3], uint32_t iterations_out[3]);
/* Robust reprojection objective sum_o rho(|r_o|) of a BAL-style scene (cameras 9 values each:
 * angle-axis w, t, f, k1, k2; X_c = exp(w) X + t) and its gradient; gcam/gpt may be NULL. */
int bao_cost_grad(size_t n_cameras, size_t n_points, size_t n_observations, const double* cameras,
                  const double* points, const uint32_t* obs_camera, const uint32_t* obs_point,
                  const double* obs_uv, int32_t loss, double loss_scale, double* cost, uint32_t* invalid,
                  double* grad_camera, double* grad_point);
/* Dense cost of a self-calibrating problem in the reference's native frame with an optional
 * radial intrinsic prior (schur/priors.rs:302-331), for the golden initial cost. */
double bao_native_cost(const ba_device_problem* problem, size_t n_priors, const uint32_t* prior_group,
                       const double* prior_mean /*3 per prior*/, const double* prior_sqrt_info /*9*/);

#ifdef __cplusplus
}
#endif
#endif
