// ba_kernels.hip — HIP kernels (gfx950) of the Levenberg–Marquardt inner loop.
//
//   (a)  k_linearize      one lane per observation: robust-weighted residual, point and camera
//                         Jacobians (schur/point.rs:239-355, self_calibration/model.rs:26-129)
//   (b)  k_point_elim     one lane per point: V, g_p, damped 3x3 inverse, V^-1 g_p and the
//                         retained rows R = V^-1 W (schur/assembly.rs:893-1039, point_block.rs)
//        k_block_gradient one workgroup per camera block: reduced/full gradient and diagonal
//        k_pairs          one workgroup per lower block pair of the reduced system: U and
//                         W V^-1 W^T by a segmented reduction over pre-sorted observation pairs
//   (a') k_candidate      one lane per point: cached back-substitution and the candidate rule
//                         (sparse/self_calibration/candidate.rs:25-233, point_geometry.rs:153-168)
//   (c)  k_pcg_*          implicit-Schur block-Jacobi PCG (self_calibration_retained.rs:512-887,
//                         pcg.rs:23-132) with its control state kept on the device
// Every reduction runs over a fixed partition in a fixed order; no floating-point atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ba_kernels.hpp"
#include "ba_math.hpp"

namespace ba {

namespace {

constexpr int RED_BLOCKS = 512;  // fixed first-stage partition of the scalar reductions
constexpr int NT = 256;

inline unsigned grid_for(size_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }

// Butterfly sum over the 64 lanes of a wavefront: every lane ends with the same total (a + b and
// b + a round alike), in an order fixed by the lane numbering.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Fixed-shape tree sum over the N threads of a block (one wavefront: the butterfly, no barrier).
template <int N>
__device__ double tree_sum(double v, double* sh) {
    if constexpr (N == 64) return wave_sum(v);
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}
template <int N>
__device__ double tree_max(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}
template <int N>
__device__ double tree_min(double v, double* sh) {
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmin(sh[t], sh[t + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Parameter index of row k (0..8) of observation o's 9-row camera Jacobian, or -1.
__device__ __forceinline__ int row_param(const Topo& tp, uint32_t cam, int k) {
    if (k < 6) return k < (int)tp.cam_dim[cam] ? (int)(tp.cam_off[cam] + k) : -1;
    uint32_t g = tp.cam_group[cam];
    int l = k - 6;
    return l < (int)tp.grp_dim[g] ? (int)(tp.grp_off[g] + l) : -1;
}

// Values of v on the 9 camera-Jacobian columns of camera c (0 beyond the free dimensions).
__device__ __forceinline__ void camera_values(const Topo& tp, uint32_t c, const double* v, double vc[9]) {
    const uint32_t pd = tp.cam_dim[c], po = tp.cam_off[c], g = tp.cam_group[c];
    const uint32_t gd = tp.grp_dim[g], go = tp.grp_off[g];
#pragma unroll
    for (int k = 0; k < 6; ++k) vc[k] = k < (int)pd ? v[po + k] : 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) vc[6 + k] = k < (int)gd ? v[go + k] : 0.0;
}

}  // namespace

namespace {
// ---- compact linearisation ----
// Per observation the linearisation keeps n = (x/z, y/z) of the camera-frame point, h = sqrt(w)/z
// and sqrt(w) (rows SOA_Q..+3).  With r2 = |n|^2, d = 1 + k1 r2 + k2 r2^2, rd = k1 + 2 k2 r2 and
// A = f (d I + 2 rd n n^T) (the derivative of the distorted projection), the robust-weighted
// camera-frame projection Jacobian is sqrt(w) J_p = h A [I | -n] (model.rs:26-129), and
//   J~_X            = s h A [I | -n] R                  (point columns, scene scale s),
//   J~_c rotation   = h A [I | -n] (-[R X]x)            (pose motion of schur/point.rs:730-762),
//   J~_c translation = s h A [I | -n]                   (baseline camera: h A [I | -n] b B),
//   J~_c intrinsic  = sqrt(w) f [d n, r2 n, r2^2 n]     (log f, k1, k2).
// So a product with J~_c or J~_X needs, per observation, these four values, R X (the rotated
// point) and the camera's R, f, k1, k2; the passes rebuild the Jacobians instead of reading them.

// Camera record (CAM_REC doubles): R (9), f, k1, k2, then a camera-frame motion: omega (3),
// tau (3) (the translation motion: s dt, or b B dt for the baseline camera, 0 without pose) and
// the intrinsic values (3).
constexpr int CAM_REC = (int)PCG_CAM_REC;
struct Cam {
    double R[9], f, k1, k2, om[3], tau[3], di[3];
};
__device__ __forceinline__ void load_cam(const double* __restrict__ src, Cam& c) {
    const double2* s = reinterpret_cast<const double2*>(src);
    double v[22];
#pragma unroll
    for (int i = 0; i < 11; ++i) {
        const double2 x = s[i];
        v[2 * i] = x.x;
        v[2 * i + 1] = x.y;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) c.R[i] = v[i];
    c.f = v[9], c.k1 = v[10], c.k2 = v[11];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        c.om[i] = v[12 + i];
        c.tau[i] = v[15 + i];
        c.di[i] = v[18 + i];
    }
}
// Record of camera c for the pose values vp (cam_dim of 6: rotation, then translation) and the
// intrinsic values vi; R the camera's rotation, g its group's (f, k1, k2), gg its gauge.
__device__ void camera_record(const Topo& tp, uint32_t c, const double* R, const double* g, const GaugeDev& gg,
                              const double vp[6], const double vi[3], double* dst) {
    const uint32_t pd = tp.cam_dim[c];
    double v[CAM_REC];
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = R[i];
    v[9] = g[0], v[10] = g[1], v[11] = g[2];
    for (int i = 0; i < 3; ++i) {
        v[12 + i] = pd > 0 ? vp[i] : 0.0;
        v[15 + i] = 0.0;
        v[18 + i] = vi[i];
    }
    if (pd > 0) {
        if ((int)c == gg.camera) {
            for (int i = 0; i < 3; ++i) v[15 + i] = gg.baseline * (gg.basis[i] * vp[3] + gg.basis[3 + i] * vp[4]);
        } else {
            for (int i = 0; i < 3; ++i) v[15 + i] = gg.scene_scale * vp[3 + i];
        }
    }
    v[21] = v[22] = v[23] = 0.0;
    double2* o = reinterpret_cast<double2*>(dst);
#pragma unroll
    for (int i = 0; i < CAM_REC / 2; ++i) o[i] = make_double2(v[2 * i], v[2 * i + 1]);
}

// World-frame record of a camera motion (WREC doubles): Om = R^T omega / s, T = R^T tau / s and
// the intrinsic terms A = f d0, B = f (k1 d0 + d1), C = f (k2 d0 + d2), so that with the
// point-major J~_X and the compact state (n, sqrt(w)):
//   J~_c v = J~_X (Om x X + T) + sqrt(w) n (A + r2 B + r2^2 C).
// Built from a camera record (camera frame) and the scene scale.
constexpr int WREC = (int)PCG_WREC;
constexpr int WREC_PAD = 10;  // a world record padded to 16-byte groups (step records)
__device__ __forceinline__ void world_from_camera(const Cam& c, double s, double w[WREC]) {
    double om[3], tr[3];
    rotate_t(c.R, c.om, om);
    rotate_t(c.R, c.tau, tr);
    const double is = 1.0 / s;
    for (int i = 0; i < 3; ++i) {
        w[i] = om[i] * is;
        w[3 + i] = tr[i] * is;
    }
    w[6] = c.f * c.di[0];
    w[7] = c.f * (c.k1 * c.di[0] + c.di[1]);
    w[8] = c.f * (c.k2 * c.di[0] + c.di[2]);
}
// e = J~_X^T (J~_c v) for one observation from J~_X (jx), n, sqrt(w), its point X and the
// camera's world record w.
__device__ __forceinline__ void world_product(const double jx[6], double n0, double n1, double sw, const double X[3],
                                              const double* w, double e[3]) {
    const double u0 = w[1] * X[2] - w[2] * X[1] + w[3];
    const double u1 = w[2] * X[0] - w[0] * X[2] + w[4];
    const double u2 = w[0] * X[1] - w[1] * X[0] + w[5];
    const double r2 = n0 * n0 + n1 * n1;
    const double si = sw * (w[6] + r2 * (w[7] + r2 * w[8]));
    const double y0 = jx[0] * u0 + jx[1] * u1 + jx[2] * u2 + si * n0;
    const double y1 = jx[3] * u0 + jx[4] * u1 + jx[5] * u2 + si * n1;
#pragma unroll
    for (int i = 0; i < 3; ++i) e[i] = jx[i] * y0 + jx[3 + i] * y1;
}

// The distortion terms of one observation.
struct Obs {
    double n0, n1, h, sw, r2, d, rd;
};
__device__ __forceinline__ Obs make_obs(double n0, double n1, double h, double sw, const Cam& c) {
    Obs o;
    o.n0 = n0, o.n1 = n1, o.h = h, o.sw = sw;
    o.r2 = n0 * n0 + n1 * n1;
    o.d = 1.0 + c.k1 * o.r2 + c.k2 * (o.r2 * o.r2);
    o.rd = c.k1 + 2.0 * (c.k2 * o.r2);
    return o;
}
// a = A b for the 2-vector b.
__device__ __forceinline__ void apply_A(const Obs& o, const Cam& c, double b0, double b1, double* a0, double* a1) {
    const double t = 2.0 * o.rd * (o.n0 * b0 + o.n1 * b1);
    *a0 = c.f * (o.d * b0 + t * o.n0);
    *a1 = c.f * (o.d * b1 + t * o.n1);
}
// y = J~_c v for the camera's motion record (om, tau, di) and the rotated point RX.
__device__ __forceinline__ void camera_product(const Obs& o, const Cam& c, const double RX[3], double* y0,
                                               double* y1) {
    const double q0 = c.om[1] * RX[2] - c.om[2] * RX[1] + c.tau[0];
    const double q1 = c.om[2] * RX[0] - c.om[0] * RX[2] + c.tau[1];
    const double q2 = c.om[0] * RX[1] - c.om[1] * RX[0] + c.tau[2];
    double a0, a1;
    apply_A(o, c, q0 - o.n0 * q2, q1 - o.n1 * q2, &a0, &a1);
    const double si = o.sw * c.f * (o.d * c.di[0] + o.r2 * c.di[1] + (o.r2 * o.r2) * c.di[2]);
    *y0 = o.h * a0 + si * o.n0;
    *y1 = o.h * a1 + si * o.n1;
}
// m = sqrt(w) J_p^T z (camera frame, 3-vector).
__device__ __forceinline__ void projection_transpose(const Obs& o, const Cam& c, double z0, double z1, double m[3]) {
    double a0, a1;
    apply_A(o, c, z0, z1, &a0, &a1);
    m[0] = o.h * a0;
    m[1] = o.h * a1;
    m[2] = -o.h * (o.n0 * a0 + o.n1 * a1);
}
// e = J~_X^T y (world frame, scaled): s R^T (sqrt(w) J_p^T y).
__device__ __forceinline__ void point_transpose(const Obs& o, const Cam& c, double s, double y0, double y1,
                                                double e[3]) {
    double m[3];
    projection_transpose(o, c, y0, y1, m);
    double rt[3];
    rotate_t(c.R, m, rt);
    for (int i = 0; i < 3; ++i) e[i] = s * rt[i];
}
// t = J~_X w for a point-space vector w (world frame).
__device__ __forceinline__ void point_product(const Obs& o, const Cam& c, double s, const double w[3], double* t0,
                                              double* t1) {
    double rw[3];
    rotate(c.R, w, rw);
    double a0, a1;
    apply_A(o, c, rw[0] - o.n0 * rw[2], rw[1] - o.n1 * rw[2], &a0, &a1);
    *t0 = s * o.h * a0;
    *t1 = s * o.h * a1;
}

// Groups of 16 bytes of a structure-of-arrays row: VN consecutive observations per load.
template <class T>
struct Vec;
template <>
struct Vec<double> {
    using V = double2;
    using I = uint2;
    static constexpr int N = 2;
    __device__ static double at(const V& v, int i) { return i == 0 ? v.x : v.y; }
    __device__ static uint32_t at(const I& v, int i) { return i == 0 ? v.x : v.y; }
};
template <>
struct Vec<float> {
    using V = float4;
    using I = uint4;
    static constexpr int N = 4;
    __device__ static double at(const V& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
    __device__ static uint32_t at(const I& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
};
template <class T>
__device__ __forceinline__ typename Vec<T>::V load_group(const T* __restrict__ rows, int row, uint32_t j) {
    return *reinterpret_cast<const typename Vec<T>::V*>(rows + soa_at(row, j));
}
}  // namespace

// ----------------------------------------------------------------------------------------------
// Gauge state (schur/layout.rs:101-113): baseline, direction and S^2 tangent basis of the
// FixedBaseline camera at the current state.
// ----------------------------------------------------------------------------------------------
template <class T>
__global__ void k_gauge(const T* cam_t, int camera, double scene_scale, GaugeDev* out) {
    double t[3] = {0.0, 0.0, 0.0};
    if (camera >= 0)
        for (int k = 0; k < 3; ++k) t[k] = (double)cam_t[3 * camera + k];
    *out = make_gauge(t, camera, scene_scale);
}

// ----------------------------------------------------------------------------------------------
// (a) k_linearize: one lane per observation.
// State is FP64; the camera-frame point, residual and robust cost are evaluated in FP64 and the
// Jacobians in T.  Outputs are robust-weighted: r~ = sqrt(w) r, J~ = sqrt(w) J.
// ----------------------------------------------------------------------------------------------
namespace {
// Linearisation of observation o (point-major index) at the state (R, t, gp, X): robust-weighted
// residual r~, point Jacobian J~_X (2x3, world frame, times the scene scale) and camera Jacobian
// J~_c (2x9: pose columns 0..5 by the pose motion of schur/point.rs:730-762, intrinsic columns
// 6..8), zero beyond the camera's and group's free dimensions.  False outside the domain.
__device__ __forceinline__ bool linearize_obs(const Topo& tp, uint32_t c, uint32_t p, double mu, double mv,
                                              const double* __restrict__ Rg, const double* __restrict__ tg,
                                              const double* __restrict__ gp, const double* __restrict__ X,
                                              const GaugeDev* __restrict__ gauges, double r[2], double jx[6],
                                              double jc[18], double q[4], double rx[3], double* cost, double* sq) {
    const uint32_t g = tp.cam_group[c];
    const GaugeDev* gauge = gauges + (tp.cam_win ? tp.cam_win[c] : 0);
    // everything read after the domain test is loaded here, with the camera and the point, so
    // that one memory round trip serves the whole observation
    const uint32_t pd = tp.cam_dim[c], idim = tp.grp_dim[g];
    const double s = gauge->scene_scale;
    const int gauge_cam = gauge->camera;
    const double* R = Rg + 9 * (size_t)c;
    double Xp[3] = {X[3 * (size_t)p], X[3 * (size_t)p + 1], X[3 * (size_t)p + 2]};
    double rot[3];
    rotate(R, Xp, rot);
    double cp[3] = {rot[0] + tg[3 * c], rot[1] + tg[3 * c + 1], rot[2] + tg[3 * c + 2]};
    const double f = gp[3 * g], k1 = gp[3 * g + 1], k2 = gp[3 * g + 2];
    const double turn = tp.gturn[g];  // first radial turn of the group (+inf: none)
    const bool has_turn = turn < INFINITY;
    double res[2], J[2][3], Ji[2][3];
    if (!radial_linearize<double>(f, k1, k2, turn, has_turn, cp, mu, mv, res, J, Ji))
        return false;
    double rn = sqrt(res[0] * res[0] + res[1] * res[1]);
    double sw = loss_sqrt_weight<double>(tp.loss, tp.loss_scale, rn);
    *cost = loss_cost<double>(tp.loss, tp.loss_scale, rn);
    *sq = rn * rn;
    r[0] = sw * res[0];
    r[1] = sw * res[1];
    // compact state: n, sqrt(w)/z, sqrt(w); and the rotated point
    const double iz = 1.0 / cp[2];
    q[0] = cp[0] * iz;
    q[1] = cp[1] * iz;
    q[2] = sw * iz;
    q[3] = sw;
    for (int k = 0; k < 3; ++k) rx[k] = rot[k];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double v = J[a][0] * R[k] + J[a][1] * R[3 + k] + J[a][2] * R[6 + k];
            jx[3 * a + k] = sw * (v * s);
        }
    // pose motion: rotation columns -[RX]_x, translation columns (scene scale, or the baseline
    // camera's S^2 tangent basis times the baseline)
    double mot[3][6] = {{0.0, rot[2], -rot[1], 0, 0, 0}, {-rot[2], 0.0, rot[0], 0, 0, 0}, {rot[1], -rot[0], 0.0, 0, 0, 0}};
    if ((int)c == gauge_cam) {
        for (int q = 0; q < 3; ++q) {
            mot[q][3] = gauge->baseline * gauge->basis[q];
            mot[q][4] = gauge->baseline * gauge->basis[3 + q];
        }
    } else {
        for (int q = 0; q < 3; ++q) mot[q][3 + q] = s;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double v = (k < (int)pd) ? (J[a][0] * mot[0][k] + J[a][1] * mot[1][k] + J[a][2] * mot[2][k]) : 0.0;
            jc[9 * a + k] = sw * v;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) jc[9 * a + 6 + k] = k < (int)idim ? sw * Ji[a][k] : 0.0;
    }
    return true;
}
}  // namespace

// ----------------------------------------------------------------------------------------------
// (a) k_linearize: one lane per observation (point-major order).
// State is FP64; the camera-frame point, residual and robust cost are evaluated in FP64 and the
// Jacobians stored in T.  Outputs are robust-weighted: r~ = sqrt(w) r, J~ = sqrt(w) J.  With
// out.jpm set, the Jacobians are also written as structure-of-arrays rows (coalesced).
// ----------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(NT) k_linearize(Topo tp, const double* __restrict__ Rg, const double* __restrict__ tg,
                                                  const double* __restrict__ gp, const double* __restrict__ X,
                                                  const double* __restrict__ uv, const GaugeDev* __restrict__ gauge,
                                                  Lin<T> out) {
    uint32_t o = blockIdx.x * NT + threadIdx.x;
    if (o >= tp.no) return;
    double r[2], jx[6], jc[18], q[4], rx[3], cost = 0.0, sq = 0.0;
    const bool ok = linearize_obs(tp, tp.obs_cam[o], tp.obs_pt[o], uv[2 * (size_t)o], uv[2 * (size_t)o + 1], Rg, tg,
                                  gp, X, gauge, r, jx, jc, q, rx, &cost, &sq);
    if (!ok) {
        for (int k = 0; k < 2; ++k) r[k] = 0.0;
        for (int k = 0; k < 6; ++k) jx[k] = 0.0;
        for (int k = 0; k < 18; ++k) jc[k] = 0.0;
        for (int k = 0; k < 4; ++k) q[k] = 0.0;
        cost = sq = 0.0;
    }
    out.valid[o] = ok ? 1 : 0;
    out.cost[o] = cost;
    out.sq[o] = sq;
    if (out.rj) {
        T* rj = out.rj + 8 * (size_t)o;
        rj[0] = (T)r[0];
        rj[1] = (T)r[1];
#pragma unroll
        for (int k = 0; k < 6; ++k) rj[2 + k] = (T)jx[k];
    }
    if (out.jc) {
        T* jo = out.jc + 18 * (size_t)o;
#pragma unroll
        for (int k = 0; k < 18; ++k) jo[k] = (T)jc[k];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) out.jpm[soa_at(SOA_R + k, o)] = (T)r[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) out.jpm[soa_at(SOA_Q + k, o)] = (T)q[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) out.jpm[soa_at(SOA_JX + k, o)] = (T)jx[k];
    if (out.pose_rows)
#pragma unroll
        for (int k = 0; k < 18; ++k) out.jpm[soa_at(SOA_JC + k, o)] = (T)jc[k];
}

// Camera-major twin of k_linearize: lane j linearises observation cm_obs[j] and writes r~, the
// compact state and the rotated point R X of the camera-major copy (coalesced reads of the
// camera-major camera, point and pixel arrays and coalesced writes; the point is gathered).
template <class T>
__global__ void __launch_bounds__(NT) k_linearize_cm(Topo tp, const double* __restrict__ Rg,
                                                     const double* __restrict__ tg, const double* __restrict__ gp,
                                                     const double* __restrict__ X, const GaugeDev* __restrict__ gauge,
                                                     Lin<T> out) {
    uint32_t j = blockIdx.x * NT + threadIdx.x;
    if (j >= tp.no) return;
    double r[2], jx[6], jc[18], q[4], rx[3], cost, sq;
    // camera, point and pixel of position j from the camera-major copies (coalesced)
    if (!linearize_obs(tp, tp.cm_cam[j], tp.cm_pt[j], tp.cm_uv[2 * (size_t)j], tp.cm_uv[2 * (size_t)j + 1], Rg, tg, gp,
                       X, gauge, r, jx, jc, q, rx, &cost, &sq)) {
        for (int k = 0; k < 2; ++k) r[k] = 0.0;
        for (int k = 0; k < 4; ++k) q[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) out.jcm[soa_at(SOA_R + k, j)] = (T)r[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) out.jcm[soa_at(SOA_Q + k, j)] = (T)q[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) out.jcm[soa_at(SOA_RX + k, j)] = (T)rx[k];
}

// ----------------------------------------------------------------------------------------------
// (b) k_point_elim: one workgroup per point chunk.  Lanes read one observation each from the
// point-major structure-of-arrays copy and form its contributions to V = J~_X^T J~_X,
// g_p = J~_X^T r~, cost and squared residual; one lane per point then sums its observations'
// contributions in observation order (a segmented reduction through LDS) and eliminates the point
// (point_block.rs:85-168, schur/assembly.rs:893-1039).  With pe.rrows set, lanes finally write the
// retained rows R_o = V^-1 W_o, W_o = J~_c,o^T J~_X,o, of the block-pair path.
// ----------------------------------------------------------------------------------------------
namespace {
constexpr int PE_V = 12;  // V (6), g (3), cost, squared residual, invalid observations
template <class T>
__device__ __forceinline__ void point_contrib(const Lin<T>& lin, size_t no, uint32_t o, double c[PE_V]) {
    const bool valid = lin.valid[o] != 0;
    const double r0 = (double)lin.jpm[soa_at(SOA_R + 0, o)], r1 = (double)lin.jpm[soa_at(SOA_R + 1, o)];
    double a[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) a[r][k] = (double)lin.jpm[soa_at(SOA_JX + 3 * r + k, o)];
    c[0] = a[0][0] * a[0][0] + a[1][0] * a[1][0];
    c[1] = a[0][1] * a[0][0] + a[1][1] * a[1][0];
    c[2] = a[0][1] * a[0][1] + a[1][1] * a[1][1];
    c[3] = a[0][2] * a[0][0] + a[1][2] * a[1][0];
    c[4] = a[0][2] * a[0][1] + a[1][2] * a[1][1];
    c[5] = a[0][2] * a[0][2] + a[1][2] * a[1][2];
    c[6] = a[0][0] * r0 + a[1][0] * r1;
    c[7] = a[0][1] * r0 + a[1][1] * r1;
    c[8] = a[0][2] * r0 + a[1][2] * r1;
    c[9] = (double)lin.cost[o];
    c[10] = (double)lin.sq[o];
    c[11] = valid ? 0.0 : 1.0;
}
}  // namespace

template <class T>
__global__ void __launch_bounds__(NT) k_point_elim(Topo tp, Lin<T> lin, T lambda, int want_rcond, PointElim<T> pe) {
    __shared__ double sh[PE_V][NT];
    __shared__ double shinv[6][NT];
    const uint32_t k = blockIdx.x, t = threadIdx.x;
    const uint32_t ob = tp.chunk_obs[k], oe = tp.chunk_obs[k + 1];
    const uint32_t pb = tp.chunk_pt[k], pend = tp.chunk_pt[k + 1];
    const size_t no = tp.no;
    const bool big = oe - ob > NT;
    double acc[PE_V];
#pragma unroll
    for (int q = 0; q < PE_V; ++q) acc[q] = 0.0;
    for (uint32_t o = ob + t; o < oe; o += NT) {
        double c[PE_V];
        point_contrib(lin, no, o, c);
#pragma unroll
        for (int q = 0; q < PE_V; ++q) acc[q] += c[q];
    }
#pragma unroll
    for (int q = 0; q < PE_V; ++q) sh[q][t] = acc[q];
    __syncthreads();
    if (big) {
        for (int s = NT / 2; s > 0; s >>= 1) {
            if ((int)t < s)
#pragma unroll
                for (int q = 0; q < PE_V; ++q) sh[q][t] += sh[q][t + s];
            __syncthreads();
        }
    }
    double s_[PE_V];
    bool have_point = t < pend - pb;
    const uint32_t p = pb + t;
    if (have_point) {
#pragma unroll
        for (int q = 0; q < PE_V; ++q) s_[q] = 0.0;
        if (big) {
#pragma unroll
            for (int q = 0; q < PE_V; ++q) s_[q] = sh[q][0];
        } else {
            for (uint32_t o = tp.pt_off[p]; o < tp.pt_off[p + 1]; ++o)
#pragma unroll
                for (int q = 0; q < PE_V; ++q) s_[q] += sh[q][o - ob];
        }
    }
    // per-point elimination; shinv[.][t] receives V^-1 of point pb + t (zero when not eliminated)
    bool eliminated = false;
    T inv[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    T ysave[3] = {T(0), T(0), T(0)};
    if (have_point) {
        double* ps = pe.pscal + 6 * (size_t)p;
        T* yp = pe.y + 3 * (size_t)p;
        const T V[6] = {(T)s_[0], (T)s_[1], (T)s_[2], (T)s_[3], (T)s_[4], (T)s_[5]};
        const T gv[3] = {(T)s_[6], (T)s_[7], (T)s_[8]};
        if (s_[11] != 0.0) {
            ps[0] = 0.0;
            ps[1] = 0.0;
            ps[2] = 0.0;
            ps[3] = 0.0;
            ps[4] = 1.0;
            ps[5] = 0.0;  // neither linearised nor eliminated
            yp[0] = yp[1] = yp[2] = T(0);
        } else {
            ps[0] = s_[9];
            ps[1] = s_[10];
            if (tp.pt_fixed[p]) {
                ps[2] = 0.0;
                ps[3] = 0.0;
                ps[4] = 1.0;
                ps[5] = 3.0;  // linearised; nothing to eliminate
                yp[0] = yp[1] = yp[2] = T(0);
            } else {
                // scaled gradient norm of the point block (schur/assembly.rs:1081-1096)
                double d3[3] = {(double)V[0], (double)V[2], (double)V[5]};
                double g3[3] = {(double)gv[0], (double)gv[1], (double)gv[2]};
                double dscale = fmax(fmax(fmax(0.0, d3[0]), d3[1]), d3[2]);
                double sgin;
                if (!(isfinite(dscale) && dscale > 0.0)) {
                    sgin = (g3[0] == 0.0 && g3[1] == 0.0 && g3[2] == 0.0) ? 0.0 : INFINITY;
                } else {
                    double fl = 64.0 * DBL_EPSILON * dscale;
                    sgin = 0.0;
                    for (int q = 0; q < 3; ++q) sgin = fmax(sgin, fabs(g3[q]) / sqrt(fmax(d3[q], fl)));
                }
                ps[3] = sgin;
                // damping diagonal and damped block (point_block.rs:90-110)
                T tscale = fmax(fmax(fmax(T(0), V[0]), V[2]), V[5]);
                bool elim = isfinite(tscale) && tscale > T(0);
                T fl = T(64) * Eps<T>::value * tscale;
                T dd[3] = {fmax(V[0], fl), fmax(V[2], fl), fmax(V[5], fl)};
                if (tp.win_lambda) lambda = (T)tp.win_lambda[tp.pt_win[p]];
                T Vd[6] = {V[0] + lambda * dd[0], V[1], V[2] + lambda * dd[1], V[3], V[4], V[5] + lambda * dd[2]};
                elim = elim && point_block_inverse(Vd, inv);
                if (!elim) {
                    for (int q = 0; q < 6; ++q) inv[q] = T(0);
                    ps[2] = 0.0;
                    ps[4] = 0.0;
                    ps[5] = (lambda == T(0)) ? 5.0 : 1.0;  // linearised, not eliminated, weak at zero damping
                    yp[0] = yp[1] = yp[2] = T(0);
                } else {
                    eliminated = true;
                    double rcond = 1.0;
                    bool weak = false;
                    if (want_rcond) {  // PointBlockFactor::reciprocal_condition_number (point_block.rs:170-183)
                        double m[9] = {(double)V[0], (double)V[1], (double)V[3], (double)V[1], (double)V[2],
                                       (double)V[4], (double)V[3], (double)V[4], (double)V[5]};
                        double iv[9] = {(double)inv[0], (double)inv[1], (double)inv[3], (double)inv[1], (double)inv[2],
                                        (double)inv[4], (double)inv[3], (double)inv[4], (double)inv[5]};
                        double nm = 0.0, ni = 0.0;
                        for (int r = 0; r < 3; ++r) {
                            double s1 = 0.0, s2 = 0.0;
                            for (int c = 0; c < 3; ++c) {
                                s1 += fabs(m[3 * r + c]);
                                s2 += fabs(iv[3 * r + c]);
                            }
                            nm = fmax(nm, s1);
                            ni = fmax(ni, s2);
                        }
                        double den = nm * ni;
                        rcond = (isfinite(den) && den > 0.0) ? 1.0 / den : 0.0;
                        weak = rcond <= 64.0 * DBL_EPSILON * 3.0;
                    }
                    T y[3];
                    sym3_mul(inv, gv, y);
                    ps[2] = 0.5 * (g3[0] * (double)y[0] + g3[1] * (double)y[1] + g3[2] * (double)y[2]);
                    ps[4] = rcond;
                    ps[5] = weak ? 7.0 : 3.0;
                    for (int q = 0; q < 3; ++q) {
                        yp[q] = y[q];
                        ysave[q] = y[q];
                        pe.ddiag[3 * (size_t)p + q] = dd[q];
                    }
                    for (int q = 0; q < 6; ++q) pe.vdamp[6 * (size_t)p + q] = Vd[q];
                }
            }
        }
        // V^-1 (zero for a point that is not eliminated: it then adds nothing to the reduced
        // system) and the packed record [V^-1, y] read by the camera-side passes
        for (int q = 0; q < 6; ++q) pe.vinv[6 * (size_t)p + q] = inv[q];
        if (pe.prec) {
            T* rec = pe.prec + PE_REC * (size_t)p;
            for (int q = 0; q < 6; ++q) rec[q] = inv[q];
            for (int q = 0; q < 3; ++q) rec[6 + q] = ysave[q];
        }
    }
    if (!pe.rrows) return;  // block-uniform
    for (int q = 0; q < 6; ++q) shinv[q][t] = eliminated ? (double)inv[q] : 0.0;
    __syncthreads();
    // retained rows of the block-pair path: R_o = V^-1 W_o (zero for points not eliminated)
    for (uint32_t o = ob + t; o < oe; o += NT) {
        const uint32_t lp = tp.obs_pt[o] - pb;
        T vi[6];
        for (int q = 0; q < 6; ++q) vi[q] = (T)shinv[q][lp];
        T x0[3], x1[3];
        for (int q = 0; q < 3; ++q) {
            x0[q] = lin.jpm[soa_at(SOA_JX + q, o)];
            x1[q] = lin.jpm[soa_at(SOA_JX + 3 + q, o)];
        }
        T* rr = pe.rrows + 27 * (size_t)o;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            T c0 = lin.jpm[soa_at(SOA_JC + q, o)], c1 = lin.jpm[soa_at(SOA_JC + 9 + q, o)];
            T w[3] = {x0[0] * c0 + x1[0] * c1, x0[1] * c0 + x1[1] * c1, x0[2] * c0 + x1[2] * c1};
            T v[3];
            sym3_mul(vi, w, v);
            rr[3 * q] = v[0];
            rr[3 * q + 1] = v[1];
            rr[3 * q + 2] = v[2];
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Deterministic scalar reductions over points.
// ----------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_reduce_points_1(const double* __restrict__ ps, uint32_t np, double* partials) {
    __shared__ double sh[NT];
    uint32_t b0 = (uint32_t)((uint64_t)np * blockIdx.x / RED_BLOCKS);
    uint32_t b1 = (uint32_t)((uint64_t)np * (blockIdx.x + 1) / RED_BLOCKS);
    double cost = 0, sq = 0, pmr = 0, sgin = 0, rc = 1.0, weak = 0, bad = 0, unproj = 0;
    for (uint32_t p = b0 + threadIdx.x; p < b1; p += NT) {
        const double* s = ps + 6 * (size_t)p;
        int flags = (int)s[5];
        if (flags & 1) {
            cost += s[0];
            sq += s[1];
        }
        if ((flags & 3) == 3) {
            pmr += s[2];
            sgin = fmax(sgin, s[3]);
        }
        if ((flags & 1) && !(flags & 2)) sgin = fmax(sgin, s[3]);
        rc = fmin(rc, s[4]);
        if (flags & 4) weak += 1.0;
        if ((flags & 3) != 3) bad += 1.0;
        if (!(flags & 1)) unproj += 1.0;
    }
    double v[8];
    v[0] = tree_sum<NT>(cost, sh);
    v[1] = tree_sum<NT>(sq, sh);
    v[2] = tree_sum<NT>(pmr, sh);
    v[3] = tree_max<NT>(sgin, sh);
    v[4] = tree_min<NT>(rc, sh);
    v[5] = tree_sum<NT>(weak, sh);
    v[6] = tree_sum<NT>(bad, sh);
    v[7] = tree_sum<NT>(unproj, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) partials[8 * blockIdx.x + k] = v[k];
}
// Inlier statistics of the final state (solver.rs:1353-1371): per observation with a valid
// linearisation, its residual norm |r| counts when |r| <= limit; the squared norms of those are
// summed.  Fixed partition into RED_BLOCKS ranges, each summed in a fixed order.
__global__ void __launch_bounds__(NT) k_inliers_1(const double* __restrict__ sq, const uint8_t* __restrict__ valid,
                                                  uint32_t no, double limit, double* partials) {
    __shared__ double sh[NT];
    const uint32_t b0 = (uint32_t)((uint64_t)no * blockIdx.x / RED_BLOCKS);
    const uint32_t b1 = (uint32_t)((uint64_t)no * (blockIdx.x + 1) / RED_BLOCKS);
    double count = 0.0, sum = 0.0;
    for (uint32_t o = b0 + threadIdx.x; o < b1; o += NT) {
        if (!valid[o]) continue;
        const double norm = sqrt(sq[o]);
        if (norm <= limit) {
            count += 1.0;
            sum += norm * norm;
        }
    }
    count = tree_sum<NT>(count, sh);
    sum = tree_sum<NT>(sum, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = count;
        partials[2 * blockIdx.x + 1] = sum;
    }
}
__global__ void __launch_bounds__(NT) k_inliers_2(const double* __restrict__ partials, double* out) {
    __shared__ double sh[NT];
    double count = 0.0, sum = 0.0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += NT) {
        count += partials[2 * b];
        sum += partials[2 * b + 1];
    }
    count = tree_sum<NT>(count, sh);
    sum = tree_sum<NT>(sum, sh);
    if (threadIdx.x == 0) {
        out[0] = count;
        out[1] = sum;
    }
}

__global__ void __launch_bounds__(NT) k_reduce_points_2(const double* __restrict__ partials, AsmScalars* out) {
    __shared__ double sh[NT];
    double a[8] = {0, 0, 0, 0, 1.0, 0, 0, 0};
    for (int b = threadIdx.x; b < RED_BLOCKS; b += NT) {
        const double* q = partials + 8 * b;
        a[0] += q[0];
        a[1] += q[1];
        a[2] += q[2];
        a[3] = fmax(a[3], q[3]);
        a[4] = fmin(a[4], q[4]);
        a[5] += q[5];
        a[6] += q[6];
        a[7] += q[7];
    }
    double v0 = tree_sum<NT>(a[0], sh), v1 = tree_sum<NT>(a[1], sh), v2 = tree_sum<NT>(a[2], sh);
    double v3 = tree_max<NT>(a[3], sh), v4 = tree_min<NT>(a[4], sh), v5 = tree_sum<NT>(a[5], sh);
    double v6 = tree_sum<NT>(a[6], sh), v7 = tree_sum<NT>(a[7], sh);
    if (threadIdx.x == 0) {
        out->cost = v0;
        out->sq = v1;
        out->pmr = v2;
        out->sgin = v3;
        out->min_rcond = v4;
        out->weak = v5;
        out->invalid = v6;
        out->unprojectable = v7;
    }
}

// ----------------------------------------------------------------------------------------------
// k_block_gradient: one workgroup per block; chunks of the block's observation list are summed
// per thread in list order and the chunk partials in chunk order.
// ----------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(NT) k_block_gradient(Topo tp, Lin<T> lin, PointElim<T> pe, double* grad,
                                                       double* full_grad, double* diag) {
    __shared__ double sh[3][NT];
    const uint32_t A = blockIdx.x;
    const int dA = (int)tp.blk_dim[A];
    if (dA == 0) return;
    const int rowbase = A < tp.nc ? 0 : 6;
    const uint32_t lb = tp.blk_obs_off[A], le = tp.blk_obs_off[A + 1];
    const uint32_t len = le - lb;
    const int chunks = NT / dA;
    const int t = threadIdx.x;
    const int r = t % dA, ch = t / dA;
    double full = 0.0, red = 0.0, dg = 0.0;
    if (ch < chunks) {
        uint32_t c0 = lb + (uint32_t)((uint64_t)len * ch / chunks);
        uint32_t c1 = lb + (uint32_t)((uint64_t)len * (ch + 1) / chunks);
        for (uint32_t k = c0; k < c1; ++k) {
            uint32_t i = tp.blk_obs[k];
            double j0 = (double)lin.jpm[soa_at(SOA_JC + rowbase + r, i)];
            double j1 = (double)lin.jpm[soa_at(SOA_JC + 9 + rowbase + r, i)];
            full += j0 * (double)lin.jpm[soa_at(SOA_R + 0, i)] + j1 * (double)lin.jpm[soa_at(SOA_R + 1, i)];
            dg += j0 * j0 + j1 * j1;
            uint32_t p = tp.obs_pt[i];
            if (!tp.pt_fixed[p]) {
                const T* y = pe.y + 3 * (size_t)p;
                double wy0 = 0.0, wy1 = 0.0;
                for (int q = 0; q < 3; ++q) {
                    wy0 += (double)lin.jpm[soa_at(SOA_JX + q, i)] * (double)y[q];
                    wy1 += (double)lin.jpm[soa_at(SOA_JX + 3 + q, i)] * (double)y[q];
                }
                red += j0 * wy0 + j1 * wy1;
            }
        }
    }
    sh[0][t] = full;
    sh[1][t] = red;
    sh[2][t] = dg;
    __syncthreads();
    if (t < dA) {
        double a = 0, bsum = 0, c = 0;
        for (int q = 0; q < chunks; ++q) {
            a += sh[0][q * dA + t];
            bsum += sh[1][q * dA + t];
            c += sh[2][q * dA + t];
        }
        uint32_t off = tp.blk_off[A] + t;
        full_grad[off] = a;
        grad[off] = a - bsum;
        diag[off] = c;
    }
}

// ----------------------------------------------------------------------------------------------
// Camera-side assembly when every group has one camera and no camera sees a point twice (every
// Schur term of a diagonal block is then one observation's), for the PCG form: per camera the
// full and reduced gradients, the damping diagonal and the diagonal blocks of the reduced system
// U - W V^-1 W^T (schur/assembly.rs:893-1039), from which the block-Jacobi preconditioner is
// factored.  Per observation, with u = J~_X y_p and M = J~_X V^-1 J~_X^T (2x2):
//   full gradient J~_c^T r~, reduced part J~_c^T u, diagonal sum of J~_c squares, and
//   J~_c^T (I - M) J~_c (= U_o - W_o V^-1 W_o^T) on the pose-pose and intrinsic-intrinsic blocks.
// The pose columns of J~_c are rebuilt from J~_X, the point and the camera rotation (the compact
// form of the PCG product), so the camera-major copy holds only r~, J~_X and the intrinsic columns.
// Kernel 1: one workgroup of two wavefronts per PCG camera segment; wavefront 0 sums the gradient
// and diagonal terms, wavefront 1 the block terms (27 values each), each over its lanes' positions
// in order and then by the wavefront butterfly.  Kernel 2: per camera, the segments' sums in
// segment order, written out (the blocks into the U values of the diagonal pairs, zero Schur
// values).
// ----------------------------------------------------------------------------------------------
namespace {
__device__ __forceinline__ int tri(int r, int c) { return r * (r + 1) / 2 + c; }  // r >= c
// per segment: 9 full, 9 reduced part, 9 diagonal | 21 pose block, 6 intrinsic block, 18 pose x
// intrinsic cross terms
constexpr int CA_V = CAM_ASM_VALUES;
static_assert(CA_V == 27 + 45, "camera assembly values");

// J~_c (2x9, camera frame) and J~_X (2x3) of an observation rebuilt from its compact state (o),
// its rotated point RX and the camera (R, f; pose dimension pd, baseline flag, gauge gg).
__device__ __forceinline__ void rebuild_jacobians(const Obs& o, const Cam& cam, const double RX[3], uint32_t pd,
                                                  bool baseline, const GaugeDev& gg, double jc[2][9],
                                                  double jx[6]) {
    const double nn[2] = {o.n0, o.n1};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        // g = sqrt(w) J_p row i = h (A [I | -n])_i
        double a0, a1;
        apply_A(o, cam, i == 0 ? 1.0 : 0.0, i == 0 ? 0.0 : 1.0, &a0, &a1);  // column i of A (A symmetric)
        const double g[3] = {o.h * a0, o.h * a1, -o.h * (o.n0 * a0 + o.n1 * a1)};
        const bool posed = pd > 0;
        jc[i][0] = posed ? RX[1] * g[2] - RX[2] * g[1] : 0.0;
        jc[i][1] = posed ? RX[2] * g[0] - RX[0] * g[2] : 0.0;
        jc[i][2] = posed ? RX[0] * g[1] - RX[1] * g[0] : 0.0;
        if (baseline) {
            jc[i][3] = posed ? gg.baseline * (gg.basis[0] * g[0] + gg.basis[1] * g[1] + gg.basis[2] * g[2]) : 0.0;
            jc[i][4] = posed ? gg.baseline * (gg.basis[3] * g[0] + gg.basis[4] * g[1] + gg.basis[5] * g[2]) : 0.0;
            jc[i][5] = 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) jc[i][3 + k] = posed ? gg.scene_scale * g[k] : 0.0;
        }
        const double si = o.sw * cam.f * nn[i];
        jc[i][6] = si * o.d;
        jc[i][7] = si * o.r2;
        jc[i][8] = si * (o.r2 * o.r2);
        double rt[3];
        rotate_t(cam.R, g, rt);
#pragma unroll
        for (int k = 0; k < 3; ++k) jx[3 * i + k] = gg.scene_scale * rt[k];
    }
}
}  // namespace

// Orders one wavefront's LDS accesses around this point (its lanes exchange data through LDS
// without a workgroup barrier; a wavefront's LDS instructions complete in issue order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__global__ void __launch_bounds__(192) k_camera_assembly(Topo tp, Lin<T> lin, PointElim<T> pe, const double* Rg,
                                                         const double* gp, const GaugeDev* gauge,
                                                         const uint32_t* __restrict__ seg_cam,
                                                         const uint32_t* __restrict__ cseg_off, double* part,
                                                         T* __restrict__ crec) {
    const uint32_t k = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63, role = threadIdx.x >> 6;  // role uniform per wavefront
    constexpr uint32_t VN = pcg_vec<T>(), SEG = pcg_seg<T>();
    const uint32_t c = seg_cam[k];
    const uint32_t cb = tp.cm_off[c], ce = tp.cm_off[c + 1];
    const uint32_t wb = cb / VN * VN + (k - cseg_off[c]) * SEG;
    const GaugeDev& gg = *gauge;
    Cam cam;
    for (int i = 0; i < 9; ++i) cam.R[i] = Rg[9 * (size_t)c + i];
    const double* g3 = gp + 3 * (size_t)tp.cam_group[c];
    cam.f = g3[0], cam.k1 = g3[1], cam.k2 = g3[2];
    const uint32_t pd = tp.cam_dim[c];
    const bool baseline = (int)c == gg.camera;
    __shared__ __align__(16) T cs[64 * CREC];  // role 0's compact records of one pass
    double acc[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) acc[q] = 0.0;
    for (uint32_t it = 0; it < SEG / 64; ++it) {
        const uint32_t j = wb + lane + 64 * it;
        const bool live = j >= cb && j < ce;
        const uint32_t p = tp.cm_pt[j];
        const double r0 = (double)lin.jcm[soa_at(SOA_R, j)], r1 = (double)lin.jcm[soa_at(SOA_R + 1, j)];
        const Obs o = make_obs((double)lin.jcm[soa_at(SOA_Q, j)], (double)lin.jcm[soa_at(SOA_Q + 1, j)],
                               (double)lin.jcm[soa_at(SOA_Q + 2, j)], (double)lin.jcm[soa_at(SOA_Q + 3, j)], cam);
        const double RX[3] = {(double)lin.jcm[soa_at(SOA_RX, j)], (double)lin.jcm[soa_at(SOA_RX + 1, j)],
                              (double)lin.jcm[soa_at(SOA_RX + 2, j)]};
        // the point's packed [V^-1, y]: zero for points that are fixed or not eliminated
        const T* prec = pe.prec + PE_REC * (size_t)p;
        double pv[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) pv[q] = (double)prec[q];
        double jc[2][9], a[6];
        rebuild_jacobians(o, cam, RX, pd, baseline, gg, jc, a);
        if (!live) {
#pragma unroll
            for (int q = 0; q < 9; ++q) jc[0][q] = jc[1][q] = 0.0;
        }
        if (role == 0) {
            double u0 = 0.0, u1 = 0.0;
            if (live) {
                const double* y = pv + 6;
                u0 = a[0] * y[0] + a[1] * y[1] + a[2] * y[2];
                u1 = a[3] * y[0] + a[4] * y[1] + a[5] * y[2];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                acc[q] += jc[0][q] * r0 + jc[1][q] * r1;
                acc[9 + q] += jc[0][q] * u0 + jc[1][q] * u1;
                acc[18 + q] += jc[0][q] * jc[0][q] + jc[1][q] * jc[1][q];
            }
            if (crec) {
                // the compact record of this position for the pair assembly, staged in LDS so that
                // the wavefront's 64 consecutive records are written as one contiguous run of
                // 16-byte stores (only live positions)
                T* mine = cs + CREC * lane;
                mine[0] = lin.jcm[soa_at(SOA_Q, j)];
                mine[1] = lin.jcm[soa_at(SOA_Q + 1, j)];
                mine[2] = lin.jcm[soa_at(SOA_Q + 2, j)];
                mine[3] = lin.jcm[soa_at(SOA_Q + 3, j)];
                mine[4] = lin.jcm[soa_at(SOA_RX, j)];
                mine[5] = lin.jcm[soa_at(SOA_RX + 1, j)];
                mine[6] = lin.jcm[soa_at(SOA_RX + 2, j)];
                mine[7] = T(0);
#pragma unroll
                for (int q = 0; q < 6; ++q) mine[8 + q] = (T)pv[q];
                mine[14] = T(0);
                mine[15] = T(0);
                wave_lds_sync();
                constexpr int CH = CREC * (int)sizeof(T) / 16;  // 16-byte chunks per record
                const uint4* src = reinterpret_cast<const uint4*>(cs);
                uint4* dst = reinterpret_cast<uint4*>(crec + (size_t)CREC * (wb + 64 * it));
#pragma unroll
                for (int i = 0; i < CH; ++i) {
                    const uint32_t e = lane + 64 * i, jr = wb + 64 * it + e / CH;
                    if (jr >= cb && jr < ce) dst[e] = src[e];
                }
                wave_lds_sync();
            }
        } else {
            // N = I - M, M = J~_X V^-1 J~_X^T
            double n00 = 1.0, n01 = 0.0, n11 = 1.0;
            if (live) {
                const double* vi = pv;
                double z0[3], z1[3];
                sym3_mul(vi, a, z0);
                sym3_mul(vi, a + 3, z1);
                n00 = 1.0 - (a[0] * z0[0] + a[1] * z0[1] + a[2] * z0[2]);
                n01 = -(a[0] * z1[0] + a[1] * z1[1] + a[2] * z1[2]);
                n11 = 1.0 - (a[3] * z1[0] + a[4] * z1[1] + a[5] * z1[2]);
            }
            double pn[2][9];  // N J~_c
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                pn[0][q] = n00 * jc[0][q] + n01 * jc[1][q];
                pn[1][q] = n01 * jc[0][q] + n11 * jc[1][q];
            }
            if (role == 1) {
#pragma unroll
                for (int x = 0; x < 6; ++x)
#pragma unroll
                    for (int y = 0; y <= x; ++y) acc[tri(x, y)] += jc[0][x] * pn[0][y] + jc[1][x] * pn[1][y];
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y <= x; ++y)
                        acc[21 + tri(x, y)] += jc[0][6 + x] * pn[0][6 + y] + jc[1][6 + x] * pn[1][6 + y];
            } else {  // role 2: pose x intrinsic cross terms
#pragma unroll
                for (int x = 0; x < 6; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) acc[3 * x + y] += jc[0][x] * pn[0][6 + y] + jc[1][x] * pn[1][6 + y];
            }
        }
    }
    const int nv = role == 2 ? 18 : 27;
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (q < nv) acc[q] = wave_sum(acc[q]);
    if (lane == 0) {
        double* out = part + CA_V * (size_t)k + 27 * role;
#pragma unroll
        for (int q = 0; q < 27; ++q)
            if (q < nv) out[q] = acc[q];
    }
}

// One wavefront per camera: lane v adds values v, v + 64, ... over the camera's segments in order.
template <class T>
__global__ void __launch_bounds__(64) k_camera_assembly_out(Topo tp, const uint32_t* __restrict__ cseg_off,
                                                            const double* part,
                                                            const int32_t* __restrict__ diag_pair,
                                                            const uint64_t* __restrict__ value_off, double* grad,
                                                            double* full_grad, double* diag, double* u_vals,
                                                            double* s_vals, SchurRows<T> S) {
    __shared__ double tot[CA_V];
    const uint32_t c = blockIdx.x, v = threadIdx.x;
    for (uint32_t u = v; u < (uint32_t)CA_V; u += 64) {
        double s = 0.0;
        const uint32_t g0 = cseg_off[c], g1 = cseg_off[c + 1];
        for (uint32_t g = g0; g < g1; g += 4) {  // four loads in flight, added in segment order
            double x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) x[i] = g + i < g1 ? part[CA_V * (size_t)(g + i) + u] : 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) s += x[i];
        }
        tot[u] = s;
    }
    __syncthreads();
    const uint32_t pd = tp.cam_dim[c], po = tp.cam_off[c];
    const uint32_t g = tp.cam_group[c], gd = tp.grp_dim[g], go = tp.grp_off[g];
    if (v < pd) {
        full_grad[po + v] = tot[v];
        grad[po + v] = tot[v] - tot[9 + v];
        diag[po + v] = tot[18 + v];
    }
    if (v < gd) {
        full_grad[go + v] = tot[6 + v];
        grad[go + v] = tot[6 + v] - tot[15 + v];
        diag[go + v] = tot[24 + v];
    }
    if (pd > 0 && diag_pair[c] >= 0) {
        double* U = u_vals + value_off[diag_pair[c]];
        double* Sv = s_vals + value_off[diag_pair[c]];
        for (uint32_t e = v; e < pd * pd; e += 64) {
            const uint32_t x = e / pd, y = e % pd;
            U[e] = tot[27 + (x >= y ? tri(x, y) : tri(y, x))];
            Sv[e] = 0.0;
        }
    }
    if (gd > 0 && diag_pair[tp.nc + g] >= 0) {
        double* U = u_vals + value_off[diag_pair[tp.nc + g]];
        double* Sv = s_vals + value_off[diag_pair[tp.nc + g]];
        for (uint32_t e = v; e < gd * gd; e += 64) {
            const uint32_t x = e / gd, y = e % gd;
            U[e] = tot[48 + (x >= y ? tri(x, y) : tri(y, x))];
            Sv[e] = 0.0;
        }
    }
    if (S.vals) {  // the camera's diagonal block of S in slot 0, 9 x 9 (zero beyond the free dimensions)
        const uint32_t nn = S.nbr_off[c + 1] - S.nbr_off[c];
        T* row = S.vals + S.row_base[c];
        for (uint32_t e = v; e < 81; e += 64) {
            const uint32_t x = e / 9, y = e % 9;
            const bool in = (x < 6 ? x < pd : x - 6 < gd) && (y < 6 ? y < pd : y - 6 < gd);
            double val = 0.0;
            if (in) {
                if (x < 6 && y < 6) val = tot[27 + (x >= y ? tri(x, y) : tri(y, x))];
                else if (x >= 6 && y >= 6) val = tot[48 + (x >= y ? tri(x - 6, y - 6) : tri(y - 6, x - 6))];
                else val = x < 6 ? tot[54 + 3 * x + (y - 6)] : tot[54 + 3 * y + (x - 6)];
            }
            row[(size_t)x * schur_row_len(nn) + y] = (T)val;
        }
    }
}

template <class T>
__global__ void __launch_bounds__(NT) k_pairs(Topo tp, Lin<T> lin, PointElim<T> pe, const uint32_t* __restrict__ pair_a,
                                              const uint32_t* __restrict__ pair_b, const uint32_t* __restrict__ term_off,
                                              const uint32_t* __restrict__ term_i, const uint32_t* __restrict__ term_j,
                                              const uint64_t* __restrict__ value_off, double* u_vals, double* s_vals) {
    __shared__ double shu[NT];
    __shared__ double shs[NT];
    const uint32_t k = blockIdx.x;
    const uint32_t A = pair_a[k], B = pair_b[k];
    const int dA = (int)tp.blk_dim[A], dB = (int)tp.blk_dim[B];
    const int nout = dA * dB;
    const int ra = A < tp.nc ? 0 : 6, rb = B < tp.nc ? 0 : 6;
    const int chunks = NT / nout;
    const int t = threadIdx.x;
    const int o = t % nout, ch = t / nout;
    const int r = o / dB, c = o % dB;
    double accu = 0.0, accs = 0.0;
    const uint32_t tb = term_off[k], te = term_off[k + 1];
    const uint32_t len = te - tb;
    if (ch < chunks) {
        uint32_t c0 = tb + (uint32_t)((uint64_t)len * ch / chunks);
        uint32_t c1 = tb + (uint32_t)((uint64_t)len * (ch + 1) / chunks);
        for (uint32_t q = c0; q < c1; ++q) {
            uint32_t i = term_i[q], jf = term_j[q];
            uint32_t j = jf & ((1u << 30) - 1);
            if (jf & (1u << 30)) {
                const T* jc = lin.jc + 18 * (size_t)i;
                accu += (double)jc[ra + r] * (double)jc[rb + c] + (double)jc[9 + ra + r] * (double)jc[9 + rb + c];
            }
            if (jf & (1u << 31)) {
                uint32_t p = tp.obs_pt[i];
                const T* vd = pe.vdamp + 6 * (size_t)p;
                const T* Ri = pe.rrows + 27 * (size_t)i + 3 * (ra + r);
                const T* Rj = pe.rrows + 27 * (size_t)j + 3 * (rb + c);
                T w[3];
                T rjv[3] = {Rj[0], Rj[1], Rj[2]};
                sym3_mul(vd, rjv, w);
                accs += (double)Ri[0] * (double)w[0] + (double)Ri[1] * (double)w[1] + (double)Ri[2] * (double)w[2];
            }
        }
    }
    shu[t] = accu;
    shs[t] = accs;
    __syncthreads();
    if (t < nout) {
        double su = 0.0, ss = 0.0;
        for (int q = 0; q < chunks; ++q) {
            su += shu[q * nout + t];
            ss += shs[q * nout + t];
        }
        uint64_t base = value_off[k];
        u_vals[base + t] = su;
        s_vals[base + t] = ss;
    }
}

__global__ void __launch_bounds__(NT) k_scatter_dense(Topo tp, uint32_t npairs, const uint32_t* __restrict__ pair_a,
                                                      const uint32_t* __restrict__ pair_b,
                                                      const uint64_t* __restrict__ value_off,
                                                      const double* __restrict__ u_vals,
                                                      const double* __restrict__ s_vals, double* S, uint32_t ld) {
    const uint32_t k = blockIdx.x;
    const uint32_t A = pair_a[k], B = pair_b[k];
    const int dA = (int)tp.blk_dim[A], dB = (int)tp.blk_dim[B];
    const int t = threadIdx.x;
    if (t >= dA * dB) return;
    const int r = t / dB, c = t % dB;
    const uint64_t at = value_off[k] + t;
    const double v = u_vals[at] - s_vals[at];
    const size_t row = tp.blk_off[A] + r, col = tp.blk_off[B] + c;
    S[row * ld + col] = v;
    S[col * ld + row] = v;
}

// The explicit system's camera rows into the dense reduced matrix (the direct step of a simple
// topology): workgroup per camera row c (slot 0 the camera's diagonal 9 x 9 block, then the pairs
// (c, j > c)).  Block coordinates 0-5 are the pose (its first cam_dim free), 6-8 the camera's
// intrinsic group (its first grp_dim free).  Workgroup (c, y) takes slots [y SCATTER_SLOTS,
// (y + 1) SCATTER_SLOTS) of the row, staged in LDS; each value goes to (row, col) with the column
// index fastest and, for a pair, to (col, row) with the row index fastest, so that both writes run
// along rows of S.
__device__ __forceinline__ int block_coordinate(const Topo& tp, uint32_t cam, uint32_t a) {
    if (a < 6) return a < tp.cam_dim[cam] ? (int)(tp.cam_off[cam] + a) : -1;
    const uint32_t g = tp.cam_group[cam];
    return a - 6 < tp.grp_dim[g] ? (int)(tp.grp_off[g] + a - 6) : -1;
}
constexpr uint32_t SCATTER_SLOTS = 8;  // slots per workgroup
template <class T>
__global__ void __launch_bounds__(NT) k_scatter_rows(Topo tp, SchurRows<T> S, double* __restrict__ Sd,
                                                     uint32_t ld) {
    __shared__ double blk[SCATTER_SLOTS][81];
    __shared__ int col[SCATTER_SLOTS][9];
    __shared__ int own[9];
    const uint32_t c = blockIdx.x;
    const uint32_t n0 = S.nbr_off[c], nn = S.nbr_off[c + 1] - n0, rl = schur_row_len(nn);
    const T* __restrict__ row = S.vals + S.row_base[c];
    const uint32_t s0 = blockIdx.y * SCATTER_SLOTS;
    if (s0 >= nn) return;
    if (threadIdx.x < 9) own[threadIdx.x] = block_coordinate(tp, c, threadIdx.x);
    {
        const uint32_t ns = nn - s0 < SCATTER_SLOTS ? nn - s0 : SCATTER_SLOTS;
        for (uint32_t e = threadIdx.x; e < 81 * ns; e += NT) {  // coalesced along the row
            const uint32_t a = e / (9 * ns), rest = e - a * 9 * ns, sl = rest / 9, b = rest - 9 * sl;
            blk[sl][9 * a + b] = (double)row[(size_t)a * rl + 9 * s0 + rest];
        }
        for (uint32_t e = threadIdx.x; e < 9 * ns; e += NT)
            col[e / 9][e % 9] = block_coordinate(tp, S.nbr[n0 + s0 + e / 9], e % 9);
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < 81 * ns; e += NT) {  // (row, col): b fastest
            const uint32_t sl = e / 81, a = (e / 9) % 9, b = e % 9;
            const int ia = own[a], ib = col[sl][b];
            if (ia >= 0 && ib >= 0) Sd[(size_t)ia * ld + ib] = blk[sl][9 * a + b];
        }
        for (uint32_t e = threadIdx.x; e < 81 * ns; e += NT) {  // (col, row) of a pair: a fastest
            const uint32_t sl = e / 81, b = (e / 9) % 9, a = e % 9;
            if (s0 + sl == 0) continue;  // the diagonal block is stored whole
            const int ia = own[a], ib = col[sl][b];
            if (ia >= 0 && ib >= 0) Sd[(size_t)ib * ld + ia] = blk[sl][9 * a + b];
        }
    }
}

// ----------------------------------------------------------------------------------------------
// k_finalize (single workgroup): damping-diagonal clamp per block (schur/assembly.rs:729-766),
// lambda * diag on the reduced diagonal, scaled gradient norm (:1098-1112), finiteness of g.
// ----------------------------------------------------------------------------------------------
// Lane per block (grid-stride over a fixed grid): the block's unclamped diagonal enters min/max, is
// clamped to 64 eps of the block's largest entry (unit for an empty block scale), lambda diag
// goes onto S or the diagonal pair's U values, and the scaled gradient max and the non-finite
// gradient count are formed on the clamped diagonal.  Per workgroup: [min, max, clamped, sg,
// nonfinite]; min/max are order-free and the counts exact, so the result is deterministic.
constexpr int FIN_WG = 256;
constexpr int FIN_GRID_MAX = 512;
__global__ void __launch_bounds__(FIN_WG) k_finalize_blocks(Topo tp, double lambda, double* diag,
                                                            const double* full_grad, const double* grad, double* S,
                                                            uint32_t ld, const int32_t* diag_pair,
                                                            const uint64_t* value_off, double* u_vals, double* part) {
    __shared__ double sh[FIN_WG];
    double mn = INFINITY, mx = 0.0, clamped = 0.0, sg = 0.0, nonfinite = 0.0;
    for (uint32_t A = blockIdx.x * FIN_WG + threadIdx.x; A < tp.nblocks; A += gridDim.x * FIN_WG) {
        const uint32_t n = tp.blk_dim[A], off = tp.blk_off[A];
        if (n == 0) continue;
        double d[9], scale = 0.0;
        for (uint32_t k = 0; k < n; ++k) {
            d[k] = diag[off + k];
            mn = fmin(mn, d[k]);
            mx = fmax(mx, d[k]);
            scale = fmax(scale, d[k]);
        }
        const double floor = scale > 0.0 ? 64.0 * DBL_EPSILON * scale : 1.0;
        for (uint32_t k = 0; k < n; ++k)
            if (d[k] < floor) {
                d[k] = floor;
                diag[off + k] = floor;
                clamped += 1.0;
            }
        if (S) {
            for (uint32_t k = 0; k < n; ++k) S[(size_t)(off + k) * ld + off + k] += lambda * d[k];
        } else if (diag_pair) {
            const int32_t pr = diag_pair[A];
            if (pr >= 0)
                for (uint32_t k = 0; k < n; ++k) u_vals[value_off[pr] + k * n + k] += lambda * d[k];
        }
        for (uint32_t k = 0; k < n; ++k) {
            const double gv = full_grad[off + k];
            const double v = (isfinite(d[k]) && d[k] > 0.0) ? fabs(gv) / sqrt(d[k]) : (gv == 0.0 ? 0.0 : INFINITY);
            sg = fmax(sg, v);
            if (!isfinite(grad[off + k])) nonfinite += 1.0;
        }
    }
    mn = tree_min<FIN_WG>(mn, sh);
    mx = tree_max<FIN_WG>(mx, sh);
    clamped = tree_sum<FIN_WG>(clamped, sh);
    sg = tree_max<FIN_WG>(sg, sh);
    nonfinite = tree_sum<FIN_WG>(nonfinite, sh);
    if (threadIdx.x == 0) {
        double* o = part + 5 * (size_t)blockIdx.x;
        o[0] = mn;
        o[1] = mx;
        o[2] = clamped;
        o[3] = sg;
        o[4] = nonfinite;
    }
}
__global__ void __launch_bounds__(FIN_WG) k_finalize_reduce(Topo tp, uint32_t nparts, const double* part,
                                                            AsmScalars* sc) {
    __shared__ double sh[FIN_WG];
    double mn = INFINITY, mx = 0.0, clamped = 0.0, sg = 0.0, nonfinite = 0.0;
    for (uint32_t i = threadIdx.x; i < nparts; i += FIN_WG) {
        const double* o = part + 5 * (size_t)i;
        mn = fmin(mn, o[0]);
        mx = fmax(mx, o[1]);
        clamped += o[2];
        sg = fmax(sg, o[3]);
        nonfinite += o[4];
    }
    mn = tree_min<FIN_WG>(mn, sh);
    mx = tree_max<FIN_WG>(mx, sh);
    clamped = tree_sum<FIN_WG>(clamped, sh);
    sg = tree_max<FIN_WG>(sg, sh);
    nonfinite = tree_sum<FIN_WG>(nonfinite, sh);
    if (threadIdx.x == 0) {
        sc->sgin_prepared = sg;
        sc->min_diag = tp.D == 0 ? 0.0 : mn;
        sc->max_diag = mx;
        sc->clamped = clamped;
        sc->nonfinite = nonfinite;
    }
}

// ----------------------------------------------------------------------------------------------
// Retraction (kernel/multi_view/optimizer.rs:386-437, self_calibration/optimizer.rs:565-586)
// ----------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(NT) k_retract(Topo tp, const double* R, const double* t, const double* gp,
                                                const double* step, const GaugeDev* gauge, double* Rc, double* tc,
                                                double* gpc) {
    uint32_t i = blockIdx.x * NT + threadIdx.x;
    if (i < tp.nc) {
        const uint32_t pd = tp.cam_dim[i], off = tp.cam_off[i];
        if (pd == 0) {
            for (int k = 0; k < 9; ++k) Rc[9 * i + k] = R[9 * i + k];
            for (int k = 0; k < 3; ++k) tc[3 * i + k] = t[3 * i + k];
        } else {
            double up[9];
            double w[3] = {step[off], step[off + 1], step[off + 2]};
            exp_so3(w, up);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double s = 0.0;
                    for (int c = 0; c < 3; ++c) s += up[3 * a + c] * R[9 * i + 3 * c + b];
                    Rc[9 * i + 3 * a + b] = s;
                }
            const GaugeDev* gw = gauge + (tp.cam_win ? tp.cam_win[i] : 0);
            if ((int)i == gw->camera) {
                double dir[3];
                retract_s2(gw->direction, gw->basis, step[off + 3], step[off + 4], dir);
                for (int k = 0; k < 3; ++k) tc[3 * i + k] = gw->baseline * dir[k];
            } else {
                for (int k = 0; k < 3; ++k) tc[3 * i + k] = t[3 * i + k] + gw->scene_scale * step[off + 3 + k];
            }
        }
    }
    if (i < tp.ng) {
        const uint32_t d = tp.grp_dim[i], off = tp.grp_off[i];
        double delta[3] = {0.0, 0.0, 0.0};
        for (uint32_t k = 0; k < d; ++k) delta[k] = step[off + k];
        if (d == 0) {
            for (int k = 0; k < 3; ++k) gpc[3 * i + k] = gp[3 * i + k];
        } else {
            gpc[3 * i] = gp[3 * i] * exp(delta[0]);
            gpc[3 * i + 1] = gp[3 * i + 1] + delta[1];
            gpc[3 * i + 2] = gp[3 * i + 2] + delta[2];
        }
    }
}

// ----------------------------------------------------------------------------------------------
// (a') k_candidate: one lane per point.  cscal per point: cost, step^2 (moved), damping norm
// (moved), rejected flag.
// ----------------------------------------------------------------------------------------------
namespace {
enum { ADMIT_PROJECTION = 0, ADMIT_COLLAPSE = 1, ADMIT_GEOMETRY = 2 };
__device__ bool admitted_cost(const Topo& tp, uint32_t b, uint32_t e, const double* R, const double* t,
                              const double* gpc, const double X[3], const double* uv, double scene,
                              int admission, double* out) {
    TrackGeometry<double> geo(scene);
    double total = 0.0;
    for (uint32_t o = b; o < e; ++o) {
        const uint32_t c = tp.obs_cam[o], g = tp.cam_group[c];
        const double* Rc = R + 9 * (size_t)c;
        double rot[3];
        rotate(Rc, X, rot);
        double cp[3] = {rot[0] + t[3 * c], rot[1] + t[3 * c + 1], rot[2] + t[3 * c + 2]};
        const double f = gpc[3 * g], k1 = gpc[3 * g + 1], k2 = gpc[3 * g + 2];
        const double turn = tp.gturn[g];  // first radial turn of the group (+inf: none)
        const bool has_turn = turn < INFINITY;
        double res[2];
        if (!radial_residual<double>(f, k1, k2, turn, has_turn, cp, uv[2 * (size_t)o], uv[2 * (size_t)o + 1], res))
            return false;
        double w[3];
        rotate_t(Rc, cp, w);
        geo.push(w);
        total += loss_cost<double>(tp.loss, tp.loss_scale, sqrt(res[0] * res[0] + res[1] * res[1]));
    }
    bool ok = admission == ADMIT_PROJECTION ? true : (admission == ADMIT_COLLAPSE ? !geo.collapsed() : !geo.degenerate());
    if (ok) *out = total;
    return ok;
}
}  // namespace

// One workgroup per point chunk:
//   1. lanes (one per observation): e_o = J~_X,o^T J~_c,o dc from the point-major copy;
//   2. lanes (one per point): dp = -(V^-1 g_p + V^-1 sum_o e_o) (the cached back-substitution of
//      candidate.rs:358-383 through the Jacobians), the stepped point X + s dp;
//   3. lanes (one per observation): residual, robust cost and camera-to-point offset at the
//      stepped point and the candidate cameras;
//   4. lanes (one per point): admitted_point_cost (schur/point.rs:706-728) in observation order and
//      the candidate rule (point_geometry.rs:153-168); a point whose stepped position is not
//      admitted is decided by the sequential hold/reject rules (admitted_cost).
// A chunk of one point with more than NT observations runs the sequential rules on one lane.
template <class T>
__global__ void __launch_bounds__(NT) k_candidate(Topo tp, const double* R, const double* t, const double* Rc,
                                                  const double* tc, const double* gpc, const double* X,
                                                  const double* uv, Lin<T> lin, PointElim<T> pe,
                                                  const double* srec, const GaugeDev* gauge, double* Xc,
                                                  double* cscal) {
    __shared__ double she[3][NT];
    __shared__ double shx[3][NT];
    __shared__ double shw[3][NT];
    __shared__ double shc[NT];
    __shared__ int shv[NT];
    const uint32_t k = blockIdx.x, lane = threadIdx.x;
    const uint32_t ob = tp.chunk_obs[k], oe = tp.chunk_obs[k + 1];
    const uint32_t pb = tp.chunk_pt[k], pend = tp.chunk_pt[k + 1];
    const bool big = oe - ob > NT;
    // 1. per-observation back-substitution terms e_o = J~_X,o^T J~_c,o step, with the cameras'
    //    world-frame step records
    double e[3] = {0.0, 0.0, 0.0};
    for (uint32_t o = ob + lane; o < oe; o += NT) {
        const uint32_t q = tp.obs_pt[o], c = tp.obs_cam[o];
        double jx[6], w[WREC], eo[3];
#pragma unroll
        for (int i = 0; i < 6; ++i) jx[i] = (double)lin.jpm[soa_at(SOA_JX + i, o)];
        const double2* wr = reinterpret_cast<const double2*>(srec + WREC_PAD * (size_t)c);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double2 x = wr[i];
            w[2 * i] = x.x;
            w[2 * i + 1] = x.y;
        }
        w[8] = wr[4].x;
        const double Xo[3] = {X[3 * (size_t)q], X[3 * (size_t)q + 1], X[3 * (size_t)q + 2]};
        world_product(jx, (double)lin.jpm[soa_at(SOA_Q, o)], (double)lin.jpm[soa_at(SOA_Q + 1, o)],
                      (double)lin.jpm[soa_at(SOA_Q + 3, o)], Xo, w, eo);
#pragma unroll
        for (int i = 0; i < 3; ++i) e[i] += eo[i];
    }
    for (int q = 0; q < 3; ++q) she[q][lane] = e[q];
    __syncthreads();
    if (big) {
        for (int s = NT / 2; s > 0; s >>= 1) {
            if ((int)lane < s)
                for (int q = 0; q < 3; ++q) she[q][lane] += she[q][lane + s];
            __syncthreads();
        }
    }
    // 2. stepped point per point
    const bool have_point = lane < pend - pb;
    const uint32_t p = pb + lane;
    const double scene = have_point ? gauge[tp.pt_win ? tp.pt_win[p] : 0].scene_scale : 0.0;
    double prev[3] = {0.0, 0.0, 0.0}, stepped[3] = {0.0, 0.0, 0.0}, dp[3] = {0.0, 0.0, 0.0};
    bool fixed = true;
    if (have_point) {
        for (int q = 0; q < 3; ++q) prev[q] = X[3 * (size_t)p + q];
        fixed = tp.pt_fixed[p] != 0;
        if (!fixed) {
            double s3[3] = {0.0, 0.0, 0.0};
            if (big) {
                for (int q = 0; q < 3; ++q) s3[q] = she[q][0];
            } else {
                for (uint32_t o = tp.pt_off[p]; o < tp.pt_off[p + 1]; ++o)
                    for (int q = 0; q < 3; ++q) s3[q] += she[q][o - ob];
            }
            const T* vi = pe.vinv + 6 * (size_t)p;
            const double m[6] = {(double)vi[0], (double)vi[1], (double)vi[2], (double)vi[3], (double)vi[4], (double)vi[5]};
            double ve[3];
            sym3_mul(m, s3, ve);
            for (int q = 0; q < 3; ++q) dp[q] = -((double)pe.y[3 * (size_t)p + q] + ve[q]);
            for (int q = 0; q < 3; ++q) stepped[q] = prev[q] + scene * dp[q];
        } else {
            for (int q = 0; q < 3; ++q) stepped[q] = prev[q];
        }
    }
    double* cs = have_point ? cscal + 4 * (size_t)p : nullptr;
    if (big) {  // one point, sequential rules on lane 0
        if (!have_point) return;
        const uint32_t b = tp.pt_off[p], en = tp.pt_off[p + 1];
        double c;
        bool ok = admitted_cost(tp, b, en, Rc, tc, gpc, stepped, uv, fixed ? 0.0 : scene,
                                fixed ? ADMIT_PROJECTION : ADMIT_GEOMETRY, &c);
        if (ok) {
            for (int q = 0; q < 3; ++q) Xc[3 * (size_t)p + q] = stepped[q];
            const T* dd = pe.ddiag + 3 * (size_t)p;
            cs[0] = c;
            cs[1] = fixed ? 0.0 : dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
            cs[2] = fixed ? 0.0
                          : (double)dd[0] * dp[0] * dp[0] + (double)dd[1] * dp[1] * dp[1] + (double)dd[2] * dp[2] * dp[2];
            cs[3] = 0.0;
            return;
        }
        bool held = false;
        if (!fixed) {
            held = admitted_cost(tp, b, en, Rc, tc, gpc, prev, uv, scene, ADMIT_COLLAPSE, &c);
            if (!held) {
                double dummy;
                bool cur_ok = admitted_cost(tp, b, en, R, t, gpc, prev, uv, scene, ADMIT_COLLAPSE, &dummy);
                held = !cur_ok && admitted_cost(tp, b, en, Rc, tc, gpc, prev, uv, scene, ADMIT_PROJECTION, &c);
            }
        }
        for (int q = 0; q < 3; ++q) Xc[3 * (size_t)p + q] = prev[q];
        cs[0] = held ? c : 0.0;
        cs[1] = 0.0;
        cs[2] = 0.0;
        cs[3] = held ? 0.0 : 1.0;
        return;
    }
    for (int q = 0; q < 3; ++q) shx[q][lane] = stepped[q];
    __syncthreads();
    // 3. residual and offset per observation at the stepped point and the candidate cameras
    if (ob + lane < oe) {
        const uint32_t o = ob + lane;
        const uint32_t lp = tp.obs_pt[o] - pb;
        const double Xs[3] = {shx[0][lp], shx[1][lp], shx[2][lp]};
        const uint32_t c = tp.obs_cam[o], g = tp.cam_group[c];
        const double* Rm = Rc + 9 * (size_t)c;
        double rot[3];
        rotate(Rm, Xs, rot);
        double cp[3] = {rot[0] + tc[3 * c], rot[1] + tc[3 * c + 1], rot[2] + tc[3 * c + 2]};
        const double f = gpc[3 * g], k1 = gpc[3 * g + 1], k2 = gpc[3 * g + 2];
        const double turn = tp.gturn[g];  // first radial turn of the group (+inf: none)
        const bool has_turn = turn < INFINITY;
        double res[2];
        const bool ok = radial_residual<double>(f, k1, k2, turn, has_turn, cp, uv[2 * (size_t)o], uv[2 * (size_t)o + 1], res);
        double w[3];
        rotate_t(Rm, cp, w);
        shv[lane] = ok ? 1 : 0;
        shc[lane] = ok ? loss_cost<double>(tp.loss, tp.loss_scale, sqrt(res[0] * res[0] + res[1] * res[1])) : 0.0;
        for (int q = 0; q < 3; ++q) shw[q][lane] = w[q];
    }
    __syncthreads();
    if (!have_point) return;
    // 4. admission of the stepped point, in observation order
    const uint32_t b = tp.pt_off[p], en = tp.pt_off[p + 1];
    TrackGeometry<double> geo(fixed ? 0.0 : scene);
    double total = 0.0;
    bool all_ok = true;
    for (uint32_t o = b; o < en; ++o) {
        const uint32_t j = o - ob;
        if (!shv[j]) {
            all_ok = false;
            break;
        }
        const double w[3] = {shw[0][j], shw[1][j], shw[2][j]};
        geo.push(w);
        total += shc[j];
    }
    const bool admitted = all_ok && (fixed || !geo.degenerate());
    if (admitted) {
        for (int q = 0; q < 3; ++q) Xc[3 * (size_t)p + q] = stepped[q];
        cs[0] = total;
        if (fixed) {
            cs[1] = 0.0;
            cs[2] = 0.0;
        } else {
            const T* dd = pe.ddiag + 3 * (size_t)p;
            cs[1] = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
            cs[2] = (double)dd[0] * dp[0] * dp[0] + (double)dd[1] * dp[1] * dp[1] + (double)dd[2] * dp[2] * dp[2];
        }
        cs[3] = 0.0;
        return;
    }
    bool held = false;
    double c = 0.0;
    if (!fixed) {
        held = admitted_cost(tp, b, en, Rc, tc, gpc, prev, uv, scene, ADMIT_COLLAPSE, &c);
        if (!held) {
            double dummy;
            bool cur_ok = admitted_cost(tp, b, en, R, t, gpc, prev, uv, scene, ADMIT_COLLAPSE, &dummy);
            held = !cur_ok && admitted_cost(tp, b, en, Rc, tc, gpc, prev, uv, scene, ADMIT_PROJECTION, &c);
        }
    }
    for (int q = 0; q < 3; ++q) Xc[3 * (size_t)p + q] = prev[q];
    cs[0] = held ? c : 0.0;
    cs[1] = 0.0;
    cs[2] = 0.0;
    cs[3] = held ? 0.0 : 1.0;
}

__global__ void __launch_bounds__(NT) k_reduce_cand_1(const double* __restrict__ cs, uint32_t np, double* partials) {
    __shared__ double sh[NT];
    uint32_t b0 = (uint32_t)((uint64_t)np * blockIdx.x / RED_BLOCKS);
    uint32_t b1 = (uint32_t)((uint64_t)np * (blockIdx.x + 1) / RED_BLOCKS);
    double a[4] = {0, 0, 0, 0};
    for (uint32_t p = b0 + threadIdx.x; p < b1; p += NT)
        for (int k = 0; k < 4; ++k) a[k] += cs[4 * (size_t)p + k];
    for (int k = 0; k < 4; ++k) {
        double v = tree_sum<NT>(a[k], sh);
        if (threadIdx.x == 0) partials[4 * blockIdx.x + k] = v;
    }
}
__global__ void __launch_bounds__(NT) k_reduce_cand_2(const double* __restrict__ partials, CandScalars* out) {
    __shared__ double sh[NT];
    double a[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < RED_BLOCKS; b += NT)
        for (int k = 0; k < 4; ++k) a[k] += partials[4 * b + k];
    double v[4];
    for (int k = 0; k < 4; ++k) v[k] = tree_sum<NT>(a[k], sh);
    if (threadIdx.x == 0) {
        out->cost = v[0];
        out->step2 = v[1];
        out->damping_norm = v[2];
        out->rejected = v[3];
    }
}

__global__ void __launch_bounds__(1024) k_step_dots(uint32_t D, const double* g, const double* diag, const double* x,
                                                    const double* product, double* dots) {
    __shared__ double sh[1024];
    double a = 0, b = 0, c = 0, d = 0;
    for (uint32_t i = threadIdx.x; i < D; i += 1024) {
        a += g[i] * x[i];
        b += diag[i] * x[i] * x[i];
        c += x[i] * x[i];
        if (product) d += x[i] * product[i];
    }
    a = tree_sum<1024>(a, sh);
    b = tree_sum<1024>(b, sh);
    c = tree_sum<1024>(c, sh);
    d = tree_sum<1024>(d, sh);
    if (threadIdx.x == 0) {
        dots[0] = a;
        dots[1] = b;
        dots[2] = c;
        dots[3] = d;
    }
}

// limit_camera_step (kernel/multi_view/optimizer.rs:677-702) of the calibrated path: the largest
// norm of a camera's rotation block (its first min(dim, 3) values) or translation block (the rest)
// is held to `maximum` by scaling the whole step, and the PCG product S x with it.  One workgroup:
// each thread takes cameras t, t + 1024, ...; each norm is the reference's sequential sum of
// squares; the maximum is exact in any order.  The scale factor (1 when not limited) goes to
// *alpha.
__global__ void __launch_bounds__(1024) k_limit_step(Topo tp, double maximum, double* step, double* product,
                                                     double* alpha) {
    __shared__ double sh[1024];
    double largest = 0.0;
    for (uint32_t c = threadIdx.x; c < tp.nc; c += 1024) {
        const uint32_t dim = tp.cam_dim[c], off = tp.cam_off[c];
        if (dim == 0) continue;
        const uint32_t first = dim < 3 ? dim : 3;
        double s1 = 0.0, s2 = 0.0;
        for (uint32_t k = 0; k < first; ++k) s1 += step[off + k] * step[off + k];
        for (uint32_t k = first; k < dim; ++k) s2 += step[off + k] * step[off + k];
        largest = fmax(largest, sqrt(s1));
        largest = fmax(largest, sqrt(s2));
    }
    largest = tree_max<1024>(largest, sh);
    const bool limit = largest > maximum;
    const double scale = limit ? maximum / largest : 1.0;
    if (threadIdx.x == 0) *alpha = scale;
    if (!limit) return;
    for (uint32_t i = threadIdx.x; i < tp.D; i += 1024) {
        step[i] *= scale;
        if (product) product[i] *= scale;
    }
}

// ----------------------------------------------------------------------------------------------
// (c) PCG.  Preconditioner blocks: Jacobi-scaled Cholesky of U_AA(damped) - Schur_AA
// (retained/block.rs:171-250); stored per block as 45 packed lower values + 9 scales.
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ int pidx(int r, int c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }

// Factors are stored structure-of-arrays: value i of block A at factors[i * nblocks + A], so
// that lane-per-block passes read them coalesced.
__global__ void __launch_bounds__(NT) k_precond_factor(Topo tp, const int32_t* diag_pair, const uint64_t* value_off,
                                                       const double* u_vals, const double* s_vals, double* factors,
                                                       int* fail) {
    uint32_t A = blockIdx.x * NT + threadIdx.x;
    if (A >= tp.nblocks) return;
    const int n = (int)tp.blk_dim[A];
    if (n == 0) return;
    int32_t pr = diag_pair[A];
    if (pr < 0) {
        *fail = 1;
        return;
    }
    const double* U = u_vals + value_off[pr];
    const double* Sv = s_vals + value_off[pr];
    double m[45];
    for (int r = 0; r < n; ++r)
        for (int c = 0; c <= r; ++c) m[pidx(r, c)] = U[r * n + c] - Sv[r * n + c];
    double ds = 0.0;
    for (int i = 0; i < n; ++i) ds = fmax(ds, m[pidx(i, i)]);
    if (!(isfinite(ds) && ds > 0.0)) {
        *fail = 1;
        return;
    }
    double floor = 64.0 * DBL_EPSILON * ds;
    double L[45], sc[9];
    for (int i = 0; i < 45; ++i) L[i] = 0.0;
    for (int i = 0; i < n; ++i) sc[i] = sqrt(fmax(m[pidx(i, i)], floor));
    for (int r = 0; r < n; ++r)
        for (int c = 0; c <= r; ++c) {
            double v = m[pidx(r, c)] / (sc[r] * sc[c]);
            for (int k = 0; k < c; ++k) v -= L[pidx(r, k)] * L[pidx(c, k)];
            if (r == c) {
                if (!(isfinite(v) && v > 64.0 * DBL_EPSILON)) {
                    *fail = 1;
                    return;
                }
                L[pidx(r, c)] = sqrt(v);
            } else {
                L[pidx(r, c)] = v / L[pidx(c, c)];
            }
        }
    const size_t st = tp.nblocks;
    for (int i = 0; i < n * (n + 1) / 2; ++i) factors[i * st + A] = L[i];
    for (int i = 0; i < n; ++i) factors[(45 + i) * st + A] = sc[i];
}

namespace {
// z = M^-1 r for block A of dimension n <= 9 (Jacobi-scaled Cholesky factor, forward then
// backward substitution in the order of retained/block.rs:171-250).  The factor is read into
// registers first so that its loads are issued together.
__device__ __forceinline__ void precond_load(const double* __restrict__ factors, size_t nblocks, uint32_t A, int n,
                                             double L[45], double sc[9]) {
    const int nl = n * (n + 1) / 2;
#pragma unroll
    for (int i = 0; i < 45; ++i) L[i] = i < nl ? factors[i * nblocks + A] : 1.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) sc[i] = i < n ? factors[(45 + i) * nblocks + A] : 1.0;
}
__device__ __forceinline__ void precond_solve(const double L[45], const double sc[9], int n, const double right[9],
                                              double sol[9]) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        double v = right[r] / sc[r];
#pragma unroll
        for (int c = 0; c < r; ++c) v -= L[pidx(r, c)] * sol[c];
        sol[r] = r < n ? v / L[pidx(r, r)] : 0.0;
    }
#pragma unroll
    for (int r = 8; r >= 0; --r) {
        double v = sol[r];
#pragma unroll
        for (int c = r + 1; c < 9; ++c) v -= L[pidx(c, r)] * sol[c];
        sol[r] = r < n ? v / L[pidx(r, r)] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) sol[i] = i < n ? sol[i] / sc[i] : 0.0;
}
__device__ __forceinline__ void precond_apply_block(const double* __restrict__ factors, size_t nblocks, uint32_t A,
                                                    int n, const double right[9], double sol[9]) {
    double L[45], sc[9];
    precond_load(factors, nblocks, A, n, L, sc);
    precond_solve(L, sc, n, right, sol);
}
// PCG_INIT: the state after k_pcg_init (r = -g, z = M^-1 r), before the first direction.
enum { PCG_NORMAL = 0, PCG_CONFIRM = 1, PCG_FINAL = 2, PCG_INIT = 3 };
constexpr int NTS = (int)PCG_BLOCK_WG;  // one lane per block
constexpr int NQ = (int)PCG_Q_WG;        // k_pcg_q: one wavefront per block

// Sum of the per-workgroup partials part[0..nparts) in a fixed order, the same in every
// workgroup (so each one derives the same step lengths and decisions).  Lane t adds
// part[t], part[t + NTS], ... in that order (sixteen loads in flight: a few thousand partials
// take one or two rounds), then a fixed tree.
__device__ double sum_parts(const double* __restrict__ part, uint32_t nparts, double* sh) {
    constexpr int U = 16;
    double acc = 0.0;
    for (uint32_t base = 0; base < nparts; base += U * NTS) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = base + u * NTS + threadIdx.x;
            v[u] = i < nparts ? part[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    return tree_sum<NTS>(acc, sh);
}


}  // namespace

// PCG (pcg.rs:23-119 with the retained confirmation of self_calibration_retained.rs:512-579).
// One iteration is five launches; the vector work runs one lane per block over many workgroups,
// and every workgroup derives the step lengths and the control decisions itself from the
// previous launch's per-workgroup partial dot products (summed in a fixed order):
//   k_pcg_points / k_pcg_cameras   the implicit Schur product (below);
//   k_pcg_q     q = S v per block, partials of d.q;
//   k_pcg_xr    alpha; x += alpha d, r -= alpha q (or the true residual r = -g - S x of a
//               confirmation); z = M^-1 r; partials of r.r and r.z;
//   k_pcg_dir   the convergence tests, beta, the next direction (into the other of the two
//               direction buffers) and the per-camera input of the next product.
// State: st is read by the product passes, k_pcg_q and k_pcg_xr and written by k_pcg_dir;
// st1 is written by k_pcg_xr (and k_pcg_init) and read by k_pcg_dir.  No kernel reads the state
// it writes, so no workgroup sees another's update.
__global__ void __launch_bounds__(NTS) k_pcg_init(Topo tp, PcgVecs V, double tol, int max_iter) {
    __shared__ double sh[NTS];
    const uint32_t A = blockIdx.x * NTS + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (A < tp.nblocks) {
        const int n = (int)tp.blk_dim[A];
        const uint32_t off = tp.blk_off[A];
        double rv[9], sol[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) rv[k] = k < n ? -V.grad[off + k] : 0.0;
        precond_apply_block(V.factors, tp.nblocks, A, n, rv, sol);
        for (int k = 0; k < n; ++k) {
            V.x[off + k] = 0.0;
            V.q[off + k] = 0.0;
            V.r[off + k] = rv[k];
            V.z[off + k] = sol[k];
            rr += rv[k] * rv[k];
            rz += rv[k] * sol[k];
        }
    }
    rr = tree_sum<NTS>(rr, sh);
    rz = tree_sum<NTS>(rz, sh);
    if (threadIdx.x == 0) {
        V.part[V.nqparts + blockIdx.x] = rr;
        V.part[V.nqparts + V.nparts + blockIdx.x] = rz;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PcgState s;
        s.tol = tol;
        s.max_iter = max_iter;
        s.iter = 0;
        s.n0 = NAN;
        s.rel = NAN;
        s.tr = NAN;
        s.rho = 0.0;
        s.done = 0;
        s.accepted = 0;
        s.mode = PCG_INIT;
        s.products = 0;
        s.par = 0;
        s._pad = 0;
        *V.st1 = s;
    }
}

// Point pass: wavefronts of a persistent grid take point chunks (whole points whose observations
// fit the 64 lanes, or one point with more) in turn, one lane per observation; each wavefront
// loads its next chunk's observations while it works on the current one.  The cameras' world
// records sit in LDS when they fit (CAM_LDS), else they are read from global memory.
// e_o = J~_X,o^T (J~_c,o v) is summed per point by a segmented scan over the lanes (a fixed
// shuffle order), and the lane of each point's last observation forms w_p = V^-1 sum e_o and
// writes the point's record (read by the camera pass).
constexpr int PT_NT = 1024;
constexpr int PT_TABLE = 2048 * WREC;  // doubles of the LDS camera table
namespace {
struct PtLoad {  // one lane's observation of a chunk: indices and compact rows
    uint32_t ob, oe, c, q;
    bool live;
    double jx[6], n0, n1, sw;
};
template <class T>
__device__ __forceinline__ void pt_load(const Topo& tp, const Lin<T>& lin, const PcgVecs& V, uint32_t k, uint32_t lane,
                                        PtLoad& L) {
    L.ob = V.pchunk_obs[k];
    L.oe = V.pchunk_obs[k + 1];
    const uint32_t o = L.ob + lane;
    L.live = o < L.oe;
    const uint32_t oo = L.live ? o : L.ob;  // in-range stand-in, its terms dropped
    L.c = tp.obs_cam[oo];
    L.q = tp.obs_pt[oo];
#pragma unroll
    for (int i = 0; i < 6; ++i) L.jx[i] = (double)lin.jpm[soa_at(SOA_JX + i, oo)];
    L.n0 = (double)lin.jpm[soa_at(SOA_Q, oo)];
    L.n1 = (double)lin.jpm[soa_at(SOA_Q + 1, oo)];
    L.sw = (double)lin.jpm[soa_at(SOA_Q + 3, oo)];
}
}  // namespace

template <class T, bool CAM_LDS>
__global__ void __launch_bounds__(PT_NT) k_pcg_points(Topo tp, Lin<T> lin, PointElim<T> pe, PcgVecs V) {
    __shared__ double tab[CAM_LDS ? PT_TABLE : 1];
    if (V.st->done) return;
    if (CAM_LDS) {
        for (uint32_t i = threadIdx.x; i < WREC * tp.nc; i += PT_NT) tab[i] = V.wcam[i];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * (PT_NT / 64);
    auto record = [&](uint32_t c, double w[WREC]) {
        const double* src = CAM_LDS ? tab + WREC * c : V.wcam + WREC * (size_t)c;
#pragma unroll
        for (int i = 0; i < WREC; ++i) w[i] = src[i];
    };
    uint32_t k = blockIdx.x * (PT_NT / 64) + (threadIdx.x >> 6);
    PtLoad cur;
    double X[3], vm[6];
    uint32_t h0 = 0, h1 = 0;
    bool fx = true;
    auto point_data = [&](uint32_t q) {
        for (int i = 0; i < 3; ++i) X[i] = V.X[3 * (size_t)q + i];
        h0 = tp.pt_off[q];
        h1 = tp.pt_off[q + 1];
        fx = tp.pt_fixed[q] != 0;
        const T* vi = pe.vinv + 6 * (size_t)q;
        for (int i = 0; i < 6; ++i) vm[i] = (double)vi[i];
    };
    if (k < V.npchunks) {
        pt_load(tp, lin, V, k, lane, cur);
        point_data(cur.q);
    }
    for (; k < V.npchunks; k += nw) {
        const bool more = k + nw < V.npchunks;
        PtLoad nxt;
        if (more) pt_load(tp, lin, V, k + nw, lane, nxt);  // in flight while this chunk is worked on
        const uint32_t ob = cur.ob, oe = cur.oe;
        const bool big = oe - ob > 64;
        double acc[3] = {0.0, 0.0, 0.0};
        uint32_t q = cur.q;
        if (!big) {
            double w[WREC], e[3];
            record(cur.c, w);
            world_product(cur.jx, cur.n0, cur.n1, cur.sw, X, w, e);
            if (cur.live)
#pragma unroll
                for (int i = 0; i < 3; ++i) acc[i] = e[i];
            const uint32_t head = cur.live ? h0 - ob : lane;
            // inclusive segmented scan: lane l ends with the sum over lanes head..l of its point
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                double sv[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) sv[i] = __shfl_up(acc[i], d, 64);
                if ((int)lane - d >= (int)head)
#pragma unroll
                    for (int i = 0; i < 3; ++i) acc[i] += sv[i];
            }
        } else {  // one point: lanes accumulate over passes, then the wavefront butterfly
            for (uint32_t o = ob + lane; o < oe; o += 64) {
                double jx[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) jx[i] = (double)lin.jpm[soa_at(SOA_JX + i, o)];
                double w[WREC], e[3];
                record(tp.obs_cam[o], w);
                world_product(jx, (double)lin.jpm[soa_at(SOA_Q, o)], (double)lin.jpm[soa_at(SOA_Q + 1, o)],
                              (double)lin.jpm[soa_at(SOA_Q + 3, o)], X, w, e);
#pragma unroll
                for (int i = 0; i < 3; ++i) acc[i] += e[i];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[i] = wave_sum(acc[i]);
            q = tp.obs_pt[ob];
        }
        const bool last = big ? lane == 0 : (cur.live && ob + lane + 1 == h1);
        if (last) {
            double out[3] = {0.0, 0.0, 0.0};
            if (big) point_data(q);
            if (!fx) sym3_mul(vm, acc, out);
            double2* rec = reinterpret_cast<double2*>(V.xw + PCG_XW * (size_t)q);
            rec[0] = make_double2(out[0], out[1]);
            rec[1] = make_double2(out[2], 0.0);
        }
        if (more) {
            cur = nxt;
            point_data(cur.q);
        }
    }
}

// Camera pass: one wavefront per camera segment.  Camera c's camera-major positions are cut into
// windows of pcg_seg<T>() starting at cm_off[c] rounded down to a multiple of VN; lanes read VN
// consecutive positions with 16-byte loads and keep those inside [cm_off[c], cm_off[c+1]).  Per
// observation z_o = J~_c,o v - J~_X,o w_p and m_o = sqrt(w) J_p^T z_o (camera frame); the
// segment's sums F = sum (R X) x m_o, M = sum m_o and I = sum J~_int,o^T z_o go to V.seg, from
// which k_pcg_q forms J~_c^T z = [F; s M; I] (baseline camera: b B^T M).
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_cameras(Topo tp, Lin<T> lin, PcgVecs V) {
    using VT = Vec<T>;
    constexpr int VN = VT::N;
    if (V.st->done) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t k = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (k >= V.nseg) return;  // whole wavefronts leave; no workgroup barrier follows
    const uint32_t c = V.seg_cam[k];
    const uint32_t cb = tp.cm_off[c], ce = tp.cm_off[c + 1];
    constexpr uint32_t SEG = pcg_seg<T>();
    const uint32_t wb = cb / VN * VN + (k - V.cseg_off[c]) * SEG;
    Cam cam;
    load_cam(V.vcam + CAM_REC * (size_t)c, cam);
    const double s = V.gauge->scene_scale;
    double acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = 0.0;
#pragma unroll
    for (uint32_t it = 0; it < SEG / (64 * VN); ++it) {
        const uint32_t j0 = wb + VN * lane + 64 * VN * it;
        typename VT::V qv[4], rv[3];
#pragma unroll
        for (int r = 0; r < 4; ++r) qv[r] = load_group(lin.jcm, SOA_Q + r, j0);
#pragma unroll
        for (int r = 0; r < 3; ++r) rv[r] = load_group(lin.jcm, SOA_RX + r, j0);
        const typename VT::I pts = *reinterpret_cast<const typename VT::I*>(tp.cm_pt + j0);
#pragma unroll
        for (int i = 0; i < VN; ++i) {
            const uint32_t j = j0 + i, p = VT::at(pts, i);
            const double2* wr = reinterpret_cast<const double2*>(V.xw + PCG_XW * (size_t)p);
            const double2 w01 = wr[0], w2 = wr[1];
            const double w[3] = {w01.x, w01.y, w2.x};
            const double RX[3] = {VT::at(rv[0], i), VT::at(rv[1], i), VT::at(rv[2], i)};
            const Obs o = make_obs(VT::at(qv[0], i), VT::at(qv[1], i), VT::at(qv[2], i), VT::at(qv[3], i), cam);
            double y0, y1, t0, t1, mm[3];
            camera_product(o, cam, RX, &y0, &y1);
            point_product(o, cam, s, w, &t0, &t1);
            const double z0 = y0 - t0, z1 = y1 - t1;
            projection_transpose(o, cam, z0, z1, mm);
            const double nz = o.sw * cam.f * (o.n0 * z0 + o.n1 * z1);
            double f[9] = {RX[1] * mm[2] - RX[2] * mm[1], RX[2] * mm[0] - RX[0] * mm[2], RX[0] * mm[1] - RX[1] * mm[0],
                           mm[0], mm[1], mm[2], nz * o.d, nz * o.r2, nz * (o.r2 * o.r2)};
            const bool live = j >= cb && j < ce;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (live) acc[q] += f[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = wave_sum(acc[q]);
    if (lane == 0) {
        double* out = V.seg + 9 * (size_t)k;
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q] = acc[q];
    }
}

// Block pass, one wavefront per block: q_A = (J~^T z rows of block A: its camera's segment
// sums, or for an intrinsic block those of the group's cameras, added lane-strided and then by
// the wavefront butterfly) + lambda diag_A v_A; one partial of v.q per workgroup.  A camera's
// rows are [F; s M] (baseline camera: [F; b B^T M]) from its segments' camera-frame sums.
__global__ void __launch_bounds__(NQ) k_pcg_q(Topo tp, PcgVecs V) {
    __shared__ double sh[NQ / 64];
    const PcgState s = *V.st;
    if (s.done) return;
    const double* __restrict__ v = s.mode == PCG_NORMAL ? V.d[s.par] : V.x;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t A = blockIdx.x * (NQ / 64) + wv;
    double vq = 0.0;
    if (A < tp.nblocks) {  // uniform over the wavefront
        const uint32_t n = tp.blk_dim[A], off = tp.blk_off[A];
        double P[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (A < tp.nc && n > 0) {
            double F[3] = {0.0, 0.0, 0.0}, M[3] = {0.0, 0.0, 0.0};
            for (uint32_t g = V.cseg_off[A] + lane; g < V.cseg_off[A + 1]; g += 64) {
                const double* sg = V.seg + 9 * (size_t)g;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    F[i] += sg[i];
                    M[i] += sg[3 + i];
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                F[i] = wave_sum(F[i]);
                M[i] = wave_sum(M[i]);
            }
            const GaugeDev& gg = *V.gauge;
            for (int i = 0; i < 3; ++i) P[i] = F[i];
            if ((int)A == gg.camera) {
                P[3] = gg.baseline * (gg.basis[0] * M[0] + gg.basis[1] * M[1] + gg.basis[2] * M[2]);
                P[4] = gg.baseline * (gg.basis[3] * M[0] + gg.basis[4] * M[1] + gg.basis[5] * M[2]);
            } else {
                for (int i = 0; i < 3; ++i) P[3 + i] = gg.scene_scale * M[i];
            }
        } else if (A >= tp.nc && n > 0) {
            const uint32_t gr = A - tp.nc;
            double I[3] = {0.0, 0.0, 0.0};
            for (uint32_t i = tp.grp_cam_off[gr]; i < tp.grp_cam_off[gr + 1]; ++i) {
                const uint32_t c = tp.grp_cams[i];
                for (uint32_t g = V.cseg_off[c] + lane; g < V.cseg_off[c + 1]; g += 64) {
                    const double* sg = V.seg + 9 * (size_t)g;
#pragma unroll
                    for (int j = 0; j < 3; ++j) I[j] += sg[6 + j];
                }
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) P[j] = wave_sum(I[j]);
        }
        double term = 0.0;
        if (lane < n) {
            double pk = 0.0;
#pragma unroll
            for (int j = 0; j < 6; ++j)
                if ((int)lane == j) pk = P[j];
            const double vk = v[off + lane];
            const double qk = pk + V.lambda * V.diag[off + lane] * vk;
            V.q[off + lane] = qk;
            term = vk * qk;
        }
        vq = wave_sum(term);
    }
    if (lane == 0) sh[wv] = vq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int i = 0; i < NQ / 64; ++i) acc += sh[i];
        V.part[blockIdx.x] = acc;
    }
}

__global__ void __launch_bounds__(NTS) k_pcg_xr(Topo tp, PcgVecs V) {
    __shared__ double sh[NTS];
    PcgState s = *V.st;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (lead) *V.st1 = s;
        return;
    }
    // every load of the block first (none depends on the step length), then the partial sums
    const bool normal = s.mode == PCG_NORMAL;
    const uint32_t A = blockIdx.x * NTS + threadIdx.x;
    const bool has = A < tp.nblocks;
    const int n = has ? (int)tp.blk_dim[A] : 0;
    const uint32_t off = has ? tp.blk_off[A] : 0u;
    double qv[9], av[9], bv[9], xv[9], L[45], sc[9];
    {
        const double* __restrict__ d = V.d[s.par];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            qv[k] = k < n ? V.q[off + k] : 0.0;
            av[k] = k < n ? (normal ? V.r[off + k] : -V.grad[off + k]) : 0.0;
            bv[k] = (k < n && normal) ? d[off + k] : 0.0;
            xv[k] = (k < n && normal) ? V.x[off + k] : 0.0;
        }
        precond_load(V.factors, tp.nblocks, has ? A : 0u, n, L, sc);
    }
    const double curv = sum_parts(V.part, V.nqparts, sh);
    s.products += 1;
    double alpha = 0.0;
    if (normal) {
        alpha = s.rho / curv;
        if (!(isfinite(curv) && curv > 0.0) || !isfinite(alpha)) {
            s.done = 1;
            if (lead) *V.st1 = s;
            return;
        }
        s.iter += 1;
    }
    double rr = 0.0, rz = 0.0;
    if (has) {
        double rv[9], sol[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (normal) {
                xv[k] += alpha * bv[k];
                rv[k] = av[k] - alpha * qv[k];
            } else {  // q = S x: the true residual
                rv[k] = av[k] - qv[k];
            }
        }
        precond_solve(L, sc, n, rv, sol);
        for (int k = 0; k < n; ++k) {
            if (normal) V.x[off + k] = xv[k];
            V.r[off + k] = rv[k];
            V.z[off + k] = sol[k];
            rr += rv[k] * rv[k];
            rz += rv[k] * sol[k];
        }
    }
    rr = tree_sum<NTS>(rr, sh);
    rz = tree_sum<NTS>(rz, sh);
    if (threadIdx.x == 0) {
        V.part[V.nqparts + blockIdx.x] = rr;
        V.part[V.nqparts + V.nparts + blockIdx.x] = rz;
    }
    if (lead) *V.st1 = s;
}

__global__ void __launch_bounds__(NTS) k_pcg_dir(Topo tp, PcgVecs V) {
    __shared__ double sh[NTS];
    PcgState s = *V.st1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (s.done) {
        if (lead) *V.st = s;
        return;
    }
    // the block's loads first (which of them count is decided below), then the partial sums
    const uint32_t A = blockIdx.x * NTS + threadIdx.x;
    const bool has = A < tp.nblocks, cam = A < tp.nc;
    const uint32_t n = has ? tp.blk_dim[A] : 0u, off = has ? tp.blk_off[A] : 0u;
    const double* __restrict__ dc = V.d[s.par];
    double zr[9], dr[9], xr[6], gzr[3], gdr[3], gxr[3];
    uint32_t pd = 0, gd = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        zr[k] = k < (int)n ? V.z[off + k] : 0.0;
        dr[k] = k < (int)n ? dc[off + k] : 0.0;
    }
    if (cam) {
        const uint32_t g = tp.cam_group[A], go = tp.grp_off[g];
        pd = tp.cam_dim[A];
        gd = tp.grp_dim[g];
#pragma unroll
        for (int k = 0; k < 6; ++k) xr[k] = k < (int)pd ? V.x[off + k] : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const bool in = k < (int)gd;
            gzr[k] = in ? V.z[go + k] : 0.0;
            gdr[k] = in ? dc[go + k] : 0.0;
            gxr[k] = in ? V.x[go + k] : 0.0;
        }
    }
    const double rr = sum_parts(V.part + V.nqparts, V.nparts, sh);
    const double rz = sum_parts(V.part + V.nqparts + V.nparts, V.nparts, sh);
    int upd = 0;  // 1: d = z + beta d, 2: d = z (restart)
    double beta = 0.0;
    if (s.mode == PCG_INIT) {
        s.n0 = sqrt(rr);
        if (!isfinite(s.n0)) {
            s.done = 1;
        } else if (s.n0 == 0.0) {
            s.done = 1;
            s.accepted = 1;
            s.rel = 0.0;
        } else {
            upd = 2;
        }
    } else if (s.mode == PCG_NORMAL) {
        const double ratio = sqrt(rr) / s.n0;
        if (!isfinite(ratio)) {
            s.done = 1;
        } else if (ratio <= s.tol) {
            s.mode = PCG_CONFIRM;
        } else {
            beta = rz / s.rho;
            if (!(isfinite(rz) && rz > 0.0) || !isfinite(beta)) {
                s.done = 1;
            } else {
                upd = 1;
                s.rho = rz;
                if (s.iter >= s.max_iter) s.mode = PCG_FINAL;
            }
        }
    } else {
        const double tr = sqrt(rr) / s.n0;
        if (s.mode == PCG_FINAL) {
            s.done = 1;
            s.rel = isfinite(tr) ? tr : NAN;
        } else if (isfinite(tr) && tr <= s.tol) {
            s.done = 1;
            s.accepted = 1;
            s.rel = tr;
        } else if (!isfinite(tr)) {
            s.done = 1;
        } else {
            s.tr = tr;
            upd = 2;
        }
    }
    if (upd == 2) {
        if (!(isfinite(rz) && rz > 0.0)) {
            s.done = 1;
            s.rel = s.tr;
        } else {
            s.rho = rz;
            s.mode = s.iter >= s.max_iter ? PCG_FINAL : PCG_NORMAL;
        }
    }
    if (!s.done) {
        double* __restrict__ dn = V.d[1 - s.par];
        const bool use_d = s.mode == PCG_NORMAL;  // else the next product is S x
        // the next direction (or x) on index i: the same expression in every lane that needs it
        auto value = [&](double z, double d, double x) { return use_d ? (upd == 1 ? z + beta * d : z) : x; };
        if (has) {
            double zv[9], dv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                zv[k] = upd ? zr[k] : 0.0;
                dv[k] = upd == 1 ? dr[k] : 0.0;
            }
            double vp[6], vi[3];
            if (cam) {
                double xp[6], gz[3], gdv[3], gx[3];
#pragma unroll
                for (int k = 0; k < 6; ++k) xp[k] = !use_d ? xr[k] : 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gz[k] = use_d ? gzr[k] : 0.0;
                    gdv[k] = (use_d && upd == 1) ? gdr[k] : 0.0;
                    gx[k] = !use_d ? gxr[k] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < 6; ++k) vp[k] = k < (int)pd ? value(zv[k], dv[k], xp[k]) : 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) vi[k] = k < (int)gd ? value(gz[k], gdv[k], gx[k]) : 0.0;
            }
            if (upd)
                for (uint32_t k = 0; k < n; ++k) dn[off + k] = upd == 1 ? zv[k] + beta * dv[k] : zv[k];
            if (cam && V.vexp) {  // explicit product: the input in camera coordinates
                double* ve = V.vexp + 9 * (size_t)A;
#pragma unroll
                for (int k = 0; k < 6; ++k) ve[k] = vp[k];
#pragma unroll
                for (int k = 0; k < 3; ++k) ve[6 + k] = vi[k];
            } else if (cam) {
                double* rec = V.vcam + CAM_REC * (size_t)A;
                camera_record(tp, A, V.R + 9 * (size_t)A, V.gp + 3 * (size_t)tp.cam_group[A], *V.gauge, vp, vi, rec);
                Cam cr;
                load_cam(rec, cr);
                double w[WREC];
                world_from_camera(cr, V.gauge->scene_scale, w);
                for (int i = 0; i < WREC; ++i) V.wcam[WREC * (size_t)A + i] = w[i];
            }
        }
        if (upd) s.par = 1 - s.par;
    }
    if (lead) *V.st = s;
}

// ---- explicit reduced camera system ----
namespace {
template <class T>
struct PairMfma;
template <>
struct PairMfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct PairMfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};
}  // namespace

// Off-diagonal blocks, one wavefront per camera pair (c, j):
//   S_cj = -sum_t J~_c,o^T K_t J~_c,o',   K_t = J~_X,o V^-1 J~_X,o'^T (2 x 2),
// i.e. -sum_t W_o V^-1 W_o'^T, over the pair's terms t (the observations o by c and o' by j of
// each common free point, in point order).  The terms go in rounds of 64: lane l rebuilds term
// l's Jacobians from the two observations' compact records (64 bytes each in FP64), the point's
// V^-1 and the two cameras, exactly as the camera-side assembly does, forms T = K J~_c,o' and
// stages the rows of J~_c,o and T in LDS; then the rank-2 updates run two terms per 16x16x4
// matrix-core step (lane l supplies value l%16 of row l/16 of the step's four rows; values 9..15
// are zero), accumulating the block in the wavefront's tile.  A term reads 2 x 64 + 48 bytes
// (instead of two 216-byte precomputed 3 x 9 records), and a camera row's own records and points
// are reread from the L2 of its XCD.  Workgroup slot b holds up to SP_GROUP pairs of one camera
// row (P.group): rows sweep in ascending order over the whole chip, so at any time every XCD
// gathers from the same band of cameras' records, which the shared last-level cache holds, while
// each row stays on one XCD.  The block goes to its slot in row c.
constexpr int SP_WAVES = (int)SP_GROUP;
// The first `n` values of a compact record (n a multiple of 4), as doubles.
template <class T, int N>
__device__ __forceinline__ void load_crec(const T* __restrict__ cr, uint32_t pos, double v[N]) {
    if constexpr (sizeof(T) == 8) {
        const double2* q = reinterpret_cast<const double2*>(cr + (size_t)CREC * pos);
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            const double2 x = q[i];
            v[2 * i] = x.x;
            v[2 * i + 1] = x.y;
        }
    } else {
        const float4* q = reinterpret_cast<const float4*>(cr + (size_t)CREC * pos);
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            const float4 x = q[i];
            v[4 * i] = x.x;
            v[4 * i + 1] = x.y;
            v[4 * i + 2] = x.z;
            v[4 * i + 3] = x.w;
        }
    }
}
template <class T>
__global__ void __launch_bounds__(64 * SP_WAVES) k_schur_pairs(SchurPairs P, const T* __restrict__ cr,
                                                             const double* __restrict__ Rg,
                                                             const double* __restrict__ gp,
                                                             const uint32_t* __restrict__ cam_group,
                                                             const uint32_t* __restrict__ cam_dim,
                                                             const GaugeDev* __restrict__ gauge, SchurRows<T> S) {
    typedef typename PairMfma<T>::v4 v4;
    __shared__ T sa[SP_WAVES][32][2][9];  // rows of J~_c,o per staged term
    __shared__ T sb[SP_WAVES][32][2][9];  // rows of T = K J~_c,o'
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loads below
    if (w >= P.group_n[blockIdx.x]) return;  // whole wavefronts leave; no workgroup barrier follows
    const uint32_t k = P.group[blockIdx.x] + w;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const uint32_t c = P.cam[2 * k], cj = P.cam[2 * k + 1];
    const GaugeDev& gg = *gauge;
    Cam camc, camj;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        camc.R[i] = Rg[9 * (size_t)c + i];
        camj.R[i] = Rg[9 * (size_t)cj + i];
    }
    {
        const double* g3 = gp + 3 * (size_t)cam_group[c];
        camc.f = g3[0], camc.k1 = g3[1], camc.k2 = g3[2];
        const double* h3 = gp + 3 * (size_t)cam_group[cj];
        camj.f = h3[0], camj.k1 = h3[1], camj.k2 = h3[2];
    }
    const uint32_t pdc = cam_dim[c], pdj = cam_dim[cj];
    const bool basec = (int)c == gg.camera, basej = (int)cj == gg.camera;
    v4 acc = v4{T(0), T(0), T(0), T(0)}, acc2 = acc;
    const uint32_t t0 = P.term_off[k], t1 = P.term_off[k + 1];
    const int r = lk & 1, lc = li < 9 ? li : 0;
    const T keep = li < 9 ? T(1) : T(0);
    for (uint32_t base = t0; base < t1; base += 64) {
        double jo[2][9], tm[2][9];
        const uint32_t t = base + (uint32_t)lane;
        if (t < t1) {
            const uint32_t pc = P.term[2 * t], pj = P.term[2 * t + 1];
            double ro[CREC], rj[CREC_HALF];
            load_crec<T, CREC>(cr, pc, ro);
            load_crec<T, CREC_HALF>(cr, pj, rj);
            const Obs oo = make_obs(ro[0], ro[1], ro[2], ro[3], camc);
            const Obs oj = make_obs(rj[0], rj[1], rj[2], rj[3], camj);
            const double RXo[3] = {ro[4], ro[5], ro[6]}, RXj[3] = {rj[4], rj[5], rj[6]};
            const double* vi = ro + 8;  // V^-1 of the term's point
            double xo[6], jcj[2][9], xj[6];
            rebuild_jacobians(oo, camc, RXo, pdc, basec, gg, jo, xo);
            rebuild_jacobians(oj, camj, RXj, pdj, basej, gg, jcj, xj);
            double u0[3], u1[3];  // V^-1 J~_X,o'^T
            sym3_mul(vi, xj, u0);
            sym3_mul(vi, xj + 3, u1);
            const double k00 = xo[0] * u0[0] + xo[1] * u0[1] + xo[2] * u0[2];
            const double k01 = xo[0] * u1[0] + xo[1] * u1[1] + xo[2] * u1[2];
            const double k10 = xo[3] * u0[0] + xo[4] * u0[1] + xo[5] * u0[2];
            const double k11 = xo[3] * u1[0] + xo[4] * u1[1] + xo[5] * u1[2];
#pragma unroll
            for (int b = 0; b < 9; ++b) {
                tm[0][b] = k00 * jcj[0][b] + k01 * jcj[1][b];
                tm[1][b] = k10 * jcj[0][b] + k11 * jcj[1][b];
            }
        } else {
#pragma unroll
            for (int b = 0; b < 9; ++b) jo[0][b] = jo[1][b] = tm[0][b] = tm[1][b] = 0.0;
        }
        // the round's 64 terms are staged 32 at a time (half the LDS of staging all of them)
        const uint32_t n = t1 - base < 64u ? t1 - base : 64u;
        for (uint32_t half = 0; half < 2 && 32 * half < n; ++half) {
            wave_lds_sync();  // the previous steps have read the staged rows
            if ((uint32_t)lane >= 32 * half && (uint32_t)lane < 32 * half + 32) {
                const uint32_t sl = (uint32_t)lane - 32 * half;
#pragma unroll
                for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                    for (int b = 0; b < 9; ++b) {
                        sa[w][sl][rr][b] = (T)jo[rr][b];
                        sb[w][sl][rr][b] = (T)tm[rr][b];
                    }
            }
            wave_lds_sync();
            // steps in groups of four (eight terms; rows past the last term are the zeros staged
            // by the lanes without a term): every LDS read of a group is issued before its
            // matrix-core steps, which alternate between two accumulators
            const uint32_t m = n - 32 * half < 32u ? n - 32 * half : 32u;
            const uint32_t groups = (m + 7) / 8;
            for (uint32_t gi = 0; gi < groups; ++gi) {
                T a[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t tl = 8 * gi + 2 * u + (uint32_t)(lk >> 1);
                    a[u] = sa[w][tl][r][lc];
                    b[u] = sb[w][tl][r][lc];
                }
#pragma unroll
                for (int u = 0; u < 4; u += 2) {
                    acc = PairMfma<T>::mac(a[u] * keep, b[u], acc);
                    acc2 = PairMfma<T>::mac(a[u + 1] * keep, b[u + 1], acc2);
                }
            }
        }
    }
    const uint32_t nnc = S.nbr_off[c + 1] - S.nbr_off[c];
    T* rc = S.vals + S.row_base[c] + 9 * (size_t)(1 + k - S.pair_off[c]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int a = PairMfma<T>::row(lk, v), b = li;
        if (a < 9 && b < 9) rc[(size_t)a * schur_row_len(nnc) + b] = -(acc[v] + acc2[v]);
    }
}

// Explicit product, first pass: one workgroup per camera c over its stored upper row.  Thread t
// takes the 16-byte column groups t, t + PU_NT, ... of all 9 row slices at once, so each stored
// value is read once and serves both of its uses: the row sums q_a += S[c][a][e] v_nbr(e) (the
// inputs gathered from V.vexp per column) and, for the columns e >= 9 of the off-diagonal blocks
// S_cj, the column sums (S_cj^T v_c)_b = sum_a S[c][a][e] v_c,a, written to S.tpair in pair
// order (a row's pairs are consecutive, so the stores are contiguous: scattering them to the
// partners' ranges instead measured 17 us slower per launch on large scenes).  The row sums are reduced over the workgroup (per thread, per wavefront by the
// butterfly, over the wavefronts in order) into S.qrow.
constexpr int PU_NT = 256;
template <class T>
__global__ void __launch_bounds__(PU_NT) k_pcg_product_upper(SchurRows<T> S, PcgVecs V) {
    __shared__ double red[PU_NT / 64][9];
    const PcgState st = *V.st;
    if (st.done) return;
    const uint32_t c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t nb = S.nbr_off[c], nn = S.nbr_off[c + 1] - nb;
    const uint32_t rl = schur_row_len(nn), len = 9 * nn;
    const T* __restrict__ sv = S.vals + S.row_base[c];
    const uint32_t p0 = S.pair_off[c];
    double vc[9], ra[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        vc[a] = V.vexp[9 * (size_t)c + a];  // zero beyond the camera's free dimensions
        ra[a] = 0.0;
    }
    using G = typename Vec<T>::V;
    constexpr uint32_t VN = Vec<T>::N;
    for (uint32_t e0 = VN * tid; e0 < len; e0 += VN * PU_NT) {
        // e0 is a multiple of VN below len, and each slice holds schur_row_len(nn) >= e0 + VN values
        G m[9];
#pragma unroll
        for (int a = 0; a < 9; ++a) m[a] = *reinterpret_cast<const G*>(sv + (size_t)a * rl + e0);
        double vin[VN];
#pragma unroll
        for (uint32_t i = 0; i < VN; ++i) {
            const uint32_t e = e0 + i;
            vin[i] = e < len ? V.vexp[9 * (size_t)S.nbr[nb + e / 9] + e % 9] : 0.0;
        }
#pragma unroll
        for (uint32_t i = 0; i < VN; ++i) {
            const uint32_t e = e0 + i;
            if (e >= len) continue;  // padding: never read as data
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < 9; ++a) {
                const double x = (double)Vec<T>::at(m[a], i);
                ra[a] += x * vin[i];
                t += x * vc[a];
            }
            if (e >= 9) S.tpair[9 * (size_t)p0 + (e - 9)] = t;  // pair p0 + e / 9 - 1, component e % 9
        }
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        const double x = wave_sum(ra[a]);
        if (lane == 0) red[w][a] = x;
    }
    __syncthreads();
    if (tid < 9) {
        double q = 0.0;
#pragma unroll
        for (int i = 0; i < PU_NT / 64; ++i) q += red[i][tid];
        S.qrow[9 * (size_t)c + tid] = q;
    }
}

// Second pass: one workgroup per camera j over the column sums of its pairs (c < j, j): thread t
// adds the 9 values of pairs rev_pair[rev_off[j] + t], + PL_NT, ... (one 72-byte record each), and
// the 9 components are reduced over the workgroup (per wavefront by the butterfly, then over the
// wavefronts in order).  q_j,b = qrow_j,b + that sum + lambda diag v_j,b on the free parameters,
// and one partial of v.q per camera (its components in order).
constexpr int PL_NT = 256;
template <class T>
__global__ void __launch_bounds__(PL_NT) k_pcg_product_lower(Topo tp, SchurRows<T> S, PcgVecs V) {
    __shared__ double red[PL_NT / 64][9];
    __shared__ double vq[9];
    const PcgState st = *V.st;
    if (st.done) return;
    const uint32_t j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t r0 = S.rev_off[j], r1 = S.rev_off[j + 1];
    double acc[9];
#pragma unroll
    for (int b = 0; b < 9; ++b) acc[b] = 0.0;
    for (uint32_t r = r0 + tid; r < r1; r += PL_NT) {
        const double* __restrict__ t = S.tpair + 9 * (size_t)S.rev_pair[r];
        double x[9];
#pragma unroll
        for (int b = 0; b < 9; ++b) x[b] = t[b];
#pragma unroll
        for (int b = 0; b < 9; ++b) acc[b] += x[b];
    }
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        const double x = wave_sum(acc[b]);
        if (lane == 0) red[w][b] = x;
    }
    __syncthreads();
    if (tid < 9) {
        const uint32_t b = tid;
        double q = S.qrow[9 * (size_t)j + b];
#pragma unroll
        for (int i = 0; i < PL_NT / 64; ++i) q += red[i][b];
        const uint32_t pd = tp.cam_dim[j], g = tp.cam_group[j], gd = tp.grp_dim[g];
        const bool live = b < 6 ? b < pd : b - 6 < gd;
        double x = 0.0;
        if (live) {
            const uint32_t par = b < 6 ? tp.cam_off[j] + b : tp.grp_off[g] + (b - 6);
            const double vb = V.vexp[9 * (size_t)j + b];
            q += V.lambda * V.diag[par] * vb;
            V.q[par] = q;
            x = vb * q;
        }
        vq[b] = x;
    }
    __syncthreads();
    if (tid == 0) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) acc += vq[i];
        V.part[j] = acc;
    }
}

// First radial turn of each group's (k1, k2) (model.rs:131-161), +inf where there is none; read by
// the kernels that project with these intrinsics.
__global__ void __launch_bounds__(NT) k_group_turns(uint32_t ng, const double* gp, double* turn) {
    const uint32_t g = blockIdx.x * NT + threadIdx.x;
    if (g >= ng) return;
    double r = 0.0;
    turn[g] = first_turn(gp[3 * g + 1], gp[3 * g + 2], &r) ? r : INFINITY;
}
void launch_group_turns(hipStream_t s, const Topo& tp, const double* gp) {
    if (tp.ng) hipLaunchKernelGGL(k_group_turns, dim3(grid_for(tp.ng, NT)), dim3(NT), 0, s, tp.ng, gp, tp.gturn);
}

// ----------------------------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------------------------
template <class T>
void launch_gauge(hipStream_t s, const T* cam_t, int camera, double scene_scale, GaugeDev* out) {
    hipLaunchKernelGGL(k_gauge<T>, dim3(1), dim3(1), 0, s, cam_t, camera, scene_scale, out);
}
template <class T>
void launch_linearize(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                      const double* X, const double* uv, const GaugeDev* gauge, Lin<T> out) {
    if (tp.no == 0) return;
    launch_group_turns(s, tp, gp);
    hipLaunchKernelGGL(k_linearize<T>, dim3(grid_for(tp.no, NT)), dim3(NT), 0, s, tp, R, t, gp, X, uv, gauge, out);
}
template <class T>
void launch_point_elim(hipStream_t s, const Topo& tp, const Lin<T> lin, T lambda, bool want_rcond, PointElim<T> out) {
    if (tp.nchunks == 0) return;
    hipLaunchKernelGGL(k_point_elim<T>, dim3(tp.nchunks), dim3(NT), 0, s, tp, lin, lambda, (int)want_rcond, out);
}
void launch_inliers(hipStream_t s, const double* sq, const uint8_t* valid, uint32_t no, double limit, double* partials,
                    double* out) {
    hipLaunchKernelGGL(k_inliers_1, dim3(RED_BLOCKS), dim3(NT), 0, s, sq, valid, no, limit, partials);
    hipLaunchKernelGGL(k_inliers_2, dim3(1), dim3(NT), 0, s, partials, out);
}
void launch_reduce_points(hipStream_t s, const double* pscal, uint32_t np, double* partials, AsmScalars* out) {
    hipLaunchKernelGGL(k_reduce_points_1, dim3(RED_BLOCKS), dim3(NT), 0, s, pscal, np, partials);
    hipLaunchKernelGGL(k_reduce_points_2, dim3(1), dim3(NT), 0, s, partials, out);
}
template <class T>
void launch_block_gradient(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, double* grad,
                           double* full_grad, double* diag) {
    if (tp.nblocks == 0) return;
    hipLaunchKernelGGL(k_block_gradient<T>, dim3(tp.nblocks), dim3(NT), 0, s, tp, lin, pe, grad, full_grad, diag);
}
template <class T>
void launch_camera_assembly(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, const double* R,
                            const double* gp, const GaugeDev* gauge, uint32_t nseg, const uint32_t* seg_cam,
                            const uint32_t* cseg_off, double* part, const int32_t* diag_pair,
                            const uint64_t* value_off, double* grad, double* full_grad, double* diag, double* u_vals,
                            double* s_vals, T* z, SchurRows<T> S) {
    if (tp.nc == 0) return;
    if (nseg)  // the third wavefront (cross terms) only for the explicit system
        hipLaunchKernelGGL(k_camera_assembly<T>, dim3(nseg), dim3(S.vals ? 192 : 128), 0, s, tp, lin, pe, R, gp, gauge,
                           seg_cam, cseg_off, part, z);
    hipLaunchKernelGGL(k_camera_assembly_out<T>, dim3(tp.nc), dim3(64), 0, s, tp, cseg_off, part, diag_pair, value_off,
                       grad, full_grad, diag, u_vals, s_vals, S);
}
template <class T>
void launch_pairs(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, uint32_t npairs,
                  const uint32_t* pair_a, const uint32_t* pair_b, const uint32_t* term_off, const uint32_t* term_i,
                  const uint32_t* term_j, const uint64_t* value_off, double* u_vals, double* s_vals) {
    if (npairs == 0) return;
    hipLaunchKernelGGL(k_pairs<T>, dim3(npairs), dim3(NT), 0, s, tp, lin, pe, pair_a, pair_b, term_off, term_i, term_j,
                       value_off, u_vals, s_vals);
}
void launch_scatter_dense(hipStream_t s, const Topo& tp, uint32_t npairs, const uint32_t* pair_a, const uint32_t* pair_b,
                          const uint64_t* value_off, const double* u_vals, const double* s_vals, double* S, uint32_t ld) {
    if (npairs == 0) return;
    hipLaunchKernelGGL(k_scatter_dense, dim3(npairs), dim3(64), 0, s, tp, npairs, pair_a, pair_b, value_off, u_vals,
                       s_vals, S, ld);
}
template <class T>
void launch_scatter_rows(hipStream_t s, const Topo& tp, const SchurRows<T> S, uint32_t max_slots, double* Sd,
                         uint32_t ld) {
    if (tp.nc && max_slots)
        hipLaunchKernelGGL(k_scatter_rows<T>, dim3(tp.nc, (max_slots + SCATTER_SLOTS - 1) / SCATTER_SLOTS), dim3(NT), 0,
                           s, tp, S, Sd, ld);
}
void launch_finalize(hipStream_t s, const Topo& tp, double lambda, double* diag, const double* full_grad,
                     const double* grad, double* S, uint32_t ld, const int32_t* diag_pair, const uint64_t* value_off,
                     double* u_vals, double* partials, AsmScalars* sc) {
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(grid_for(tp.nblocks, FIN_WG), FIN_GRID_MAX));
    hipLaunchKernelGGL(k_finalize_blocks, dim3(g), dim3(FIN_WG), 0, s, tp, lambda, diag, full_grad, grad, S, ld,
                       diag_pair, value_off, u_vals, partials);
    hipLaunchKernelGGL(k_finalize_reduce, dim3(1), dim3(FIN_WG), 0, s, tp, g, (const double*)partials, sc);
}
void launch_retract(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                    const double* step, const GaugeDev* gauge, double* Rc, double* tc, double* gpc) {
    uint32_t n = tp.nc > tp.ng ? tp.nc : tp.ng;
    if (n == 0) return;
    hipLaunchKernelGGL(k_retract<double>, dim3(grid_for(n, NT)), dim3(NT), 0, s, tp, R, t, gp, step, gauge, Rc, tc, gpc);
}
// Per-camera world-frame records of the camera step (at the current rotations and intrinsics) for
// the candidate back-substitution.
__global__ void __launch_bounds__(NT) k_step_records(Topo tp, const double* R, const double* gp, const double* step,
                                                     const GaugeDev* gauges, double* srec) {
    const uint32_t c = blockIdx.x * NT + threadIdx.x;
    if (c >= tp.nc) return;
    double vc[9], rec[CAM_REC], w[WREC];
    camera_values(tp, c, step, vc);
    const GaugeDev& gg = gauges[tp.cam_win ? tp.cam_win[c] : 0];
    camera_record(tp, c, R + 9 * (size_t)c, gp + 3 * (size_t)tp.cam_group[c], gg, vc, vc + 6, rec);
    Cam cr;
    load_cam(rec, cr);
    world_from_camera(cr, gg.scene_scale, w);
    double2* o = reinterpret_cast<double2*>(srec + WREC_PAD * (size_t)c);
    for (int i = 0; i < 4; ++i) o[i] = make_double2(w[2 * i], w[2 * i + 1]);
    o[4] = make_double2(w[8], 0.0);
}
template <class T>
void launch_candidate(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                      const double* Rc, const double* tc, const double* gpc, const double* X, const double* uv,
                      const Lin<T> lin, const PointElim<T> pe, const double* step, const GaugeDev* gauge, double* srec,
                      double* Xc, double* cscal) {
    if (tp.nchunks == 0) return;
    if (tp.nc)
        hipLaunchKernelGGL(k_step_records, dim3(grid_for(tp.nc, NT)), dim3(NT), 0, s, tp, R, gp, step, gauge, srec);
    launch_group_turns(s, tp, gpc);
    hipLaunchKernelGGL(k_candidate<T>, dim3(tp.nchunks), dim3(NT), 0, s, tp, R, t, Rc, tc, gpc, X, uv, lin, pe, srec,
                       gauge, Xc, cscal);
}
void launch_reduce_candidates(hipStream_t s, const double* cscal, uint32_t np, double* partials, CandScalars* out) {
    hipLaunchKernelGGL(k_reduce_cand_1, dim3(RED_BLOCKS), dim3(NT), 0, s, cscal, np, partials);
    hipLaunchKernelGGL(k_reduce_cand_2, dim3(1), dim3(NT), 0, s, partials, out);
}
void launch_step_dots(hipStream_t s, uint32_t D, const double* g, const double* diag, const double* x,
                      const double* product, double* dots) {
    hipLaunchKernelGGL(k_step_dots, dim3(1), dim3(1024), 0, s, D, g, diag, x, product, dots);
}
void launch_limit_step(hipStream_t s, const Topo& tp, double maximum, double* step, double* product, double* alpha) {
    hipLaunchKernelGGL(k_limit_step, dim3(1), dim3(1024), 0, s, tp, maximum, step, product, alpha);
}
void launch_precond_factor(hipStream_t s, const Topo& tp, const int32_t* diag_pair, const uint64_t* value_off,
                           const double* u_vals, const double* s_vals, double* factors, int* fail) {
    if (tp.nblocks == 0) return;
    hipLaunchKernelGGL(k_precond_factor, dim3(grid_for(tp.nblocks, NT)), dim3(NT), 0, s, tp, diag_pair, value_off,
                       u_vals, s_vals, factors, fail);
}
uint32_t pcg_partials(const Topo& tp) { return grid_for(tp.nblocks, NTS); }
template <class T>
void launch_pcg_init(hipStream_t s, const Topo& tp, const PcgVecs& V, double tol, int max_iter) {
    const unsigned g = pcg_partials(tp);
    hipLaunchKernelGGL(k_pcg_init, dim3(g), dim3(NTS), 0, s, tp, V, tol, max_iter);
    hipLaunchKernelGGL(k_pcg_dir, dim3(g), dim3(NTS), 0, s, tp, V);
}
uint32_t pcg_q_partials(const Topo& tp) { return grid_for(tp.nblocks, NQ / 64); }
template <class T>
void launch_pcg_step(hipStream_t s, const Topo& tp, const PcgVecs& V) {
    const unsigned g = pcg_partials(tp);
    hipLaunchKernelGGL(k_pcg_q, dim3(pcg_q_partials(tp)), dim3(NQ), 0, s, tp, V);
    hipLaunchKernelGGL(k_pcg_xr, dim3(g), dim3(NTS), 0, s, tp, V);
    hipLaunchKernelGGL(k_pcg_dir, dim3(g), dim3(NTS), 0, s, tp, V);
}
template <class T>
void launch_pcg_points(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, const PcgVecs& V) {
    if (V.npchunks == 0) return;
    const unsigned waves = PT_NT / 64, need = grid_for(V.npchunks, waves);
    if ((size_t)WREC * tp.nc <= (size_t)PT_TABLE) {  // one resident workgroup per compute unit
        const unsigned g = need < V.ncu ? need : V.ncu;
        hipLaunchKernelGGL((k_pcg_points<T, true>), dim3(g), dim3(PT_NT), 0, s, tp, lin, pe, V);
    } else {
        const unsigned g = need < 8 * V.ncu ? need : 8 * V.ncu;
        hipLaunchKernelGGL((k_pcg_points<T, false>), dim3(g), dim3(PT_NT), 0, s, tp, lin, pe, V);
    }
}
uint32_t pcg_product_partials(const Topo& tp) { return tp.nc; }
template <class T>
void launch_pcg_product_upper(hipStream_t s, const Topo& tp, const SchurRows<T> S, const PcgVecs& V) {
    if (tp.nc) hipLaunchKernelGGL(k_pcg_product_upper<T>, dim3(tp.nc), dim3(PU_NT), 0, s, S, V);
}
template <class T>
void launch_pcg_product_lower(hipStream_t s, const Topo& tp, const SchurRows<T> S, const PcgVecs& V) {
    if (tp.nc) hipLaunchKernelGGL(k_pcg_product_lower<T>, dim3(tp.nc), dim3(PL_NT), 0, s, tp, S, V);
}
template <class T>
void launch_pcg_step_explicit(hipStream_t s, const Topo& tp, const PcgVecs& V) {
    const unsigned g = pcg_partials(tp);
    hipLaunchKernelGGL(k_pcg_xr, dim3(g), dim3(NTS), 0, s, tp, V);
    hipLaunchKernelGGL(k_pcg_dir, dim3(g), dim3(NTS), 0, s, tp, V);
}
template <class T>
void launch_schur_pairs(hipStream_t s, const SchurPairs& P, const T* crec, const double* R, const double* gp,
                        const uint32_t* cam_group, const uint32_t* cam_dim, const GaugeDev* gauge, SchurRows<T> S) {
    if (P.npairs && P.ngroups)
        hipLaunchKernelGGL(k_schur_pairs<T>, dim3(P.ngroups), dim3(64 * SP_WAVES), 0, s, P, crec, R, gp, cam_group,
                           cam_dim, gauge, S);
}
template <class T>
void launch_pcg_cameras(hipStream_t s, const Topo& tp, const Lin<T> lin, const PcgVecs& V) {
    if (V.nseg) hipLaunchKernelGGL(k_pcg_cameras<T>, dim3(grid_for(V.nseg, NT / 64)), dim3(NT), 0, s, tp, lin, V);
}
template <class T>
void launch_linearize_cm(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                         const double* X, const double* uv, const GaugeDev* gauge, Lin<T> out) {
    if (tp.no == 0) return;
    (void)uv;  // the camera-major pixels are in tp.cm_uv
    launch_group_turns(s, tp, gp);
    hipLaunchKernelGGL(k_linearize_cm<T>, dim3(grid_for(tp.no, NT)), dim3(NT), 0, s, tp, R, t, gp, X, gauge, out);
}

__global__ void k_negate(uint32_t n, const double* in, double* out) {
    uint32_t i = blockIdx.x * NT + threadIdx.x;
    if (i < n) out[i] = -in[i];
}
void launch_negate(hipStream_t s, uint32_t n, const double* in, double* out) {
    if (n) hipLaunchKernelGGL(k_negate, dim3(grid_for(n, NT)), dim3(NT), 0, s, n, in, out);
}

template void launch_gauge<double>(hipStream_t, const double*, int, double, GaugeDev*);

#define BA_INSTANTIATE(T)                                                                                          \
    template void launch_point_elim<T>(hipStream_t, const Topo&, const Lin<T>, T, bool, PointElim<T>);               \
    template void launch_block_gradient<T>(hipStream_t, const Topo&, const Lin<T>, const PointElim<T>, double*,      \
                                           double*, double*);                                                      \
    template void launch_camera_assembly<T>(hipStream_t, const Topo&, const Lin<T>, const PointElim<T>,              \
                                            const double*, const double*, const GaugeDev*, uint32_t,               \
                                            const uint32_t*, const uint32_t*, double*, const int32_t*,             \
                                            const uint64_t*, double*, double*, double*, double*, double*, T*,      \
                                            SchurRows<T>);                                                         \
    template void launch_schur_pairs<T>(hipStream_t, const SchurPairs&, const T*, const double*, const double*,      \
                                        const uint32_t*, const uint32_t*, const GaugeDev*, SchurRows<T>);           \
    template void launch_scatter_rows<T>(hipStream_t, const Topo&, const SchurRows<T>, uint32_t, double*, uint32_t);            \
    template void launch_pcg_product_upper<T>(hipStream_t, const Topo&, const SchurRows<T>, const PcgVecs&);                  \
    template void launch_pcg_product_lower<T>(hipStream_t, const Topo&, const SchurRows<T>, const PcgVecs&);                  \
    template void launch_pcg_step_explicit<T>(hipStream_t, const Topo&, const PcgVecs&);                            \
    template void launch_pairs<T>(hipStream_t, const Topo&, const Lin<T>, const PointElim<T>, uint32_t,               \
                                  const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*,              \
                                  const uint32_t*, const uint64_t*, double*, double*);                             \
    template void launch_linearize<T>(hipStream_t, const Topo&, const double*, const double*, const double*,        \
                                      const double*, const double*, const GaugeDev*, Lin<T>);                      \
    template void launch_linearize_cm<T>(hipStream_t, const Topo&, const double*, const double*, const double*,     \
                                         const double*, const double*, const GaugeDev*, Lin<T>);                   \
    template void launch_candidate<T>(hipStream_t, const Topo&, const double*, const double*, const double*,        \
                                      const double*,                                                               \
                                      const double*, const double*, const double*, const double*, const Lin<T>,    \
                                      const PointElim<T>, const double*, const GaugeDev*, double*, double*,        \
                                      double*);                                                                    \
    template void launch_pcg_points<T>(hipStream_t, const Topo&, const Lin<T>, const PointElim<T>, const PcgVecs&);  \
    template void launch_pcg_cameras<T>(hipStream_t, const Topo&, const Lin<T>, const PcgVecs&);                     \
    template void launch_pcg_init<T>(hipStream_t, const Topo&, const PcgVecs&, double, int);                       \
    template void launch_pcg_step<T>(hipStream_t, const Topo&, const PcgVecs&);

BA_INSTANTIATE(double)
BA_INSTANTIATE(float)

}  // namespace ba
