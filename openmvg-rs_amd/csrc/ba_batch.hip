// ba_batch.hip — batched local bundle adjustment: many independent windows solved together.
//
// The windows are concatenated into one problem (cameras, groups, points and observations in
// window order), so the per-observation, per-point and per-block kernels of the single solve run
// over every window in one launch each.  What is per window lives here:
//   k_win_gauge, k_win_reduce_points, k_win_finalize, k_win_step_dots, k_win_reduce_cand
//                     segmented versions of the single-problem scalar passes (fixed order);
//   k_win_scatter     block pairs into one dense reduced system per window;
//   k_win_cholesky    one workgroup per window: blocked Cholesky of S_w on the matrix cores
//                     (FP64, or a Jacobi-scaled FP32 copy with FP64 iterative refinement), the
//                     two triangular solves and the refinement loop, all inside the launch;
//   k_lm_begin/solved/end, k_win_accept
//                     the LM control of each window (sparse/self_calibration/mod.rs:212-639 for a
//                     single-stage loss and the direct step) kept on the device, so an iteration
//                     is a fixed sequence of launches that a HIP graph can replay.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../../include/ba_hip.h"
#include "ba_kernels.hpp"

namespace ba {
namespace {

constexpr int NT = 256;

template <int N>
__device__ double wg_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
template <int N>
__device__ double wg_max(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
template <int N>
__device__ double wg_min(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = N / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmin(sh[t], sh[t + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void k_win_gauge(WinTopo wt, const double* t, GaugeDev* out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= wt.W) return;
    const int c = wt.gauge_cam[w];
    double tc[3] = {0.0, 0.0, 0.0};
    if (c >= 0)
        for (int k = 0; k < 3; ++k) tc[k] = t[3 * (size_t)c + k];
    out[w] = make_gauge(tc, c, wt.scene[w]);
}

// Per-window sums of the per-point scalars (the fields of k_reduce_points, same flag rules).
__global__ void __launch_bounds__(NT) k_win_reduce_points(WinTopo wt, const double* __restrict__ ps, AsmScalars* out) {
    __shared__ double sh[NT];
    const uint32_t w = blockIdx.x;
    double cost = 0, sq = 0, pmr = 0, sgin = 0, rc = 1.0, weak = 0, bad = 0, unproj = 0;
    for (uint32_t p = wt.pt_off[w] + threadIdx.x; p < wt.pt_off[w + 1]; p += NT) {
        const double* s = ps + 6 * (size_t)p;
        const int flags = (int)s[5];
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
    const double v0 = wg_sum<NT>(cost, sh), v1 = wg_sum<NT>(sq, sh), v2 = wg_sum<NT>(pmr, sh);
    const double v3 = wg_max<NT>(sgin, sh), v4 = wg_min<NT>(rc, sh), v5 = wg_sum<NT>(weak, sh);
    const double v6 = wg_sum<NT>(bad, sh), v7 = wg_sum<NT>(unproj, sh);
    if (threadIdx.x == 0) {
        AsmScalars& a = out[w];
        a.cost = v0;
        a.sq = v1;
        a.pmr = v2;
        a.sgin = v3;
        a.min_rcond = v4;
        a.weak = v5;
        a.invalid = v6;
        a.unprojectable = v7;
    }
}

__global__ void k_win_zero(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) p[i] = 0.0;
}

// One workgroup per block pair: U - Schur into the window's dense system (both triangles).
__global__ void k_win_scatter(Topo tp, WinTopo wt, const uint32_t* __restrict__ pair_a,
                              const uint32_t* __restrict__ pair_b, const uint64_t* __restrict__ value_off,
                              const double* __restrict__ u_vals, const double* __restrict__ s_vals, double* S) {
    const uint32_t k = blockIdx.x;
    const uint32_t A = pair_a[k], B = pair_b[k];
    const int dA = (int)tp.blk_dim[A], dB = (int)tp.blk_dim[B];
    const int t = threadIdx.x;
    if (t >= dA * dB) return;
    const uint32_t w = wt.blk_win[A];
    const uint32_t base = wt.param_off[w], n = wt.param_off[w + 1] - base;
    double* Sw = S + wt.S_off[w];
    const int r = t / dB, c = t % dB;
    const uint64_t at = value_off[k] + t;
    const double v = u_vals[at] - s_vals[at];
    const size_t row = tp.blk_off[A] - base + r, col = tp.blk_off[B] - base + c;
    Sw[row * n + col] = v;
    Sw[col * n + row] = v;
}

// One workgroup per window: damping-diagonal clamp per block (schur/assembly.rs:729-766), lambda_w
// diag on the window's dense diagonal (when damp), scaled gradient norm and finiteness.
__global__ void __launch_bounds__(NT) k_win_finalize(Topo tp, WinTopo wt, const double* lam, double* diag,
                                                     const double* full_grad, const double* grad, double* S,
                                                     AsmScalars* out, int damp) {
    __shared__ double sh[NT];
    const uint32_t w = blockIdx.x;
    const int t = threadIdx.x;
    const uint32_t p0 = wt.param_off[w], p1 = wt.param_off[w + 1], n = p1 - p0;
    double mn = INFINITY, mx = 0.0;
    for (uint32_t i = p0 + t; i < p1; i += NT) {
        mn = fmin(mn, diag[i]);
        mx = fmax(mx, diag[i]);
    }
    mn = wg_min<NT>(mn, sh);
    mx = wg_max<NT>(mx, sh);
    // blocks of the window: its pose blocks, then its intrinsic blocks
    const uint32_t c0 = wt.cam_off[w], nc_w = wt.cam_off[w + 1] - c0;
    const uint32_t g0 = wt.grp_off[w], ng_w = wt.grp_off[w + 1] - g0;
    const double lambda = damp ? lam[w] : 0.0;
    double* Sw = S + wt.S_off[w];
    double clamped = 0.0;
    for (uint32_t k = t; k < nc_w + ng_w; k += NT) {
        const uint32_t A = k < nc_w ? c0 + k : tp.nc + g0 + (k - nc_w);
        const uint32_t d = tp.blk_dim[A], off = tp.blk_off[A];
        if (d == 0) continue;
        double scale = 0.0;
        for (uint32_t q = 0; q < d; ++q) scale = fmax(scale, diag[off + q]);
        const double floor = scale > 0.0 ? 64.0 * DBL_EPSILON * scale : 1.0;
        for (uint32_t q = 0; q < d; ++q)
            if (diag[off + q] < floor) {
                diag[off + q] = floor;
                clamped += 1.0;
            }
        if (damp)
            for (uint32_t q = 0; q < d; ++q) {
                const size_t l = off + q - p0;
                Sw[l * n + l] += lambda * diag[off + q];
            }
    }
    clamped = wg_sum<NT>(clamped, sh);
    double sg = 0.0, nonfinite = 0.0;
    for (uint32_t i = p0 + t; i < p1; i += NT) {
        const double cv = diag[i], gv = full_grad[i];
        const double v = (isfinite(cv) && cv > 0.0) ? fabs(gv) / sqrt(cv) : (gv == 0.0 ? 0.0 : INFINITY);
        sg = fmax(sg, v);
        if (!isfinite(grad[i])) nonfinite += 1.0;
    }
    sg = wg_max<NT>(sg, sh);
    nonfinite = wg_sum<NT>(nonfinite, sh);
    if (t == 0) {
        AsmScalars& a = out[w];
        a.sgin_prepared = sg;
        a.min_diag = n == 0 ? 0.0 : mn;
        a.max_diag = mx;
        a.clamped = clamped;
        a.nonfinite = nonfinite;
    }
}

// ---------------------------------------------------------------------------------------------
// Per-window dense Cholesky, one workgroup (four waves) per window, tile edge TB in {32, 64}.
// ---------------------------------------------------------------------------------------------
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
    static constexpr double eps = DBL_EPSILON;
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
    static constexpr double eps = FLT_EPSILON;
};

// C(TB x TB) -= or = X Y^T over K = kn, X and Y tiles in LDS (row-major, ld TB + 1); four waves in
// a 2 x 2 arrangement, each (TB/2)^2 as (TB/32)^2 16x16 matrix-core tiles.
template <class T, int TB>
__device__ void wg_tile_product(const T (*X)[TB + 1], const T (*Y)[TB + 1], int kn,
                                typename Mfma<T>::v4 (&acc)[TB / 32][TB / 32]) {
    typedef typename Mfma<T>::v4 v4;
    constexpr int M = TB / 32;
    const int t = threadIdx.x, wave = t / 64, lane = t % 64;
    const int qr = (wave / 2) * (TB / 2), qc = (wave % 2) * (TB / 2);
    const int li = lane % 16, lk = lane / 16;
    for (int a = 0; a < M; ++a)
        for (int b = 0; b < M; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
    for (int k = 0; k < kn; k += 4) {
        const bool kin = k + lk < kn;
        for (int a = 0; a < M; ++a) {
            const T av = kin ? X[qr + 16 * a + li][k + lk] : T(0);
            for (int b = 0; b < M; ++b) {
                const T bv = kin ? Y[qc + 16 * b + li][k + lk] : T(0);
                acc[a][b] = Mfma<T>::mac(av, bv, acc[a][b]);
            }
        }
    }
}
template <class T, int TB>
__device__ void wg_tile_coords(int a, int b, int v, int* r, int* c) {
    const int t = threadIdx.x, wave = t / 64, lane = t % 64;
    const int qr = (wave / 2) * (TB / 2), qc = (wave % 2) * (TB / 2);
    *r = qr + 16 * a + Mfma<T>::row(lane / 16, v);
    *c = qc + 16 * b + lane % 16;
}

// Factor A (n x n, row-major ld n, lower triangle) in place; Linv receives L_kk^-1 per diagonal
// tile (TB x TB each).  Returns false after a pivot at or below tol.
template <class T, int TB>
__device__ bool wg_factor(T* A, int n, T* Linv, double tol_d, T (*a)[TB + 1], T (*X)[TB + 1], T (*Y)[TB + 1]) {
    constexpr int TPR = NT / TB;  // threads per row of the diagonal tile
    const int t = threadIdx.x, i = t / TPR, q = t % TPR;
    const T tol = (T)tol_d;
    const int ntile = (n + TB - 1) / TB;
    for (int kt = 0; kt < ntile; ++kt) {
        const int k0 = kt * TB, nb = n - k0 < TB ? n - k0 : TB;
        const bool row_in = i < nb;
        for (int c = q; c < TB; c += TPR)
            a[i][c] = (row_in && c <= i) ? A[(size_t)(k0 + i) * n + k0 + c] : (i == c ? T(1) : T(0));
        __syncthreads();
        for (int j = 0; j < nb; ++j) {
            const T piv = a[j][j];
            if (!(isfinite(piv) && piv > tol)) return false;  // uniform
            const T l = sqrt(piv);
            __syncthreads();
            if (q == 0 && row_in && i > j) a[i][j] /= l;
            if (t == 0) a[j][j] = l;
            __syncthreads();
            if (row_in && i > j) {
                const T lij = a[i][j];
                int c = j + 1 + ((q - (j + 1)) % TPR + TPR) % TPR;
                for (; c <= i; c += TPR) a[i][c] -= lij * a[c][j];
            }
            __syncthreads();
        }
        if (row_in)
            for (int c = q; c <= i; c += TPR) A[(size_t)(k0 + i) * n + k0 + c] = a[i][c];
        for (int j = nb - 1; j >= 0; --j) {
            const T xjj = T(1) / a[j][j];
            T s = T(0);
            if (row_in && i > j) {
                int k = j + 1 + ((q - (j + 1)) % TPR + TPR) % TPR;
                for (; k <= i; k += TPR) s += a[i][k] * a[k][j];
            }
            for (int m = 1; m < TPR; m <<= 1) s += __shfl_xor(s, m);  // TPR consecutive lanes
            __syncthreads();
            if (q == 0 && row_in && i > j) a[i][j] = -s * xjj;
            if (t == 0) a[j][j] = xjj;
            __syncthreads();
        }
        T* Li = Linv + (size_t)kt * TB * TB;
        for (int c = q; c < TB; c += TPR) {
            const T v = (row_in && c < nb) ? (c <= i ? a[i][c] : T(0)) : (i == c ? T(1) : T(0));
            Li[i * TB + c] = v;
            Y[i][c] = v;  // keep L_kk^-1 for the panel products
        }
        __syncthreads();
        // panel: X_ik = A_ik L_kk^-T for every row tile below
        for (int ti = kt + 1; ti < ntile; ++ti) {
            const int r0 = ti * TB;
            for (int e = t; e < TB * TB; e += NT) {
                const int r = e / TB, c = e % TB;
                X[r][c] = (r0 + r < n && c < nb) ? A[(size_t)(r0 + r) * n + k0 + c] : T(0);
            }
            __syncthreads();
            typename Mfma<T>::v4 acc[TB / 32][TB / 32];
            wg_tile_product<T, TB>(X, Y, TB, acc);
            for (int x = 0; x < TB / 32; ++x)
                for (int y = 0; y < TB / 32; ++y)
                    for (int v = 0; v < 4; ++v) {
                        int r, c;
                        wg_tile_coords<T, TB>(x, y, v, &r, &c);
                        if (r0 + r < n && c < nb) A[(size_t)(r0 + r) * n + k0 + c] = acc[x][y][v];
                    }
            __syncthreads();
        }
        // trailing update: A_ij -= L_ik L_jk^T for ti >= tj > kt
        for (int ti = kt + 1; ti < ntile; ++ti)
            for (int tj = kt + 1; tj <= ti; ++tj) {
                const int r0 = ti * TB, c0 = tj * TB;
                for (int e = t; e < TB * TB; e += NT) {
                    const int r = e / TB, c = e % TB;
                    X[r][c] = (r0 + r < n && c < nb) ? A[(size_t)(r0 + r) * n + k0 + c] : T(0);
                    a[r][c] = (c0 + r < n && c < nb) ? A[(size_t)(c0 + r) * n + k0 + c] : T(0);
                }
                __syncthreads();
                typename Mfma<T>::v4 acc[TB / 32][TB / 32];
                wg_tile_product<T, TB>(X, (const T(*)[TB + 1])a, nb, acc);
                for (int x = 0; x < TB / 32; ++x)
                    for (int y = 0; y < TB / 32; ++y)
                        for (int v = 0; v < 4; ++v) {
                            int r, c;
                            wg_tile_coords<T, TB>(x, y, v, &r, &c);
                            if (r0 + r < n && c0 + c < n && c0 + c <= r0 + r)
                                A[(size_t)(r0 + r) * n + c0 + c] -= acc[x][y][v];
                        }
                __syncthreads();
            }
    }
    return true;
}

// x = (L L^T)^-1 b by tiles (FP64 vectors; L and L_kk^-1 of type T); y is scratch of n doubles.
template <class T, int TB>
__device__ void wg_solve(const T* L, const T* Linv, int n, const double* b, double* y, double* x, double* part) {
    constexpr int G = NT / TB;  // partial sums per row
    const int t = threadIdx.x;
    const int ntile = (n + TB - 1) / TB;
    // forward: y_k = Linv_kk (b_k - L_k,<k y_<k); thread (r, g) = (t / G, t % G)
    for (int kt = 0; kt < ntile; ++kt) {
        const int k0 = kt * TB, r = t / G, g = t % G;
        double s = 0.0;
        if (k0 + r < n) {
            const T* Lr = L + (size_t)(k0 + r) * n;
            for (int j = g; j < k0; j += G) s += (double)Lr[j] * y[j];
        }
        part[g * TB + r] = s;
        __syncthreads();
        if (t < TB) {
            double v = k0 + t < n ? b[k0 + t] : 0.0;
            for (int q = 0; q < G; ++q) v -= part[q * TB + t];
            part[G * TB + t] = v;
        }
        __syncthreads();
        if (t < TB && k0 + t < n) {
            const T* Li = Linv + (size_t)kt * TB * TB + t * TB;
            double v = 0.0;
            for (int c = 0; c <= t; ++c) v += (double)Li[c] * part[G * TB + c];
            y[k0 + t] = v;
        }
        __syncthreads();
    }
    // backward: x_k = Linv_kk^T (y_k - L_>k,k^T x_>k); thread (g, c) = (t / TB, t % TB)
    for (int kt = ntile - 1; kt >= 0; --kt) {
        const int k0 = kt * TB, c = t % TB, g = t / TB;
        double s = 0.0;
        if (k0 + c < n)
            for (int i = k0 + TB + g; i < n; i += G) s += (double)L[(size_t)i * n + k0 + c] * x[i];
        part[g * TB + c] = s;
        __syncthreads();
        if (t < TB) {
            double v = k0 + t < n ? y[k0 + t] : 0.0;
            for (int q = 0; q < G; ++q) v -= part[q * TB + t];
            part[G * TB + t] = v;
        }
        __syncthreads();
        if (t < TB && k0 + t < n) {
            const T* Li = Linv + (size_t)kt * TB * TB;
            double v = 0.0;
            for (int rr = t; rr < TB; ++rr) v += (double)Li[rr * TB + t] * part[G * TB + rr];
            x[k0 + t] = v;
        }
        __syncthreads();
    }
}

// One workgroup per window.  vecs: 4 doubles per parameter (y, r, dx, scale) at the window's
// parameter range.
// fallback != null (T = double only): run only for windows whose mixed solve failed, factoring the
// FP64 S_w in place, and count the fallback.
template <class T, int TB>
__global__ void __launch_bounds__(NT) k_win_cholesky(WinTopo wt, double* S, T* F, T* linv, const double* grad,
                                                     double* x, double* vecs, const LmWin* lm, int* fail,
                                                     double* refine_rel, uint32_t* fallback) {
    __shared__ T a[TB][TB + 1];
    __shared__ T X[TB][TB + 1];
    __shared__ T Y[TB][TB + 1];
    __shared__ double part[(NT / TB + 1) * TB];
    __shared__ double sh[NT];
    const uint32_t w = blockIdx.x;
    const int t = threadIdx.x;
    if (!lm[w].go) return;
    if (fallback) {
        if (!fail[w]) return;  // uniform
        if (t == 0) fallback[w] += 1;
    }
    const uint32_t p0 = wt.param_off[w];
    const int n = (int)(wt.param_off[w + 1] - p0);
    double* Sw = S + wt.S_off[w];
    T* Fw = F + wt.S_off[w];
    T* Lw = linv + wt.linv_off[w];
    const size_t D = wt.param_off[wt.W];
    double* y = vecs + p0;
    double* r = vecs + D + p0;
    double* dx = vecs + 2 * D + p0;
    double* scale = vecs + 3 * D + p0;
    double* xw = x + p0;
    const double* gw = grad + p0;
    if (n == 0) {
        if (t == 0) fail[w] = 0;
        return;
    }
    // pivot floor 64 eps max|diag| of the matrix factored
    double m = 0.0;
    constexpr bool mixed = sizeof(T) == 4;
    for (int i = t; i < n; i += NT) {
        const double d = Sw[(size_t)i * n + i];
        scale[i] = mixed ? ((isfinite(d) && d > 0.0) ? 1.0 / sqrt(d) : 1.0) : 1.0;
        m = fmax(m, fabs(mixed ? d * scale[i] * scale[i] : d));
    }
    m = wg_max<NT>(m, sh);
    const double tol = 64.0 * Mfma<T>::eps * fmax(m, DBL_MIN);
    T* A = reinterpret_cast<T*>(Sw);
    if (mixed) {
        A = Fw;
        for (size_t e = t; e < (size_t)n * n; e += NT) {
            const int i = (int)(e / n), c = (int)(e % n);
            if (c <= i) Fw[e] = (T)(Sw[e] * scale[i] * scale[c]);
        }
        __syncthreads();
    }
    const bool ok = wg_factor<T, TB>(A, n, Lw, tol, a, X, Y);
    if (!ok) {
        if (t == 0) {
            fail[w] = 1;
            refine_rel[w] = NAN;
        }
        return;
    }
    // right-hand side -g (scaled for the FP32 factor)
    for (int i = t; i < n; i += NT) r[i] = -gw[i] * scale[i];
    __syncthreads();
    wg_solve<T, TB>(A, Lw, n, r, y, dx, part);
    for (int i = t; i < n; i += NT) xw[i] = dx[i] * scale[i];
    __syncthreads();
    if (!mixed) {
        if (t == 0) fail[w] = 0;
        return;
    }
    // iterative refinement against the FP64 S_w: r = -g - S x, stop at ||r|| <= 1e-11 ||g|| or
    // when ||r|| no longer halves
    double prev = INFINITY, rel = NAN;
    for (int step = 0; step < 5; ++step) {
        double rr = 0.0, bb = 0.0;
        for (int i = t; i < n; i += NT) {
            const double* Si = Sw + (size_t)i * n;
            double s = 0.0;
            for (int c = 0; c < n; ++c) s += Si[c] * xw[c];
            const double ri = -gw[i] - s;
            r[i] = ri * scale[i];
            rr += ri * ri;
            bb += gw[i] * gw[i];
        }
        rr = wg_sum<NT>(rr, sh);
        bb = wg_sum<NT>(bb, sh);
        rel = bb > 0.0 ? sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
        if (!(rr > 1e-22 * bb) || !(rr < 0.25 * prev)) break;  // uniform
        prev = rr;
        wg_solve<T, TB>(A, Lw, n, r, y, dx, part);
        for (int i = t; i < n; i += NT) xw[i] += dx[i] * scale[i];
        __syncthreads();
    }
    if (t == 0) {
        refine_rel[w] = rel;
        fail[w] = (isfinite(rel) && rel <= 1e-10) ? 0 : 1;
    }
}

__global__ void __launch_bounds__(NT) k_win_step_dots(WinTopo wt, const double* g, const double* diag,
                                                      const double* x, double* dots) {
    __shared__ double sh[NT];
    const uint32_t w = blockIdx.x;
    double a = 0, b = 0, c = 0;
    for (uint32_t i = wt.param_off[w] + threadIdx.x; i < wt.param_off[w + 1]; i += NT) {
        a += g[i] * x[i];
        b += diag[i] * x[i] * x[i];
        c += x[i] * x[i];
    }
    a = wg_sum<NT>(a, sh);
    b = wg_sum<NT>(b, sh);
    c = wg_sum<NT>(c, sh);
    if (threadIdx.x == 0) {
        dots[4 * w] = a;
        dots[4 * w + 1] = b;
        dots[4 * w + 2] = c;
        dots[4 * w + 3] = 0.0;
    }
}

__global__ void __launch_bounds__(NT) k_win_reduce_cand(WinTopo wt, const double* __restrict__ cs, CandScalars* out) {
    __shared__ double sh[NT];
    const uint32_t w = blockIdx.x;
    double a[4] = {0, 0, 0, 0};
    for (uint32_t p = wt.pt_off[w] + threadIdx.x; p < wt.pt_off[w + 1]; p += NT)
        for (int k = 0; k < 4; ++k) a[k] += cs[4 * (size_t)p + k];
    double v[4];
    for (int k = 0; k < 4; ++k) v[k] = wg_sum<NT>(a[k], sh);
    if (threadIdx.x == 0) {
        out[w].cost = v[0];
        out[w].step2 = v[1];
        out[w].damping_norm = v[2];
        out[w].rejected = v[3];
    }
}

// NielsenController (numerics/refine/nonlinear/controller.rs:15-91) on one window's state.
__device__ void lm_increase(LmWin& s, double& lambda) {
    lambda = fmin(lambda * s.nu, 1e12);
    s.nu = fmin(s.nu * 2.0, 1e12);
}
__device__ void lm_reject(LmWin& s, double& lambda) {
    s.has_gain = 0;
    lm_increase(s, lambda);
}
__device__ bool lm_accept(LmWin& s, double& lambda, double actual, double predicted) {
    const double g = (isfinite(actual) && isfinite(predicted) && predicted > 0.0) ? actual / predicted : NAN;
    s.has_gain = isfinite(g) ? 1 : 0;
    s.gain = g;
    if (!(isfinite(g) && g > 0.0)) {
        lm_increase(s, lambda);
        return false;
    }
    const double c = 2.0 * g - 1.0;
    lambda = fmax(lambda * fmax(1.0 - c * c * c, 1.0 / 3.0), 1e-12);
    s.nu = 2.0;
    return true;
}
__device__ void lm_out_of_iterations(LmWin& s, const LmParams& P) {
    if (!s.done && (uint32_t)s.iterations >= P.max_iterations) s.done = 1;
}

// After the assembly at lambda_w: count the iteration, reject an invalid linearisation, stop on
// the gradient test, else let the window's step run (go).
__global__ void k_lm_begin(WinTopo wt, LmWin* lm, double* lam, const AsmScalars* as, LmParams P, WinTrace tr) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= wt.W) return;
    LmWin s = lm[w];
    double lambda = lam[w];
    s.go = 0;
    s.accepted_now = 0;
    if (s.done) {
        lm[w] = s;
        return;
    }
    const AsmScalars a = as[w];
    const uint32_t it = (uint32_t)s.iterations;
    s.iterations += 1;
    if (!s.have_initial && isfinite(a.cost) && isfinite(a.sq)) {
        s.initial_cost = a.cost;
        s.initial_sq = a.sq;
        s.have_initial = 1;
    }
    if (tr.cost) {
        tr.cost[(size_t)it * wt.W + w] = a.cost;
        tr.lambda[(size_t)it * wt.W + w] = lambda;
        tr.candidate_cost[(size_t)it * wt.W + w] = NAN;
        tr.accepted[(size_t)it * wt.W + w] = 0;
    }
    const bool valid = a.invalid == 0.0 && a.nonfinite == 0.0;
    if (!valid) {
        s.rej[0] += 1;
        lm_reject(s, lambda);
        lm_out_of_iterations(s, P);
        lm[w] = s;
        lam[w] = lambda;
        return;
    }
    const double sgin = fmax(a.sgin, a.sgin_prepared);
    if (isfinite(sgin) && sgin >= 0.0 && isfinite(a.cost) && sgin <= P.gtol * (1.0 + fabs(a.cost))) {
        s.status = BA_STATUS_CONVERGED_GRADIENT;
        s.done = 1;
        lm[w] = s;
        return;
    }
    s.current = a.cost;
    s.pmr = a.pmr;
    s.go = 1;
    lm[w] = s;
}

__global__ void k_lm_solved(WinTopo wt, LmWin* lm, double* lam, const int* fail, LmParams P) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= wt.W) return;
    LmWin s = lm[w];
    if (!s.go || !fail[w]) return;
    double lambda = lam[w];
    s.rej[1] += 1;
    lm_reject(s, lambda);
    s.go = 0;
    lm_out_of_iterations(s, P);
    lm[w] = s;
    lam[w] = lambda;
}

// After the candidate pass: model reduction (self_calibration/optimizer.rs:30-52), gain ratio and
// damping update, the step and cost stop tests (run_state.rs:15-181).
__global__ void k_lm_end(WinTopo wt, LmWin* lm, double* lam, const CandScalars* cs, const double* dots, LmParams P,
                         WinTrace tr) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= wt.W) return;
    LmWin s = lm[w];
    if (!s.go) return;
    double lambda = lam[w];
    const CandScalars c = cs[w];
    const uint32_t it = (uint32_t)s.iterations - 1;
    if (c.rejected != 0.0) {
        s.rej[2] += 1;
        lm_reject(s, lambda);
        s.go = 0;
        lm_out_of_iterations(s, P);
        lm[w] = s;
        lam[w] = lambda;
        return;
    }
    const double npts = (double)(wt.pt_off[w + 1] - wt.pt_off[w]);
    const double D = (double)(wt.param_off[w + 1] - wt.param_off[w]);
    const double point_rms = sqrt(c.step2 / npts);
    const double gdot = dots[4 * w], cdn = dots[4 * w + 1], xx = dots[4 * w + 2];
    const double predicted = (s.pmr - 0.5 * gdot) + 0.5 * lambda * (cdn + c.damping_norm);
    const double reduction = s.current - c.cost;
    const bool acc = lm_accept(s, lambda, reduction, predicted);
    lam[w] = lambda;
    if (tr.cost) {
        tr.candidate_cost[(size_t)it * wt.W + w] = c.cost;
        tr.accepted[(size_t)it * wt.W + w] = acc ? 1 : 0;
    }
    s.go = 0;
    if (!acc) {
        s.rej[3] += 1;
        lm_out_of_iterations(s, P);
        lm[w] = s;
        return;
    }
    s.accepted_now = 1;
    s.accepted += 1;
    s.target_accepted += 1;
    const double rms = sqrt(xx / fmax(D, 1.0));
    const double sn = sqrt(rms * rms + point_rms * point_rms);
    s.step_norm = sn;
    const double relred = reduction / (1.0 + fabs(s.current));
    if (sn <= P.ptol) {
        s.status = BA_STATUS_CONVERGED_STEP;
        s.done = 1;
    } else if (relred <= P.ftol) {
        s.status = BA_STATUS_CONVERGED_COST;
        s.done = 1;
    }
    lm_out_of_iterations(s, P);
    lm[w] = s;
}

// Accepted windows take their candidate state.
__global__ void k_win_accept(Topo tp, WinTopo wt, const LmWin* lm, double* R0, double* t0, double* gp0, double* X0,
                             const double* R1, const double* t1, const double* gp1, const double* X1) {
    const uint32_t i = blockIdx.x * NT + threadIdx.x;
    if (i < tp.nc && lm[tp.cam_win[i]].accepted_now) {
        for (int k = 0; k < 9; ++k) R0[9 * (size_t)i + k] = R1[9 * (size_t)i + k];
        for (int k = 0; k < 3; ++k) t0[3 * (size_t)i + k] = t1[3 * (size_t)i + k];
    }
    if (i < tp.ng && lm[wt.grp_win[i]].accepted_now)
        for (int k = 0; k < 3; ++k) gp0[3 * (size_t)i + k] = gp1[3 * (size_t)i + k];
    if (i < tp.np && lm[tp.pt_win[i]].accepted_now)
        for (int k = 0; k < 3; ++k) X0[3 * (size_t)i + k] = X1[3 * (size_t)i + k];
}

inline unsigned grid_for(size_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

void launch_win_gauge(hipStream_t s, const WinTopo& wt, const double* t, GaugeDev* out) {
    hipLaunchKernelGGL(k_win_gauge, dim3(grid_for(wt.W, 64)), dim3(64), 0, s, wt, t, out);
}
void launch_win_reduce_points(hipStream_t s, const WinTopo& wt, const double* pscal, AsmScalars* out) {
    hipLaunchKernelGGL(k_win_reduce_points, dim3(wt.W), dim3(NT), 0, s, wt, pscal, out);
}
void launch_win_zero(hipStream_t s, double* p, size_t n) {
    if (n) hipLaunchKernelGGL(k_win_zero, dim3(grid_for(n, NT) < 4096 ? grid_for(n, NT) : 4096), dim3(NT), 0, s, p, n);
}
void launch_win_scatter(hipStream_t s, const Topo& tp, const WinTopo& wt, uint32_t npairs, const uint32_t* pair_a,
                        const uint32_t* pair_b, const uint64_t* value_off, const double* u_vals, const double* s_vals,
                        double* S) {
    if (npairs) hipLaunchKernelGGL(k_win_scatter, dim3(npairs), dim3(128), 0, s, tp, wt, pair_a, pair_b, value_off,
                                   u_vals, s_vals, S);
}
void launch_win_finalize(hipStream_t s, const Topo& tp, const WinTopo& wt, const double* lambda, double* diag,
                         const double* full_grad, const double* grad, double* S, AsmScalars* out, bool damp) {
    hipLaunchKernelGGL(k_win_finalize, dim3(wt.W), dim3(NT), 0, s, tp, wt, lambda, diag, full_grad, grad, S, out,
                       (int)damp);
}
template <class T>
void launch_win_cholesky(hipStream_t s, const WinTopo& wt, int tb, double* S, T* F, T* linv, const double* grad,
                         double* x, double* vecs, const LmWin* lm, int* fail, double* refine_rel,
                         double* linv64, uint32_t* fallbacks) {
    if (tb == 32)
        hipLaunchKernelGGL((k_win_cholesky<T, 32>), dim3(wt.W), dim3(NT), 0, s, wt, S, F, linv, grad, x, vecs, lm,
                           fail, refine_rel, (uint32_t*)nullptr);
    else
        hipLaunchKernelGGL((k_win_cholesky<T, 64>), dim3(wt.W), dim3(NT), 0, s, wt, S, F, linv, grad, x, vecs, lm,
                           fail, refine_rel, (uint32_t*)nullptr);
    if (sizeof(T) == 4 && linv64 && fallbacks) {  // FP64 factor for the windows whose refinement failed
        if (tb == 32)
            hipLaunchKernelGGL((k_win_cholesky<double, 32>), dim3(wt.W), dim3(NT), 0, s, wt, S, S, linv64, grad, x,
                               vecs, lm, fail, refine_rel, fallbacks);
        else
            hipLaunchKernelGGL((k_win_cholesky<double, 64>), dim3(wt.W), dim3(NT), 0, s, wt, S, S, linv64, grad, x,
                               vecs, lm, fail, refine_rel, fallbacks);
    }
}
template void launch_win_cholesky<double>(hipStream_t, const WinTopo&, int, double*, double*, double*, const double*,
                                          double*, double*, const LmWin*, int*, double*, double*, uint32_t*);
template void launch_win_cholesky<float>(hipStream_t, const WinTopo&, int, double*, float*, float*, const double*,
                                         double*, double*, const LmWin*, int*, double*, double*, uint32_t*);
void launch_win_step_dots(hipStream_t s, const WinTopo& wt, const double* g, const double* diag, const double* x,
                          double* dots) {
    hipLaunchKernelGGL(k_win_step_dots, dim3(wt.W), dim3(NT), 0, s, wt, g, diag, x, dots);
}
void launch_win_reduce_cand(hipStream_t s, const WinTopo& wt, const double* cscal, CandScalars* out) {
    hipLaunchKernelGGL(k_win_reduce_cand, dim3(wt.W), dim3(NT), 0, s, wt, cscal, out);
}
void launch_lm_begin(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const AsmScalars* a, LmParams P,
                     WinTrace tr) {
    hipLaunchKernelGGL(k_lm_begin, dim3(grid_for(wt.W, 64)), dim3(64), 0, s, wt, lm, lambda, a, P, tr);
}
void launch_lm_solved(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const int* fail, LmParams P) {
    hipLaunchKernelGGL(k_lm_solved, dim3(grid_for(wt.W, 64)), dim3(64), 0, s, wt, lm, lambda, fail, P);
}
void launch_lm_end(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const CandScalars* c,
                   const double* dots, LmParams P, WinTrace tr) {
    hipLaunchKernelGGL(k_lm_end, dim3(grid_for(wt.W, 64)), dim3(64), 0, s, wt, lm, lambda, c, dots, P, tr);
}
void launch_win_accept(hipStream_t s, const Topo& tp, const WinTopo& wt, const LmWin* lm, double* R0, double* t0,
                       double* gp0, double* X0, const double* R1, const double* t1, const double* gp1,
                       const double* X1) {
    const uint32_t n = tp.np > tp.nc ? (tp.np > tp.ng ? tp.np : tp.ng) : (tp.nc > tp.ng ? tp.nc : tp.ng);
    if (n) hipLaunchKernelGGL(k_win_accept, dim3(grid_for(n, NT)), dim3(NT), 0, s, tp, wt, lm, R0, t0, gp0, X0, R1, t1,
                              gp1, X1);
}

}  // namespace ba
