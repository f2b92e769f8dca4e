// ba_cholesky.hip — kernel (d): dense blocked Cholesky of the reduced camera system on gfx950.
//
// Right-looking blocked LL^T on a row-major lower triangle with 64x64 tiles:
//   k_panel        (tile 0) every workgroup factors the diagonal tile in FP64 (blocked by 16:
//                  register-resident 16x16 diagonal blocks, matrix-core panel and trailing
//                  products), checking every pivot against 64 eps max|diag(A)| as the reference's
//                  dense camera solve does (kernel/multi_view/dense/linear_solver.rs:9-61), and forms
//                  L_kk^-1; workgroup b > 0 then forms its panel tile L_ik = A_ik L_kk^-T on the
//                  matrix cores and advances the forward solve L y = b by one tile;
//   k_syrk_update  (one launch per further tile) the trailing update A_ij -= L_ik L_jk^T on the
//                  matrix cores, with a look-ahead: the workgroups of the next column update,
//                  factor and invert the next diagonal tile themselves, as k_panel does, and form
//                  the panel below it and the next forward-solve step.
// (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for the FP32 factor; four waves per 64x64
// tile.)  Only the off-diagonal tiles of L stay in A; the solves use the stored L_kk^-1.  Further
// solves (iterative refinement) and L^T x = y run one launch per tile step across workgroups.
// The mixed form factors a Jacobi-scaled FP32 copy and recovers the FP64 solution by iterative
// refinement against the FP64 matrix (residuals r = b - S x in FP64).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ba_kernels.hpp"


namespace ba {
namespace {

constexpr int TB = 64;  // tile edge
constexpr int TT = TB * TB;

// Matrix-core step C += A B on 16x16x4 tiles: lane l holds A[l%16][l/16] and B[l/16][l%16];
// element v of the accumulator holds D[row(l/16, v)][l%16] (layouts measured on gfx950).
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

template <class T>
__global__ void __launch_bounds__(1024) k_diag_tol(const T* A, uint32_t n, uint32_t ld, double rel, double* work,
                                                   int* fail) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs((double)A[(size_t)i * ld + i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        work[0] = rel * fmax(sh[0], DBL_MIN);
        *fail = 0;
    }
}

// Orders the LDS accesses of one wavefront around this point (lanes exchange data through LDS
// without a workgroup barrier; a wavefront's LDS instructions complete in issue order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copies the 64x64 tile at src (row-major, leading dimension ld) into LDS, converted to the
// destination's type; entries in rows >= nr
// or columns >= nc (or above the diagonal, with `lower`) read as zero.  Every thread issues all
// of its loads before the first LDS store, so the tile arrives in one memory round trip.
template <class T, class S>
__device__ __forceinline__ void load_tile(T (*dst)[TB + 1], const S* src, size_t ld, uint32_t nr, uint32_t nc,
                                          bool lower = false) {
    constexpr int PER = TT / 256;
    T v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = (int)threadIdx.x + 256 * q, r = e / TB, c = e % TB;
        v[q] = ((uint32_t)r < nr && (uint32_t)c < nc && (!lower || c <= r)) ? (T)src[(size_t)r * ld + c] : T(0);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        dst[e / TB][e % TB] = v[q];
    }
}

// The value of v on lane `lane` (uniform), read across lanes without LDS.
template <class T>
__device__ __forceinline__ T lane_value(T v, int lane) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), lane));
    }
}

// 1 / sqrt(p) from the hardware estimate refined by two Newton steps y (3 - p y^2) / 2: a short
// dependent chain in place of a correctly rounded square root and two divisions (within an ulp
// or two of them).
template <class T>
__device__ __forceinline__ T inv_sqrt(T p) {
    T y;
    if constexpr (sizeof(T) == 8)
        y = __builtin_amdgcn_rsq(p);
    else
        y = __builtin_amdgcn_rsqf(p);
    y = y * (T(1.5) - T(0.5) * (p * y) * y);
    y = y * (T(1.5) - T(0.5) * (p * y) * y);
    return y;
}

// Matrix-core product of 16x16 blocks by one wavefront: acc += A op(B), A at `a` (row stride
// lda), B at `b` (row stride ldb), op(B) = B or, with `bt`, B^T.  Lane l supplies A[l%16][k + l/16]
// and op(B)[k + l/16][l%16] for k = 0, 4, 8, 12; element v of acc is C[row(l/16, v)][l%16].
template <class T>
__device__ __forceinline__ typename Mfma<T>::v4 mm16(const T* a, int lda, const T* b, int ldb, bool bt,
                                                     typename Mfma<T>::v4 acc) {
    const int lane = threadIdx.x & 63, li = lane % 16, lk = lane / 16;
#pragma unroll
    for (int k = 0; k < 16; k += 4) {
        const T av = a[li * lda + k + lk];
        const T bv = bt ? b[li * ldb + k + lk] : b[(k + lk) * ldb + li];
        acc = Mfma<T>::mac(av, bv, acc);
    }
    return acc;
}
// Stores sign * acc (a 16x16 block in the matrix-core layout) at `d` (row stride ld).
template <class T>
__device__ __forceinline__ void store16(T* d, int ld, typename Mfma<T>::v4 acc, T sign) {
    const int lane = threadIdx.x & 63, li = lane % 16, lk = lane / 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) d[Mfma<T>::row(lk, v) * ld + li] = sign * acc[v];
}

// Factor and inverse of the 16x16 diagonal block at (p0, p0) of t by one wavefront (see
// factor_diag_blocked); X_pp goes to xs.  Clears *ok_flag when a pivot fails.
template <class T>
__device__ __forceinline__ void factor_block16(T (*t)[TB + 1], T (*xs)[TB + 1], T (*sc)[16][17], T* colb,
                                                         int p0, uint32_t nb, T tol, const double* tol_row,
                                                         int* ok_flag) {
    // lane i < 16 holds row p0 + i of the block (lanes 16..63 repeat rows 0..15; entries above the
    // diagonal are carried along and never read)
    const int i = threadIdx.x % 64, ri = i & 15;
    T r[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) r[c] = t[p0 + ri][p0 + c];
    // a failed pivot is recorded and the (then meaningless) elimination runs on, so that the loops
    // stay fully unrolled with r in registers
    bool ok = true;
    // lane i: column i of X_pp = L_pp^-1 by column substitution from the unit vector e_i, advanced
    // with the factor: column j of L is the broadcast column times 1 / L_jj
    T x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = sc[3][ri][k];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        // column j through LDS (two alternating buffers, so that a column's reads never meet the
        // next column's writes), its values read together after one wait
        T* cb = colb + 16 * (j & 1);
        const T aj = r[j];
        if (i < 16) cb[i] = aj;
        wave_sync();
        T cv[16];
#pragma unroll
        for (int c = j; c < 16; ++c) cv[c] = cb[c];
        const T piv = cv[j];
        // the identity padding of a partial tile is not held to the matrix's tolerance
        const T tj = tol_row && p0 + j < (int)nb ? (T)tol_row[p0 + j] : tol;
        if (p0 + j < (int)nb && !(isfinite(piv) && piv > tj)) ok = false;  // uniform pivot
        const T il = inv_sqrt(piv);
        const T sj = aj * (il * il);
        r[j] = aj * il;  // L_ij below the diagonal, sqrt(a_jj) on lane j
#pragma unroll
        for (int c = j + 1; c < 16; ++c) r[c] -= sj * cv[c];
        x[j] *= il;
        const T xj = x[j] * il;
#pragma unroll
        for (int m = j + 1; m < 16; ++m) x[m] -= cv[m] * xj;  // L_mj x_j = (cv[m] / L_jj) x_j
    }
    if (!ok) {
        if (i == 0) *ok_flag = 0;
        return;
    }
    if (i < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) xs[p0 + c][p0 + i] = x[c];
    }
}

// Factor and inverse of the diagonal tile (size nb <= 64; identity beyond nb) by the four
// wavefronts of the workgroup, blocked by 16:  for each 16-column block p, wave 0 factors the
// 16x16 diagonal block with lanes 0..15 holding its rows in registers (column j broadcast through
// LDS; every pivot checked against 64 eps max|diag(A)| as the reference's dense
// camera solve does, kernel/multi_view/dense/linear_solver.rs:9-61) and inverts it (lane j: column
// j of X_pp = L_pp^-1 by elimination), then the rows below form L_qp = A_qp X_pp^T and the
// trailing blocks A_qs -= L_qp L_sp^T on the matrix cores, one 16x16 block per wavefront.  The
// off-diagonal blocks of X = L^-1 follow by distance from the diagonal:
// X_qp = -X_qq sum_{m=p}^{q-1} L_qm X_mp.  On return `xs` holds X (row-major, zero above the
// diagonal) and `t` the factor; returns false, in every thread, when a pivot fails.  Every
// product sums in a fixed order, so the result does not depend on scheduling.
template <class T>
__device__ __forceinline__ bool factor_diag_blocked(T (*t)[TB + 1], T (*xs)[TB + 1], T (*sc)[16][17], T* colb, uint32_t nb, T tol,
                                    const double* tol_row, int* ok_flag) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    for (int e = tid; e < TT; e += 256) xs[e / TB][e % TB] = T(0);
    if (tid >= (int)nb && tid < TB) t[tid][tid] = T(1);  // identity padding of a partial tile
    sc[3][tid / 16][tid % 16] = tid / 16 == tid % 16 ? T(1) : T(0);  // unit vectors (sc[3] is free until the end)
    if (tid == 0) *ok_flag = 1;
    __syncthreads();
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
        const int p0 = 16 * p;
        if (wave == 0) factor_block16<T>(t, xs, sc, colb, p0, nb, tol, tol_row, ok_flag);
        __syncthreads();
        if (!*ok_flag) return false;
        // rows below the block: L_qp = A_qp X_pp^T, one block per wavefront
        const int below = 3 - p;
        if (wave < below) {
            const int q0 = p0 + 16 * (wave + 1);
            v4 acc = v4{T(0), T(0), T(0), T(0)};
            acc = mm16<T>(&t[q0][p0], TB + 1, &xs[p0][p0], TB + 1, true, acc);
            wave_sync();  // every lane has read the block before any lane overwrites it
            store16<T>(&t[q0][p0], TB + 1, acc, T(1));
        }
        __syncthreads();
        // trailing blocks A_qs -= L_qp L_sp^T (p < s <= q), dealt to the wavefronts in turn
        for (int b = wave; b < below * (below + 1) / 2; b += 4) {
            int q = 0;
            while ((q + 1) * (q + 2) / 2 <= b) ++q;
            const int sblk = b - q * (q + 1) / 2;
            const int q0 = p0 + 16 * (q + 1), s0 = p0 + 16 * (sblk + 1);
            v4 acc = v4{T(0), T(0), T(0), T(0)};
            acc = mm16<T>(&t[q0][p0], TB + 1, &t[s0][p0], TB + 1, true, acc);
            const int li = lane % 16, lk = lane / 16;
#pragma unroll
            for (int v = 0; v < 4; ++v) t[q0 + Mfma<T>::row(lk, v)][s0 + li] -= acc[v];
        }
        __syncthreads();
    }
    // off-diagonal blocks of X by distance from the diagonal
#pragma unroll 1
    for (int dist = 1; dist < 4; ++dist) {
        const int p = wave;  // block (p + dist, p)
        if (p + dist < 4) {
            const int q = p + dist, p0 = 16 * p, q0 = 16 * q;
            v4 acc = v4{T(0), T(0), T(0), T(0)};
            for (int m = p; m < q; ++m) acc = mm16<T>(&t[q0][16 * m], TB + 1, &xs[16 * m][p0], TB + 1, false, acc);
            store16<T>(&sc[wave][0][0], 17, acc, T(1));
            wave_sync();
            v4 out = v4{T(0), T(0), T(0), T(0)};
            out = mm16<T>(&xs[q0][q0], TB + 1, &sc[wave][0][0], 17, false, out);
            store16<T>(&xs[q0][p0], TB + 1, out, T(-1));
        }
        __syncthreads();
    }
    return true;
}

// Tile step kt of the factor, one launch (grid 1 + the row tiles below kt, 256 threads):
// every workgroup factors the diagonal tile A_kk (read-only here) and forms
// X = L_kk^-1 (the same arithmetic in every workgroup); workgroup 0 stores X, workgroup b > 0 the
// panel tile L_ik = A_ik X^T (i = kt + b) in place, as a 64x64x64 matrix-core product.  With `rhs`
// set the forward solve L y = rhs advances with the factor: every workgroup forms
// y_k = X rhs_k, workgroup 0 stores it, workgroup b updates rhs_i -= L_ik y_k.  A failing pivot
// sets *fail, and later steps return at once.
// Panel tile: L_ik = A_ik X_kk^T from the tile A_ik in t and X_kk = L_kk^-1 in xs (LDS, four
// waves on the matrix cores), stored in A; with rhs, rhs_i -= L_ik y_k (y_k in LDS).
template <class T>
__device__ __forceinline__ void panel_tile(T (*t)[TB + 1], T (*xs)[TB + 1], const double* yk, T* A, uint32_t ld,
                                           uint32_t n, uint32_t r0, uint32_t k0, uint32_t nb, double* rhs) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int k = 0; k < TB; k += 4) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const T av = t[qr + 16 * a + li][k + lk];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const T bv = xs[qc + 16 * b + li][k + lk];  // B[k][j] = X[j][k]
                acc[a][b] = Mfma<T>::mac(av, bv, acc[a][b]);
            }
        }
    }
    __syncthreads();  // every wave has read t: overwrite it with L_ik
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int r = qr + 16 * a + Mfma<T>::row(lk, v), c = qc + 16 * b + li;
                t[r][c] = acc[a][b][v];
                if (r0 + r < n && c < (int)nb) A[(size_t)(r0 + r) * ld + k0 + c] = acc[a][b][v];
            }
    if (!rhs) return;
    __syncthreads();
    // rhs_i -= L_ik y_k: thread (row, quarter) sums 16 columns, quarters combined in lane order
    const int row = tid / 4, q = tid % 4;
    double s = 0.0;
    for (int c = 16 * q; c < 16 * q + 16; ++c) s += (double)t[row][c] * yk[c];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0 && r0 + row < n) rhs[r0 + row] -= s;
}

template <class T>
__global__ void __launch_bounds__(256) k_panel(T* A, uint32_t n, uint32_t ld, uint32_t kt, const double* work,
                                               int* fail, T* Linv, double* rhs, double* y, const double* tol_row) {
    // the diagonal tile is factored and inverted in FP64 for either storage type, so that the FP32
    // factor of the mixed form keeps only its panel and trailing products in FP32
    __shared__ double td[TB][TB + 1];
    __shared__ double xd[TB][TB + 1];
    __shared__ double sc[4][16][17];
    __shared__ double colb[32];
    __shared__ double yk[TB];
    __shared__ int ok_flag;
    T (*t)[TB + 1];
    T (*xs)[TB + 1];
    if constexpr (sizeof(T) == sizeof(double)) {
        t = reinterpret_cast<T (*)[TB + 1]>(td);
        xs = reinterpret_cast<T (*)[TB + 1]>(xd);
    } else {
        __shared__ T tf[TB][TB + 1];
        __shared__ T xf[TB][TB + 1];
        t = tf;
        xs = xf;
    }
    if (*fail) return;
    const uint32_t k0 = kt * TB;
    const uint32_t nb = n - k0 < TB ? n - k0 : TB;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    load_tile<double>(td, A + (size_t)k0 * ld + k0, ld, nb, nb, true);
    __syncthreads();
    if (!factor_diag_blocked<double>(td, xd, sc, colb, nb, work[0], tol_row ? tol_row + k0 : nullptr, &ok_flag)) {
        if (tid == 0) *fail = 1;
        return;
    }
    if constexpr (sizeof(T) != sizeof(double)) {
        for (int e = tid; e < TT; e += 256) xs[e / TB][e % TB] = (T)xd[e / TB][e % TB];
        __syncthreads();
    }
    if (rhs && wave == 0) {  // y_k = X rhs_k, rows in a fixed order
        yk[lane] = lane < (int)nb ? rhs[k0 + lane] : 0.0;
        wave_sync();
        double v = 0.0;
        for (int c = 0; c <= lane; ++c) v += (double)xs[lane][c] * yk[c];
        wave_sync();
        yk[lane] = v;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int e = tid; e < TT; e += 256) Linv[(size_t)kt * TT + e] = xs[e / TB][e % TB];
        if (rhs && tid < (int)nb) y[k0 + tid] = yk[tid];
        return;
    }
    const uint32_t r0 = k0 + TB * blockIdx.x;
    load_tile<T>(t, A + (size_t)r0 * ld + k0, ld, n - r0, nb);
    __syncthreads();
    panel_tile<T>(t, xs, yk, A, ld, n, r0, k0, nb, rhs);
}

// Trailing update of tile pair (ti, tj), tj <= ti, both after the panel at k0:
// A_ij -= L_ik L_jk^T with K = nb, both panel tiles staged whole in LDS.  Four waves, each a
// 32x32 quadrant of 2x2 16x16x4 matrix-core tiles.
// Update product of the trailing tile (rows of Li, columns of Lj): acc = Li Lj^T over the panel
// column on the matrix cores, four waves of 2 x 2 16x16 blocks.
template <class T>
__device__ __forceinline__ void update_product(T (*Li)[TB + 1], T (*Lj)[TB + 1], typename Mfma<T>::v4 acc[2][2]) {
    typedef typename Mfma<T>::v4 v4;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int k = 0; k < TB; k += 4) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            T av = Li[qr + 16 * a + li][k + lk];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                T bv = Lj[qc + 16 * b + li][k + lk];  // B[k][j] = L_jk[j][k]
                acc[a][b] = Mfma<T>::mac(av, bv, acc[a][b]);
            }
        }
    }
}
// The updated values A_rc - acc of tile (r0, c0) in the lower triangle (zero elsewhere), in the
// lanes' matrix-core layout.
template <class T>
__device__ __forceinline__ void updated_values(const T* A, uint32_t n, uint32_t ld, uint32_t r0, uint32_t c0,
                                               const typename Mfma<T>::v4 acc[2][2], T nv[2][2][4]) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    T old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t r = r0 + qr + 16 * a + Mfma<T>::row(lk, v), c = c0 + qc + 16 * b + li;
                old[a][b][v] = (r < n && c <= r) ? A[(size_t)r * ld + c] : T(0);
            }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t r = r0 + qr + 16 * a + Mfma<T>::row(lk, v), c = c0 + qc + 16 * b + li;
                nv[a][b][v] = (r < n && c <= r) ? T(old[a][b][v] - acc[a][b][v]) : T(0);
            }
}

// Trailing update A_ij -= L_ik L_jk^T (k the panel column k0), tile pair per workgroup, with a
// look-ahead when Linv is given: the workgroups of the next column (tj = 0: the diagonal tile
// k + 1 and the tiles below it) each update the diagonal tile k + 1 themselves (the same products
// in the same order, so bit-identical in every workgroup), factor and invert it in FP64 (as the
// panel pass does for tile 0), and workgroup 0 stores the inverse and y_{k+1} while the others form
// their panel tiles L_i,k+1 and the forward-solve update.  The diagonal tile k + 1 of A is only
// read in this launch (its inverse is what the solves use), so no workgroup waits for another.
template <class T>
__global__ void __launch_bounds__(256) k_syrk_update(T* A, uint32_t n, uint32_t ld, uint32_t k0, uint32_t nb,
                                                     uint32_t first_tile, int* fail, const double* work, T* Linv,
                                                     double* rhs, double* y, const double* tol_row) {
    typedef typename Mfma<T>::v4 v4;
    // the two panel tiles; then this workgroup's updated tile (storage type), the diagonal tile
    // and its inverse (FP64)
    __shared__ double raw0[TB][TB + 1];
    __shared__ double raw1[TB][TB + 1];
    __shared__ double raw2[TB][TB + 1];
    __shared__ double sc[4][16][17];
    __shared__ double colb[32];
    __shared__ double yk[TB];
    __shared__ int ok_flag;
    T (*Li)[TB + 1] = reinterpret_cast<T (*)[TB + 1]>(raw0);
    T (*Lj)[TB + 1] = reinterpret_cast<T (*)[TB + 1]>(raw1);
    if (*fail) return;
    // blockIdx.x enumerates the lower-triangular tile pairs (ti, tj) of the trailing matrix
    const uint32_t idx = blockIdx.x;
    uint32_t ti = (uint32_t)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= idx) ++ti;
    while (ti * (ti + 1) / 2 > idx) --ti;
    const uint32_t tj = idx - ti * (ti + 1) / 2;
    const uint32_t r0 = (first_tile + ti) * TB, c0 = (first_tile + tj) * TB;
    load_tile<T>(Li, A + (size_t)r0 * ld + k0, ld, n - r0, nb);
    load_tile<T>(Lj, A + (size_t)c0 * ld + k0, ld, n - c0, nb);
    __syncthreads();
    const int t = threadIdx.x;
    const int wave = t / 64, lane = t % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    v4 acc[2][2];
    T nv[2][2][4];
    update_product<T>(Li, Lj, acc);
    updated_values<T>(A, n, ld, r0, c0, acc, nv);
    const bool look = Linv != nullptr && tj == 0;  // uniform in the workgroup
    if (!look) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t r = r0 + qr + 16 * a + Mfma<T>::row(lk, v), c = c0 + qc + 16 * b + li;
                    if (r < n && c <= r) A[(size_t)r * ld + c] = nv[a][b][v];
                }
        return;
    }
    // the diagonal tile k + 1 updated (for workgroup 0 it is its own tile)
    T nd_v[2][2][4];
    if (ti == 0) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) nd_v[a][b][v] = nv[a][b][v];
    } else {
        update_product<T>(Lj, Lj, acc);
        updated_values<T>(A, n, ld, c0, c0, acc, nd_v);
    }
    __syncthreads();  // every wave has read Li and Lj
    T (*tile)[TB + 1] = reinterpret_cast<T (*)[TB + 1]>(raw0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int lr = qr + 16 * a + Mfma<T>::row(lk, v), lc = qc + 16 * b + li;
                raw1[lr][lc] = (double)nd_v[a][b][v];  // lower triangle, as load_tile gives it
                if (ti > 0) tile[lr][lc] = nv[a][b][v];
            }
    __syncthreads();
    const uint32_t nd = n - c0 < TB ? n - c0 : TB;  // the diagonal tile's width
    if (!factor_diag_blocked<double>(raw1, raw2, sc, colb, nd, work[0], tol_row ? tol_row + c0 : nullptr,
                                     &ok_flag)) {
        if (ti == 0 && t == 0) *fail = 1;
        return;
    }
    // X in the storage type (for FP32, into raw1, free after the factor)
    T (*xs)[TB + 1] = reinterpret_cast<T (*)[TB + 1]>(raw2);
    if constexpr (sizeof(T) != sizeof(double)) {
        xs = reinterpret_cast<T (*)[TB + 1]>(raw1);
        for (int e = t; e < TT; e += 256) xs[e / TB][e % TB] = (T)raw2[e / TB][e % TB];
        __syncthreads();
    }
    if (rhs && wave == 0) {  // y_{k+1} = X rhs_{k+1}, rows in a fixed order
        yk[lane] = lane < (int)nd ? rhs[c0 + lane] : 0.0;
        wave_sync();
        double v = 0.0;
        for (int c = 0; c <= lane; ++c) v += (double)xs[lane][c] * yk[c];
        wave_sync();
        yk[lane] = v;
    }
    __syncthreads();
    if (ti == 0) {
        for (int e = t; e < TT; e += 256) Linv[(size_t)first_tile * TT + e] = xs[e / TB][e % TB];
        if (rhs && t < (int)nd) y[c0 + t] = yk[t];
        return;
    }
    panel_tile<T>(tile, xs, yk, A, ld, n, r0, c0, nd, rhs);
}

// Forward solve step kt of L y = r (FP64 vectors; grid 1 + the row tiles below kt): every
// workgroup forms y_k = X_k r_k (X_k = L_kk^-1, rows in a fixed order), workgroup 0 stores it and
// workgroup b updates r_i -= L_ik y_k (i = kt + b).  `skip` (optional) makes the step a no-op.
template <class T>
__global__ void __launch_bounds__(256) k_fwd_step(const T* L, const T* Linv, uint32_t n, uint32_t ld, uint32_t kt,
                                                  double* r, double* y, const int* skip) {
    __shared__ T xs[TB][TB + 1];
    __shared__ T lt[TB][TB + 1];
    __shared__ double rk[TB];
    __shared__ double yk[TB];
    if (skip && *skip) return;
    const uint32_t k0 = kt * TB;
    const uint32_t nb = n - k0 < TB ? n - k0 : TB;
    const uint32_t r0 = k0 + TB * blockIdx.x;
    const int tid = threadIdx.x;
    load_tile<T>(xs, Linv + (size_t)kt * TT, TB, TB, TB);
    if (blockIdx.x) load_tile<T>(lt, L + (size_t)r0 * ld + k0, ld, n - r0, nb);
    if (tid < TB) rk[tid] = tid < (int)nb ? r[k0 + tid] : 0.0;
    __syncthreads();
    if (tid < TB) {
        double v = 0.0;
        for (int c = 0; c <= tid; ++c) v += (double)xs[tid][c] * rk[c];
        yk[tid] = v;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        if (tid < (int)nb) y[k0 + tid] = yk[tid];
        return;
    }
    const int row = tid / 4, q = tid % 4;
    double s = 0.0;
    for (int c = 16 * q; c < 16 * q + 16; ++c) s += (double)lt[row][c] * yk[c];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0 && r0 + row < n) r[r0 + row] -= s;
}

// Backward solve step kt of L^T x = r (grid kt + 1, from the last tile down): every workgroup
// forms x_k = X_k^T r_k, workgroup 0 stores it and workgroup b updates r_j -= L_kj^T x_k
// (j = b - 1 < kt).
template <class T>
__global__ void __launch_bounds__(256) k_bwd_step(const T* L, const T* Linv, uint32_t n, uint32_t ld, uint32_t kt,
                                                  double* r, double* x, const int* skip) {
    __shared__ T xs[TB][TB + 1];
    __shared__ T lt[TB][TB + 1];
    __shared__ double rk[TB];
    __shared__ double xk[TB];
    __shared__ double part[4][TB];
    if (skip && *skip) return;
    const uint32_t k0 = kt * TB;
    const uint32_t nb = n - k0 < TB ? n - k0 : TB;
    const int tid = threadIdx.x;
    const uint32_t j0 = blockIdx.x ? (blockIdx.x - 1) * TB : 0;
    load_tile<T>(xs, Linv + (size_t)kt * TT, TB, TB, TB);
    if (blockIdx.x) load_tile<T>(lt, L + (size_t)k0 * ld + j0, ld, nb, TB);
    if (tid < TB) rk[tid] = tid < (int)nb ? r[k0 + tid] : 0.0;
    __syncthreads();
    if (tid < TB) {  // x_c = sum_{m >= c} X[m][c] r_m
        double v = 0.0;
        for (int m = tid; m < TB; ++m) v += (double)xs[m][tid] * rk[m];
        xk[tid] = v;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        if (tid < (int)nb) x[k0 + tid] = xk[tid];
        return;
    }
    const int c = tid % TB, g = tid / TB;  // column c of tile (kt, j), rows m = g (mod 4)
    double s = 0.0;
    for (int m = g; m < TB; m += 4) s += (double)lt[m][c] * xk[m];
    part[g][c] = s;
    __syncthreads();
    if (tid < TB) r[j0 + tid] -= ((part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]));
}

// Per-pivot tolerances of the rank certificate.  The reference certifies a reduced system of
// dimension <= 2048 by its spectral rank after diagonal equilibration (direct.rs:342-457): with
// s_i = 1 / sqrt(max(A_ii, 64 eps max_j A_jj)), A' = S A S is rank deficient when an eigenvalue
// is at most 64 eps n lambda_max(A').  The factor here applies the same scale-aware test to the
// pivots: the k-th pivot of A' is d_k s_k^2 (d_k the pivot of A), and A' is taken as deficient when
// one is at most tau = 64 eps n g, with g = max_i sum_j |A'_ij| (the Gershgorin bound of lambda_max),
// i.e. when d_k <= tau max(A_kk, floor).  A pivot above it bounds lambda_min(A') from below only
// loosely, so the two rules agree except for eigenvalues within a factor g / lambda_max of the
// tolerance; an exactly singular system fails both.  scratch: s (n), row sums (n), then the
// tolerances (n).  A is the lower triangle (row-major, leading dimension ld).
__global__ void __launch_bounds__(1024) k_rank_scale(const double* A, uint32_t n, uint32_t ld, double* scratch) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(A[(size_t)i * ld + i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const double floor = 64.0 * DBL_EPSILON * sh[0];
    for (uint32_t i = threadIdx.x; i < n; i += 1024) scratch[i] = 1.0 / sqrt(fmax(A[(size_t)i * ld + i], floor));
}
__global__ void __launch_bounds__(256) k_rank_rows(const double* A, uint32_t n, uint32_t ld, double* scratch) {
    __shared__ double sh[256];
    const uint32_t i = blockIdx.x;
    const double* sc = scratch;
    double v = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += 256) {
        const double a = j <= i ? A[(size_t)i * ld + j] : A[(size_t)j * ld + i];
        v += fabs(a) * sc[j];
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) scratch[n + i] = sh[0] * sc[i];
}
__global__ void __launch_bounds__(1024) k_rank_tol(const double* A, uint32_t n, uint32_t ld, double* scratch) {
    __shared__ double sh[1024];
    double g = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) g = fmax(g, scratch[n + i]);
    sh[threadIdx.x] = g;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const double tau = 64.0 * DBL_EPSILON * (double)n * sh[0];
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        const double si = scratch[i];
        scratch[2 * (size_t)n + i] = tau / (si * si);  // tau max(A_ii, floor)
    }
}

__global__ void __launch_bounds__(256) k_copy_vec(uint32_t n, const double* a, double* b, const int* skip) {
    if (skip && *skip) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Iterative refinement: r = b - S x from the FP64 matrix (one wave per row, fixed lane order),
// then the convergence test (one workgroup): stop when ||r|| <= 1e-11 ||b|| or when ||r|| no
// longer halves; state[0] = done flag, state[1] = ||r||^2 of the last step, state[2] = ||b||^2.
__global__ void __launch_bounds__(256) k_refine_residual(const double* S, uint32_t n, uint32_t ld, const double* b,
                                                         const double* x, double* r, const int* done) {
    if (*done) return;
    const uint32_t row = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (row >= n) return;
    const double* Sr = S + (size_t)row * ld;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;  // four chains: four loads in flight
    uint32_t c = lane;
    for (; c + 192 < n; c += 256) {
        s0 += Sr[c] * x[c];
        s1 += Sr[c + 64] * x[c + 64];
        s2 += Sr[c + 128] * x[c + 128];
        s3 += Sr[c + 192] * x[c + 192];
    }
    for (; c < n; c += 64) s0 += Sr[c] * x[c];
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) r[row] = b[row] - s;
}
__global__ void __launch_bounds__(1024) k_refine_check(uint32_t n, const double* b, const double* r, double* state,
                                                       int* done, int first) {
    __shared__ double sh[2][1024];
    if (*done) return;
    double rr = 0.0, bb = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        rr += r[i] * r[i];
        bb += b[i] * b[i];
    }
    sh[0][threadIdx.x] = rr;
    sh[1][threadIdx.x] = bb;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double r2 = sh[0][0], b2 = sh[1][0];
    const bool small = r2 <= 1e-22 * b2;  // ||r|| <= 1e-11 ||b||
    const bool stalled = !first && !(r2 < 0.25 * state[1]);
    if (small || stalled || !isfinite(r2)) *done = 1;
    if (!stalled || first) state[1] = r2;
    state[2] = b2;
}
__global__ void __launch_bounds__(256) k_refine_axpy(uint32_t n, const double* dx, double* x, const int* done) {
    if (*done) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += dx[i];
}
// Jacobi scaling of the FP32 copy: F = D^-1/2 S D^-1/2 (unit diagonal), scale[i] = S_ii^-1/2 (1
// when S_ii is not positive; the factor then reports the breakdown).  The FP32 factor of the
// scaled matrix keeps the units of pose, intrinsic and gauge columns apart from its rounding.
__global__ void __launch_bounds__(256) k_jacobi_scale(const double* S, uint32_t n, uint32_t ld, double* scale) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = S[(size_t)i * ld + i];
    scale[i] = (isfinite(d) && d > 0.0) ? 1.0 / sqrt(d) : 1.0;
}
__global__ void __launch_bounds__(256) k_lower_to_f32(const double* S, uint32_t n, uint32_t ld, const double* scale,
                                                      float* F) {
    for (uint32_t row = blockIdx.y; row < n; row += gridDim.y)
        for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c <= row && c < n; c += 256 * gridDim.x)
            F[(size_t)row * ld + c] = (float)(S[(size_t)row * ld + c] * scale[row] * scale[c]);
}
__global__ void k_clear_flag(int* flag) { *flag = 0; }
// y = scale * v (elementwise), skipped once `done` is set
__global__ void __launch_bounds__(256) k_scale_vec(uint32_t n, const double* scale, const double* v, double* y,
                                                   const int* done) {
    if (done && *done) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = scale[i] * v[i];
}

// Workspace (doubles): [0, 16) scalars | y n | r n | dx n | scale n | sv n | w n | L_kk^-1 tiles.
inline size_t linv_offset(uint32_t n) { return 16 + 6 * (size_t)n; }
inline double* work_y(double* work) { return work + 16; }
inline double* work_w(double* work, uint32_t n) { return work + 16 + 5 * (size_t)n; }

// Factor A (lower triangle, in place) tile step by tile step; with `rhs` the forward solve
// L y = rhs advances with it (rhs is consumed).
// Pivots must exceed rel * max|diag(A)| (and be finite).
template <class T>
void factor(hipStream_t s, T* A, uint32_t n, uint32_t ld, double* work, int* fail, double* rhs, double* y,
            double rel = 64.0 * Mfma<T>::eps, const double* tol_row = nullptr) {
    hipLaunchKernelGGL(k_diag_tol<T>, dim3(1), dim3(1024), 0, s, A, n, ld, rel, work, fail);
    T* Linv = reinterpret_cast<T*>(work + linv_offset(n));
    const uint32_t ntile = (n + TB - 1) / TB;
    // Tile 0 and its panel by k_panel; then each trailing update k also factors tile k + 1 (its
    // workgroup 0) and forms the panel below it (the workgroups of its column, once workgroup 0 has
    // published the inverse), while its other workgroups update the rest of the trailing matrix.
    const uint32_t ntile0 = ntile;
    hipLaunchKernelGGL(k_panel<T>, dim3(ntile0), dim3(256), 0, s, A, n, ld, 0u, (const double*)work, fail, Linv, rhs,
                       y, tol_row);
    for (uint32_t kt = 0; kt + 1 < ntile; ++kt) {
        const uint32_t k0 = kt * TB;
        const uint32_t nb = n - k0 < TB ? n - k0 : TB;
        const uint32_t rest = ntile - kt - 1;
        const uint32_t pairs = rest * (rest + 1) / 2;
        hipLaunchKernelGGL(k_syrk_update<T>, dim3(pairs), dim3(256), 0, s, A, n, ld, k0, nb, kt + 1, fail, work, Linv,
                           rhs, y, tol_row);
    }
}
// x = L^-T y (y is left unchanged).
template <class T>
void backward(hipStream_t s, const T* L, uint32_t n, uint32_t ld, const double* y, double* x, double* work,
              const int* skip) {
    const T* Linv = reinterpret_cast<const T*>(work + linv_offset(n));
    double* w = work_w(work, n);
    const uint32_t ntile = (n + TB - 1) / TB;
    hipLaunchKernelGGL(k_copy_vec, dim3((n + 255) / 256), dim3(256), 0, s, n, y, w, skip);
    for (int kt = (int)ntile - 1; kt >= 0; --kt)
        hipLaunchKernelGGL(k_bwd_step<T>, dim3(kt + 1), dim3(256), 0, s, L, Linv, n, ld, (uint32_t)kt, w, x, skip);
}
// x = L^-T L^-1 b with a factored L.
template <class T>
void apply(hipStream_t s, const T* L, uint32_t n, uint32_t ld, const double* b, double* x, double* work, const int* skip) {
    const T* Linv = reinterpret_cast<const T*>(work + linv_offset(n));
    double* y = work_y(work);
    double* w = work_w(work, n);
    const uint32_t ntile = (n + TB - 1) / TB;
    hipLaunchKernelGGL(k_copy_vec, dim3((n + 255) / 256), dim3(256), 0, s, n, b, w, skip);
    for (uint32_t kt = 0; kt < ntile; ++kt)
        hipLaunchKernelGGL(k_fwd_step<T>, dim3(ntile - kt), dim3(256), 0, s, L, Linv, n, ld, kt, w, y, skip);
    backward<T>(s, L, n, ld, y, x, work, skip);
}

// Inverse columns of the trailing factor, for the marginals: R starts as the unit columns e_j,
// j0 <= j < n, restricted to the rows [k0, n) (k0 = the tile row of j0; row-major, m = n - j0
// columns, leading dimension ldr).  Step kt (k0 / 64 <= kt < ntile): every workgroup forms
// Y_k = X_k R_k (X_k = L_kk^-1) for its column block; workgroup 0 stores it in Y, workgroup b > 0
// updates the rows of tile kt + b by R_i -= L_ik Y_k (R_k itself is only read in this step).
// After the last step Y = L^-1 E on rows [k0, n) (the rows above are zero).  Grid (1 + the tiles
// below kt, column blocks of 32); every sum runs in a fixed order.
constexpr int INV_COLS = 32;
__global__ void __launch_bounds__(256) k_inverse_step(const double* L, const double* Linv, uint32_t n, uint32_t ld,
                                                      uint32_t kt, uint32_t k0, double* R, double* Y, uint32_t m,
                                                      uint32_t ldr, const int* fail) {
    __shared__ double xs[TB][TB + 1];
    __shared__ double yb[TB][INV_COLS + 1];
    if (*fail) return;
    const uint32_t t0 = kt * TB;                     // first row of tile kt
    const uint32_t nb = n - t0 < TB ? n - t0 : TB;
    const uint32_t c0 = blockIdx.y * INV_COLS;       // first column of this block
    const int tid = threadIdx.x;
    load_tile<double>(xs, Linv + (size_t)kt * TT, TB, TB, TB);
    for (int e = tid; e < TB * INV_COLS; e += 256) {
        const int r = e / INV_COLS, c = e % INV_COLS;
        yb[r][c] = ((uint32_t)r < nb && c0 + c < m) ? R[(size_t)(t0 - k0 + r) * ldr + c0 + c] : 0.0;
    }
    __syncthreads();
    // Y_k = X_k R_k: thread (row, 8 columns), sums over q <= row ascending
    const int row = tid / 4, cb = (tid % 4) * 8;
    double acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0;
    for (int q = 0; q <= row; ++q) {
        const double x = xs[row][q];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] += x * yb[q][cb + c];
    }
    __syncthreads();  // every thread has read R_k
#pragma unroll
    for (int c = 0; c < 8; ++c) yb[row][cb + c] = acc[c];
    __syncthreads();
    if (blockIdx.x == 0) {
        if ((uint32_t)row < nb)
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c0 + cb + c < m) Y[(size_t)(t0 - k0 + row) * ldr + c0 + cb + c] = yb[row][cb + c];
        return;
    }
    const uint32_t r0 = t0 + TB * blockIdx.x;         // first row of tile kt + b
    load_tile<double>(xs, L + (size_t)r0 * ld + t0, ld, n - r0, nb);
    __syncthreads();
    if (r0 + row >= n) return;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0;
    for (int q = 0; q < TB; ++q) {
        const double l = xs[row][q];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] += l * yb[q][cb + c];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (c0 + cb + c < m) R[(size_t)(r0 - k0 + row) * ldr + c0 + cb + c] -= acc[c];
}

// R = the unit columns e_{j0 + c} on the rows [k0, n).
__global__ void __launch_bounds__(256) k_unit_columns(double* R, uint32_t rows, uint32_t m, uint32_t ldr,
                                                      uint32_t diag0) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)rows * ldr) return;
    const uint32_t r = (uint32_t)(i / ldr), c = (uint32_t)(i % ldr);
    R[i] = (c < m && r == diag0 + c) ? 1.0 : 0.0;
}

// Per group g (one workgroup each): cov_g[a][b] = w * sum_r Y[r][col(g, a)] Y[r][col(g, b)]
// (= w (S^-1)_gg, since S^-1 = L^-T L^-1 and the group's columns are among the trailing ones; R
// here is the Y of k_inverse_step),
// the rows summed in a fixed partition and order.  col(g, a) = off[g] - j0 + a.
__global__ void __launch_bounds__(256) k_group_marginals(const double* R, uint32_t rows, uint32_t ldr,
                                                         const uint32_t* goff, const uint32_t* gdim, uint32_t j0,
                                                         double w, double* cov) {
    __shared__ double part[256][6];
    const uint32_t g = blockIdx.x, d = gdim[g];
    const int tid = threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (d) {
        const uint32_t c = goff[g] - j0;
        for (uint32_t r = tid; r < rows; r += 256) {
            double y[3];
            for (uint32_t a = 0; a < 3; ++a) y[a] = a < d ? R[(size_t)r * ldr + c + a] : 0.0;
            acc[0] += y[0] * y[0];
            acc[1] += y[1] * y[0];
            acc[2] += y[1] * y[1];
            acc[3] += y[2] * y[0];
            acc[4] += y[2] * y[1];
            acc[5] += y[2] * y[2];
        }
    }
    for (int k = 0; k < 6; ++k) part[tid][k] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
            for (int k = 0; k < 6; ++k) part[tid][k] += part[tid + s][k];
        __syncthreads();
    }
    if (tid == 0) {
        const double v[6] = {part[0][0], part[0][1], part[0][2], part[0][3], part[0][4], part[0][5]};
        const int idx[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}};
        for (uint32_t a = 0; a < 3; ++a)
            for (uint32_t b = 0; b < 3; ++b)
                cov[9 * (size_t)g + 3 * a + b] = (a < d && b < d) ? w * v[idx[a][b]] : 0.0;
    }
}

}  // namespace

size_t cholesky_work_doubles(uint32_t n) {
    const size_t ntile = (n + TB - 1) / TB;
    return linv_offset(n) + ntile * TT;
}

void launch_cholesky_solve(hipStream_t s, double* A, uint32_t n, uint32_t ld, const double* b, double* x, double* work,
                           int* fail) {
    if (n == 0) return;
    double* w = work_w(work, n);
    hipLaunchKernelGGL(k_copy_vec, dim3((n + 255) / 256), dim3(256), 0, s, n, b, w, (const int*)nullptr);
    factor<double>(s, A, n, ld, work, fail, w, work_y(work));
    backward<double>(s, A, n, ld, work_y(work), x, work, (const int*)fail);
}

void launch_cholesky_solve_mixed(hipStream_t s, const double* S, float* F, uint32_t n, uint32_t ld, const double* b,
                                 double* x, double* work, int* fail, int* done, int max_steps) {
    if (n == 0) return;
    double* r = work + 16 + n;
    double* dx = r + n;
    double* scale = dx + n;
    double* sv = scale + n;
    const unsigned nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_jacobi_scale, dim3(nb), dim3(256), 0, s, S, n, ld, scale);
    hipLaunchKernelGGL(k_lower_to_f32, dim3(nb, n < 65535u ? n : 65535u), dim3(256), 0, s, S, n, ld,
                       (const double*)scale, F);
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, done);  // a kernel: no memset node in a captured graph
    // x = D^-1/2 F^-1 D^-1/2 b, the forward solve advancing with the factor
    hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, b, work_w(work, n),
                       (const int*)nullptr);
    factor<float>(s, F, n, ld, work, fail, work_w(work, n), work_y(work));
    backward<float>(s, F, n, ld, work_y(work), dx, work, (const int*)fail);
    hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)dx, x,
                       (const int*)nullptr);
    const unsigned rows = (n + 3) / 4;
    for (int k = 0; k < max_steps; ++k) {
        hipLaunchKernelGGL(k_refine_residual, dim3(rows), dim3(256), 0, s, S, n, ld, b, (const double*)x, r,
                           (const int*)done);
        hipLaunchKernelGGL(k_refine_check, dim3(1), dim3(1024), 0, s, n, b, (const double*)r, work + 8, done, (int)(k == 0));
        hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)r, sv,
                           (const int*)done);
        apply<float>(s, F, n, ld, sv, dx, work, done);
        hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)dx, dx,
                           (const int*)done);
        hipLaunchKernelGGL(k_refine_axpy, dim3(nb), dim3(256), 0, s, n, (const double*)dx, x, (const int*)done);
    }
}

void launch_cholesky_factor(hipStream_t s, double* A, uint32_t n, uint32_t ld, double* work, int* fail,
                            double rel_floor, const double* tol_row) {
    if (n == 0) return;
    factor<double>(s, A, n, ld, work, fail, nullptr, nullptr, rel_floor, tol_row);
}

void launch_rank_tolerances(hipStream_t s, const double* A, uint32_t n, uint32_t ld, double* scratch) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_rank_scale, dim3(1), dim3(1024), 0, s, A, n, ld, scratch);
    hipLaunchKernelGGL(k_rank_rows, dim3(n), dim3(256), 0, s, A, n, ld, scratch);
    hipLaunchKernelGGL(k_rank_tol, dim3(1), dim3(1024), 0, s, A, n, ld, scratch);
}

size_t trailing_inverse_doubles(uint32_t n, uint32_t j0) {
    const uint32_t k0 = (j0 / TB) * TB, m = n - j0;
    return 2 * (size_t)(n - k0) * (m ? m : 1);
}

void launch_group_marginals(hipStream_t s, const double* L, uint32_t n, uint32_t ld, const double* work,
                            const int* fail, uint32_t j0, uint32_t ngroups, const uint32_t* goff,
                            const uint32_t* gdim, double w, double* R, double* cov) {
    if (j0 >= n || ngroups == 0) return;
    const uint32_t k0 = (j0 / TB) * TB, m = n - j0, rows = n - k0, ldr = m;
    const double* Linv = work + linv_offset(n);
    double* Y = R + (size_t)rows * ldr;
    hipLaunchKernelGGL(k_unit_columns, dim3((unsigned)(((size_t)rows * ldr + 255) / 256)), dim3(256), 0, s, R, rows, m,
                       ldr, j0 - k0);
    const uint32_t ntile = (n + TB - 1) / TB;
    for (uint32_t kt = k0 / TB; kt < ntile; ++kt)
        hipLaunchKernelGGL(k_inverse_step, dim3(ntile - kt, (m + INV_COLS - 1) / INV_COLS), dim3(256), 0, s, L, Linv,
                           n, ld, kt, k0, R, Y, m, ldr, fail);
    hipLaunchKernelGGL(k_group_marginals, dim3(ngroups), dim3(256), 0, s, (const double*)Y, rows, ldr, goff, gdim, j0,
                       w, cov);
}

}  // namespace ba
