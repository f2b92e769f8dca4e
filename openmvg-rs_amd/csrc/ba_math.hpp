// ba_math.hpp — per-observation mathematics shared by the HIP kernels (FP64 and FP32 forms).
//
// Radial camera model of crates/openmvg-ba/src/kernel/multi_view/self_calibration/model.rs:26-129
// and :308-324, robust losses of numerics/robust.rs:8-44, pose motion of schur/point.rs:730-762
// and the manifold maps of numerics/manifold.rs and numerics/manifold/so3.rs.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#define BA_HD __host__ __device__ __forceinline__

namespace ba {

template <class T>
struct Eps;
template <>
struct Eps<double> {
    static constexpr double value = DBL_EPSILON;
};
template <>
struct Eps<float> {
    static constexpr float value = FLT_EPSILON;
};

template <class T>
BA_HD bool finite(T x) {
    return __builtin_isfinite(x);
}

enum : int { LOSS_L2 = 0, LOSS_HUBER = 1, LOSS_CAUCHY = 2 };

// first_positive_radial_turn, model.rs:266-292: the first positive root of 1 + 3 k1 r^2 + 5 k2 r^4
// (in r^2); returns false when the increasing branch never ends.
template <class T>
BA_HD bool first_turn(T k1, T k2, T* root) {
    T scale = fmax(fmax(fabs(k1), fabs(k2)), T(1));
    T a = T(5) * (k2 / scale);
    T b = T(3) * (k1 / scale);
    T c = T(1) / scale;
    if (a == T(0)) {
        T r = -c / b;
        if (b < T(0) && finite(r) && r > T(0)) {
            *root = r;
            return true;
        }
        return false;
    }
    T disc = fma(b, b, T(-4) * a * c);
    if (disc < T(0) || !finite(disc)) return false;
    T sq = sqrt(disc);
    T q = T(-0.5) * (b + copysign(sq, b));
    T r0, r1;
    if (q == T(0)) {
        r0 = -b / (T(2) * a);
        r1 = T(NAN);
    } else {
        r0 = q / a;
        r1 = c / q;
    }
    bool found = false;
    T best = T(0);
    if (finite(r0) && r0 > T(0)) {
        best = r0;
        found = true;
    }
    if (finite(r1) && r1 > T(0) && (!found || r1 < best)) {
        best = r1;
        found = true;
    }
    *root = best;
    return found;
}

// Residual (prediction minus measurement), 2x3 point Jacobian in the camera frame, and the 2x3
// intrinsic Jacobian d/d(log f, k1, k2).  Returns false outside the projection domain
// (model.rs:26-75): depth <= 64 eps max|X|, r^2 beyond the first radial turn, or a radial
// Jacobian determinant <= 1e-12, or any non-finite value.
template <class T>
BA_HD bool radial_linearize(T f, T k1, T k2, T turn, bool has_turn, const T p[3], T u, T v, T res[2],
                            T J[2][3], T Ji[2][3]) {
    T scale = fmax(fmax(fabs(p[0]), fabs(p[1])), fabs(p[2]));
    if (!finite(p[0]) || !finite(p[1]) || !finite(p[2]) || !(p[2] > T(64) * Eps<T>::value * scale))
        return false;
    T iz = T(1) / p[2];
    T n0 = p[0] * iz, n1 = p[1] * iz;
    T r2 = n0 * n0 + n1 * n1;
    if (has_turn && r2 >= turn) return false;
    T r4 = r2 * r2;
    T d = T(1) + k1 * r2 + k2 * r4;
    T rd = k1 + T(2) * (k2 * r2);
    T det = d * (d + T(2) * r2 * rd);
    if (!finite(det) || det <= T(1e-12)) return false;
    T fd = f * d;
    T pr0 = fd * n0, pr1 = fd * n1;
    T dx = T(2) * n0 * rd, dy = T(2) * n1 * rd;
    T a00 = f * (d + n0 * dx), a01 = f * n0 * dy;
    T a10 = f * n1 * dx, a11 = f * (d + n1 * dy);
    J[0][0] = a00 * iz;
    J[0][1] = a01 * iz;
    J[0][2] = -(a00 * n0 + a01 * n1) * iz;
    J[1][0] = a10 * iz;
    J[1][1] = a11 * iz;
    J[1][2] = -(a10 * n0 + a11 * n1) * iz;
    res[0] = pr0 - u;
    res[1] = pr1 - v;
    Ji[0][0] = pr0;
    Ji[1][0] = pr1;
    Ji[0][1] = f * r2 * n0;
    Ji[1][1] = f * r2 * n1;
    Ji[0][2] = f * r4 * n0;
    Ji[1][2] = f * r4 * n1;
    bool ok = finite(res[0]) && finite(res[1]) && finite(r4) && finite(d);
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) ok = ok && finite(J[r][c]) && finite(Ji[r][c]);
    return ok;
}

// Residual only: radial_projection (model.rs:81-87) without the Jacobians.
template <class T>
BA_HD bool radial_residual(T f, T k1, T k2, T turn, bool has_turn, const T p[3], T u, T v, T res[2]) {
    T scale = fmax(fmax(fabs(p[0]), fabs(p[1])), fabs(p[2]));
    if (!finite(p[0]) || !finite(p[1]) || !finite(p[2]) || !(p[2] > T(64) * Eps<T>::value * scale))
        return false;
    T iz = T(1) / p[2];
    T n0 = p[0] * iz, n1 = p[1] * iz;
    T r2 = n0 * n0 + n1 * n1;
    if (has_turn && r2 >= turn) return false;
    T r4 = r2 * r2;
    T d = T(1) + k1 * r2 + k2 * r4;
    T rd = k1 + T(2) * (k2 * r2);
    T det = d * (d + T(2) * r2 * rd);
    if (!finite(det) || det <= T(1e-12)) return false;
    T fd = f * d;
    res[0] = fd * n0 - u;
    res[1] = fd * n1 - v;
    // the Jacobian entries must be finite too (model.rs:123-128); they are when these are
    T dx = T(2) * n0 * rd, dy = T(2) * n1 * rd;
    T a00 = f * (d + n0 * dx), a11 = f * (d + n1 * dy);
    return finite(res[0]) && finite(res[1]) && finite(a00 * iz) && finite(a11 * iz) &&
           finite(f * n0 * dy * iz) && finite(f * n1 * dx * iz);
}

// Robust loss of a residual norm: cost and IRLS square-root weight (numerics/robust.rs:8-44).
template <class T>
BA_HD T loss_cost(int kind, T scale, T r) {
    if (kind == LOSS_HUBER) {
        T m = fabs(r);
        return m <= scale ? T(0.5) * r * r : scale * (m - T(0.5) * scale);
    }
    if (kind == LOSS_CAUCHY) {
        T x = r / scale;
        return log1p(T(0.5) * (x * x));
    }
    return T(0.5) * r * r;
}
template <class T>
BA_HD T loss_sqrt_weight(int kind, T scale, T r) {
    if (kind == LOSS_HUBER) {
        T m = fabs(r);
        return (m <= scale || m == T(0)) ? T(1) : sqrt(scale / m);
    }
    if (kind == LOSS_CAUCHY) {
        T x = r / scale;
        return sqrt(T(1) / (scale * scale * (T(0.5) * (x * x) + T(1))));
    }
    return T(1);
}

// 3x3 helpers on row-major arrays.
template <class T>
BA_HD void rotate(const T* R, const T X[3], T o[3]) {
    o[0] = R[0] * X[0] + R[1] * X[1] + R[2] * X[2];
    o[1] = R[3] * X[0] + R[4] * X[1] + R[5] * X[2];
    o[2] = R[6] * X[0] + R[7] * X[1] + R[8] * X[2];
}
template <class T>
BA_HD void rotate_t(const T* R, const T X[3], T o[3]) {
    o[0] = R[0] * X[0] + R[3] * X[1] + R[6] * X[2];
    o[1] = R[1] * X[0] + R[4] * X[1] + R[7] * X[2];
    o[2] = R[2] * X[0] + R[5] * X[1] + R[8] * X[2];
}

// exp_so3 (so3.rs:254-265): I + sin K + (1 - cos) K^2 on the normalised axis.
template <class T>
BA_HD void exp_so3(const T v[3], T R[9]) {
    T th = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    T a[3], lin, quad;
    if (th < T(1e-10)) {
        a[0] = v[0];
        a[1] = v[1];
        a[2] = v[2];
        lin = T(1);
        quad = T(0.5);
    } else {
        a[0] = v[0] / th;
        a[1] = v[1] / th;
        a[2] = v[2] / th;
        lin = sin(th);
        quad = T(1) - cos(th);
    }
    T K[9] = {T(0), -a[2], a[1], a[2], T(0), -a[0], -a[1], a[0], T(0)};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T k2 = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
            R[3 * i + j] = ((i == j ? T(1) : T(0)) + K[3 * i + j] * lin) + k2 * quad;
        }
}

// retract_s2 (manifold.rs:144-158) in the tangent basis rows b (6 values).
template <class T>
BA_HD void retract_s2(const T t[3], const T* b, T d0, T d1, T out[3]) {
    T v[3] = {b[0] * d0 + b[3] * d1, b[1] * d0 + b[4] * d1, b[2] * d0 + b[5] * d1};
    T th = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    T tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    T tv[3] = {t[0] / tn, t[1] / tn, t[2] / tn};
    if (th < T(1e-10)) {
        T s[3] = {tv[0] + v[0], tv[1] + v[1], tv[2] + v[2]};
        T n = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int k = 0; k < 3; ++k) out[k] = s[k] / n;
    } else {
        T c = cos(th), s = sin(th) / th;
        for (int k = 0; k < 3; ++k) out[k] = tv[k] * c + v[k] * s;
    }
}

// Jacobi-scaled Cholesky inverse of a symmetric positive definite 3x3 block (point_block.rs:112-157);
// m packed lower [m00, m10, m11, m20, m21, m22]; writes the packed inverse.
template <class T>
BA_HD bool point_block_inverse(const T m[6], T inv[6]) {
    T diag[3] = {m[0], m[2], m[5]};
    T ds = fmax(fmax(fmax(T(0), diag[0]), diag[1]), diag[2]);
    T floor = T(64) * Eps<T>::value * ds;
    T s0 = sqrt(fmax(diag[0], floor)), s1 = sqrt(fmax(diag[1], floor)), s2 = sqrt(fmax(diag[2], floor));
    T tiny = T(64) * Eps<T>::value;
    T a00 = m[0] / (s0 * s0);
    if (!(finite(a00) && a00 > tiny)) return false;
    T l00 = sqrt(a00);
    T l10 = (m[1] / (s1 * s0)) / l00;
    T a11 = m[2] / (s1 * s1) - l10 * l10;
    if (!(finite(a11) && a11 > tiny)) return false;
    T l11 = sqrt(a11);
    T l20 = (m[3] / (s2 * s0)) / l00;
    T l21 = (m[4] / (s2 * s1) - l20 * l10) / l11;
    T a22 = m[5] / (s2 * s2) - l20 * l20 - l21 * l21;
    if (!(finite(a22) && a22 > tiny)) return false;
    T l22 = sqrt(a22);
    T m00 = T(1) / l00, m11 = T(1) / l11, m22 = T(1) / l22;
    T m10 = -l10 * m00 * m11;
    T m21 = -l21 * m11 * m22;
    T m20 = -(l20 * m00 + l21 * m10) * m22;
    T d0 = T(1) / s0, d1 = T(1) / s1, d2 = T(1) / s2;
    inv[0] = (m00 * m00 + m10 * m10 + m20 * m20) * (d0 * d0);
    inv[1] = (m11 * m10 + m21 * m20) * (d1 * d0);
    inv[2] = (m11 * m11 + m21 * m21) * (d1 * d1);
    inv[3] = (m22 * m20) * (d2 * d0);
    inv[4] = (m22 * m21) * (d2 * d1);
    inv[5] = (m22 * m22) * (d2 * d2);
    return true;
}

template <class T>
BA_HD void sym3_mul(const T m[6], const T x[3], T o[3]) {
    o[0] = m[0] * x[0] + m[1] * x[1] + m[3] * x[2];
    o[1] = m[1] * x[0] + m[2] * x[1] + m[4] * x[2];
    o[2] = m[3] * x[0] + m[4] * x[1] + m[5] * x[2];
}

// Track geometry of point_geometry.rs:29-101.
template <class T>
struct TrackGeometry {
    T near2, far2, scene2;
    T first[3];
    int count;
    bool has_first, parallax, invalid;
    BA_HD explicit TrackGeometry(T scene_length) {
        near2 = T(INFINITY);
        far2 = scene_length * scene_length;
        scene2 = scene_length * scene_length;
        count = 0;
        has_first = parallax = invalid = false;
        first[0] = first[1] = first[2] = T(0);
    }
    BA_HD void push(const T o[3]) {
        T sq = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
        if (!(finite(sq) && sq > T(0))) {
            invalid = true;
            return;
        }
        count += 1;
        near2 = fmin(near2, sq);
        far2 = fmax(far2, sq);
        T inv = T(1) / sqrt(sq);
        T ray[3] = {o[0] * inv, o[1] * inv, o[2] * inv};
        if (!has_first) {
            has_first = true;
            first[0] = ray[0];
            first[1] = ray[1];
            first[2] = ray[2];
            return;
        }
        if (parallax) return;
        T dot = ray[0] * first[0] + ray[1] * first[1] + ray[2] * first[2];
        T c0 = ray[1] * first[2] - ray[2] * first[1];
        T c1 = ray[2] * first[0] - ray[0] * first[2];
        T c2 = ray[0] * first[1] - ray[1] * first[0];
        T s2 = c0 * c0 + c1 * c1 + c2 * c2;
        const T ps = T(1.7e-5);
        parallax = dot < T(0) || s2 >= ps * ps;
    }
    BA_HD bool degenerate() const {
        const T r = T(1e-4);
        return invalid || (count >= 2 && (!parallax || near2 < r * r * far2));
    }
    BA_HD bool collapsed() const {
        const T r = T(1e-4);
        return invalid || (count >= 2 && near2 < r * r * scene2);
    }
};

}  // namespace ba
