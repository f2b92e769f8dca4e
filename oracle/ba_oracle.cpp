// ba_oracle.cpp — CPU oracle (TEST INFRASTRUCTURE ONLY) of openmvg-ba's self-calibrating
// Levenberg–Marquardt path.  A plain sequential FP64 restatement of the reference algorithm;
// every function cites the reference file it follows (paths under crates/openmvg-ba/src).
// Built with -ffp-contract=off so that, like the reference's Rust, no multiply-add is fused
// except where the reference calls mul_add.
//
// The reference's direct camera solve is faer's AMD-ordered sparse LLT (numerics/sparse.rs);
// faer 0.24.4 is not in this image, so this oracle uses the reference's own dense Cholesky
// (kernel/multi_view/dense/linear_solver.rs:9-61), which the reference's tests hold to the
// sparse result within 5e-8 relative cost (tests/sparse_self_calibration/dense_parity.rs).
#include "ba_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

namespace bao {

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<Vec3, 3>;
constexpr double EPS = std::numeric_limits<double>::epsilon();
constexpr double MINPOS = std::numeric_limits<double>::min();
constexpr double INF = std::numeric_limits<double>::infinity();
constexpr double NaN = std::numeric_limits<double>::quiet_NaN();

struct Radial {
    double focal, k1, k2;
};
struct Pose {
    Mat3 R;
    Vec3 t;
};

static inline double fmax_(double a, double b) { return std::fmax(a, b); }
static inline double fmin_(double a, double b) { return std::fmin(a, b); }

// ---------------------------------------------------------------------------------------
// Radial camera model: kernel/multi_view/self_calibration/model.rs
// ---------------------------------------------------------------------------------------
struct Terms {
    double n[2];
    double r2, r4, d;
    double pred[2];
    double invz;
};

// first_positive_radial_turn, model.rs:266-292
static bool first_positive_radial_turn(Radial c, double* root_out) {
    double scale = fmax_(fmax_(std::fabs(c.k1), std::fabs(c.k2)), 1.0);
    double a = 5.0 * (c.k2 / scale);
    double b = 3.0 * (c.k1 / scale);
    double cc = 1.0 / scale;
    if (a == 0.0) {
        double root = -cc / b;
        if (b < 0.0 && std::isfinite(root) && root > 0.0) {
            *root_out = root;
            return true;
        }
        return false;
    }
    double disc = std::fma(b, b, -4.0 * a * cc);
    if (disc < 0.0 || !std::isfinite(disc)) return false;
    double sq = std::sqrt(disc);
    double q = -0.5 * (b + std::copysign(sq, b));
    double roots[2];
    if (q == 0.0) {
        roots[0] = -b / (2.0 * a);
        roots[1] = NaN;
    } else {
        roots[0] = q / a;
        roots[1] = cc / q;
    }
    bool found = false;
    double best = 0.0;
    for (double r : roots) {
        if (std::isfinite(r) && r > 0.0) {
            if (!found || r < best) best = r;
            found = true;
        }
    }
    if (found) *root_out = best;
    return found;
}

// radial_terms_with_depth (positive depth), model.rs:26-75
static bool radial_terms(Radial c, const double p[3], Terms& t) {
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale = fmax_(scale, std::fabs(p[i]));
    double depth_floor = 64.0 * EPS * scale;
    bool valid_depth = p[2] > depth_floor;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || !valid_depth)
        return false;
    t.invz = 1.0 / p[2];
    t.n[0] = p[0] * t.invz;
    t.n[1] = p[1] * t.invz;
    t.r2 = t.n[0] * t.n[0] + t.n[1] * t.n[1];
    double bound;
    if (first_positive_radial_turn(c, &bound) && t.r2 >= bound) return false;
    t.r4 = t.r2 * t.r2;
    t.d = 1.0 + c.k1 * t.r2 + c.k2 * t.r4;
    double rd = c.k1 + 2.0 * (c.k2 * t.r2);
    double slope = t.d + 2.0 * t.r2 * rd;
    double det = t.d * slope;
    if (!std::isfinite(det) || det <= 1e-12) return false;
    t.pred[0] = c.focal * t.d * t.n[0];
    t.pred[1] = c.focal * t.d * t.n[1];
    return std::isfinite(t.pred[0]) && std::isfinite(t.pred[1]) && std::isfinite(t.r2) &&
           std::isfinite(t.r4) && std::isfinite(t.d);
}

// radial_projection_from_terms, model.rs:89-129
static bool projection_from_terms(Radial c, const double obs[2], const Terms& t, double res[2],
                                  double J[2][3]) {
    double rd = c.k1 + 2.0 * (c.k2 * t.r2);
    double dx = 2.0 * t.n[0] * rd;
    double dy = 2.0 * t.n[1] * rd;
    double nj[2][2] = {{c.focal * (t.d + t.n[0] * dx), c.focal * t.n[0] * dy},
                       {c.focal * t.n[1] * dx, c.focal * (t.d + t.n[1] * dy)}};
    for (int r = 0; r < 2; ++r) {
        J[r][0] = nj[r][0] * t.invz;
        J[r][1] = nj[r][1] * t.invz;
        J[r][2] = -(nj[r][0] * t.n[0] + nj[r][1] * t.n[1]) * t.invz;
    }
    res[0] = t.pred[0] - obs[0];
    res[1] = t.pred[1] - obs[1];
    for (int r = 0; r < 2; ++r) {
        if (!std::isfinite(res[r])) return false;
    }
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(J[r][k])) return false;
    return true;
}

// radial_intrinsic_jacobian_from_terms, model.rs:308-324: d/d(log f, k1, k2)
static bool intrinsic_from_terms(Radial c, const Terms& t, double Ji[2][3]) {
    for (int r = 0; r < 2; ++r) {
        Ji[r][0] = t.pred[r];
        Ji[r][1] = c.focal * t.r2 * t.n[r];
        Ji[r][2] = c.focal * t.r4 * t.n[r];
    }
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(Ji[r][k])) return false;
    return true;
}

// radial_linearization, model.rs:143-172
static bool radial_linearization(Radial c, const double obs[2], const double p[3], double res[2],
                                 double J[2][3], double Ji[2][3]) {
    Terms t;
    if (!radial_terms(c, p, t)) return false;
    if (!projection_from_terms(c, obs, t, res, J)) return false;
    return intrinsic_from_terms(c, t, Ji);
}

// radial_projection, model.rs:81-87
static bool radial_projection(Radial c, const double obs[2], const double p[3], double res[2],
                              double J[2][3]) {
    Terms t;
    if (!radial_terms(c, p, t)) return false;
    return projection_from_terms(c, obs, t, res, J);
}

static double radial_radius(Radial c, double radius) {
    double r2 = radius * radius;
    return radius * (1.0 + c.k1 * r2 + c.k2 * r2 * r2);
}

// RadialCamera::is_valid: positive finite focal, finite coefficients
static bool radial_valid(Radial c) {
    return std::isfinite(c.focal) && c.focal > 0.0 && std::isfinite(c.k1) && std::isfinite(c.k2);
}

// radial_bearing, model.rs:179-259
static bool radial_bearing(Radial c, const double m[2], double out[3]) {
    if (!radial_valid(c) || !std::isfinite(m[0]) || !std::isfinite(m[1])) return false;
    double dist[2] = {m[0] / c.focal, m[1] / c.focal};
    double dr = std::hypot(dist[0], dist[1]);
    if (!std::isfinite(dr)) return false;
    if (dr == 0.0) {
        out[0] = 0.0;
        out[1] = 0.0;
        out[2] = 1.0;
        return true;
    }
    double turn2;
    bool has_turn = first_positive_radial_turn(c, &turn2);
    double upper = has_turn ? std::sqrt(turn2) : fmax_(dr, 1.0);
    if (has_turn) {
        double maximum = radial_radius(c, upper);
        double tol = 64.0 * EPS * (1.0 + std::fabs(maximum));
        if (!std::isfinite(maximum) || dr >= maximum - tol) return false;
    } else {
        for (;;) {
            double projected = radial_radius(c, upper);
            if (!std::isfinite(projected)) return false;
            if (projected >= dr) break;
            upper *= 2.0;
            if (!std::isfinite(upper)) return false;
        }
    }
    double tol = 1.0e-14 * (1.0 + dr);
    double lower = 0.0;
    double radius = fmin_(dr, 0.5 * upper);
    for (int it = 0; it < 48; ++it) {
        double value = radial_radius(c, radius) - dr;
        if (std::fabs(value) <= tol) {
            double s = radius / dr;
            double b[3] = {dist[0] * s, dist[1] * s, 1.0};
            double nrm = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            if (!(std::isfinite(nrm) && nrm > MINPOS)) return false;
            for (int k = 0; k < 3; ++k) out[k] = b[k] / nrm;
            return true;
        }
        if (value < 0.0)
            lower = radius;
        else
            upper = radius;
        double rr = radius * radius;
        double deriv = 1.0 + 3.0 * c.k1 * rr + 5.0 * c.k2 * rr * rr;
        double newton = radius - value / deriv;
        radius = (std::isfinite(deriv) && deriv > 0.0 && newton > lower && newton < upper)
                     ? newton
                     : 0.5 * (lower + upper);
    }
    return false;
}

// ---------------------------------------------------------------------------------------
// Robust losses: numerics/robust.rs:8-44, numerics/refine/api.rs:351-386
// ---------------------------------------------------------------------------------------
enum LossKind { L2 = 0, HUBER = 1, CAUCHY = 2 };
struct Loss {
    int kind;
    double scale;  // Huber delta or Cauchy kernel scale
};
static double loss_cost(Loss l, double r) {
    switch (l.kind) {
        case HUBER: {
            double m = std::fabs(r);
            return m <= l.scale ? 0.5 * r * r : l.scale * (m - 0.5 * l.scale);
        }
        case CAUCHY: {
            double x2 = (r / l.scale) * (r / l.scale);
            return std::log1p(0.5 * x2);
        }
        default:
            return 0.5 * r * r;
    }
}
static double loss_sqrt_weight(Loss l, double r) {
    switch (l.kind) {
        case HUBER: {
            double m = std::fabs(r);
            return (m <= l.scale || m == 0.0) ? 1.0 : std::sqrt(l.scale / m);
        }
        case CAUCHY: {
            double x2 = (r / l.scale) * (r / l.scale);
            return std::sqrt(1.0 / (l.scale * l.scale * (0.5 * x2 + 1.0)));
        }
        default:
            return 1.0;
    }
}

// ---------------------------------------------------------------------------------------
// Manifold helpers: numerics/manifold.rs, numerics/manifold/so3.rs
// ---------------------------------------------------------------------------------------
static Mat3 mat_mul(const Mat3& a, const Mat3& b) {  // manifold.rs:53-63
    Mat3 c{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += a[i][k] * b[k][j];
            c[i][j] = v;
        }
    return c;
}
static Mat3 skew(const double v[3]) {  // manifold.rs:47-49
    Mat3 m{};
    m[0] = {0.0, -v[2], v[1]};
    m[1] = {v[2], 0.0, -v[0]};
    m[2] = {-v[1], v[0], 0.0};
    return m;
}
// matrix_polynomial, so3.rs:420-423: I*identity + K*linear + (K*K)*quadratic
static Mat3 matrix_polynomial(const double v[3], double id, double lin, double quad) {
    Mat3 K = skew(v);
    Mat3 K2 = mat_mul(K, K);
    Mat3 out{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double e = (i == j) ? 1.0 : 0.0;
            out[i][j] = (e * id + K[i][j] * lin) + K2[i][j] * quad;
        }
    return out;
}
// exp_so3, so3.rs:254-265
static Mat3 exp_so3(const double v[3]) {
    double th2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double th = std::sqrt(th2);
    if (th < 1e-10) return matrix_polynomial(v, 1.0, 1.0, 0.5);
    double axis[3] = {v[0] / th, v[1] / th, v[2] / th};
    double s = std::sin(th), c = std::cos(th);
    return matrix_polynomial(axis, 1.0, s, 1.0 - c);
}
static double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void normalize3(const double v[3], double o[3]) {
    double n = norm3(v);
    for (int k = 0; k < 3; ++k) o[k] = v[k] / n;
}
// s2_tangent_basis, manifold.rs:129-141
static void s2_tangent_basis(const double t[3], double b0[3], double b1[3]) {
    double tv[3];
    normalize3(t, tv);
    double ref[3] = {0, 0, 0};
    if (std::fabs(tv[0]) <= std::fabs(tv[1]) && std::fabs(tv[0]) <= std::fabs(tv[2]))
        ref[0] = 1.0;
    else if (std::fabs(tv[1]) <= std::fabs(tv[2]))
        ref[1] = 1.0;
    else
        ref[2] = 1.0;
    double c0[3], c1[3];
    cross3(tv, ref, c0);
    normalize3(c0, b0);
    cross3(tv, b0, c1);
    normalize3(c1, b1);
}
// retract_s2, manifold.rs:144-158
static void retract_s2(const double t[3], const double b[2][3], const double d[2], double out[3]) {
    double v[3] = {b[0][0] * d[0] + b[1][0] * d[1], b[0][1] * d[0] + b[1][1] * d[1],
                   b[0][2] * d[0] + b[1][2] * d[1]};
    double th = norm3(v);
    double tv[3];
    normalize3(t, tv);
    if (th < 1e-10) {
        double s[3] = {tv[0] + v[0], tv[1] + v[1], tv[2] + v[2]};
        normalize3(s, out);
    } else {
        double c = std::cos(th), sn = std::sin(th) / th;
        for (int k = 0; k < 3; ++k) out[k] = tv[k] * c + v[k] * sn;
    }
}

// ---------------------------------------------------------------------------------------
// Point block factor: kernel/point_block.rs:85-203
// ---------------------------------------------------------------------------------------
struct PointFactor {
    double inv[6];  // packed lower [i00, i10, i11, i20, i21, i22]
    void solve(const double x[3], double o[3]) const {
        o[0] = inv[0] * x[0] + inv[1] * x[1] + inv[3] * x[2];
        o[1] = inv[1] * x[0] + inv[2] * x[1] + inv[4] * x[2];
        o[2] = inv[3] * x[0] + inv[4] * x[1] + inv[5] * x[2];
    }
};
static bool damping_diagonal(const Mat3& m, double out[3]) {  // point_block.rs:99-110
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale = fmax_(scale, m[i][i]);
    if (!(std::isfinite(scale) && scale > 0.0)) return false;
    double floor = 64.0 * EPS * scale;
    for (int i = 0; i < 3; ++i) out[i] = fmax_(m[i][i], floor);
    return true;
}
static bool damped_matrix(const Mat3& m, double lambda, Mat3& out) {  // :90-97
    double dd[3];
    if (!damping_diagonal(m, dd)) return false;
    out = m;
    for (int i = 0; i < 3; ++i) out[i][i] += lambda * dd[i];
    return true;
}
static bool factor_damped(const Mat3& dm, PointFactor& f) {  // :112-157
    double ds = 0.0;
    for (int i = 0; i < 3; ++i) ds = fmax_(ds, dm[i][i]);
    double floor = 64.0 * EPS * ds;
    double scale[3];
    for (int i = 0; i < 3; ++i) scale[i] = std::sqrt(fmax_(dm[i][i], floor));
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c <= r; ++c) {
            double v = 0.5 * (dm[r][c] + dm[c][r]) / (scale[r] * scale[c]);
            for (int k = 0; k < c; ++k) v -= L[r][k] * L[c][k];
            if (r == c) {
                if (!(std::isfinite(v) && v > 64.0 * EPS)) return false;
                L[r][c] = std::sqrt(v);
            } else {
                L[r][c] = v / L[c][c];
            }
        }
    double m00 = 1.0 / L[0][0], m11 = 1.0 / L[1][1], m22 = 1.0 / L[2][2];
    double m10 = -L[1][0] * m00 * m11;
    double m21 = -L[2][1] * m11 * m22;
    double m20 = -(L[2][0] * m00 + L[2][1] * m10) * m22;
    double d0 = 1.0 / scale[0], d1 = 1.0 / scale[1], d2 = 1.0 / scale[2];
    f.inv[0] = (m00 * m00 + m10 * m10 + m20 * m20) * (d0 * d0);
    f.inv[1] = (m11 * m10 + m21 * m20) * (d1 * d0);
    f.inv[2] = (m11 * m11 + m21 * m21) * (d1 * d1);
    f.inv[3] = (m22 * m20) * (d2 * d0);
    f.inv[4] = (m22 * m21) * (d2 * d1);
    f.inv[5] = (m22 * m22) * (d2 * d2);
    return true;
}
static bool point_factor_new(const Mat3& m, double lambda, PointFactor& f) {  // :85-88
    Mat3 dm;
    if (!damped_matrix(m, lambda, dm)) return false;
    return factor_damped(dm, f);
}
static double point_rcond(const PointFactor& f, const Mat3& m) {  // :170-183
    Mat3 inv{};
    for (int c = 0; c < 3; ++c) {
        double e[3] = {0, 0, 0};
        e[c] = 1.0;
        double o[3];
        f.solve(e, o);
        for (int r = 0; r < 3; ++r) inv[r][c] = o[r];  // transpose of the column list
    }
    auto infn = [](const Mat3& a) {
        double best = 0.0;
        for (int r = 0; r < 3; ++r) {
            double s = 0.0;
            for (int c = 0; c < 3; ++c) s += std::fabs(a[r][c]);
            best = fmax_(best, s);
        }
        return best;
    };
    double den = infn(m) * infn(inv);
    return (std::isfinite(den) && den > 0.0) ? 1.0 / den : 0.0;
}
constexpr double POINT_BLOCK_RCOND_THRESHOLD = 64.0 * EPS * 3.0;  // point_block.rs:10

// ---------------------------------------------------------------------------------------
// Track geometry: kernel/point_geometry.rs
// ---------------------------------------------------------------------------------------
constexpr double MIN_NEAR_FAR_RATIO = 1e-4;
constexpr double MIN_PARALLAX_SINE = 1.7e-5;
struct TrackGeometry {
    double near2, far2, scene2;
    double first[3] = {0.0, 0.0, 0.0};
    bool has_first = false, parallax = false, invalid = false;
    size_t count = 0;
    explicit TrackGeometry(double scene_length) {
        near2 = INF;
        far2 = scene_length * scene_length;
        scene2 = scene_length * scene_length;
    }
    void push(const double o[3]) {  // :56-82
        double sq = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
        if (!(std::isfinite(sq) && sq > 0.0)) {
            invalid = true;
            return;
        }
        count += 1;
        near2 = fmin_(near2, sq);
        far2 = fmax_(far2, sq);
        double inv = 1.0 / std::sqrt(sq);
        double ray[3] = {o[0] * inv, o[1] * inv, o[2] * inv};
        if (!has_first) {
            has_first = true;
            for (int k = 0; k < 3; ++k) first[k] = ray[k];
            return;
        }
        if (parallax) return;
        double dot = ray[0] * first[0] + ray[1] * first[1] + ray[2] * first[2];
        double c[3] = {ray[1] * first[2] - ray[2] * first[1], ray[2] * first[0] - ray[0] * first[2],
                       ray[0] * first[1] - ray[1] * first[0]};
        double s2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        parallax = dot < 0.0 || s2 >= MIN_PARALLAX_SINE * MIN_PARALLAX_SINE;
    }
    bool degenerate() const {  // :86-92
        return invalid || (count >= 2 && (!parallax || near2 < MIN_NEAR_FAR_RATIO * MIN_NEAR_FAR_RATIO * far2));
    }
    bool collapsed() const {  // :96-100
        return invalid || (count >= 2 && near2 < MIN_NEAR_FAR_RATIO * MIN_NEAR_FAR_RATIO * scene2);
    }
};
enum Admission { ADMIT_PROJECTION, ADMIT_COLLAPSE, ADMIT_GEOMETRY };

// ---------------------------------------------------------------------------------------
// Native problem, gauge state and layout
// ---------------------------------------------------------------------------------------
struct GaugeState {  // schur/layout.rs:79-125 (Free and FixedBaseline)
    int fixed_camera = -1;
    double baseline = 0.0, scene_scale = 1.0;
    double direction[3] = {0, 0, 0};
    double basis[2][3] = {{0, 0, 0}, {0, 0, 0}};
};

struct View {  // one state of a lowered problem
    size_t nc, ng, np, no, D;
    const double* R;   // nc*9
    const double* t;   // nc*3
    const double* gp;  // ng*3
    const uint32_t* cam_group;
    const double* X;  // np*3
    const uint32_t* obs_cam;
    const double* uv;
    const uint32_t* off;
    const uint8_t* point_fixed;
    const uint32_t* cam_off;
    const uint32_t* cam_dim;
    const uint32_t* grp_off;
    const uint32_t* grp_dim;
    Loss loss;
    GaugeState gauge;
    Radial group(size_t g) const { return Radial{gp[3 * g], gp[3 * g + 1], gp[3 * g + 2]}; }
};

static GaugeState gauge_state(int camera, const double* t, double scene_scale) {
    GaugeState g;
    g.scene_scale = scene_scale;
    if (camera >= 0) {
        const double* tc = t + 3 * camera;
        g.fixed_camera = camera;
        g.baseline = norm3(tc);
        for (int k = 0; k < 3; ++k) g.direction[k] = tc[k] / g.baseline;
        s2_tangent_basis(g.direction, g.basis[0], g.basis[1]);
    }
    return g;
}

static View view_from(const ba_device_problem* p) {
    View v;
    v.nc = p->n_cameras;
    v.ng = p->n_groups;
    v.np = p->n_points;
    v.no = p->n_observations;
    v.D = p->dimension;
    v.R = p->camera_rotation;
    v.t = p->camera_translation;
    v.gp = p->group_params;
    v.cam_group = p->camera_group;
    v.X = p->point_xyz;
    v.obs_cam = p->obs_camera;
    v.uv = p->obs_uv;
    v.off = p->point_offsets;
    v.point_fixed = p->point_fixed;
    v.cam_off = p->camera_offset;
    v.cam_dim = p->camera_dim;
    v.grp_off = p->group_offset;
    v.grp_dim = p->group_dim;
    v.loss = Loss{p->loss, p->loss_scale};
    v.gauge.fixed_camera = p->gauge_camera;
    v.gauge.scene_scale = p->scene_scale;
    v.gauge.baseline = p->baseline;
    for (int k = 0; k < 3; ++k) {
        v.gauge.basis[0][k] = p->basis[k];
        v.gauge.basis[1][k] = p->basis[3 + k];
    }
    return v;
}

static void rotate(const double* R, const double X[3], double o[3]) {  // multi_view/model.rs:695-702
    o[0] = R[0] * X[0] + R[1] * X[1] + R[2] * X[2];
    o[1] = R[3] * X[0] + R[4] * X[1] + R[5] * X[2];
    o[2] = R[6] * X[0] + R[7] * X[1] + R[8] * X[2];
}

// pose_motion, schur/point.rs:730-762 (Free / FixedBaseline forms)
static void pose_motion(const double rot[3], size_t camera, const GaugeState& g, double m[3][6]) {
    double base[3][6] = {{0.0, rot[2], -rot[1], 0, 0, 0},
                         {-rot[2], 0.0, rot[0], 0, 0, 0},
                         {rot[1], -rot[0], 0.0, 0, 0, 0}};
    std::memcpy(m, base, sizeof(base));
    if (g.fixed_camera == (int)camera) {
        for (int c = 0; c < 3; ++c) {
            m[c][3] = g.baseline * g.basis[0][c];
            m[c][4] = g.baseline * g.basis[1][c];
        }
    } else {
        for (int c = 0; c < 3; ++c) m[c][3 + c] = g.scene_scale;
    }
}

// Per-observation linearisation: the first half of accumulate_observation
// (schur/point.rs:260-335) with the radial model.
struct ObsLin {
    double res[2];
    double pj[2][3];  // scene-scaled point Jacobian J_p R s (unweighted)
    double cj[2][9];  // camera Jacobian, pose run then intrinsic run (unweighted)
    double sw;        // robust sqrt weight
    double cost, sq;  // robust cost, squared residual norm
    int pose_dim, intr_dim;
    size_t pose_off, intr_off;
};
static bool linearize_obs(const View& v, size_t obs, const double X[3], ObsLin& o) {
    size_t cam = v.obs_cam[obs];
    const double* R = v.R + 9 * cam;
    double rot[3];
    rotate(R, X, rot);
    double cp[3] = {rot[0] + v.t[3 * cam], rot[1] + v.t[3 * cam + 1], rot[2] + v.t[3 * cam + 2]};
    double J[2][3], Ji[2][3];
    size_t grp = v.cam_group[cam];
    if (!radial_linearization(v.group(grp), v.uv + 2 * obs, cp, o.res, J, Ji)) return false;
    double rn = std::sqrt(o.res[0] * o.res[0] + o.res[1] * o.res[1]);
    o.sw = 1.0 * loss_sqrt_weight(v.loss, rn * 1.0);
    o.cost = loss_cost(v.loss, rn * 1.0);
    o.sq = rn * rn;
    // multiply_2x3_3x3 then scale_2x3 (point.rs:813-830)
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += J[r][k] * R[3 * k + c];
            o.pj[r][c] = s * v.gauge.scene_scale;
        }
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 9; ++c) o.cj[r][c] = 0.0;
    o.pose_dim = (int)v.cam_dim[cam];
    o.pose_off = v.cam_off[cam];
    o.intr_dim = (int)v.grp_dim[grp];
    o.intr_off = v.grp_off[grp];
    if (o.pose_dim > 0) {
        double m[3][6];
        pose_motion(rot, cam, v.gauge, m);
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < o.pose_dim; ++c) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += J[r][k] * m[k][c];
                o.cj[r][c] = s;
            }
    }
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < o.intr_dim; ++c) o.cj[r][o.pose_dim + c] = Ji[r][c];
    return true;
}

// ---------------------------------------------------------------------------------------
// Assembly: schur/assembly.rs:832-1069 and schur/point.rs:176-448
// ---------------------------------------------------------------------------------------
// Reduced-system storage: dense lower triangle when the Schur complement is materialised,
// otherwise the retained Hessian's blocks and pose-intrinsic edges
// (sparse/self_calibration_retained.rs:78-205).
struct HStore {
    bool dense = true;
    size_t D = 0, nc = 0;
    std::vector<double> full;              // D*D, entries (r, c) with r >= c
    std::vector<uint32_t> pblock, plocal;  // parameter -> block, local index
    std::vector<std::vector<double>> blocks;  // dim*dim per block
    std::vector<std::vector<double>> edges;   // per camera: pose_dim * 3 (pose_local * 3 + intrinsic_local)
    std::vector<uint32_t> bdim, cam_group;
    bool valid = true;
    void init_dense(size_t d) {
        dense = true;
        D = d;
        full.assign(d * d, 0.0);
    }
    void init_blocks(const uint32_t* cam_off, const uint32_t* cam_dim, const uint32_t* grp_off,
                     const uint32_t* grp_dim, const uint32_t* cgroup, size_t ncams, size_t ngroups, size_t d) {
        dense = false;
        D = d;
        nc = ncams;
        pblock.assign(d, 0);
        plocal.assign(d, 0);
        bdim.assign(ncams + ngroups, 0);
        blocks.assign(ncams + ngroups, {});
        edges.assign(ncams, {});
        cam_group.assign(cgroup, cgroup + ncams);
        for (size_t c = 0; c < ncams; ++c) {
            bdim[c] = cam_dim[c];
            for (uint32_t k = 0; k < cam_dim[c]; ++k) {
                pblock[cam_off[c] + k] = (uint32_t)c;
                plocal[cam_off[c] + k] = k;
            }
            blocks[c].assign(cam_dim[c] * cam_dim[c], 0.0);
            edges[c].assign(cam_dim[c] * 3, 0.0);
        }
        for (size_t g = 0; g < ngroups; ++g) {
            bdim[ncams + g] = grp_dim[g];
            for (uint32_t k = 0; k < grp_dim[g]; ++k) {
                pblock[grp_off[g] + k] = (uint32_t)(ncams + g);
                plocal[grp_off[g] + k] = k;
            }
            blocks[ncams + g].assign(grp_dim[g] * grp_dim[g], 0.0);
        }
    }
    // add `v` at (r, c) of the symmetric matrix, r >= c in storage
    void add(size_t r, size_t c, double v) {
        if (r < c) std::swap(r, c);
        if (dense) {
            full[r * D + c] += v;
            return;
        }
        uint32_t br = pblock[r], bc = pblock[c];
        if (br == bc) {
            uint32_t n = bdim[br];
            blocks[br][plocal[r] * n + plocal[c]] += v;
            return;
        }
        // r is the intrinsic parameter (groups follow poses), c the pose parameter
        if (br >= nc && bc < nc && cam_group[bc] == br - nc) {
            edges[bc][plocal[c] * 3 + plocal[r]] += v;
            return;
        }
        valid = false;
    }
    double get(size_t r, size_t c) const {
        if (r < c) std::swap(r, c);
        if (dense) return full[r * D + c];
        uint32_t br = pblock[r], bc = pblock[c];
        if (br == bc) return blocks[br][plocal[r] * bdim[br] + plocal[c]];
        if (br >= nc && bc < nc && cam_group[bc] == br - nc) return edges[bc][plocal[c] * 3 + plocal[r]];
        return 0.0;
    }
    // this += o, entry by entry (the same layout)
    void add_from(const HStore& o) {
        valid = valid && o.valid;
        for (size_t i = 0; i < full.size(); ++i) full[i] += o.full[i];
        for (size_t b = 0; b < blocks.size(); ++b)
            for (size_t k = 0; k < blocks[b].size(); ++k) blocks[b][k] += o.blocks[b][k];
        for (size_t c = 0; c < edges.size(); ++c)
            for (size_t k = 0; k < edges[c].size(); ++k) edges[c][k] += o.edges[c][k];
    }
    size_t bytes() const {
        size_t n = full.size();
        for (auto& b : blocks) n += b.size();
        for (auto& e : edges) n += e.size();
        return 8 * n;
    }
    bool finite() const {
        if (!valid) return false;
        if (dense) {
            for (size_t r = 0; r < D; ++r)
                for (size_t c = 0; c <= r; ++c)
                    if (!std::isfinite(full[r * D + c])) return false;
            return true;
        }
        for (auto& b : blocks)
            for (double v : b)
                if (!std::isfinite(v)) return false;
        for (auto& e : edges)
            for (double v : e)
                if (!std::isfinite(v)) return false;
        return true;
    }
};

struct Assembly {
    bool valid = true;
    size_t weak = 0;
    double min_rcond = 1.0, sgin = 0.0, pmr = 0.0, cost = 0.0, sq = 0.0;
    double min_diag = 0.0, max_diag = 0.0;
    size_t clamp = 0;
    HStore H;                         // S when materialised, else the retained U blocks
    std::vector<double> g, fg, diag;  // reduced gradient, full gradient, damping diagonal
    // point caches (the Cached back-substitution and retained PCG)
    std::vector<uint8_t> cache_valid;
    std::vector<double> inv_grad, damp_diag, damped_packed, inv_packed;  // 3,3,6,6 per point
    std::vector<double> inv_cross;                                       // 3 per cross entry
};

static inline void add_sym(HStore& H, size_t /*D*/, size_t r, size_t c, double v) { H.add(r, c, v); }

static double scaled_gradient_inf_norm(const double g[3], const double d[3]) {  // :1081-1096
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale = fmax_(scale, d[i]);
    if (!(std::isfinite(scale) && scale > 0.0))
        return (g[0] == 0.0 && g[1] == 0.0 && g[2] == 0.0) ? 0.0 : INF;
    double floor = 64.0 * EPS * scale;
    double best = 0.0;
    for (int i = 0; i < 3; ++i) best = fmax_(best, std::fabs(g[i]) / std::sqrt(fmax_(d[i], floor)));
    return best;
}

static size_t clamp_block(double* d, size_t n) {  // :744-766
    if (n == 0) return 0;
    double scale = 0.0;
    for (size_t i = 0; i < n; ++i) scale = fmax_(scale, d[i]);
    double floor = scale > 0.0 ? 64.0 * EPS * scale : 1.0;
    size_t clamped = 0;
    for (size_t i = 0; i < n; ++i)
        if (d[i] < floor) {
            d[i] = floor;
            ++clamped;
        }
    return clamped;
}

// ---------------------------------------------------------------------------------------
// Deterministic point shards (kernel/multi_view/sparse/retained/parallel.rs:9-31, 94-122).
// A pass over points runs in contiguous point shards balanced by observations.  Their number
// depends on the problem alone (the observation count, and the private accumulator bytes each
// shard needs), never on the thread count: each shard sums in its own storage in point order,
// and the shards are added in shard order, so any number of OpenMP threads gives the same bits.
// Below 32,768 observations there is one shard: the serial sums.
// ---------------------------------------------------------------------------------------
static std::vector<size_t> point_shards(const uint32_t* off, size_t np, size_t worker_bytes) {
    const size_t no = np ? off[np] : 0;
    size_t want = 1;
    if (no >= 32768) {
        want = std::min<size_t>(std::max<size_t>((no + 8191) / 8192, 1), 32);
        const size_t mem = worker_bytes ? ((size_t)128 << 20) / worker_bytes : want;
        want = std::max<size_t>(1, std::min(want, mem));
    }
    std::vector<size_t> r{0};
    if (np == 0) return r;
    const size_t workers = std::min(want, np);
    size_t previous = 0;
    for (size_t part = 1; part < workers; ++part) {
        const size_t lo = previous + 1, hi = np - (workers - part);
        size_t boundary;
        if (no == 0) {
            boundary = np * part / workers;
        } else {
            const size_t target = (size_t)((unsigned __int128)no * part / workers);
            boundary = (size_t)(std::lower_bound(off, off + np + 1, (uint32_t)std::min<size_t>(target, UINT32_MAX)) - off);
        }
        boundary = std::min(std::max(boundary, lo), hi);
        r.push_back(boundary);
        previous = boundary;
    }
    r.push_back(np);
    return r;
}

// assemble_with_hessian_problem_impl + finalize_assembly; `materialize` forms S, else U.
static Assembly assemble(const View& v, double lambda, bool materialize, const uint32_t* cross_off) {
    const size_t D = v.D;
    Assembly a;
    if (materialize)
        a.H.init_dense(D);
    else
        a.H.init_blocks(v.cam_off, v.cam_dim, v.grp_off, v.grp_dim, v.cam_group, v.nc, v.ng, D);
    a.g.assign(D, 0.0);
    a.fg.assign(D, 0.0);
    a.diag.assign(D, 0.0);
    a.cache_valid.assign(v.np, 0);
    a.inv_grad.assign(3 * v.np, 0.0);
    a.damp_diag.assign(3 * v.np, 0.0);
    a.damped_packed.assign(6 * v.np, 0.0);
    a.inv_packed.assign(6 * v.np, 0.0);
    if (cross_off) a.inv_cross.assign(3 * (size_t)cross_off[v.np], 0.0);
    // per-shard sums: the Hessian store, the three D-vectors and the scalars
    struct Acc {
        HStore H;
        std::vector<double> g, fg, diag;
        double cost = 0.0, sq = 0.0, pmr = 0.0, sgin = 0.0, min_rcond = 1.0;
        size_t weak = 0;
        bool valid = true;
    };
    const std::vector<size_t> shards = point_shards(v.off, v.np, a.H.bytes() + 6 * 8 * D);
    const size_t nshards = shards.size() - 1;
    std::vector<Acc> accs(nshards);
    for (size_t sh = 0; sh < nshards; ++sh) {
        accs[sh].H = sh == 0 ? std::move(a.H) : HStore{};
        if (sh > 0) {
            if (materialize)
                accs[sh].H.init_dense(D);
            else
                accs[sh].H.init_blocks(v.cam_off, v.cam_dim, v.grp_off, v.grp_dim, v.cam_group, v.nc, v.ng, D);
        }
        accs[sh].g.assign(D, 0.0);
        accs[sh].fg.assign(D, 0.0);
        accs[sh].diag.assign(D, 0.0);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t sh = 0; sh < nshards; ++sh) {
    Acc& acc = accs[sh];
    std::vector<double> cross(3 * D, 0.0);
    std::vector<uint8_t> marks(D, 0);
    std::vector<size_t> active;
    active.reserve(D);
    for (size_t p = shards[sh]; p < shards[sh + 1]; ++p) {
        Mat3 H{};
        double gp[3] = {0, 0, 0}, pcost = 0.0, psq = 0.0;
        bool ok = true;
        active.clear();
        for (uint32_t obs = v.off[p]; obs < v.off[p + 1]; ++obs) {
            ObsLin o;
            if (!linearize_obs(v, obs, v.X + 3 * p, o)) {
                ok = false;
                break;
            }
            double sw = o.sw;
            pcost += o.cost;
            psq += o.sq;
            // accumulate_point, point.rs:357-372
            for (int r = 0; r < 2; ++r) {
                double wr = sw * o.res[r];
                double wp[3] = {sw * o.pj[r][0], sw * o.pj[r][1], sw * o.pj[r][2]};
                for (int i = 0; i < 3; ++i) {
                    gp[i] += wp[i] * wr;
                    for (int j = 0; j < 3; ++j) H[i][j] += wp[i] * wp[j];
                }
            }
            if (o.pose_dim == 0 && o.intr_dim == 0) continue;
            // ActiveParameterTracker::register, layout.rs:238-256
            size_t runs[2][2] = {{o.pose_off, (size_t)o.pose_dim}, {o.intr_off, (size_t)o.intr_dim}};
            for (auto& run : runs) {
                if (run[1] == 0) continue;
                if (marks[run[0]]) continue;
                marks[run[0]] = 1;
                for (size_t k = 0; k < run[1]; ++k) {
                    active.push_back(run[0] + k);
                    for (int c = 0; c < 3; ++c) cross[3 * (run[0] + k) + c] = 0.0;
                }
            }
            // accumulate_camera, point.rs:380-448
            int cd = o.pose_dim + o.intr_dim;
            double first[9], second[9];
            for (int k = 0; k < cd; ++k) {
                first[k] = sw * o.cj[0][k];
                second[k] = sw * o.cj[1][k];
            }
            double pf[3], ps[3];
            for (int c = 0; c < 3; ++c) {
                pf[c] = sw * o.pj[0][c];
                ps[c] = sw * o.pj[1][c];
            }
            size_t param[9];
            for (int k = 0; k < o.pose_dim; ++k) param[k] = o.pose_off + k;
            for (int k = 0; k < o.intr_dim; ++k) param[o.pose_dim + k] = o.intr_off + k;
            for (int k = 0; k < cd; ++k) {
                double* row = &cross[3 * param[k]];
                for (int c = 0; c < 3; ++c) {
                    row[c] += pf[c] * first[k];
                    row[c] += ps[c] * second[k];
                }
            }
            double wres[2] = {sw * o.res[0], sw * o.res[1]};
            for (int k = 0; k < cd; ++k) {
                size_t q = param[k];
                acc.g[q] += first[k] * wres[0];
                acc.g[q] += second[k] * wres[1];
                acc.fg[q] += first[k] * wres[0];
                acc.fg[q] += second[k] * wres[1];
                acc.diag[q] += first[k] * first[k];
                acc.diag[q] += second[k] * second[k];
            }
            // Gram: intrinsic rows against the pose run, then each run's lower triangle
            int run_begin[2] = {0, o.pose_dim}, run_end[2] = {o.pose_dim, cd};
            for (int ri = 0; ri < 2; ++ri) {
                if (run_end[ri] == run_begin[ri]) continue;
                for (int ci = 0; ci < ri; ++ci) {
                    if (run_end[ci] == run_begin[ci]) continue;
                    for (int i = run_begin[ri]; i < run_end[ri]; ++i)
                        for (int j = run_begin[ci]; j < run_end[ci]; ++j) {
                            add_sym(acc.H, D, param[i], param[j], first[i] * first[j]);
                            add_sym(acc.H, D, param[i], param[j], second[i] * second[j]);
                        }
                }
                for (int i = run_begin[ri]; i < run_end[ri]; ++i)
                    for (int j = run_begin[ri]; j <= i; ++j) {
                        add_sym(acc.H, D, param[i], param[j], first[i] * first[j]);
                        add_sym(acc.H, D, param[i], param[j], second[i] * second[j]);
                    }
            }
        }
        for (size_t q : active) marks[q] = 0;
        if (!ok) {
            acc.valid = false;
            continue;
        }
        acc.cost += pcost;
        acc.sq += psq;
        if (v.point_fixed && v.point_fixed[p]) continue;
        double pd[3] = {H[0][0], H[1][1], H[2][2]};
        acc.sgin = fmax_(acc.sgin, scaled_gradient_inf_norm(gp, pd));
        PointFactor f;
        if (!point_factor_new(H, lambda, f)) {
            acc.valid = false;
            if (lambda == 0.0) {
                acc.weak += 1;
                acc.min_rcond = 0.0;
            }
            continue;
        }
        double dd[3] = {0.0, 0.0, 0.0};
        damping_diagonal(H, dd);
        for (int k = 0; k < 3; ++k) a.damp_diag[3 * p + k] = dd[k];
        Mat3 dm{};
        damped_matrix(H, lambda, dm);
        double packed[6] = {dm[0][0], dm[1][0], dm[1][1], dm[2][0], dm[2][1], dm[2][2]};
        for (int k = 0; k < 6; ++k) {
            a.damped_packed[6 * p + k] = packed[k];
            a.inv_packed[6 * p + k] = f.inv[k];
        }
        if (lambda == 0.0) {
            double rc = point_rcond(f, H);
            acc.min_rcond = fmin_(acc.min_rcond, rc);
            if (rc <= POINT_BLOCK_RCOND_THRESHOLD) acc.weak += 1;
        }
        double ig[3];
        f.solve(gp, ig);
        acc.pmr += 0.5 * (gp[0] * ig[0] + gp[1] * ig[1] + gp[2] * ig[2]);
        if (cross_off && cross_off[p + 1] - cross_off[p] != active.size()) {
            acc.valid = false;
            continue;
        }
        a.cache_valid[p] = 1;
        for (int k = 0; k < 3; ++k) a.inv_grad[3 * p + k] = ig[k];
        std::vector<std::array<double, 3>> inv_rows(active.size());
        for (size_t i = 0; i < active.size(); ++i) {
            const double* w = &cross[3 * active[i]];
            acc.g[active[i]] -= w[0] * ig[0] + w[1] * ig[1] + w[2] * ig[2];
            f.solve(w, inv_rows[i].data());
            if (cross_off)
                for (int c = 0; c < 3; ++c) a.inv_cross[3 * (cross_off[p] + i) + c] = inv_rows[i][c];
        }
        if (materialize) {
            for (size_t i = 0; i < active.size(); ++i)
                for (size_t j = 0; j <= i; ++j) {
                    const double* w = &cross[3 * active[j]];
                    double val = -(inv_rows[i][0] * w[0] + inv_rows[i][1] * w[1] + inv_rows[i][2] * w[2]);
                    add_sym(acc.H, D, active[i], active[j], val);
                }
        }
    }
    }
    // shard sums in shard order
    {
        Acc& first = accs[0];
        for (size_t sh = 1; sh < nshards; ++sh) {
            const Acc& o = accs[sh];
            first.H.add_from(o.H);
            for (size_t i = 0; i < D; ++i) {
                first.g[i] += o.g[i];
                first.fg[i] += o.fg[i];
                first.diag[i] += o.diag[i];
            }
            first.cost += o.cost;
            first.sq += o.sq;
            first.pmr += o.pmr;
            first.sgin = fmax_(first.sgin, o.sgin);
            first.min_rcond = fmin_(first.min_rcond, o.min_rcond);
            first.weak += o.weak;
            first.valid = first.valid && o.valid;
        }
        a.H = std::move(first.H);
        a.g = std::move(first.g);
        a.fg = std::move(first.fg);
        a.diag = std::move(first.diag);
        a.cost = first.cost;
        a.sq = first.sq;
        a.pmr = first.pmr;
        a.sgin = first.sgin;
        a.min_rcond = first.min_rcond;
        a.weak = first.weak;
        a.valid = first.valid;
    }
    // finalize_assembly, assembly.rs:691-727
    double scale = 0.0, mn = INF;
    for (size_t i = 0; i < D; ++i) {
        scale = fmax_(scale, a.diag[i]);
        mn = fmin_(mn, a.diag[i]);
    }
    a.min_diag = D == 0 ? 0.0 : mn;
    a.max_diag = scale;
    size_t clamped = 0;
    for (size_t c = 0; c < v.nc; ++c) clamped += clamp_block(&a.diag[v.cam_off[c]], v.cam_dim[c]);
    for (size_t g = 0; g < v.ng; ++g) clamped += clamp_block(&a.diag[v.grp_off[g]], v.grp_dim[g]);
    a.clamp = clamped;
    for (size_t i = 0; i < D; ++i) a.H.add(i, i, lambda * a.diag[i]);
    double prepared = 0.0;
    for (size_t i = 0; i < D; ++i) {
        double cv = a.diag[i], gv = a.fg[i], s;
        if (std::isfinite(cv) && cv > 0.0)
            s = std::fabs(gv) / std::sqrt(cv);
        else
            s = gv == 0.0 ? 0.0 : INF;
        prepared = fmax_(prepared, s);
    }
    a.sgin = fmax_(a.sgin, prepared);
    bool finite = a.H.finite();
    for (size_t i = 0; i < D; ++i)
        if (!std::isfinite(a.g[i])) finite = false;
    a.valid = a.valid && finite;
    return a;
}

// ---------------------------------------------------------------------------------------
// Dense Cholesky camera solve: kernel/multi_view/dense/linear_solver.rs:9-61
// ---------------------------------------------------------------------------------------
static double dot4(const double* l, const double* r, size_t n) {  // linear_solver.rs:67+
    double s[4] = {0, 0, 0, 0};
    size_t blocks = n / 4;
    for (size_t b = 0; b < blocks; ++b)
        for (int k = 0; k < 4; ++k) s[k] += l[4 * b + k] * r[4 * b + k];
    double tail = 0.0;
    for (size_t i = 4 * blocks; i < n; ++i) tail += l[i] * r[i];
    return (s[0] + s[1]) + (s[2] + s[3]) + tail;
}
// Factor the lower triangle of `A` (row-major n*n) in place and solve A x = rhs.
static bool dense_cholesky_solve(std::vector<double>& F, size_t n, const double* rhs, double* x) {
    double scale = 0.0;
    for (size_t i = 0; i < n; ++i) scale = fmax_(scale, std::fabs(F[i * n + i]));
    double tol = 64.0 * EPS * fmax_(scale, MINPOS);
    for (size_t c = 0; c < n; ++c) {
        // Row c's pivot first, then rows below: identical arithmetic to the row-by-row loop (each
        // row's entry depends only on rows c and r, so the rows below run in parallel).
        {
            double val = F[c * n + c] - dot4(&F[c * n], &F[c * n], c);
            if (!(std::isfinite(val) && val > tol)) return false;
            F[c * n + c] = std::sqrt(val);
        }
        const double pivot = F[c * n + c];
#pragma omp parallel for schedule(static) if (n - c > 256)
        for (size_t r = c + 1; r < n; ++r) F[r * n + c] = (F[r * n + c] - dot4(&F[r * n], &F[c * n], c)) / pivot;
    }
    for (size_t r = 0; r < n; ++r) {
        double v = rhs[r];
        for (size_t c = 0; c < r; ++c) v -= F[r * n + c] * x[c];
        x[r] = v / F[r * n + r];
    }
    for (size_t rr = n; rr-- > 0;) {
        double v = x[rr];
        for (size_t c = rr + 1; c < n; ++c) v -= F[c * n + rr] * x[c];
        x[rr] = v / F[rr * n + rr];
    }
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i])) return false;
    return true;
}

// ---------------------------------------------------------------------------------------
// Retained Schur PCG: sparse/self_calibration_retained.rs, retained/block.rs, pcg.rs
// ---------------------------------------------------------------------------------------
struct Packed {
    double v[45];  // lower triangle, index r(r+1)/2 + c (retained/block.rs:31-38)
    static size_t idx(size_t r, size_t c) {
        if (r < c) std::swap(r, c);
        return r * (r + 1) / 2 + c;
    }
    double get(size_t r, size_t c) const { return v[idx(r, c)]; }
};
struct BlockFactor {  // retained/block.rs:171-250
    double lower[45];
    double scale[9];
    bool factor(const Packed& m, size_t n) {
        if (n == 0 || n > 9) return false;
        std::memset(lower, 0, sizeof(lower));
        double ds = 0.0;
        for (size_t i = 0; i < n; ++i) ds = fmax_(ds, m.get(i, i));
        if (!(std::isfinite(ds) && ds > 0.0)) return false;
        double floor = 64.0 * EPS * ds;
        for (size_t i = 0; i < n; ++i) scale[i] = std::sqrt(fmax_(m.get(i, i), floor));
        for (size_t r = 0; r < n; ++r)
            for (size_t c = 0; c <= r; ++c) {
                double val = m.get(r, c) / (scale[r] * scale[c]);
                for (size_t k = 0; k < c; ++k) val -= lower[Packed::idx(r, k)] * lower[Packed::idx(c, k)];
                if (r == c) {
                    if (!(std::isfinite(val) && val > 64.0 * EPS)) return false;
                    lower[Packed::idx(r, c)] = std::sqrt(val);
                } else {
                    lower[Packed::idx(r, c)] = val / lower[Packed::idx(c, c)];
                }
            }
        return true;
    }
    void solve(const double* right, double* sol, size_t n) const {
        for (size_t r = 0; r < n; ++r) {
            double val = right[r] / scale[r];
            for (size_t c = 0; c < r; ++c) val -= lower[Packed::idx(r, c)] * sol[c];
            sol[r] = val / lower[Packed::idx(r, r)];
        }
        for (size_t rr = n; rr-- > 0;) {
            double val = sol[rr];
            for (size_t c = rr + 1; c < n; ++c) val -= lower[Packed::idx(c, rr)] * sol[c];
            sol[rr] = val / lower[Packed::idx(rr, rr)];
        }
        for (size_t i = 0; i < n; ++i) sol[i] /= scale[i];
    }
};

struct Retained {
    size_t D, nblocks, nc;
    std::vector<size_t> boff, bdim;   // blocks: cameras then groups
    std::vector<size_t> param_block;  // D
    std::vector<Packed> blocks;       // U blocks (damped)
    struct Edge {
        size_t pose_block, intr_block;
        double v[18];  // pose_local*3 + intrinsic_local
    };
    std::vector<Edge> edges;
    const Assembly* asmb;
    const uint32_t* cross_off;
    const uint32_t* params;
    size_t np;
    const uint8_t* point_fixed;
    std::vector<size_t> shards;  // point shards of the coupling (one private D-vector each)
    mutable std::vector<std::vector<double>> shard_out;

    // apply_base (operator.rs:563-594) then the fused point coupling (operator.rs:597-630)
    void apply(const double* in, double* out) const {
        std::fill(out, out + D, 0.0);
        for (size_t b = 0; b < nblocks; ++b) {
            size_t o = boff[b], n = bdim[b];
            for (size_t r = 0; r < n; ++r)
                for (size_t c = 0; c <= r; ++c) {
                    double val = 1.0 * blocks[b].get(r, c);
                    out[o + r] += val * in[o + c];
                    if (r != c) out[o + c] += val * in[o + r];
                }
        }
        for (const Edge& e : edges) {
            size_t po = boff[e.pose_block], io = boff[e.intr_block];
            size_t pd = bdim[e.pose_block], id = bdim[e.intr_block];
            for (size_t pl = 0; pl < pd; ++pl)
                for (size_t il = 0; il < id; ++il) {
                    double val = e.v[pl * 3 + il];
                    out[po + pl] += val * in[io + il];
                    out[io + il] += val * in[po + pl];
                }
        }
        // shard 0 adds into `out` after the base terms, shard s > 0 into its own vector; the
        // shard vectors are then added in shard order
        const size_t nshards = shards.size() - 1;
        if (shard_out.size() < nshards) shard_out.resize(nshards);
#pragma omp parallel for schedule(dynamic, 1)
        for (size_t sh = 0; sh < nshards; ++sh) {
            double* dst = out;
            if (sh > 0) {
                shard_out[sh].assign(D, 0.0);
                dst = shard_out[sh].data();
            }
            for (size_t p = shards[sh]; p < shards[sh + 1]; ++p) {
                uint32_t b = cross_off[p], en = cross_off[p + 1];
                double sum[3] = {0, 0, 0};
                for (uint32_t k = b; k < en; ++k) {
                    double val = in[params[k]];
                    const double* rc = &asmb->inv_cross[3 * k];
                    sum[0] += rc[0] * val;
                    sum[1] += rc[1] * val;
                    sum[2] += rc[2] * val;
                }
                const double* h = &asmb->damped_packed[6 * p];
                double w[3] = {h[0] * sum[0] + h[1] * sum[1] + h[3] * sum[2],
                               h[1] * sum[0] + h[2] * sum[1] + h[4] * sum[2],
                               h[3] * sum[0] + h[4] * sum[1] + h[5] * sum[2]};
                for (uint32_t k = b; k < en; ++k) {
                    const double* rc = &asmb->inv_cross[3 * k];
                    dst[params[k]] -= rc[0] * w[0] + rc[1] * w[1] + rc[2] * w[2];
                }
            }
        }
        for (size_t sh = 1; sh < nshards; ++sh)
            for (size_t i = 0; i < D; ++i) out[i] += shard_out[sh][i];
    }
};

static double vdot(const double* a, const double* b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

struct PcgResult {
    bool accepted;
    size_t terms;
    double rel;  // NaN when absent
};

// pcg::solve, pcg.rs:23-119, with solve_blocks (self_calibration_retained.rs:870-887)
template <class Op, class Pre>
static PcgResult pcg_solve(const double* grad, size_t D, size_t max_it, double tol, Op&& op, Pre&& pre,
                           std::vector<double>& x, std::vector<double>& product) {
    std::vector<double> r(D), z(D), d(D);
    x.assign(D, 0.0);
    product.assign(D, 0.0);
    for (size_t i = 0; i < D; ++i) r[i] = -grad[i];
    double n0 = std::sqrt(vdot(r.data(), r.data(), D));
    if (!std::isfinite(n0)) return {false, 0, NaN};
    if (n0 == 0.0) return {true, 0, 0.0};
    pre(r.data(), z.data());
    d = z;
    double rho = vdot(r.data(), z.data(), D);
    if (!(std::isfinite(rho) && rho > 0.0)) return {false, 0, NaN};
    auto recompute = [&]() {
        op(x.data(), product.data());
        for (size_t i = 0; i < D; ++i) r[i] = -grad[i] - product[i];
    };
    for (size_t it = 1; it <= max_it; ++it) {
        op(d.data(), product.data());
        double curv = vdot(d.data(), product.data(), D);
        if (!(std::isfinite(curv) && curv > 0.0)) return {false, it - 1, NaN};
        double alpha = rho / curv;
        if (!std::isfinite(alpha)) return {false, it - 1, NaN};
        for (size_t i = 0; i < D; ++i) {
            x[i] += alpha * d[i];
            r[i] -= alpha * product[i];
        }
        double ratio = std::sqrt(vdot(r.data(), r.data(), D)) / n0;
        if (!std::isfinite(ratio)) return {false, it, NaN};
        if (ratio <= tol) {
            recompute();
            double tr = std::sqrt(vdot(r.data(), r.data(), D)) / n0;
            if (std::isfinite(tr) && tr <= tol) return {true, it, tr};
            if (!std::isfinite(tr)) return {false, it, NaN};
            pre(r.data(), z.data());
            d = z;
            rho = vdot(r.data(), z.data(), D);
            if (!(std::isfinite(rho) && rho > 0.0)) return {false, it, tr};
            continue;
        }
        pre(r.data(), z.data());
        double next = vdot(r.data(), z.data(), D);
        if (!(std::isfinite(next) && next > 0.0)) return {false, it, NaN};
        double beta = next / rho;
        if (!std::isfinite(beta)) return {false, it, NaN};
        for (size_t i = 0; i < D; ++i) d[i] = z[i] + beta * d[i];
        rho = next;
    }
    recompute();
    double rel = std::sqrt(vdot(r.data(), r.data(), D)) / n0;
    return {false, max_it, std::isfinite(rel) ? rel : NaN};
}

// Build the retained operator from an unmaterialised assembly (U blocks and edges), then
// factor the block-Jacobi preconditioner (self_calibration_retained.rs:742-868).
static bool build_retained(const View& v, const Assembly& a, const uint32_t* cross_off,
                           const uint32_t* params, Retained& ret, std::vector<BlockFactor>& factors) {
    ret.D = v.D;
    ret.nc = v.nc;
    ret.nblocks = v.nc + v.ng;
    ret.boff.resize(ret.nblocks);
    ret.bdim.resize(ret.nblocks);
    for (size_t c = 0; c < v.nc; ++c) {
        ret.boff[c] = v.cam_off[c];
        ret.bdim[c] = v.cam_dim[c];
    }
    for (size_t g = 0; g < v.ng; ++g) {
        ret.boff[v.nc + g] = v.grp_off[g];
        ret.bdim[v.nc + g] = v.grp_dim[g];
    }
    ret.param_block.assign(v.D, 0);
    for (size_t b = 0; b < ret.nblocks; ++b)
        for (size_t k = 0; k < ret.bdim[b]; ++k) ret.param_block[ret.boff[b] + k] = b;
    ret.blocks.assign(ret.nblocks, Packed{});
    for (size_t b = 0; b < ret.nblocks; ++b) {
        std::memset(ret.blocks[b].v, 0, sizeof(ret.blocks[b].v));
        size_t o = ret.boff[b];
        for (size_t r = 0; r < ret.bdim[b]; ++r)
            for (size_t c = 0; c <= r; ++c) ret.blocks[b].v[Packed::idx(r, c)] = a.H.get(o + r, o + c);
    }
    ret.edges.clear();
    for (size_t c = 0; c < v.nc; ++c) {
        size_t g = v.cam_group[c];
        if (v.cam_dim[c] == 0 || v.grp_dim[g] == 0) continue;
        Retained::Edge e;
        e.pose_block = c;
        e.intr_block = v.nc + g;
        std::memset(e.v, 0, sizeof(e.v));
        for (size_t pl = 0; pl < v.cam_dim[c]; ++pl)
            for (size_t il = 0; il < v.grp_dim[g]; ++il) {
                size_t r = v.grp_off[g] + il, col = v.cam_off[c] + pl;  // intrinsic rows follow poses
                e.v[pl * 3 + il] = a.H.get(r, col);
            }
        ret.edges.push_back(e);
    }
    ret.asmb = &a;
    ret.cross_off = cross_off;
    ret.params = params;
    ret.np = v.np;
    ret.point_fixed = v.point_fixed;
    ret.shards = point_shards(v.off, v.np, v.D * sizeof(double));
    // accumulate_point_coupling_range (self_calibration_retained.rs:827-868)
    std::vector<Packed> coupling(ret.nblocks);
    for (auto& p : coupling) std::memset(p.v, 0, sizeof(p.v));
    for (size_t p = 0; p < v.np; ++p) {
        uint32_t b0 = cross_off[p], len = cross_off[p + 1] - cross_off[p];
        const double* h = &a.damped_packed[6 * p];
        size_t ls = 0;
        while (ls < len) {
            size_t blk = ret.param_block[params[b0 + ls]];
            size_t le = ls + 1;
            while (le < len && ret.param_block[params[b0 + le]] == blk) ++le;
            for (size_t r = ls; r < le; ++r) {
                size_t rl = params[b0 + r] - ret.boff[blk];
                const double* rc = &a.inv_cross[3 * (b0 + r)];
                for (size_t c = ls; c <= r; ++c) {
                    size_t cl = params[b0 + c] - ret.boff[blk];
                    const double* cc = &a.inv_cross[3 * (b0 + c)];
                    double w[3] = {h[0] * cc[0] + h[1] * cc[1] + h[3] * cc[2],
                                   h[1] * cc[0] + h[2] * cc[1] + h[4] * cc[2],
                                   h[3] * cc[0] + h[4] * cc[1] + h[5] * cc[2]};
                    double val = rc[0] * w[0] + rc[1] * w[1] + rc[2] * w[2];
                    coupling[blk].v[Packed::idx(rl, cl)] += val;
                }
            }
            ls = le;
        }
    }
    factors.assign(ret.nblocks, BlockFactor{});
    for (size_t b = 0; b < ret.nblocks; ++b) {
        if (ret.bdim[b] == 0) continue;
        Packed pre = ret.blocks[b];
        for (int k = 0; k < 45; ++k) pre.v[k] -= coupling[b].v[k];
        if (!factors[b].factor(pre, ret.bdim[b])) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Lowering: native/self_calibration.rs:47-197, native/gauge.rs, native/world_rebase.rs
// ---------------------------------------------------------------------------------------
struct Rebase {
    Mat3 Ra;
    Vec3 Ca;
    double scale;
};

struct Lowered {
    size_t n_input_points = 0;
    std::vector<uint32_t> kept_points;  // native point -> input point
    std::vector<uint8_t> quarantined;   // input point
    Rebase rebase;
    std::vector<uint32_t> cam_source;  // native camera -> input pose
    std::vector<double> R, t, gp, X, uv;
    std::vector<uint32_t> cam_group, obs_cam, obs_src, obs_point, off;
    std::vector<uint8_t> cam_fixed, point_fixed, grp_fixed;
    std::vector<uint32_t> grp_free;
    int gauge_camera = -1;
    // the calibrated (fixed-intrinsic) path of native/calibrated.rs; its anchor stays free when
    // survey controls fix the frame without a fixed observed pose (native/gauge.rs:153-162)
    bool calibrated = false;
    bool anchor_free = false;
    // layout
    size_t D = 0;
    std::vector<uint32_t> cam_off, cam_dim, grp_off, grp_dim, cross_off, params, col_ptr, row_idx;
    Loss loss{0, 0.0};
};

struct Err {
    int code = BA_OK;
    std::string msg;
};
static Err fail(int code, const std::string& m) { return Err{code, m}; }

static Vec3 mat_vec(const Mat3& m, const Vec3& v) {
    Vec3 o;
    for (int i = 0; i < 3; ++i) o[i] = (m[i][0] * v[0] + m[i][1] * v[1]) + m[i][2] * v[2];
    return o;
}
static Mat3 transpose(const Mat3& m) {
    Mat3 o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o[i][j] = m[j][i];
    return o;
}
static Mat3 pose_R(const ba_problem* p, size_t i) {
    Mat3 m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = p->pose_rotation[9 * i + 3 * r + c];
    return m;
}
static Vec3 pose_C(const ba_problem* p, size_t i) {
    return Vec3{p->pose_center[3 * i], p->pose_center[3 * i + 1], p->pose_center[3 * i + 2]};
}
static bool pose_fixed(const ba_problem* p, size_t i) {
    return !p->pose_var_rotation[i] && !p->pose_var_translation[i];
}
static double dist2(const Vec3& a, const Vec3& b) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}
static bool point_is_fixed(const ba_problem* p, size_t i) { return p->point_fixed && p->point_fixed[i]; }

// Problem checks of BundleProblemBuilder::build (problem.rs:492-576) and SolveConfig::validate.
static Err validate(const ba_problem* p) {
    if (p->n_poses == 0 || p->n_intrinsics == 0 || p->n_points == 0 || p->n_observations == 0)
        return fail(BA_ERR_INVALID_PROBLEM, "empty problem block");
    for (size_t i = 0; i < p->n_poses; ++i) {
        Mat3 R = pose_R(p, i);
        Vec3 C = pose_C(p, i);
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(C[k])) return fail(BA_ERR_INVALID_PROBLEM, "invalid pose");
        Mat3 RRt = mat_mul(R, transpose(R));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                if (!(std::fabs(RRt[r][c] - (r == c ? 1.0 : 0.0)) <= 9.5367431640625e-7))
                    return fail(BA_ERR_INVALID_PROBLEM, "invalid pose");
    }
    for (size_t i = 0; i < 3 * p->n_points; ++i)
        if (!std::isfinite(p->point_xyz[i])) return fail(BA_ERR_INVALID_PROBLEM, "non-finite point");
    for (size_t o = 0; o < p->n_observations; ++o) {
        if (p->obs_pose[o] >= p->n_poses) return fail(BA_ERR_INVALID_PROBLEM, "observation pose out of bounds");
        if (p->obs_point[o] >= p->n_points) return fail(BA_ERR_INVALID_PROBLEM, "observation point out of bounds");
        if (p->obs_intrinsic[o] >= p->n_intrinsics)
            return fail(BA_ERR_INVALID_PROBLEM, "intrinsic index out of bounds");
        if (!std::isfinite(p->obs_pixel[2 * o]) || !std::isfinite(p->obs_pixel[2 * o + 1]))
            return fail(BA_ERR_INVALID_PROBLEM, "non-finite observation");
    }
    if (p->robust_loss != BA_LOSS_SQUARED && p->residual_space != BA_RESIDUAL_PIXELS)
        return fail(BA_ERR_INVALID_PROBLEM, "loss unit mismatch");
    if (p->gauge == BA_GAUGE_FIXED_BASELINE) {
        if (p->gauge_first >= p->n_poses || p->gauge_second >= p->n_poses)
            return fail(BA_ERR_INVALID_PROBLEM, "gauge pose out of bounds");
        if (p->gauge_first == p->gauge_second) return fail(BA_ERR_INVALID_PROBLEM, "gauge uses one pose");
        if (!std::isfinite(p->gauge_distance) || p->gauge_distance <= 0.0)
            return fail(BA_ERR_INVALID_PROBLEM, "invalid baseline distance");
    }
    return Err{};
}
static Err validate_config(const ba_solve_config* c) {  // solver.rs:122-141
    if (c->max_iterations == 0) return fail(BA_ERR_INVALID_CONFIG, "max_iterations must be positive");
    const double v[3] = {c->function_tolerance, c->gradient_tolerance, c->parameter_tolerance};
    const char* names[3] = {"function_tolerance", "gradient_tolerance", "parameter_tolerance"};
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(v[i]) || v[i] <= 0.0) return fail(BA_ERR_INVALID_CONFIG, names[i]);
    return Err{};
}

// quarantine::degenerate_input_points, quarantine.rs:33-96
static std::vector<uint8_t> degenerate_input_points(const ba_problem* p) {
    std::vector<size_t> counts(p->n_points, 0);
    for (size_t o = 0; o < p->n_observations; ++o) counts[p->obs_point[o]] += 1;
    std::vector<uint8_t> weak(p->n_points, 0);
    for (size_t i = 0; i < p->n_points; ++i) weak[i] = counts[i] < 2;
    std::vector<double> offs(3 * p->n_observations), dist(p->n_observations);
    for (size_t o = 0; o < p->n_observations; ++o) {
        Vec3 C = pose_C(p, p->obs_pose[o]);
        const double* X = p->point_xyz + 3 * p->obs_point[o];
        for (int k = 0; k < 3; ++k) offs[3 * o + k] = X[k] - C[k];
        dist[o] = std::sqrt(offs[3 * o] * offs[3 * o] + offs[3 * o + 1] * offs[3 * o + 1] +
                            offs[3 * o + 2] * offs[3 * o + 2]);
    }
    double scene = 0.0;
    if (!dist.empty()) {
        size_t mid = dist.size() / 2;
        std::nth_element(dist.begin(), dist.begin() + mid, dist.end());
        scene = dist[mid];
    }
    std::vector<TrackGeometry> geom(p->n_points, TrackGeometry(scene));
    for (size_t o = 0; o < p->n_observations; ++o) geom[p->obs_point[o]].push(&offs[3 * o]);
    for (size_t i = 0; i < p->n_points; ++i)
        if (geom[i].degenerate()) weak[i] = 1;
    for (size_t i = 0; i < p->n_points; ++i)
        if (point_is_fixed(p, i)) weak[i] = 0;
    return weak;
}

static bool intrinsic_is_fixed(const ba_intrinsic& in) {
    return !in.var_focal && !in.var_principal_point && !in.var_radial && !in.var_tangential;
}

// quarantine: the input pass of solve_with_quarantine (off for the problems of the re-solves,
// whose degenerate inputs were held out before); hold: intrinsic groups held at their values
// (the unconstrained re-solve of native/self_calibration.rs:240-247), or NULL.
static Err lower(const ba_problem* p, Lowered& L, bool quarantine = true, const std::vector<uint8_t>* hold = nullptr) {
    // solve_with_quarantine input pass (native/mod.rs:271-290)
    L.n_input_points = p->n_points;
    L.quarantined = quarantine ? degenerate_input_points(p) : std::vector<uint8_t>(p->n_points, 0);
    size_t nq = 0;
    for (auto q : L.quarantined) nq += q;
    if (nq == p->n_points)
        return fail(BA_ERR_UNSUPPORTED,
                    "no point can be adjusted: every point is seen fewer than twice or without parallax");
    if (p->residual_space != BA_RESIDUAL_PIXELS)
        return fail(BA_ERR_UNSUPPORTED, "self-calibration requires pixel residuals");
    // the restricted problem keeps the observations of the kept points (problem.rs:686-693)
    std::vector<uint32_t> kobs;
    kobs.reserve(p->n_observations);
    for (size_t o = 0; o < p->n_observations; ++o)
        if (!L.quarantined[p->obs_point[o]]) kobs.push_back((uint32_t)o);
    bool free_intr = false;  // has_free_intrinsics, native/mod.rs:457-463
    for (uint32_t o : kobs)
        if (!intrinsic_is_fixed(p->intrinsics[p->obs_intrinsic[o]])) free_intr = true;
    L.calibrated = !free_intr;
    // single_pose_intrinsics, native/mod.rs:482-497
    std::vector<size_t> pose_intr(p->n_poses, SIZE_MAX);
    for (uint32_t o : kobs) {
        size_t& s = pose_intr[p->obs_pose[o]];
        if (s != SIZE_MAX && s != p->obs_intrinsic[o])
            return fail(BA_ERR_UNSUPPORTED, L.calibrated
                                                ? "calibrated pixel BA here requires one camera model per observed pose"
                                                : "native self-calibration requires one intrinsic group per observed pose");
        s = p->obs_intrinsic[o];
    }
    for (auto& s : pose_intr)
        if (s == SIZE_MAX) s = 0;
    std::vector<uint8_t> observed(p->n_poses, 0);
    for (uint32_t o : kobs) observed[p->obs_pose[o]] = 1;
    if (L.calibrated) {
        // PreparedCalibratedProblem::try_new (native/calibrated.rs:200-226): fixed pinhole and
        // Brown-Conrady cameras only; the lowering restated here covers the radial k1/k2 models
        // with one focal length (the radial projection of the self-calibrating path)
        for (uint32_t o : kobs) {
            const ba_intrinsic& in = p->intrinsics[p->obs_intrinsic[o]];
            if (in.model == BA_CAMERA_FISHEYE || in.model == BA_CAMERA_SPHERICAL)
                return fail(BA_ERR_UNSUPPORTED,
                            "pixel BA supports fixed pinhole and Brown-Conrady cameras; use angular residuals for "
                            "wide-angle cameras");
        }
        // can_use_two_view (native/two_view.rs:14-50): the reference's two-view solver
        size_t kept_points = 0;
        for (size_t i = 0; i < p->n_points; ++i) kept_points += !L.quarantined[i];
        bool any_fixed_point = false, pinholes = true, repeated = false;
        for (size_t i = 0; i < p->n_points; ++i) any_fixed_point |= !L.quarantined[i] && point_is_fixed(p, i);
        {
            std::vector<std::pair<uint32_t, uint32_t>> pp;
            for (uint32_t o : kobs) pp.emplace_back(p->obs_pose[o], p->obs_point[o]);
            std::sort(pp.begin(), pp.end());
            repeated = std::adjacent_find(pp.begin(), pp.end()) != pp.end();
        }
        for (size_t i = 0; i < p->n_poses; ++i) {
            const ba_intrinsic& in = p->intrinsics[pose_intr[i]];
            const bool distortion = in.model != BA_CAMERA_PINHOLE &&
                                    ((in.model >= BA_CAMERA_BROWN_K1 && in.k1 != 0.0) ||
                                     (in.model >= BA_CAMERA_BROWN_K1K2 && in.k2 != 0.0) ||
                                     (in.model >= BA_CAMERA_BROWN_K1K2K3 && in.k3 != 0.0) || in.p1 != 0.0 ||
                                     in.p2 != 0.0);
            pinholes &= in.model != BA_CAMERA_FISHEYE && in.model != BA_CAMERA_SPHERICAL && !distortion;
        }
        if (p->n_poses == 2 && !any_fixed_point && kobs.size() == 2 * kept_points && !repeated && pinholes) {
            long a = -1;
            if (p->gauge == BA_GAUGE_FIXED_BASELINE)
                a = (long)p->gauge_first;
            else
                for (size_t i = 0; i < 2 && a < 0; ++i)
                    if (pose_fixed(p, i)) a = (long)i;
            if (a >= 0 && pose_fixed(p, (size_t)a) && p->pose_var_rotation[1 - a] && p->pose_var_translation[1 - a])
                return fail(BA_ERR_UNSUPPORTED, "two-view refinement (native/two_view.rs) is not this adjuster's");
        }
        // pose masks: every observed camera full-pose or fixed (the 6-dimensional blocks)
        for (size_t i = 0; i < p->n_poses; ++i)
            if (observed[i] && !pose_fixed(p, i) && !(p->pose_var_rotation[i] && p->pose_var_translation[i]))
                return fail(BA_ERR_UNSUPPORTED, "calibrated pixel BA here requires full-pose or fixed camera blocks");
    } else {
        // common_pose_variables, native/gauge.rs:9-40
        for (size_t i = 0; i < p->n_poses; ++i) {
            if (!observed[i] || pose_fixed(p, i)) continue;
            if (!p->pose_var_translation[i])
                return fail(BA_ERR_UNSUPPORTED, "direct native lowering cannot optimize rotation without translation");
            if (!p->pose_var_rotation[i])
                return fail(BA_ERR_UNSUPPORTED, "self-calibration requires full-pose or fixed camera blocks");
        }
        // intrinsic model checks, native/self_calibration.rs:63-85, 483-512
        for (size_t g = 0; g < p->n_intrinsics; ++g) {
            const ba_intrinsic& in = p->intrinsics[g];
            if (in.var_radial && in.model == BA_CAMERA_BROWN_K1K2K3)
                return fail(BA_ERR_UNSUPPORTED, "native self-calibration refines at most two declared radial coefficients");
            if (!intrinsic_is_fixed(in) && (!in.var_focal || in.var_principal_point || in.var_tangential))
                return fail(BA_ERR_UNSUPPORTED,
                            "self-calibration refines the focal length and radial coefficients with a fixed principal point");
        }
    }
    bool fixed_point_observed = false;
    for (uint32_t o : kobs)
        if (point_is_fixed(p, p->obs_point[o])) fixed_point_observed = true;
    size_t anchor = 0;
    double scale = 1.0;
    int native_gauge = -1;  // input pose index of the FixedBaseline camera
    if (L.calibrated && p->gauge == BA_GAUGE_FREE) {
        // calibrated_gauge_plan (native/gauge.rs:123-229) for full-pose or fixed cameras: held
        // centres and known rotations are then exactly the fixed observed cameras
        size_t first_observed = SIZE_MAX;
        long fixed_anchor = -1;
        for (size_t i = 0; i < p->n_poses; ++i) {
            if (observed[i] && first_observed == SIZE_MAX) first_observed = i;
            if (observed[i] && pose_fixed(p, i) && fixed_anchor < 0) fixed_anchor = (long)i;
        }
        if (first_observed == SIZE_MAX) return fail(BA_ERR_INTERNAL, "checked bundle has no observed pose");
        anchor = fixed_anchor >= 0 ? (size_t)fixed_anchor : first_observed;
        const Vec3 Ca = pose_C(p, anchor);
        if (fixed_point_observed && fixed_anchor < 0) {
            // survey controls fix the frame: the anchor only sets the native frame
            L.anchor_free = true;
        } else {
            bool distinct_held = false;
            for (size_t i = 0; i < p->n_poses; ++i)
                if (observed[i] && pose_fixed(p, i) && dist2(pose_C(p, i), Ca) > 0.0) distinct_held = true;
            bool control_sets_scale = false;
            for (uint32_t o : kobs) {
                size_t q = p->obs_point[o];
                if (point_is_fixed(p, q)) {
                    Vec3 X{p->point_xyz[3 * q], p->point_xyz[3 * q + 1], p->point_xyz[3 * q + 2]};
                    if (dist2(X, Ca) > 0.0) control_sets_scale = true;
                }
            }
            if (!(distinct_held || control_sets_scale)) {
                long best = -1;
                double best_d = 0.0;
                for (size_t i = 0; i < p->n_poses; ++i) {
                    if (i == anchor || !observed[i] || pose_fixed(p, i)) continue;
                    double d = dist2(pose_C(p, i), Ca);
                    if (!std::isfinite(d)) continue;
                    if (best < 0 || d > best_d) {
                        best = (long)i;
                        best_d = d;
                    }
                }
                if (best < 0 || !(best_d > 0.0))
                    return fail(BA_ERR_UNSUPPORTED, "free-scale BA has no nonzero active baseline");
                native_gauge = (int)best;
            }
        }
    } else if (p->gauge == BA_GAUGE_FREE) {
        // gauge_plan, native/gauge.rs:231-330 (validate_anchor_policy first)
        if (fixed_point_observed) {
            bool any = false;
            for (size_t i = 0; i < p->n_poses; ++i)
                if (observed[i] && pose_fixed(p, i)) any = true;
            if (!any)
                return fail(BA_ERR_UNSUPPORTED,
                            "native BA with fixed world points requires an explicitly fixed observed pose or fixed-baseline anchor");
        }
        // free_scale_anchor, gauge.rs:57-68
        size_t first_observed = SIZE_MAX;
        bool found = false;
        for (size_t i = 0; i < p->n_poses; ++i) {
            if (observed[i] && pose_fixed(p, i)) {
                anchor = i;
                found = true;
                break;
            }
            if (observed[i] && first_observed == SIZE_MAX) first_observed = i;
        }
        if (!found) anchor = first_observed == SIZE_MAX ? 0 : first_observed;
        Vec3 Ca = pose_C(p, anchor);
        bool pose_sets = false, point_sets = false;
        for (size_t i = 0; i < p->n_poses; ++i)
            if (i != anchor && observed[i] && pose_fixed(p, i) && dist2(pose_C(p, i), Ca) > 0.0) pose_sets = true;
        for (uint32_t o : kobs) {
            size_t q = p->obs_point[o];
            if (point_is_fixed(p, q)) {
                Vec3 X{p->point_xyz[3 * q], p->point_xyz[3 * q + 1], p->point_xyz[3 * q + 2]};
                if (dist2(X, Ca) > 0.0) point_sets = true;
            }
        }
        if (!(pose_sets || point_sets)) {
            // farthest_camera, gauge.rs:72-82: max squared distance, ties to the lower index
            long best = -1;
            double best_d = 0.0;
            for (size_t i = 0; i < p->n_poses; ++i) {
                if (i == anchor || !observed[i] || pose_fixed(p, i)) continue;
                double d = dist2(pose_C(p, i), Ca);
                if (!std::isfinite(d)) continue;
                if (best < 0 || d > best_d) {
                    best = (long)i;
                    best_d = d;
                }
            }
            if (best < 0 || !(best_d > 0.0))
                return fail(BA_ERR_UNSUPPORTED, "free-scale BA has no nonzero active baseline");
            native_gauge = (int)best;
        }
    } else {
        size_t first = p->gauge_first, second = p->gauge_second;
        Vec3 d{pose_C(p, second)[0] - pose_C(p, first)[0], pose_C(p, second)[1] - pose_C(p, first)[1],
               pose_C(p, second)[2] - pose_C(p, first)[2]};
        double current = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!std::isfinite(current) || current <= 0.0)
            return fail(BA_ERR_UNSUPPORTED, "fixed-baseline cameras have no usable separation");
        double raw = p->gauge_distance / current;
        double tol = 256.0 * EPS * fmax_(fmax_(std::fabs(current), std::fabs(p->gauge_distance)), 1.0);
        scale = std::fabs(current - p->gauge_distance) <= tol ? 1.0 : raw;
        if (scale != 1.0) {
            bool moves = false;
            for (size_t i = 0; i < p->n_poses; ++i)
                if (i != first && !p->pose_var_translation[i]) moves = true;
            for (size_t i = 0; i < p->n_points; ++i)
                if (point_is_fixed(p, i)) moves = true;
            if (moves) return fail(BA_ERR_UNSUPPORTED, "fixed-baseline scaling would move fixed state");
            if (nq > 0)
                return fail(BA_ERR_UNSUPPORTED, "native BA cannot change its physical baseline while holding out points");
        }
        native_gauge = p->pose_var_translation[second] ? (int)second : -1;
        anchor = first;
    }
    // WorldRebase (world_rebase.rs:19-59)
    L.rebase.Ra = pose_R(p, anchor);
    L.rebase.Ca = pose_C(p, anchor);
    L.rebase.scale = scale;
    const Mat3& Ra = L.rebase.Ra;
    const Vec3& Ca = L.rebase.Ca;
    size_t nc = p->n_poses;
    L.cam_source.clear();
    L.cam_source.push_back((uint32_t)anchor);
    for (size_t i = 0; i < nc; ++i)
        if (i != anchor) L.cam_source.push_back((uint32_t)i);
    std::vector<uint32_t> dest(nc);
    for (size_t d = 0; d < nc; ++d) dest[L.cam_source[d]] = (uint32_t)d;
    L.R.assign(9 * nc, 0.0);
    L.t.assign(3 * nc, 0.0);
    Mat3 RaT = transpose(Ra);
    for (size_t d = 0; d < nc; ++d) {
        size_t s = L.cam_source[d];
        Mat3 R = pose_R(p, s);
        Vec3 C = pose_C(p, s);
        Mat3 Rn;
        Vec3 tn;
        if (R == Ra && C == Ca) {
            Rn = Mat3{Vec3{1, 0, 0}, Vec3{0, 1, 0}, Vec3{0, 0, 1}};
            tn = Vec3{0, 0, 0};
        } else {
            Rn = mat_mul(R, RaT);
            Vec3 diff{Ca[0] - C[0], Ca[1] - C[1], Ca[2] - C[2]};
            Vec3 rd = mat_vec(R, diff);
            tn = Vec3{rd[0] * scale, rd[1] * scale, rd[2] * scale};
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) L.R[9 * d + 3 * r + c] = Rn[r][c];
        for (int k = 0; k < 3; ++k) L.t[3 * d + k] = tn[k];
    }
    // points without the quarantined ones (problem.rs:650-701)
    std::vector<uint32_t> new_index(p->n_points, UINT32_MAX);
    L.kept_points.clear();
    for (size_t i = 0; i < p->n_points; ++i)
        if (!L.quarantined[i]) {
            new_index[i] = (uint32_t)L.kept_points.size();
            L.kept_points.push_back((uint32_t)i);
        }
    size_t np = L.kept_points.size();
    L.X.assign(3 * np, 0.0);
    L.point_fixed.assign(np, 0);
    for (size_t q = 0; q < np; ++q) {
        size_t s = L.kept_points[q];
        Vec3 X{p->point_xyz[3 * s], p->point_xyz[3 * s + 1], p->point_xyz[3 * s + 2]};
        Vec3 diff{X[0] - Ca[0], X[1] - Ca[1], X[2] - Ca[2]};
        Vec3 rd = mat_vec(Ra, diff);
        for (int k = 0; k < 3; ++k) L.X[3 * q + k] = rd[k] * scale;
        L.point_fixed[q] = point_is_fixed(p, s);
    }
    // intrinsics -> radial groups (radial_camera, native/self_calibration.rs:483-512)
    size_t ng = p->n_intrinsics;
    L.gp.assign(3 * ng, 0.0);
    L.grp_fixed.assign(ng, 0);
    L.grp_free.assign(ng, 0);
    std::vector<uint8_t> used(ng, 0);
    for (uint32_t o : kobs) used[p->obs_intrinsic[o]] = 1;
    for (size_t g = 0; g < ng; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        if (L.calibrated) {
            // fixed cameras: the radial projection with f = fx and the declared k1, k2; the
            // models of unobserved intrinsics never enter the objective
            L.grp_fixed[g] = 1;
            L.grp_free[g] = 0;
            L.gp[3 * g] = 1.0;
            if (!used[g]) continue;
            const bool brown = in.model != BA_CAMERA_PINHOLE;
            const double k3 = brown && in.model >= BA_CAMERA_BROWN_K1K2K3 ? in.k3 : 0.0;
            if (in.fx != in.fy || k3 != 0.0 || (brown && (in.p1 != 0.0 || in.p2 != 0.0)))
                return fail(BA_ERR_UNSUPPORTED,
                            "calibrated pixel BA here supports one focal length and k1+k2 radial distortion without "
                            "tangential terms");
            L.gp[3 * g] = in.fx;
            L.gp[3 * g + 1] = brown && in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
            L.gp[3 * g + 2] = brown && in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
            continue;
        }
        if (in.model == BA_CAMERA_FISHEYE || in.model == BA_CAMERA_SPHERICAL)
            return fail(BA_ERR_UNSUPPORTED, "self-calibration supports perspective cameras only");
        double k1 = 0, k2 = 0, k3 = 0, p1 = 0, p2 = 0;
        if (in.model != BA_CAMERA_PINHOLE) {
            k1 = in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
            k2 = in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
            k3 = in.model >= BA_CAMERA_BROWN_K1K2K3 ? in.k3 : 0.0;
            p1 = in.p1;
            p2 = in.p2;
        }
        if (in.fx != in.fy || k3 != 0.0 || p1 != 0.0 || p2 != 0.0)
            return fail(BA_ERR_UNSUPPORTED,
                        "self-calibration requires isotropic focal length, k1+k2 radial distortion, and no tangential distortion");
        L.gp[3 * g] = in.fx;
        L.gp[3 * g + 1] = k1;
        L.gp[3 * g + 2] = k2;
        L.grp_fixed[g] = intrinsic_is_fixed(in) || (hold && (*hold)[g]);
        // free_radial_dimension, native/self_calibration.rs:442-451
        uint32_t dim = 1;
        if (in.model != BA_CAMERA_PINHOLE && in.var_radial) {
            if (in.model == BA_CAMERA_BROWN_NONE) dim = 1;
            else if (in.model == BA_CAMERA_BROWN_K1) dim = 2;
            else dim = 3;
        }
        L.grp_free[g] = dim;
    }
    L.cam_group.assign(nc, 0);
    L.cam_fixed.assign(nc, 0);
    for (size_t d = 0; d < nc; ++d) {
        size_t s = L.cam_source[d];
        L.cam_group[d] = (uint32_t)pose_intr[s];
        L.cam_fixed[d] = (s == anchor && !L.anchor_free) || pose_fixed(p, s);
    }
    L.gauge_camera = native_gauge < 0 ? -1 : (int)dest[native_gauge];
    // observations sorted by (point, source) over the kept points (:119-137)
    std::vector<uint32_t> order = kobs;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return new_index[p->obs_point[a]] < new_index[p->obs_point[b]];
    });
    size_t no = order.size();
    L.obs_cam.assign(no, 0);
    L.obs_src.assign(no, 0);
    L.obs_point.assign(no, 0);
    L.uv.assign(2 * no, 0.0);
    for (size_t k = 0; k < no; ++k) {
        uint32_t s = order[k];
        const ba_intrinsic& in = p->intrinsics[p->obs_intrinsic[s]];
        L.obs_cam[k] = dest[p->obs_pose[s]];
        L.obs_src[k] = s;
        L.obs_point[k] = new_index[p->obs_point[s]];
        L.uv[2 * k] = p->obs_pixel[2 * s] - in.cx;
        L.uv[2 * k + 1] = p->obs_pixel[2 * s + 1] - in.cy;
    }
    L.off.assign(np + 1, 0);
    for (size_t k = 0; k < no; ++k) L.off[L.obs_point[k] + 1] += 1;
    for (size_t q = 0; q < np; ++q) L.off[q + 1] += L.off[q];
    // ReducedLayout::try_radial, workspace/topology.rs:351-397
    L.cam_off.assign(nc, 0);
    L.cam_dim.assign(nc, 0);
    size_t D = 0;
    for (size_t c = 0; c < nc; ++c) {
        uint32_t local = ((c == 0 && !L.anchor_free) || L.cam_fixed[c]) ? 0 : (L.gauge_camera == (int)c ? 5 : 6);
        L.cam_off[c] = (uint32_t)D;
        L.cam_dim[c] = local;
        D += local;
    }
    L.grp_off.assign(ng, 0);
    L.grp_dim.assign(ng, 0);
    for (size_t g = 0; g < ng; ++g) {
        uint32_t local = L.grp_fixed[g] ? 0 : L.grp_free[g];
        L.grp_off[g] = (uint32_t)D;
        L.grp_dim[g] = local;
        D += local;
    }
    L.D = D;
    // radial_point_parameters, workspace/pattern.rs:306-379
    L.cross_off.assign(np + 1, 0);
    L.params.clear();
    std::vector<uint8_t> marks(D, 0);
    std::vector<uint32_t> act;
    for (size_t q = 0; q < np; ++q) {
        act.clear();
        if (!L.point_fixed[q]) {
            for (uint32_t o = L.off[q]; o < L.off[q + 1]; ++o) {
                uint32_t c = L.obs_cam[o];
                for (uint32_t k = 0; k < L.cam_dim[c]; ++k) {
                    uint32_t pr = L.cam_off[c] + k;
                    if (!marks[pr]) {
                        marks[pr] = 1;
                        act.push_back(pr);
                    }
                }
                uint32_t g = L.cam_group[c];
                for (uint32_t k = 0; k < L.grp_dim[g]; ++k) {
                    uint32_t pr = L.grp_off[g] + k;
                    if (!marks[pr]) {
                        marks[pr] = 1;
                        act.push_back(pr);
                    }
                }
            }
        }
        for (uint32_t pr : act) {
            L.params.push_back(pr);
            marks[pr] = 0;
        }
        L.cross_off[q + 1] = (uint32_t)L.params.size();
    }
    // radial_reduced_block_columns + radial_reduced_pattern, pattern.rs:381-468
    size_t nb = nc + ng;
    auto bdim = [&](size_t b) { return b < nc ? L.cam_dim[b] : L.grp_dim[b - nc]; };
    auto boff = [&](size_t b) { return b < nc ? L.cam_off[b] : L.grp_off[b - nc]; };
    std::vector<std::vector<uint8_t>> cols(nb);  // dense membership; sets are small
    std::vector<std::vector<size_t>> col_rows(nb);
    auto insert = [&](size_t col, size_t row) {
        if (cols[col].empty()) cols[col].assign(nb, 0);
        if (!cols[col][row]) {
            cols[col][row] = 1;
            col_rows[col].push_back(row);
        }
    };
    for (size_t b = 0; b < nb; ++b)
        if (bdim(b) != 0) insert(b, b);
    std::vector<size_t> ab;
    for (size_t q = 0; q < np; ++q) {
        if (L.point_fixed[q]) {
            for (uint32_t o = L.off[q]; o < L.off[q + 1]; ++o) {
                size_t c = L.obs_cam[o], g = L.cam_group[c];
                if (L.cam_dim[c] > 0 && L.grp_dim[g] > 0) insert(c, nc + g);
            }
            continue;
        }
        ab.clear();
        for (uint32_t o = L.off[q]; o < L.off[q + 1]; ++o) {
            size_t c = L.obs_cam[o];
            if (L.cam_dim[c] > 0) ab.push_back(c);
            size_t g = L.cam_group[c];
            if (L.grp_dim[g] > 0) ab.push_back(nc + g);
        }
        std::sort(ab.begin(), ab.end());
        ab.erase(std::unique(ab.begin(), ab.end()), ab.end());
        for (size_t ri = 0; ri < ab.size(); ++ri)
            for (size_t ci = 0; ci <= ri; ++ci) insert(ab[ci], ab[ri]);
    }
    L.col_ptr.assign(1, 0);
    L.row_idx.clear();
    for (size_t cb = 0; cb < nb; ++cb) {
        std::sort(col_rows[cb].begin(), col_rows[cb].end());
        for (uint32_t cl = 0; cl < bdim(cb); ++cl) {
            for (size_t rb : col_rows[cb]) {
                uint32_t start = rb == cb ? cl : 0;
                for (uint32_t k = start; k < bdim(rb); ++k) L.row_idx.push_back(boff(rb) + k);
            }
            L.col_ptr.push_back((uint32_t)L.row_idx.size());
        }
    }
    // loss: RobustLoss::kernel_loss, problem.rs:372-382
    if (p->robust_loss == BA_LOSS_HUBER)
        L.loss = Loss{HUBER, p->loss_scale};
    else if (p->robust_loss == BA_LOSS_CAUCHY)
        L.loss = Loss{CAUCHY, p->loss_scale / std::sqrt(2.0)};
    else
        L.loss = Loss{L2, 0.0};
    return Err{};
}

// ---------------------------------------------------------------------------------------
// LM driver: sparse/self_calibration/mod.rs:212-639 with the Nielsen controller
// (numerics/refine/nonlinear/controller.rs), run state (run_state.rs) and stages
// (continuation.rs).
// ---------------------------------------------------------------------------------------
struct Nielsen {
    double value, nu = 2.0;
    bool has_gain = false;
    double gain = NaN;
    void increase() {
        value = fmin_(value * nu, 1e12);
        nu = fmin_(nu * 2.0, 1e12);
    }
    void reject() {
        has_gain = false;
        increase();
    }
    bool accept(double actual, double predicted) {
        double g = (std::isfinite(actual) && std::isfinite(predicted) && predicted > 0.0) ? actual / predicted : NaN;
        has_gain = std::isfinite(g);
        gain = g;
        if (!(std::isfinite(g) && g > 0.0)) {
            increase();
            return false;
        }
        double c = 2.0 * g - 1.0;
        double f = fmax_(1.0 - c * c * c, 1.0 / 3.0);
        value = fmax_(value * f, 1e-12);
        nu = 2.0;
        return true;
    }
};

struct Stage {
    Loss loss;
    size_t iters;
};
static void stages(Loss loss, size_t max_it, Stage out[3]) {  // continuation.rs:26-74
    Loss l2{L2, 0.0};
    if (loss.kind == CAUCHY) {
        size_t s = std::min<size_t>(max_it / 4, 5);
        out[0] = {l2, s};
        out[1] = {loss, max_it - s};
        out[2] = {l2, 0};
    } else {
        out[0] = {loss, max_it};
        out[1] = {l2, 0};
        out[2] = {l2, 0};
    }
}

enum Conv { CONV_NONE = -1, CONV_GRADIENT = 0, CONV_STEP = 1, CONV_COST = 2 };

struct State {
    std::vector<double> R, t, gp, X;
};

static View make_view(const Lowered& L, const State& s, Loss loss, int gauge_camera, double scene_scale) {
    View v;
    v.nc = L.cam_source.size();
    v.ng = L.gp.size() / 3;
    v.np = L.kept_points.size();
    v.no = L.obs_cam.size();
    v.D = L.D;
    v.R = s.R.data();
    v.t = s.t.data();
    v.gp = s.gp.data();
    v.cam_group = L.cam_group.data();
    v.X = s.X.data();
    v.obs_cam = L.obs_cam.data();
    v.uv = L.uv.data();
    v.off = L.off.data();
    v.point_fixed = L.point_fixed.data();
    v.cam_off = L.cam_off.data();
    v.cam_dim = L.cam_dim.data();
    v.grp_off = L.grp_off.data();
    v.grp_dim = L.grp_dim.data();
    v.loss = loss;
    v.gauge = gauge_state(gauge_camera, s.t.data(), scene_scale);
    return v;
}

// candidate_point + admitted_point_cost (schur/point.rs:659-728) under candidate_rule
// (point_geometry.rs:153-168).  Returns false to reject the step.
static bool candidate_point(const View& cur, const View& cand, size_t p, const double prev[3],
                            const double stepped[3], double* pos, double* cost, bool* moved) {
    auto admitted = [&](bool cand_cams, bool cand_point, Admission adm, double* out) {
        const View& v = cand_cams ? cand : cur;
        const double* X = cand_point ? stepped : prev;
        TrackGeometry geo(cur.gauge.scene_scale);
        double total = 0.0;
        for (uint32_t obs = cur.off[p]; obs < cur.off[p + 1]; ++obs) {
            size_t c = cur.obs_cam[obs];
            const double* R = v.R + 9 * c;
            double rot[3];
            rotate(R, X, rot);
            double cp[3] = {rot[0] + v.t[3 * c], rot[1] + v.t[3 * c + 1], rot[2] + v.t[3 * c + 2]};
            double res[2], J[2][3];
            // the candidate intrinsics are used for every evaluation (candidate.rs:202-214)
            if (!radial_projection(cand.group(cur.cam_group[c]), cur.uv + 2 * obs, cp, res, J)) return false;
            double w[3] = {R[0] * cp[0] + R[3] * cp[1] + R[6] * cp[2], R[1] * cp[0] + R[4] * cp[1] + R[7] * cp[2],
                           R[2] * cp[0] + R[5] * cp[1] + R[8] * cp[2]};
            geo.push(w);
            total += loss_cost(cur.loss, std::sqrt(res[0] * res[0] + res[1] * res[1]) * 1.0);
        }
        bool ok = adm == ADMIT_PROJECTION ? true : (adm == ADMIT_COLLAPSE ? !geo.collapsed() : !geo.degenerate());
        if (!ok) return false;
        *out = total;
        return true;
    };
    double c;
    if (admitted(true, true, ADMIT_GEOMETRY, &c)) {
        for (int k = 0; k < 3; ++k) pos[k] = stepped[k];
        *cost = c;
        *moved = true;
        return true;
    }
    if (admitted(true, false, ADMIT_COLLAPSE, &c)) {
        // held
    } else {
        double dummy;
        if (admitted(false, false, ADMIT_COLLAPSE, &dummy)) return false;
        if (!admitted(true, false, ADMIT_PROJECTION, &c)) return false;
    }
    for (int k = 0; k < 3; ++k) pos[k] = prev[k];
    *cost = c;
    *moved = false;
    return true;
}

struct SolveOut {
    ba_solve_report rep;
    State state;
};

static double g_loop_start = 0.0;
static double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static void trace_push(ba_iteration_trace* tr, double cost, double damping, double cand, double pred, double gain,
                       bool acc, uint32_t pcg) {
    if (!tr || tr->count >= tr->capacity) return;
    uint32_t i = tr->count++;
    if (tr->seconds) tr->seconds[i] = wall() - g_loop_start;
    if (tr->cost) tr->cost[i] = cost;
    if (tr->damping) tr->damping[i] = damping;
    if (tr->candidate_cost) tr->candidate_cost[i] = cand;
    if (tr->predicted_reduction) tr->predicted_reduction[i] = pred;
    if (tr->gain_ratio) tr->gain_ratio[i] = gain;
    if (tr->accepted) tr->accepted[i] = acc;
    if (tr->pcg_iterations) tr->pcg_iterations[i] = pcg;
}

// scene_scale (optimizer.rs:620-627), or observed_scene_scale (optimizer.rs:629-668) for the
// calibrated workspace, whose explicit pose masks make the length come from observed points only
// (a scaled running sum of squares)
static double scene_scale_of(const Lowered& L, const double* X, const double* t) {
    const size_t np = L.kept_points.size();
    if (L.gauge_camera >= 0) return norm3(&t[3 * L.gauge_camera]);
    if (L.calibrated) {
        double sc = 0.0, sqs = 1.0;
        size_t count = 0;
        for (size_t q = 0; q < np; ++q) {
            if (L.off[q] == L.off[q + 1]) continue;
            count += 1;
            for (int k = 0; k < 3; ++k) {
                const double m = std::fabs(X[3 * q + k]);
                if (m == 0.0) continue;
                if (sc < m) {
                    const double ratio = sc / m;
                    sqs = 1.0 + sqs * ratio * ratio;
                    sc = m;
                } else {
                    const double ratio = m / sc;
                    sqs += ratio * ratio;
                }
            }
        }
        if (sc == 0.0 || count == 0) return MINPOS;
        const double rms = sc * std::sqrt(sqs / (double)count);
        return std::isfinite(rms) ? fmax_(rms, MINPOS) : sc;
    }
    double s = 0.0;
    for (size_t q = 0; q < np; ++q) s += X[3 * q] * X[3 * q] + X[3 * q + 1] * X[3 * q + 1] + X[3 * q + 2] * X[3 * q + 2];
    return fmax_(std::sqrt(s / (double)np), MINPOS);
}

static Err refine(const Lowered& L, const ba_solve_config* cfg, State& st, ba_solve_report& rep,
                  ba_iteration_trace* trace) {
    const size_t D = L.D;
    const size_t np = L.kept_points.size();
    const bool iterative = cfg->linear_solver == BA_SOLVER_ITERATIVE;
    // validation::projection_domain: every observation projects at the initial state
    {
        View v0 = make_view(L, st, L.loss, L.gauge_camera, 1.0);
        for (size_t q = 0; q < np; ++q)
            for (uint32_t o = L.off[q]; o < L.off[q + 1]; ++o) {
                size_t c = L.obs_cam[o];
                double rot[3];
                rotate(v0.R + 9 * c, v0.X + 3 * q, rot);
                double cp[3] = {rot[0] + v0.t[3 * c], rot[1] + v0.t[3 * c + 1], rot[2] + v0.t[3 * c + 2]};
                double res[2], J[2][3];
                if (!radial_projection(v0.group(L.cam_group[c]), L.uv.data() + 2 * o, cp, res, J))
                    return fail(BA_ERR_NUMERICAL, "point not projectable at the initial state");
            }
    }
    const double scene_scale = scene_scale_of(L, st.X.data(), st.t.data());
    Stage sched[3];
    stages(L.loss, cfg->max_iterations, sched);
    int target = 0;
    for (int i = 2; i >= 0; --i)
        if (sched[i].iters > 0) {
            target = i;
            break;
        }
    g_loop_start = wall();
    double initial_damping = iterative ? 1e-1 : 1e-4;  // api.rs:300-305
    Nielsen damping{initial_damping};
    size_t iterations = 0, accepted = 0, target_accepted = 0;
    size_t rej[4] = {0, 0, 0, 0};
    double step_norm = 0.0;
    int conv = CONV_NONE;
    bool have_initial = false;
    double initial_cost = INF, initial_rmse = INF;
    {
        // reuse_initial_assembly is true when the first running stage uses the target loss
        bool reuse = false;
        for (int i = 0; i < 3; ++i)
            if (sched[i].iters > 0) {
                reuse = sched[i].loss.kind == L.loss.kind;
                break;
            }
        if (!reuse) {  // schur::metrics at the initial state with the target loss
            View v = make_view(L, st, L.loss, L.gauge_camera, scene_scale);
            double cost = 0.0, sq = 0.0;
            for (size_t q = 0; q < np; ++q)
                for (uint32_t o = L.off[q]; o < L.off[q + 1]; ++o) {
                    ObsLin ol;
                    if (linearize_obs(v, o, v.X + 3 * q, ol)) {
                        cost += ol.cost;
                        sq += ol.sq;
                    }
                }
            initial_cost = cost;
            initial_rmse = std::sqrt(sq / (2.0 * (double)L.obs_cam.size()));
            have_initial = true;
        }
    }
    const double iterative_conv_tol = fmin_(std::sqrt(cfg->function_tolerance), 1e-6);
    const size_t pcg_cap = cfg->pcg_max_iterations ? cfg->pcg_max_iterations : 600;
    size_t pcg_total = 0;
    double pcg_max_rel = NaN;
    State cand = st;
    std::vector<double> step(D, 0.0);
    bool stop = false;
    for (int si = 0; si < 3 && !stop; ++si) {
        bool confirm = false;
        for (size_t it = 0; it < sched[si].iters; ++it) {
            iterations += 1;
            double lambda = damping.value;
            View v = make_view(L, st, sched[si].loss, L.gauge_camera, scene_scale);
            Assembly a = assemble(v, lambda, !iterative, L.cross_off.data());
            if (!have_initial && a.cost == a.cost) {
                if (std::isfinite(a.cost) && std::isfinite(a.sq) && !L.obs_cam.empty()) {
                    initial_cost = a.cost;
                    initial_rmse = std::sqrt(a.sq / (2.0 * (double)L.obs_cam.size()));
                    have_initial = true;
                }
            }
            if (!a.valid) {
                rej[0]++;
                damping.reject();
                trace_push(trace, a.cost, lambda, NaN, NaN, NaN, false, 0);
                continue;
            }
            double current = a.cost;
            // run.gradient_converged, run_state.rs:116-133
            if (si == target && conv == CONV_NONE && std::isfinite(a.sgin) && a.sgin >= 0.0 &&
                std::isfinite(current) && a.sgin <= cfg->gradient_tolerance * (1.0 + std::fabs(current))) {
                conv = CONV_GRADIENT;
                trace_push(trace, current, lambda, NaN, NaN, NaN, false, 0);
                stop = true;
                break;
            }
            // camera step
            bool solved;
            uint32_t pcg_terms = 0;
            double pcg_rel = NaN;
            std::vector<double> product;
            Retained ret;
            std::vector<BlockFactor> factors;
            if (!iterative) {
                std::vector<double> F = a.H.full;
                std::vector<double> rhs(D);
                for (size_t i = 0; i < D; ++i) rhs[i] = -a.g[i];
                solved = D == 0 ? true : dense_cholesky_solve(F, D, rhs.data(), step.data());
            } else {
                double base_tol = cfg->pcg_relative_tolerance > 0.0 ? cfg->pcg_relative_tolerance : 1e-1;
                double tol = confirm ? fmin_(base_tol, iterative_conv_tol) : base_tol;
                size_t limit = confirm ? pcg_cap : std::min(pcg_cap, D);
                if (!build_retained(v, a, L.cross_off.data(), L.params.data(), ret, factors)) {
                    solved = false;
                } else {
                    auto pre = [&](const double* r, double* z) {
                        std::fill(z, z + D, 0.0);
                        for (size_t b = 0; b < ret.nblocks; ++b)
                            if (ret.bdim[b] > 0) factors[b].solve(r + ret.boff[b], z + ret.boff[b], ret.bdim[b]);
                    };
                    auto op = [&](const double* x, double* y) { ret.apply(x, y); };
                    PcgResult r = pcg_solve(a.g.data(), D, limit, tol, op, pre, step, product);
                    pcg_terms = (uint32_t)r.terms;
                    pcg_total += r.terms;
                    pcg_rel = r.rel;
                    if (std::isfinite(r.rel)) pcg_max_rel = std::isfinite(pcg_max_rel) ? fmax_(pcg_max_rel, r.rel) : r.rel;
                    solved = r.accepted;
                }
            }
            if (!solved) {
                rej[1]++;
                damping.reject();
                trace_push(trace, current, lambda, NaN, NaN, NaN, false, pcg_terms);
                continue;
            }
            if (L.calibrated) {
                // limit_camera_step (optimizer.rs:677-702) at the start of the calibrated candidate
                // evaluation (calibrated/candidate.rs:36-44): the largest rotation or translation
                // block norm of any camera is held to 0.5 by scaling the whole step
                double largest = 0.0;
                for (size_t c = 0; c < L.cam_dim.size(); ++c) {
                    const size_t off = L.cam_off[c], dim = L.cam_dim[c];
                    if (dim == 0) continue;
                    const size_t first = std::min<size_t>(dim, 3);
                    double s1 = 0.0, s2 = 0.0;
                    for (size_t k = 0; k < first; ++k) s1 += step[off + k] * step[off + k];
                    for (size_t k = first; k < dim; ++k) s2 += step[off + k] * step[off + k];
                    largest = fmax_(largest, std::sqrt(s1));
                    largest = fmax_(largest, std::sqrt(s2));
                }
                if (largest > 0.5) {
                    const double sc = 0.5 / largest;
                    for (size_t i = 0; i < D; ++i) step[i] *= sc;
                }
                // the retained reduction applies the operator to the step it predicts for
                // (retained/mod.rs:781-800)
                if (iterative) ret.apply(step.data(), product.data());
            }
            // retract_poses (optimizer.rs:386-437) and retract_cameras (self_calibration/optimizer.rs:565-586)
            const GaugeState& gs = v.gauge;
            for (size_t c = 0; c < v.nc; ++c) {
                size_t off = L.cam_off[c], pd = L.cam_dim[c];
                if (pd == 0) {
                    for (int k = 0; k < 9; ++k) cand.R[9 * c + k] = st.R[9 * c + k];
                    for (int k = 0; k < 3; ++k) cand.t[3 * c + k] = st.t[3 * c + k];
                    continue;
                }
                Mat3 up = exp_so3(&step[off]);
                Mat3 Rc;
                for (int r = 0; r < 3; ++r)
                    for (int k = 0; k < 3; ++k) Rc[r][k] = st.R[9 * c + 3 * r + k];
                Mat3 Rn = mat_mul(up, Rc);
                double tn[3];
                if (gs.fixed_camera == (int)c) {
                    double dir[3];
                    retract_s2(gs.direction, gs.basis, &step[off + 3], dir);
                    for (int k = 0; k < 3; ++k) tn[k] = gs.baseline * dir[k];
                } else {
                    for (int k = 0; k < 3; ++k) tn[k] = st.t[3 * c + k] + gs.scene_scale * step[off + 3 + k];
                }
                for (int r = 0; r < 3; ++r)
                    for (int k = 0; k < 3; ++k) cand.R[9 * c + 3 * r + k] = Rn[r][k];
                for (int k = 0; k < 3; ++k) cand.t[3 * c + k] = tn[k];
            }
            for (size_t g = 0; g < v.ng; ++g) {
                if (L.grp_dim[g] == 0) {
                    for (int k = 0; k < 3; ++k) cand.gp[3 * g + k] = st.gp[3 * g + k];
                    continue;
                }
                double delta[3] = {0, 0, 0};
                for (uint32_t k = 0; k < L.grp_dim[g]; ++k) delta[k] = step[L.grp_off[g] + k];
                cand.gp[3 * g] = st.gp[3 * g] * std::exp(delta[0]);
                cand.gp[3 * g + 1] = st.gp[3 * g + 1] + delta[1];
                cand.gp[3 * g + 2] = st.gp[3 * g + 2] + delta[2];
            }
            // candidate_points, cached back-substitution (candidate.rs:25-233, 358-391)
            View cv = make_view(L, cand, sched[si].loss, L.gauge_camera, scene_scale);
            cv.gauge = gs;  // the candidate uses the linearisation's gauge state
            // per point shard: its sums in point order, stopping at its first failing point; the
            // shards' sums are added in shard order (a failing shard rejects the step)
            const std::vector<size_t> cshards = point_shards(L.off.data(), np, 0);
            const size_t ncs = cshards.size() - 1;
            std::vector<double> csq(ncs, 0.0), cpdn(ncs, 0.0), ccost_s(ncs, 0.0);
            std::vector<uint8_t> cok(ncs, 1);
#pragma omp parallel for schedule(dynamic, 1)
            for (size_t sh = 0; sh < ncs; ++sh) {
            double sq = 0.0, pdn = 0.0, ccost = 0.0;
            bool cand_ok = true;
            for (size_t q = cshards[sh]; q < cshards[sh + 1] && cand_ok; ++q) {
                const double* prev = &st.X[3 * q];
                if (L.point_fixed[q]) {
                    for (int k = 0; k < 3; ++k) cand.X[3 * q + k] = prev[k];
                    // fixed_point_cost: projection-only admission at the candidate cameras
                    double tot = 0.0;
                    for (uint32_t o = L.off[q]; o < L.off[q + 1] && cand_ok; ++o) {
                        size_t c = L.obs_cam[o];
                        double rot[3];
                        rotate(cv.R + 9 * c, prev, rot);
                        double cp[3] = {rot[0] + cv.t[3 * c], rot[1] + cv.t[3 * c + 1], rot[2] + cv.t[3 * c + 2]};
                        double res[2], J[2][3];
                        if (!radial_projection(cv.group(L.cam_group[c]), L.uv.data() + 2 * o, cp, res, J)) {
                            cand_ok = false;
                            break;
                        }
                        tot += loss_cost(cv.loss, std::sqrt(res[0] * res[0] + res[1] * res[1]));
                    }
                    ccost += tot;
                    continue;
                }
                if (!a.cache_valid[q]) {
                    cand_ok = false;
                    break;
                }
                double s3[3] = {a.inv_grad[3 * q], a.inv_grad[3 * q + 1], a.inv_grad[3 * q + 2]};
                for (uint32_t k = L.cross_off[q]; k < L.cross_off[q + 1]; ++k) {
                    double val = step[L.params[k]];
                    for (int c = 0; c < 3; ++c) s3[c] += a.inv_cross[3 * k + c] * val;
                }
                double dp[3] = {-s3[0], -s3[1], -s3[2]};
                double dn = a.damp_diag[3 * q] * dp[0] * dp[0] + a.damp_diag[3 * q + 1] * dp[1] * dp[1] +
                            a.damp_diag[3 * q + 2] * dp[2] * dp[2];
                double stepped[3] = {prev[0] + gs.scene_scale * dp[0], prev[1] + gs.scene_scale * dp[1],
                                     prev[2] + gs.scene_scale * dp[2]};
                double pos[3], pc;
                bool mv;
                if (!candidate_point(v, cv, q, prev, stepped, pos, &pc, &mv)) {
                    cand_ok = false;
                    break;
                }
                for (int k = 0; k < 3; ++k) cand.X[3 * q + k] = pos[k];
                ccost += pc;
                if (mv) {
                    pdn += dn;
                    sq += dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                }
            }
            csq[sh] = sq;
            cpdn[sh] = pdn;
            ccost_s[sh] = ccost;
            cok[sh] = cand_ok;
            }
            double sq = csq[0], pdn = cpdn[0], ccost = ccost_s[0];
            bool cand_ok = cok[0] != 0;
            for (size_t sh = 1; sh < ncs; ++sh) {
                sq += csq[sh];
                pdn += cpdn[sh];
                ccost += ccost_s[sh];
                cand_ok = cand_ok && cok[sh] != 0;
            }
            if (!cand_ok) {
                rej[2]++;
                damping.reject();
                trace_push(trace, current, lambda, NaN, NaN, NaN, false, pcg_terms);
                continue;
            }
            double point_rms = std::sqrt(sq / (double)np);
            // self_calibration_model_reduction (camera_step.rs:33-67)
            double predicted;
            double cdn = 0.0;
            for (size_t i = 0; i < D; ++i) cdn += a.diag[i] * step[i] * step[i];
            if (!iterative && L.calibrated) {
                // calibrated_model_reduction (calibrated/camera_step.rs:40-62): the quadratic form
                // of the damped reduced matrix over its lower pattern (direct.rs:235-257), then
                // reduced_model_reduction and undamped_model_reduction (optimizer.rs:590-618)
                double quad = 0.0;
                for (size_t col = 0; col < D; ++col)
                    for (size_t row = col; row < D; ++row) {
                        const double hv = a.H.full[row * D + col];
                        if (hv == 0.0) continue;
                        const double contribution = hv * step[row] * step[col];
                        quad += row == col ? contribution : 2.0 * contribution;
                    }
                double gs_dot = 0.0;
                for (size_t i = 0; i < D; ++i) gs_dot += a.g[i] * step[i];
                const double damped = a.pmr - gs_dot - 0.5 * quad;
                predicted = damped + 0.5 * lambda * (cdn + pdn);
            } else if (!iterative) {
                double gs_dot = 0.0;
                for (size_t i = 0; i < D; ++i) gs_dot += a.g[i] * step[i];
                double damped = a.pmr - 0.5 * gs_dot;
                predicted = damped + 0.5 * lambda * (cdn + pdn);
            } else {
                double camred = -vdot(a.g.data(), step.data(), D) - 0.5 * vdot(step.data(), product.data(), D);
                double pr = a.pmr + camred;
                predicted = (std::isfinite(pr) && pr > 0.0) ? pr + 0.5 * lambda * (cdn + pdn) : NaN;
            }
            double reduction = current - ccost;
            bool acc = damping.accept(reduction, predicted);
            trace_push(trace, current, lambda, ccost, predicted, damping.gain, acc, pcg_terms);
            if (!acc) {
                rej[3]++;
                continue;
            }
            st = cand;
            accepted += 1;
            double rms = std::sqrt(vdot(step.data(), step.data(), D) / (double)std::max<size_t>(D, 1));
            double sn = std::sqrt(rms * rms + point_rms * point_rms);
            step_norm = sn;
            bool is_target = si == target;
            int candidate_conv = CONV_NONE;
            if (is_target) {
                target_accepted += 1;
                double rel = reduction / (1.0 + std::fabs(current));
                if (sn <= cfg->parameter_tolerance)
                    candidate_conv = CONV_STEP;
                else if (rel <= cfg->function_tolerance)
                    candidate_conv = CONV_COST;
                if (iterative && candidate_conv != CONV_NONE && pcg_rel > iterative_conv_tol) {
                    confirm = true;
                    conv = CONV_NONE;
                } else {
                    confirm = false;
                    conv = candidate_conv;
                    if (conv != CONV_NONE) {
                        stop = true;
                        break;
                    }
                }
            }
        }
    }
    // final undamped assembly (mod.rs:498-560)
    View fv = make_view(L, st, L.loss, L.gauge_camera, scene_scale);
    Assembly fa = assemble(fv, 0.0, false, L.cross_off.data());
    double final_cost, final_rmse;
    if (fa.valid && std::isfinite(fa.cost) && std::isfinite(fa.sq) && !L.obs_cam.empty()) {
        final_cost = fa.cost;
        final_rmse = std::sqrt(fa.sq / (2.0 * (double)L.obs_cam.size()));
    } else {
        final_cost = INF;
        final_rmse = INF;
        if (!fa.valid) {
            final_cost = fa.cost;
            final_rmse = std::sqrt(fa.sq / (2.0 * (double)L.obs_cam.size()));
        }
    }
    std::memset(&rep, 0, sizeof(rep));
    // RunState::status, run_state.rs:170-180 (rank certificate not assessed here)
    if (conv != CONV_NONE)
        rep.status = conv;
    else if (iterations > 0 && target_accepted == 0)
        rep.status = BA_STATUS_STALLED;
    else
        rep.status = BA_STATUS_MAX_ITERATIONS;
    rep.linear_solver_used = iterative ? BA_SOLVER_ITERATIVE : BA_SOLVER_DIRECT;
    rep.iterations = (uint32_t)iterations;
    rep.accepted_steps = (uint32_t)accepted;
    rep.rejected_linearization = (uint32_t)rej[0];
    rep.rejected_step_solve = (uint32_t)rej[1];
    rep.rejected_candidate = (uint32_t)rej[2];
    rep.rejected_insufficient_reduction = (uint32_t)rej[3];
    rep.rejected_steps = (uint32_t)(rej[0] + rej[1] + rej[2] + rej[3]);
    rep.initial_cost = initial_cost;
    rep.final_cost = final_cost;
    rep.initial_rmse = initial_rmse * std::sqrt(2.0);
    rep.final_rmse = final_rmse * std::sqrt(2.0);
    rep.gradient_inf_norm = fa.sgin;
    rep.step_norm = step_norm;
    rep.initial_damping = initial_damping;
    rep.final_damping = damping.value;
    rep.last_gain_ratio = damping.has_gain ? damping.gain : NaN;
    rep.reduced_dimension = (uint32_t)D;
    rep.weak_point_count = (uint32_t)fa.weak;
    rep.minimum_point_rcond = fa.min_rcond;
    rep.pcg_iterations = (uint32_t)pcg_total;
    rep.maximum_pcg_relative_residual = pcg_max_rel;
    rep.normalization_scale = scene_scale;
    rep.observation_count = L.obs_cam.size();
    rep.point_count = np;
    rep.pose_count = L.cam_source.size();
    return Err{};
}

// raise_self_calibrating_solution, native/self_calibration.rs:375-436 and world_rebase.rs:61-86
static void raise_solution(const ba_problem* p, const Lowered& L, const State& st, ba_solution* out) {
    const Mat3& Ra = L.rebase.Ra;
    Mat3 RaT = transpose(Ra);
    for (size_t i = 0; i < p->n_poses; ++i) {
        for (int k = 0; k < 9; ++k) out->pose_rotation[9 * i + k] = p->pose_rotation[9 * i + k];
        for (int k = 0; k < 3; ++k) out->pose_center[3 * i + k] = p->pose_center[3 * i + k];
    }
    for (size_t d = 0; d < L.cam_source.size(); ++d) {
        size_t s = L.cam_source[d];
        if (pose_fixed(p, s)) continue;
        Mat3 Rn;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Rn[r][c] = st.R[9 * d + 3 * r + c];
        Mat3 R = mat_mul(Rn, Ra);
        Vec3 t{st.t[3 * d], st.t[3 * d + 1], st.t[3 * d + 2]};
        Vec3 rc = mat_vec(transpose(Rn), t);
        Vec3 rel{-rc[0], -rc[1], -rc[2]};
        Vec3 wc = mat_vec(RaT, rel);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out->pose_rotation[9 * s + 3 * r + c] = R[r][c];
        for (int k = 0; k < 3; ++k) out->pose_center[3 * s + k] = L.rebase.Ca[k] + wc[k];
    }
    for (size_t i = 0; i < p->n_points; ++i)
        for (int k = 0; k < 3; ++k) out->point_xyz[3 * i + k] = p->point_xyz[3 * i + k];
    for (size_t q = 0; q < L.kept_points.size(); ++q) {
        size_t s = L.kept_points[q];
        if (point_is_fixed(p, s)) continue;
        Vec3 X{st.X[3 * q], st.X[3 * q + 1], st.X[3 * q + 2]};
        Vec3 w = mat_vec(RaT, X);
        for (int k = 0; k < 3; ++k) out->point_xyz[3 * s + k] = L.rebase.Ca[k] + w[k];
    }
    if (out->point_quarantined)
        for (size_t i = 0; i < p->n_points; ++i) out->point_quarantined[i] = L.quarantined[i];
    for (size_t g = 0; g < p->n_intrinsics; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        if (L.grp_fixed[g]) {
            out->intrinsic_focal[g] = in.fx;
            out->intrinsic_k1[g] = in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
            out->intrinsic_k2[g] = in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
        } else {
            out->intrinsic_focal[g] = st.gp[3 * g];
            out->intrinsic_k1[g] = st.gp[3 * g + 1];
            out->intrinsic_k2[g] = st.gp[3 * g + 2];
        }
        if (out->intrinsic_unconstrained) out->intrinsic_unconstrained[g] = 0;
    }
}

// LinearSolver::Auto: kernel/multi_view/sparse/execution_policy.rs:190-227 and 280-286, with the
// estimates of workspace/self_calibration.rs:466-535 (estimated_direct_bytes with a serial direct
// assembly: execution_policy.rs:68-120; estimated_self_calibrating_retained_bytes without its
// worker term: self_calibration_retained.rs:233-290).  The lower nonzero count is the length of
// the reduced pattern (pattern.rs:441-468), which lower() builds.
static void plan_solver(const Lowered& L, const ba_solve_config* c, ba_execution_plan& out) {
    std::memset(&out, 0, sizeof(out));
    out.reduced_dimension = (uint32_t)L.D;
    if (c->linear_solver != BA_SOLVER_AUTO) {
        out.linear_solver = c->linear_solver;
        out.reason = BA_PLAN_EXPLICIT_SOLVER;
        return;
    }
    const uint64_t limit = c->workspace_budget_bytes ? c->workspace_budget_bytes : (uint64_t)1 << 30;
    const uint64_t np = L.kept_points.size(), cross = L.params.size(), D = L.D;
    const uint64_t nnz = L.row_idx.size();
    if (L.calibrated) {
        // select_calibrated (execution_policy.rs:123-183) over the serial estimates of
        // workspace/calibrated.rs:480-530: the cached point cache keeps point steps and f64 cross
        // rows (execution_policy.rs:68-96); the direct estimate adds the dense rank certificate up
        // to dimension 2048 (workspace.rs:126-132); the retained estimate is
        // estimated_calibrated_retained_bytes (retained/mod.rs:156-191) without worker shards
        const uint64_t nc = L.cam_dim.size(), no = L.obs_cam.size();
        const uint64_t cached_points = (np + 1) * 8 + np * (1 + 2 * 24) + cross * 24;
        const uint64_t dense_rank = D <= 2048 ? (D + 2) * D * 8 : 0;
        out.budget_bytes = limit;
        out.reduced_nonzeros = nnz;
        out.estimated_direct_bytes = nnz * 8 * 4 + dense_rank + cached_points;
        out.estimated_retained_bytes = nc * (2 * 168 + 216) + np * 48 + D * 48 + cached_points;
        const bool prefers_cached = no >= 250000 && D >= 768;  // execution_policy.rs:14-22
        const uint64_t guarded = limit / 5 * 4 + limit % 5 * 4 / 5;
        if (prefers_cached && out.estimated_retained_bytes <= limit) {
            out.linear_solver = BA_SOLVER_ITERATIVE;
            out.reason = BA_PLAN_WORKLOAD_PREFERS_RETAINED;
        } else if (out.estimated_direct_bytes <= guarded) {
            out.linear_solver = BA_SOLVER_DIRECT;
            out.reason = BA_PLAN_DIRECT_FITS_BUDGET;
        } else {
            out.linear_solver = BA_SOLVER_ITERATIVE;
            out.reason = out.estimated_retained_bytes <= limit ? BA_PLAN_DIRECT_EXCEEDS_BUDGET : BA_PLAN_LOWEST_MEMORY;
        }
        return;
    }
    const uint64_t point_cache = (np + 1) * sizeof(uint64_t) + cross * 3 * sizeof(double);
    out.budget_bytes = limit;
    out.reduced_nonzeros = nnz;
    out.estimated_direct_bytes = nnz * sizeof(double) * 4 + point_cache;
    const uint64_t blocks = L.cam_dim.size() + L.grp_dim.size();
    uint64_t edges = 0;
    for (size_t k = 0; k < L.cam_dim.size(); ++k) edges += L.cam_dim[k] > 0 && L.grp_dim[L.cam_group[k]] > 0;
    const uint64_t packed_block = 21 * 8, block_factor = (21 + 6) * 8, edge = 2 * 8 + 18 * 8;
    out.estimated_retained_bytes = blocks * (2 * packed_block + block_factor) + edges * edge + blocks * 3 * 8 +
                                   np * 6 * 8 + cross * 4 + D * (5 * 8 + 8) + (np + 1) * 8 + np * (1 + 6 * 8) +
                                   cross * 3 * 8;
    const uint64_t guarded = limit / 5 * 4 + limit % 5 * 4 / 5;
    if (out.estimated_direct_bytes <= guarded) {
        out.linear_solver = BA_SOLVER_DIRECT;
        out.reason = BA_PLAN_DIRECT_FITS_BUDGET;
    } else {
        out.linear_solver = BA_SOLVER_ITERATIVE;
        out.reason = out.estimated_retained_bytes <= limit ? BA_PLAN_DIRECT_EXCEEDS_BUDGET : BA_PLAN_LOWEST_MEMORY;
    }
}

// inlier_metrics / visit_residual_norms / residual_norm (solver.rs:1353-1430): pixel residual
// norms of the returned state in the caller's frame, over the observations of kept points.
static void inlier_metrics(const ba_problem* p, const Lowered& L, const ba_solution* s, uint64_t* count,
                           double* rmse) {
    const double limit = p->robust_loss == BA_LOSS_SQUARED ? INFINITY : p->loss_scale;
    uint64_t n = 0;
    double sum = 0.0;
    for (size_t o = 0; o < p->n_observations; ++o) {
        const size_t q = p->obs_point[o], i = p->obs_pose[o], g = p->obs_intrinsic[o];
        if (L.quarantined[q]) continue;
        const double* R = s->pose_rotation + 9 * i;
        const double* C = s->pose_center + 3 * i;
        const double* X = s->point_xyz + 3 * q;
        const double d[3] = {X[0] - C[0], X[1] - C[1], X[2] - C[2]};
        double xc[3];
        for (int r = 0; r < 3; ++r) xc[r] = R[3 * r] * d[0] + R[3 * r + 1] * d[1] + R[3 * r + 2] * d[2];
        if (!(xc[2] > 0.0)) continue;
        const double nx = xc[0] / xc[2], ny = xc[1] / xc[2], r2 = nx * nx + ny * ny;
        const ba_intrinsic& in = p->intrinsics[g];
        const double dist = 1.0 + s->intrinsic_k1[g] * r2 + s->intrinsic_k2[g] * r2 * r2;
        const double f = s->intrinsic_focal[g];
        const double ex = f * dist * nx + in.cx - p->obs_pixel[2 * o];
        const double ey = f * dist * ny + in.cy - p->obs_pixel[2 * o + 1];
        const double norm = std::sqrt(ex * ex + ey * ey);
        if (norm <= limit) {
            n += 1;
            sum += norm * norm;
        }
    }
    *count = n;
    *rmse = n > 0 ? std::sqrt(sum / (double)n) : NaN;
    if (!std::isfinite(*rmse)) *rmse = NaN;
}

static void copy_msg(const Err& e, char* msg, size_t len) {
    if (msg && len) {
        std::snprintf(msg, len, "%s", e.msg.c_str());
    }
}

}  // namespace bao

using namespace bao;

extern "C" {

// ---------------------------------------------------------------------------------------
// Post-loop certification: the final rank certificate (sparse/self_calibration/mod.rs:498-560,
// direct.rs:288-334), the intrinsic marginals and the held re-solve (native/self_calibration.rs:
// 223-303), the weak-point re-solves (native/mod.rs:252-350, solver.rs:1149-1219) and the round-off
// canonicalisation (solver.rs:1030-1101).
// ---------------------------------------------------------------------------------------
struct OwnedProblemO {
    std::vector<double> rot, cen, xyz, pix;
    std::vector<uint8_t> vr, vt, fixed;
    std::vector<ba_intrinsic> intr;
    std::vector<uint32_t> pose, group, point;
    ba_problem p{};
};

struct SolutionO {
    std::vector<double> rot, cen, f, k1, k2, xyz, cov;
    std::vector<uint8_t> quarantined, unconstrained;
    ba_solve_report rep{};
    void init(const ba_problem* p) {
        rot.assign(9 * p->n_poses, 0.0);
        cen.assign(3 * p->n_poses, 0.0);
        f.assign(p->n_intrinsics, 0.0);
        k1.assign(p->n_intrinsics, 0.0);
        k2.assign(p->n_intrinsics, 0.0);
        xyz.assign(3 * p->n_points, 0.0);
        cov.assign(9 * p->n_intrinsics, NaN);
        quarantined.assign(p->n_points, 0);
        unconstrained.assign(p->n_intrinsics, 0);
        std::memset(&rep, 0, sizeof(rep));
    }
    ba_solution view() {
        ba_solution s{};
        s.pose_rotation = rot.data();
        s.pose_center = cen.data();
        s.intrinsic_focal = f.data();
        s.intrinsic_k1 = k1.data();
        s.intrinsic_k2 = k2.data();
        s.point_xyz = xyz.data();
        s.point_quarantined = quarantined.data();
        s.intrinsic_unconstrained = unconstrained.data();
        s.intrinsic_covariance = cov.data();
        return s;
    }
};

// BundleProblem::without_points (problem.rs:650-701)
static void restrict_problem(const ba_problem* p, const SolutionO* warm, const std::vector<uint8_t>& drop,
                             OwnedProblemO& o, std::vector<size_t>& kept) {
    o.rot = warm ? warm->rot : std::vector<double>(p->pose_rotation, p->pose_rotation + 9 * p->n_poses);
    o.cen = warm ? warm->cen : std::vector<double>(p->pose_center, p->pose_center + 3 * p->n_poses);
    o.vr.assign(p->pose_var_rotation, p->pose_var_rotation + p->n_poses);
    o.vt.assign(p->pose_var_translation, p->pose_var_translation + p->n_poses);
    o.intr.assign(p->intrinsics, p->intrinsics + p->n_intrinsics);
    if (warm) {
        for (size_t g = 0; g < o.intr.size(); ++g) {
            ba_intrinsic& in = o.intr[g];
            in.fx = warm->f[g];
            in.fy = warm->f[g];
            if (in.model == BA_CAMERA_BROWN_K1 || in.model == BA_CAMERA_BROWN_K1K2 || in.model == BA_CAMERA_BROWN_K1K2K3)
                in.k1 = warm->k1[g];
            if (in.model == BA_CAMERA_BROWN_K1K2 || in.model == BA_CAMERA_BROWN_K1K2K3) in.k2 = warm->k2[g];
        }
    }
    std::vector<size_t> index(p->n_points, SIZE_MAX);
    kept.clear();
    o.xyz.clear();
    o.fixed.clear();
    for (size_t i = 0; i < p->n_points; ++i) {
        if (drop[i]) continue;
        index[i] = kept.size();
        kept.push_back(i);
        for (int k = 0; k < 3; ++k) o.xyz.push_back(warm ? warm->xyz[3 * i + k] : p->point_xyz[3 * i + k]);
        o.fixed.push_back(point_is_fixed(p, i) ? 1 : 0);
    }
    o.pose.clear();
    o.group.clear();
    o.point.clear();
    o.pix.clear();
    for (size_t k = 0; k < p->n_observations; ++k) {
        const size_t q = index[p->obs_point[k]];
        if (q == SIZE_MAX) continue;
        o.pose.push_back(p->obs_pose[k]);
        o.group.push_back(p->obs_intrinsic[k]);
        o.point.push_back((uint32_t)q);
        o.pix.push_back(p->obs_pixel[2 * k]);
        o.pix.push_back(p->obs_pixel[2 * k + 1]);
    }
    o.p = *p;
    o.p.pose_rotation = o.rot.data();
    o.p.pose_center = o.cen.data();
    o.p.pose_var_rotation = o.vr.data();
    o.p.pose_var_translation = o.vt.data();
    o.p.intrinsics = o.intr.data();
    o.p.n_points = kept.size();
    o.p.point_xyz = o.xyz.data();
    o.p.point_fixed = o.fixed.data();
    o.p.n_observations = o.point.size();
    o.p.obs_pose = o.pose.data();
    o.p.obs_intrinsic = o.group.data();
    o.p.obs_point = o.point.data();
    o.p.obs_pixel = o.pix.data();
}

// Eigenvalues of a symmetric 3x3 matrix from its characteristic cubic (trigonometric form).
static void sym3_eigenvalues(const double a[3][3], double ev[3]) {
    const double p1 = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double q = (a[0][0] + a[1][1] + a[2][2]) / 3.0;
    if (p1 == 0.0) {
        ev[0] = a[0][0];
        ev[1] = a[1][1];
        ev[2] = a[2][2];
        return;
    }
    const double p2 = (a[0][0] - q) * (a[0][0] - q) + (a[1][1] - q) * (a[1][1] - q) + (a[2][2] - q) * (a[2][2] - q) +
                      2.0 * p1;
    const double pp = std::sqrt(p2 / 6.0);
    double b[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) b[i][j] = (a[i][j] - (i == j ? q : 0.0)) / pp;
    const double det = b[0][0] * (b[1][1] * b[2][2] - b[1][2] * b[2][1]) - b[0][1] * (b[1][0] * b[2][2] - b[1][2] * b[2][0]) +
                       b[0][2] * (b[1][0] * b[2][1] - b[1][1] * b[2][0]);
    const double r = std::fmin(std::fmax(det / 2.0, -1.0), 1.0);
    const double phi = std::acos(r) / 3.0;
    ev[0] = q + 2.0 * pp * std::cos(phi);
    ev[2] = q + 2.0 * pp * std::cos(phi + 2.0 * M_PI / 3.0);
    ev[1] = 3.0 * q - ev[0] - ev[2];
}

// CameraModel::project_with_jacobian of the pinhole and the radial Brown-Conrady camera
// (openmvg-core camera/pinhole.rs:301-322, brown_conrady.rs:240-267, 334-360,
// distortion.rs:530-633) at a camera-frame point: pixel and d pixel / d point.
static bool camera_project(const ba_intrinsic& in, double f, double k1, double k2, const double pc[3], double pix[2],
                           double J[2][3]) {
    if (!(std::isfinite(pc[0]) && std::isfinite(pc[1]) && std::isfinite(pc[2]))) return false;
    if (pc[2] <= 0.0) return false;
    const double iz = 1.0 / pc[2];
    if (in.model == BA_CAMERA_PINHOLE) {
        pix[0] = std::fma(f, pc[0] * iz, in.cx);
        pix[1] = std::fma(f, pc[1] * iz, in.cy);
        J[0][0] = f * iz;
        J[0][1] = 0.0;
        J[0][2] = -f * pc[0] * (iz * iz);
        J[1][0] = 0.0;
        J[1][1] = f * iz;
        J[1][2] = -f * pc[1] * (iz * iz);
    } else {
        const double x = pc[0] / pc[2], y = pc[1] / pc[2];
        if (!(std::isfinite(x) && std::isfinite(y))) return false;
        // distortion_jacobian: (j11, j12, j22) and its determinant
        auto dj = [&](double out[4]) {
            const double q = std::fma(x, x, y * y);
            const double rf = 1.0 + k1 * q + k2 * q * q;
            const double rd = k1 + 2.0 * (k2 * q);
            out[0] = rf + 2.0 * x * x * rd;
            out[1] = 2.0 * x * y * rd;
            out[2] = rf + 2.0 * y * y * rd;
            out[3] = std::fma(out[0], out[2], -out[1] * out[1]);
            return std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2]) && std::isfinite(out[3]);
        };
        double d[4];
        if (k1 != 0.0 || k2 != 0.0) {
            const double q = std::fma(x, x, y * y);
            if (!std::isfinite(q)) return false;
            double turn;
            if (first_positive_radial_turn(Radial{f, k1, k2}, &turn) && q >= turn) return false;
            if (!dj(d) || d[3] <= 1e-12) return false;
        }
        const double q = x * x + y * y;
        const double rf = 1.0 + k1 * q + k2 * q * q;
        pix[0] = std::fma(f, x * rf, in.cx);
        pix[1] = std::fma(f, y * rf, in.cy);
        if (!dj(d)) return false;
        J[0][0] = f * d[0] * iz;
        J[0][1] = f * d[1] * iz;
        J[0][2] = -f * (d[0] * x + d[1] * y) * iz;
        J[1][0] = f * d[1] * iz;
        J[1][1] = f * d[2] * iz;
        J[1][2] = -f * (d[1] * x + d[2] * y) * iz;
    }
    if (!(std::isfinite(pix[0]) && std::isfinite(pix[1]))) return false;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(J[r][c])) return false;
    return true;
}

// Solution::weak_points (solver.rs:1149-1219)
static std::vector<uint8_t> solution_weak_points(const ba_problem* p, const SolutionO& s) {
    const size_t N = p->n_points;
    Loss loss{L2, 0.0};
    if (p->robust_loss == BA_LOSS_HUBER) loss = Loss{HUBER, p->loss_scale};
    if (p->robust_loss == BA_LOSS_CAUCHY) loss = Loss{CAUCHY, p->loss_scale / std::sqrt(2.0)};
    std::vector<std::array<double, 9>> info(N);
    for (auto& b : info) b.fill(0.0);
    std::vector<uint8_t> missing(N, 0);
    for (size_t o = 0; o < p->n_observations; ++o) {
        const size_t q = p->obs_point[o], i = p->obs_pose[o], g = p->obs_intrinsic[o];
        if (missing[q]) continue;
        const double* R = &s.rot[9 * i];
        const double* C = &s.cen[3 * i];
        double pc[3];
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += R[3 * r + k] * (s.xyz[3 * q + k] - C[k]);
            pc[r] = acc;
        }
        double pix[2], J[2][3];
        if (!camera_project(p->intrinsics[g], s.f[g], s.k1[g], s.k2[g], pc, pix, J)) {
            missing[q] = 1;
            continue;
        }
        const double res = std::hypot(pix[0] - p->obs_pixel[2 * o], pix[1] - p->obs_pixel[2 * o + 1]);
        if (!std::isfinite(res)) {
            missing[q] = 1;
            continue;
        }
        double A[2][3];
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 3; ++c) A[r][c] = J[r][0] * R[c] + J[r][1] * R[3 + c] + J[r][2] * R[6 + c];
        const double w = loss_sqrt_weight(loss, res);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) info[q][3 * a + b] += w * w * (A[0][a] * A[0][b] + A[1][a] * A[1][b]);
    }
    // TrackGeometry at the solution (quarantine.rs:57-96)
    std::vector<double> offs(3 * p->n_observations), dist(p->n_observations);
    for (size_t o = 0; o < p->n_observations; ++o) {
        for (int k = 0; k < 3; ++k)
            offs[3 * o + k] = s.xyz[3 * p->obs_point[o] + k] - s.cen[3 * p->obs_pose[o] + k];
        dist[o] = std::sqrt(offs[3 * o] * offs[3 * o] + offs[3 * o + 1] * offs[3 * o + 1] + offs[3 * o + 2] * offs[3 * o + 2]);
    }
    double scene = 0.0;
    if (!dist.empty()) {
        std::nth_element(dist.begin(), dist.begin() + dist.size() / 2, dist.end());
        scene = dist[dist.size() / 2];
    }
    std::vector<TrackGeometry> geom(N, TrackGeometry(scene));
    for (size_t o = 0; o < p->n_observations; ++o) geom[p->obs_point[o]].push(&offs[3 * o]);
    std::vector<uint8_t> weak(N, 0);
    for (size_t i = 0; i < N; ++i) {
        if (s.quarantined.size() == N && s.quarantined[i]) {
            weak[i] = 1;
            continue;
        }
        if (point_is_fixed(p, i)) continue;
        if (geom[i].degenerate() || missing[i]) {
            weak[i] = 1;
            continue;
        }
        double a[3][3], ev[3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) a[r][c] = info[i][3 * r + c];
        sym3_eigenvalues(a, ev);
        const double hi = fmax_(fmax_(ev[0], ev[1]), ev[2]), lo = fmin_(fmin_(ev[0], ev[1]), ev[2]);
        weak[i] = !(lo > 1e-5 * hi);
    }
    return weak;
}

// The exact certificate and marginals at the final state: S of the undamped final assembly,
// factored with every pivot above the rank rule's scale-aware tolerance, and block_marginal
// (covariance.rs:562-587) of every free group by solves with unit right-hand sides.
static void certify_view(const View& v, const uint32_t* cross_off, bool exact, bool want, int32_t* obs,
                         bool* marginals, std::vector<double>& cov) {
    *obs = BA_OBSERVABILITY_NOT_ASSESSED;
    *marginals = false;
    cov.assign(9 * v.ng, NaN);
    const size_t D = v.D;
    if (!exact && !want) return;
    if (D == 0) {
        if (exact) *obs = BA_OBSERVABILITY_FULL_RANK;
        *marginals = want;
        return;
    }
    Assembly a = assemble(v, 0.0, true, cross_off);
    bool ok = a.valid && a.H.finite();
    std::vector<double> F;
    std::vector<double> tol_row(D, 0.0);
    if (ok) {
        // the rank rule: the reference certifies a reduced system by its spectral rank after
        // diagonal equilibration (direct.rs:342-457: A' = S A S with s_i = 1 / sqrt(max(A_ii,
        // 64 eps max_j A_jj)); deficient when an eigenvalue is at most 64 eps D lambda_max(A')).
        // Restated on the pivots of the LLT: pivot k of A' is d_k s_k^2, and the system is taken
        // as deficient when one is at most tau = 64 eps D g, g = max_i sum_j |A'_ij| (the
        // Gershgorin bound of lambda_max), i.e. when d_k <= tau max(A_kk, floor)
        const std::vector<double>& H = a.H.full;
        double maxd = 0.0;
        for (size_t i = 0; i < D; ++i) maxd = fmax_(maxd, std::fabs(H[i * D + i]));
        const double floor = 64.0 * EPS * maxd;
        std::vector<double> sc(D);
        for (size_t i = 0; i < D; ++i) sc[i] = 1.0 / std::sqrt(fmax_(H[i * D + i], floor));
        double g = 0.0;
        for (size_t i = 0; i < D; ++i) {
            double row = 0.0;
            for (size_t j = 0; j < D; ++j) row += std::fabs(j <= i ? H[i * D + j] : H[j * D + i]) * sc[j];
            g = fmax_(g, row * sc[i]);
        }
        const double tau = 64.0 * EPS * (double)D * g;
        for (size_t i = 0; i < D; ++i) tol_row[i] = tau / (sc[i] * sc[i]);
    }
    if (ok) {  // LLT of the lower triangle, every pivot above its tolerance
        F = a.H.full;
        for (size_t c = 0; c < D && ok; ++c) {
            double d = F[c * D + c];
            for (size_t k = 0; k < c; ++k) d -= F[c * D + k] * F[c * D + k];
            if (!(std::isfinite(d) && d > tol_row[c])) {
                ok = false;
                break;
            }
            const double l = std::sqrt(d);
            F[c * D + c] = l;
#pragma omp parallel for schedule(static) if (D - c > 256)
            for (size_t r = c + 1; r < D; ++r) {
                double x = F[r * D + c];
                for (size_t k = 0; k < c; ++k) x -= F[r * D + k] * F[c * D + k];
                F[r * D + c] = x / l;
            }
        }
    }
    if (exact) *obs = ok ? BA_OBSERVABILITY_FULL_RANK : BA_OBSERVABILITY_RANK_DEFICIENT;
    if (!ok || !want) return;
    const double w0 = loss_sqrt_weight(v.loss, 0.0);
    const double wn = w0 * w0;
    std::vector<double> x(D);
    bool finite = true;
    for (size_t g = 0; g < v.ng; ++g) {
        const size_t d = v.grp_dim[g], off = v.grp_off[g];
        if (!d) continue;
        double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (size_t col = 0; col < d; ++col) {
            std::fill(x.begin(), x.end(), 0.0);
            x[off + col] = 1.0;
            for (size_t r = 0; r < D; ++r) {  // L y = e
                double v2 = x[r];
                for (size_t k = 0; k < r; ++k) v2 -= F[r * D + k] * x[k];
                x[r] = v2 / F[r * D + r];
            }
            for (size_t r = D; r-- > 0;) {  // L^T x = y
                double v2 = x[r];
                for (size_t k = r + 1; k < D; ++k) v2 -= F[k * D + r] * x[k];
                x[r] = v2 / F[r * D + r];
            }
            for (size_t row = 0; row < d; ++row) m[row][col] = x[off + row] * wn;
        }
        for (size_t r = 0; r < d; ++r)  // symmetrize
            for (size_t c = r + 1; c < d; ++c) m[r][c] = m[c][r] = 0.5 * (m[r][c] + m[c][r]);
        for (size_t r = 0; r < 3; ++r)
            for (size_t c = 0; c < 3; ++c) {
                cov[9 * g + 3 * r + c] = m[r][c];
                finite = finite && std::isfinite(m[r][c]);
            }
    }
    if (!finite) {
        std::fill(cov.begin(), cov.end(), NaN);
        return;
    }
    *marginals = true;
}

static void certify_state(const Lowered& L, const State& st, double scene_scale, bool exact, bool want, int32_t* obs,
                          bool* marginals, std::vector<double>& cov) {
    View v = make_view(L, st, L.loss, L.gauge_camera, scene_scale);
    certify_view(v, L.cross_off.data(), exact, want, obs, marginals, cov);
}

// unconstrained_intrinsics (native/self_calibration.rs:255-303)
static std::vector<uint8_t> unconstrained(const ba_problem* p, const Lowered& L, const State& st, const double* cov) {
    std::vector<uint8_t> out(p->n_intrinsics, 0);
    for (size_t g = 0; g < L.grp_dim.size(); ++g) {
        const size_t d = L.grp_dim[g];
        if (!d) continue;
        if (!cov) {
            out[g] = 1;
            continue;
        }
        const ba_intrinsic& in = p->intrinsics[g];
        const double f = st.gp[3 * g];
        const double ax = fmax_(in.cx, (double)in.width - in.cx) / f, ay = fmax_(in.cy, (double)in.height - in.cy) / f;
        const double rr = ax * ax + ay * ay;
        const double scale[3] = {1.0, rr, rr * rr};
        for (size_t k = 0; k < d; ++k) {
            const double dev = std::sqrt(cov[9 * g + 3 * k + k]) * scale[k];
            if (std::isnan(dev) || (!in.certification_waived && dev > 0.01)) out[g] = 1;
        }
    }
    return out;
}

int bao_weak_points(const ba_problem* p, const ba_solution* s, uint8_t* weak) {
    SolutionO so;
    so.init(p);
    so.rot.assign(s->pose_rotation, s->pose_rotation + 9 * p->n_poses);
    so.cen.assign(s->pose_center, s->pose_center + 3 * p->n_poses);
    so.f.assign(s->intrinsic_focal, s->intrinsic_focal + p->n_intrinsics);
    so.k1.assign(s->intrinsic_k1, s->intrinsic_k1 + p->n_intrinsics);
    so.k2.assign(s->intrinsic_k2, s->intrinsic_k2 + p->n_intrinsics);
    so.xyz.assign(s->point_xyz, s->point_xyz + 3 * p->n_points);
    if (s->point_quarantined) so.quarantined.assign(s->point_quarantined, s->point_quarantined + p->n_points);
    const std::vector<uint8_t> w = solution_weak_points(p, so);
    std::memcpy(weak, w.data(), w.size());
    return BA_OK;
}

static bool direct_fits_o(uint64_t est, uint64_t limit) { return est <= limit / 5 * 4 + limit % 5 * 4 / 5; }

// The limits of the product's dense certificate (BA_CERTIFY_MAX_DIMENSION, BA_CERTIFY_MAX_TERMS),
// restated so that the oracle certifies exactly the solves the product does; beyond them nothing
// is certified and nothing is held.
static bool within_limits(const Lowered& L) {
    if (L.D > BA_CERTIFY_MAX_DIMENSION) return false;
    uint64_t terms = 0;
    for (size_t q = 0; q + 1 < L.off.size(); ++q) {
        const uint64_t k = 2ull * (L.off[q + 1] - L.off[q]);
        terms += L.point_fixed[q] ? k : k * k;
    }
    return terms <= BA_CERTIFY_MAX_TERMS;
}

// solve_once (native/mod.rs:318-335) with solve_self_calibrating (native/self_calibration.rs:223-249)
// and canonicalize_roundoff_non_descent.
static Err solve_once_o(const ba_problem* p, const ba_solve_config* cfg, ba_iteration_trace* trace, bool quarantine,
                        SolutionO& out) {
    const bool full = cfg->certification == BA_CERTIFY_FULL;
    Lowered L;
    Err e = lower(p, L, quarantine);
    if (e.code) return e;
    ba_execution_plan plan;
    plan_solver(L, cfg, plan);
    ba_solve_config planned = *cfg;
    planned.linear_solver = plan.linear_solver;
    const uint32_t trace_start = trace ? trace->count : 0;
    State st{L.R, L.t, L.gp, L.X};
    ba_solve_report r;
    e = refine(L, &planned, st, r, trace);
    if (e.code) return e;
    r.plan_reason = plan.reason;
    r.observability = BA_OBSERVABILITY_NOT_ASSESSED;
    out.init(p);
    std::vector<uint8_t> held(p->n_intrinsics, 0);
    size_t nheld = 0;
    auto certify = [&](const Lowered& LL, const ba_execution_plan& pl, const State& s, ba_solve_report& rr,
                       bool want, bool* mok, std::vector<double>& cv) {
        // exact certificate unless Auto chose PCG and retained + certificate bytes exceed 4/5 of the budget
        const bool auto_pcg = cfg->linear_solver == BA_SOLVER_AUTO && pl.linear_solver == BA_SOLVER_ITERATIVE;
        bool exact = true;
        if (auto_pcg) {
            const uint64_t nnz = pl.reduced_nonzeros ? pl.reduced_nonzeros : (uint64_t)LL.row_idx.size();
            exact = direct_fits_o(pl.estimated_retained_bytes + nnz * 32, pl.budget_bytes);
        }
        int32_t obs;
        certify_state(LL, s, rr.normalization_scale, exact, want && !auto_pcg, &obs, mok, cv);
        rr.observability = obs;
        if (obs == BA_OBSERVABILITY_FULL_RANK)
            rr.pose_rank = rr.pose_dimension = (uint32_t)LL.D;
        else if (obs == BA_OBSERVABILITY_RANK_DEFICIENT) {
            rr.pose_rank = 0;
            rr.pose_dimension = (uint32_t)LL.D;
        }
        if (rr.pose_dimension == 0) rr.pose_rank = 0;
        if (obs == BA_OBSERVABILITY_RANK_DEFICIENT || rr.weak_point_count > 0) rr.status = BA_STATUS_DEGENERATE;
        return auto_pcg;
    };
    if (full && within_limits(L)) {
        bool mok = false;
        std::vector<double> cv;
        const bool auto_pcg = certify(L, plan, st, r, true, &mok, cv);
        out.cov = cv;
        held = unconstrained(p, L, st, mok && !auto_pcg ? cv.data() : nullptr);
        for (auto h : held) nheld += h;
    }
    if (nheld == 0) {
        ba_solution v = out.view();
        raise_solution(p, L, st, &v);
        inlier_metrics(p, L, &v, &r.inlier_count, &r.final_inlier_rmse);
        out.rep = r;
    } else {
        Lowered L2;
        e = lower(p, L2, quarantine, &held);
        if (e.code) return e;
        ba_execution_plan plan2;
        plan_solver(L2, cfg, plan2);
        ba_solve_config planned2 = *cfg;
        planned2.linear_solver = plan2.linear_solver;
        if (trace) trace->count = trace_start;
        State st2{L2.R, L2.t, L2.gp, L2.X};
        ba_solve_report r2;
        e = refine(L2, &planned2, st2, r2, trace);
        if (e.code) return e;
        r2.plan_reason = plan2.reason;
        bool mok2;
        std::vector<double> cv2;
        if (within_limits(L2)) certify(L2, plan2, st2, r2, false, &mok2, cv2);
        ba_solution v = out.view();
        raise_solution(p, L2, st2, &v);
        inlier_metrics(p, L2, &v, &r2.inlier_count, &r2.final_inlier_rmse);
        out.rep = r2;
        L = std::move(L2);
    }
    if (L.calibrated) {
        // solve_calibrated reports the unobserved free intrinsics as unconstrained
        // (native/calibrated.rs:485-487, native/mod.rs:465-477); nothing is re-solved
        std::vector<uint8_t> seen(p->n_intrinsics, 0);
        for (size_t o = 0; o < p->n_observations; ++o) seen[p->obs_intrinsic[o]] = 1;
        for (size_t g = 0; g < p->n_intrinsics; ++g)
            if (!seen[g] && !intrinsic_is_fixed(p->intrinsics[g])) {
                held[g] = 1;
                nheld += 1;
            }
    }
    for (size_t g = 0; g < held.size(); ++g) out.unconstrained[g] = held[g];
    out.rep.unconstrained_intrinsics = (uint32_t)nheld;
    if (!full) return Err{};
    // canonicalize_roundoff_non_descent (solver.rs:1030-1101)
    ba_solve_report& rr = out.rep;
    if (p->gauge == BA_GAUGE_FIXED_BASELINE) {
        const Vec3 a = pose_C(p, p->gauge_first), b = pose_C(p, p->gauge_second);
        const double initial = std::sqrt(dist2(a, b));
        if (!std::isfinite(initial) ||
            std::fabs(initial - p->gauge_distance) > 512.0 * EPS * fmax_(initial, p->gauge_distance))
            return Err{};
    }
    const double increase = rr.final_cost - rr.initial_cost;
    const double tol = 2.0 * std::sqrt(EPS) * fmax_(std::fabs(rr.initial_cost), 1.0);
    const bool conv = rr.status <= BA_STATUS_CONVERGED_COST;
    if (!(increase > 0.0 && increase <= tol && conv && rr.observability != BA_OBSERVABILITY_RANK_DEFICIENT))
        return Err{};
    SolutionO input;
    input.init(p);
    input.rot.assign(p->pose_rotation, p->pose_rotation + 9 * p->n_poses);
    input.cen.assign(p->pose_center, p->pose_center + 3 * p->n_poses);
    for (size_t g = 0; g < p->n_intrinsics; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        input.f[g] = in.fx;
        input.k1[g] = (in.model == BA_CAMERA_BROWN_K1 || in.model == BA_CAMERA_BROWN_K1K2 || in.model == BA_CAMERA_BROWN_K1K2K3) ? in.k1 : 0.0;
        input.k2[g] = (in.model == BA_CAMERA_BROWN_K1K2 || in.model == BA_CAMERA_BROWN_K1K2K3) ? in.k2 : 0.0;
    }
    input.xyz.assign(p->point_xyz, p->point_xyz + 3 * p->n_points);
    const bool same = input.rot == out.rot && input.cen == out.cen && input.f == out.f && input.k1 == out.k1 &&
                      input.k2 == out.k2 && input.xyz == out.xyz;
    if (!same) {
        out.rot = input.rot;
        out.cen = input.cen;
        out.f = input.f;
        out.k1 = input.k1;
        out.k2 = input.k2;
        out.xyz = input.xyz;
        rr.observability = BA_OBSERVABILITY_NOT_ASSESSED;
        rr.pose_rank = rr.pose_dimension = 0;
        rr.weak_point_count = 0;
    }
    rr.final_cost = rr.initial_cost;
    rr.final_rmse = rr.initial_rmse;
    rr.status = BA_STATUS_SKIPPED;
    ba_solution v = input.view();
    inlier_metrics(p, L, &v, &rr.inlier_count, &rr.final_inlier_rmse);
    return Err{};
}

static void continued_from(ba_solve_report& later, const ba_solve_report& first) {  // solver.rs:527-538
    later.initial_cost = first.initial_cost;
    later.initial_rmse = first.initial_rmse;
    later.iterations += first.iterations;
    later.accepted_steps += first.accepted_steps;
    later.rejected_steps += first.rejected_steps;
    later.rejected_linearization += first.rejected_linearization;
    later.rejected_step_solve += first.rejected_step_solve;
    later.rejected_candidate += first.rejected_candidate;
    later.rejected_insufficient_reduction += first.rejected_insufficient_reduction;
    later.pcg_iterations += first.pcg_iterations;
    later.weak_point_resolves += first.weak_point_resolves;
}

// solve_with_quarantine (native/mod.rs:252-316)
static Err solve_with_quarantine_o(const ba_problem* p, const ba_solve_config* cfg, ba_iteration_trace* trace,
                                   SolutionO& result) {
    Lowered L0;
    Err e = lower(p, L0);  // the input pass and every lowering check
    if (e.code) return e;
    const size_t N = p->n_points;
    size_t quarantined = 0;
    for (auto q : L0.quarantined) quarantined += q;
    OwnedProblemO working_own;
    const ba_problem* working = p;
    std::vector<size_t> kept(N);
    for (size_t i = 0; i < N; ++i) kept[i] = i;
    if (quarantined) {
        restrict_problem(p, nullptr, L0.quarantined, working_own, kept);
        working = &working_own.p;
    }
    const bool changes_scale = p->gauge == BA_GAUGE_FIXED_BASELINE && std::fabs(L0.rebase.scale - 1.0) > 1e-9;
    SolutionO sol;
    e = solve_once_o(working, cfg, trace, false, sol);
    if (e.code) return e;
    ba_solve_config retry = *cfg;
    retry.max_iterations = cfg->max_iterations < 8 ? cfg->max_iterations : 8;  // RESOLVE_ITERATIONS
    OwnedProblemO owned[2];
    const ba_problem* current = working;
    bool have_origin = false;
    ba_solve_report origin{};
    uint32_t rounds = 0;
    for (int round = 0; round < 4; ++round) {  // MAX_RESOLVES
        const ba_solve_report& rr = sol.rep;
        const bool needs = rr.weak_point_count > 0 || rr.observability == BA_OBSERVABILITY_RANK_DEFICIENT ||
                           rr.status == BA_STATUS_DEGENERATE;
        if (!needs) break;
        std::vector<uint8_t> weak = solution_weak_points(current, sol);
        size_t n = 0;
        for (size_t i = 0; i < weak.size(); ++i) {
            if (point_is_fixed(current, i)) weak[i] = 0;
            n += weak[i];
        }
        if (!n) break;
        OwnedProblemO& next = owned[round % 2];
        std::vector<size_t> next_kept;
        restrict_problem(current, &sol, weak, next, next_kept);
        SolutionO resolved;
        if (solve_once_o(&next.p, &retry, trace, false, resolved).code) break;
        std::vector<size_t> composed(next_kept.size());
        for (size_t i = 0; i < next_kept.size(); ++i) composed[i] = kept[next_kept[i]];
        kept.swap(composed);
        quarantined += n;
        if (have_origin) continued_from(sol.rep, origin);
        origin = sol.rep;
        have_origin = true;
        sol = std::move(resolved);
        current = &next.p;
        ++rounds;
    }
    if (have_origin) continued_from(sol.rep, origin);
    sol.rep.weak_point_resolves = rounds;
    if (changes_scale && quarantined > 0)
        return fail(BA_ERR_UNSUPPORTED, "native BA cannot change its physical baseline while holding out points");
    if (kept.size() == N) {
        result = std::move(sol);
        result.rep.quarantined_points = 0;
        return Err{};
    }
    result.init(p);  // Solution::lifted (solver.rs:891-935)
    result.rot = sol.rot;
    result.cen = sol.cen;
    result.f = sol.f;
    result.k1 = sol.k1;
    result.k2 = sol.k2;
    result.cov = sol.cov;
    result.unconstrained = sol.unconstrained;
    result.rep = sol.rep;
    result.xyz.assign(p->point_xyz, p->point_xyz + 3 * N);
    std::fill(result.quarantined.begin(), result.quarantined.end(), 1);
    for (size_t i = 0; i < kept.size(); ++i) {
        for (int k = 0; k < 3; ++k) result.xyz[3 * kept[i] + k] = sol.xyz[3 * i + k];
        result.quarantined[kept[i]] = 0;
    }
    result.rep.quarantined_points = (uint32_t)quarantined;
    result.rep.point_count = N;
    return Err{};
}

static void copy_solution(const SolutionO& s, ba_solution* out) {
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(out->pose_rotation, s.rot.data(), s.rot.size() * 8);
    cp(out->pose_center, s.cen.data(), s.cen.size() * 8);
    cp(out->intrinsic_focal, s.f.data(), s.f.size() * 8);
    cp(out->intrinsic_k1, s.k1.data(), s.k1.size() * 8);
    cp(out->intrinsic_k2, s.k2.data(), s.k2.size() * 8);
    cp(out->point_xyz, s.xyz.data(), s.xyz.size() * 8);
    cp(out->point_quarantined, s.quarantined.data(), s.quarantined.size());
    cp(out->intrinsic_unconstrained, s.unconstrained.data(), s.unconstrained.size());
    cp(out->intrinsic_covariance, s.cov.data(), s.cov.size() * 8);
}

int bao_prepare(const ba_problem* p, ba_native_sizes* sizes, ba_native_problem* n, char* msg, size_t len) {
    Err e = validate(p);
    if (e.code) {
        copy_msg(e, msg, len);
        return e.code;
    }
    Lowered L;
    e = lower(p, L);
    if (e.code) {
        copy_msg(e, msg, len);
        return e.code;
    }
    size_t nc = L.cam_source.size(), ng = L.gp.size() / 3, np = L.kept_points.size(), no = L.obs_cam.size();
    if (sizes) {
        sizes->n_cameras = nc;
        sizes->n_groups = ng;
        sizes->n_points = np;
        sizes->n_observations = no;
        sizes->dimension = L.D;
        sizes->n_point_parameters = L.params.size();
        sizes->n_pattern_entries = L.row_idx.size();
        size_t nq = 0;
        for (auto q : L.quarantined) nq += q;
        sizes->n_quarantined = nq;
    }
    if (!n) return BA_OK;
    std::memcpy(n->camera_rotation, L.R.data(), 9 * nc * sizeof(double));
    std::memcpy(n->camera_translation, L.t.data(), 3 * nc * sizeof(double));
    std::memcpy(n->group_params, L.gp.data(), 3 * ng * sizeof(double));
    std::memcpy(n->camera_group, L.cam_group.data(), nc * 4);
    std::memcpy(n->camera_source, L.cam_source.data(), nc * 4);
    std::memcpy(n->point_xyz, L.X.data(), 3 * np * sizeof(double));
    std::memcpy(n->point_source, L.kept_points.data(), np * 4);
    std::memcpy(n->obs_camera, L.obs_cam.data(), no * 4);
    std::memcpy(n->obs_source, L.obs_src.data(), no * 4);
    std::memcpy(n->obs_uv, L.uv.data(), 2 * no * sizeof(double));
    std::memcpy(n->point_offsets, L.off.data(), (np + 1) * 4);
    std::memcpy(n->camera_fixed, L.cam_fixed.data(), nc);
    std::memcpy(n->point_fixed, L.point_fixed.data(), np);
    std::memcpy(n->camera_offset, L.cam_off.data(), nc * 4);
    std::memcpy(n->camera_dim, L.cam_dim.data(), nc * 4);
    std::memcpy(n->group_offset, L.grp_off.data(), ng * 4);
    std::memcpy(n->group_dim, L.grp_dim.data(), ng * 4);
    std::memcpy(n->cross_offsets, L.cross_off.data(), (np + 1) * 4);
    std::memcpy(n->point_parameters, L.params.data(), L.params.size() * 4);
    std::memcpy(n->pattern_col_ptr, L.col_ptr.data(), L.col_ptr.size() * 4);
    std::memcpy(n->pattern_row_idx, L.row_idx.data(), L.row_idx.size() * 4);
    n->gauge_camera = L.gauge_camera;
    n->scene_scale = scene_scale_of(L, L.X.data(), L.t.data());
    return BA_OK;
}

int bao_solve(const ba_problem* p, const ba_solve_config* cfg, ba_solution* sol, ba_solve_report* rep,
              ba_iteration_trace* trace, char* msg, size_t len) {
    Err e = validate_config(cfg);
    if (!e.code) e = validate(p);
    if (e.code) {
        copy_msg(e, msg, len);
        return e.code;
    }
    if (trace) trace->count = 0;
    SolutionO out;
    if (cfg->certification == BA_CERTIFY_FULL) {
        e = solve_with_quarantine_o(p, cfg, trace, out);
    } else {  // the LM loop alone, input quarantine in the lowering
        e = solve_once_o(p, cfg, trace, true, out);
        if (!e.code) {
            size_t nq = 0;
            for (auto q : out.quarantined) nq += q;
            out.rep.quarantined_points = (uint32_t)nq;
        }
    }
    if (e.code) {
        copy_msg(e, msg, len);
        return e.code;
    }
    copy_solution(out, sol);
    if (rep) *rep = out.rep;
    return BA_OK;
}

int bao_plan(const ba_problem* p, const ba_solve_config* cfg, ba_execution_plan* plan, char* msg, size_t len) {
    Err e = validate_config(cfg);
    if (!e.code) e = validate(p);
    Lowered L;
    if (!e.code) e = lower(p, L);
    if (e.code) {
        copy_msg(e, msg, len);
        return e.code;
    }
    plan_solver(L, cfg, *plan);
    return BA_OK;
}

int bao_linearize(const ba_device_problem* dp, double* residual, double* pj, double* cj, double* cost,
                  uint8_t* valid) {
    View v = view_from(dp);
    for (size_t q = 0; q < v.np; ++q)
        for (uint32_t o = v.off[q]; o < v.off[q + 1]; ++o) {
            ObsLin ol;
            bool ok = linearize_obs(v, o, v.X + 3 * q, ol);
            valid[o] = ok;
            if (!ok) {
                for (int k = 0; k < 2; ++k) residual[2 * o + k] = 0.0;
                for (int k = 0; k < 6; ++k) pj[6 * o + k] = 0.0;
                for (int k = 0; k < 18; ++k) cj[18 * o + k] = 0.0;
                cost[o] = 0.0;
                continue;
            }
            for (int r = 0; r < 2; ++r) {
                residual[2 * o + r] = ol.sw * ol.res[r];
                for (int c = 0; c < 3; ++c) pj[6 * o + 3 * r + c] = ol.sw * ol.pj[r][c];
                // fixed slots: pose columns 0..5, intrinsic columns 6..8, zero beyond each run
                for (int c = 0; c < 9; ++c) cj[18 * o + 9 * r + c] = 0.0;
                for (int c = 0; c < ol.pose_dim; ++c) cj[18 * o + 9 * r + c] = ol.sw * ol.cj[r][c];
                for (int c = 0; c < ol.intr_dim; ++c) cj[18 * o + 9 * r + 6 + c] = ol.sw * ol.cj[r][ol.pose_dim + c];
            }
            cost[o] = ol.cost;
        }
    return BA_OK;
}

// Cross offsets of a view in the order of radial_point_parameters (pattern.rs:306-379).
static std::vector<uint32_t> cross_offsets(const View& v) {
    std::vector<uint32_t> cross_off(v.np + 1, 0);
    std::vector<uint8_t> marks(v.D, 0);
    std::vector<uint32_t> act;
    for (size_t q = 0; q < v.np; ++q) {
        act.clear();
        if (!(v.point_fixed && v.point_fixed[q]))
            for (uint32_t o = v.off[q]; o < v.off[q + 1]; ++o) {
                uint32_t c = v.obs_cam[o], g = v.cam_group[c];
                for (uint32_t k = 0; k < v.cam_dim[c]; ++k)
                    if (!marks[v.cam_off[c] + k]) {
                        marks[v.cam_off[c] + k] = 1;
                        act.push_back(v.cam_off[c] + k);
                    }
                for (uint32_t k = 0; k < v.grp_dim[g]; ++k)
                    if (!marks[v.grp_off[g] + k]) {
                        marks[v.grp_off[g] + k] = 1;
                        act.push_back(v.grp_off[g] + k);
                    }
            }
        for (auto pr : act) marks[pr] = 0;
        cross_off[q + 1] = cross_off[q] + (uint32_t)act.size();
    }
    return cross_off;
}

int bao_certify(const ba_device_problem* dp, int32_t* observability, double* covariance) {
    View v = view_from(dp);
    const std::vector<uint32_t> cross_off = cross_offsets(v);
    int32_t obs;
    bool ok;
    std::vector<double> cov;
    certify_view(v, cross_off.data(), true, true, &obs, &ok, cov);
    if (observability) *observability = obs;
    if (covariance) std::memcpy(covariance, cov.data(), cov.size() * sizeof(double));
    return ok ? BA_OK : BA_ERR_NUMERICAL;
}

int bao_assemble(const ba_device_problem* dp, double lambda, double* schur, double* gradient, double* full_gradient,
                 double* diagonal, double* point_inverse, double* inverse_gradient, double* inverse_cross,
                 ba_schur_scalars* sc) {
    View v = view_from(dp);
    const std::vector<uint32_t> cross_off = cross_offsets(v);
    Assembly a = assemble(v, lambda, true, cross_off.data());
    size_t D = v.D;
    for (size_t r = 0; r < D; ++r)
        for (size_t c = 0; c < D; ++c) schur[r * D + c] = a.H.get(r, c);
    std::memcpy(gradient, a.g.data(), D * sizeof(double));
    std::memcpy(full_gradient, a.fg.data(), D * sizeof(double));
    std::memcpy(diagonal, a.diag.data(), D * sizeof(double));
    if (point_inverse) std::memcpy(point_inverse, a.inv_packed.data(), a.inv_packed.size() * sizeof(double));
    if (inverse_gradient) std::memcpy(inverse_gradient, a.inv_grad.data(), a.inv_grad.size() * sizeof(double));
    if (inverse_cross) std::memcpy(inverse_cross, a.inv_cross.data(), a.inv_cross.size() * sizeof(double));
    if (sc) {
        sc->cost = a.cost;
        sc->squared_residual = a.sq;
        sc->point_model_reduction = a.pmr;
        sc->scaled_gradient_inf_norm = a.sgin;
        sc->minimum_point_rcond = a.min_rcond;
        sc->weak_point_count = (uint32_t)a.weak;
        sc->valid = a.valid;
    }
    return BA_OK;
}

int bao_pcg(const ba_device_problem* dp, double lambda, double tol, uint32_t max_it, double* step, uint32_t* iters,
            double* rel, int32_t* accepted) {
    View v = view_from(dp);
    std::vector<uint32_t> cross_off(v.np + 1, 0), params;
    {
        std::vector<uint8_t> marks(v.D, 0);
        std::vector<uint32_t> act;
        for (size_t q = 0; q < v.np; ++q) {
            act.clear();
            if (!(v.point_fixed && v.point_fixed[q]))
                for (uint32_t o = v.off[q]; o < v.off[q + 1]; ++o) {
                    uint32_t c = v.obs_cam[o], g = v.cam_group[c];
                    for (uint32_t k = 0; k < v.cam_dim[c]; ++k)
                        if (!marks[v.cam_off[c] + k]) {
                            marks[v.cam_off[c] + k] = 1;
                            act.push_back(v.cam_off[c] + k);
                        }
                    for (uint32_t k = 0; k < v.grp_dim[g]; ++k)
                        if (!marks[v.grp_off[g] + k]) {
                            marks[v.grp_off[g] + k] = 1;
                            act.push_back(v.grp_off[g] + k);
                        }
                }
            for (auto pr : act) {
                marks[pr] = 0;
                params.push_back(pr);
            }
            cross_off[q + 1] = (uint32_t)params.size();
        }
    }
    Assembly a = assemble(v, lambda, false, cross_off.data());
    Retained ret;
    std::vector<BlockFactor> factors;
    if (!a.valid || !build_retained(v, a, cross_off.data(), params.data(), ret, factors)) {
        *iters = 0;
        *rel = NaN;
        *accepted = 0;
        return BA_OK;
    }
    size_t D = v.D;
    auto pre = [&](const double* r, double* z) {
        std::fill(z, z + D, 0.0);
        for (size_t b = 0; b < ret.nblocks; ++b)
            if (ret.bdim[b] > 0) factors[b].solve(r + ret.boff[b], z + ret.boff[b], ret.bdim[b]);
    };
    auto op = [&](const double* x, double* y) { ret.apply(x, y); };
    std::vector<double> x, product;
    PcgResult r = pcg_solve(a.g.data(), D, max_it, tol, op, pre, x, product);
    std::memcpy(step, x.data(), D * sizeof(double));
    *iters = (uint32_t)r.terms;
    *rel = r.rel;
    *accepted = r.accepted;
    return BA_OK;
}

int bao_cholesky_solve(size_t n, const double* a, const double* b, double* x) {
    std::vector<double> F(a, a + n * n);
    return dense_cholesky_solve(F, n, b, x) ? 1 : 0;
}

int bao_radial_linearization(const double cam[3], const double obs[2], const double pt[3], double res[2], double J[6],
                             double Ji[6]) {
    double j[2][3], ji[2][3];
    if (!radial_linearization(Radial{cam[0], cam[1], cam[2]}, obs, pt, res, j, ji)) return 0;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
            J[3 * r + c] = j[r][c];
            Ji[3 * r + c] = ji[r][c];
        }
    return 1;
}
int bao_radial_projection(const double cam[3], const double obs[2], const double pt[3], double res[2], double J[6]) {
    double j[2][3];
    if (!radial_projection(Radial{cam[0], cam[1], cam[2]}, obs, pt, res, j)) return 0;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) J[3 * r + c] = j[r][c];
    return 1;
}
int bao_radial_bearing(const double cam[3], const double m[2], double b[3]) {
    return radial_bearing(Radial{cam[0], cam[1], cam[2]}, m, b) ? 1 : 0;
}
double bao_first_positive_radial_turn(const double cam[3]) {
    double r;
    return first_positive_radial_turn(Radial{cam[0], cam[1], cam[2]}, &r) ? r : NaN;
}
// Robust reprojection objective of a BAL-style scene and its gradient (differentiation form):
// per observation the radial residual and Jacobians of model.rs:26-129, the robust weight of
// robust.rs:8-44 (d rho/dr = w r with w = sqrt_weight^2), the pose motion of schur/point.rs:730-762
// (d X_c / d omega = -[R X]x for the left perturbation exp(omega) R) and, for the angle-axis
// parameter itself, the chain rule through the SO(3) left Jacobian
// J_l(w) = I + (1 - cos t)/t^2 [w]x + (t - sin t)/t^3 [w]x^2.  Sums in input order.
int bao_cost_grad(size_t nc, size_t np, size_t no, const double* cams, const double* pts, const uint32_t* oc,
                  const uint32_t* op, const double* uv, int32_t loss, double scale, double* cost,
                  uint32_t* invalid, double* gcam, double* gpt) {
    std::vector<Mat3> R(nc);
    for (size_t c = 0; c < nc; ++c) R[c] = exp_so3(cams + 9 * c);
    std::vector<double> gomega(3 * nc, 0.0);
    if (gcam) std::fill(gcam, gcam + 9 * nc, 0.0);
    if (gpt) std::fill(gpt, gpt + 3 * np, 0.0);
    double total = 0.0;
    uint32_t bad = 0;
    for (size_t i = 0; i < no; ++i) {
        const uint32_t c = oc[i], p = op[i];
        const double* q = cams + 9 * c;
        const double* X = pts + 3 * p;
        double RX[3];
        for (int r = 0; r < 3; ++r) RX[r] = R[c][r][0] * X[0] + R[c][r][1] * X[1] + R[c][r][2] * X[2];
        const double Xc[3] = {RX[0] + q[3], RX[1] + q[4], RX[2] + q[5]};
        double res[2], J[2][3], Ji[2][3];
        if (!radial_linearization(Radial{q[6], q[7], q[8]}, uv + 2 * i, Xc, res, J, Ji)) {
            ++bad;
            continue;
        }
        const double nrm = std::sqrt(res[0] * res[0] + res[1] * res[1]);
        total += loss_cost(Loss{loss, scale}, nrm);
        if (!gcam) continue;
        const double sw = loss_sqrt_weight(Loss{loss, scale}, nrm);
        const double psi[2] = {sw * sw * res[0], sw * sw * res[1]};
        double e[3];
        for (int k = 0; k < 3; ++k) e[k] = J[0][k] * psi[0] + J[1][k] * psi[1];
        for (int k = 0; k < 3; ++k) gpt[3 * p + k] += R[c][0][k] * e[0] + R[c][1][k] * e[1] + R[c][2][k] * e[2];
        gomega[3 * c + 0] += RX[1] * e[2] - RX[2] * e[1];
        gomega[3 * c + 1] += RX[2] * e[0] - RX[0] * e[2];
        gomega[3 * c + 2] += RX[0] * e[1] - RX[1] * e[0];
        for (int k = 0; k < 3; ++k) gcam[9 * c + 3 + k] += e[k];
        gcam[9 * c + 6] += (Ji[0][0] * psi[0] + Ji[1][0] * psi[1]) / q[6];
        gcam[9 * c + 7] += Ji[0][1] * psi[0] + Ji[1][1] * psi[1];
        gcam[9 * c + 8] += Ji[0][2] * psi[0] + Ji[1][2] * psi[1];
    }
    if (gcam)
        for (size_t c = 0; c < nc; ++c) {
            const double* w = cams + 9 * c;
            const double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2], t = std::sqrt(t2);
            double A, B;
            if (t < 1e-4) {
                A = 0.5 - t2 / 24.0;
                B = 1.0 / 6.0 - t2 / 120.0;
            } else {
                A = (1.0 - std::cos(t)) / t2;
                B = (t - std::sin(t)) / (t2 * t);
            }
            const double K[3][3] = {{0, -w[2], w[1]}, {w[2], 0, -w[0]}, {-w[1], w[0], 0}};
            double Jl[3][3];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) {
                    double k2 = K[r][0] * K[0][k] + K[r][1] * K[1][k] + K[r][2] * K[2][k];
                    Jl[r][k] = (r == k ? 1.0 : 0.0) + A * K[r][k] + B * k2;
                }
            for (int k = 0; k < 3; ++k)
                gcam[9 * c + k] =
                    Jl[0][k] * gomega[3 * c] + Jl[1][k] * gomega[3 * c + 1] + Jl[2][k] * gomega[3 * c + 2];
        }
    if (cost) *cost = total;
    if (invalid) *invalid = bad;
    return BA_OK;
}
double bao_loss_cost(int32_t loss, double scale, double r) { return loss_cost(Loss{loss, scale}, r); }
double bao_loss_sqrt_weight(int32_t loss, double scale, double r) { return loss_sqrt_weight(Loss{loss, scale}, r); }
void bao_exp_so3(const double v[3], double R[9]) {
    Mat3 m = exp_so3(v);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[3 * r + c] = m[r][c];
}
void bao_s2_tangent_basis(const double t[3], double basis[6]) { s2_tangent_basis(t, basis, basis + 3); }
void bao_retract_s2(const double t[3], const double basis[6], const double d[2], double out[3]) {
    double b[2][3];
    for (int k = 0; k < 3; ++k) {
        b[0][k] = basis[k];
        b[1][k] = basis[3 + k];
    }
    retract_s2(t, b, d, out);
}
int bao_point_block_factor(const double m[9], double damping, double inverse[6]) {
    Mat3 M;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = m[3 * r + c];
    PointFactor f;
    if (!point_factor_new(M, damping, f)) return 0;
    std::memcpy(inverse, f.inv, sizeof(f.inv));
    return 1;
}
double bao_point_block_rcond(const double m[9]) {
    Mat3 M;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = m[3 * r + c];
    PointFactor f;
    if (!point_factor_new(M, 0.0, f)) return 0.0;
    return point_rcond(f, M);
}
int bao_track_geometry(size_t count, const double* offsets, double scene_length) {
    TrackGeometry g(scene_length);
    for (size_t i = 0; i < count; ++i) g.push(offsets + 3 * i);
    return (g.degenerate() ? 1 : 0) | (g.collapsed() ? 2 : 0);
}
int bao_nielsen_accept(double* value, double* nu, double actual, double predicted, double* gain) {
    Nielsen n{*value};
    n.nu = *nu;
    bool a = n.accept(actual, predicted);
    *value = n.value;
    *nu = n.nu;
    if (gain) *gain = n.has_gain ? n.gain : NaN;
    return a ? 1 : 0;
}
void bao_nielsen_reject(double* value, double* nu) {
    Nielsen n{*value};
    n.nu = *nu;
    n.reject();
    *value = n.value;
    *nu = n.nu;
}
void bao_stages(int32_t loss, uint32_t max_it, int32_t lo[3], uint32_t io[3]) {
    Stage s[3];
    stages(Loss{loss, 1.0}, max_it, s);
    for (int i = 0; i < 3; ++i) {
        lo[i] = s[i].loss.kind;
        io[i] = (uint32_t)s[i].iters;
    }
}
double bao_native_cost(const ba_device_problem* dp, size_t n_priors, const uint32_t* prior_group,
                       const double* prior_mean, const double* prior_sqrt_info) {
    View v = view_from(dp);
    double cost = 0.0;
    for (size_t q = 0; q < v.np; ++q)
        for (uint32_t o = v.off[q]; o < v.off[q + 1]; ++o) {
            ObsLin ol;
            if (!linearize_obs(v, o, v.X + 3 * q, ol)) return NaN;
            cost += ol.cost;
        }
    // schur::prior_cost intrinsic term, schur/priors.rs:320-329
    double prior = 0.0;
    for (size_t i = 0; i < n_priors; ++i) {
        Radial c = v.group(prior_group[i]);
        double r[3] = {std::log(c.focal / prior_mean[3 * i]), c.k1 - prior_mean[3 * i + 1],
                       c.k2 - prior_mean[3 * i + 2]};
        const double* S = prior_sqrt_info + 9 * i;
        double w[3];
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += S[3 * k + j] * r[j];
            w[k] = s;
        }
        prior += 0.5 * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    }
    return cost + prior;
}

}  // extern "C"
