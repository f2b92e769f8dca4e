// Host checks of the reduced-system index lists (csrc/ba_host.cpp), built and run by
// tests/test_host_lists.py; prints "ok" or the failed check.  The builders count and place their
// terms per point range on several threads; here they are compared with one straightforward pass
// that collects the same terms into ordered maps:
//   build_pair_lists: the local pairs (diagonal blocks and pose-intrinsic edges), with and
//     without their term lists, and every pair of the pattern, on shared intrinsic groups,
//     repeated observations of a point by one camera, fixed points and a fixed pose;
//   build_schur_lists: the camera pairs of the explicit reduced system on a scene with one
//     camera per intrinsic group.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <utility>
#include <vector>

#include "../openmvg-rs_amd/csrc/ba_host.hpp"
#include "../openmvg-rs_amd/csrc/ba_kernels.hpp"

#define CHECK(cond)                                                  \
    do {                                                             \
        if (!(cond)) {                                               \
            std::printf("failed: %s (line %d)\n", #cond, __LINE__);  \
            std::exit(1);                                            \
        }                                                            \
    } while (0)

namespace {

struct Scene {
    std::vector<double> rot, cen, xyz, pix;
    std::vector<uint8_t> vr, vt, fixed;
    std::vector<ba_intrinsic> intr;
    std::vector<uint32_t> pose, group, point;
    ba_problem p{};
};

// Cameras on the plane z = -10 looking along +z at points in the unit cube.  shared: cameras
// share intrinsic groups (camera c in group c % 5) and a point may be seen twice by one camera;
// otherwise one camera per group and at most one observation per point and camera.
void make_scene(Scene& s, size_t P, size_t N, bool shared, uint32_t seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const size_t G = shared ? 5 : P;
    for (size_t c = 0; c < P; ++c) {
        const double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        s.rot.insert(s.rot.end(), R, R + 9);
        s.cen.push_back(4.0 * u(rng));
        s.cen.push_back(4.0 * u(rng));
        s.cen.push_back(-10.0);
    }
    s.vr.assign(P, 1);
    s.vt.assign(P, 1);
    s.vr[0] = s.vt[0] = 0;  // a fixed pose: the gauge of the fixed points
    for (size_t g = 0; g < G; ++g) {
        ba_intrinsic in{};
        in.model = BA_CAMERA_BROWN_K1K2;
        in.fx = in.fy = 800.0 + g;
        in.cx = 320.0;
        in.cy = 240.0;
        in.width = 640;
        in.height = 480;
        in.var_focal = 1;
        in.var_radial = 1;
        in.certification_waived = 1;
        s.intr.push_back(in);
    }
    s.fixed.assign(N, 0);
    for (size_t q = 0; q < N; ++q) {
        for (int k = 0; k < 3; ++k) s.xyz.push_back(u(rng));
        s.fixed[q] = rng() % 23 == 0;
        const size_t k = 2 + rng() % 5;
        std::vector<uint32_t> seen;
        for (size_t m = 0; m < k; ++m) {
            uint32_t c = (uint32_t)(rng() % P);
            bool again = false;
            for (uint32_t x : seen) again |= x == c;
            if (again && !shared) continue;
            seen.push_back(c);
            s.pose.push_back(c);
            s.group.push_back((uint32_t)(c % G));
            s.point.push_back((uint32_t)q);
            s.pix.push_back(300.0 + 50.0 * u(rng));
            s.pix.push_back(200.0 + 50.0 * u(rng));
        }
    }
    // observations in a shuffled input order
    std::vector<size_t> perm(s.pose.size());
    for (size_t o = 0; o < perm.size(); ++o) perm[o] = o;
    std::shuffle(perm.begin(), perm.end(), rng);
    std::vector<uint32_t> a(perm.size()), b(perm.size()), c(perm.size());
    std::vector<double> px(2 * perm.size());
    for (size_t o = 0; o < perm.size(); ++o) {
        a[o] = s.pose[perm[o]];
        b[o] = s.group[perm[o]];
        c[o] = s.point[perm[o]];
        px[2 * o] = s.pix[2 * perm[o]];
        px[2 * o + 1] = s.pix[2 * perm[o] + 1];
    }
    s.pose = a;
    s.group = b;
    s.point = c;
    s.pix = px;
    ba_problem& p = s.p;
    p.n_poses = P;
    p.pose_rotation = s.rot.data();
    p.pose_center = s.cen.data();
    p.pose_var_rotation = s.vr.data();
    p.pose_var_translation = s.vt.data();
    p.n_intrinsics = G;
    p.intrinsics = s.intr.data();
    p.n_points = N;
    p.point_xyz = s.xyz.data();
    p.point_fixed = s.fixed.data();
    p.n_observations = s.pose.size();
    p.obs_pose = s.pose.data();
    p.obs_intrinsic = s.group.data();
    p.obs_point = s.point.data();
    p.obs_pixel = s.pix.data();
    p.residual_space = BA_RESIDUAL_PIXELS;
    p.robust_loss = BA_LOSS_SQUARED;
    p.loss_scale = 1.0;
    p.gauge = BA_GAUGE_FREE;
}

using PairMap = std::map<std::pair<uint32_t, uint32_t>, std::vector<std::pair<uint32_t, uint32_t>>>;

// The terms of the reduced system, point by point, into a map ordered by (A, B).
PairMap expected_pairs(const ba::Lowered& L, bool all) {
    PairMap m;
    const uint32_t P = (uint32_t)L.n_cameras();
    auto blocks = [&](uint32_t o) {  // the pose block, then the intrinsic block
        std::vector<uint32_t> v;
        const uint32_t c = L.obs_cam[o], g = L.cam_group[c];
        if (L.cam_dim[c]) v.push_back(c);
        if (L.grp_dim[g]) v.push_back(P + g);
        return v;
    };
    for (size_t q = 0; q < L.n_points(); ++q) {
        const bool fixed = L.point_fixed[q];
        for (uint32_t i = L.point_off[q]; i < L.point_off[q + 1]; ++i)
            for (uint32_t j = L.point_off[q]; j < L.point_off[q + 1]; ++j) {
                if (fixed && i != j) continue;
                for (uint32_t A : blocks(i))
                    for (uint32_t B : blocks(j)) {
                        if (A < B) continue;
                        uint32_t flags = (fixed ? 0u : ba::TERM_SCHUR) | (i == j ? ba::TERM_U : 0u);
                        if (!all) {
                            const bool edge = i == j && A >= P && B < P && L.cam_group[B] == A - P;
                            if (A != B && !edge) continue;
                            if (A != B) flags = ba::TERM_U;
                        }
                        if (flags) m[{A, B}].push_back({i, j | flags});
                    }
            }
    }
    return m;
}

void check_pairs(const ba::Lowered& L, const ba::BlockLists& B, bool all, bool with_terms) {
    const PairMap want = expected_pairs(L, all);
    const ba::PairLists got = ba::build_pair_lists(L, B, all, with_terms);
    CHECK(got.block_a.size() == want.size() && got.block_b.size() == want.size());
    CHECK(got.term_off.size() == want.size() + 1 && got.value_off.size() == want.size() + 1);
    CHECK(got.term_off[0] == 0 && got.value_off[0] == 0);
    size_t k = 0, t = 0;
    uint64_t v = 0;
    for (const auto& e : want) {
        CHECK(got.block_a[k] == e.first.first && got.block_b[k] == e.first.second);
        t += e.second.size();
        v += (uint64_t)B.block_dim[e.first.first] * B.block_dim[e.first.second];
        CHECK(got.term_off[k + 1] == t && got.value_off[k + 1] == v);
        if (with_terms)
            for (size_t m = 0; m < e.second.size(); ++m) {
                CHECK(got.term_i[got.term_off[k] + m] == e.second[m].first);
                CHECK(got.term_j[got.term_off[k] + m] == e.second[m].second);
            }
        ++k;
    }
    if (with_terms) CHECK(got.term_i.size() == t && got.term_j.size() == t);
    else CHECK(got.term_i.empty() && got.term_j.empty());
}

void check_schur(const ba::Lowered& L) {
    const size_t nc = L.n_cameras(), no = L.n_obs();
    std::vector<uint32_t> pos(no), cur(nc + 1, 0);
    for (size_t o = 0; o < no; ++o) cur[L.obs_cam[o] + 1]++;
    for (size_t c = 0; c < nc; ++c) cur[c + 1] += cur[c];
    for (size_t o = 0; o < no; ++o) pos[o] = cur[L.obs_cam[o]]++;
    PairMap want;  // (c < j) -> (pos of c's observation, pos of j's), point order
    for (size_t q = 0; q < L.n_points(); ++q) {
        if (L.point_fixed[q]) continue;
        for (uint32_t a = L.point_off[q]; a < L.point_off[q + 1]; ++a)
            for (uint32_t b = a + 1; b < L.point_off[q + 1]; ++b) {
                const uint32_t ca = L.obs_cam[a], cb = L.obs_cam[b];
                if (ca < cb) want[{ca, cb}].push_back({pos[a], pos[b]});
                else want[{cb, ca}].push_back({pos[b], pos[a]});
            }
    }
    const ba::SchurLists S = ba::build_schur_lists(L, pos);
    CHECK(S.pair_cam.size() == 2 * want.size() && S.term_off.size() == want.size() + 1);
    std::vector<std::vector<uint32_t>> partners(nc), rev(nc);
    size_t k = 0, t = 0;
    for (const auto& e : want) {
        const uint32_t c = e.first.first, j = e.first.second;
        CHECK(S.pair_cam[2 * k] == c && S.pair_cam[2 * k + 1] == j);
        partners[c].push_back(j);
        rev[j].push_back((uint32_t)k);
        for (const auto& x : e.second) {
            CHECK(S.term[2 * t] == x.first && S.term[2 * t + 1] == x.second);
            ++t;
        }
        CHECK(S.term_off[k + 1] == t);
        ++k;
    }
    CHECK(S.term.size() == 2 * t);
    uint64_t base = 0;
    for (size_t c = 0; c < nc; ++c) {
        CHECK(S.pair_off[c + 1] - S.pair_off[c] == partners[c].size());
        CHECK(S.nbr_off[c + 1] - S.nbr_off[c] == 1 + partners[c].size());
        CHECK(S.nbr[S.nbr_off[c]] == c);
        for (size_t m = 0; m < partners[c].size(); ++m) CHECK(S.nbr[S.nbr_off[c] + 1 + m] == partners[c][m]);
        CHECK(S.rev_off[c + 1] - S.rev_off[c] == rev[c].size());
        for (size_t m = 0; m < rev[c].size(); ++m) CHECK(S.rev_pair[S.rev_off[c] + m] == rev[c][m]);
        CHECK(S.row_base[c] == base);
        base += 9ull * ba::schur_row_len((uint32_t)(1 + partners[c].size()));
    }
    CHECK(S.n_values == base);
}

}  // namespace

int main() {
    {  // shared groups, repeated observations, fixed points; enough points for several ranges
        Scene s;
        make_scene(s, 40, 40000, true, 7);
        CHECK(ba::validate_problem(&s.p).ok());
        ba::Lowered L;
        CHECK(ba::lower_problem(&s.p, L, false).ok());
        const ba::BlockLists B = ba::build_block_lists(L);
        check_pairs(L, B, false, true);
        check_pairs(L, B, false, false);
    }
    {  // every pair of the pattern, a smaller scene
        Scene s;
        make_scene(s, 12, 3000, true, 11);
        ba::Lowered L;
        CHECK(ba::lower_problem(&s.p, L, false).ok());
        check_pairs(L, ba::build_block_lists(L), true, true);
    }
    {  // one camera per group, at most one observation per point and camera
        Scene s;
        make_scene(s, 64, 40000, false, 13);
        ba::Lowered L;
        CHECK(ba::lower_problem(&s.p, L, false).ok());
        check_pairs(L, ba::build_block_lists(L), false, true);
        check_schur(L);
    }
    std::printf("ok\n");
    return 0;
}
