"""Host-side checks that need no GPU.

* the C-ABI library loads and exports every function include/ba_hip.h declares, and the ctypes
  mirrors have the C layouts (sizes and field offsets from a C program compiled against the header);
* the product's lowering (ba_hip_prepare: anchor-first cameras, point-major observations, layout,
  first-encounter point parameters and the block pattern of S) is bit-identical to the oracle's;
* the golden initial cost of the reference's self-calibration contract fixture
  (crates/openmvg-ba/tests/self_calibration.rs:27-139: 863.644_893_653_298_3 within 1e-8);
* invalid problems fail the same way in the product and in the oracle.
"""

from __future__ import annotations

import ctypes as C
import pathlib
import re
import subprocess

import numpy as np
import pytest

from openmvg_ba import _abi as A
from openmvg_ba import _lib, synthetic
from openmvg_ba.problem import BundleProblem, Intrinsics, ProblemError, device_problem, prepare_with

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "ba_hip.h"


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ba_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_function(lib):
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.ba_hip_name() == b"openmvg-ba-hip"


STRUCTS = {
    "ba_intrinsic": A.Intrinsic, "ba_problem": A.Problem, "ba_solve_config": A.SolveConfigC,
    "ba_solve_report": A.SolveReportC, "ba_iteration_trace": A.IterationTrace, "ba_solution": A.SolutionC,
    "ba_native_sizes": A.NativeSizes, "ba_native_problem": A.NativeProblem, "ba_device_problem": A.DeviceProblem,
    "ba_schur_scalars": A.SchurScalars, "ba_batch_trace": A.BatchTrace, "ba_execution_plan": A.ExecutionPlanC,
}


def test_ctypes_mirrors_match_c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, cls in STRUCTS.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def _same_lowering(lib, oracle, problem):
    a = prepare_with(lib, "ba_hip_prepare", problem)
    b = prepare_with(oracle, "bao_prepare", problem)
    assert a.sizes == b.sizes
    assert a.gauge_camera == b.gauge_camera
    assert a.scene_scale == b.scene_scale
    for k in a.arrays:
        assert a.arrays[k].dtype == b.arrays[k].dtype
        assert np.array_equal(a.arrays[k], b.arrays[k]), k
    return a


@pytest.mark.parametrize("name", ["ring", "reference_fixture", "medium"])
def test_index_maps_are_bit_identical(lib, oracle, name):
    scene = {"ring": synthetic.ring_scene, "reference_fixture": synthetic.reference_fixture_scene,
             "medium": synthetic.medium_scene}[name]()
    _same_lowering(lib, oracle, scene.problem)


def _tiny_problem(obs, n_poses=3, n_points=2, fixed_pose=None):
    """Poses on a line looking down +z at points near (0, 0, 5); observation pixels projected."""
    R = np.stack([np.eye(3)] * n_poses)
    C_ = np.stack([[0.3 * i - 0.3, 0.05 * i, 0.0] for i in range(n_poses)])
    X = np.array([[0.1 * j, -0.05 * j, 5.0 + 0.2 * j] for j in range(n_points)])
    intr = [Intrinsics(focal=800.0, cx=320.0, cy=240.0, width=640, height=480, var_focal=True)
            for _ in range(n_poses)]
    pose, point = np.array([o[0] for o in obs]), np.array([o[1] for o in obs])
    pc = X[point] - C_[pose]
    pix = 800.0 * pc[:, :2] / pc[:, 2:3] + np.array([320.0, 240.0])
    vr = np.ones(n_poses, dtype=np.uint8)
    if fixed_pose is not None:
        vr[fixed_pose] = 0
    return BundleProblem(R, C_, intr, X, pose, pose, point, pix, pose_var_rotation=vr, pose_var_translation=vr.copy())


def test_quarantine_and_point_major_order(lib, oracle):
    """Point 1 is seen once and point 2 twice by one pose along one ray (no parallax): both are
    held out.  The kept observations, given in a shuffled order, come out point by point with the
    input order kept inside each point."""
    obs = [(2, 3), (0, 0), (2, 2), (1, 0), (1, 1), (0, 3), (2, 0), (2, 2)]
    nat = _same_lowering(lib, oracle, _tiny_problem(obs, n_points=4))
    assert nat.sizes["n_quarantined"] == 2
    assert list(nat.arrays["point_source"]) == [0, 3]
    assert list(nat.arrays["obs_source"]) == [1, 3, 6, 0, 5]
    assert list(nat.arrays["point_offsets"]) == [0, 3, 5]


def test_repeated_track_rows_keep_first_encounter_order(lib, oracle):
    """The reference's incidence example (workspace/topology.rs:611-629) with a second view added to
    point 1 so that it is not quarantined: point 0 seen by poses [2, 0, 2, 1], point 1 by [1, 1, 0];
    the unique poses of each track in first-encounter order are [2, 0, 1] and [1, 0]."""
    obs = [(2, 0), (0, 0), (2, 0), (1, 0), (1, 1), (1, 1), (0, 1)]
    nat = _same_lowering(lib, oracle, _tiny_problem(obs))
    arr = nat.arrays
    src = arr["camera_source"]
    off, cams = arr["point_offsets"], arr["obs_camera"]
    assert list(off) == [0, 4, 7]
    assert [int(src[c]) for c in cams] == [2, 0, 2, 1, 1, 1, 0]
    params = arr["point_parameters"]
    cross = arr["cross_offsets"]
    cam_off, cam_dim = arr["camera_offset"], arr["camera_dim"]
    grp_off, grp_dim, grp = arr["group_offset"], arr["group_dim"], arr["camera_group"]
    for p, order in ((0, [2, 0, 1]), (1, [1, 0])):
        want = []
        for s in order:
            c = int(np.nonzero(src == s)[0][0])
            want += list(range(cam_off[c], cam_off[c] + cam_dim[c]))
            g = grp[c]
            want += list(range(grp_off[g], grp_off[g] + grp_dim[g]))
        got = list(params[cross[p]:cross[p + 1]])
        assert sorted(got) == sorted(want)
        # first encounter: the first parameter of each pose appears in the track's pose order
        firsts = [got.index(cam_off[int(np.nonzero(src == s)[0][0])]) for s in order
                  if cam_dim[int(np.nonzero(src == s)[0][0])] > 0]
        assert firsts == sorted(firsts)


def test_invalid_problems_fail_like_the_oracle(lib, oracle):
    good = _tiny_problem([(0, 0), (1, 0), (1, 1), (2, 1)])
    bad_pose = _tiny_problem([(0, 0), (1, 0), (1, 1), (2, 1)])
    bad_pose.pose_rotation[1] = np.diag([1.0, 1.0, 1.01])
    bad_obs = _tiny_problem([(0, 0), (1, 0), (1, 1), (2, 1)])
    bad_obs.obs_point[2] = 7
    _same_lowering(lib, oracle, good)
    for bad in (bad_pose, bad_obs):
        for impl, fn in ((lib, "ba_hip_prepare"), (oracle, "bao_prepare")):
            with pytest.raises(ProblemError):
                prepare_with(impl, fn, bad)


def test_golden_initial_cost_of_the_self_calibration_contract_fixture(oracle):
    """tests/self_calibration.rs:27-139: 3 cameras, 30 points, squared loss, an intrinsic prior on
    camera 1 (log-focal, k1, k2 with sqrt-information diag(0.2, 5, 5)); the initial cost is
    863.644_893_653_298_3 within 1e-8."""
    truth_R = [np.eye(3), synthetic.exp_so3(np.array([0.01, 0.05, -0.01])),
               synthetic.exp_so3(np.array([-0.03, -0.04, 0.02]))]
    truth_t = [np.zeros(3), np.array([0.7, 0.03, 0.08]), np.array([-0.6, 0.07, 0.12])]
    truth_cam = [(680.0, 0.03, -0.01), (720.0, -0.025, 0.008), (700.0, 0.015, -0.006)]
    pts = np.array([[-1.2 + 0.42 * (i % 6), -0.8 + 0.38 * (i // 6), 3.6 + 0.19 * (i % 5)] for i in range(30)])

    def project(cam, p):
        n = p[:2] / p[2]
        r2 = n @ n
        return cam[0] * (1.0 + cam[1] * r2 + cam[2] * r2 * r2) * n

    uv = np.array([project(truth_cam[c], truth_R[c] @ X + truth_t[c]) for X in pts for c in range(3)])
    R = [r.copy() for r in truth_R]
    t = [v.copy() for v in truth_t]
    R[1] = synthetic.exp_so3(np.array([0.004, -0.003, 0.002])) @ R[1]
    t[2][1] -= 0.012
    X0 = np.array([[p[0] + 0.01 * np.sin(i), p[1], p[2]] for i, p in enumerate(pts)])
    cams = np.array([truth_cam[0], (735.0, -0.02, 0.006), (685.0, 0.01, -0.004)])
    arrays = dict(
        camera_rotation=np.ascontiguousarray(np.stack(R).reshape(3, 9)),
        camera_translation=np.ascontiguousarray(np.stack(t)),
        group_params=np.ascontiguousarray(cams), camera_group=np.arange(3, dtype=np.uint32),
        point_xyz=np.ascontiguousarray(X0), obs_camera=np.tile(np.arange(3, dtype=np.uint32), 30),
        obs_uv=np.ascontiguousarray(uv), point_offsets=(3 * np.arange(31)).astype(np.uint32),
        point_fixed=np.zeros(30, dtype=np.uint8), camera_offset=np.array([0, 6, 12], dtype=np.uint32),
        camera_dim=np.array([6, 6, 6], dtype=np.uint32), group_offset=np.array([18, 21, 24], dtype=np.uint32),
        group_dim=np.array([3, 3, 3], dtype=np.uint32),
        obs_point=np.repeat(np.arange(30, dtype=np.uint32), 3))
    sizes = dict(n_cameras=3, n_groups=3, n_points=30, n_observations=90, dimension=27)
    view = device_problem(arrays, sizes, A.LOSS_SQUARED, 0.0, -1, 1.0)
    grp = np.array([1], dtype=np.uint32)
    mean = np.array([720.0, -0.025, 0.008])
    sqrt_info = np.diag([0.2, 5.0, 5.0]).reshape(-1).copy()
    oracle.bao_native_cost.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    cost = oracle.bao_native_cost(C.byref(view), 1, grp.ctypes.data, mean.ctypes.data, sqrt_info.ctypes.data)
    assert abs(cost - 863.644_893_653_298_3) <= 1e-8, cost


def test_dense_factor_kernels_keep_their_registers(tmp_path):
    """The one-wavefront diagonal factor keeps a 64-entry row per lane in registers through fully
    unrolled loops; a change that defeats the unrolling (an early return inside the loop did)
    moves the row to scratch memory and slows the factor several times over.  Compiles the dense
    factor for gfx950 and requires zero scratch in every kernel of the file."""
    import shutil

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(hipcc).exists():
        pytest.skip("hipcc not available")
    src = ROOT / "openmvg-rs_amd" / "csrc" / "ba_cholesky.hip"
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-fPIC", "--cuda-device-only", "-c", str(src), "-o",
                        str(tmp_path / "chol.o"), "-Rpass-analysis=kernel-resource-usage"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    names = re.findall(r"Function Name: (\S+)", r.stderr)
    scratch = [int(x) for x in re.findall(r"ScratchSize \[bytes/lane\]: (\d+)", r.stderr)]
    assert names and len(names) == len(scratch)
    spilled = [(n, s) for n, s in zip(names, scratch) if s]
    assert not spilled, spilled
