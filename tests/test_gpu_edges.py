"""GPU edge cases against the oracle: ragged image sizes that leave partial
ray tiles, views whose rays all miss the cube, and a cube edge of 1024 voxels
(the largest the brick plan takes) traversed end to end.  Forward projections
bit-exact; backprojections (atomic scatter and brick plan) to 1e-5 relative L2
in float32 and 1e-12 in float64."""
import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import siddon_lo

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-12}


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den else 1.0)


def check_pair(oracle, ch, hs, shape, dtype, seed=0):
    """Forward (functional and operator paths) bit-exact, backward (atomic and
    plan) within TOL, against the oracle on the same inputs."""
    rng = np.random.default_rng(seed)
    n1, n2 = int(hs[0]["NAXIS1"]), int(hs[0]["NAXIS2"])
    cube = rng.uniform(0.0, 1.0, shape).astype(dtype)
    ref = np.zeros((n1, n2, len(hs)), dtype=dtype)
    oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), cube, ref)
    xd = torch.from_numpy(cube).cuda()
    xd.header = dict(ch)
    yd = torch.zeros((n1, n2, len(hs)), dtype=xd.dtype, device="cuda")
    yd.header = hs
    T.projector(yd, xd)
    assert np.array_equal(yd.cpu().numpy(), ref)
    P = siddon_lo(hs, ch, dtype=dtype)
    assert np.array_equal(P.matvec(torch.from_numpy(cube).cuda()).cpu().numpy(), ref)
    y = rng.normal(size=(n1, n2, len(hs))).astype(dtype)
    bref = np.zeros(shape, dtype=dtype)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref, y, nthread=1)
    for mode in ("atomic", "plan"):
        Pm = siddon_lo(hs, ch, dtype=dtype, bpj_mode=mode)
        b = Pm.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
        T.check_status()
        assert rel_l2(b, bref) <= TOL[dtype], mode
    return ref, bref


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ragged_images(oracle, dtype):
    """37 x 53 images (column tiles of 64 pixels, the last rows and columns
    partial) of a 29^3 cube from 5 views."""
    print("ragged images", np.dtype(dtype).name, flush=True)
    ch = T.centered_cubic_map_header(3.0, 29, dtype=dtype)
    ih = T.centered_image_header(T.fov(ch, 9.0), (37, 53), dtype=dtype)
    hs = T.circular_trajectory_header(radius=9.0, longitudes=0.2 + 1.1 * np.arange(5),
                                      latitudes=0.1 * np.sin(np.arange(5)), **ih)
    check_pair(oracle, ch, hs, (29, 29, 29), dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_views_missing_the_cube(oracle, dtype):
    """Views pointed 0.3 rad off the cube: no ray meets it, so projections are
    zero, backprojections leave the cube at zero, the plan has no records and
    every ray has zero steps."""
    print("views missing the cube", np.dtype(dtype).name, flush=True)
    ch = T.centered_cubic_map_header(3.0, 24, dtype=dtype)
    ih = T.centered_image_header(T.fov(ch, 20.0), 16, dtype=dtype)
    ih["CRVAL1"] = 0.3
    hs = T.circular_trajectory_header(radius=20.0, longitudes=0.5 * np.arange(4), **ih)
    ref, bref = check_pair(oracle, ch, hs, (24, 24, 24), dtype)
    assert not ref.any() and not bref.any()
    P = siddon_lo(hs, ch, dtype=dtype, bpj_mode="plan")
    assert int(P.build_plan().records) == 0
    assert int(P.ray_steps().sum()) == 0 and P.count_steps() == 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cube_edge_1024(oracle, dtype):
    """A 1024 x 8 x 16 slab seen end-on (rays cross up to 1024 voxels along
    the first axis): voxel indices and brick coordinates at the plan's largest
    edge."""
    print("cube edge 1024", np.dtype(dtype).name, flush=True)
    shape = (1024, 8, 16)
    ch = T.centered_cubic_map_header((3.0, 3.0 * 8 / 1024, 3.0 * 16 / 1024), shape, dtype=dtype)
    ih = T.centered_image_header((2.0 * np.arctan(0.03 / 214.0), 2.0 * np.arctan(0.05 / 214.0)),
                                 (12, 20), dtype=dtype)
    hs = T.circular_trajectory_header(radius=214.0, longitudes=np.array([0.0, np.pi + 0.01]),
                                      **ih)
    check_pair(oracle, ch, hs, shape, dtype)
    P = siddon_lo(hs, ch, dtype=dtype)
    assert int(P.ray_steps().max()) >= 1024  # the rays inside the slab cross all of it


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ["c1_sun", "masked", "time_bins", "axis_views", "missing"])
def test_plan_count_step_counts_and_single_trace_fill(oracle, dtype, case):
    """The plan's count pass writes every ray's step count as tg_count_steps
    does (the projector runs on them), and the fill that reuses the count
    pass's per-workgroup record counts builds the plan the two-trace fill
    builds: same record count, backprojection equal to the ulp (float64 sums)
    and within tolerance of the oracle."""
    from tomograpy_amd import workloads as W
    from tomograpy_amd.lo_wrapper import Siddon4d
    print("plan count", case, np.dtype(dtype).name, flush=True)
    kw, cls = {}, siddon_lo
    if case == "c1_sun":
        ch, _, hs = W.config1(dtype)
        kw["obstacle"] = "sun"
    elif case == "axis_views":
        ch, hs, _ = W.config2_geometry(dtype, n=64, n_views=8, n_pix=96)  # LON = k pi / 4
    else:
        ch = T.centered_cubic_map_header(3.0, 40, dtype=dtype)
        ih = T.centered_image_header(T.fov(ch, 30.0), 48, dtype=dtype)
        if case == "missing":
            ih["CRVAL1"] = 0.3
        hs = T.circular_trajectory_header(radius=30.0, longitudes=0.3 + 0.9 * np.arange(6), **ih)
        if case == "masked":
            rng = np.random.default_rng(3)
            kw["mask"] = rng.uniform(size=(48, 48, 6)) < 0.3
        if case == "time_bins":
            cls = lambda h, c, **k: Siddon4d(h, c, tindex=np.arange(6) % 3, **k)
    P1 = cls(hs, ch, dtype=dtype, bpj_mode="plan", **kw)
    P2 = cls(hs, ch, dtype=dtype, bpj_mode="plan", plan_single_trace=False, **kw)
    # the fill that tests every step against the map and the obstacle again
    P3 = cls(hs, ch, dtype=dtype, bpj_mode="plan", plan_counted_fill=False, **kw)
    P0 = cls(hs, ch, dtype=dtype, bpj_mode="atomic", **kw)
    steps_plan = P1.ray_steps().clone()  # written by the plan's count pass
    assert P1._plan_count is not None
    assert torch.equal(steps_plan, P0.ray_steps())
    p1, p2, p3 = P1.build_plan(), P2.build_plan(), P3.build_plan()
    assert p1.records == p2.records == p3.records and p1.chunks == p2.chunks == p3.chunks
    rng = np.random.default_rng(4)
    y = torch.from_numpy(rng.uniform(size=P1.shapeout).astype(dtype)).cuda()
    b1 = P1.rmatvec(y).cpu().numpy()
    for P in (P2, P3):
        b2 = P.rmatvec(y).cpu().numpy()
        if P1.acc64:  # float64 sums rounded once: equal but for one-ulp ties
            d = b1 != b2
            assert d.mean() <= 1e-3
            assert np.all(np.abs(b1 - b2)[d] <= np.spacing(np.maximum(np.abs(b1), np.abs(b2)))[d])
        else:  # float64 data: the brick sums' order follows the record order, which may differ
            assert np.linalg.norm(b1 - b2) <= 1e-13 * max(np.linalg.norm(b2), 1e-300)
    if case == "missing":
        assert p1.records == 0 and not b1.any()


def test_plan_too_large_for_free_memory_falls_back(oracle):
    """In "auto" mode an operator whose counted plan would not fit the free
    device memory backprojects with the atomic scatter (forced here by a zero
    share); the count pass's per-ray step counts still drive the forward."""
    from tomograpy_amd import workloads as W
    print("plan memory fallback", flush=True)
    ch, cube, hs = W.config1(np.float32)
    P = siddon_lo(hs, ch, dtype=np.float32)
    P.PLAN_MAX_FREE_SHARE = 0.0
    assert not P.use_plan()
    assert P._ray_steps is not None and P._plan_count is None
    x = torch.from_numpy(np.ascontiguousarray(cube)).cuda()
    y_ref = np.zeros((128, 128, len(hs)), np.float32)
    oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), np.asarray(cube),
                  T.InfoArray(y_ref, header=hs))
    assert np.array_equal(P.matvec(x).cpu().numpy(), y_ref)
    rng = np.random.default_rng(41)
    y = rng.uniform(-1, 1, y_ref.shape).astype(np.float32)
    ref = np.zeros(cube.shape, np.float32)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), ref, T.InfoArray(y, header=hs),
                  nthread=1)
    b = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert P._plan is None
    assert np.linalg.norm(b - ref) <= 1e-5 * np.linalg.norm(ref)
