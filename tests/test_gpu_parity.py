"""GPU parity: the HIP library against the CPU oracle on identical inputs.

Tolerances (written per test):
  * forward projection (pj/pjt): bit-exact -- each pixel's sum is sequential
    in the reference (C_siddon.c.template:915-920) and on the GPU;
  * backprojection, atomic mode: relative L2 <= 1e-5 (fp32), 1e-12 (fp64)
    against the oracle's single-thread order;
  * backprojection, deterministic mode: bit-exact to the single-thread order;
  * step counts, unit vectors, intersection parameters: exact.
"""
import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd import workloads as W
from tomograpy_amd.headers import map_record, view_records
from tests import cases as C

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den else 1.0)


def c1_problem(dtype, n_views=32):
    ch, cube, hs = W.config1(dtype)
    hs = hs[:n_views]
    data = T.InfoArray(np.zeros((128, 128, len(hs)), dtype=dtype), header=hs)
    return ch, T.InfoArray(cube, header=dict(ch)), data


def to_dev(x):
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(x))).cuda()
    t.header = x.header
    return t


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("obstacle", [None, "sun"])
def test_forward_c1_bit_exact(oracle, dtype, obstacle):
    print("forward C1", np.dtype(dtype).name, obstacle, flush=True)
    ch, cube, data = c1_problem(dtype)
    rng = np.random.default_rng(5)
    data[:] = rng.uniform(-1, 1, data.shape).astype(dtype)  # accumulate into nonzero data
    ref = data.copy()
    oracle.siddon(oracle.PJ, view_records(data.header), map_record(ch), cube, ref,
                  sun=obstacle == "sun")
    out = data.copy()
    T.projector(out, cube, obstacle=obstacle)
    assert np.array_equal(np.asarray(out), np.asarray(ref))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("obstacle", [None, "sun"])
def test_backward_c1_atomic_within_tolerance(oracle, dtype, obstacle):
    print("backward atomic C1", np.dtype(dtype).name, obstacle, flush=True)
    ch, cube, data = c1_problem(dtype)
    rng = np.random.default_rng(6)
    data[:] = rng.uniform(0, 1, data.shape).astype(dtype)
    ref = T.InfoArray(np.zeros(cube.shape, dtype), header=dict(ch))
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, data,
                  sun=obstacle == "sun", nthread=1)
    out = T.InfoArray(np.zeros(cube.shape, dtype), header=dict(ch))
    T.backprojector(data, out, obstacle=obstacle)
    tol = 1e-5 if dtype == np.float32 else 1e-12
    assert rel_l2(out, ref) <= tol


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("obstacle", [None, "sun"])
def test_backward_c1_deterministic_bit_exact(oracle, dtype, obstacle):
    print("backward deterministic C1", np.dtype(dtype).name, obstacle, flush=True)
    ch, cube, data = c1_problem(dtype)
    rng = np.random.default_rng(7)
    data[:] = rng.uniform(-1, 1, data.shape).astype(dtype)
    init = rng.uniform(0, 1, cube.shape).astype(dtype)  # accumulate into nonzero cube
    ref = T.InfoArray(init.copy(), header=dict(ch))
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, data,
                  sun=obstacle == "sun", nthread=1)
    out = T.InfoArray(init.copy(), header=dict(ch))
    T.backprojector(data, out, obstacle=obstacle, deterministic=True)
    assert np.array_equal(np.asarray(out), np.asarray(ref))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_deterministic_small_workspace_chunks_views(oracle, dtype):
    """Workspaces of one pixel row up to a few views still give the exact order.

    The rays are taken in chunks of pixel rows, so small workspaces start and
    end chunks inside a view."""
    import ctypes
    from tomograpy_amd import _native as N, siddon as S
    print("deterministic small workspace", np.dtype(dtype).name, flush=True)
    ch, x, data = C.small_case(n=16, npix=32, nv=9, dtype=dtype)
    rng = np.random.default_rng(8)
    y = T.InfoArray(rng.normal(size=data.shape).astype(dtype), header=data.header)
    tgd = N.TG_F64 if dtype == np.float64 else N.TG_F32
    yd = torch.from_numpy(np.asarray(y)).cuda()
    for sun in (False, True):
        ref = np.zeros(x.shape, dtype)
        oracle.siddon(oracle.BPJ, view_records(y.header), map_record(ch), ref, y, sun=sun, nthread=1)
        out = torch.zeros(x.shape, dtype=yd.dtype, device="cuda")
        views = S.views_tensor(view_records(y.header), tgd, 32, 32, out.device)
        flags = N.TG_DETERMINISTIC | (N.TG_OBSTACLE_SUN if sun else 0)
        call = S.make_call(N.TG_BPJ, tgd, yd.shape, yd.stride(), out.shape, out.stride(),
                           map_record(ch), views, out.data_ptr(), yd.data_ptr(), flags=flags)
        need = int(N.lib().tg_det_workspace_min_bytes(ctypes.byref(call)))
        for size in (need, 3 * need, 40 * need):
            out.zero_()
            ws = torch.empty(size, dtype=torch.uint8, device="cuda")
            call.workspace = ws.data_ptr()
            call.workspace_bytes = size
            N.check(N.lib().tg_conic_full_projector(ctypes.byref(call), S.stream_handle()))
            assert np.array_equal(out.cpu().numpy(), ref), (sun, size)
    # too small a workspace is an error, not a wrong answer
    call.workspace_bytes = 1024
    assert N.lib().tg_conic_full_projector(ctypes.byref(call), S.stream_handle()) == N.TG_EWORKSPACE


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_step_counts_match_oracle(oracle, dtype):
    ch, cube, data = c1_problem(dtype, n_views=8)
    ref = oracle.ray_steps(map_record(ch), cube.shape, view_records(data.header), 128, 128,
                           dtype=dtype)
    total, per_ray = T.count_steps(data, cube, per_ray=True)
    assert np.array_equal(per_ray.cpu().numpy().astype(np.uint32), ref)
    assert total == int(ref.sum())
    ref_sun = oracle.ray_steps(map_record(ch), cube.shape, view_records(data.header), 128, 128,
                               sun=True, dtype=dtype)
    assert T.count_steps(data, cube, obstacle="sun") == int(ref_sun.sum())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_unit_vectors_and_intersections_exact(oracle, dtype):
    ch, cube, data = C.small_case(n=16, npix=32, nv=6, dtype=dtype)
    data[3, 4, 1] = np.nan  # NaN pixels are skipped (C_siddon.c.template:638)
    recs = view_records(data.header)
    u_ref = oracle.full_unit_vector(recs, 32, 32, dtype, data=np.asarray(data))
    u = T.C_full_unit_vector(data)
    assert np.array_equal(u, u_ref)
    assert np.all(u[3, 4, 1] == 0)
    a1_ref, an_ref = oracle.full_intersection_parameters(map_record(ch), recs, u_ref, dtype)
    a1, an = T.C_full_intersection_parameters(data, cube, u)
    assert np.array_equal(a1, a1_ref) and np.array_equal(an, an_ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mask_and_time_variants(oracle, dtype):
    ch, cube, data = C.small_case(n=16, npix=32, nv=6, dtype=dtype)
    rng = np.random.default_rng(11)
    mask = rng.uniform(size=data.shape) < 0.3
    recs = view_records(data.header)
    # masked forward, bit-exact
    ref = data.copy()
    oracle.siddon(oracle.PJ, recs, map_record(ch), cube, ref, mask=mask)
    out = data.copy()
    T.projector(out, cube, mask=mask)
    assert np.array_equal(np.asarray(out), np.asarray(ref))
    # 4-D: image t reads bin t; and an explicit view -> bin map
    cube4 = T.InfoArray(rng.uniform(size=(16, 16, 16, 6)).astype(dtype), header=dict(ch))
    ref = data.copy()
    oracle.siddon(oracle.PJT, recs, map_record(ch), np.asarray(cube4), ref)
    out = data.copy()
    T.projector4d(out, cube4)
    assert np.array_equal(np.asarray(out), np.asarray(ref))
    tindex = np.array([2, 0, 1, 2, 0, 1], np.int32)
    cube3b = T.InfoArray(np.asarray(cube4)[..., :3].copy(), header=dict(ch))
    ref = data.copy()
    oracle.siddon(oracle.PJT, recs, map_record(ch), np.asarray(cube3b), ref, tindex=tindex)
    out = data.copy()
    T.projector4d(out, cube3b, tindex=tindex)
    assert np.array_equal(np.asarray(out), np.asarray(ref))
    # 4-D backprojection, deterministic: bit-exact
    y = T.InfoArray(rng.normal(size=data.shape).astype(dtype), header=data.header)
    ref4 = np.zeros((16, 16, 16, 3), dtype)
    oracle.siddon(oracle.BPJT, recs, map_record(ch), ref4, y, tindex=tindex, nthread=1)
    out4 = T.InfoArray(np.zeros((16, 16, 16, 3), dtype), header=dict(ch))
    T.backprojector4d(y, out4, tindex=tindex, deterministic=True)
    assert np.array_equal(np.asarray(out4), ref4)
    # masked backprojection, atomic
    refm = np.zeros(cube.shape, dtype)
    oracle.siddon(oracle.BPJ, recs, map_record(ch), refm, y, mask=mask, nthread=1)
    outm = T.InfoArray(np.zeros(cube.shape, dtype), header=dict(ch))
    T.backprojector(y, outm, mask=mask)
    assert rel_l2(outm, refm) <= (1e-5 if dtype == np.float32 else 1e-12)


def test_image_and_ray_projectors(oracle):
    ch, cube, data = C.small_case(n=16, npix=32, nv=6)
    recs = view_records(data.header)
    ref = data.copy()
    oracle.siddon(oracle.PJ, recs, map_record(ch), cube, ref, t0=2, t1=3)
    out = data.copy()
    T.conic_image_projector(out, cube, 2)
    assert np.array_equal(np.asarray(out), np.asarray(ref))
    out = data.copy()
    T.ray_projector(out, cube, (5, 7, 4))
    iv, ln, _ = oracle.ray_trace(map_record(ch), cube.shape, recs[4], 5, 7)
    v = 0.0
    for k, l in zip(iv, ln):
        v = v + l * np.asarray(cube)[tuple(k)]
    assert np.asarray(out)[5, 7, 4] == v
    assert np.count_nonzero(np.asarray(out)) == 1


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_adjoint_identity_gpu(dtype, tol):
    ch, x, data = C.small_case(n=32, npix=48, nv=12, dtype=dtype, radius=6.0)
    rng = np.random.default_rng(12)
    xd = to_dev(x)
    y = torch.from_numpy(rng.normal(size=data.shape).astype(dtype)).cuda()
    y.header = data.header
    px = torch.zeros(data.shape, dtype=xd.dtype, device="cuda")
    px.header = data.header
    T.projector(px, xd, obstacle="sun")
    pty = torch.zeros(x.shape, dtype=xd.dtype, device="cuda")
    pty.header = dict(ch)
    T.backprojector(y, pty, obstacle="sun")
    lhs = float((px.double() * y.double()).sum())
    rhs = float((xd.double() * pty.double()).sum())
    assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs))


def test_reference_invariants_on_gpu():
    """tests/test_siddon.py:88-115 on the GPU path."""
    im_h = dict(C.image_headers(np.float64)[0])
    im_h["CRPIX1"] = 1.0
    im_h["CRPIX2"] = 1.0
    data = C.stack(im_h, np.float64, n_images=1, fill=1.0)
    for ch in C.object_headers(np.float64):
        cube = T.InfoArray(np.zeros((ch["NAXIS1"],) * 3), header=dict(ch))
        T.backprojector(data, cube)
        assert np.sum(cube) == ch["PSHAPE1"]
    data = C.stack(C.image_headers(np.float64)[0], np.float64, n_images=1, fill=1.0)
    h1, h16 = C.object_headers(np.float64)
    c16 = T.InfoArray(np.zeros((16, 16, 16)), header=dict(h16))
    c1 = T.InfoArray(np.zeros((1, 1, 1)), header=dict(h1))
    T.backprojector(data, c16, deterministic=True)
    T.backprojector(data, c1, deterministic=True)
    assert np.sum(c16) == c1[0, 0, 0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_support_special_cases_gpu(dtype):
    """tests/test_siddon.py:141-155: data != 0 exactly where the ray meets the cube."""
    for ch in C.object_headers(dtype):
        for im_h in C.image_headers(dtype):
            data = C.stack(im_h, dtype, n_images=17)
            cube = T.InfoArray(np.ones((ch["NAXIS1"],) * 3, dtype=dtype), header=dict(ch))
            T.projector(data, cube)
            u = T.C_full_unit_vector(data)
            a1, an = T.C_full_intersection_parameters(data, cube, u)
            flag = np.minimum(a1, an).max(-1) < np.maximum(a1, an).min(-1)
            assert np.array_equal(np.asarray(data) != 0, flag)


def test_c2_shape_views_bit_exact(oracle):
    """Full C2 shapes (256^3, 512^2 images) on three views: forward bit-exact."""
    ch, hs, rng = W.config2_geometry(np.float32)
    sel = [0, 37, 101]
    hs = [hs[k] for k in sel]
    cube = T.InfoArray(W.blob_phantom(ch, 40, rng), header=dict(ch))
    data = T.InfoArray(np.zeros((512, 512, 3), np.float32), header=hs)
    ref = data.copy()
    oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), cube, ref)
    out = data.copy()
    T.projector(out, cube)
    assert np.array_equal(np.asarray(out), np.asarray(ref))


def test_torch_inputs_accumulate_in_place():
    ch, cube, data = C.small_case(n=16, npix=32, nv=4, dtype=np.float32)
    x = to_dev(cube)
    y = torch.full(data.shape, 2.0, dtype=torch.float32, device="cuda")
    y.header = data.header
    y0 = y.clone()
    T.projector(y, x)
    y1 = y.clone()
    T.projector(y, x)
    torch.cuda.synchronize()
    T.check_status()
    # second call adds the same projection again (in-place accumulation)
    assert torch.allclose(y - y1, y1 - y0, rtol=1e-5, atol=1e-5)


# ---- brick-privatised backprojection (plan) ---------------------------------------------
@pytest.mark.parametrize("dtype,acc64", [(np.float32, True), (np.float32, False),
                                         (np.float64, False)])
@pytest.mark.parametrize("obstacle", [None, "sun"])
def test_plan_backprojection_c1(oracle, dtype, acc64, obstacle):
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("plan bpj C1", np.dtype(dtype).name, acc64, obstacle, flush=True)
    ch, cube, data = c1_problem(dtype)
    rng = np.random.default_rng(21)
    y = rng.uniform(-1, 1, data.shape).astype(dtype)
    ref = np.zeros(cube.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), sun=obstacle == "sun", nthread=1)
    P = siddon_lo(data.header, ch, dtype=dtype, obstacle=obstacle, bpj_mode="plan", acc64=acc64)
    out = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    T.check_status()
    assert P._plan.records > 0
    assert rel_l2(out, ref) <= (1e-5 if dtype == np.float32 else 1e-12)


def test_plan_backprojection_any_stack_layout():
    """A stack in the reference's (n1, n2, views) C order, a view-major one and a
    strided view give the same cube (float64 sums do not depend on the order)."""
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("plan bpj stack layouts", flush=True)
    ch, cube, data = c1_problem(np.float32)
    rng = np.random.default_rng(22)
    y = rng.uniform(-1, 1, data.shape).astype(np.float32)
    P = siddon_lo(data.header, ch, dtype=np.float32, bpj_mode="plan")
    yc = torch.from_numpy(y).cuda()  # views last, C order
    yv = torch.from_numpy(np.ascontiguousarray(y.transpose(2, 0, 1))).cuda().permute(1, 2, 0)
    big = torch.zeros((y.shape[0], 2 * y.shape[1], y.shape[2]), device="cuda")
    big[:, ::2] = yc
    ys = big[:, ::2]  # strided
    outs = [P.rmatvec(t).clone() for t in (yv, yc, ys, yv)]
    T.check_status()
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    # the C ABI reads any stack layout itself (ray ids decoded to pixel offsets)
    import ctypes
    from tomograpy_amd import _native as N, siddon as S
    for t in (yc, ys):
        out = torch.zeros_like(outs[0])
        c = P._call(N.TG_BPJ, out, t)
        N.check(N.lib().tg_bpj_plan_backproject(ctypes.byref(c), ctypes.byref(P.build_plan()),
                                                S.stream_handle()))
        assert torch.equal(out, outs[0])


# ---- float64 accumulation of float32 backprojections (TG_ACC_F64) ----------------------
def assert_ulp_close(a, b, max_frac=1e-3):
    """Equal except, in at most max_frac of the entries, by one unit in the last place."""
    a = np.asarray(a)
    b = np.asarray(b)
    diff = a != b
    if diff.any():
        sp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
        assert np.all(np.abs(a - b)[diff] <= sp[diff]), np.abs(a - b)[diff].max()
    assert diff.mean() <= max_frac, diff.mean()


def test_acc64_backprojection_is_order_independent(oracle):
    """Atomic scatter, brick plan and batched launch sum the same float32 products
    in float64: the three cubes agree up to rare one-ulp ties, each is within
    1e-5 of the oracle, and the accumulator is zero again afterwards."""
    from tomograpy_amd import batch as B
    from tomograpy_amd.lo_wrapper import siddon_lo
    ch, cube, data = c1_problem(np.float32)
    rng = np.random.default_rng(23)
    y = rng.uniform(-1, 1, data.shape).astype(np.float32)
    ref = np.zeros(cube.shape, np.float32)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), nthread=1)
    yt = torch.from_numpy(y).cuda()
    Pa = siddon_lo(data.header, ch, dtype=np.float32, bpj_mode="atomic")
    Pp = siddon_lo(data.header, ch, dtype=np.float32, bpj_mode="plan")
    assert Pa.acc64 and Pp.acc64
    oa = Pa.rmatvec(yt).cpu().numpy()
    oa2 = Pa.rmatvec(yt).cpu().numpy()
    op = Pp.rmatvec(yt).cpu().numpy()
    ob = torch.zeros_like(Pa.xin)
    B.run([Pa, Pp], [yt, yt], [ob, torch.zeros_like(ob)], backward=True)
    ob = ob.cpu().numpy()
    T.check_status()
    for o in (oa, op, ob):
        assert rel_l2(o, ref) <= 1e-5
    assert_ulp_close(oa2, oa)
    assert_ulp_close(op, oa)
    assert_ulp_close(ob, oa)
    assert float(Pa.acc_buffer().abs().max()) == 0.0
    assert float(Pp.acc_buffer().abs().max()) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("spread", [0, 3, 6, 10])
def test_plan_tie_spread_only_reorders(oracle, spread):
    """Plans whose sorted passes order ties differently (tg_bpj_plan.tie_spread)
    give the same float64-summed cube, within 1e-5 of the oracle."""
    print("plan tie spread", spread, flush=True)
    from tomograpy_amd.lo_wrapper import siddon_lo
    ch, cube, data = c1_problem(np.float32)
    rng = np.random.default_rng(29)
    y = rng.uniform(-1, 1, data.shape).astype(np.float32)
    ref = np.zeros(cube.shape, np.float32)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), nthread=1)
    yt = torch.from_numpy(y).cuda()
    P0 = siddon_lo(data.header, ch, dtype=np.float32, bpj_mode="plan", plan_tie_spread=0)
    Ps = siddon_lo(data.header, ch, dtype=np.float32, bpj_mode="plan", plan_tie_spread=spread)
    o0 = P0.rmatvec(yt).cpu().numpy()
    os_ = Ps.rmatvec(yt).cpu().numpy()
    T.check_status()
    assert Ps.build_plan().tie_spread == spread
    assert rel_l2(os_, ref) <= 1e-5
    assert_ulp_close(os_, o0)


@pytest.mark.parametrize("n,brick,chunk", [(40, (16, 16, 16), 0), (24, (8, 16, 32), 100),
                                           (33, (32, 8, 8), 7)])
def test_plan_backprojection_edges(oracle, n, brick, chunk):
    """Cube edges that are not brick multiples, anisotropic bricks, tiny chunks,
    with pixel and voxel masks."""
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("plan bpj edges", n, brick, chunk, flush=True)
    ch, x, data = C.small_case(n=n, npix=32, nv=6, dtype=np.float64, radius=9.0)
    rng = np.random.default_rng(22)
    y = rng.normal(size=data.shape)
    mask = rng.uniform(size=data.shape) < 0.2
    vm = rng.uniform(size=(n, n, n)) < 0.1
    ref = np.zeros((n, n, n))
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), sun=True, nthread=1, mask=mask)
    ref[vm] = 0.0
    P = siddon_lo(data.header, ch, dtype=np.float64, obstacle="sun", mask=mask, voxmask=vm,
                  bpj_mode="plan", brick=brick, chunk_records=chunk)
    out = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    T.check_status()
    assert rel_l2(out, ref) <= 1e-12


@pytest.mark.parametrize("brick,order", [((2, 2, 2), 2), ((8, 8, 8), 1)])
def test_plan_variants_global_counts_and_permuted_order(oracle, brick, order):
    """2^3 bricks of a 40^3 cube have 20 480 brick keys, more than the count and
    fill passes keep in LDS: the counts and cursors are global atomics.  Plan
    order 1 places a brick's records through a permutation of its range, from
    the brick offsets.  Both within 1e-12 of the oracle (fp64), with the Sun."""
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("plan variants", brick, order, flush=True)
    ch, x, data = C.small_case(n=40, npix=32, nv=6, dtype=np.float64, radius=9.0)
    y = np.random.default_rng(27).normal(size=data.shape)
    ref = np.zeros((40, 40, 40))
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), sun=True, nthread=1)
    P = siddon_lo(data.header, ch, dtype=np.float64, obstacle="sun", bpj_mode="plan", brick=brick,
                  plan_order=order)
    out = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    T.check_status()
    assert int(P.build_plan().order) == order
    assert rel_l2(out, ref) <= 1e-12


def test_plan_backprojection_time_bins(oracle):
    from tomograpy_amd.lo_wrapper import Siddon4d
    ch, x, data = C.small_case(n=20, npix=24, nv=8, dtype=np.float64, radius=9.0)
    tindex = np.array([0, 2, 1, 2, 0, 1, 1, 2], np.int32)
    rng = np.random.default_rng(23)
    y = rng.normal(size=data.shape)
    ref = np.zeros((20, 20, 20, 3))
    oracle.siddon(oracle.BPJT, view_records(data.header), map_record(ch), ref,
                  T.InfoArray(y, header=data.header), nthread=1, tindex=tindex)
    P = Siddon4d(data.header, ch, tindex=tindex, dtype=np.float64, bpj_mode="plan", brick=(8, 8, 8))
    out = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert rel_l2(out, ref) <= 1e-12
