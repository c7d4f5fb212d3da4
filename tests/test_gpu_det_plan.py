"""GPU checks of the deterministic backprojection through a canonical brick plan
(tg_bpj_plan order 3, TG_DETERMINISTIC): bit for bit the oracle's
single-thread order (C_siddon.c.template:489-494 with nthread = 1, C:923-928),
through the operator, through the reference-compatible function (which builds
the plan in its workspace) and at the full C2 shapes."""
import ctypes

import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd import _native as N
from tomograpy_amd import siddon as S
from tomograpy_amd import workloads as W
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import Siddon, Siddon4d, siddon_lo
from tests import cases as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("obstacle", [None, "sun"])
def test_operator_deterministic_c1_bit_exact(oracle, dtype, obstacle):
    """siddon_lo(deterministic=True) at C1 (64^3, 32 views of 128^2), summing
    into a nonzero cube (accumulate in place) with images in the reference's
    (n1, n2, views) C order: equal to the oracle's nthread = 1 run."""
    print("operator deterministic C1", np.dtype(dtype).name, obstacle, flush=True)
    ch, cube, hs = W.config1(dtype)
    P = siddon_lo(hs, ch, dtype=dtype, obstacle=obstacle, deterministic=True)
    assert P.det_plan_ok()
    rng = np.random.default_rng(17)
    y = rng.uniform(-1, 1, (128, 128, 32)).astype(dtype)
    init = rng.uniform(0, 1, cube.shape).astype(dtype)
    ref = init.copy()
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), ref, y, sun=obstacle == "sun",
                  nthread=1)
    out = torch.from_numpy(init.copy()).cuda()
    P.backproject_into(torch.from_numpy(y).cuda(), out)
    assert np.array_equal(out.cpu().numpy(), ref)
    # a second call reuses the plan; rmatvec starts from zero
    ref0 = np.zeros(cube.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), ref0, y, sun=obstacle == "sun",
                  nthread=1)
    assert np.array_equal(P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy(), ref0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_operator_deterministic_masks_bins_weights(oracle, dtype):
    """Pixel mask, voxel mask (masked voxels untouched), time bins (bpjt) and
    zero pixels, through the canonical plan: bit-exact to the oracle."""
    print("operator deterministic variants", np.dtype(dtype).name, flush=True)
    ch, x, data = C.small_case(n=40, npix=56, nv=10, dtype=dtype, radius=12.0)
    rng = np.random.default_rng(18)
    y = rng.uniform(-1, 1, data.shape).astype(dtype)
    y[:, :, 3] = 0.0  # a zero image adds nothing
    pm = rng.uniform(size=data.shape) < 0.2
    vm = W.voxel_radius(ch) < 0.6
    recs, mrec = view_records(data.header), map_record(ch)
    # pixel and voxel masks
    P = Siddon(data.header, ch, dtype=dtype, mask=pm, voxmask=vm, obstacle="sun",
               deterministic=True)
    assert P.det_plan_ok()
    ref = np.zeros(x.shape, dtype)
    oracle.siddon(oracle.BPJ, recs, mrec, ref, y, sun=True, nthread=1, mask=pm)
    ref[vm] = 0
    got = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert np.array_equal(got, ref)
    # time bins: view k -> bin k mod 3
    tindex = (np.arange(10) % 3).astype(np.int32)
    P4 = Siddon4d(data.header, ch, tindex=tindex, dtype=dtype, deterministic=True)
    assert P4.det_plan_ok()
    ref4 = np.zeros(x.shape + (3,), dtype)
    oracle.siddon(oracle.BPJT, recs, mrec, ref4, y, nthread=1, tindex=tindex)
    got4 = P4.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert np.array_equal(got4, ref4)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_full_call_builds_canonical_plan_in_workspace(oracle, dtype):
    """tg_conic_full_projector with TG_DETERMINISTIC and a workspace of
    tg_det_plan_workspace_bytes builds the canonical plan in it; with the
    chunked passes' minimum it takes those.  Both bit-exact to the oracle."""
    print("deterministic full call", np.dtype(dtype).name, flush=True)
    ch, x, data = C.small_case(n=24, npix=40, nv=7, dtype=dtype)
    rng = np.random.default_rng(19)
    y = rng.normal(size=data.shape).astype(dtype)
    tgd = N.TG_F64 if dtype == np.float64 else N.TG_F32
    yd = torch.from_numpy(y).cuda()
    ref = np.zeros(x.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, y, sun=True, nthread=1)
    out = torch.zeros(x.shape, dtype=yd.dtype, device="cuda")
    views = S.views_tensor(view_records(data.header), tgd, 40, 40, out.device)
    call = S.make_call(N.TG_BPJ, tgd, yd.shape, yd.stride(), out.shape, out.stride(),
                       map_record(ch), views, out.data_ptr(), yd.data_ptr(),
                       flags=N.TG_DETERMINISTIC | N.TG_OBSTACLE_SUN)
    planned = int(N.lib().tg_det_plan_workspace_bytes(ctypes.byref(call)))
    minimum = int(N.lib().tg_det_workspace_min_bytes(ctypes.byref(call)))
    assert planned > 0
    for size in (planned, minimum):
        out.zero_()
        ws = torch.empty(size, dtype=torch.uint8, device="cuda")
        call.workspace = ws.data_ptr()
        call.workspace_bytes = size
        N.check(N.lib().tg_conic_full_projector(ctypes.byref(call), S.stream_handle()))
        assert np.array_equal(out.cpu().numpy(), ref), size


def test_c2_deterministic_16_views_bit_exact(oracle):
    """The C2 shapes (256^3, 512^2) on 16 of its views, views 32, 64, 96 with
    observers on an axis among them: the operator's deterministic backward is
    the oracle's single-thread run bit for bit."""
    print("C2 deterministic 16 views", flush=True)
    ch, hs, rng = W.config2_geometry(np.float32)
    hs = [hs[k] for k in range(0, 128, 8)]
    P = siddon_lo(hs, ch, dtype=np.float32, deterministic=True)
    assert P.det_plan_ok()
    y = rng.uniform(0, 1, (512, 512, 16)).astype(np.float32)
    ref = np.zeros((256, 256, 256), np.float32)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), ref, y, nthread=1)
    got = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_deterministic_long_buckets(oracle, dtype):
    """Many rays per voxel (a 12^3 cube under 160^2 pixels, about 13 pixels
    per voxel edge): a batch gives a voxel far more than 8 contributions, so
    the in-place sort of long buckets runs beside the register sort of short
    ones.  Bit-exact to the oracle's single-thread order."""
    print("deterministic long buckets", np.dtype(dtype).name, flush=True)
    ch, x, data = C.small_case(n=12, npix=160, nv=5, dtype=dtype, seed=3)
    P = Siddon(data.header, ch, dtype=dtype, deterministic=True)
    assert P.det_plan_ok()
    rng = np.random.default_rng(20)
    y = rng.uniform(-1, 1, data.shape).astype(dtype)
    ref = np.zeros(x.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, y, nthread=1)
    got = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    assert np.array_equal(got, ref)


def test_canonical_fill_of_tiles_without_brick_lists(oracle):
    """The count pass of a canonical plan lists, per trace tile, up to 255
    bricks its rays enter, and the fill keeps the lane masks of those bricks in
    LDS slots; a tile whose rays enter more bricks is filled with a mask per
    brick of the map instead.  No tile of C1 enters that many, so the test marks
    every other tile's list as overflowed between the count and the build: the
    plan must still give the oracle's nthread = 1 sums bit for bit."""
    print("canonical fill without brick lists", flush=True)
    ch, cube, hs = W.config1(np.float32)
    P = siddon_lo(hs, ch, dtype=np.float32, deterministic=True)
    lib = N.lib()
    c = P._call_plain_bpj()
    p = N.TgBpjPlan()
    p.order = 3
    N.check(lib.tg_bpj_plan_init(ctypes.byref(c), ctypes.byref(p)))
    meta = torch.empty(int(p.meta_bytes), dtype=torch.uint8, device="cuda")
    p.meta, p.meta_bytes = meta.data_ptr(), meta.numel()
    scratch = torch.empty(int(p.scratch_bytes), dtype=torch.uint8, device="cuda")
    p.scratch, p.scratch_bytes = scratch.data_ptr(), scratch.numel()
    steps = torch.empty(P.nv * P.n1 * P.n2, dtype=torch.int32, device="cuda")
    p.ray_steps_out = steps.data_ptr()
    N.check(lib.tg_bpj_plan_count(ctypes.byref(c), ctypes.byref(p), S.stream_handle()))
    torch.cuda.synchronize()
    # the scratch ends with one list of 512 words per trace workgroup (256
    # pixels of one image row); word 0 counts its bricks, all ones = overflow
    tiles = P.nv * P.n1 * ((P.n2 + 255) // 256)
    lists = scratch.view(torch.int32)[-512 * tiles:].view(tiles, 512)
    assert bool((lists[:, 0] >= 0).all())
    lists[::2, 0] = -1
    rec = torch.empty(int(p.rec_bytes), dtype=torch.uint8, device="cuda")
    p.rec, p.rec_bytes = rec.data_ptr(), rec.numel()
    N.check(lib.tg_bpj_plan_build(ctypes.byref(c), ctypes.byref(p), S.stream_handle()))
    torch.cuda.synchronize()
    p.scratch, p.scratch_bytes, p.pairs_ready, p.ray_steps_out = None, 0, 0, None
    P._det_plan, P._det_plan_bufs = p, (meta, rec)
    y = np.random.default_rng(20).uniform(-1, 1, (128, 128, len(hs))).astype(np.float32)
    ref = np.zeros(cube.shape, np.float32)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), ref, y, nthread=1)
    out = P.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    T.check_status()
    assert np.array_equal(out, ref)
