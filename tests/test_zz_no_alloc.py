"""The entry point that allocates nothing (tg_conic_full_projector_static):
its float64 accumulator is a device array of the library, not a caller
buffer.  Checked through the C ABI against tg_conic_full_projector with a
caller-owned accumulator and against the oracle, on two calls in a row, the
second on a smaller map (the array keeps the first call's contents, which the
second must not see)."""
import ctypes

import numpy as np
import pytest
import torch

from tomograpy_amd import _native as N
from tomograpy_amd import siddon as S
from tomograpy_amd import workloads as W
from tomograpy_amd.headers import map_record, view_records

pytestmark = pytest.mark.gpu


def _case(n, npix, nv, seed):
    rng = np.random.default_rng(seed)
    ch = W.cube_header(n, dtype=np.float32)
    hs = W.view_headers(npix, rng.uniform(0, 2 * np.pi, nv), np.deg2rad(rng.uniform(-7, 7, nv)),
                        d=float(rng.uniform(150, 250)), dtype=np.float32)
    y = rng.uniform(0.0, 1.0, (npix, npix, nv)).astype(np.float32)
    return ch, hs, y


def _call(ch, hs, yd, cube, flags, acc=None, kind=N.TG_BPJ, nt=1, tindex=None):
    dev = yd.device
    views = S.views_tensor(view_records(hs), N.TG_F32, yd.shape[0], yd.shape[1], dev)
    c = S.make_call(kind, N.TG_F32, yd.shape, yd.stride(), cube.shape, cube.stride(),
                    map_record(ch), views, cube.data_ptr(), yd.data_ptr(), flags=flags, nt=nt,
                    tindex=tindex, status=S._status_tensor(dev))
    if acc is not None:
        c.acc = acc.data_ptr()
    return c, views


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _ulp_equal(a, b):
    d = a != b
    return d.mean() <= 1e-3 and np.all(np.abs(a - b)[d] <=
                                       np.spacing(np.maximum(np.abs(a), np.abs(b)))[d])


def test_static_entry_point_matches_caller_accumulator(oracle):
    lib = N.lib()
    assert lib.tg_static_acc_capacity() == 1 << 24
    s = S.stream_handle()
    # second call smaller than the first: the static array holds the first call's
    # range, which must not leak into the second's result
    for n, npix, nv, seed in ((72, 96, 12, 1), (40, 48, 7, 2)):
        print("static entry point", n, npix, nv, flush=True)
        ch, hs, y = _case(n, npix, nv, seed)
        yd = torch.from_numpy(y).cuda()
        ref_cube = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        acc = torch.zeros(n * n * n, dtype=torch.float64, device="cuda")
        c, keep = _call(ch, hs, yd, ref_cube, N.TG_OBSTACLE_SUN | N.TG_ACC_F64, acc=acc)
        N.check(lib.tg_conic_full_projector(ctypes.byref(c), s))
        st_cube = torch.full((n, n, n), 0.5, dtype=torch.float32, device="cuda")
        c2, keep2 = _call(ch, hs, yd, st_cube, N.TG_OBSTACLE_SUN)
        N.check(lib.tg_conic_full_projector_static(ctypes.byref(c2), s))
        torch.cuda.synchronize()
        S.check_status()
        got = st_cube.cpu().numpy() - np.float32(0.5)  # accumulates into the cube in place
        want = ref_cube.cpu().numpy()
        assert float(acc.abs().max()) == 0.0  # the caller's accumulator is left zero
        bref = np.zeros((n, n, n), np.float32)
        oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref, y, sun=True, nthread=1)
        err = rel_l2(want, bref)
        err_static = rel_l2(got, bref)
        print("rel L2 to the oracle: caller accumulator", err, "static", err_static, flush=True)
        assert err <= 1e-5 and err_static <= 1e-5
        # same float64 sums rounded once: equal but for one-ulp ties (and the
        # rounding of adding them to 0.5)
        st2 = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        c3, keep3 = _call(ch, hs, yd, st2, N.TG_OBSTACLE_SUN)
        N.check(lib.tg_conic_full_projector_static(ctypes.byref(c3), s))
        torch.cuda.synchronize()
        assert _ulp_equal(st2.cpu().numpy(), want)


def test_static_entry_point_time_bins_and_limits(oracle):
    """bpjt through the static array (views to bins by tindex), and the size and
    mode checks that reject a call before anything runs."""
    lib = N.lib()
    s = S.stream_handle()
    n, npix, nv, nt = 32, 40, 8, 3
    ch, hs, y = _case(n, npix, nv, 3)
    tindex = np.arange(nv, dtype=np.int32) % nt
    yd = torch.from_numpy(y).cuda()
    ti = torch.from_numpy(tindex).cuda()
    cube = torch.zeros((nt, n, n, n), dtype=torch.float32, device="cuda").permute(1, 2, 3, 0)
    c, keep = _call(ch, hs, yd, cube, 0, kind=N.TG_BPJT, nt=nt, tindex=ti)
    N.check(lib.tg_conic_full_projector_static(ctypes.byref(c), s))
    torch.cuda.synchronize()
    ref = np.zeros((n, n, n, nt), np.float32)
    oracle.siddon(oracle.BPJT, view_records(hs), map_record(ch), ref, y, nthread=1, tindex=tindex)
    assert rel_l2(cube.cpu().numpy(), ref) <= 1e-5
    c.flags |= N.TG_DETERMINISTIC
    assert lib.tg_conic_full_projector_static(ctypes.byref(c), s) == N.TG_EINVAL
    c.flags &= ~N.TG_DETERMINISTIC
    c.map.nt = (1 << 24) // (n * n * n) + 1  # more voxels than the array holds
    assert lib.tg_conic_full_projector_static(ctypes.byref(c), s) == N.TG_EINVAL
    assert b"static accumulator" in lib.tg_last_error()
