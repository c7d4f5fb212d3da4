"""GPU checks of the brick plan's pass sort (tg_bpj_plan order 2): inside
every pass of 1024 records of a chunk the records run by decreasing step
count, and each brick holds exactly the records of the unsorted plan (order
0) of the same geometry.  Usual bricks (every record under 256 steps) take
the counting sort; bricks 256 voxels long, crossed end to end, take the
bitonic network.  Both plans backproject to the oracle's result."""
import numpy as np
import pytest
import torch

from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import Siddon
from tests import cases as C

pytestmark = pytest.mark.gpu

PASS = 1024


def _align(x):
    return (x + 255) // 256 * 256


def plan_records(P, dtype):
    """The plan's brick offsets, chunk size and record fields, read back from
    its device buffers (layout of tg_bpj_plan's meta and rec)."""
    p = P.build_plan()
    meta, rec = P._plan_bufs
    nb, R = int(p.nbricks), int(p.records)
    m = meta.cpu().numpy().view(np.uint8)
    at_off = _align(4 * nb)
    offsets = m[at_off:at_off + 8 * (nb + 1)].view(np.uint64).astype(np.int64)
    r = rec.cpu().numpy().view(np.uint8)
    isz = np.dtype(dtype).itemsize
    fields, o = [], 0
    for dt, size in ((np.uint32, 4), (np.uint32, 4), (dtype, isz), (dtype, isz), (dtype, isz)):
        fields.append(r[o:o + size * R].view(dt))
        o = _align(o + size * R)
    return offsets, int(p.chunk_records), fields


def check_sorted_and_same(P2, P0, dtype):
    off2, chunk, (rid2, lq2, a2, b2, c2) = plan_records(P2, dtype)
    off0, _, (rid0, lq0, a0, b0, c0) = plan_records(P0, dtype)
    assert np.array_equal(off2, off0)
    long_passes = 0
    for b in range(len(off2) - 1):
        s, e = off2[b], off2[b + 1]
        if s == e:
            continue
        steps = lq2[s:e] >> 16
        for c0_ in range(s, e, chunk):
            for w in range(c0_, min(c0_ + chunk, e), PASS):
                win = steps[w - s:min(w + PASS, c0_ + chunk, e) - s]
                assert np.all(win[:-1] >= win[1:]), (b, w)
                long_passes += int(win.max() > 255)
        i2 = np.argsort(rid2[s:e], kind="stable") + s
        i0 = np.argsort(rid0[s:e], kind="stable") + s
        for x2, x0 in ((rid2, rid0), (lq2, lq0), (a2, a0), (b2, b0), (c2, c0)):
            assert np.array_equal(x2[i2], x0[i0]), b
    return long_passes


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pass_sort_usual_bricks(oracle, dtype):
    """64^3 cube, 20 views of 112^2, default bricks: counting-sorted passes."""
    print("plan sort usual bricks", np.dtype(dtype).name, flush=True)
    ch, x, data = C.small_case(n=64, npix=112, nv=20, dtype=dtype, seed=4)
    P2 = Siddon(data.header, ch, dtype=dtype, bpj_mode="plan", plan_order=2)
    P0 = Siddon(data.header, ch, dtype=dtype, bpj_mode="plan", plan_order=0)
    assert check_sorted_and_same(P2, P0, dtype) == 0
    rng = np.random.default_rng(21)
    y = rng.uniform(-1, 1, data.shape).astype(dtype)
    ref = np.zeros(x.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, y, nthread=1)
    got = P2.rmatvec(torch.from_numpy(y).cuda()).cpu().numpy()
    tol = 1e-5 if dtype == np.float32 else 1e-12
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


def test_pass_sort_long_records():
    """256^3 cube seen along axis 0 (2 views, 64^2) with 256 x 8 x 8 bricks:
    records of more than 255 steps, whose passes take the bitonic network,
    beside passes of short records in the same plan."""
    print("plan sort long records", flush=True)
    dtype = np.float32
    ch, x, data = C.small_case(n=256, npix=64, nv=2, dtype=dtype, seed=5, lon_offset=0.0, lat=0.0)
    P2 = Siddon(data.header, ch, dtype=dtype, bpj_mode="plan", plan_order=2, brick=(256, 8, 8))
    P0 = Siddon(data.header, ch, dtype=dtype, bpj_mode="plan", plan_order=0, brick=(256, 8, 8))
    assert check_sorted_and_same(P2, P0, dtype) > 0
    rng = np.random.default_rng(22)
    y = torch.from_numpy(rng.uniform(-1, 1, data.shape).astype(dtype)).cuda()
    a, b = P2.rmatvec(y).cpu().numpy(), P0.rmatvec(y).cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
