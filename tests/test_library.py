"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
packs view descriptors exactly, and rejects bad arguments -- no GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

import tomograpy_amd as T
from tomograpy_amd import _native as N
from tomograpy_amd.headers import view_records
from tests import cases as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    lib = N.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    bound = {s[0] for s in N.SIGNATURES}
    assert set(names) <= bound, set(names) - bound
    assert b"gfx950" in lib.tg_version()


def test_integration_notes_list_every_declared_symbol():
    """The integration notes name every entry point the header declares."""
    notes = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    missing = [n for n in declared_functions() if n not in notes]
    assert not missing, missing


def test_missing_library_raises_instead_of_falling_back(monkeypatch):
    """Without the HIP library the package refuses to compute: no CPU path."""
    from tomograpy_amd import siddon as S
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", os.path.join(ROOT, "no-such-dir", "libtomograpy_amd.so"))
    with pytest.raises(ImportError, match="missing"):
        N.lib()
    # a product helper reaches the library before any device work
    with pytest.raises(ImportError, match="missing"):
        S.views_tensor(np.zeros((1, 18)), N.TG_F32, 4, 4, "cpu")


def test_struct_layout_matches_header():
    # sizes of the POD structs as the C compiler lays them out
    assert ctypes.sizeof(N.TgMap) == 12 * 8 + 4 * 4 + 4 * 8
    assert ctypes.sizeof(N.TgCall) % 8 == 0
    assert N.TgCall.map.offset == 16


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_views_pack_matches_oracle_unit_vectors(oracle, dtype):
    """The packed cos/sin tables + CT-rounded headers give the oracle's unit vectors
    when the kernel's arithmetic is replayed in numpy in the same order."""
    ch, cube, data = C.small_case(npix=13, nv=5, dtype=dtype)
    recs = view_records(data.header)
    code = N.TG_F32 if dtype == np.float32 else N.TG_F64
    n1, n2, nv = data.shape
    nbytes = N.lib().tg_views_bytes(code, nv, n1, n2)
    buf = np.zeros(nbytes, dtype=np.uint8)
    N.check(N.lib().tg_views_pack(code, nv, n1, n2, recs.ctypes.data, buf.ctypes.data, nbytes))
    hdr_bytes = (np.dtype(dtype).itemsize * 18 * nv + 255) // 256 * 256
    hdr = buf[: 18 * nv * np.dtype(dtype).itemsize].view(dtype).reshape(nv, 18)
    np.testing.assert_array_equal(hdr, recs.astype(dtype))
    trig = buf[hdr_bytes:].view(np.float64).reshape(nv, n1 + n2, 2)
    u_ref = oracle.full_unit_vector(recs, n1, n2, dtype)
    for t in range(nv):
        R = hdr[t, 6:15].reshape(3, 3)
        cg, sg = trig[t, :n1, 0][:, None], trig[t, :n1, 1][:, None]
        cl, sl = trig[t, n1:, 0][None, :], trig[t, n1:, 1][None, :]
        u2 = [(cl * cg).astype(dtype), (cl * sg).astype(dtype), np.broadcast_to(sl, (n1, n2)).astype(dtype)]
        for k in range(3):
            uk = (R[k, 0] * u2[0] + R[k, 1] * u2[1]) + R[k, 2] * u2[2]
            np.testing.assert_array_equal(uk, u_ref[:, :, t, k])


def test_views_pack_rejects_small_buffer():
    recs = np.zeros((2, 18))
    buf = np.zeros(16, dtype=np.uint8)
    rc = N.lib().tg_views_pack(N.TG_F32, 2, 4, 4, recs.ctypes.data, buf.ctypes.data, 16)
    assert rc == N.TG_EINVAL
    assert b"too small" in N.lib().tg_last_error()
    with pytest.raises(N.NativeError):
        N.check(rc)


def test_bad_dtype_and_headers_raise_before_device():
    ch, cube, data = C.small_case(dtype=np.float64)
    d32 = T.InfoArray(np.zeros(data.shape, np.float32), header=data.header)
    with pytest.raises(ValueError, match="same data-type"):
        T.projector(d32, cube)
    bad = T.InfoArray(np.zeros(data.shape), header=tuple(data.header))
    with pytest.raises(ValueError, match="should be a list"):
        T.projector(bad, cube)
    di = T.InfoArray(np.zeros(data.shape, np.int32), header=data.header)
    ci = T.InfoArray(np.zeros(cube.shape, np.int32), header=cube.header)
    with pytest.raises(KeyError):
        T.projector(di, ci)
    with pytest.raises(KeyError):
        T.projector(data, cube, obstacle="moon")


def test_call_validation_errors_without_gpu():
    """make_args rejects bad shapes before any launch."""
    call = N.TgCall()
    call.dtype = N.TG_F32
    call.kind = N.TG_PJ
    call.nv = 1
    call.n1 = call.n2 = 4
    rc = N.lib().tg_conic_full_projector(ctypes.byref(call), None)
    assert rc == N.TG_EINVAL
    call.map.n[0] = call.map.n[1] = call.map.n[2] = 4
    call.kind = 7
    assert N.lib().tg_conic_full_projector(ctypes.byref(call), None) == N.TG_EINVAL
    assert b"kind" in N.lib().tg_last_error()
    call.kind = N.TG_PJ
    call.tile_j = 3
    assert N.lib().tg_conic_full_projector(ctypes.byref(call), None) == N.TG_EINVAL
    call.tile_j = 16
    assert N.lib().tg_conic_image_projector(ctypes.byref(call), 5, None) == N.TG_EINVAL


@pytest.mark.parametrize("dtype_code", [N.TG_F32, N.TG_F64])
@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("mixed", [False, True])
def test_batch_job_table_layout(dtype_code, schedule, mixed):
    """The staged job table has the stride and fields the persistent kernels read
    (entries JobDesc<T>[k] in queue order, grouped by variant -- here the
    obstacle -- with first-tile offsets counted from the group's start)."""
    rng = np.random.default_rng(0)
    njobs = 7
    calls = (N.TgCall * njobs)()
    expect_tiles = []
    sizes = []
    for k in range(njobs):
        c = calls[k]
        c.dtype = dtype_code
        c.kind = N.TG_PJ
        c.flags = N.TG_OBSTACLE_SUN if (mixed and k % 3 == 0) else 0
        n = int(rng.integers(8, 40))
        c.map.n[0] = c.map.n[1] = c.map.n[2] = n
        c.map.nt = 1
        c.n1 = c.n2 = int(rng.choice([16, 24, 40]))
        c.nv = int(rng.integers(1, 6))
        c.views = 0x1000 + k
        c.cube = 0x100000 * (k + 1)
        c.data = 0x200000 * (k + 1)
        sizes.append(3 * n)
        tj = 16
        bi = 256 // tj
        expect_tiles.append(c.nv * ((c.n1 + bi - 1) // bi) * ((c.n2 + tj - 1) // tj))
    plan = N.TgBatchPlan()
    plan.tile_j = 16
    plan.schedule = schedule
    nbytes = N.lib().tg_batch_queue_bytes(njobs)
    stage = np.zeros(nbytes, dtype=np.uint8)
    N.check(N.lib().tg_batch_plan_table(calls, njobs, ctypes.byref(plan), stage.ctypes.data))
    jb, to, co = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    N.check(N.lib().tg_batch_job_layout(dtype_code, ctypes.byref(jb), ctypes.byref(to), ctypes.byref(co)))
    assert 256 + njobs * jb.value <= nbytes
    group = [int(calls[k].flags & N.TG_OBSTACLE_SUN) for k in range(njobs)]
    # stable: by group, then (schedule 1) longest rays first
    order = sorted(range(njobs), key=lambda k: (group[k], -sizes[k] if schedule == 1 else 0))
    hdr = stage[:256].view(np.uint64)
    assert hdr[1] == sum(expect_tiles) and hdr[2] == njobs
    begin = {}
    for q, k in enumerate(order):
        base = 256 + q * jb.value
        tb = int(stage[base + to.value: base + to.value + 8].view(np.uint64)[0])
        cube = int(stage[base + co.value: base + co.value + 8].view(np.uint64)[0])
        assert tb == begin.get(group[k], 0), (q, k)
        assert cube == calls[k].cube, (q, k)
        begin[group[k]] = begin.get(group[k], 0) + expect_tiles[k]


@pytest.mark.parametrize("dtype_code,n,brick", [(N.TG_F32, 40, (0, 0, 0)), (N.TG_F64, 70, (0, 0, 0)),
                                                (N.TG_F32, 130, (0, 0, 0)),
                                                (N.TG_F32, 33, (8, 16, 32))])
def test_bpj_plan_init_fields(dtype_code, n, brick):
    """tg_bpj_plan_init (host only) fills the fields where the ctypes mirror expects them."""
    call = N.TgCall()
    call.dtype = dtype_code
    call.kind = N.TG_BPJT
    call.map.n[0] = call.map.n[1] = call.map.n[2] = n
    call.map.nt = 3
    call.n1 = call.n2 = 16
    call.nv = 4
    p = N.TgBpjPlan()
    for k in range(3):
        p.brick[k] = brick[k]
    p.order = 1
    N.check(N.lib().tg_bpj_plan_init(ctypes.byref(call), ctypes.byref(p)))
    b = list(p.brick)
    if brick == (0, 0, 0):
        # 128 KiB of double accumulators, 64 KiB for cubes of edge <= 96
        assert b == ([16, 16, 32] if n <= 96 else [32, 16, 32])
    nb = [(n + b[k] - 1) // b[k] for k in range(3)]
    assert list(p.nb_axis) == nb
    assert p.nbricks == nb[0] * nb[1] * nb[2] * 3
    assert p.order == 1 and p.records == 0 and p.chunks == 0
    assert p.chunk_records == 0  # chosen by the count
    assert p.meta_bytes >= p.nbricks * (4 + 8 + 4 + 8 + 4)
    p.brick[0] = 12
    assert N.lib().tg_bpj_plan_init(ctypes.byref(call), ctypes.byref(p)) == N.TG_EINVAL
    p.brick[0] = b[0]
    p.tie_spread = 11
    assert N.lib().tg_bpj_plan_init(ctypes.byref(call), ctypes.byref(p)) == N.TG_EINVAL
    assert b"tie_spread" in N.lib().tg_last_error()
    p.tie_spread = 10
    N.check(N.lib().tg_bpj_plan_init(ctypes.byref(call), ctypes.byref(p)))


def test_plan_batch_rejects_bad_arguments_without_gpu():
    """The plan batch validates its arguments before any device call."""
    lib = N.lib()
    assert lib.tg_bpj_plan_batch_bytes(4) > lib.tg_bpj_plan_batch_bytes(1) > 0
    calls = (N.TgCall * 1)()
    plans = (N.TgBpjPlan * 1)()
    assert lib.tg_bpj_plan_batch(calls, plans, 0, None, None, None) == N.TG_EINVAL
    assert b"empty" in lib.tg_last_error()
    assert lib.tg_bpj_plan_batch(calls, plans, 1, None, None, None) == N.TG_EINVAL
    assert b"staging" in lib.tg_last_error()
    stage = np.zeros(int(lib.tg_bpj_plan_batch_bytes(1)), dtype=np.uint8)
    calls[0].dtype = N.TG_F32
    calls[0].kind = N.TG_PJ  # not a backprojection: rejected while the table is staged,
    # before the table copy (the queue pointer is never used here)
    rc = lib.tg_bpj_plan_batch(calls, plans, 1, stage.ctypes.data, stage.ctypes.data, None)
    assert rc == N.TG_EINVAL
    assert b"bpj" in lib.tg_last_error()


def test_plan_trace_tile_and_tie_spread_fields_follow_order():
    """tg_bpj_plan.trace_tile, tie_spread and pad sit between order and records,
    as in the header."""
    assert N.TgBpjPlan.trace_tile.offset == N.TgBpjPlan.order.offset + 4
    assert N.TgBpjPlan.tie_spread.offset == N.TgBpjPlan.trace_tile.offset + 4
    assert N.TgBpjPlan.records.offset == N.TgBpjPlan.tie_spread.offset + 8


def test_auto_tie_spread_follows_the_voxel_to_pixel_ratio():
    """Plans spread ties only when a voxel spans several pixels (about 1.9 pixels
    per voxel at 256^3 seen in 512^2, 3.7 at 128^3, 8.7 at 55^3)."""
    import types
    from tomograpy_amd import workloads as W
    from tomograpy_amd.headers import map_record
    from tomograpy_amd.lo_wrapper import Siddon

    def spread(n, npix):
        ch, hs, _ = W.config2_geometry(np.float32, n=n, n_views=8, n_pix=npix)
        return Siddon.auto_tie_spread(types.SimpleNamespace(data_header=hs, map_rec=map_record(ch)))

    assert spread(256, 512) == 0
    assert spread(128, 256) == 0
    assert spread(128, 512) == 6
    assert spread(55, 512) == 10


def test_smoothness_rejects_bad_arguments_without_gpu():
    """tg_smoothness validates its arguments before any device call."""
    lib = N.lib()
    shape = (ctypes.c_int32 * 4)(4, 4, 1, 1)
    w = (ctypes.c_double * 4)(1.0, 1.0, 0.0, 0.0)
    assert lib.tg_smoothness(N.TG_F32, shape, w, None, None, None, None, 0, None) == N.TG_EINVAL
    assert b"tg_smoothness" in lib.tg_last_error()
    bad = (ctypes.c_int32 * 4)(4, 0, 1, 1)
    x = np.zeros(16, np.float32)
    scal = np.zeros(N.TG_CG_SCAL, np.float64)
    assert lib.tg_smoothness(N.TG_F32, bad, w, x.ctypes.data, None, None, scal.ctypes.data, 0,
                             None) == N.TG_EINVAL
    assert lib.tg_smoothness(N.TG_F32, shape, w, x.ctypes.data, None, None, scal.ctypes.data, 16,
                             None) == N.TG_EINVAL


@pytest.mark.parametrize("share", ["acc", "cube", "overlap", "none"])
def test_batch_rejects_shared_accumulators(share):
    """A backward batch's float64 accumulators are finished by one launch, each
    job reading and clearing its own and adding it to its cube without atomics:
    tg_batch_plan_table (host only) rejects two jobs whose accumulators, or whose
    cubes, overlap."""
    calls = (N.TgCall * 3)()
    for k in range(3):
        c = calls[k]
        c.dtype = N.TG_F32
        c.kind = N.TG_BPJ
        c.flags = N.TG_ACC_F64
        c.map.n[0] = c.map.n[1] = c.map.n[2] = 10
        c.map.nt = 1
        c.map.stride[0], c.map.stride[1], c.map.stride[2] = 100, 10, 1
        c.n1 = c.n2 = 16
        c.nv = 2
        c.data_stride[0], c.data_stride[1], c.data_stride[2] = 16, 1, 256
        c.views = 0x1000 + k
        c.cube = 0x100000 * (k + 1)
        c.data = 0x900000 * (k + 1)
        c.acc = 0x4000000 * (k + 1)
    if share == "acc":
        calls[2].acc = calls[0].acc
    elif share == "cube":
        calls[1].cube = calls[0].cube
    elif share == "overlap":
        calls[1].acc = calls[0].acc + 8 * 999  # the last double of job 0's range
    stage = np.zeros(N.lib().tg_batch_queue_bytes(3), dtype=np.uint8)
    plan = N.TgBatchPlan()
    rc = N.lib().tg_batch_plan_table(calls, 3, ctypes.byref(plan), stage.ctypes.data)
    if share == "none":
        assert rc == N.TG_OK
    else:
        assert rc == N.TG_EINVAL
        assert b"share" in N.lib().tg_last_error() or b"same cube" in N.lib().tg_last_error()
