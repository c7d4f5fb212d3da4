"""GPU checks of what sits on the projector pair: the fused Thomson operator
against its unfused composition, the time-binned operator against the
reference's per-group loop, the regulariser and CG against numpy/scipy, and
the autograd Function against finite differences and the oracle."""
import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd import cg as CG
from tomograpy_amd import models as Mo
from tomograpy_amd import workloads as W
from tomograpy_amd.autograd import project
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import Siddon, Siddon4d, siddon_lo
from tests import cases as C
from tests.test_models import regularizer_reference

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_thomson_fused_equals_composition(oracle, dtype):
    """O.P_sun.M fused in the kernels == diag(O) (P_sun (diag(M) x)) computed unfused;
    forward bit-exact (same roundings), adjoint via the oracle to the bpj tolerance."""
    ch, x, data = C.small_case(n=24, npix=32, nv=6, dtype=dtype, radius=10.0)
    Pt = Mo.thomson_lo(data.header, ch, u=0.5, dtype=dtype)
    O = Mo.pb_data_coef(data.header, 32, 32).astype(dtype)
    M = Mo.pb_map_coef(ch, 0.5).astype(dtype)
    xd = torch.from_numpy(np.asarray(x)).cuda()
    y = Pt.matvec(xd).cpu().numpy()
    mx = (M * np.asarray(x)).astype(dtype)
    ref = np.zeros(data.shape, dtype)
    oracle.siddon(oracle.PJ, view_records(data.header), map_record(ch), mx, ref, sun=True)
    ref = (O * ref).astype(dtype)
    assert np.array_equal(y, ref)
    rng = np.random.default_rng(1)
    yy = rng.normal(size=data.shape).astype(dtype)
    g = Pt.rmatvec(torch.from_numpy(yy).cuda()).cpu().numpy()
    oy = T.InfoArray((O * yy).astype(dtype), header=data.header)
    gref = np.zeros(x.shape, dtype)
    oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), gref, oy, sun=True, nthread=1)
    gref = (M * gref).astype(dtype)
    assert rel_l2(g, gref) <= (1e-5 if dtype == np.float32 else 1e-12)


def test_siddon4d_matches_reference_group_loop(oracle):
    """Siddon4d(ng) == the reference's loop of projector4d over data[..., i::ng]
    (lo_wrapper.py:49-57): image i + j*ng reads time bin j."""
    ch, x, data = C.small_case(n=16, npix=24, nv=8, dtype=np.float64)
    ng = 2
    nt = 4
    rng = np.random.default_rng(2)
    cube4 = rng.uniform(size=(16, 16, 16, nt))
    P4 = Siddon4d(data.header, ch, ng=ng, dtype=np.float64)
    y = P4.matvec(torch.from_numpy(cube4).cuda()).cpu().numpy()
    ref = np.zeros(data.shape)
    recs = view_records(data.header)
    for i in range(ng):
        sub = np.ascontiguousarray(ref[..., i::ng])
        oracle.siddon(oracle.PJT, recs[i::ng], map_record(ch), cube4, sub)
        ref[..., i::ng] = sub
    assert np.array_equal(y, ref)


def test_regularizer_matches_numpy():
    rng = np.random.default_rng(3)
    for shape, axes in (((6, 5, 4, 1), (0, 1, 2)), ((4, 5, 6, 3), (0, 1, 2, 3)), ((7, 3, 5, 1), (1,))):
        x = rng.normal(size=shape)
        out0 = rng.normal(size=shape)
        import ctypes
        from tomograpy_amd import _native as N, siddon as S
        xd = torch.from_numpy(x).cuda()
        od = torch.from_numpy(out0.copy()).cuda()
        bits = sum(1 << a for a in axes)
        shp = (ctypes.c_int32 * 4)(*shape)
        N.check(N.lib().tg_regularize(N.TG_F64, shp, bits, 0.3, 1, xd.data_ptr(), od.data_ptr(),
                                      None, None, S.stream_handle()))
        ref = out0 + regularizer_reference(x, axes, 0.3, True)
        np.testing.assert_allclose(od.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)


def test_cg_matches_scipy_fp64(oracle):
    """Device CG == scipy.sparse.linalg.cg on the oracle operator, same iterations."""
    from scipy.sparse.linalg import LinearOperator, cg as scipy_cg
    ch, x, data = C.small_case(n=12, npix=16, nv=6, dtype=np.float64, radius=8.0)
    P = siddon_lo(data.header, ch, dtype=np.float64, obstacle="sun")
    rng = np.random.default_rng(4)
    y = rng.uniform(size=data.shape)
    lam = 1e-2
    recs = view_records(data.header)
    mrec = map_record(ch)

    def A(v):
        v = v.reshape(12, 12, 12)
        pv = np.zeros(data.shape)
        oracle.siddon(oracle.PJ, recs, mrec, np.ascontiguousarray(v), pv, sun=True)
        out = np.zeros((12, 12, 12))
        oracle.siddon(oracle.BPJ, recs, mrec, out, pv, sun=True, nthread=1)
        return (out + regularizer_reference(v, (0, 1, 2), lam, True)).ravel()

    b = np.zeros((12, 12, 12))
    oracle.siddon(oracle.BPJ, recs, mrec, b, np.ascontiguousarray(y), sun=True, nthread=1)
    # the operator itself first: one application against the oracle's
    A_gpu = CG.NormalOperator(P, lam=lam, identity=True)
    v = rng.uniform(size=(12, 12, 12))
    av = A_gpu.apply_into(torch.from_numpy(v).cuda(), torch.zeros((12, 12, 12), dtype=torch.float64,
                                                                    device="cuda"))
    err_a = rel_l2(av.cpu().numpy().ravel(), A(v.ravel()))
    print("operator rel err", err_a, flush=True)
    assert err_a < 1e-13
    for its in (1, 2, 8):
        op = LinearOperator((b.size, b.size), matvec=A, dtype=np.float64)
        xs, _ = scipy_cg(op, b.ravel(), maxiter=its, rtol=1e-30, atol=0)
        xg, info = CG.reconstruct(P, torch.from_numpy(y).cuda(), lam=lam, maxiter=its)
        err = rel_l2(xg.cpu().numpy().ravel(), xs)
        print("cg iterations", its, "rel err", err, flush=True)
        assert info["iterations"] == its
        assert err < 1e-9, (its, err)


def test_cg_converges_to_tolerance():
    ch, x, data = C.small_case(n=16, npix=24, nv=12, dtype=np.float32, radius=10.0)
    P = siddon_lo(data.header, ch, dtype=np.float32)
    y = P.matvec(torch.from_numpy(np.asarray(x)).cuda()).clone()
    xr, info = CG.reconstruct(P, y, lam=1e-2, maxiter=200, tol=1e-4, check_every=5)
    assert info["residual"] <= 1e-4
    assert info["iterations"] < 200


def test_autograd_gradient_fd_and_oracle(oracle):
    """dt + masked operator: autograd gradient of ||P x - y||^2 + priors against
    central finite differences (fp64, h = 1e-3) on 64 voxels and against the
    oracle backprojector."""
    n, nt, nv, npix = 16, 4, 8, 24
    ch, _, data = C.small_case(n=n, npix=npix, nv=nv, dtype=np.float64, radius=8.0)
    tindex = np.arange(nv) % nt
    voxmask = W.voxel_radius(ch) < 1.05
    P = Siddon4d(data.header, ch, tindex=tindex, voxmask=voxmask, obstacle="sun",
                 dtype=np.float64)
    rng = np.random.default_rng(5)
    x0 = torch.from_numpy(rng.uniform(size=(n, n, n, nt))).cuda()
    y = torch.from_numpy(rng.uniform(size=(npix, npix, nv))).cuda()

    def loss_fn(x):
        r = project(x, P) - y
        l = (r * r).sum()
        for ax in range(3):
            l = l + 1e-2 * (torch.diff(x, dim=ax) ** 2).sum()
        l = l + 1e-2 * (torch.diff(x, dim=3) ** 2).sum()
        return l

    x = x0.clone().requires_grad_(True)
    loss_fn(x).backward()
    g = x.grad.detach()
    idx = rng.integers(0, n, size=(64, 3))
    bins = rng.integers(0, nt, size=64)
    h = 1e-3
    for (a, b, c), t in zip(idx, bins):
        xp = x0.clone()
        xm = x0.clone()
        xp[a, b, c, t] += h
        xm[a, b, c, t] -= h
        fd = (float(loss_fn(xp)) - float(loss_fn(xm))) / (2 * h)
        assert abs(fd - float(g[a, b, c, t])) <= 1e-6 * max(1.0, abs(fd)), (a, b, c, t)
    # data-term gradient vs the oracle backprojector: 2 P^T (P x - y), masked voxels zero
    r = (project(x0, P) - y).cpu().numpy()
    gref = np.zeros((n, n, n, nt))
    oracle.siddon(oracle.BPJT, view_records(data.header), map_record(ch), gref,
                  np.ascontiguousarray(2 * r), sun=True, nthread=1, tindex=tindex)
    gref[voxmask] = 0
    x1 = x0.clone().requires_grad_(True)
    ((project(x1, P) - y) ** 2).sum().backward()
    assert rel_l2(x1.grad.cpu().numpy(), gref) <= 1e-12


@pytest.mark.parametrize("layout", ["bins_fastest", "bin_major", "cube"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_vox_apply_layouts(layout, dtype):
    """tg_vox_apply (f * x, zero where masked, into the operator's bin-major
    buffer) for a source with the bins fastest (the tiled transpose path), a
    bin-major source and a plain cube, against the same product in torch."""
    print("vox_apply", layout, dtype, flush=True)
    ch = W.cube_header(37, dtype=np.float32 if dtype == torch.float32 else np.float64)
    hs = W.view_headers(16, np.arange(6) * 0.9, d=214.0,
                        dtype=np.float32 if dtype == torch.float32 else np.float64)
    g = torch.Generator().manual_seed(3)
    w = torch.rand((37, 37, 37), generator=g, dtype=torch.float64).to(dtype)
    mask = torch.rand((37, 37, 37), generator=g) < 0.3
    if layout == "cube":
        P = Siddon(hs, ch, vox_weight=w.numpy(), voxmask=mask.numpy(),
                   dtype=np.float32 if dtype == torch.float32 else np.float64)
        x = torch.randn((37, 37, 37), generator=g, dtype=torch.float64).to(dtype).cuda()
        ref = torch.where(mask.cuda(), torch.zeros_like(x), w.cuda() * x)
    else:
        tindex = np.arange(6) % 3
        P = Siddon4d(hs, ch, tindex=tindex, voxmask=mask.numpy(),
                     dtype=np.float32 if dtype == torch.float32 else np.float64)
        x = torch.randn((37, 37, 37, 3), generator=g, dtype=torch.float64).to(dtype).cuda()
        if layout == "bin_major":
            x = x.permute(3, 0, 1, 2).contiguous().permute(1, 2, 3, 0)
        ref = torch.where(mask.cuda()[..., None], torch.zeros_like(x), x)
    out = P.vox_apply(x)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
