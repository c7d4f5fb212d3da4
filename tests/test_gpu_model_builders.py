"""GPU checks of the model builders end to end (reference tests/test_models.py:14-37)
and of the height prior (models.py:85-104) in the regulariser, the CG solver
and the differentiable prior."""
import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd import cg as CG
from tomograpy_amd import models as Mo
from tomograpy_amd.autograd import smoothness
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import siddon_lo
from tests import cases as C
from tests.test_models import weighted_regularizer_reference

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _reference_case():
    """The fixture of tests/test_models.py:17-24 at its float64 sizes
    (test_cases.py: a 16^3 cube of side 3 and 32 x 32 images of fov(cube, 200)):
    100 views on a circle of radius 200 over 3 pi (linspace, endpoint included),
    all taken at one time, and an object of ones."""
    obj_h = C.object_headers(np.float64)[1]
    im_h = dict(C.image_headers(np.float64)[1])
    im_h["DATE_OBS"] = "2010-01-01T00:00:00Z"
    im_h["n_images"] = 100
    im_h["max_lon"] = 3 * np.pi
    data = T.circular_trajectory_data(**im_h)
    obj = T.object_from_header(obj_h)
    obj[:] = 1.0
    return obj, data


@pytest.mark.parametrize("model", ["srt", "stsrt", "thomson"])
def test_model_reconstructs_constant_object(model):
    """check_model of tests/test_models.py:14-33: with obj_rmin = 1, data
    simulated from a constant object outside the Sun, CG on the model with
    tiny smoothness hyper-parameters gives the object back on every voxel a
    ray sees, to one decimal."""
    print("model", model, flush=True)
    obj, data = _reference_case()
    kw = dict(obj_rmin=1.0, decimate=True)
    if model == "srt":
        P, D, obj_mask, data_mask = Mo.srt(data, obj, **kw)
    elif model == "stsrt":
        P, D, obj_mask, data_mask = Mo.stsrt(data, obj, dt_min=1.0, **kw)
    else:
        P, D, obj_mask, data_mask = Mo.thomson(data, obj, **kw)
    assert obj_mask is not None and obj_mask.any()
    ones = torch.ones(P.shapeout, dtype=P.tdtype, device="cuda")
    w = P.rmatvec(ones).clone()
    if model == "stsrt":
        assert P.nt == 1  # one acquisition time: one bin
        w = w[..., 0]
    is_seen = (w.cpu().numpy() != 0)
    assert is_seen.sum() > 0.5 * (~obj_mask).sum()
    new_obj = (1.0 - obj_mask).astype(np.float64)
    x_in = torch.from_numpy(new_obj[..., None] if model == "stsrt" else new_obj).cuda()
    y = P.matvec(x_in).clone()
    x, info = CG.solve(P, y, lam=1e-10, identity=False, axes=D, maxiter=1000, tol=1e-13,
                       check_every=50)
    sol = x.detach().cpu().numpy()
    if model == "stsrt":
        sol = sol[..., 0]
    print(model, "iterations", info["iterations"], "residual", info["residual"],
          "max error on seen voxels", float(np.abs(sol[is_seen] - 1.0).max()), flush=True)
    np.testing.assert_almost_equal(sol[is_seen], new_obj[is_seen], decimal=1)
    assert np.all(sol[obj_mask] == 0.0)  # masked unknowns stay out of the solve


@pytest.mark.parametrize("identity", [True, False])
def test_weighted_regularizer_matches_matrix(identity):
    """tg_regularize with the height-prior weight == lam (I + sum_k (D_k W)^T (D_k W)) x
    built from explicit difference matrices."""
    import ctypes
    from tomograpy_amd import _native as N, siddon as S
    rng = np.random.default_rng(11)
    shape = (6, 5, 4)
    x = rng.normal(size=shape)
    w = rng.uniform(0.5, 3.0, size=shape)
    out0 = rng.normal(size=shape)
    xd = torch.from_numpy(x).cuda()
    wd = torch.from_numpy(w).cuda()
    od = torch.from_numpy(out0.copy()).cuda()
    shp = (ctypes.c_int32 * 4)(*shape, 1)
    N.check(N.lib().tg_regularize(N.TG_F64, shp, 0b111, 0.3, int(identity), xd.data_ptr(),
                                  od.data_ptr(), None, wd.data_ptr(), S.stream_handle()))
    ref = out0 + weighted_regularizer_reference(x, w, (0, 1, 2), 0.3, identity)
    np.testing.assert_allclose(od.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_weighted_smoothness_matches_torch(dtype):
    """autograd.smoothness with a voxel weight v == sum_k w_k ||diff_k(v x)||^2 in torch,
    value and gradient."""
    g = torch.Generator().manual_seed(5)
    x0 = torch.randn((7, 6, 5), generator=g, dtype=torch.float64).to(dtype).cuda()
    v = (0.5 + torch.rand((7, 6, 5), generator=g, dtype=torch.float64)).to(dtype).cuda()
    wts = (1.0, 0.5, 2.0)
    xa = x0.clone().requires_grad_(True)
    xb = x0.clone().requires_grad_(True)
    va = smoothness(xa, wts, voxel_weight=v)
    vb = sum(w * (torch.diff(v * xb, dim=ax) ** 2).sum() for ax, w in enumerate(wts))
    va.backward()
    vb.backward()
    rtol = 1e-12 if dtype == torch.float64 else 1e-5
    assert abs(float(va.detach()) - float(vb.detach())) <= rtol * abs(float(vb.detach()))
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert torch.allclose(xa.grad, xb.grad, rtol=tol, atol=tol * float(xb.grad.abs().max()))


def test_cg_height_prior_matches_scipy(oracle):
    """CG on (P^T P + lam sum_k (D_k R)^T (D_k R)) x = P^T y with R = diag(r^2)
    (srt with height_prior=True, models.py:52 and 85-104) == scipy CG on the
    oracle operator with the explicit matrices, iteration for iteration."""
    from scipy.sparse.linalg import LinearOperator, cg as scipy_cg
    n = 10
    ch, x, data = C.small_case(n=n, npix=14, nv=6, dtype=np.float64, radius=8.0)
    P = siddon_lo(data.header, ch, dtype=np.float64, obstacle="sun")
    prior = Mo.smoothness_prior(ch, 3, height_prior=True)
    rng = np.random.default_rng(6)
    y = rng.uniform(size=data.shape)
    lam = 1e-2
    recs = view_records(data.header)
    mrec = map_record(ch)
    wt = prior.weight

    def A(v):
        v = v.reshape(n, n, n)
        pv = np.zeros(data.shape)
        oracle.siddon(oracle.PJ, recs, mrec, np.ascontiguousarray(v), pv, sun=True)
        out = np.zeros((n, n, n))
        oracle.siddon(oracle.BPJ, recs, mrec, out, pv, sun=True, nthread=1)
        return (out + weighted_regularizer_reference(v, wt, (0, 1, 2), lam, False)).ravel()

    b = np.zeros((n, n, n))
    oracle.siddon(oracle.BPJ, recs, mrec, b, np.ascontiguousarray(y), sun=True, nthread=1)
    op = LinearOperator((b.size, b.size), matvec=A, dtype=np.float64)
    for its in (1, 3, 10):
        xs, _ = scipy_cg(op, b.ravel(), maxiter=its, rtol=1e-30, atol=0)
        xg, info = CG.solve(P, torch.from_numpy(y).cuda(), lam=lam, identity=False, axes=prior,
                            maxiter=its)
        err = rel_l2(xg.cpu().numpy().ravel(), xs)
        print("height prior cg iterations", its, "rel err", err, flush=True)
        assert info["iterations"] == its
        assert err < 1e-9, (its, err)
    # the prior changes the solve: without the weight the iterates differ
    xp, _ = CG.solve(P, torch.from_numpy(y).cuda(), lam=lam, identity=False,
                     axes=Mo.smoothness_prior(ch, 3), maxiter=10)
    assert rel_l2(xp.cpu().numpy().ravel(), xs) > 1e-6
