"""GPU checks at the full sizes of the differentiation workload (128^3 x 8
time bins, 64 views of 256^2, voxel mask) and of the batched-jobs workload
(64 uneven projector + backprojector jobs, half Thomson-weighted)."""
import numpy as np
import pytest
import torch

from tomograpy_amd import batch as B
from tomograpy_amd import workloads as W
from tomograpy_amd.headers import map_record, view_records

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_c3_gradient_against_finite_differences_fp64():
    """Autograd gradient of the full-size loss against central differences
    (h = 1e-3) on 64 random voxel/bin pairs, in float64.  For a quadratic loss
    the central difference is exact but for rounding: each loss value carries
    about eps * L of error, so the quotient carries about eps * L / h; the
    tolerance is 1e-7 relative plus 16 eps L / h."""
    print("C3 fp64 finite differences", flush=True)
    P, y, x0 = W.config3_problem(np.float64)
    loss = W.config3_loss(P, y)
    x = x0.clone().requires_grad_(True)
    L = loss(x)
    L.backward()
    g = x.grad.detach()
    h = 1e-3
    rounding = 16 * np.finfo(np.float64).eps * abs(float(L.detach())) / h
    rng = np.random.default_rng(31)
    idx = rng.integers(0, 128, size=(64, 3))
    bins = rng.integers(0, 8, size=64)
    for (a, b, c), t in zip(idx, bins):
        xp = x0.clone()
        xm = x0.clone()
        xp[a, b, c, t] += h
        xm[a, b, c, t] -= h
        fd = (float(loss(xp)) - float(loss(xm))) / (2 * h)
        gv = float(g[a, b, c, t])
        assert abs(fd - gv) <= 1e-7 * abs(gv) + rounding, (a, b, c, t, fd, gv, rounding)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_c3_data_gradient_against_oracle_backprojector(oracle, dtype, tol):
    """The data-term gradient 2 P^T (P x - y) of the full-size workload against
    the oracle's time-binned backprojector, voxels with r < 1.05 zero."""
    from tomograpy_amd.autograd import project
    print("C3 oracle", np.dtype(dtype).name, flush=True)
    P, y, x0 = W.config3_problem(dtype)
    ch, hs, tindex, voxmask, _ = W.config3_geometry(dtype)
    x = x0.clone().requires_grad_(True)
    ((project(x, P) - y) ** 2).sum().backward()
    r = (project(x0, P) - y).cpu().numpy()
    gref = np.zeros(tuple(P.shapein), dtype=dtype)
    oracle.siddon(oracle.BPJT, view_records(hs), map_record(ch), gref,
                  np.ascontiguousarray(2 * r), tindex=tindex)
    gref[voxmask] = 0
    assert rel_l2(x.grad.cpu().numpy(), gref) <= tol


def test_c4_batched_jobs_match_one_job_calls():
    """64 uneven jobs in the persistent launch: forward bit-exact to each job
    alone, backward equal to it up to one-ulp ties (float64 accumulation)."""
    print("C4 batch of 64 jobs", flush=True)
    ops, xs = W.config4_problem()
    single = [P.matvec(x).clone() for P, x in zip(ops, xs)]
    outs = [torch.zeros_like(s) for s in single]
    B.run(ops, xs, outs, backward=False, zero_out=True)
    torch.cuda.synchronize()
    for k, (s, o) in enumerate(zip(single, outs)):
        assert torch.equal(s, o), k
    ys = [torch.rand_like(s) for s in single]
    sb = [P.rmatvec(y).clone() for P, y in zip(ops, ys)]
    ob = [torch.zeros_like(b) for b in sb]
    B.run(ops, ys, ob, backward=True)
    torch.cuda.synchronize()
    for k, (s, o) in enumerate(zip(sb, ob)):
        a, b = o.cpu().numpy(), s.cpu().numpy()
        d = a != b
        assert d.mean() <= 1e-3, k
        assert np.all(np.abs(a - b)[d] <= np.spacing(np.maximum(np.abs(a), np.abs(b)))[d]), k
