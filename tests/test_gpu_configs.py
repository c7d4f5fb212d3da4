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


def test_c4_batched_jobs_match_oracle(oracle):
    """Four of the 64 batched jobs (the three with the fewest rays x cube edge and
    the cheapest Thomson job) against the oracle: batched forward bit-exact,
    batched backward within 1e-5 of the oracle's single-thread backprojection.
    A Thomson job is O . P_sun . M (models.py:223-235): the oracle projects
    M x with the Sun obstacle and scales by O, and backprojects O y and scales
    by M, each product rounded to float32 as the fused kernels round it."""
    print("C4 batch vs oracle", flush=True)
    jobs = W.config4_jobs()
    ops, xs = W.config4_problem()
    cost = [j["npix"] ** 2 * j["nv"] * j["n"] for j in jobs]
    order = sorted(range(len(jobs)), key=lambda k: cost[k])
    plain = [k for k in order if not jobs[k]["thomson"]][:3]
    thom = [k for k in order if jobs[k]["thomson"]][:1]
    picked = plain + thom
    outs = [torch.zeros((P.n1, P.n2, P.nv), dtype=torch.float32, device="cuda") for P in ops]
    B.run(ops, xs, outs, backward=False, zero_out=True)
    rng = np.random.default_rng(41)
    ys = [torch.from_numpy(rng.uniform(0, 1, (P.n1, P.n2, P.nv)).astype(np.float32)).cuda()
          for P in ops]
    ob = [torch.zeros(P.cube_shape, dtype=torch.float32, device="cuda") for P in ops]
    B.run(ops, ys, ob, backward=True)
    torch.cuda.synchronize()
    for k in picked:
        job, P = jobs[k], ops[k]
        print("job", k, "n", job["n"], "views", job["nv"], "pixels", job["npix"],
              "thomson", job["thomson"], flush=True)
        hs, ch = job["headers"], job["cube_header"]
        x = xs[k].cpu().numpy()
        y = ys[k].cpu().numpy()
        if job["thomson"]:
            O = P.pix_weight.cpu().numpy()
            M = P.vox_weight.cpu().numpy()
            ref = np.zeros((P.n1, P.n2, P.nv), np.float32)
            oracle.siddon(oracle.PJ, view_records(hs), map_record(ch),
                          np.ascontiguousarray(M * x), ref, sun=True)
            ref = O * ref
            bref = np.zeros(P.cube_shape, np.float32)
            oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref,
                          np.ascontiguousarray(O * y), sun=True, nthread=1)
            bref = M * bref
        else:
            ref = np.zeros((P.n1, P.n2, P.nv), np.float32)
            oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), x, ref)
            bref = np.zeros(P.cube_shape, np.float32)
            oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref, y, nthread=1)
        assert np.array_equal(outs[k].cpu().numpy(), ref), k
        err = rel_l2(ob[k].cpu().numpy(), bref)
        print("job", k, "backward rel L2", err, flush=True)
        assert err <= 1e-5, (k, err)


def _c2_subset(dtype, views):
    ch, hs, rng = W.config2_geometry(dtype)
    return ch, [hs[k] for k in views], rng


# every 8th view of the 128: includes views 32, 64 and 96, whose observers sit on
# a coordinate axis (LON = pi/2, pi, 3 pi/2)
C2_VIEWS = list(range(0, 128, 8))


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_c2_operator_pair_on_16_views(oracle, dtype, tol):
    """The C2 shapes (256^3 cube, 512^2 images) on 16 of its 128 views through
    siddon_lo with its defaults -- precounted forward, brick-plan backward with
    the default 32 x 16 x 32 bricks (float64 sums in float32): forward bit-exact
    against the oracle, backward (rmatvec, lo_wrapper.py:29-34) within tol of the
    oracle's single-thread backprojection (C:923-928), and the adjoint identity
    <P x, y> = <x, P^T y> within tol."""
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("C2 pair on 16 views", np.dtype(dtype).name, flush=True)
    ch, hs, rng = _c2_subset(dtype, C2_VIEWS)
    P = siddon_lo(hs, ch, dtype=dtype)
    assert P.use_plan()
    plan = P.build_plan()
    assert tuple(plan.brick) == (32, 16, 32)
    x = W.blob_phantom(ch, 40, rng, dtype=dtype)
    y = P.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = np.zeros((512, 512, len(hs)), dtype)
    oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), x, ref)
    assert np.array_equal(y, ref)
    yy = (ref + 0.01 * float(np.abs(ref).max())
          * rng.standard_normal(ref.shape)).astype(dtype)
    b = P.rmatvec(torch.from_numpy(yy).cuda()).cpu().numpy()
    bref = np.zeros((256, 256, 256), dtype)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref, yy, nthread=1)
    err = rel_l2(b, bref)
    print("C2 backward rel L2", err, "records", plan.records, "chunks", plan.chunks, flush=True)
    assert err <= tol
    # the adjoint identity <P x, y> = <x, P^T y> on the same pair (sums in float64)
    lhs = float(np.dot(y.astype(np.float64).ravel(), yy.astype(np.float64).ravel()))
    rhs = float(np.dot(x.astype(np.float64).ravel(), b.astype(np.float64).ravel()))
    print("C2 adjoint", lhs, rhs, flush=True)
    assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs))


def test_c2_cg_fp64_matches_scipy(oracle):
    """Two fp64 CG iterations of the C2 normal equations
    (P^T P + 1e-2 (I + sum_k D_k^T D_k)) x = P^T y on 16 of its views (full
    256^3 / 512^2 shapes) against scipy.sparse.linalg.cg on the oracle operator."""
    from scipy.sparse.linalg import LinearOperator, cg as scipy_cg
    from tomograpy_amd import cg as CG
    from tomograpy_amd.lo_wrapper import siddon_lo
    from tests.test_models import regularizer_reference
    print("C2 CG fp64 vs scipy", flush=True)
    ch, hs, rng = _c2_subset(np.float64, C2_VIEWS)
    P = siddon_lo(hs, ch, dtype=np.float64)
    x = W.blob_phantom(ch, 40, rng, dtype=np.float64)
    recs, mrec = view_records(hs), map_record(ch)
    y = np.zeros((512, 512, len(hs)))
    oracle.siddon(oracle.PJ, recs, mrec, x, y)
    y += 0.01 * float(np.abs(y).max()) * rng.standard_normal(y.shape)
    lam = 1e-2
    shape = (256, 256, 256)

    def A(v):
        v = np.ascontiguousarray(v.reshape(shape))
        pv = np.zeros(y.shape)
        oracle.siddon(oracle.PJ, recs, mrec, v, pv)
        out = np.zeros(shape)
        oracle.siddon(oracle.BPJ, recs, mrec, out, pv)
        return (out + regularizer_reference(v, (0, 1, 2), lam, True)).ravel()

    b = np.zeros(shape)
    oracle.siddon(oracle.BPJ, recs, mrec, b, y)
    op = LinearOperator((b.size, b.size), matvec=A, dtype=np.float64)
    xs, _ = scipy_cg(op, b.ravel(), x0=np.zeros(b.size), maxiter=2, rtol=1e-30, atol=0)
    xg, info = CG.solve(P, torch.from_numpy(y).cuda(), lam=lam, identity=True, maxiter=2)
    err = rel_l2(xg.cpu().numpy().ravel(), xs)
    print("C2 CG 2 iterations rel err", err, flush=True)
    assert info["iterations"] == 2
    assert err < 1e-9
