"""GPU checks of the batched persistent launch, the autotuner and the
continuous-batching server: every job/request gives the result it gives alone."""
import numpy as np
import pytest
import torch

from tomograpy_amd import batch as B
from tomograpy_amd import models as Mo
from tomograpy_amd import serve as SV
from tomograpy_amd import workloads as W
from tomograpy_amd.lo_wrapper import Siddon

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def small_jobs(n_jobs=10, seed=0):
    rng = np.random.default_rng(seed)
    ops, xs = [], []
    for k in range(n_jobs):
        n = int(rng.integers(12, 41))
        nv = int(rng.integers(3, 9))
        npx = int(rng.choice([16, 24, 48]))
        d = float(rng.uniform(150, 250))
        ch = W.cube_header(n)
        hs = W.view_headers(npx, rng.uniform(0, 2 * np.pi, nv), np.deg2rad(rng.uniform(-7, 7, nv)), d=d)
        if k % 2:
            P = Mo.thomson_lo(hs, ch, dtype=np.float32)
        else:
            P = Siddon(hs, ch, dtype=np.float32, obstacle="sun" if k % 4 == 0 else None)
        ops.append(P)
        xs.append(torch.from_numpy(W.blob_phantom(ch, 4, rng)).cuda())
    return ops, xs


@pytest.mark.parametrize("tile_j,schedule", [(16, 0), (8, 1), (32, 1), (64, 0)])
def test_batched_forward_bit_exact_and_backward_close(tile_j, schedule):
    print("batch", tile_j, schedule, flush=True)
    ops, xs = small_jobs()
    singles = [P.matvec(x).clone() for P, x in zip(ops, xs)]
    outs = [torch.zeros_like(s) for s in singles]
    B.run(ops, xs, outs, backward=False, plan=B.make_plan(tile_j=tile_j, schedule=schedule))
    torch.cuda.synchronize()
    for s, o in zip(singles, outs):
        assert torch.equal(s, o)
    # backward: images -> cubes, against the one-job calls
    ys = [torch.rand_like(s) for s in singles]
    sb = [P.rmatvec(y).clone() for P, y in zip(ops, ys)]
    ob = [torch.zeros_like(b) for b in sb]
    B.run(ops, ys, ob, backward=True, plan=B.make_plan(tile_j=tile_j, schedule=schedule))
    torch.cuda.synchronize()
    for P, s, o in zip(ops, sb, ob):
        assert rel_l2(o.cpu(), s.cpu()) <= 1e-5
        if P.acc64:  # float64 sums: the batched result is the one-job result up to ulp ties
            a, b = o.cpu().numpy(), s.cpu().numpy()
            d = a != b
            assert d.mean() <= 1e-3
            assert np.all(np.abs(a - b)[d] <= np.spacing(np.maximum(np.abs(a), np.abs(b)))[d])


def test_zero_out_matches_zeroed_accumulate():
    ops, xs = small_jobs(4, seed=3)
    ref = [P.matvec(x).clone() for P, x in zip(ops, xs)]
    outs = [torch.full_like(r, 7.0) for r in ref]  # garbage that zero_out must overwrite
    B.run(ops, xs, outs, backward=False, zero_out=True)
    for r, o in zip(ref, outs):
        assert torch.equal(r, o)


def test_autotuner_picks_a_plan_and_keeps_results():
    ops, xs = small_jobs(6, seed=1)
    tuner = B.Autotuner(tiles=(8, 16, 32), schedules=(0, 1), repeats=1)
    outs = [torch.zeros((P.n1, P.n2, P.nv), dtype=torch.float32, device="cuda") for P in ops]
    plan = tuner.plan_for(ops, xs, outs)
    assert plan.tile_j in (8, 16, 32) and plan.schedule in (0, 1)
    assert len(tuner.trials[B.signature(ops, False)]) == 6
    assert all(float(o.abs().sum()) == 0.0 for o in outs)  # tuning restores the outputs
    B.run_tuned(tuner, ops, xs, outs)
    for P, x, o in zip(ops, xs, outs):
        assert torch.equal(P.matvec(x), o)


def _requests(n, seed, rate=None):
    rng = np.random.default_rng(seed)
    mix = W.config5_mix(n, rng, n_range=(16, 40), v_range=(4, 12), pix=(16, 32, 48), max_bins=3)
    return SV.make_requests(mix, rng, rate=rate, tol=1e-4, maxiter=300)


@pytest.mark.parametrize("speculative", [False, True, "auto"])
def test_served_results_match_plain(speculative):
    """Each reconstruction is the same alone (plain) and in the mix (served),
    with residual checks read at once or one step late: within 1e-4 relative
    L2 (identical in fact, with float64 backprojection sums), same iteration
    counts, both converged to the tolerance."""
    print("served vs plain, speculative", speculative, flush=True)
    reqs = _requests(10, seed=2)
    plain = _requests(10, seed=2)
    server = SV.Server(budget_bytes=1 << 30, max_active=8, check_every=5, speculative=speculative)
    s_sum = server.run(reqs)
    p_sum = SV.serve_plain(plain, check_every=5)
    assert s_sum["requests"] == p_sum["requests"] == 10
    for a, b in zip(reqs, plain):
        assert a.residual <= 1e-4 and b.residual <= 1e-4
        assert a.iterations == b.iterations, (a.id, a.iterations, b.iterations)
        assert rel_l2(a.x, b.x) <= 1e-4, (a.id, rel_l2(a.x, b.x))
    assert s_sum["p99_s"] >= s_sum["p50_s"] > 0


def test_admission_rejects_oversize_request():
    reqs = _requests(1, seed=4)
    server = SV.Server(budget_bytes=1024, max_active=2)
    with pytest.raises(ValueError, match="budget"):
        server.run(reqs)


def test_plan_batch_mixed_jobs_match_one_job_calls():
    """One plan launch over jobs of different kinds (time bins, the Sun, Thomson
    weights, float64 accumulation on and off) equals each job's own
    backprojection: to the ulp with float64 sums, to 1e-5 without."""
    from tomograpy_amd.lo_wrapper import Siddon4d
    print("plan batch mixed", flush=True)
    rng = np.random.default_rng(11)
    ops = []
    for k in range(5):
        n = int(rng.integers(20, 48))
        nv = int(rng.integers(4, 10))
        npx = int(rng.choice([32, 48]))
        ch = W.cube_header(n)
        hs = W.view_headers(npx, rng.uniform(0, 2 * np.pi, nv), np.deg2rad(rng.uniform(-7, 7, nv)),
                            d=float(rng.uniform(150, 250)))
        if k == 0:
            P = Siddon4d(hs, ch, tindex=np.arange(nv) % 3, dtype=np.float32, bpj_mode="plan")
        elif k == 1:
            P = Mo.thomson_lo(hs, ch, dtype=np.float32, bpj_mode="plan")
        elif k == 2:
            P = Siddon(hs, ch, dtype=np.float32, obstacle="sun", bpj_mode="plan", acc64=False)
        else:
            P = Siddon(hs, ch, dtype=np.float32, bpj_mode="plan", brick=(8, 16, 16))
        ops.append(P)
    ys = [torch.rand((P.nv, P.n1, P.n2), device="cuda").permute(1, 2, 0) for P in ops]
    single = [P.rmatvec(y).clone() for P, y in zip(ops, ys)]
    outs = [torch.zeros_like(s) for s in single]
    B.run(ops, ys, outs, backward=True)
    torch.cuda.synchronize()
    for P, s, o in zip(ops, single, outs):
        a, b = o.cpu().numpy(), s.cpu().numpy()
        if P.acc64:
            d = a != b
            assert d.mean() <= 1e-3
            assert np.all(np.abs(a - b)[d] <= np.spacing(np.maximum(np.abs(a), np.abs(b)))[d])
        else:
            assert rel_l2(a, b) <= 1e-5


@pytest.mark.parametrize("mode", ["plan", "atomic"])
def test_batched_backward_repeated_operator(mode):
    """One operator listed twice in a backward batch: each job sums in its own
    float64 accumulator, so B.run([P, P], [y1, y2]) equals the two rmatvec
    calls (to one-ulp ties), for brick-plan and atomic jobs alike."""
    print("repeated operator", mode, flush=True)
    ch = W.cube_header(40)
    hs = W.view_headers(48, np.linspace(0.1, 5.0, 9), d=200.0)
    P = Siddon(hs, ch, dtype=np.float32, bpj_mode=mode)
    assert P.acc64
    g = torch.Generator().manual_seed(9)
    y1 = torch.rand((48, 48, 9), generator=g).cuda()
    y2 = torch.rand((48, 48, 9), generator=g).cuda()
    r1 = P.rmatvec(y1).clone()
    r2 = P.rmatvec(y2).clone()
    q1, q2 = torch.zeros_like(r1), torch.zeros_like(r2)
    B.run([P, P], [y1, y2], [q1, q2], backward=True)
    torch.cuda.synchronize()
    for s, o in ((r1, q1), (r2, q2)):
        a, b = o.cpu().numpy(), s.cpu().numpy()
        d = a != b
        assert d.mean() <= 1e-3
        assert np.all(np.abs(a - b)[d] <= np.spacing(np.maximum(np.abs(a), np.abs(b)))[d])
    with pytest.raises(ValueError, match="twice"):
        B.run([P, P], [y1, y2], [q1, q1], backward=True)


def test_served_poisson_arrivals_match_plain():
    """Poisson arrivals at a rate above what one request at a time can serve,
    so that requests arrive while others are on the GPU and join the running
    set: every result and iteration count equals the request's plain solve."""
    print("served vs plain, Poisson arrivals", flush=True)
    reqs = _requests(16, seed=7, rate=400.0)
    plain = _requests(16, seed=7, rate=400.0)
    assert reqs[-1].arrival > 0
    server = SV.Server(budget_bytes=1 << 30, max_active=6, check_every=5)
    s_sum = server.run(reqs)
    p_sum = SV.serve_plain(plain, check_every=5, order="shortest")
    assert s_sum["requests"] == p_sum["requests"] == 16
    overlap = sum(1 for a in reqs for b in reqs
                  if a is not b and b.arrival > 0 and a.t_admit < b.t_admit < a.t_done)
    print("served", s_sum, "plain", p_sum, "admissions during another solve", overlap,
          flush=True)
    assert overlap > 0
    for a, b in zip(reqs, plain):
        assert a.status == b.status == "done"
        assert a.residual <= 1e-4 and b.residual <= 1e-4
        assert a.iterations == b.iterations, (a.id, a.iterations, b.iterations)
        assert rel_l2(a.x, b.x) <= 1e-4, (a.id, rel_l2(a.x, b.x))


def test_server_cancel_and_timeout():
    """A request cancelled while on the GPU leaves the active set and frees its
    slot; one whose timeout passes while it waits is never admitted; the others
    complete with their plain results."""
    print("cancel and timeout", flush=True)
    reqs = _requests(6, seed=8)
    plain = _requests(6, seed=8)
    reqs[1].timeout = 0.0  # already past when the server first looks at it
    victim = reqs[0]

    def on_step(step, rs):
        if step == 2:
            victim.cancel()

    # one slot, arrival order: request 0 runs first and is cancelled after two
    # of its iterations (a check every 5 cannot have ended it)
    server = SV.Server(budget_bytes=1 << 30, max_active=1, check_every=5, order="arrival")
    out = server.run(reqs, on_step=on_step)
    assert reqs[1].status == "timed out" and reqs[1].t_admit is None
    assert victim.status == "cancelled" and victim.iterations == 2 and victim.x is None
    assert out["dropped"] == 2 and out["requests"] == 4
    assert sorted(server.free_rows) == [0]  # the cancelled request's slot came back
    SV.serve_plain(plain, check_every=5)
    for a, b in zip(reqs, plain):
        if a.status == "done":
            assert a.iterations == b.iterations
            assert rel_l2(a.x, b.x) <= 1e-4
