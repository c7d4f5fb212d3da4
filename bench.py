"""Benchmark: Siddon projector + backprojector pair on one MI355X.

Workload (configuration C2): 256^3 fp32 cube over [-1.5, 1.5]^3, 128 conic
views of 512 x 512 pixels from 214 R_sun, longitudes 2 pi k / 128, latitudes
U(-7, 7) deg, a seeded 40-blob phantom, images = P(phantom) + 1% noise.  A
"step" is one application of the normal operator's projector pair, as a CG
iteration makes it: zero the image stack, project the cube into it, zero the
cube buffer, backproject the images (siddon_lo matvec + rmatvec,
lo_wrapper.py:23-34).  Inputs are resident in HBM before timing starts.

value = Siddon steps per second (voxel segments traversed by the pair / time).
roofline: the dominant kernel against the HBM roofline: its DRAM bytes per
launch from the request-size-resolved TCC read/write counters of the committed
summary profiles/siddon_c2_pair_counters.json (when this run is the workload it
measured) over its average launch time (HIP events on the launch stream),
against 8 TB/s.  Beside it: its wave64 VALU instructions per launch
(SQ_INSTS_VALU) over the same time against 1228.8 G instructions/s (256 CUs x
2.4 GHz x 2: four SIMDs of 32 lanes per CU for 32-bit operations, the
157.3 TFLOP/s FP32 vector peak over 128 flops per wave64 FMA), the L2 hit rate,
and the per-step access bytes of the algorithmic model (access_bytes):
  forward  access = S*b + R_t*b + V*18*b   (voxel gather per step, pixel write per ray)
  backward access = 2*S*b + R_t*b + V*18*b (voxel read-modify-write per step, pixel read per ray)
which the brick backprojection serves from LDS and the forward's gathers from
L2, so that access rate exceeds the HBM peak and is not a fraction of it.
adjoint: <P x, y> against <x, P^T y> on the timed pair's own outputs (sums in
float64), against a tolerance of 1e-5 in fp32 (1e-12 in fp64).
cpu_baseline: the oracle (C/OpenMP restatement of the reference) on a bounded
sample of the same workload's views, on the host cores; "parity" compares the
GPU pair with the oracle's output on that sample.
serving: the mixed-size reconstruction stream (see serving_leg), served by
the continuous-batching server and solved one request at a time: p50/p99
latency, throughput, and the largest relative L2 between the two results.
cg: 50 CG iterations of the regularised normal equations on the workload, in
fp32 and in fp64 (with the fp64 pair time).  differentiation, batch: the
autograd workload (grad_leg) and the 64 batched jobs (batch_leg).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions per second: the 157.3 TFLOP/s FP32 vector peak over
# 128 flops per wave64 FMA instruction = 256 CUs x 2.4 GHz x 2 per CU per clock
# (4 SIMDs, each 32 lanes wide for 32-bit operations)
VALU_PEAK_GIPS = 256 * 2.4 * 2
COUNTERS = os.path.join(ROOT, "profiles", "siddon_c2_pair_counters.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", default="float32", choices=["float32", "float64"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--views", type=int, default=128)
    ap.add_argument("--pix", type=int, default=512)
    ap.add_argument("--tile-j", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cg-iters", type=int, default=50,
                    help="CG iterations timed on the workload in fp32 and fp64 (0 = skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="reconstruction stream served vs plain (0 = skip)")
    ap.add_argument("--serve-max-active", type=int, default=32)
    ap.add_argument("--serve-step-work", type=float, default=0.0,
                    help="work per engine step of the server (0 = its default)")
    ap.add_argument("--serve-rate", type=float, default=12.0,
                    help="Poisson arrival rate of the mixed stream in requests/s (0 = all at "
                         "once); the default is above what one request at a time serves")
    ap.add_argument("--serve-small-requests", type=int, default=64,
                    help="requests of the small-request stream (0 = skip)")
    ap.add_argument("--serve-small-rate", type=float, default=1000.0,
                    help="Poisson arrival rate of the small-request stream (0 = all at once)")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-jobs leg")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip the leg that times one GPU's share of a view-sharded run")
    ap.add_argument("--no-det", action="store_true",
                    help="skip the deterministic (fixed summation order) backward leg")
    return ap.parse_args()


def _event_ms(fn, reps):
    import torch
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def cg_run(P, y, iters, keep=None):
    """Times `iters` CG iterations on (P^T P + 1e-2 (I + sum D_k^T D_k)) x = P^T y;
    with a dict `keep`, stores the solution there (host, float64)."""
    import torch
    from tomograpy_amd import cg as CG
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    x, info = CG.solve(P, y, lam=1e-2, identity=True, maxiter=iters)
    e1.record(stream)
    torch.cuda.synchronize()
    if keep is not None:
        keep["x"] = x.detach().double().cpu()
    return {"iterations": info["iterations"], "ms_per_iteration": e0.elapsed_time(e1) / iters,
            "relative_residual": info["residual"]}


def _rel(a, b):
    a = a.double().flatten()
    b = b.double().flatten()
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def axis_views(hs):
    """Views whose observer sits on a coordinate plane up to rounding: a
    coordinate of M tiny but not zero (LON = pi/2, pi, 3 pi/2 evaluated in
    floating point).  There the reference's float32 ray set-up gives a few rays
    steps of about +-212 (C:723-735 with C:804-809); float64 does not."""
    out = []
    for t, h in enumerate(hs):
        m = [float(h["M%d" % (k + 1)]) for k in range(3)]
        d = float(np.sqrt(sum(v * v for v in m)))
        if any(0.0 < abs(v) < 1e-9 * d for v in m):
            out.append(t)
    return out


def view_mask(hs, n1, n2, views):
    """Pixel mask (True = skip) covering the given views."""
    m = np.zeros((n1, n2, len(hs)), dtype=bool)
    m[:, :, views] = True
    return m


def fp64_leg(args, ref32=None):
    """The same workload in float64: the pair and the CG iterations.  With the
    float32 run's inputs and results (ref32: x, y, forward, backward, CG
    solution; host tensors), also the float32 error against float64 on the
    same inputs."""
    import torch
    from tomograpy_amd import workloads as W
    from tomograpy_amd.lo_wrapper import siddon_lo
    ch, hs, rng = W.config2_geometry(np.float64, n=args.n, n_views=args.views, n_pix=args.pix)
    phantom = W.blob_phantom(ch, 40, rng, dtype=np.float64)
    P = siddon_lo(hs, ch, dtype=np.float64)
    x = torch.from_numpy(phantom).cuda()
    y = P.matvec(x).clone()
    y += 0.01 * float(y.max()) * torch.from_numpy(rng.normal(size=tuple(y.shape))).cuda()
    if P.use_plan():
        P.build_plan()

    def pair():
        P.project_into(x, P.xout)
        P.backproject_into(y, P.xin)

    out = {"ms_pair": _event_ms(pair, 5), "siddon_steps_per_projection": P.count_steps()}
    out.update(cg_run(P, y, args.cg_iters))
    if ref32 is None:
        return out, None
    # the float32 run's inputs, in float64
    x32 = ref32["x"].cuda()
    y32 = ref32["y"].cuda()
    ok = [t for t in range(len(hs)) if t not in set(ref32["axis_views"])]
    f64 = P.matvec(x32).cpu()
    err = {"axis_views": ref32["axis_views"],
           "forward_rel_l2": _rel(f64, ref32["forward"]),
           "forward_rel_l2_other_views": _rel(f64[:, :, ok], ref32["forward"][:, :, ok]),
           "backward_rel_l2": _rel(P.rmatvec(y32).cpu(), ref32["backward"]),
           "backward_rel_l2_other_views": _rel(P.rmatvec(ref32["y_other"].cuda()).cpu(),
                                               ref32["backward_other"])}
    keep = {}
    cg_run(P, y32, args.cg_iters, keep)
    err["cg_solution_rel_l2"] = _rel(ref32["cg_x"], keep["x"])
    del P
    torch.cuda.empty_cache()
    Pm = siddon_lo(hs, ch, dtype=np.float64,
                   mask=view_mask(hs, args.pix, args.pix, ref32["axis_views"]))
    keep = {}
    cg_run(Pm, y32, args.cg_iters, keep)
    err["cg_solution_rel_l2_other_views"] = _rel(ref32["cg_x_other"], keep["x"])
    return out, err


def grad_leg():
    """Differentiation workload: 128^3 x 8 time bins, 64 views of 256^2, voxel
    mask r < 1.05; one evaluation of ||P x - y||^2 + 1e-2 ||grad x||^2 +
    1e-2 ||dx/dt||^2 and its gradient (forward through the projector, backward
    through the hand-written backprojector).  Also the float32 gradient's error
    against the float64 one at the same point and data."""
    import torch
    from tomograpy_amd import workloads as W
    P, y, x0 = W.config3_problem(np.float32)
    loss = W.config3_loss(P, y)
    x = x0.clone().requires_grad_(True)

    def step():
        x.grad = None
        loss(x).backward()

    ms = _event_ms(step, 5)
    steps = P.count_steps()
    _, hs, _, _, _ = W.config3_geometry(np.float32)
    keepv = torch.ones(len(hs), dtype=torch.float64)
    keepv[axis_views(hs)] = 0.0

    def grads(P, y, x0, dt):
        """The loss gradient at x0, and the same with the data term of the
        axis views left out."""
        out = []
        for w in (None, keepv.to(P.dev, dt)):
            xg = x0.clone().requires_grad_(True)
            if w is None:
                L = W.config3_loss(P, y)(xg)
            else:
                from tomograpy_amd.autograd import project, smoothness
                r = (project(xg, P) - y) * w
                L = (r * r).sum() + smoothness(xg, (1e-2, 1e-2, 1e-2, 1e-2))
            L.backward()
            out.append(xg.grad.detach().double().cpu())
        return out

    g32, g32o = grads(P, y, x0, torch.float32)
    y32, x32 = y.double(), x0.double()
    del P, loss, x
    torch.cuda.empty_cache()
    P64, _, _ = W.config3_problem(np.float64)
    g64, g64o = grads(P64, y32, x32, torch.float64)
    return {"workload": "128^3 x 8 bins, 64 views of 256^2, voxel mask r<1.05, fp32",
            "ms_per_loss_and_gradient": ms, "siddon_steps_per_projection": steps,
            "gradient_rel_l2_vs_fp64": _rel(g32, g64),
            "gradient_rel_l2_vs_fp64_other_views": _rel(g32o, g64o),
            "axis_views": axis_views(hs)}


def batch_leg():
    """Batched-jobs workload: 64 uneven projector + backprojector jobs (cubes
    32^3-192^3, 8-96 views of 64^2-512^2, D 150-250, half Thomson-weighted),
    in the persistent launch (autotuned tile and schedule) against the same
    jobs one launch at a time."""
    import torch
    from tomograpy_amd import batch as B
    from tomograpy_amd import workloads as W
    ops, xs = W.config4_problem()
    outs = [torch.zeros((P.nv, P.n1, P.n2), dtype=P.tdtype, device=P.dev).permute(1, 2, 0)
            for P in ops]
    cubes = [torch.zeros(tuple(P.shapein), dtype=P.tdtype, device=P.dev) for P in ops]
    for P in ops:  # geometry set-up (step counts, plans) outside the timing
        P.ray_steps()
        if P.use_plan():
            P.build_plan()
    # each (tile width, schedule) point timed 3 times after a warm-up; the
    # fastest of its 3 decides
    tuner = B.Autotuner(repeats=3)

    def batched():
        B.run_tuned(tuner, ops, xs, outs)
        B.run(ops, outs, cubes, backward=True)

    def one_by_one():
        for P, x, o, c in zip(ops, xs, outs, cubes):
            P.project_into(x, o)
            P.backproject_into(o, c)

    batched()  # tunes the forward launch
    torch.cuda.synchronize()
    trials = list(tuner.trials.values())[0]
    # marker ranges (roctx) that a trace of this run can cut its phases by
    torch.cuda.nvtx.range_push("batch batched")
    ms_b = _event_ms(batched, 3)
    torch.cuda.nvtx.range_pop()
    torch.cuda.nvtx.range_push("batch one-by-one")
    ms_1 = _event_ms(one_by_one, 3)
    torch.cuda.nvtx.range_pop()
    steps = sum(P.count_steps() for P in ops)
    # one batched pair from zero against the same jobs run one by one in float64
    for o, c in zip(outs, cubes):
        o.zero_()
        c.zero_()
    batched()
    ops64, xs64 = W.config4_problem(np.float64)
    err_f, err_b = [], []
    for P64, x64, o, c in zip(ops64, xs64, outs, cubes):
        o64 = P64.matvec(x64)
        err_f.append(_rel(o, o64))
        err_b.append(_rel(c, P64.rmatvec(o64)))
    del ops64, xs64
    return {"workload": "64 jobs, cubes 32^3-192^3, 8-96 views of 64^2-512^2, half Thomson, fp32",
            "jobs": len(ops), "ms_batched_pair": ms_b, "ms_one_by_one_pair": ms_1,
            "speedup": ms_1 / ms_b, "siddon_steps_per_pair": 2 * steps,
            "forward_plan": {"tile_j": trials[0][1], "schedule": trials[0][2],
                             "repeats": tuner.repeats,
                             "trials_ms": [[round(t, 4), tj, sch] for t, tj, sch in trials]},
            "steps_per_s_batched": 2 * steps / (ms_b * 1e-3),
            "error_vs_fp64": {"forward_rel_l2_max": max(err_f),
                              "forward_rel_l2_median": float(np.median(err_f)),
                              "backward_rel_l2_max": max(err_b),
                              "backward_rel_l2_median": float(np.median(err_b))}}


def serving_leg(args, small=False):
    """The mixed-size reconstruction stream (32^3-256^3 cubes, 8-128 views of
    64^2-512^2, pixel mask on/off, 1-8 time bins, CG to a relative residual of
    1e-4), solved once by the continuous-batching server and once one request
    at a time in the server's admission order (shortest work first among the
    requests that have arrived); same requests, same seeds, Poisson arrivals at
    --serve-rate requests/s (0: all at once).  small=True: the small end of
    the mix (32^3-64^3 cubes, 8-24 views of 64^2-128^2, 1-4 bins) with 32
    slots, where a request alone does not fill the GPU, --serve-small-requests
    of them at --serve-small-rate."""
    import torch
    from tomograpy_amd import serve as SV
    from tomograpy_amd import workloads as W

    n_req = args.serve_small_requests if small else args.serve_requests
    slots = 32 if small else args.serve_max_active

    rate = args.serve_small_rate if small else args.serve_rate

    def requests(dtype=np.float32):
        rng = np.random.default_rng(W.SEED0 + (6 if small else 5))
        if small:
            mix = W.config5_mix(n_req, rng, n_range=(32, 64), v_range=(8, 24), pix=(64, 128),
                                max_bins=4)
        else:
            mix = W.config5_mix(n_req, rng)
        out, t = [], 0.0
        for k in range(0, len(mix), 100):
            part = SV.make_requests(mix[k:k + 100], rng, rate=rate or None, tol=1e-4,
                                    maxiter=200, dtype=dtype, start=t, first_id=k)
            out += part
            t = part[-1].arrival
            if len(mix) > 100:
                print("serving: %d of %d requests made" % (len(out), len(mix)), file=sys.stderr,
                      flush=True)
        return out

    reqs = requests()
    torch.cuda.synchronize()
    print("serving: %d requests made" % len(reqs), file=sys.stderr, flush=True)
    tag = "serving_small" if small else "serving"
    # the plain run's solutions: a device fingerprint of each (two sums over its
    # bit patterns, taken before the download in both paths), and the host
    # solutions of a sample (every request of a short stream, 1 in 50 of a
    # long one) for the error; nothing is copied or hashed inside the runs
    keep_every = 1 if len(reqs) <= 64 else 50
    digest, plain_x, plain_its, plain_res = {}, {}, {}, {}
    last = [time.perf_counter()]
    for r in reqs:
        r.fingerprint = True

    def progress(what, r):
        if time.perf_counter() - last[0] > 20.0:
            last[0] = time.perf_counter()
            print("serving: %s, request %d done at %.1f s" % (what, r.id, r.t_done),
                  file=sys.stderr, flush=True)

    def plain_done(r):
        digest[r.id] = r.fingerprint
        plain_its[r.id], plain_res[r.id] = r.iterations, r.residual
        if r.id % keep_every == 0:
            plain_x[r.id] = r.x  # the pinned buffer itself, kept
        r.x = None
        progress("plain", r)

    served_fp, served_x = {}, {}

    def served_done(r):
        served_fp[r.id] = r.fingerprint
        if r.id in plain_x:
            served_x[r.id] = r.x
        r.x = None
        progress("served", r)

    # marker ranges (roctx) that a trace of this run can cut its phases by
    torch.cuda.nvtx.range_push(tag + " plain")
    plain = SV.serve_plain(reqs, order="shortest", on_complete=plain_done)
    torch.cuda.synchronize()
    torch.cuda.nvtx.range_pop()
    print("serving: plain done in %.1f s" % plain["makespan_s"], file=sys.stderr, flush=True)
    converged = all(r.residual <= 1e-4 for r in reqs)
    for r in reqs:
        r.reset()
        r.fingerprint = True
    torch.cuda.nvtx.range_push(tag + " served")
    server = SV.Server(budget_bytes=128 << 30, max_active=slots,
                       step_work=args.serve_step_work or None)
    served = server.run(reqs, on_complete=served_done)
    torch.cuda.synchronize()
    torch.cuda.nvtx.range_pop()
    print("serving: served done in %.1f s" % served["makespan_s"], file=sys.stderr, flush=True)
    converged = converged and all(r.residual <= 1e-4 for r in reqs)
    same = {"bits": sum(int(torch.equal(served_fp[k], digest[k])) for k in digest),
            "iterations": sum(int(r.iterations == plain_its[r.id]) for r in reqs)}
    errs = [float(np.linalg.norm(served_x[k] - plain_x[k]) / max(np.linalg.norm(plain_x[k]), 1e-300))
            for k in plain_x]
    # the served float32 solutions of the 8 lightest sampled requests against
    # float64 solves of the same requests (same images, upcast)
    light = sorted(served_x, key=lambda k: reqs[k].work())[:8]
    light_x = {k: np.array(served_x[k]) for k in light}
    reqs64 = [SV.Request(k, reqs[k].cube_header, reqs[k].data_header,
                         torch.as_tensor(reqs[k].y).double(), lam=reqs[k].lam, tol=reqs[k].tol,
                         maxiter=reqs[k].maxiter, mask=reqs[k].mask, tindex=reqs[k].tindex,
                         obstacle=reqs[k].obstacle, dtype=np.float64) for k in light]
    # release the stream's pinned images and the pinned solution buffers that
    # torch's host allocator keeps cached (tens of GB for a long stream)
    keep_y = {k: None for k in light}
    for r in reqs:
        r.y = keep_y.get(r.id, None)
    plain_x = served_x = None
    if hasattr(torch._C, "_host_emptyCache"):
        torch._C._host_emptyCache()
    SV.serve_plain(reqs64)
    err64 = [float(np.linalg.norm(light_x[k] - r.x) / np.linalg.norm(r.x))
             for k, r in zip(light, reqs64)]
    del reqs64
    out = {"workload": ("%d requests, %s, CG to a relative residual of 1e-4, arrivals %s"
                        % (len(reqs),
                           "32^3-64^3 cubes, 8-24 views of 64^2-128^2, mask on/off, 1-4 bins"
                           if small else
                           "32^3-256^3 cubes, 8-128 views of 64^2-512^2, mask on/off, 1-8 bins",
                           "Poisson at %g/s" % rate if rate else "all at t = 0")),
           "requests": len(reqs), "arrival_rate_rps": rate or None,
           "max_active": slots, "step_work": server.step_work, "tol": 1e-4,
           "plain_order": "shortest work first",
           "served": served, "plain": plain,
           "throughput_ratio": served["throughput_rps"] / plain["throughput_rps"],
           "p50_ratio": plain["p50_s"] / served["p50_s"],
           "p99_ratio": plain["p99_s"] / served["p99_s"],
           "bit_identical_to_plain": same["bits"], "same_iterations": same["iterations"],
           "max_rel_l2_served_vs_plain_sampled": max(errs) if errs else None,
           "sampled": len(errs),
           "error_vs_fp64": {"requests": len(light), "solution_rel_l2_max": max(err64),
                             "solution_rel_l2_median": float(np.median(err64))},
           "all_converged": converged,
           "not_converged_at_maxiter": sum(1 for r in reqs if r.residual > 1e-4)}
    return out


def shard_leg(args, nshards=8, reps=10):
    """The pair on one GPU's share of a view-sharded run: view k on shard k mod
    nshards, every shard at the full cube and image shape, each timed alone."""
    import torch
    from tomograpy_amd import workloads as W
    from tomograpy_amd.lo_wrapper import siddon_lo
    dtype = np.dtype(args.dtype)
    ch, hs, rng = W.config2_geometry(dtype, n=args.n, n_views=args.views, n_pix=args.pix)
    x = torch.from_numpy(rng.uniform(size=(args.n,) * 3).astype(dtype)).cuda()
    ms, steps = [], []
    for s in range(nshards):
        hs_s = hs[s::nshards]
        P = siddon_lo(hs_s, ch, dtype=dtype, tile_j=args.tile_j)
        y = torch.from_numpy(rng.normal(size=(args.pix, args.pix, len(hs_s))).astype(dtype)).cuda()
        if P.use_plan():
            P.build_plan()

        def pair():
            P.xout.zero_()
            P.project_into(x, P.xout)
            P.xin.zero_()
            P.backproject_into(y, P.xin)

        ms.append(_event_ms(pair, reps))
        steps.append(int(P.count_steps()))
        del P, y
        torch.cuda.empty_cache()
    k = int(np.argmax(ms))
    return {"shards": nshards, "views_per_shard": len(hs[::nshards]),
            "assignment": "view k on shard k mod %d" % nshards,
            "ms_pair": [round(m, 4) for m in ms], "ms_pair_max": ms[k],
            "siddon_steps_per_s_slowest_shard": 2 * steps[k] / (ms[k] * 1e-3)}


def det_leg(args, reps=2):
    """The deterministic backward (bit-identical to the reference's single-thread
    order) at the workload's shape, against the operator's default backward."""
    import torch
    from tomograpy_amd import workloads as W
    from tomograpy_amd.lo_wrapper import siddon_lo
    dtype = np.dtype(args.dtype)
    ch, hs, rng = W.config2_geometry(dtype, n=args.n, n_views=args.views, n_pix=args.pix)
    y = torch.from_numpy(rng.normal(size=(args.pix, args.pix, len(hs))).astype(dtype)).cuda()
    Pd = siddon_lo(hs, ch, dtype=dtype, deterministic=True)
    out = {"method": "canonical brick plan" if Pd.det_plan_ok() else "chunked passes"}
    if Pd.det_plan_ok():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Pd.build_det_plan()
        torch.cuda.synchronize()
        out["plan_build_ms"] = (time.perf_counter() - t0) * 1e3
        out["plan_records"] = int(Pd._det_plan.records)
        out["plan_bytes"] = int(sum(b.numel() for b in Pd._det_plan_bufs))
    out["ms_backward"] = _event_ms(lambda: Pd.rmatvec(y), reps)
    out["workspace_bytes"] = int(Pd._ws.numel()) if Pd._ws is not None else 0
    del Pd
    torch.cuda.empty_cache()
    return out


def measured_counters(kernel, args, plan_info):
    """Per-launch counter figures of `kernel` (VALU instructions, HBM bytes, L2
    hits and misses, atomics) from the committed rocprofv3 summary, when this
    run is the workload that summary measured (counters cannot be read from
    inside the timed process); None otherwise."""
    if not os.path.exists(COUNTERS):
        return None
    rec = json.load(open(COUNTERS))
    w = rec["workload"]
    if (w["n"], w["views"], w["pix"], w["dtype"]) != (args.n, args.views, args.pix, args.dtype):
        return None
    if plan_info is not None and list(w.get("brick", [])) != plan_info["brick"]:
        return None
    # one backprojection call launches the brick kernel and the float64
    # accumulator finish, which the bench's events time together
    return rec["kernels"].get(kernel + "_call", rec["kernels"].get(kernel))


def roofline(kernel, t_launch, access_bytes, args, plan_info):
    """The roofline object of the bench line for `kernel` at t_launch seconds
    per launch (see the module docstring): HBM bytes per launch from the
    committed counter summary against 8 TB/s, with VALU issue beside it."""
    c = measured_counters(kernel, args, plan_info)
    out = {"bound": "hbm", "kernel": kernel, "unit": "GB/s", "peak": HBM_PEAK_GBS,
           "achieved": None, "frac": None, "traffic": None,
           "access_bytes": access_bytes, "access_GBs": access_bytes / t_launch / 1e9}
    if c is None:
        out["note"] = "no counter summary for this workload (profiles/ covers the default)"
        return out
    valu = c["valu_instructions"]
    hbm = c["hbm_read_bytes"] + c["hbm_write_bytes"]
    out.update({"achieved": hbm / t_launch / 1e9,
                "frac": hbm / t_launch / 1e9 / HBM_PEAK_GBS,
                "traffic": hbm,
                "read_bytes": c["hbm_read_bytes"], "write_bytes": c["hbm_write_bytes"],
                "valu": {"instructions_per_launch": valu,
                         "achieved_GIPS": valu / t_launch / 1e9, "peak_GIPS": VALU_PEAK_GIPS,
                         "frac": valu / t_launch / 1e9 / VALU_PEAK_GIPS},
                "l2_hit_rate": c["l2_hit_rate"],
                "counters": os.path.relpath(COUNTERS, ROOT)})
    for k in ("lds_instructions", "l2_atomics", "clock_GHz", "wait_inst_lds_frac"):
        if k in c:
            out[k] = c[k]
    if "clock_GHz" in c:
        out["valu"]["frac_at_measured_clock"] = (valu / t_launch / 1e9
                                                 / (VALU_PEAK_GIPS * c["clock_GHz"] / 2.4))
    return out


def cpu_baseline(ch, hs, cube_np, data_np, target_s):
    """Oracle fwd+bwd on a bounded subset of the views, all host threads."""
    from oracle import oracle as O
    from tomograpy_amd.headers import map_record, view_records
    recs = view_records(hs)
    mrec = map_record(ch)
    threads = O.num_threads()
    nv = len(hs)
    # calibrate on one view per thread, then size the sample to about target_s of CPU time
    k = max(1, min(threads, nv))
    t0 = time.perf_counter()
    d = np.zeros_like(data_np)
    c = np.zeros_like(cube_np)
    s1 = O.siddon(O.PJ, recs, mrec, cube_np, d, t0=0, t1=k, nthread=threads)
    s2 = O.siddon(O.BPJ, recs, mrec, c, data_np, t0=0, t1=k, nthread=threads)
    dt = time.perf_counter() - t0
    per_view = dt / k
    k2 = int(max(k, min(nv, target_s / max(per_view, 1e-6))))
    # OpenMP splits views over threads: use a multiple of the thread count when possible
    if k2 >= threads:
        k2 = (k2 // threads) * threads
    d[:] = 0
    c[:] = 0
    t0 = time.perf_counter()
    s1 = O.siddon(O.PJ, recs, mrec, cube_np, d, t0=0, t1=k2, nthread=threads)
    s2 = O.siddon(O.BPJ, recs, mrec, c, data_np, t0=0, t1=k2, nthread=threads)
    dt = time.perf_counter() - t0
    cpu = {"value": (s1 + s2) / dt, "unit": "Siddon steps/s", "cores": threads,
           "kind": "port",
           "sample": "oracle pj+bpj on views 0..%d of the workload (%d of %d), %.1f s" %
                     (k2 - 1, k2, nv, dt)}
    # one thread, as the reference runs with nthread=1: two views
    k1 = min(2, nv)
    d1 = np.zeros_like(data_np)
    c1 = np.zeros_like(cube_np)
    t0 = time.perf_counter()
    s1 = O.siddon(O.PJ, recs, mrec, cube_np, d1, t0=0, t1=k1, nthread=1)
    s2 = O.siddon(O.BPJ, recs, mrec, c1, data_np, t0=0, t1=k1, nthread=1)
    dt1 = time.perf_counter() - t0
    cpu["single_thread"] = {"value": (s1 + s2) / dt1, "cores": 1,
                            "sample": "oracle pj+bpj on views 0..%d, %.1f s" % (k1 - 1, dt1)}
    cpu["config1"] = cpu_config1(threads)
    return cpu, d[..., :k2], c, k2


def cpu_config1(threads):
    """The oracle on the whole C1 workload (64^3, 32 views of 128^2): steps/s of
    the forward and of the backward alone, fp32 and fp64, on all host threads
    and on one; median of 5 runs each."""
    from oracle import oracle as O
    from tomograpy_amd import workloads as W
    from tomograpy_amd.headers import map_record, view_records
    out = {"workload": "C1: 64^3 cube, 32 views of 128^2", "runs": 5}
    rng = np.random.default_rng(W.SEED0 + 9)
    for dtype in (np.float32, np.float64):
        ch, cube, hs = W.config1(dtype)
        recs, mrec = view_records(hs), map_record(ch)
        x = np.ascontiguousarray(cube, dtype=dtype)
        y = rng.uniform(0, 1, (128, 128, len(hs))).astype(dtype)
        res = {}
        for nt in (threads, 1):
            for name, kind in (("forward", O.PJ), ("backward", O.BPJ)):
                rates = []
                for _ in range(5):
                    d = np.zeros_like(y)
                    c = np.zeros_like(x)
                    t0 = time.perf_counter()
                    steps = (O.siddon(kind, recs, mrec, x, d, nthread=nt) if kind == O.PJ
                             else O.siddon(kind, recs, mrec, c, y, nthread=nt))
                    rates.append(steps / (time.perf_counter() - t0))
                res["%s_%s" % (name, "all_threads" if nt == threads else "one_thread")] = \
                    float(np.median(rates))
        out[np.dtype(dtype).name] = res
    out["threads"] = threads
    out["unit"] = "Siddon steps/s"
    return out


def sample_parity(ch, hs, dtype, cube_np, data_np, d_ref, c_ref, k):
    """The GPU pair on the CPU baseline's sample of views against the oracle's
    output for it: forward bit-exact, backward relative L2."""
    import torch
    from tomograpy_amd.lo_wrapper import siddon_lo
    Ps = siddon_lo(hs[:k], ch, dtype=dtype)
    fwd = Ps.matvec(torch.from_numpy(cube_np).cuda()).cpu().numpy()
    bwd = Ps.rmatvec(torch.from_numpy(np.ascontiguousarray(data_np[..., :k])).cuda()).cpu().numpy()
    c64 = c_ref.astype(np.float64)
    return {"views": k, "forward_bit_exact": bool(np.array_equal(fwd, d_ref)),
            "backward_rel_l2": float(np.linalg.norm(bwd - c64) / np.linalg.norm(c64)),
            "tolerance": 1e-5 if dtype == np.float32 else 1e-12}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only; --gpus %d is not supported (the path would shard "
              "by views with one all-reduce of the backprojected cube)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    import tomograpy_amd as T
    from tomograpy_amd import workloads as W
    from tomograpy_amd.lo_wrapper import siddon_lo

    torch.cuda.set_device(0)
    dtype = np.dtype(args.dtype)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    b = dtype.itemsize
    ch, hs, rng = W.config2_geometry(dtype, n=args.n, n_views=args.views, n_pix=args.pix)
    phantom = W.blob_phantom(ch, 40, rng, dtype=dtype)
    P = siddon_lo(hs, ch, dtype=dtype, tile_j=args.tile_j)
    x_true = torch.from_numpy(phantom).cuda()
    y = P.matvec(x_true).clone()
    noise = torch.from_numpy(rng.normal(size=tuple(y.shape)).astype(dtype)).cuda()
    y += 0.01 * float(y.max()) * noise
    S = P.count_steps()
    R_t = P._rays_traced
    V = len(hs)
    torch.cuda.synchronize()
    plan_info = None
    if P.use_plan():
        t0 = time.perf_counter()
        plan = P.build_plan()
        torch.cuda.synchronize()
        plan_info = {"build_ms": (time.perf_counter() - t0) * 1e3, "records": int(plan.records),
                     "bytes": int(P.plan_bytes()), "brick": list(plan.brick),
                     "chunks": int(plan.chunks)}

    x = x_true.clone()
    stream = torch.cuda.current_stream()
    if args.copy_calibration_mib:
        # copies that read and write exactly this many bytes once: the runtime's
        # buffer copy (dwordx4 loads) and an elementwise kernel (out = src * 1)
        n = args.copy_calibration_mib << 18
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    def step():
        P.xout.zero_()
        P.project_into(x, P.xout)
        P.xin.zero_()
        P.backproject_into(y, P.xin)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms_step = ev0.elapsed_time(ev1) / args.steps

    # per-kernel average launch time, events on the launch stream
    fe = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    be = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    for k in range(args.steps):
        P.xout.zero_()
        fe[2 * k].record(stream)
        P.project_into(x, P.xout)
        fe[2 * k + 1].record(stream)
        P.xin.zero_()
        be[2 * k].record(stream)
        P.backproject_into(y, P.xin)
        be[2 * k + 1].record(stream)
    torch.cuda.synchronize()
    t_fwd = np.mean([fe[2 * k].elapsed_time(fe[2 * k + 1]) for k in range(args.steps)]) * 1e-3
    t_bwd = np.mean([be[2 * k].elapsed_time(be[2 * k + 1]) for k in range(args.steps)]) * 1e-3
    T.check_status()
    B_fwd = S * b + R_t * b + V * 18 * b
    B_bwd = 2 * S * b + R_t * b + V * 18 * b
    if t_bwd >= t_fwd:
        dom, B_dom, t_dom = "siddon_backproject", B_bwd, t_bwd
    else:
        dom, B_dom, t_dom = "siddon_project", B_fwd, t_fwd
    roof = roofline(dom, t_dom, B_dom, args, plan_info)
    roof_other = roofline("siddon_project" if dom == "siddon_backproject" else "siddon_backproject",
                          t_fwd if dom == "siddon_backproject" else t_bwd,
                          B_fwd if dom == "siddon_backproject" else B_bwd, args, plan_info)

    # the adjoint identity <P x, y> = <x, P^T y> on the timed pair's outputs
    lhs = float((P.xout.double() * y.double()).sum())
    rhs = float((x.double() * P.xin.double()).sum())
    adjoint = {"lhs": lhs, "rhs": rhs, "rel_diff": abs(lhs - rhs) / max(abs(lhs), abs(rhs)),
               "tolerance": 1e-5 if dtype == np.float32 else 1e-12}

    # the pair from a cold operator: per-ray step counts, plan build, one pair
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Pc = siddon_lo(hs, ch, dtype=dtype, tile_j=args.tile_j)
    Pc.ray_steps()
    if Pc.use_plan():
        Pc.build_plan()
    Pc.xout.zero_()
    Pc.project_into(x, Pc.xout)
    Pc.xin.zero_()
    Pc.backproject_into(y, Pc.xin)
    torch.cuda.synchronize()
    ms_cold = (time.perf_counter() - t0) * 1e3
    del Pc
    torch.cuda.empty_cache()

    extra = {"adjoint": adjoint}
    # the float32 results whose error against float64 the fp64 leg states
    ref32 = None
    if dtype == np.float32 and args.cg_iters:
        av = axis_views(hs)
        y_other = y.clone()
        y_other[:, :, av] = 0
        ref32 = {"x": x_true.double().cpu(), "y": y.double().cpu(), "axis_views": av,
                 "forward": P.xout.double().cpu(), "backward": P.xin.double().cpu(),
                 "y_other": y_other.double().cpu(),
                 "backward_other": P.rmatvec(y_other).double().cpu()}
        del y_other
    if args.cg_iters:
        keep = {} if ref32 is not None else None
        extra["cg"] = {"fp32": cg_run(P, y, args.cg_iters, keep)}
        # the same solve from a cold operator: step counts and plan build included
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Pc = siddon_lo(hs, ch, dtype=dtype, tile_j=args.tile_j)
        cg_run(Pc, y, args.cg_iters)
        torch.cuda.synchronize()
        extra["cg"]["fp32"]["ms_total_with_setup"] = (time.perf_counter() - t0) * 1e3
        del Pc
        torch.cuda.empty_cache()
        if keep:
            ref32["cg_x"] = keep["x"]
            # the same solve with the axis views' pixels masked out
            Pm = siddon_lo(hs, ch, dtype=dtype, mask=view_mask(hs, args.pix, args.pix, av))
            masked = cg_run(Pm, y, args.cg_iters, keep)
            ref32["cg_x_other"] = keep["x"]
            del Pm
            if av:
                f32 = extra["cg"]["fp32"]
                f32["relative_residual_axis_views_masked"] = masked["relative_residual"]
                f32["note"] = ("float32 CG stalls near a residual of 0.1 because views %s put the "
                               "observer on a coordinate plane, where the reference's float32 ray "
                               "set-up gives a few rays steps of about +-212 (error_vs_fp64); with "
                               "those views' pixels masked it reaches the residual above; float64 "
                               "does not stall" % av)

    y_host = y.cpu().numpy() if not args.no_cpu else None
    if args.serve_requests or args.serve_small_requests or not args.no_grad \
            or not args.no_batch or args.cg_iters \
            or not args.no_shard or not args.no_det:
        # release the pair's operator (its brick plan is 9.4 GB at 256^3) first
        del P, x, x_true, y, noise
        torch.cuda.empty_cache()
    if args.cg_iters and dtype == np.float32:
        extra["cg"]["fp64"], err64 = fp64_leg(args, ref32)
        extra["error_vs_fp64"] = err64
        ref32 = None
        torch.cuda.empty_cache()
    if not args.no_grad:
        extra["differentiation"] = grad_leg()
        torch.cuda.empty_cache()
    if not args.no_batch:
        extra["batch"] = batch_leg()
        torch.cuda.empty_cache()
    if not args.no_shard:
        extra["shard_of_8"] = shard_leg(args)
        torch.cuda.empty_cache()
    if not args.no_det:
        extra["deterministic"] = det_leg(args)
    if args.serve_requests:
        extra["serving"] = serving_leg(args)
    if args.serve_small_requests:
        extra["serving_small"] = serving_leg(args, small=True)
    cpu = None
    if not args.no_cpu:
        y_host = np.ascontiguousarray(y_host)
        cpu, d_ref, c_ref, k = cpu_baseline(ch, hs, phantom, y_host, args.cpu_seconds)
        extra["parity"] = sample_parity(ch, hs, dtype, phantom, y_host, d_ref, c_ref, k)

    value = 2 * S / (ms_step * 1e-3)
    out = {
        "metric": "Siddon steps/s + recon p50/p99 latency, served vs plain, at <=1e-5 rel err",
        "value": value,
        "unit": "Siddon steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if dtype == np.float32 else "f64",
        "data": "synthetic",
        "config": {"workload": "C2 pair: %d^3 cube, %d views of %d^2, D=214, LAT U(-7,7) deg"
                               % (args.n, V, args.pix),
                   "siddon_steps_per_projection": S, "rays_traced": R_t,
                   "ms_forward": t_fwd * 1e3, "ms_backward": t_bwd * 1e3,
                   "ms_cold_pair": ms_cold,
                   "backprojection": "brick plan" if plan_info else "atomic scatter",
                   "plan": plan_info},
        "roofline": roof,
        "roofline_other_kernel": roof_other,
        "cpu_baseline": cpu,
        "wall_s_timed": t_wall,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
