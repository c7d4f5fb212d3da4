"""Continuous-batching reconstruction server.

Requests (an image stack with its headers, and CG settings) arrive over time.
The server admits a waiting request as soon as its HBM footprint (CG
vectors, images and, for a large cube, its brick plan) fits the budget and a
slot is free, keeps its cube, images and CG state on the device, and
advances every admitted request by one CG iteration per engine step: one
batched projector launch over all of them (``tg_batch_run``), one brick-plan
backprojection launch over the requests with a plan (``tg_bpj_plan_batch``),
one batched backprojector launch over the others, and the segmented CG vector
kernels (``tg_cg_batch``).  A request leaves when its relative residual reaches
its tolerance (checked every ``check_every`` of its iterations) or at maxiter;
requests that arrive meanwhile join at the next engine step.  Image stacks
come from pinned host memory and solutions go back to pinned buffers with
stream-ordered copies, so admissions and completions do not stall the GPU work
of the other requests; a request is complete when its solution is on the host.

``serve_plain`` is the per-request baseline: the same requests solved one at
a time by ``cg.solve`` (the same kernels, one job at a time), in arrival order
or in the server's shortest-first order.  Both report p50/p99 latency
(completion - arrival) and throughput.  A request can be cancelled
(``Request.cancel``) or given a timeout; either drops it from the queue or the
active set.
"""
import collections
import time

import numpy as np
import torch

from . import _native as N
from . import cg as CG
from . import siddon as S
from .lo_wrapper import Siddon, Siddon4d


class Request:
    """One reconstruction: headers, image stack (host), solver settings.

    ``y`` may be a numpy array or a host tensor; ``pin()`` moves it to pinned
    memory and gives the request a pinned buffer for its solution, so that the
    server moves both without blocking.
    """

    def __init__(self, rid, cube_header, data_header, y, arrival=0.0, lam=1e-2, tol=1e-4,
                 maxiter=200, mask=None, tindex=None, obstacle=None, dtype=np.float32,
                 timeout=None):
        self.id = rid
        self.cube_header = cube_header
        self.data_header = data_header
        self.y = y
        self.arrival = float(arrival)
        self.lam = lam
        self.tol = tol
        self.maxiter = maxiter
        self.mask = mask
        self.tindex = tindex
        self.obstacle = obstacle
        self.dtype = np.dtype(dtype)
        # seconds after arrival by which the request must be complete; the
        # server drops it (status "timed out") once that has passed
        self.timeout = timeout
        self.status = "queued"  # -> "done", "cancelled" or "timed out"
        self._cancel = False
        self.x = None
        self._pinned = False
        # True: keep a device fingerprint of the solution (solution_fingerprint)
        self.fingerprint = None
        self.iterations = 0
        self.residual = None
        self.t_admit = None
        self.t_done = None
        self._op = None
        self._plan_bytes = None
        self._reserved = 0
        self._work = None
        self.x_host = None

    def footprint(self):
        """Device bytes while admitted, without the brick plan: 4 CG vectors, 2
        operator buffers, the float64 backprojection accumulator, images, the
        per-ray step counts (4 bytes per pixel) and the pixel mask."""
        n = np.prod([self.cube_header["NAXIS%d" % (k + 1)] for k in range(3)])
        nt = 1 if self.tindex is None else int(np.max(self.tindex)) + 1
        n1 = self.data_header[0]["NAXIS1"]
        n2 = self.data_header[0]["NAXIS2"]
        pix = n1 * n2 * len(self.data_header)
        b = self.dtype.itemsize
        trig = len(self.data_header) * (n1 + n2) * 16
        acc = 8 * n * nt if self.dtype == np.float32 else 0
        steps = 4 * pix
        return int((6 * n * nt + 2 * pix) * b + acc + trig + steps
                   + (pix if self.mask is not None else 0) + (1 << 16))

    def work(self):
        """Relative cost of one CG iteration: rays x mean cube edge."""
        if self._work is None:
            n = sum(self.cube_header["NAXIS%d" % (k + 1)] for k in range(3)) / 3.0
            n1 = self.data_header[0]["NAXIS1"]
            n2 = self.data_header[0]["NAXIS2"]
            self._work = float(n1 * n2 * len(self.data_header) * n)
        return self._work

    def result_shape(self):
        n = tuple(int(self.cube_header["NAXIS%d" % (k + 1)]) for k in range(3))
        return n if self.tindex is None else n + (int(np.max(self.tindex)) + 1,)

    def pin(self):
        """Keeps the image stack in pinned host memory (asynchronous upload); the
        solution goes to a pinned buffer taken when the request completes (from
        torch's caching host allocator, so that buffers freed by earlier
        requests are reused)."""
        if torch.cuda.is_available():
            y = torch.as_tensor(np.asarray(self.y))
            self.y = y if y.is_pinned() else y.pin_memory()
            self._pinned = True
        return self

    def reset(self):
        """Forgets the outcome of a run (solution, counts, times, status), so that
        the same request can be served again."""
        self.x = None
        self.x_host = None
        self.iterations = 0
        self.residual = None
        self.t_admit = None
        self.t_done = None
        self.status = "queued"
        self._cancel = False
        self._op = None
        self._plan_bytes = None
        self._reserved = 0
        return self

    def operator(self):
        if self.tindex is not None:
            return Siddon4d(self.data_header, self.cube_header, tindex=self.tindex, mask=self.mask,
                            obstacle=self.obstacle, dtype=self.dtype)
        return Siddon(self.data_header, self.cube_header, mask=self.mask, obstacle=self.obstacle,
                      dtype=self.dtype)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    def cancel(self):
        """Asks the server to drop this request: a waiting request is never
        admitted, an admitted one leaves at the next engine step and frees its
        slot and memory; a completed one is unaffected."""
        self._cancel = True

    def _dropped(self, now):
        """The status to end with if cancelled or past its timeout at `now`."""
        if self._cancel:
            return "cancelled"
        if self.timeout is not None and now - self.arrival > self.timeout:
            return "timed out"
        return None


def make_requests(mix, rng, rate=None, tol=1e-4, lam=1e-2, maxiter=200, dtype=np.float32,
                  noise=0.01, start=0.0, first_id=0):
    """Requests of the served stream from a size mix (workloads.config5_mix).

    Each request images a seeded blob corona (time bins scale it by 1 + 0.05 t)
    from random observers at the mix's distance; y = P x + noise * max * N(0, 1).
    Arrival times are a Poisson process of the given rate (all at 0 if None)
    starting at `start` seconds; ids count from first_id.  Calling it on
    consecutive slices of a mix with the same rng, each call starting at the
    previous call's last arrival, gives the requests of one call on the whole.
    """
    from . import models as Mo
    from . import workloads as W
    reqs = []
    t = float(start)
    for k, spec in enumerate(mix, start=first_id):
        ch = W.cube_header(spec["n"], dtype=dtype)
        lons = rng.uniform(0, 2 * np.pi, spec["nv"])
        lats = np.deg2rad(rng.uniform(-7.0, 7.0, spec["nv"]))
        hs = W.view_headers(spec["npix"], lons, lats, d=spec["d"], dtype=dtype)
        x = W.blob_phantom_torch(ch, 8, rng)
        tindex = None
        if spec["bins"] > 1:
            tindex = (np.arange(spec["nv"]) % spec["bins"]).astype(np.int32)
            x = torch.stack([x * (1 + 0.05 * b) for b in range(spec["bins"])], dim=-1)
        mask = None
        if spec["mask"]:
            mask = Mo.radius_data_mask(hs, spec["npix"], spec["npix"], data_rmin=1.05)
        r = Request(k, ch, hs, None, lam=lam, tol=tol, maxiter=maxiter, mask=mask, tindex=tindex,
                    dtype=dtype)
        P = r.operator()
        y = P.matvec(x.to(P.dev, P.tdtype)).clone()
        nz = torch.from_numpy(rng.standard_normal(tuple(y.shape))).to(y.device, y.dtype)
        y += noise * float(y.abs().max()) * nz
        r.y = y.cpu()
        r.pin()
        if rate:
            t += float(rng.exponential(1.0 / rate))
        r.arrival = t
        reqs.append(r)
    return reqs


def _y_device(req, P, non_blocking=False):
    """The request's image stack on the device, in the operator's layout.  From
    pinned host memory with non_blocking=True the upload is stream-ordered and
    does not block the host."""
    y = torch.as_tensor(np.asarray(req.y)) if not isinstance(req.y, torch.Tensor) else req.y
    buf = torch.empty((P.nv, P.n1, P.n2), dtype=P.tdtype, device=P.dev).permute(1, 2, 0)
    if y.dtype != P.tdtype:
        y = y.to(P.tdtype)
    buf.copy_(y, non_blocking=non_blocking and y.is_pinned())
    return buf


def _torch_dtype(dtype):
    return torch.float32 if np.dtype(dtype) == np.float32 else torch.float64


def solution_fingerprint(x):
    """Two int64 sums over the bit patterns of the solution x (device): of the
    words, and of their squares (wrapping), computed on the device and left
    there.  Equal solutions have equal fingerprints; a difference of one ulp
    anywhere changes them.  Lets a caller compare many served and plain
    solutions bit for bit without keeping or hashing them on the host."""
    w = x.detach().contiguous().view(-1)
    bits = w.view(torch.int32 if w.dtype == torch.float32 else torch.int64).to(torch.int64)
    return torch.stack([bits.sum(), (bits * bits).sum()])


def _download(req, x):
    """Starts the copy of the solution x (device) to the request's host buffer
    (pinned if the request was pinned, so that the copy does not block);
    returns (host tensor, completion event).  Takes the solution's fingerprint
    first when the request asks for one (request.fingerprint = True)."""
    if req.fingerprint is True:
        req.fingerprint = solution_fingerprint(x)
    host = req.x_host
    if host is None or tuple(host.shape) != tuple(x.shape):
        host = torch.empty(tuple(x.shape), dtype=x.dtype, pin_memory=req._pinned)
    host.copy_(x, non_blocking=host.is_pinned())
    ev = torch.cuda.Event()
    ev.record(S.current_stream())
    return host, ev


def summarize(requests, t_first, t_end):
    """Latency percentiles and throughput over the completed requests."""
    done = [r for r in requests if r.status == "done"]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    span = t_end - t_first
    out = {"requests": len(done), "p50_s": float(np.percentile(lat, 50)),
           "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
           "makespan_s": span, "throughput_rps": len(done) / span if span > 0 else float("inf"),
           "iterations": int(sum(r.iterations for r in done))}
    dropped = len(requests) - len(done)
    if dropped:
        out["dropped"] = dropped
    return out


class Server:
    """Continuous batching over CG reconstructions on one GPU."""

    def __init__(self, budget_bytes=64 << 30, max_active=64, check_every=5, tuner=None,
                 order="shortest", speculative="auto", step_work=None):
        """step_work: the work (Request.work units) one engine step may take.  Each
        step advances the admitted requests with the least work per iteration,
        as many as fit in step_work (at least one); the others keep their state
        on the device and wait.  Small requests are then batched together until
        they fill the GPU, and a large request, which fills it alone, does not
        hold them back: a request's iterations are as fast as its batch
        allows.  None = STEP_WORK; float("inf") advances every admitted request
        at every step.
        order: "shortest" admits the waiting request with the least work first
        (traced steps per iteration, estimated from the geometry); "arrival" admits
        in arrival order.  speculative: read each residual check while the next
        step runs (a request ended by it has taken one more step, discarded: its
        solution is the one snapshotted at the check), instead of waiting for the
        check before the next step.  That keeps the GPU fed when steps are short
        (small requests: 2.14x -> 2.33x plain throughput on 64 requests of
        32^3-64^3) and costs the discarded steps when they are long (0.98x
        instead of 1.00x on the 32^3-256^3 mix).  "auto" speculates while the
        active set's work per step is under SPECULATE_WORK."""
        if order not in ("shortest", "arrival"):
            raise ValueError("order must be 'shortest' or 'arrival'")
        self.order = order
        if speculative not in ("auto", True, False):
            raise ValueError("speculative must be True, False or 'auto'")
        self.speculative = speculative
        self.step_work = self.STEP_WORK if step_work is None else float(step_work)
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.check_every = int(check_every)
        # the active set changes every few steps: a fixed plan unless a tuner is given
        self.tuner = tuner
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.scal_pool = torch.zeros((self.max_active, 16), dtype=torch.float64, device=self.dev)
        self.free_rows = list(range(self.max_active))
        # two pinned copies of the scalars: one read by the host, one in flight
        self._scal_host = [torch.empty((self.max_active, 16), dtype=torch.float64,
                                       pin_memory=True) for _ in range(2)]
        self._scal_turn = 0

    # Work (Request.work: rays x mean cube edge) of the active set under which
    # residual checks are read one step late: about 4 ms of GPU time per step.
    SPECULATE_WORK = 1e9
    # Work of one engine step's batch: about 0.3 ms of GPU time, more than a
    # batch needs to fill the 256 CUs with its projector and brick-plan launches.
    STEP_WORK = 2.5e8

    def _run_set(self, active):
        """The admitted requests one engine step advances: the least work per
        iteration first, as many as fit in step_work, at least one."""
        out, w = [], 0.0
        for a in sorted(active, key=lambda a: a[0].work()):
            if out and w + a[0].work() > self.step_work:
                break
            out.append(a)
            w += a[0].work()
        return out

    def _read_scal(self):
        """Starts a copy of every slot's CG scalars to pinned host memory;
        returns (host tensor, completion event)."""
        host = self._scal_host[self._scal_turn]
        self._scal_turn ^= 1
        host.copy_(self.scal_pool, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(S.current_stream())
        return host, ev

    def run(self, requests, on_step=None, on_complete=None):
        """Serves `requests` (arrival times in seconds from now); returns a summary.
        ``on_step(step, requests)``, if given, is called after every engine step
        (between steps a caller may cancel requests); ``on_complete(request)``
        when a request's solution has reached the host."""
        with S.fixed_stream():
            return self._run(requests, on_step, on_complete)

    def _run(self, requests, on_step=None, on_complete=None):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active = []  # (req, solve, row)
        finishing = []  # (req, host solution, copy event): results on their way to the host
        used = 0
        table = None
        t0 = time.perf_counter()
        steps = 0
        run = []  # the requests the current table covers (see _run_set)
        checks = []  # residual checks not yet read (at most the last step's)
        ended = set()
        while pending or waiting or active or finishing or checks:
            finishing = _collect(finishing, t0,
                                 block=not (pending or waiting or active or checks),
                                 on_complete=on_complete)
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # cancellations and timeouts: waiting requests leave the queue, admitted
            # ones leave the active set and give back their slot and memory
            for r in [r for r in waiting if r._dropped(now)]:
                waiting.remove(r)
                r.status, r.t_done = r._dropped(now), now
            gone = [(r, s, row) for r, s, row in active if r._dropped(now)]
            for r, s, row in gone:
                ended.add(id(r))
                r.status, r.t_done = r._dropped(now), now
                r.iterations = s.iterations
                used -= r._reserved
                self.free_rows.append(row)
            if gone:
                active = [a for a in active if id(a[0]) not in ended]
                table = None
            admitted = []
            if self.order == "shortest" and len(waiting) > 1:
                waiting = collections.deque(sorted(waiting, key=lambda r: r.work()))
            while waiting and len(active) < self.max_active:
                req = waiting[0]
                fp = req.footprint()
                if fp > self.budget:
                    raise ValueError("request %s needs %d bytes, budget %d" % (req.id, fp, self.budget))
                if req._plan_bytes is None:
                    # the plan's size needs its record count: build the operator
                    # (kept if the request is admitted now) and count once
                    req._op = req.operator()
                    req._plan_bytes = req._op.plan_reserve_bytes()
                fp += req._plan_bytes
                req._reserved = fp
                if fp > self.budget:
                    raise ValueError("request %s needs %d bytes, budget %d" % (req.id, fp, self.budget))
                if used + fp > self.budget:
                    # a request that waits holds no device memory: its operator's
                    # buffers are not counted in `used`
                    req._op = None
                    break
                waiting.popleft()
                used += fp
                P = req._op if req._op is not None else req.operator()
                req._op = None
                row = self.free_rows.pop()
                s = CG.CGSolve(P, _y_device(req, P, non_blocking=True), lam=req.lam, tol=req.tol,
                               maxiter=req.maxiter, scal=self.scal_pool[row])
                req.t_admit = time.perf_counter() - t0
                active.append((req, s, row))
                admitted.append(s)
            if admitted:
                CG.SegTable(admitted, self.dev).op(N.TG_CG_OP_INIT)
                table = None
            if not active and not checks:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
                continue
            fresh = None
            if active and table is None:
                run = self._run_set(active)
                table = CG.SegTable([s for _, s, _ in run], self.dev)
            spec = self.speculative
            if spec == "auto":
                spec = bool(active) and sum(r.work() for r, _, _ in run) < self.SPECULATE_WORK
            if active:
                CG.iterate([s for _, s, _ in run], table, self.tuner if len(run) > 1 else None)
                steps += 1
                if on_step is not None:
                    on_step(steps, requests)
                due = [(r, s, row) for r, s, row in run
                       if s.iterations % self.check_every == 0 or s.iterations >= r.maxiter]
                if due:
                    # speculative: the solution at the check is kept in case the
                    # check ends the request, which by the time the host reads the
                    # residual has taken one more (discarded) step
                    snaps = [(r, s, row, s.iterations, s.x.clone() if spec else s.x)
                             for r, s, row in due]
                    fresh = (snaps,) + self._read_scal()
            # read every earlier check, and this step's unless speculating
            ready, checks = checks, []
            if fresh is not None:
                if spec:
                    checks = [fresh]
                else:
                    ready = ready + [fresh]
            for snaps, scal_host, ev in ready:
                ev.synchronize()
                scal = scal_host.numpy()
                done = set()
                for r, s, row, its, x in snaps:
                    if id(r) in ended:
                        continue
                    res = s.residual(scal[row])
                    if res <= r.tol or its >= r.maxiter:
                        ended.add(id(r))
                        r.iterations = its
                        r.residual = res
                        host, cev = _download(r, x)
                        finishing.append((r, host, cev))
                        used -= r._reserved
                        self.free_rows.append(row)
                        done.add(id(r))
                if done:
                    active = [(r, s, row) for r, s, row in active if id(r) not in done]
                    table = None
        t_end = time.perf_counter() - t0
        out = summarize(requests, min(r.arrival for r in requests), t_end)
        out["engine_steps"] = steps
        return out


def _collect(finishing, t0, block=False, on_complete=None):
    """Completes the requests whose solutions have reached the host; with
    block=True waits for the oldest.  Returns the ones still in flight."""
    if finishing and block:
        finishing[0][2].synchronize()
    still = []
    for r, host, ev in finishing:
        if ev.query():
            r.x = host.numpy()
            r.t_done = time.perf_counter() - t0
            r.status = "done"
            if on_complete is not None:
                on_complete(r)
        else:
            still.append((r, host, ev))
    return still


def serve_plain(requests, check_every=5, order="arrival", on_complete=None):
    """Per-request baseline: solve each request alone, one after the other.
    order: "arrival" takes them in arrival order; "shortest" takes, among the
    requests that have arrived, the one with the least work per iteration
    first -- the server's admission order, for a like-for-like latency
    comparison.  A request past its timeout when its turn comes is dropped."""
    if order not in ("shortest", "arrival"):
        raise ValueError("order must be 'shortest' or 'arrival'")
    with S.fixed_stream():
        return _serve_plain(requests, check_every, order, on_complete)


def _serve_plain(requests, check_every, order, on_complete=None):
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    while pending:
        now = time.perf_counter() - t0
        if pending[0].arrival > now:
            time.sleep(pending[0].arrival - now)
            now = time.perf_counter() - t0
        req = pending[0]
        if order == "shortest":
            arrived = [r for r in pending if r.arrival <= now]
            req = min(arrived, key=lambda r: r.work())
        pending.remove(req)
        gone = req._dropped(now)
        if gone:
            req.status, req.t_done = gone, now
            continue
        req.t_admit = now
        P = req.operator()
        x, info = CG.solve(P, _y_device(req, P), lam=req.lam, maxiter=req.maxiter, tol=req.tol,
                           check_every=check_every)
        host, ev = _download(req, x)
        ev.synchronize()
        req.x = host.numpy()
        req.iterations = info["iterations"]
        req.residual = info["residual"]
        req.t_done = time.perf_counter() - t0
        req.status = "done"
        if on_complete is not None:
            on_complete(req)
    t_end = time.perf_counter() - t0
    return summarize(requests, min(r.arrival for r in requests), t_end)
