"""Many uneven projector / backprojector jobs in one persistent launch, with
an autotuner for the ray-tile shape and the schedule.

A job is a ``lo_wrapper.Siddon`` operator plus its operands; jobs may differ
in cube size, view count, image size, observer distance, obstacle and
Thomson weights.  ``run`` launches all of them at once (``tg_batch_run``);
each job's output equals its one-job call's (the forward projection bit for
bit).  ``Autotuner`` times the candidate plans on the actual batch and keeps
the fastest per batch signature.
"""
import ctypes

import torch

from . import _native as N
from . import siddon as S

TILE_CHOICES = (0, 16, 64, 256)  # 0: each job's default (its image width, at most 256)
SCHEDULES = (0, 1)


class _Staging:
    """A device job table and its pinned host staging buffer, with an event
    recorded after the table's host-to-device copy is enqueued: the staging
    buffer may be written again once that event has completed."""

    def __init__(self, nbytes, dev):
        self.nbytes = nbytes
        self.queue = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.stage = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        self.copied = None


class _Ring:
    """A few staging buffers used in turn, so that a launch rarely waits for the
    copy of the previous one (a stream synchronisation per launch otherwise)."""

    DEPTH = 4

    def __init__(self):
        self.slots = []
        self.i = 0

    def take(self, nbytes, dev):
        if len(self.slots) < self.DEPTH:
            self.slots.append(None)
        k = self.i % len(self.slots)
        self.i += 1
        b = self.slots[k]
        if b is None or b.nbytes < nbytes:
            if b is not None and b.copied is not None:
                b.copied.synchronize()
            b = _Staging(max(nbytes, 1 << 16), dev)
            self.slots[k] = b
        elif b.copied is not None:
            b.copied.synchronize()
        return b

    def release(self, b):
        """Marks the launch that read b.stage as enqueued on the current stream."""
        ev = torch.cuda.Event()
        ev.record(S.current_stream())
        b.copied = ev


_RINGS = {}


def _ring(kind, dev):
    key = (kind, str(dev))
    r = _RINGS.get(key)
    if r is None:
        r = _RINGS[key] = _Ring()
    return r


def make_plan(tile_j=0, blocks=0, schedule=1):
    p = N.TgBatchPlan()
    p.tile_j = int(tile_j)
    p.blocks = int(blocks)
    p.schedule = int(schedule)
    return p


class Launch:
    """A batched launch prepared once and run any number of times.

    The job calls (and, for a backward batch, the split between brick-plan and
    atomic jobs) are packed at construction; every ``run`` launches them again
    on the operands' current contents.  A CG engine step over a fixed set of
    solves therefore packs its calls once, not once per iteration.

    ops: Siddon operators (same dtype); ins/outs: device tensors shaped like the
    operator's cube / image stack.  Outputs are accumulated into, as in the
    one-job calls; with zero_out=True a forward launch overwrites them as if
    they had been zeroed.  Voxel operands of a forward job are applied to its
    map before each launch; voxel weights of a backward Thomson job are applied
    to its output after it.
    """

    def __init__(self, ops, ins, outs, backward=False, zero_out=False, packs=None):
        """packs: optional per-job dicts that keep each job's packed call between
        launches over the same operands (a CG solve's fixed buffers), so that a
        new set of jobs reuses the calls of the jobs it shares with an earlier
        one.  A job whose operator repeats an earlier job's in this batch is
        packed afresh (it needs accumulators of its own)."""
        n = len(ops)
        self.n = n
        self.backward = bool(backward)
        self.zero_out = bool(zero_out)
        self.dev = ops[0].dev if n else None
        self.refresh = []    # (operator-layout copy, caller's operand): refreshed per run
        self.copies = []     # (operator-layout output, caller's output): copied back per run
        self.post = []       # (operator, temporary, output): backward voxel weights
        self.prepass = []    # (operator, map, weighted-map buffer): forward voxel operands
        self.alive = []      # every tensor the packed calls point to
        calls = (N.TgCall * n)()
        seen = set()
        accs = set()
        targets = set()
        for k, (P, x, y) in enumerate(zip(ops, ins, outs)):
            if y.data_ptr() in targets:
                raise ValueError("job %d: a batch cannot list one output twice" % k)
            targets.add(y.data_ptr())
            key = (self.backward, self.zero_out, x.data_ptr(), tuple(x.shape), tuple(x.stride()),
                   y.data_ptr(), tuple(y.shape), tuple(y.stride()))
            first = id(P) not in seen
            seen.add(id(P))
            pack = packs[k].get(key) if (packs is not None and first) else None
            if pack is None:
                pack = self._pack(P, x, y, first)
                if packs is not None and first:
                    packs[k][key] = pack
            if self.backward and pack[0].flags & N.TG_ACC_F64 and pack[0].acc in accs:
                # another operator's job sums into this accumulator: one of its own
                pack = self._pack(P, x, y, False)
            call, refresh, copies, post, prepass, alive = pack
            if self.backward and call.flags & N.TG_ACC_F64:
                accs.add(call.acc)
            self.refresh += refresh
            self.copies += copies
            self.post += post
            self.prepass += prepass
            self.alive += alive + [P]
            ctypes.memmove(ctypes.byref(calls[k]), ctypes.byref(call), ctypes.sizeof(N.TgCall))
        self.calls = calls
        self.pcalls = self.plans = self.acalls = None
        if self.backward:
            # brick-plan jobs in one plan launch, the others in one atomic launch
            planned = [k for k, P in enumerate(ops) if P.use_plan()]
            rest = sorted(set(range(n)) - set(planned))
            if planned:
                self.pcalls = (N.TgCall * len(planned))()
                self.plans = (N.TgBpjPlan * len(planned))()
                for q, k in enumerate(planned):
                    ctypes.memmove(ctypes.byref(self.pcalls[q]), ctypes.byref(calls[k]),
                                   ctypes.sizeof(N.TgCall))
                    p = ops[k].build_plan()
                    ctypes.memmove(ctypes.byref(self.plans[q]), ctypes.byref(p),
                                   ctypes.sizeof(N.TgBpjPlan))
            if rest:
                self.acalls = (N.TgCall * len(rest))()
                for q, k in enumerate(rest):
                    ctypes.memmove(ctypes.byref(self.acalls[q]), ctypes.byref(calls[k]),
                                   ctypes.sizeof(N.TgCall))

    def _pack(self, P, x, y, first):
        """One job's packed call and what its runs need: (call, refresh, copies,
        post, prepass, alive), as the lists of the same names above.  `first`:
        P has no earlier job in this batch (it then uses its own float64
        accumulator and weighted-map buffer; a repeat gets fresh ones)."""
        refresh, copies, post, prepass, alive = [], [], [], [], []

        def operand(t, xx):
            # the operator-layout operand t of the caller's xx (a copy when
            # layouts, devices or dtypes differ, then refreshed on every run)
            alive.append(t)
            if not isinstance(xx, torch.Tensor) or t.data_ptr() != xx.data_ptr():
                refresh.append((t, xx))
            return t

        if self.backward:
            target = y
            if P.vox_weight is not None:
                target = torch.zeros_like(y)
                post.append((P, target, y))
            data = operand(P._as_stack(x), x)
            c = P._call(N.TG_BPJ, target, data)
            if (c.flags & N.TG_ACC_F64) and not first:
                # the operator's float64 accumulator serves its first job; a
                # repeat of the operator sums into a buffer of its own
                own = torch.zeros_like(P.acc_buffer())
                alive.append(own)
                c.acc = own.data_ptr()
        else:
            target = y
            if P.pix_weight is not None and y.stride() != P.pix_weight.stride():
                # pixel weights live in the operator's view-major layout
                target = P._as_stack(y)
                copies.append((target, y))
            cube = operand(P._as_cube(x), x)
            if P.vox_weight is not None or P.voxmask is not None:
                # an operator's weighted-map buffer serves its first job only
                wmap = P.vox_buffer(None if first else torch.empty_like(P.xin))
                prepass.append((P, cube, wmap))
                alive.append(wmap)
                cube = wmap
            c = P._call(N.TG_PJ, cube, target, apply_vox=False)
            if self.zero_out:
                c.flags |= N.TG_ZERO_OUT
        alive.append(target)
        return c, refresh, copies, post, prepass, alive

    def run(self, plan=None):
        if self.n == 0:
            return
        for t, x in self.refresh:
            t.copy_(torch.as_tensor(x).reshape(t.shape))
        for P, cube, wmap in self.prepass:
            P.vox_apply(cube, wmap)
        if not self.zero_out:
            for tmp, y in self.copies:
                tmp.copy_(y)
        for P, tmp, y in self.post:
            tmp.zero_()
        s = S.stream_handle()
        if self.backward:
            if self.pcalls is not None:
                m = len(self.pcalls)
                ring = _ring("plan", self.dev)
                b = ring.take(int(N.lib().tg_bpj_plan_batch_bytes(m)), self.dev)
                N.check(N.lib().tg_bpj_plan_batch(self.pcalls, self.plans, m, b.queue.data_ptr(),
                                                  b.stage.data_ptr(), s))
                ring.release(b)
            if self.acalls is not None:
                _batch_launch(self.acalls, len(self.acalls), plan, s, self.dev)
        else:
            _batch_launch(self.calls, self.n, plan, s, self.dev)
        for tmp, y in self.copies:
            y.copy_(tmp)
        for P, tmp, y in self.post:
            sc = S.make_call(N.TG_BPJ, P.code, P.xout.shape, P.xout.stride(), tmp.shape,
                             tmp.stride(), P.map_rec, P.views, tmp.data_ptr(), P.xout.data_ptr(),
                             vox_weight=P.vox_weight)
            N.check(N.lib().tg_scale_cube(ctypes.byref(sc), s))
            y.add_(tmp)


def run(ops, ins, outs, backward=False, plan=None, zero_out=False):
    """outs[k] += P_k ins[k] (or P_k^T ins[k] when backward) for all k, on the
    current stream: one persistent launch for a forward batch; for a backward
    batch, one launch over the brick plans of the jobs that use one and one
    persistent atomic launch over the others.  See ``Launch`` for the operands."""
    Launch(ops, ins, outs, backward, zero_out).run(plan)
    return outs


def _batch_launch(calls, n, plan, s, dev):
    ring = _ring("batch", dev)
    b = ring.take(int(N.lib().tg_batch_queue_bytes(n)), dev)
    plan = plan if plan is not None else make_plan()
    N.check(N.lib().tg_batch_run(calls, n, ctypes.byref(plan), b.queue.data_ptr(),
                                 b.stage.data_ptr(), s))
    ring.release(b)


def signature(ops, backward):
    return (bool(backward),) + tuple((P.cube_shape, P.shapeout, P.flags, P.pix_weight is not None,
                                      P.vox_weight is not None, P.voxmask is not None) for P in ops)


class Autotuner:
    """Picks the (tile_j, schedule) plan with the lowest measured launch time."""

    def __init__(self, tiles=TILE_CHOICES, schedules=SCHEDULES, repeats=2):
        self.tiles = tiles
        self.schedules = schedules
        self.repeats = repeats
        self.best = {}
        self.trials = {}

    def plan_for(self, ops, ins, outs, backward=False, zero_out=False):
        key = signature(ops, backward)
        if key in self.best:
            return self.best[key]
        saved = [o.clone() for o in outs]
        results = []
        for tj in self.tiles:
            for sch in self.schedules:
                plan = make_plan(tile_j=tj, schedule=sch)
                run(ops, ins, outs, backward, plan, zero_out=zero_out)  # warm-up
                torch.cuda.synchronize()
                t = []
                for _ in range(self.repeats):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    run(ops, ins, outs, backward, plan, zero_out=zero_out)
                    e1.record()
                    torch.cuda.synchronize()
                    t.append(e0.elapsed_time(e1))
                results.append((min(t), tj, sch))
        for o, s0 in zip(outs, saved):
            o.copy_(s0)
        results.sort()
        _, tj, sch = results[0]
        self.trials[key] = results
        self.best[key] = make_plan(tile_j=tj, schedule=sch)
        return self.best[key]


def run_tuned(tuner, ops, ins, outs, backward=False):
    return run(ops, ins, outs, backward, tuner.plan_for(ops, ins, outs, backward))
