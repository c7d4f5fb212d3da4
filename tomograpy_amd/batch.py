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

import numpy as np
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
        ev.record(torch.cuda.current_stream())
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


def run(ops, ins, outs, backward=False, plan=None, zero_out=False):
    """outs[k] += P_k ins[k] (or P_k^T ins[k] when backward) for all k, on the
    current stream: one persistent launch for a forward batch; for a backward
    batch, one launch over the brick plans of the jobs that use one and one
    persistent atomic launch over the others.

    ops: Siddon operators (same dtype); ins/outs: device tensors shaped like the
    operator's cube / image stack (outputs are accumulated into, as in the
    one-job calls; with zero_out=True a forward launch overwrites them as if
    they had been zeroed).  Voxel weights of a backward Thomson job are applied
    to its output after the launch.
    """
    n = len(ops)
    if n == 0:
        return outs
    calls = (N.TgCall * n)()
    post = []
    copies = []
    alive = []  # temporaries the launches read: referenced until they are enqueued
    weighted = set()
    for k, (P, x, y) in enumerate(zip(ops, ins, outs)):
        if backward:
            target = y
            if P.vox_weight is not None:
                target = torch.zeros_like(y)
                post.append((P, target, y))
            data = P._as_stack(x)
            alive.append(data)
            c = P._call(N.TG_BPJ, target, data)
        else:
            target = y
            if P.pix_weight is not None and y.stride() != P.pix_weight.stride():
                # pixel weights live in the operator's view-major layout
                target = P._as_stack(y)
                copies.append((target, y))
            cube = P._as_cube(x)
            alive.append(cube)
            vox_out = None
            if P.vox_weight is not None or P.voxmask is not None:
                # an operator's weighted-map buffer serves its first job only
                if id(P) in weighted:
                    vox_out = torch.empty_like(P.xin)
                    alive.append(vox_out)
                weighted.add(id(P))
            c = P._call(N.TG_PJ, cube, target, vox_out=vox_out)
            if zero_out:
                c.flags |= N.TG_ZERO_OUT
        ctypes.memmove(ctypes.byref(calls[k]), ctypes.byref(c), ctypes.sizeof(N.TgCall))
    s = S.stream_handle()
    dev = ops[0].dev
    if backward:
        # brick-plan jobs in one plan launch, the others in one atomic launch
        planned = [k for k, P in enumerate(ops) if P.use_plan()]
        rest = sorted(set(range(n)) - set(planned))
        if planned:
            m = len(planned)
            pcalls = (N.TgCall * m)()
            plans = (N.TgBpjPlan * m)()
            for q, k in enumerate(planned):
                ctypes.memmove(ctypes.byref(pcalls[q]), ctypes.byref(calls[k]), ctypes.sizeof(N.TgCall))
                p = ops[k].build_plan()
                ctypes.memmove(ctypes.byref(plans[q]), ctypes.byref(p), ctypes.sizeof(N.TgBpjPlan))
            ring = _ring("plan", dev)
            b = ring.take(int(N.lib().tg_bpj_plan_batch_bytes(m)), dev)
            N.check(N.lib().tg_bpj_plan_batch(pcalls, plans, m, b.queue.data_ptr(),
                                              b.stage.data_ptr(), s))
            ring.release(b)
        if rest:
            m = len(rest)
            acalls = (N.TgCall * m)()
            for q, k in enumerate(rest):
                ctypes.memmove(ctypes.byref(acalls[q]), ctypes.byref(calls[k]), ctypes.sizeof(N.TgCall))
            _batch_launch(acalls, m, plan, s, dev)
    else:
        _batch_launch(calls, n, plan, s, dev)
    del alive
    for tmp, y in copies:
        y.copy_(tmp)
    for P, tmp, y in post:
        sc = S.make_call(N.TG_BPJ, P.code, P.xout.shape, P.xout.stride(), tmp.shape, tmp.stride(),
                         P.map_rec, P.views, tmp.data_ptr(), P.xout.data_ptr(),
                         vox_weight=P.vox_weight)
        N.check(N.lib().tg_scale_cube(ctypes.byref(sc), S.stream_handle()))
        y.add_(tmp)
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
