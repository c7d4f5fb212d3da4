"""Device-resident conjugate gradient on the regularised normal equations

    (P^T P + lam (I + sum_k D_k^T D_k)) x = P^T y

the system lo.acg solves for the reference's models (srt.py:225,
exemples/test_siddon_simu_lo.py:31; the priors lo.diff of models.py:85-94).
``lo`` is not vendored, so its stopping rule and hyper-parameter scaling are
unpinned; this is textbook CG (Hestenes-Stiefel) stopped on ||r|| <= tol ||r0||.

Every vector and scalar stays on the GPU.  One solve and many concurrent
solves (the reconstruction server) run the same kernels: the segmented CG
kernels (``tg_cg_batch``) take a table of solves, and the projector pair runs
as one-job launches for one solve or batched launches for many.  A solve
computes the same iterates alone or in a mix: the forward projection is
bit-identical either way, and float32 backprojections sum in float64
(``TG_ACC_F64``), which makes them independent of the summation order.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from . import batch as B
from . import siddon as S
from .lo_wrapper import Siddon4d


class CGSolve:
    """Vectors, scalars and regulariser of one solve on the device.

    Vectors are stored contiguously in the operator's storage layout: the cube
    (N0, N1, N2), or (T, N0, N1, N2) for a time-binned operator (presented to
    the projector as (N0, N1, N2, T)).
    """

    def __init__(self, P, y, lam=1e-2, identity=True, axes=None, tol=1e-4, maxiter=50,
                 scal=None, weight=None):
        self.P = P
        self.lam = float(lam)
        self.identity = bool(identity)
        self.tol = tol
        self.maxiter = int(maxiter)
        self.iterations = 0
        self.done = False
        dt, dev = P.tdtype, P.dev
        self.time = isinstance(P, Siddon4d)
        if self.time:
            n0, n1, n2 = P.cube_shape
            self.store_shape = (P.nt, n0, n1, n2)
            # user axes refer to (N0, N1, N2, T); storage puts T first
            ax = tuple(range(4)) if axes is None else tuple(axes)
            self.store_axes = tuple((a + 1) if a < 3 else 0 for a in ax)
        else:
            self.store_shape = tuple(P.cube_shape)
            self.store_axes = tuple(range(3)) if axes is None else tuple(axes)
        self.n = int(np.prod(self.store_shape))
        mk = lambda: torch.zeros(self.store_shape, dtype=dt, device=dev)
        self.xs, self.rs, self.ps, self.qs = mk(), mk(), mk(), mk()
        self.scal = scal if scal is not None else torch.zeros(16, dtype=torch.float64, device=dev)
        self.scal.zero_()
        self.mask = None
        vm = getattr(P, "voxmask", None)
        if vm is not None:
            self.mask = vm.expand(self.store_shape).contiguous() if self.time else vm
        self.weight = _prior_weight(P, axes, weight, self.store_shape)
        self.y = P._as_stack(y)
        self.out = P.xout  # forward output buffer of this solve
        # the solve's packed batched-launch calls (its operands never change)
        self.packs = {}
        # r0 = P^T y (x0 = 0), p0 = r0
        P.backproject_into(self.y, self.view(self.rs))
        self.ps.copy_(self.rs)

    def view(self, store):
        """The projector's view of a storage tensor."""
        return store.permute(1, 2, 3, 0) if self.time else store

    @property
    def x(self):
        return self.view(self.xs)

    def seg(self, block0):
        g = N.TgCgSeg()
        g.x, g.r, g.p, g.q = (t.data_ptr() for t in (self.xs, self.rs, self.ps, self.qs))
        if self.mask is not None:
            g.voxmask = self.mask.data_ptr()
        if self.weight is not None:
            g.weight = self.weight.data_ptr()
        g.scal = self.scal.data_ptr()
        g.n = self.n
        g.block0 = int(block0)
        shp = list(self.store_shape) + [1] * (4 - len(self.store_shape))
        for k in range(4):
            g.shape[k] = int(shp[k])
        bits = 0
        for a in self.store_axes:
            bits |= 1 << a
        g.axes = bits
        g.identity = int(self.identity)
        g.lam = self.lam
        return g

    def residual(self, scal_host=None):
        v = self.scal.cpu().numpy() if scal_host is None else scal_host
        return float(np.sqrt(v[0] / v[5])) if v[5] > 0 else 0.0


def _prior_weight(P, axes, weight, shape):
    """The height prior's per-voxel weight on the device (operator dtype,
    `shape`), from `weight` or from a SmoothnessPrior passed as `axes`; None
    without one.  Time-binned maps have none (the reference's _radius_map,
    models.py:96-104, unpacks three axes and fails on a 4-D map)."""
    if weight is None:
        weight = getattr(axes, "weight", None)
    if weight is None:
        return None
    if isinstance(P, Siddon4d):
        raise NotImplementedError("the height prior is defined for 3-D maps only")
    w = weight if isinstance(weight, torch.Tensor) else torch.from_numpy(np.asarray(weight))
    w = w.to(P.dev, P.tdtype).contiguous()
    if tuple(w.shape) != tuple(shape):
        raise ValueError("prior weight of shape %s for a map of shape %s"
                         % (tuple(w.shape), tuple(shape)))
    return w


class SegTable:
    """Device table of tg_cg_seg for a list of solves, plus reduction scratch."""

    def __init__(self, solves, dev):
        lib = N.lib()
        self.nseg = len(solves)
        segs = (N.TgCgSeg * self.nseg)()
        block = 0
        for k, s in enumerate(solves):
            g = s.seg(block)
            ctypes.memmove(ctypes.byref(segs[k]), ctypes.byref(g), ctypes.sizeof(N.TgCgSeg))
            block += int(lib.tg_cg_seg_blocks(s.n))
        self.nblocks = block
        # staged in pinned memory (a ring of staging buffers reused once their
        # copies are done), so the upload is stream-ordered and does not wait
        # for the work already queued
        nbytes = ctypes.sizeof(segs)
        ring = B._ring("cgseg", dev)
        st = ring.take(nbytes, dev)
        ctypes.memmove(st.stage.data_ptr(), ctypes.addressof(segs), nbytes)
        self.dev_table = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.dev_table.copy_(st.stage[:nbytes], non_blocking=True)
        ring.release(st)
        self.partials = torch.empty(max(block, 1), dtype=torch.float64, device=dev)
        self.code = solves[0].P.code
        self.launches = None  # the batched forward and backward of a step, packed once

    def op(self, op):
        N.check(N.lib().tg_cg_batch(self.code, op, self.dev_table.data_ptr(), self.nseg,
                                    self.nblocks, self.partials.data_ptr(), S.stream_handle()))


def _forward_one(s):
    P = s.P
    c = P._call(N.TG_PJ, s.view(s.ps), s.out)
    c.flags |= N.TG_ZERO_OUT
    N.check(N.lib().tg_conic_full_projector(ctypes.byref(c), S.stream_handle()))


def _backward_one(s):
    s.P.backproject_into(s.out, s.view(s.qs))


def iterate(solves, table, tuner=None):
    """One CG iteration of every solve in `solves` (table built over the same list)."""
    table.op(N.TG_CG_OP_ZERO_Q)
    if len(solves) == 1:
        _forward_one(solves[0])
        _backward_one(solves[0])
    else:
        if table.launches is None:
            # the operands of a step are the same buffers every iteration: pack
            # the batched calls (and pick the tuned plans) once per set of solves
            ops = [s.P for s in solves]
            ins = [s.view(s.ps) for s in solves]
            outs = [s.out for s in solves]
            qs = [s.view(s.qs) for s in solves]
            plan_f = tuner.plan_for(ops, ins, outs, False, zero_out=True) if tuner else None
            plan_b = tuner.plan_for(ops, outs, qs, True) if tuner else None
            packs = [s.packs for s in solves]
            table.launches = (B.Launch(ops, ins, outs, backward=False, zero_out=True, packs=packs),
                              plan_f, B.Launch(ops, outs, qs, backward=True, packs=packs), plan_b)
        fwd, plan_f, bwd, plan_b = table.launches
        fwd.run(plan_f)
        # geometries with a brick plan backproject in one plan launch, the
        # others in one persistent atomic launch
        bwd.run(plan_b)
    table.op(N.TG_CG_OP_REGULARIZE_DOT)
    table.op(N.TG_CG_OP_UPDATE)
    table.op(N.TG_CG_OP_DIRECTION)
    for s in solves:
        s.iterations += 1


def solve(P, y, lam=1e-2, identity=True, axes=None, maxiter=50, tol=None, check_every=10,
          callback=None, weight=None):
    """Solves the regularised normal equations for data y (device, P.shapeout).

    ``axes`` are the smoothness priors' axes (default: every axis), or a
    ``models.SmoothnessPrior``, which also carries the per-voxel weight of the
    height prior; ``weight`` gives that weight explicitly.

    Returns (x, info): x in the projector's layout, info = {"iterations",
    "residual" (||r|| / ||r0||)}.
    """
    s = CGSolve(P, y, lam=lam, identity=identity, axes=axes, tol=tol, maxiter=maxiter,
                weight=weight)
    table = SegTable([s], P.dev)
    table.op(N.TG_CG_OP_INIT)
    for it in range(1, maxiter + 1):
        iterate([s], table)
        if callback is not None:
            callback(it, s.x)
        if tol is not None and (it % check_every == 0 or it == maxiter):
            if s.residual() <= tol:
                break
    return s.x, {"iterations": s.iterations, "residual": s.residual()}


reconstruct = solve


class NormalOperator:
    """x -> P^T P x + lam ((x if identity) + sum_k D_k^T D_k x): the operator
    ``solve`` applies, for use outside a solve (benchmarks, checks)."""

    def __init__(self, P, lam=1e-2, identity=True, axes=None, weight=None):
        self.P = P
        self.lam = float(lam)
        self.identity = bool(identity)
        nd = len(P.shapein)
        self.axes = tuple(range(nd)) if axes is None else tuple(axes)
        self.weight = _prior_weight(P, axes, weight, tuple(P.shapein))
        bits = 0
        for a in self.axes:
            bits |= 1 << a
        self.mask_bits = bits
        self.shape4 = (ctypes.c_int32 * 4)(*(list(P.shapein) + [1] * (4 - nd)))

    def apply_into(self, x, out):
        """out = A x for contiguous x, out shaped like P.shapein."""
        P = self.P
        c = P._call(N.TG_PJ, x, P.xout)
        c.flags |= N.TG_ZERO_OUT
        N.check(N.lib().tg_conic_full_projector(ctypes.byref(c), S.stream_handle()))
        out.zero_()
        P.backproject_into(P.xout, out)
        if self.lam != 0.0:
            vm = P.voxmask.data_ptr() if getattr(P, "voxmask", None) is not None else None
            w = self.weight.data_ptr() if self.weight is not None else None
            N.check(N.lib().tg_regularize(P.code, self.shape4, self.mask_bits, self.lam,
                                          int(self.identity), x.data_ptr(), out.data_ptr(), vm,
                                          w, S.stream_handle()))
        return out
