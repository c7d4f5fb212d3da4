"""Device-resident conjugate gradient on the regularised normal equations

    (P^T P + lam (I + sum_k D_k^T D_k)) x = P^T y

the system lo.acg solves for the reference's models (srt.py:225,
exemples/test_siddon_simu_lo.py:31, models.py:85-94 for the lo.diff priors).
``lo`` is not vendored, so its stopping rule and hyper-parameter scaling are
unpinned; this is textbook CG (Hestenes-Stiefel) with the residual test
||r|| <= tol ||r0||.  Every vector stays on the GPU; the CG scalars live in a
device array, so a fixed-iteration solve makes no host round trip and a
residual-stopping solve synchronises once every ``check_every`` iterations.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from . import siddon as S


class NormalOperator:
    """x -> P^T P x + lam ((x if identity) + sum_k D_k^T D_k x) on the device."""

    def __init__(self, P, lam=1e-2, identity=True, axes=None):
        self.P = P
        self.lam = float(lam)
        self.identity = bool(identity)
        nd = len(P.shapein)
        self.axes = tuple(range(nd)) if axes is None else tuple(axes)
        self.mask_bits = 0
        for a in self.axes:
            self.mask_bits |= 1 << a
        self.shape4 = (ctypes.c_int32 * 4)(*(list(P.shapein) + [1] * (4 - nd)))
        self.code = P.code
        self._y = None

    def apply_into(self, x, out):
        """out = A x (out is overwritten)."""
        P = self.P
        y = P.xout
        y.zero_()
        P.project_into(x, y)
        out.zero_()
        P.backproject_into(y, out)
        if self.lam != 0.0:
            xc = x if x.is_contiguous() else x.contiguous()
            oc = out
            if not out.is_contiguous():
                raise ValueError("regularizer output must be contiguous")
            vm = P.voxmask.data_ptr() if getattr(P, "voxmask", None) is not None else None
            N.check(N.lib().tg_regularize(self.code, self.shape4, self.mask_bits, self.lam,
                                          int(self.identity), xc.data_ptr(), oc.data_ptr(),
                                          vm, S.stream_handle()))
        return out


def _flat(t):
    return t.reshape(-1)


def cg(A, b, x0=None, maxiter=50, tol=None, check_every=10, callback=None):
    """Solves A x = b with A a NormalOperator.  b: device tensor shaped like the cube.

    Returns (x, info) with info = {"iterations", "residual" (||r||/||r0||)}.
    """
    P = A.P
    code = A.code
    dev = b.device
    n = b.numel()
    shape = tuple(P.shapein)
    x = torch.zeros(shape, dtype=b.dtype, device=dev) if x0 is None else x0.clone().reshape(shape).contiguous()
    # contiguous working vectors in the operator's cube layout
    r = torch.empty(shape, dtype=b.dtype, device=dev)
    q = torch.empty(shape, dtype=b.dtype, device=dev)
    if x0 is None:
        r.copy_(b.reshape(shape))
    else:
        A.apply_into(x, q)
        r.copy_(b.reshape(shape) - q)
    p = r.clone()
    scal = torch.zeros(N.TG_CG_SCAL, dtype=torch.float64, device=dev)
    lib = N.lib()
    s = S.stream_handle()
    N.check(lib.tg_cg_dot(code, n, r.data_ptr(), r.data_ptr(), scal.data_ptr(), 0, s))
    N.check(lib.tg_cg_dot(code, n, r.data_ptr(), r.data_ptr(), scal.data_ptr(), 5, s))
    it = 0
    res = 1.0
    for it in range(1, maxiter + 1):
        A.apply_into(p, q)
        N.check(lib.tg_cg_dot(code, n, p.data_ptr(), q.data_ptr(), scal.data_ptr(), 1, s))
        N.check(lib.tg_cg_update(code, n, x.data_ptr(), r.data_ptr(), p.data_ptr(), q.data_ptr(),
                                 scal.data_ptr(), s))
        N.check(lib.tg_cg_direction(code, n, r.data_ptr(), p.data_ptr(), scal.data_ptr(), s))
        if callback is not None:
            callback(it, x)
        if tol is not None and (it % check_every == 0 or it == maxiter):
            v = scal[[0, 5]].cpu().numpy()
            res = float(np.sqrt(v[0] / v[1])) if v[1] > 0 else 0.0
            if res <= tol:
                break
    v = scal[[0, 5]].cpu().numpy()
    res = float(np.sqrt(v[0] / v[1])) if v[1] > 0 else 0.0
    return x, {"iterations": it, "residual": res}


def reconstruct(P, y, lam=1e-2, identity=True, maxiter=50, tol=None, x0=None, axes=None,
                check_every=10):
    """Solves the regularised normal equations for data y (device, P.shapeout)."""
    A = NormalOperator(P, lam=lam, identity=identity, axes=axes)
    b = torch.zeros(tuple(P.shapein), dtype=P.tdtype, device=P.dev)
    P.backproject_into(P._as_stack(y), b)
    return cg(A, b, x0=x0, maxiter=maxiter, tol=tol, check_every=check_every)
