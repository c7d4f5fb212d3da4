"""The projector as a differentiable torch operation.

``project(x, P)`` returns P x; its hand-written backward is the adjoint,
the (masked, time-binned, Thomson-weighted -- whatever P carries) Siddon
backprojector: grad_x = P^T grad_y.  P is a ``lo_wrapper.Siddon`` or
``Siddon4d`` operator; the output is a fresh tensor (not P's cached buffer),
so several projections can live in one autograd graph.

``smoothness(x, weights)`` is the finite-difference prior of the regularised
models (lo.diff, models.py:85-94; the temporal term of stsrt, models.py:140)
as one differentiable HIP pass.
"""
import ctypes

import torch

from . import _native as N
from . import siddon as S
from .lo_wrapper import stack_buffer


class SiddonProject(torch.autograd.Function):
    """y = P x with dy/dx^T v = P^T v (the backprojector)."""

    @staticmethod
    def forward(ctx, x, P):
        ctx.P = P
        y = stack_buffer(P.n1, P.n2, P.nv, P.tdtype, P.dev)
        P.project_into(x.detach().reshape(P.shapein), y)
        return y

    @staticmethod
    def backward(ctx, gy):
        P = ctx.P
        # the operator's own layout (bin-major for time bins): the brick plan then
        # sums in float64, independent of the order of summation
        gx = P.zeros_in()
        P.backproject_into(gy.contiguous() if gy.stride() != P.xout.stride() else gy, gx)
        return gx, None


def project(x, P):
    """Differentiable conic projection of x through operator P."""
    return SiddonProject.apply(x, P)


class SiddonBackproject(torch.autograd.Function):
    """x = P^T y with the projector as its backward."""

    @staticmethod
    def forward(ctx, y, P):
        ctx.P = P
        x = P.zeros_in()
        P.backproject_into(y.detach(), x)
        return x

    @staticmethod
    def backward(ctx, gx):
        P = ctx.P
        gy = stack_buffer(P.n1, P.n2, P.nv, P.tdtype, P.dev)
        P.project_into(gx.contiguous(), gy)
        return gy, None


def backproject(y, P):
    """Differentiable backprojection of y through operator P."""
    return SiddonBackproject.apply(y, P)


class Smoothness(torch.autograd.Function):
    """sum_k w_k ||D_k x||^2 (D_k: numpy.diff along axis k) with the gradient
    2 sum_k w_k D_k^T D_k x, both from one pass of ``tg_smoothness``.  With a
    per-voxel weight v (the height prior r^2, models.py:90-93) the differences
    are those of v x and the gradient is 2 v sum_k w_k D_k^T D_k (v x)."""

    @staticmethod
    def forward(ctx, x, weights, voxel_weight=None):
        if x.dim() > 4:
            raise ValueError("smoothness takes arrays of up to 4 axes")
        if x.dtype not in (torch.float32, torch.float64):
            raise ValueError("smoothness takes float32 or float64 arrays")
        w = [float(v) for v in weights]
        if len(w) != x.dim():
            raise ValueError("one weight per axis: %d weights for %d axes" % (len(w), x.dim()))
        xc = x.detach().contiguous()
        vw = None
        if voxel_weight is not None:
            vw = torch.as_tensor(voxel_weight).to(xc.device, xc.dtype).contiguous()
            if tuple(vw.shape) != tuple(xc.shape):
                raise ValueError("voxel weight of shape %s for an array of shape %s"
                                 % (tuple(vw.shape), tuple(xc.shape)))
        shape = (ctypes.c_int32 * 4)(*(list(xc.shape) + [1] * (4 - xc.dim())))
        wts = (ctypes.c_double * 4)(*(w + [0.0] * (4 - len(w))))
        grad = torch.empty_like(xc)
        scal = torch.empty(N.TG_CG_SCAL, dtype=torch.float64, device=xc.device)
        code = N.TG_F32 if xc.dtype == torch.float32 else N.TG_F64
        N.check(N.lib().tg_smoothness(code, shape, wts, ctypes.c_void_p(xc.data_ptr()),
                                      ctypes.c_void_p(vw.data_ptr() if vw is not None else None),
                                      ctypes.c_void_p(grad.data_ptr()),
                                      ctypes.c_void_p(scal.data_ptr()), 0, S.stream_handle()))
        ctx.save_for_backward(grad)
        return scal[0].to(x.dtype)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return grad * gout, None, None


def smoothness(x, weights, voxel_weight=None):
    """Differentiable sum_k weights[k] ||D_k (v x)||^2 over the axes of x, with
    v = voxel_weight (the height prior) or 1."""
    return Smoothness.apply(x, tuple(weights), voxel_weight)
