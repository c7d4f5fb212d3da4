"""The projector as a differentiable torch operation.

``project(x, P)`` returns P x; its hand-written backward is the adjoint,
the (masked, time-binned, Thomson-weighted -- whatever P carries) Siddon
backprojector: grad_x = P^T grad_y.  P is a ``lo_wrapper.Siddon`` or
``Siddon4d`` operator; the output is a fresh tensor (not P's cached buffer),
so several projections can live in one autograd graph.
"""
import torch

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
        gx = torch.zeros(P.shapein, dtype=P.tdtype, device=P.dev)
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
        x = torch.zeros(P.shapein, dtype=P.tdtype, device=P.dev)
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
