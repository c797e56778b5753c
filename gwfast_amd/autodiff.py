"""Differentiable strain: the forward-mode tangent kernel behind a
torch.autograd.Function.

``strain(sig, rows, f, arm)`` returns the detector strain h (N_f, N_ev, 2)
(real, imaginary) of one arm as a differentiable function of the parameter
rows (P, N_ev) in Fisher order.  The forward pass runs gwf_strain, which
emits h and all P derivative rows dh/dtheta in one pass (the Jacobian the
reference builds with vmap(jacrev), gwfast/signal.py:1174-1189, plus its
closed-form rows, 1303-1584); the backward pass contracts the saved Jacobian
with the incoming gradient in gwf_strain_vjp.  The tcoal row of the saved
Jacobian is per second (signal.py:919-920), so the VJP rescales it to the
input's unit (GMST days).
"""
import torch

from . import _engine as E
from . import _native as N
from . import signal as S


class _Strain(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rows, f, sig, arm):
        h, J = S.strain_device(sig, rows.detach().contiguous(), f.contiguous(), arm)
        ctx.save_for_backward(J)
        ctx.P = J.shape[0]
        return h

    @staticmethod
    def backward(ctx, gh):
        (J,) = ctx.saved_tensors
        P, n, nf = J.shape[0], J.shape[1], J.shape[2]
        g = torch.empty((P, n), dtype=torch.float64, device=J.device)
        E.call('gwf_strain_vjp', N.ptr(J), N.ptr(gh.contiguous()), P, n, nf, 86400.0, N.ptr(g), N.stream_ptr())
        return g, None, None, None


def strain(sig, rows, f, arm=0):
    """Differentiable h (N_f, N_ev, 2) of detector ``sig``'s arm ``arm`` on the
    grid f (N_f, N_ev) w.r.t. ``rows`` (P, N_ev, float64)."""
    return _Strain.apply(rows, f, sig, arm)
