"""Differentiable fermion action and log det M for torch.autograd.

The reference differentiates by hand: ``HubbardFermiAction.force`` is the adjoint of
the action, computed by the backward sweep of ``forceDirectSinglePart``
(src/isle/cpp/action/hubbardFermiAction.cpp:16-60, docs/algorithm/hubbardFermiAction.tex:
288-299).  Here that sweep (the GPU force kernels) is the ``backward`` of a
``torch.autograd.Function`` whose ``forward`` is the action (or one species' log det M)
evaluated on the GPU; no autograd graph is recorded through the kernels.

Derivative conventions.  S(phi) and log det M(phi) are holomorphic in phi, with
    dS/dphi = -force(phi)                         (hubbardFermiAction.hpp:96-142)
    d logdetM_p/dphi = -i (f_p - 1),  d logdetM_h/dphi = +i (f_h - 1)      particle/hole
    d logdetM_p/dphi = -(f_p(aux) - 1), d logdetM_h/dphi = +(f_h(aux) - 1) spin, aux = -i phi
where f_s is forceDirectSinglePart of species s (the terms combined in
hubbardFermiAction.cpp:243-272; the -1 is the derivative of the explicit -+i sum(phi) of
logdetM, hubbardFermiMatrixExp.cpp:594-612, hubbardFermiMatrixDia.cpp:588-593, and cancels
in the action's force).  For an upstream gradient g, backward returns
g * conj(d/dphi) (PyTorch's rule for holomorphic functions); for a real phi it returns the
real part of that, i.e. Re(conj(g) d/dphi), the gradient of the real loss.
"""
import numpy as np
import torch

from . import _lib
from .action import HFABasis, HubbardFermiAction
from .core import as_field


def _grad_in(phi, g, deriv):
    grad = g * deriv.conj()
    return grad if phi.is_complex() else grad.real.to(phi.dtype)


class _ActionFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, action):
        ctx.action = action
        ctx.save_for_backward(phi)
        if phi.dim() == 1:
            s = np.array([action.eval(phi)])
        elif isinstance(action, HubbardFermiAction):
            s = action.eval_batch(phi)
        else:
            s = np.array([action.eval(row) for row in phi])
        out = torch.as_tensor(s, dtype=torch.complex128, device=phi.device)
        return out[0] if phi.dim() == 1 else out

    @staticmethod
    def backward(ctx, g):
        (phi,) = ctx.saved_tensors
        act = ctx.action
        if phi.dim() == 1:
            deriv = -act.force(phi)
        elif isinstance(act, HubbardFermiAction):
            deriv = -act.force_batch(phi)
        else:
            deriv = -torch.stack([act.force(row) for row in phi])
        if phi.dim() > 1:
            g = g.reshape(-1, 1)
        return _grad_in(phi, g, deriv.to(phi.device)), None


def action_value(action, phi):
    """S(phi) (complex) for a field [nx*nt] or a batch [b, nx*nt], real or complex; its
    backward runs ``action.force`` (the adjoint sweep)."""
    return _ActionFunction.apply(phi, action)


class _LogdetMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, hfa, species):
        ctx.hfa = hfa
        ctx.species = int(species)
        ctx.save_for_backward(phi)
        ld = hfa.logdets(phi)  # [batch][nspecies]; particle = 0, hole = 1
        col = 0 if hfa.shortcut else ctx.species
        vals = ld[:, col]
        if hfa.shortcut and ctx.species == 1:
            vals = np.conj(vals)
        out = torch.as_tensor(vals, dtype=torch.complex128, device=phi.device)
        return out[0] if phi.dim() == 1 else out

    @staticmethod
    def backward(ctx, g):
        (phi,) = ctx.saved_tensors
        hfa, species = ctx.hfa, ctx.species
        parts = hfa.force_parts(phi)  # [batch][nspecies][nx*nt]
        if hfa.shortcut:
            f = parts[:, 0] if species == 0 else parts[:, 0].conj()
        else:
            f = parts[:, species]
        if hfa.basis == HFABasis.SPIN:
            coef = -1.0 if species == 0 else 1.0
        else:
            coef = -1j if species == 0 else 1j
        deriv = coef * (f - 1.0)
        if phi.dim() == 1:
            deriv = deriv[0]
        else:
            g = g.reshape(-1, 1)
        return _grad_in(phi, g, deriv), None, None


def logdetM_value(hfa, phi, species):
    """log det M of one species (0 particle, 1 hole; aux = -i phi in the spin basis) as a
    differentiable tensor, for a HubbardFermiAction ``hfa``; backward runs the force
    kernels of that species."""
    return _LogdetMFunction.apply(phi, hfa, species)
