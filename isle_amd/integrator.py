"""Leapfrog integrator on the GPU (src/isle/cpp/integrator.cpp:9-36).

``leapfrog(phi, pi, action, length, nsteps, direction=1)`` keeps the reference's
semantics: eps = direction*length/nsteps; half kick with Re(force); nsteps drifts with
nsteps-1 full kicks in between; final half kick; returns (phi, pi, action.eval(phi)).
Kicks and drifts are streaming kernels; when the action is a SumAction of a
HubbardGaugeAction and a HubbardFermiAction the gauge force -phi/utilde is fused into
the kick.  ``leapfrog_batch`` advances a [batch, nx*nt] stack of independent chains.

``rungeKutta4Flow`` integrates the holomorphic flow d phi / dt = conj(dS / dphi)
(= -conj(force)) with adaptive RK4 steps (src/isle/cpp/integrator.cpp:39-195).
"""
import cmath
import math

import torch

from . import _lib
from .action import HFAAlgorithm, HubbardFermiAction, HubbardGaugeAction, SumAction
from .core import as_field


def _split_fused(action):
    """(fermion action, utilde) if ``action`` is gauge + fermion, else None."""
    if isinstance(action, SumAction) and len(action) == 2:
        a, b = action[0], action[1]
        if isinstance(a, HubbardGaugeAction) and isinstance(b, HubbardFermiAction):
            return b, a.utilde
        if isinstance(b, HubbardGaugeAction) and isinstance(a, HubbardFermiAction):
            return a, b.utilde
    if isinstance(action, HubbardFermiAction):
        return action, 0.0
    return None


def _kick(pi, force, phi, eps, utilde):
    _lib.check(_lib.load().isle_leapfrog_kick(_lib.ptr(pi), _lib.ptr(force) if force is not None else None,
                                              _lib.ptr(phi) if utilde > 0 else None, pi.numel(), eps, utilde,
                                              _lib.stream_ptr()))


def _drift(phi, pi, eps):
    _lib.check(_lib.load().isle_leapfrog_drift(_lib.ptr(phi), _lib.ptr(pi), phi.numel(), eps, _lib.stream_ptr()))


def leapfrog_batch(phis, pis, action, length, nsteps, direction=1.0, evaluate=True):
    """Leapfrog for a batch of chains ([batch, nx*nt] complex128, device).

    Returns (phis', pis', actVals) with actVals a numpy array (None if not evaluate)."""
    if nsteps < 1:
        raise ValueError("nsteps must be at least 1")
    phi = as_field(phis).clone()
    pi = as_field(pis).clone()
    if phi.dim() == 1:
        phi = phi.reshape(1, -1)
        pi = pi.reshape(1, -1)
    eps = direction * length / float(nsteps)
    fused = _split_fused(action)
    if fused is not None:
        ferm, utilde = fused
        fbuf = torch.empty_like(phi)

        def kick(scale):
            ferm.force_batch(phi, out=fbuf)
            _kick(pi, fbuf, phi, eps * scale, utilde)
    else:
        def kick(scale):
            f = torch.stack([as_field(action.force(phi[b])) for b in range(phi.shape[0])])
            _kick(pi, f.contiguous(), phi, eps * scale, 0.0)

    kick(0.5)
    _drift(phi, pi, eps)
    for _ in range(nsteps - 1):
        kick(1.0)
        _drift(phi, pi, eps)
    kick(0.5)
    acts = eval_batch(action, phi) if evaluate else None
    return phi, pi, acts


def leapfrog_multi(phis, pis, fermion, utildes, length, nsteps, direction=1.0):
    """Leapfrog for several batches of chains on one lattice with different time extents:
    lists of [batch_i, nx*nt_i] fields and momenta, one HubbardFermiAction for all of them,
    the gauge action's utilde of each batch.  Every kick evaluates the fermion force of all
    batches in shared launches (HubbardFermiAction.force_batch_multi) and adds each batch's
    gauge force in its kick kernel.  Returns the lists (phis', pis')."""
    if nsteps < 1:
        raise ValueError("nsteps must be at least 1")
    phi = [as_field(p).clone().reshape(-1, p.shape[-1]) for p in phis]
    pi = [as_field(p).clone().reshape(-1, p.shape[-1]) for p in pis]
    eps = direction * length / float(nsteps)
    fbufs = [torch.empty_like(p) for p in phi]

    def kick(scale):
        fermion.force_batch_multi(phi, fbufs)
        for p, q, f, ut in zip(phi, pi, fbufs, utildes):
            _kick(q, f, p, eps * scale, ut)

    def drift():
        for p, q in zip(phi, pi):
            _drift(p, q, eps)

    kick(0.5)
    drift()
    for _ in range(nsteps - 1):
        kick(1.0)
        drift()
    kick(0.5)
    return phi, pi


def eval_batch(action, phi):
    """Action of each field of a [batch, nx*nt] stack as a numpy complex array; gauge +
    fermion sums are evaluated with the batched kernels."""
    import numpy as np
    fused = _split_fused(action)
    if fused is None:
        return np.array([complex(action.eval(phi[b])) for b in range(phi.shape[0])])
    ferm, utilde = fused
    acts = ferm.eval_batch(phi)
    if utilde > 0:
        sums = np.empty((phi.shape[0], 4))
        n = phi.shape[1]
        _lib.check(_lib.load().isle_field_sums(_lib.ptr(phi), n, n, phi.shape[0], _lib.dptr(sums),
                                               _lib.stream_ptr()))
        acts = acts + (sums[:, 2] + 1j * sums[:, 3]) / 2.0 / utilde
    return acts


def leapfrog(phi, pi, action, length, nsteps, direction=1):
    """Single-chain leapfrog with the reference's signature (bind_integrator.cpp:10-11)."""
    phi = as_field(phi)
    pi = as_field(pi)
    p1, q1, acts = leapfrog_batch(phi.reshape(1, -1), pi.reshape(1, -1), action, length, nsteps, direction)
    return p1[0], q1[0], complex(acts[0])


class LeapfrogGraph:
    """A leapfrog trajectory of a fixed [batch, nx*nt] stack of chains captured once in a
    HIP graph (torch.cuda.CUDAGraph) and replayed: the 2*nsteps kicks/drifts and the
    nsteps+1 batched force evaluations become one graph launch.  The action's plans are
    built and tuned by an eager warm-up before the capture; the force path then makes no
    host synchronisation, allocation or memset.  The action value is not part of the
    graph (its log-determinants are read back to the host).

    ``split``: optional list of k actions equal to ``action`` but with workspaces of their
    own.  The batch is then cut into k contiguous parts, part i integrated with split[i] on a
    stream of its own inside the graph, so that the latency-bound phases of one part (the
    LU and solve of the top level, the bandwidth-bound diagonals) overlap the MFMA-bound
    chain products of the others.  The graph keeps references to every action it captured,
    whose plans must not be destroyed while it can be replayed."""

    def __init__(self, action, batch, n, length, nsteps, direction=1.0, split=None):
        fused = _split_fused(action)
        if fused is not None and fused[0].algorithm == HFAAlgorithm.DIRECT_SQUARE:
            raise ValueError("HFAAlgorithm.DIRECT_SQUARE (dense LU with its own allocations) cannot be captured "
                             "in a graph; integrate it eagerly")
        dev = _lib.device()
        self.phi = torch.zeros((batch, n), dtype=torch.complex128, device=dev)
        self.pi = torch.zeros_like(self.phi)
        actions = list(split) if split else [action]
        # the graph holds device pointers into these actions' plans (descriptor arrays and
        # workspaces): keep the actions alive as long as the graph can be replayed
        self._actions = actions
        k = len(actions)
        if k > batch:
            raise ValueError("more parts than chains")
        bounds = [round(i * batch / k) for i in range(k + 1)]
        parts = [(actions[i], bounds[i], bounds[i + 1]) for i in range(k)]
        main = torch.cuda.current_stream()
        streams = [torch.cuda.Stream() for _ in parts]
        for s, (act, lo, hi) in zip(streams, parts):  # warm-up: builds and tunes every plan
            s.wait_stream(main)
            with torch.cuda.stream(s):
                leapfrog_batch(self.phi[lo:hi], self.pi[lo:hi], act, length, nsteps, direction, evaluate=False)
            main.wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            if k == 1:
                self.out_phi, self.out_pi, _ = leapfrog_batch(self.phi, self.pi, action, length, nsteps, direction,
                                                              evaluate=False)
            else:
                cap = torch.cuda.current_stream()
                outs = []
                for s, (act, lo, hi) in zip(streams, parts):
                    s.wait_stream(cap)
                    with torch.cuda.stream(s):
                        outs.append(leapfrog_batch(self.phi[lo:hi], self.pi[lo:hi], act, length, nsteps, direction,
                                                   evaluate=False))
                for s in streams:
                    cap.wait_stream(s)
                self.out_phi = torch.cat([o[0] for o in outs])
                self.out_pi = torch.cat([o[1] for o in outs])
        self._streams = streams

    def __call__(self, phis, pis):
        """Returns (phis', pis') as new tensors."""
        self.phi.copy_(as_field(phis).reshape(self.phi.shape))
        self.pi.copy_(as_field(pis).reshape(self.pi.shape))
        self.graph.replay()
        return self.out_phi.clone(), self.out_pi.clone()


# classic RK4 (n = 0) and the 3/8 rule (n = 1) as (weights b1..b4, stage coefficients
# beta21, beta31, beta32, beta41, beta42, beta43), integrator.cpp:39-69.  "reference" keeps
# the reference's beta21 = 1/30 in the 3/8 rule; "textbook" has the rule's 1/3.
_RK4_TABLES = {
    "reference": {0: ((1 / 6, 1 / 3, 1 / 3, 1 / 6), (0.5, 0.0, 0.5, 0.0, 0.0, 1.0)),
                  1: ((1 / 8, 3 / 8, 3 / 8, 1 / 8), (1 / 30, -1 / 3, 1.0, 1.0, -1.0, 1.0))},
    "textbook": {0: ((1 / 6, 1 / 3, 1 / 3, 1 / 6), (0.5, 0.0, 0.5, 0.0, 0.0, 1.0)),
                 1: ((1 / 8, 3 / 8, 3 / 8, 1 / 8), (1 / 3, -1 / 3, 1.0, 1.0, -1.0, 1.0))},
}


def _rk4_step(phi, action, eps, direction, n, variant):
    w, b = _RK4_TABLES[variant][n]
    e = eps * direction
    # the reference's fourth stage takes -eps, not -eps * direction (integrator.cpp:96)
    e4 = eps if variant == "reference" else e

    def k(scale, arg):
        return (-scale) * torch.conj(as_field(action.force(arg)))
    k1 = k(e, phi)
    k2 = k(e, phi + b[0] * k1)
    k3 = k(e, phi + b[1] * k1 + b[2] * k2)
    k4 = k(e4, phi + b[3] * k1 + b[4] * k2 + b[5] * k3)
    out = phi + w[0] * k1 + w[1] * k2 + w[2] * k3 + w[3] * k4
    return out, complex(action.eval(out))


def rungeKutta4Flow(phi, action, flowTime, stepSize, actVal=complex(math.nan, math.nan), n=0, direction=+1,
                    adaptAttenuation=0.9, adaptThreshold=1.0e-8, minStepSize=math.nan, imActTolerance=0.001,
                    *, variant="reference"):
    """Adaptive RK4 integration of the holomorphic flow for ``flowTime``
    (integrator.hpp:45-84, integrator.cpp:144-193).  Returns (phi, action value, flow time
    reached).

    A step is kept when |e^{i (Im S_before - Im S_after)} - 1| <= imActTolerance; after a
    rejected step the size becomes max(step adaptAttenuation (tol/err)^(1/5), minStepSize)
    (the flow stops if it already was minStepSize), after a kept step with
    err < adaptThreshold tol it grows by min(max(adaptAttenuation (tol/err)^(1/4), 1), 2).
    The last step is shortened to end at flowTime, or skipped if shorter than minStepSize
    (default max(stepSize / 1000, 1e-12)).  n = 0: classic RK4, n = 1: the 3/8 rule.

    ``variant="reference"`` (default) gives the reference's results: its fourth stage
    drops ``direction`` (-epsilon where the other stages use -epsilon direction) and its
    3/8 rule has beta21 = 1/30, which leaves that rule first order.  ``variant="textbook"``
    keeps the direction in every stage and uses beta21 = 1/3 (fourth order for both n and
    both directions).  The two agree for n = 0 and direction = +1."""
    if variant not in _RK4_TABLES:
        raise ValueError(f"variant must be one of {sorted(_RK4_TABLES)}, not {variant!r}")
    if n not in (0, 1):
        raise ValueError("n must be 0 or 1")
    phi = as_field(phi).clone()
    if cmath.isnan(actVal):
        actVal = complex(action.eval(phi))
    if math.isnan(minStepSize):
        minStepSize = max(stepSize / 1000.0, 1e-12)
    t = 0.0
    while t < flowTime:
        if t + stepSize > flowTime:
            stepSize = flowTime - t
            if stepSize < minStepSize:
                break
        new_phi, new_act = _rk4_step(phi, action, stepSize, direction, n, variant)
        err = abs(cmath.exp(1j * (actVal.imag - new_act.imag)) - 1.0)
        if err > imActTolerance:
            if stepSize == minStepSize:
                break
            stepSize = max(stepSize * adaptAttenuation * (imActTolerance / err) ** 0.2, minStepSize)
        else:
            t += stepSize
            phi, actVal = new_phi, new_act
            if err < adaptThreshold * imActTolerance:
                grow = 2.0 if err == 0 else min(max(adaptAttenuation * (imActTolerance / err) ** 0.25, 1.0), 2.0)
                stepSize *= grow
    return phi, actVal, t
