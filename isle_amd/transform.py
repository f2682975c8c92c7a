"""Maps between the manifold the molecular dynamics runs on and the one Metropolis
samples (src/isle/evolver/transform/transform.py, identity.py, constantShift.py).

A transform has ``forward(phi, actVal) -> (phi', actVal', logdetJ)`` from the MD field to
the Monte Carlo field, and ``backward(phi, jacobian=False) -> (phi', logdetJ or None)``
the other way, ``logdetJ`` being log det of the forward map's Jacobian.  The leapfrog
evolvers call ``pull_back`` before integrating and ``push_forward`` after it; both accept
``None`` for "no transform", which costs nothing and contributes logdetJ = 0.

Fields may be numpy vectors or device tensors: the transforms here only add and subtract.
"""
import numpy as np
import torch


class Transform:
    """Interface of an evolver transform; subclasses provide forward, backward, save and
    fromH5 (the checkpoint pair used by runfile.EvolverManager)."""

    def forward(self, phi, actVal):
        raise NotImplementedError

    def backward(self, phi, jacobian=False):
        raise NotImplementedError

    def save(self, group, manager):
        raise NotImplementedError

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        raise NotImplementedError


class Identity(Transform):
    """Leaves the field and the action value alone; log det J = 0."""

    def forward(self, phi, actVal):
        return phi, actVal, 0

    def backward(self, phi, jacobian=False):
        return phi, (0 if jacobian else None)

    def save(self, group, manager):
        return None

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls()


class ConstantShift(Transform):
    """The Monte Carlo field is the MD field plus a fixed shift; the action value carried
    along is the action of the MD field, S(phi_MC - shift) (constantShift.py:39-58), and
    log det J = 0.  ``shift`` is a vector of the lattice size or a scalar with ``lattSize``."""

    def __init__(self, shift, action, lattSize=None):
        if np.ndim(shift) == 0 and not isinstance(shift, torch.Tensor):
            if lattSize is None:
                raise ValueError("a scalar shift needs lattSize, the number of field entries")
            vec = np.full(int(lattSize), complex(shift))
        else:
            vec = shift.detach().cpu().numpy() if isinstance(shift, torch.Tensor) else np.asarray(shift)
        self.shift = np.asarray(vec, dtype=np.complex128).reshape(-1)
        self.action = action
        self._device_shift = {}

    def _like(self, phi):
        """The shift in the array kind of ``phi``."""
        if isinstance(phi, torch.Tensor):
            key = phi.device
            if key not in self._device_shift:
                self._device_shift[key] = torch.as_tensor(self.shift, device=phi.device)
            return self._device_shift[key]
        return self.shift

    def forward(self, phi, actVal):
        return phi + self._like(phi), self.action.eval(phi), 0

    def backward(self, phi, jacobian=False):
        return phi - self._like(phi), (0 if jacobian else None)

    def save(self, group, manager):
        group["shift"] = self.shift

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(group["shift"][()], action)


def pull_back(transform, stage):
    """(MD field, log det J of the stage's field) before a trajectory
    (transform.py:73-90): the stored logdetJ is reused when the stage carries one."""
    if transform is None:
        return stage.phi, 0
    known = stage.logWeights.get("logdetJ")
    if known is not None:
        return transform.backward(stage.phi, False)[0], known
    return transform.backward(stage.phi, True)


def push_forward(transform, phi, actVal):
    """(Monte Carlo field, its action value, log det J) after a trajectory."""
    if transform is None:
        return phi, actVal, 0
    return transform.forward(phi, actVal)
