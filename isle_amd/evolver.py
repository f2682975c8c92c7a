"""HMC evolution: EvolutionStage, BinarySelector, ConstStepLeapfrog, NumpyRNG and the
trajectory loop, with the reference's semantics and random-number consumption.

* ``NumpyRNG`` wraps numpy's MT19937 RandomState (src/isle/random.py:80-185), its state
  stored in a run-file group by ``writeH5`` / ``readH5`` / ``readStateH5``;
* ``EvolutionStage`` (src/isle/evolver/stage.py:10-117) with ``save`` / ``fromH5``;
* ``BinarySelector.selectTrajPoint``: accept iff dE < 0 or exp(-dE) > uniform, the
  uniform drawn only when dE >= 0 (src/isle/evolver/selector.py:20-30);
* ``ConstStepLeapfrog.evolve``: pi = rng.normal(0, 1, N) + 0j, leapfrog, Metropolis with
  E = sum(logWeights) + |pi|^2/2 (src/isle/evolver/leapfrog.py:38-66);
* ``LinearStepLeapfrog``: the same with length and step count moving linearly between
  two values over ``ninterp`` trajectories (leapfrog.py:105-229); both evolvers store
  their state in checkpoints (``save`` / ``fromH5``);
* ``hmc``: the HMC loop of src/isle/drivers/hmc.py:41-84 without HDF5 output; a start
  configuration that is not a stage is evaluated first (hmc.py:64-66).

``TwoPiJumps`` and ``UniformJump`` (src/isle/evolver/hubbard.py:15-184) shift the field by
2 pi at random sites or by 2 pi / Nt everywhere; the particle-hole fermion action is
invariant under both, so with only gauge and such fermion actions the new action value
is the old one plus the change of the gauge action (no evaluation on the GPU).

``BatchedConstStepLeapfrog`` advances many independent chains together (one field tensor
[batch, nx*nt]); each chain has its own NumpyRNG, consumed exactly as a lone chain would.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from .core import as_field
from .integrator import LeapfrogGraph, eval_batch as _eval_batch, leapfrog_batch


class NumpyRNG:
    """MT19937 via numpy.random.RandomState (src/isle/random.py:80-124)."""

    NAME = "np.MT19937"

    def __init__(self, seed):
        self._state = np.random.RandomState(seed)

    def seed(self, seed):
        self._state.seed(seed)

    def uniform(self, low=0.0, high=1.0, size=None, cmplx=False):
        if cmplx:
            return self._state.uniform(low, high, size) + 1j * self._state.uniform(low, high, size)
        return self._state.uniform(low, high, size)

    def normal(self, mean=0.0, std=1.0, size=None, cmplx=False):
        if cmplx:
            return self._state.normal(mean, std, size) + 1j * self._state.normal(mean, std, size)
        return self._state.normal(mean, std, size)

    def choice(self, a, size=None, replace=True, p=None):
        return self._state.choice(a, size, replace, p)

    def get_state(self):
        return self._state.get_state()

    def set_state(self, state):
        self._state.set_state(state)

    def writeH5(self, group):
        """Store the generator state in a run-file group (random.py:140-154)."""
        kind, keys, pos, has_gauss, cached_gaussian = self._state.get_state()
        assert kind == self.NAME.rsplit(".")[-1]
        group["name"] = self.NAME
        group["keys"] = keys
        group.attrs["pos"] = pos
        group.attrs["has_gauss"] = has_gauss
        group.attrs["cached_gaussian"] = cached_gaussian

    def readH5(self, group):
        """Restore the generator state written by writeH5 (random.py:156-166)."""
        name = str(group["name"][()])
        if name != self.NAME:
            raise RuntimeError(f"Wrong kind of RNG. Expected {self.NAME}, got {name}")
        self._state.set_state((self.NAME.rsplit(".")[-1], np.asarray(group["keys"][()], dtype=np.uint32),
                               int(group.attrs["pos"]), int(group.attrs["has_gauss"]),
                               float(group.attrs["cached_gaussian"])))


def readStateH5(group):
    """A generator with the state stored in a run-file group (random.py:168-185)."""
    name = str(group["name"][()])
    if name != NumpyRNG.NAME:
        raise ValueError(f"Unknown RNG in file: {name}")
    rng = NumpyRNG(0)
    rng.readH5(group)
    return rng


def _host(x):
    """A field as a complex128 numpy vector."""
    if isinstance(x, torch.Tensor):
        return x.detach().to("cpu", torch.complex128).numpy()
    return np.asarray(x, dtype=np.complex128)


class EvolutionStage:
    """Configuration + log weights at one point of the Markov chain (stage.py:10-68)."""

    __slots__ = "phi", "trajPoint", "logWeights", "extra", "phi_RTP"

    def __init__(self, phi, actVal, trajPoint=1, logWeights=None, extra=None, phi_RTP=None):
        self.phi = phi
        self.phi_RTP = phi_RTP
        self.trajPoint = trajPoint
        self.logWeights = logWeights if logWeights is not None else dict()
        self.logWeights["actVal"] = actVal
        self.extra = extra if extra is not None else dict()

    @property
    def actVal(self):
        return self.logWeights["actVal"]

    @actVal.setter
    def actVal(self, actVal):
        self.logWeights["actVal"] = actVal

    def accept(self, phi, actVal, logWeights=None, extra=None, phi_RTP=None):
        return self.__class__(phi, actVal, 1, logWeights, extra, phi_RTP)

    def reject(self, extra=None):
        return self.__class__(self.phi, self.actVal, 0, self.logWeights, extra, self.phi_RTP)

    def sumLogWeights(self):
        return sum(self.logWeights.values())

    def save(self, group):
        """Store phi, actVal, trajPoint and any further log weights and extras in a
        run-file group (stage.py:74-93)."""
        group["phi"] = _host(self.phi)
        if self.phi_RTP is not None:
            group["phi_RT-P"] = _host(self.phi_RTP)
        group["actVal"] = complex(self.actVal)
        group["trajPoint"] = int(self.trajPoint)
        if len(self.logWeights) > 1:
            weights = group.create_group("logWeights")
            for key, val in self.logWeights.items():
                if key != "actVal":
                    weights[key] = complex(val)
        if self.extra:
            extra = group.create_group("extra")
            for key, val in self.extra.items():
                extra[key] = val

    @classmethod
    def fromH5(cls, group):
        """The stage stored by save; phi comes back as a numpy vector (stage.py:95-117)."""
        extra = {k: v[()] for k, v in group["extra"].items()} if "extra" in group else None
        weights = {k: complex(v[()]) for k, v in group["logWeights"].items()} if "logWeights" in group else None
        phi_rtp = group["phi_RT-P"][()] if "phi_RT-P" in group else None
        return cls(group["phi"][()], complex(group["actVal"][()]), int(group["trajPoint"][()]), weights, extra,
                   phi_rtp)


class BinarySelector:
    """Metropolis accept/reject (selector.py:9-44)."""

    def __init__(self, rng):
        self.rng = rng

    def selectTrajPoint(self, energy0, energy1):
        deltaE = np.real(energy1 - energy0)
        return 1 if deltaE < 0 or np.exp(-deltaE) > self.rng.uniform(0, 1) else 0

    def selectTrajectory(self, energy0, data0, energy1, data1):
        return (energy1, data1, 1) if self.selectTrajPoint(energy0, energy1) == 1 else (energy0, data0, 0)


class ConstStepLeapfrog:
    """Leapfrog evolver with constant length and step count (leapfrog.py:15-103)."""

    def __init__(self, action, length, nstep, rng, transform=None):
        if transform is not None:
            raise NotImplementedError("transforms are not supported")
        self.action = action
        self.length = length
        self.nstep = nstep
        self.rng = rng
        self.selector = BinarySelector(rng)
        self.transform = None
        self.trajPoints = []
        self.last_energies = None

    def evolve(self, stage):
        return _leapfrog_trajectory(self, stage, self.length, self.nstep)

    def save(self, group, manager):
        """Store length and nstep in a checkpoint (leapfrog.py:68-77)."""
        group["length"] = self.length
        group["nstep"] = self.nstep

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, float(group["length"][()]), int(group["nstep"][()]), rng)

    def report(self):
        return f"""<ConstStepLeapfrog> (0x{id(self):x})
  length = {self.length}, nstep = {self.nstep}
  acceptance rate = {np.mean(self.trajPoints)}"""


def _leapfrog_trajectory(ev, stage, length, nstep):
    """One HMC trajectory of evolver ``ev``: momenta from ev.rng, leapfrog, Metropolis."""
    phi = as_field(stage.phi)
    n = phi.numel()
    pi_host = ev.rng.normal(0, 1, n) + 0j
    pi = torch.as_tensor(pi_host, device=phi.device)
    phi1, pi1, acts = leapfrog_batch(phi.reshape(1, -1), pi.reshape(1, -1), ev.action, length, nstep)
    actVal1 = complex(acts[0])
    energy0 = stage.sumLogWeights() + np.linalg.norm(pi_host) ** 2 / 2
    energy1 = actVal1 + 0 + np.linalg.norm(pi1.cpu().numpy()) ** 2 / 2
    ev.last_energies = (energy0, energy1)
    trajPoint = ev.selector.selectTrajPoint(energy0, energy1)
    ev.trajPoints.append(trajPoint)
    return stage.accept(phi1[0], actVal1) if trajPoint == 1 else stage.reject()


def _hinge_range(start, end, step):
    """start, start + step, ... while short of end (the sum accumulated step by step),
    then end for ever (src/isle/collection.py:10-30)."""
    cur = start
    while (step > 0 and cur < end) or (step < 0 and cur > end):
        yield cur
        cur += step
    while True:
        yield end


class LinearStepLeapfrog:
    """Leapfrog whose trajectory length and step count move linearly from the first to the
    second value of their ranges over ``ninterp`` trajectories and then stay at the second
    (leapfrog.py:105-229).  ``startPoint`` skips that many trajectories of the schedule.

    The position in the schedule that a checkpoint stores counts every trajectory since
    the schedule began, including those before a resumption (the reference restarts that
    count at zero when it resumes, so a second resumption would rewind the schedule)."""

    def __init__(self, action, lengthRange, nstepRange, ninterp, rng, startPoint=0, transform=None):
        if transform is not None:
            raise NotImplementedError("transforms are not supported")
        self.action = action
        self.lengthRange = tuple(lengthRange)
        self.nstepRange = tuple(nstepRange)
        self.ninterp = ninterp
        self.rng = rng
        self.selector = BinarySelector(rng)
        self.transform = None
        self.trajPoints = []
        self.last_energies = None
        self._lengths = _hinge_range(*self.lengthRange, (self.lengthRange[1] - self.lengthRange[0]) / ninterp)
        self._nsteps = _hinge_range(*self.nstepRange, (self.nstepRange[1] - self.nstepRange[0]) / ninterp)
        self._current = 0
        for _ in range(int(startPoint)):
            next(self._lengths)
            next(self._nsteps)
            self._current += 1

    def evolve(self, stage):
        self._current += 1
        return _leapfrog_trajectory(self, stage, next(self._lengths), int(next(self._nsteps)))

    def save(self, group, manager):
        """Store the ranges and the position in the schedule (leapfrog.py:183-196)."""
        group["minLength"] = self.lengthRange[0]
        group["maxLength"] = self.lengthRange[1]
        group["minNstep"] = self.nstepRange[0]
        group["maxNstep"] = self.nstepRange[1]
        group["ninterp"] = self.ninterp
        group["current"] = self._current

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, (group["minLength"][()], group["maxLength"][()]),
                   (group["minNstep"][()], group["maxNstep"][()]), int(group["ninterp"][()]), rng,
                   startPoint=int(group["current"][()]))

    def report(self):
        return f"""<LinearStepLeapfrog> (0x{id(self):x})
  lengthRange = {self.lengthRange}, nstepRange = {self.nstepRange}, ninterp = {self.ninterp}
  acceptance rate = {np.mean(self.trajPoints)}"""


class _HubbardActionShortcut:
    """Action values after jumps that leave the fermion determinant invariant
    (src/isle/evolver/hubbard.py:187-287).

    The shortcut applies when the action is a sum of HubbardGaugeActions and particle-hole
    basis HubbardFermiActions: the fermion matrices hold the field only through e^{+-i phi},
    so a shift by 2 pi at any site, or by 2 pi / Nt at every site (the product of the Nt
    slices gains e^{+-2 pi i}), leaves them unchanged, and only the gauge action
    sum phi^2 / (2 Utilde) moves.  In the spin basis the field enters as e^{+-phi}: no
    shortcut, the action is evaluated.  (The reference names fermion-action classes that
    do not exist here, so its shortcut never applies to a fermion action, and its global
    update divides by the sum of 1 / (2 Utilde) where the local one multiplies.)"""

    def __init__(self, action):
        utildes = self._find_utildes(action)
        if utildes is None:
            self._action = action
            self._half_inv_u = None
        else:
            self._action = None
            self._half_inv_u = sum(1.0 / u for u in utildes) / 2

    @classmethod
    def _find_utildes(cls, action):
        from .action import HFABasis, HubbardFermiAction, HubbardGaugeAction, SumAction
        if isinstance(action, SumAction):
            found = []
            for sub in action:
                aux = cls._find_utildes(sub)
                if aux is None:
                    return None
                found.extend(aux)
            return found
        if isinstance(action, HubbardGaugeAction):
            return [action.utilde]
        if isinstance(action, HubbardFermiAction) and action.basis == HFABasis.PARTICLE_HOLE:
            return []
        return None

    @property
    def shortcut(self):
        return self._action is None

    def evalLocal(self, newPhi, sites, shifts, actVal):
        """Action at newPhi = old phi shifted by ``shifts`` at ``sites``."""
        if self._action is None:
            return actVal + np.dot(2 * newPhi[sites] - shifts, shifts) * self._half_inv_u
        return complex(self._action.eval(newPhi))

    def evalGlobal(self, newPhi, shift, actVal):
        """Action at newPhi = old phi + shift at every site."""
        if self._action is None:
            return actVal + shift * (2 * np.sum(newPhi) - shift * len(newPhi)) * self._half_inv_u
        return complex(self._action.eval(newPhi))


class TwoPiJumps:
    """Shift the field by +-2 pi at random sites, momenta untouched (hubbard.py:15-116).
    Per batch: ``batchSize`` distinct sites rng.choice(N, batchSize, replace=False), signs
    rng.choice((-2 pi, 2 pi), batchSize), then accept / reject of that batch alone.  The
    returned stage holds the field after the last accepted batch (the reference returns the
    last batch's proposal, accepted or not)."""

    def __init__(self, nBatches, batchSize, action, lattice, rng):
        self.nBatches = int(nBatches)
        self.batchSize = int(batchSize)
        self.rng = rng
        self.latSize = lattice.lattSize()
        self.selector = BinarySelector(rng)
        self.acceptedPerBatch = []
        self._action = _HubbardActionShortcut(action)

    def evolve(self, stage):
        phi = _host(stage.phi)
        actVal = complex(stage.actVal)
        accepted = []
        for _ in range(self.nBatches):
            sites = self.rng.choice(self.latSize, self.batchSize, replace=False)
            shifts = self.rng.choice((-2 * np.pi, +2 * np.pi), self.batchSize, replace=True)
            newPhi = phi.copy()
            newPhi[sites] += shifts
            newActVal = self._action.evalLocal(newPhi, sites, shifts, actVal)
            if self.selector.selectTrajPoint(actVal, newActVal) == 1:
                phi, actVal = newPhi, newActVal
                accepted.append(1)
            else:
                accepted.append(0)
        self.acceptedPerBatch.append(accepted)
        return stage.accept(phi, actVal) if any(accepted) else stage.reject()

    def save(self, group, manager):
        group["nBatches"] = self.nBatches
        group["batchSize"] = self.batchSize

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(int(group["nBatches"][()]), int(group["batchSize"][()]), action, lattice, rng)

    def report(self):
        return f"""<TwoPiJumps> (0x{id(self):x})
  nBatches = {self.nBatches}, batchSize = {self.batchSize}
  acceptance rate per batch = {np.mean(self.acceptedPerBatch)}"""


class UniformJump:
    """Shift the field by +-2 pi / Nt at every site, momenta untouched (hubbard.py:119-184):
    sign rng.choice((-1, 1)), then accept / reject."""

    def __init__(self, action, lattice, rng):
        self.rng = rng
        self.selector = BinarySelector(rng)
        self.accepted = []
        self._action = _HubbardActionShortcut(action)
        self._absShift = 2 * np.pi / lattice.nt()

    def evolve(self, stage):
        shift = self.rng.choice((-1, +1)) * self._absShift
        newPhi = _host(stage.phi) + shift
        newActVal = self._action.evalGlobal(newPhi, shift, complex(stage.actVal))
        if self.selector.selectTrajPoint(stage.actVal, newActVal) == 1:
            self.accepted.append(1)
            return stage.accept(newPhi, newActVal)
        self.accepted.append(0)
        return stage.reject()

    def save(self, group, manager):
        pass

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, lattice, rng)

    def report(self):
        return f"""<UniformJump> (0x{id(self):x})
  shift = 2 pi / Nt = {self._absShift}
  acceptance rate = {np.mean(self.accepted)}"""


class BatchedConstStepLeapfrog:
    """ConstStepLeapfrog over independent chains advanced together on the GPU.

    ``rngs`` holds one NumpyRNG per chain; chain b draws its momenta and its Metropolis
    uniform from rngs[b] exactly as ConstStepLeapfrog would, so each chain reproduces a
    lone run with the same seed (up to the arithmetic of the batched kernels)."""

    def __init__(self, action, length, nstep, rngs, graph=False, split=None):
        self.action = action
        self.split = split  # LeapfrogGraph's split: actions with workspaces of their own
        self.length = length
        self.nstep = nstep
        self.rngs = list(rngs)
        self.selectors = [BinarySelector(r) for r in self.rngs]
        self.trajPoints = []
        self.deltaH = []
        self.graph = bool(graph)  # replay the integration from a captured HIP graph
        self._lf = None
        self._pool = None

    def momenta(self, n, device):
        """Each chain's momenta rng.normal(0, 1, n) + 0j (leapfrog.py:51), drawn on a thread
        pool (the chains' generators are independent and numpy releases the GIL while
        drawing); returns the real parts on the host and the complex momenta on the device."""
        batch = len(self.rngs)
        if self._pool is None:
            self._pool = ThreadPoolExecutor(max_workers=min(16, batch))
        real = np.empty((batch, n))

        def draw(b):
            real[b] = self.rngs[b].normal(0, 1, n)
        list(self._pool.map(draw, range(batch)))
        pi = torch.zeros((batch, n), dtype=torch.complex128, device=device)
        pi.real.copy_(torch.from_numpy(real))
        return real, pi

    def evolve(self, phis, actVals):
        """phis: [batch, nx*nt] device tensor, actVals: numpy complex [batch].
        Returns (phis', actVals', trajPoints)."""
        batch, n = phis.shape
        real0, pi = self.momenta(n, phis.device)
        if self.graph:
            if self._lf is None:
                self._lf = LeapfrogGraph(self.action, batch, n, self.length, self.nstep, split=self.split)
            phi1, pi1 = self._lf(phis, pi)
            acts1 = _eval_batch(self.action, phi1)
        else:
            phi1, pi1, acts1 = leapfrog_batch(phis, pi, self.action, self.length, self.nstep)
        kin0 = np.einsum("ij,ij->i", real0, real0)  # |pi|^2 of each chain
        kin1 = (pi1.real.square() + pi1.imag.square()).sum(dim=1).cpu().numpy()
        new_phi = phis.clone()
        new_act = np.array(actVals, dtype=complex)
        points = []
        dh = []
        for b in range(batch):
            e0 = actVals[b] + kin0[b] / 2
            e1 = acts1[b] + kin1[b] / 2
            dh.append(np.real(e1 - e0))
            tp = self.selectors[b].selectTrajPoint(e0, e1)
            points.append(tp)
            if tp == 1:
                new_phi[b] = phi1[b]
                new_act[b] = acts1[b]
        self.trajPoints.append(points)
        self.deltaH.append(dh)
        return new_phi, new_act, points


def hmc(stage, evolver, ntr):
    """HMC loop (src/isle/drivers/hmc.py:41-84) without file output.

    Returns the list of stages after each trajectory."""
    if not isinstance(stage, EvolutionStage):
        stage = EvolutionStage(stage, evolver.action.eval(stage), 1)
    out = []
    for _ in range(ntr):
        try:
            stage = evolver.evolve(stage)
        except StopIteration:
            break
        out.append(stage)
    return out
