"""Markov-chain evolution over the GPU action: generators, chain states, the Metropolis
choice, leapfrog HMC evolvers, 2 pi jump evolvers and batched chains.

Public names follow the reference's evolver package (src/isle/evolver/, src/isle/random.py)
so that driver code written against it runs unchanged; the random numbers are consumed
in the reference's order, which fixed-seed parity of the accept decisions needs:

* one trajectory draws its momenta rng.normal(0, 1, N) (real, as complex with zero
  imaginary part) and then, only if the energy rises, one rng.uniform(0, 1);
* a 2 pi jump batch draws its sites (rng.choice without replacement), then its signs,
  then the Metropolis uniform as above.

Fields live on the GPU during a trajectory; stages may hold numpy vectors or device
tensors.  ``BatchedConstStepLeapfrog`` advances many independent chains in one set of
launches, each with its own generator consumed exactly as a lone chain would consume it.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import transform
from .core import as_field
from .integrator import LeapfrogGraph, eval_batch as _eval_batch, leapfrog_batch
from .transform import pull_back, push_forward


# ---------------------------------------------------------------------------------------
# random numbers

class NumpyRNG:
    """Mersenne Twister from numpy's legacy ``RandomState``: a seed gives the stream of the
    reference's NumpyRNG (src/isle/random.py:80-185).

    ``uniform`` and ``normal`` take (first parameter, second parameter, size, cmplx); a
    complex draw is the real parts followed by the imaginary parts, ``size`` numbers each."""

    NAME = "np.MT19937"
    _BITGEN = "MT19937"

    def __init__(self, seed):
        self._mt = np.random.RandomState(seed)

    def seed(self, seed):
        self._mt.seed(seed)

    def _sample(self, method, a, b, size, cmplx):
        draw = getattr(self._mt, method)
        values = draw(a, b, size)
        if cmplx:
            values = values + 1j * draw(a, b, size)
        return values

    def uniform(self, low=0.0, high=1.0, size=None, cmplx=False):
        return self._sample("uniform", low, high, size, cmplx)

    def normal(self, mean=0.0, std=1.0, size=None, cmplx=False):
        return self._sample("normal", mean, std, size, cmplx)

    def choice(self, a, size=None, replace=True, p=None):
        return self._mt.choice(a, size=size, replace=replace, p=p)

    def get_state(self):
        return self._mt.get_state()

    def set_state(self, state):
        self._mt.set_state(state)

    def writeH5(self, group):
        """Put the generator into a checkpoint group: datasets ``name`` and ``keys`` (the
        624 state words), attributes ``pos``, ``has_gauss`` and ``cached_gaussian``."""
        snap = self._mt.get_state(legacy=False)
        group["name"] = self.NAME
        group["keys"] = np.asarray(snap["state"]["key"], dtype=np.uint32)
        group.attrs["pos"] = int(snap["state"]["pos"])
        group.attrs["has_gauss"] = int(snap["has_gauss"])
        group.attrs["cached_gaussian"] = float(snap["gauss"])

    def readH5(self, group):
        """Take the state stored by ``writeH5``."""
        stored = str(group["name"][()])
        if stored != self.NAME:
            raise RuntimeError(f"checkpoint holds a {stored!r} generator; {self.NAME} cannot take its state")
        attrs = group.attrs
        self._mt.set_state({
            "bit_generator": self._BITGEN,
            "state": {"key": np.asarray(group["keys"][()], dtype=np.uint32), "pos": int(attrs["pos"])},
            "has_gauss": int(attrs["has_gauss"]),
            "gauss": float(attrs["cached_gaussian"]),
        })


_GENERATORS = {NumpyRNG.NAME: NumpyRNG}


def readStateH5(group):
    """A new generator in the state a checkpoint group holds."""
    kind = str(group["name"][()])
    if kind not in _GENERATORS:
        raise ValueError(f"no generator type is known by the name {kind!r}")
    rng = _GENERATORS[kind](0)
    rng.readH5(group)
    return rng


def _host(x):
    """A field as a complex128 numpy vector."""
    if isinstance(x, torch.Tensor):
        return x.detach().to("cpu", torch.complex128).numpy()
    return np.asarray(x, dtype=np.complex128)


# ---------------------------------------------------------------------------------------
# chain states and the Metropolis choice

class EvolutionStage:
    """One point of a Markov chain (src/isle/evolver/stage.py:10-117).

    ``phi`` is the field; ``logWeights`` maps names to log weights whose sum enters the
    Metropolis energy, ``logWeights["actVal"]`` (also ``actVal``) being the action;
    ``trajPoint`` is 1 if the step that produced the stage moved the chain, else 0;
    ``extra`` holds any further data and ``phi_RTP`` an optional second field.

    ``accept`` and ``reject`` build new stages; a new stage gets its own copy of the
    weights, so later changes to one stage's weights never reach another."""

    __slots__ = ("phi", "trajPoint", "logWeights", "extra", "phi_RTP")

    def __init__(self, phi, actVal, trajPoint=1, logWeights=None, extra=None, phi_RTP=None):
        self.phi = phi
        self.trajPoint = trajPoint
        self.phi_RTP = phi_RTP
        self.logWeights = dict(logWeights) if logWeights else {}
        self.logWeights["actVal"] = actVal
        self.extra = dict(extra) if extra else {}

    @property
    def actVal(self):
        return self.logWeights["actVal"]

    @actVal.setter
    def actVal(self, value):
        self.logWeights["actVal"] = value

    def accept(self, phi, actVal, logWeights=None, extra=None, phi_RTP=None):
        """The stage after a move to ``phi`` (trajPoint 1); weights other than the action
        come from ``logWeights`` only."""
        return type(self)(phi, actVal, 1, logWeights, extra, phi_RTP)

    def reject(self, extra=None):
        """The stage after a refused move: same field and weights, trajPoint 0."""
        return type(self)(self.phi, self.actVal, 0, self.logWeights, extra, self.phi_RTP)

    def sumLogWeights(self):
        return sum(self.logWeights.values())

    def save(self, group):
        """Write the stage into a run-file group: ``phi``, ``actVal``, ``trajPoint``,
        ``phi_RT-P`` when set, ``logWeights/<name>`` for each weight besides the action and
        ``extra/<name>``."""
        group["phi"] = _host(self.phi)
        group["actVal"] = complex(self.actVal)
        group["trajPoint"] = int(self.trajPoint)
        if self.phi_RTP is not None:
            group["phi_RT-P"] = _host(self.phi_RTP)
        subgroups = (("logWeights", {k: complex(v) for k, v in self.logWeights.items() if k != "actVal"}),
                     ("extra", self.extra))
        for title, entries in subgroups:
            if entries:
                sub = group.create_group(title)
                for key, value in entries.items():
                    sub[key] = value

    @classmethod
    def fromH5(cls, group):
        """The stage that ``save`` wrote; fields come back as numpy vectors."""
        def entries(title, convert):
            if title not in group:
                return None
            return {key: convert(ds[()]) for key, ds in group[title].items()}
        second = group["phi_RT-P"][()] if "phi_RT-P" in group else None
        return cls(group["phi"][()], complex(group["actVal"][()]), int(group["trajPoint"][()]),
                   entries("logWeights", complex), entries("extra", lambda v: v), second)


class BinarySelector:
    """Metropolis choice between the start (0) and the end (1) of a proposal
    (src/isle/evolver/selector.py:9-44): the end is taken if its energy is lower;
    otherwise one uniform u is drawn and the end is taken if u < exp(-dE), dE the real
    part of the energy difference.  A downhill proposal draws nothing."""

    def __init__(self, rng):
        self.rng = rng

    def selectTrajPoint(self, energy0, energy1):
        rise = float(np.real(energy1 - energy0))
        if rise < 0:
            return 1
        return int(self.rng.uniform(0, 1) < math.exp(-rise))

    def selectTrajectory(self, energy0, data0, energy1, data1):
        if self.selectTrajPoint(energy0, energy1):
            return energy1, data1, 1
        return energy0, data0, 0


def _acceptance_line(name, ev, settings, decisions):
    taken = int(np.sum(decisions)) if len(decisions) else 0
    total = int(np.size(decisions))
    rate = f"{taken / total:.4f}" if total else "n/a"
    return f"{name} @0x{id(ev):x}: {settings}; accepted {taken} of {total} ({rate})"


# ---------------------------------------------------------------------------------------
# leapfrog HMC

class _LeapfrogHMC:
    """One HMC trajectory per ``evolve`` (src/isle/evolver/leapfrog.py:38-66): the stage's
    field is pulled back by the transform (if any), integrated by the GPU leapfrog with
    fresh Gaussian momenta, pushed forward, and the end point is accepted or rejected on
    E = sum of log weights (+ log det J) + |pi|^2 / 2.  Subclasses choose each
    trajectory's (length, nstep).  ``trajPoints`` records the decisions and
    ``last_energies`` the last trajectory's (E0, E1)."""

    def __init__(self, action, rng, transform):
        self.action = action
        self.rng = rng
        self.transform = transform
        self.selector = BinarySelector(rng)
        self.trajPoints = []
        self.last_energies = None

    def _trajectory_parameters(self):
        raise NotImplementedError

    def evolve(self, stage):
        length, nstep = self._trajectory_parameters()
        phi_md, logdet0 = pull_back(self.transform, stage)
        if self.transform is not None:
            stage.logWeights.setdefault("logdetJ", logdet0)
        phi_md = as_field(phi_md).reshape(1, -1)
        momenta = self.rng.normal(0, 1, phi_md.shape[1])
        pi = torch.as_tensor(momenta + 0j, device=phi_md.device).reshape(1, -1)
        ends, pi_end, acts = leapfrog_batch(phi_md, pi, self.action, length, nstep)
        phi1, act1, logdet1 = push_forward(self.transform, ends[0], complex(acts[0]))
        energy0 = stage.sumLogWeights() + float(np.dot(momenta, momenta)) / 2
        energy1 = act1 + logdet1 + float((pi_end.real.square() + pi_end.imag.square()).sum()) / 2
        self.last_energies = (energy0, energy1)
        moved = self.selector.selectTrajPoint(energy0, energy1)
        self.trajPoints.append(moved)
        if not moved:
            return stage.reject()
        return stage.accept(phi1, act1, None if self.transform is None else {"logdetJ": logdet1})

    def _save_transform(self, group, manager):
        if self.transform is not None:
            manager.save(self.transform, group.create_group("transform"))

    @staticmethod
    def _load_transform(group, manager, action, lattice, rng):
        return manager.load(group["transform"], action, lattice, rng) if "transform" in group else None


class ConstStepLeapfrog(_LeapfrogHMC):
    """Every trajectory has length ``length`` in ``nstep`` leapfrog steps
    (src/isle/evolver/leapfrog.py:15-103).  Checkpoints store ``length``, ``nstep`` and
    the transform."""

    def __init__(self, action, length, nstep, rng, transform=None):
        super().__init__(action, rng, transform)
        self.length = length
        self.nstep = nstep

    def _trajectory_parameters(self):
        return self.length, self.nstep

    def save(self, group, manager):
        group["length"] = self.length
        group["nstep"] = self.nstep
        self._save_transform(group, manager)

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, float(group["length"][()]), int(group["nstep"][()]), rng,
                   cls._load_transform(group, manager, action, lattice, rng))

    def report(self):
        return _acceptance_line("ConstStepLeapfrog", self, f"length {self.length}, {self.nstep} steps",
                                self.trajPoints)


def _hinge_range(start, end, step):
    """start, start + step, ... while short of end (the sum accumulated step by step),
    then end for ever (src/isle/collection.py:10-30)."""
    cur = start
    while (step > 0 and cur < end) or (step < 0 and cur > end):
        yield cur
        cur += step
    while True:
        yield end


class LinearStepLeapfrog(_LeapfrogHMC):
    """Leapfrog whose trajectory length and step count move linearly from the first to the
    second value of their ranges over ``ninterp`` trajectories and then stay at the second
    (leapfrog.py:105-229).  ``startPoint`` skips that many trajectories of the schedule.

    The position in the schedule that a checkpoint stores counts every trajectory since
    the schedule began, including those before a resumption (the reference restarts that
    count at zero when it resumes, so a second resumption would rewind the schedule)."""

    def __init__(self, action, lengthRange, nstepRange, ninterp, rng, startPoint=0, transform=None):
        super().__init__(action, rng, transform)
        self.lengthRange = tuple(lengthRange)
        self.nstepRange = tuple(nstepRange)
        self.ninterp = ninterp
        self._lengths = _hinge_range(*self.lengthRange, (self.lengthRange[1] - self.lengthRange[0]) / ninterp)
        self._nsteps = _hinge_range(*self.nstepRange, (self.nstepRange[1] - self.nstepRange[0]) / ninterp)
        self._current = 0
        for _ in range(int(startPoint)):
            self._trajectory_parameters()

    def _trajectory_parameters(self):
        self._current += 1
        return next(self._lengths), int(next(self._nsteps))

    def save(self, group, manager):
        """Ranges and schedule position under the names minLength, maxLength, minNstep,
        maxNstep, ninterp and current (leapfrog.py:183-196)."""
        for key, value in (("minLength", self.lengthRange[0]), ("maxLength", self.lengthRange[1]),
                           ("minNstep", self.nstepRange[0]), ("maxNstep", self.nstepRange[1]),
                           ("ninterp", self.ninterp), ("current", self._current)):
            group[key] = value
        self._save_transform(group, manager)

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, (group["minLength"][()], group["maxLength"][()]),
                   (group["minNstep"][()], group["maxNstep"][()]), int(group["ninterp"][()]), rng,
                   startPoint=int(group["current"][()]),
                   transform=cls._load_transform(group, manager, action, lattice, rng))

    def report(self):
        return _acceptance_line("LinearStepLeapfrog", self,
                                f"length {self.lengthRange[0]} to {self.lengthRange[1]}, steps {self.nstepRange[0]} "
                                f"to {self.nstepRange[1]} over {self.ninterp} trajectories", self.trajPoints)


# ---------------------------------------------------------------------------------------
# jumps by 2 pi

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
    """Moves the field by +-2 pi at random sites, momenta untouched
    (src/isle/evolver/hubbard.py:15-116).

    One ``evolve`` makes ``nBatches`` proposals in turn.  A proposal picks ``batchSize``
    distinct sites (rng.choice without replacement) and a sign for each (rng.choice of
    -2 pi and +2 pi) and is accepted or rejected by itself on the change of the action.
    The stage returned holds the field after the last accepted proposal (trajPoint 1 if
    any was accepted); ``acceptedPerBatch`` records each evolve's decisions."""

    def __init__(self, nBatches, batchSize, action, lattice, rng):
        self.nBatches = int(nBatches)
        self.batchSize = int(batchSize)
        self.rng = rng
        self.selector = BinarySelector(rng)
        self._nsites = lattice.lattSize()
        self._action = _HubbardActionShortcut(action)
        self.acceptedPerBatch = []

    def _proposal(self):
        sites = self.rng.choice(self._nsites, self.batchSize, replace=False)
        shifts = self.rng.choice((-2 * np.pi, +2 * np.pi), self.batchSize, replace=True)
        return sites, shifts

    def evolve(self, stage):
        current, value = _host(stage.phi), complex(stage.actVal)
        decisions = []
        for _ in range(self.nBatches):
            sites, shifts = self._proposal()
            trial = current.copy()
            trial[sites] += shifts
            trial_value = self._action.evalLocal(trial, sites, shifts, value)
            took = self.selector.selectTrajPoint(value, trial_value)
            decisions.append(took)
            if took:
                current, value = trial, trial_value
        self.acceptedPerBatch.append(decisions)
        return stage.accept(current, value) if any(decisions) else stage.reject()

    def save(self, group, manager):
        group["nBatches"] = self.nBatches
        group["batchSize"] = self.batchSize

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(int(group["nBatches"][()]), int(group["batchSize"][()]), action, lattice, rng)

    def report(self):
        return _acceptance_line("TwoPiJumps", self, f"{self.nBatches} proposals of {self.batchSize} sites",
                                self.acceptedPerBatch)


class UniformJump:
    """Moves every entry of the field by the same +-2 pi / Nt, momenta untouched
    (src/isle/evolver/hubbard.py:119-184): the sign is rng.choice((-1, +1)), then the
    Metropolis choice on the change of the action.  ``accepted`` records the decisions."""

    def __init__(self, action, lattice, rng):
        self.rng = rng
        self.selector = BinarySelector(rng)
        self._action = _HubbardActionShortcut(action)
        self._absShift = 2 * np.pi / lattice.nt()
        self.accepted = []

    def evolve(self, stage):
        shift = self._absShift * self.rng.choice((-1, +1))
        trial = _host(stage.phi) + shift
        trial_value = self._action.evalGlobal(trial, shift, complex(stage.actVal))
        took = self.selector.selectTrajPoint(stage.actVal, trial_value)
        self.accepted.append(took)
        return stage.accept(trial, trial_value) if took else stage.reject()

    def save(self, group, manager):
        return None

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, lattice, rng)

    def report(self):
        return _acceptance_line("UniformJump", self, f"shift 2 pi / Nt = {self._absShift:.6g}", self.accepted)


# ---------------------------------------------------------------------------------------
# batched chains

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
    """Evolve ``ntr`` steps without writing files and return every stage reached; a field
    given instead of a stage is evaluated first.  Stops early when the evolver raises
    StopIteration."""
    if not isinstance(stage, EvolutionStage):
        stage = EvolutionStage(stage, evolver.action.eval(stage), 1)
    visited = []
    for _ in range(ntr):
        try:
            stage = evolver.evolve(stage)
        except StopIteration:
            return visited
        visited.append(stage)
    return visited
