"""HMC runs that write their chain into a run directory and resume from it (the driver of
src/isle/drivers/hmc.py:13-350; the directory format is described in isle_amd.runfile).

* ``newRun(lattice, params, rng, makeAction, outfile, overwrite)`` creates the run
  directory and returns an ``HMC`` driver at trajectory 0;
* ``continueRun(infile, outfile, startIdx, overwrite)`` returns (driver, stage, evolver,
  save stride, checkpoint stride) restored from checkpoint ``startIdx`` (negative counts
  back from the newest checkpoint, -1 being the newest);
* ``driver(stage, evolver, ntr, saveFreq, checkpointFreq)`` evolves ``ntr`` trajectories
  and writes configuration i when i % saveFreq == 0 and, with it, checkpoint i (generator
  state, evolver state) when i % checkpointFreq == 0; 0 turns either off.

A run resumed at checkpoint i continues as the uninterrupted run would have from
trajectory i, because the checkpoint holds everything the next trajectory draws on.

``BatchedHMC`` drives independent chains advanced together by a BatchedConstStepLeapfrog;
every chain writes an ordinary single-chain run directory of its own.
"""
import inspect
import itertools
from logging import getLogger

import numpy as np

from .. import runfile
from ..evolver import BatchedConstStepLeapfrog, ConstStepLeapfrog, EvolutionStage, readStateH5

_log = getLogger(__name__)


class _Cadence:
    """Which trajectory indices get a configuration and which also a checkpoint."""

    def __init__(self, saveFreq, checkpointFreq):
        self.save_every = int(saveFreq)
        self.checkpoint_every = int(checkpointFreq)
        if self.checkpoint_every and (not self.save_every or self.checkpoint_every % self.save_every):
            raise ValueError(f"a checkpoint every {self.checkpoint_every} trajectories needs the configuration of "
                             f"that trajectory, but configurations are written every {self.save_every}")

    def at(self, idx):
        """(write the configuration, write a checkpoint) after trajectory ``idx``."""
        config = self.save_every != 0 and idx % self.save_every == 0
        return config, config and self.checkpoint_every != 0 and idx % self.checkpoint_every == 0


def _trajectory_budget(ntr, maxNtr):
    """Iterations of a run: ``ntr`` of them, or with ntr None at most ``maxNtr`` (at least
    one), or without end when maxNtr is None too."""
    if ntr is not None:
        return range(ntr)
    if maxNtr is None:
        return itertools.count()
    return range(max(int(maxNtr), 1))


class HMC:
    """Evolves a Markov chain and writes it into the run directory ``outfname``.

    ``trajIdx`` is the index of the last trajectory made (``startIdx`` at first); the
    evolver manager stores evolvers in checkpoints (built from ``definitions`` when not
    given)."""

    def __init__(self, lattice, params, rng, action, outfname, startIdx, definitions={}, evManager=None):
        self.lattice = lattice
        self.params = params
        self.rng = rng
        self.action = action
        self.outfname = str(outfname)
        self._index = int(startIdx)
        self._manager = evManager or runfile.EvolverManager(outfname, definitions=definitions)

    @property
    def trajIdx(self):
        return self._index

    def advance(self, amount=1):
        """Count ``amount`` more trajectories without writing anything."""
        self._index += amount

    def resetIndex(self, idx=0):
        self._index = idx

    def save(self, stage):
        """Write ``stage`` as the configuration of the current trajectory."""
        runfile.writeTrajectory(self.outfname, self._index, stage)

    def saveFieldAndCheckpoint(self, stage, evolver):
        """Write the configuration and a checkpoint that refers to it."""
        cfg = runfile.writeTrajectory(self.outfname, self._index, stage)
        runfile.writeCheckpoint(self.outfname, self._index, self.rng, cfg, evolver, self._manager)

    def __call__(self, stage, evolver, ntr, saveFreq, checkpointFreq, maxNtr=None):
        """Evolve ``ntr`` trajectories (ntr None: until the evolver raises StopIteration or
        ``maxNtr`` are done) and return the final stage.  A field passed as ``stage`` is
        evaluated with the driver's action first."""
        cadence = _Cadence(saveFreq, checkpointFreq)
        if not isinstance(stage, EvolutionStage):
            stage = EvolutionStage(stage, self.action.eval(stage), 1)
        budget = _trajectory_budget(ntr, maxNtr)
        done = 0
        for _ in budget:
            try:
                stage = evolver.evolve(stage)
            except StopIteration:
                break
            done += 1
            self._index += 1
            config, checkpoint = cadence.at(self._index)
            if checkpoint:
                self.saveFieldAndCheckpoint(stage, evolver)
            elif config:
                self.save(stage)
        if ntr is None and isinstance(budget, range) and done == len(budget):
            _log.warning("open-ended HMC run stopped at its limit of %d trajectories", done)
        _log.info("HMC run at trajectory %d after %d new ones; %s", self._index, done, evolver.report())
        return stage


# ---------------------------------------------------------------------------------------
# the action of a run: built from makeAction, its source kept in the run's metadata

def _source_of(makeAction):
    if isinstance(makeAction, str):
        return makeAction
    try:
        src = inspect.getsource(makeAction)
    except (OSError, TypeError):
        return ""  # a function defined interactively: continueRun then needs makeAction
    return src if src.lstrip().startswith("def") else ""


def _function_from_source(src):
    """The one function defined by ``src`` (src/isle/meta.py:45-118)."""
    scope = {}
    exec(src, {}, scope)  # noqa: S102 -- the run's own makeAction, as the reference does
    funcs = [v for v in scope.values() if inspect.isfunction(v)]
    if len(scope) != 1 or len(funcs) != 1:
        raise ValueError("the stored action source must define exactly one function")
    return funcs[0]


def _build_action(makeAction, lattice, params):
    build = makeAction if callable(makeAction) else _function_from_source(makeAction)
    return build(lattice, params)


def _stored_run(fname, makeAction):
    """(lattice, params, action source, action) of an existing run directory."""
    lattice, params, source, _versions = runfile.readMetadata(fname)
    if makeAction is None:
        if not source:
            raise ValueError(f"run {fname} keeps no source for its action; pass makeAction")
        makeAction = source
    return lattice, params, source, _build_action(makeAction, lattice, params)


# ---------------------------------------------------------------------------------------
# starting and resuming

def newRun(lattice, params, rng, makeAction, outfile, overwrite, definitions={}):
    """Create the run directory ``outfile`` (replacing an existing run only if
    ``overwrite``) and return a driver at trajectory 0.  ``makeAction(lattice, params)``
    builds the action; its source goes into the run so that continueRun can rebuild it."""
    if outfile is None:
        raise ValueError("newRun needs a run directory to write to (outfile is None)")
    runfile.initializeNewFile(outfile, overwrite, lattice, params, _source_of(makeAction))
    return HMC(lattice, params, rng, _build_action(makeAction, lattice, params), outfile, 0, definitions)


def _last_gap(indices):
    """Difference of the two newest indices, None with fewer than two."""
    return indices[-1] - indices[-2] if len(indices) > 1 else None


def _checkpoint_index(requested, available):
    """The checkpoint to resume from; negative ``requested`` counts back from the newest."""
    if not available:
        raise ValueError("the run holds no checkpoint to resume from")
    newest = available[-1]
    idx = newest + 1 + requested if requested < 0 else requested
    if idx in available:
        return idx
    if 0 <= idx <= newest:
        reason = f"no checkpoint was written at trajectory {idx}"
    else:
        reason = f"trajectory {idx} is outside the checkpointed range 0..{newest}"
    _log.error("cannot resume: %s", reason)
    raise ValueError(f"cannot resume: {reason}")


def _drop_later_records(fname, start, overwrite):
    """Checkpoints and configurations after trajectory ``start`` would be rewritten by the
    resumed run: refuse, or with ``overwrite`` delete them."""
    for group in ("checkpoint", "configuration"):
        later = [i for i in runfile.indices(fname, group) if i > start]
        if not later:
            continue
        what = f"{fname} has {len(later)} {group} record(s) after trajectory {start} (newest {later[-1]})"
        if not overwrite:
            _log.error("%s", what)
            raise RuntimeError(f"{what}; resume from the newest checkpoint or pass overwrite=True")
        _log.warning("%s; deleting them", what)
        runfile.removeGreaterThan(fname, group, start)


def continueRun(infile, outfile, startIdx, overwrite, definitions={}, makeAction=None):
    """Resume the run ``infile`` at checkpoint ``startIdx``.

    Returns (driver, stage, evolver, saveFreq, checkpointFreq): the frequencies are the
    gaps between the two newest configurations / checkpoints (None with fewer than two).
    The action is ``makeAction(lattice, params)`` when given, else it is rebuilt from the
    source kept in the run (executing that source, as the reference does).

    With ``outfile`` None or equal to ``infile`` the run continues in place: records newer
    than the start are an error, or deleted with ``overwrite``.  With another ``outfile`` a
    new run directory is created there (replacing an existing run only if ``overwrite``)."""
    if infile is None:
        raise ValueError("continueRun needs the run directory to resume (infile is None)")
    target = infile if outfile is None else outfile
    in_place = str(target) == str(infile)
    lattice, params, source, action = _stored_run(infile, makeAction)
    if not in_place:
        runfile.initializeNewFile(target, overwrite, lattice, params, source)
    configs = runfile.indices(infile, "configuration")
    checkpoints = runfile.indices(infile, "checkpoint")
    manager = runfile.EvolverManager(infile, definitions=definitions)
    start = _checkpoint_index(startIdx, checkpoints)
    rng, cfg, evolver = runfile.loadCheckpoint(infile, start, manager, action, lattice)
    stage = EvolutionStage.fromH5(cfg)
    if in_place:
        _drop_later_records(infile, start, overwrite)
    driver = HMC(lattice, params, rng, action, target, start, definitions, manager if in_place else None)
    return driver, stage, evolver, _last_gap(configs), _last_gap(checkpoints)


# ---------------------------------------------------------------------------------------
# independent chains advanced together

class BatchedHMC:
    """HMC over independent chains advanced together, chain b written to ``outfnames[b]``
    with its own generator ``rngs[b]`` (each directory is an ordinary single-chain run:
    continueRun resumes any one chain alone, continueBatchedRun all of them together)."""

    def __init__(self, lattice, params, rngs, action, outfnames, startIdx):
        self.lattice = lattice
        self.params = params
        self.rngs = list(rngs)
        self.action = action
        self.outfnames = [str(f) for f in outfnames]
        if len(self.outfnames) != len(self.rngs):
            raise ValueError("one output run per chain")
        self._index = startIdx
        self._manager = runfile.EvolverManager()

    @property
    def trajIdx(self):
        return self._index

    def _write(self, phis, actVals, points, evolver, checkpoint):
        host = phis.detach().cpu().numpy()
        for b, fname in enumerate(self.outfnames):
            cfg = runfile.writeTrajectory(fname, self._index, EvolutionStage(host[b], actVals[b], points[b]))
            if checkpoint:  # recorded as the ConstStepLeapfrog that would run this chain alone
                alone = ConstStepLeapfrog(self.action, evolver.length, evolver.nstep, self.rngs[b])
                runfile.writeCheckpoint(fname, self._index, self.rngs[b], cfg, alone, self._manager)

    def __call__(self, phis, actVals, evolver, ntr, saveFreq, checkpointFreq):
        """Evolve the chains ([batch, nx*nt] device fields, their action values) ``ntr``
        trajectories; returns (phis, actVals, trajPoints of the last trajectory)."""
        cadence = _Cadence(saveFreq, checkpointFreq)
        if not isinstance(evolver, BatchedConstStepLeapfrog) or evolver.rngs != self.rngs:
            raise ValueError("the evolver must be a BatchedConstStepLeapfrog drawing from this driver's rngs")
        points = [1] * len(self.rngs)
        actVals = np.asarray(actVals, dtype=complex)
        for _ in range(ntr):
            phis, actVals, points = evolver.evolve(phis, actVals)
            self._index += 1
            config, checkpoint = cadence.at(self._index)
            if config:
                self._write(phis, actVals, points, evolver, checkpoint)
        return phis, actVals, points


def newBatchedRun(lattice, params, rngs, makeAction, outfiles, overwrite):
    """Start one run directory per chain (same lattice, parameters and action)."""
    src = _source_of(makeAction)
    for fname in outfiles:
        runfile.initializeNewFile(fname, overwrite, lattice, params, src)
    return BatchedHMC(lattice, params, rngs, _build_action(makeAction, lattice, params), outfiles, 0)


def continueBatchedRun(infiles, startIdx, overwrite, makeAction=None, graph=False, split=None):
    """Resume every chain of a batched run from checkpoint ``startIdx`` (negative: from the
    last common checkpoint).  Returns (driver, phis [batch, nx*nt] numpy, actVals,
    evolver): the evolver is a BatchedConstStepLeapfrog over the restored generators."""
    lattice, params, _source, action = _stored_run(infiles[0], makeAction)
    common = sorted(set.intersection(*[set(runfile.indices(f, "checkpoint")) for f in infiles]))
    start = _checkpoint_index(startIdx, common)
    phis, acts, rngs, settings = [], [], [], set()
    for fname in infiles:
        grp = runfile.readCheckpoint(fname, start)
        rngs.append(readStateH5(grp["rngState"]))
        ev = grp["evolver"]
        if str(ev["__name__"][()]) != "ConstStepLeapfrog":
            raise ValueError("batched runs resume ConstStepLeapfrog checkpoints only")
        settings.add((float(ev["length"][()]), int(ev["nstep"][()])))
        stage = EvolutionStage.fromH5(runfile.readTrajectory(fname, start))
        phis.append(stage.phi)
        acts.append(stage.actVal)
        _drop_later_records(fname, start, overwrite)
    if len(settings) != 1:
        raise ValueError("the chains' checkpoints disagree on the trajectory length or step count")
    length, nstep = settings.pop()
    evolver = BatchedConstStepLeapfrog(action, length, nstep, rngs, graph=graph, split=split)
    driver = BatchedHMC(lattice, params, rngs, action, infiles, start)
    return driver, np.stack(phis), np.array(acts), evolver
