"""HMC driver with run files: trajectories, checkpoints and resumption
(src/isle/drivers/hmc.py:13-350).

``newRun(lattice, params, rng, makeAction, outfile, overwrite)`` starts a run and
``continueRun(infile, outfile, startIdx, overwrite)`` resumes one from a checkpoint,
returning the driver, the stage, the evolver and the save and checkpoint strides found
in the file.  Calling the driver, ``hmc(stage, evolver, ntr, saveFreq, checkpointFreq)``,
evolves ``ntr`` trajectories, saving the configuration every ``saveFreq`` trajectories
and a checkpoint (RNG state, evolver) every ``checkpointFreq`` (a multiple of
saveFreq; 0 disables either).  A run resumed from checkpoint i continues exactly as the
uninterrupted run would have from trajectory i: the RNG state, configuration, action
value and evolver state are those at the end of trajectory i.

``BatchedHMC`` does the same for independent chains advanced together by a
BatchedConstStepLeapfrog, one run directory (and one RNG) per chain.

The run format is described in isle_amd.runfile.
"""
import inspect
from logging import getLogger

import numpy as np

from .. import runfile
from ..evolver import BatchedConstStepLeapfrog, ConstStepLeapfrog, EvolutionStage, readStateH5


class HMC:
    """Driver that evolves a configuration and writes it to a run directory (hmc.py:13-146)."""

    def __init__(self, lattice, params, rng, action, outfname, startIdx, definitions={}, evManager=None):
        self.lattice = lattice
        self.params = params
        self.rng = rng
        self.action = action
        self.outfname = str(outfname)
        self._trajIdx = startIdx
        self._evManager = evManager if evManager else runfile.EvolverManager(outfname, definitions=definitions)

    def __call__(self, stage, evolver, ntr, saveFreq, checkpointFreq, maxNtr=None):
        """Evolve ``ntr`` trajectories (``None``: until the evolver raises StopIteration or
        ``maxNtr`` is reached) and return the last stage (hmc.py:41-84)."""
        if checkpointFreq != 0 and checkpointFreq % saveFreq != 0:
            getLogger(__name__).error("checkpointFreq must be a multiple of saveFreq."
                                      " Got %d and %d, resp.", checkpointFreq, saveFreq)
            raise ValueError("checkpointFreq must be a multiple of saveFreq.")
        if not isinstance(stage, EvolutionStage):
            stage = EvolutionStage(stage, self.action.eval(stage), 1)
        for _ in _iterTrajectories(ntr, maxNtr):
            try:
                stage = evolver.evolve(stage)
            except StopIteration:
                break
            self.advance()
            self._saveConditionally(stage, evolver, saveFreq, checkpointFreq)
        getLogger(__name__).info("Evolution stopped, evolver:\n%s", evolver.report())
        return stage

    def saveFieldAndCheckpoint(self, stage, evolver):
        """Write the trajectory endpoint and a checkpoint."""
        name = runfile.writeTrajectory(self.outfname, self._trajIdx, stage)
        runfile.writeCheckpoint(self.outfname, self._trajIdx, self.rng, name, evolver, self._evManager)

    def save(self, stage):
        """Write the trajectory endpoint."""
        runfile.writeTrajectory(self.outfname, self._trajIdx, stage)

    def advance(self, amount=1):
        """Advance the trajectory counter without saving."""
        self._trajIdx += amount

    def resetIndex(self, idx=0):
        """Set the trajectory counter to ``idx``."""
        self._trajIdx = idx

    @property
    def trajIdx(self):
        return self._trajIdx

    def _saveConditionally(self, stage, evolver, saveFreq, checkpointFreq):
        if saveFreq != 0 and self._trajIdx % saveFreq == 0:
            if checkpointFreq != 0 and self._trajIdx % checkpointFreq == 0:
                self.saveFieldAndCheckpoint(stage, evolver)
            else:
                self.save(stage)


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


def newRun(lattice, params, rng, makeAction, outfile, overwrite, definitions={}):
    """Start a run: create the run directory with its metadata and return the driver
    (hmc.py:149-179).  ``makeAction(lattice, params)`` builds the action; its source is
    stored so that continueRun can rebuild it."""
    if outfile is None:
        getLogger(__name__).error("No output file given for HMC driver")
        raise ValueError("No output file")
    runfile.initializeNewFile(outfile, overwrite, lattice, params, _source_of(makeAction))
    action = makeAction(lattice, params) if callable(makeAction) else \
        _function_from_source(makeAction)(lattice, params)
    return HMC(lattice, params, rng, action, outfile, 0, definitions)


def continueRun(infile, outfile, startIdx, overwrite, definitions={}, makeAction=None):
    """Resume a run from checkpoint ``startIdx`` (negative: counted from the last
    checkpoint) (hmc.py:182-323).

    Returns (driver, stage, evolver, saveFreq, checkpointFreq); the frequencies are the
    strides of the last two stored configurations / checkpoints, or None.  The action is
    ``makeAction(lattice, params)`` when given, else it is rebuilt from the source stored
    in the run's metadata (which executes that source, as the reference does).

    With ``outfile`` None or equal to ``infile``, configurations and checkpoints newer than
    the start are an error unless ``overwrite``, which deletes them; with another
    ``outfile``, the new run directory is created (replacing an existing run only if
    ``overwrite``)."""
    if infile is None:
        getLogger(__name__).error("No input file given for HMC driver in continuation run")
        raise ValueError("No input file")
    if outfile is None:
        outfile = infile
    lattice, params, makeActionSrc, _versions = runfile.readMetadata(infile)
    if makeAction is None:
        if not makeActionSrc:
            raise ValueError("the run stores no action source; pass makeAction")
        makeAction = _function_from_source(makeActionSrc)
    action = makeAction(lattice, params)
    same = str(outfile) == str(infile)
    if not same:
        runfile.initializeNewFile(outfile, overwrite, lattice, params, makeActionSrc)

    configurations = runfile.indices(infile, "configuration")
    checkpoints = runfile.indices(infile, "checkpoint")
    evManager = runfile.EvolverManager(infile, definitions=definitions)
    checkpointIdx, rng, stage, evolver = _loadCheckpoint(infile, startIdx, checkpoints, evManager, action, lattice)
    if same:
        _ensureNoNewerConfigs(infile, checkpointIdx, checkpoints, configurations, overwrite)
    driver = HMC(lattice, params, rng, action, outfile, checkpointIdx, definitions, evManager if same else None)
    return driver, stage, evolver, _stride(configurations), _stride(checkpoints)


def _stride(values):
    return values[-1] - values[-2] if len(values) >= 2 else None


def _resolve_start(startIdx, checkpoints):
    if not checkpoints:
        raise ValueError("The run has no checkpoint to continue from")
    if startIdx < 0:
        startIdx = checkpoints[-1] + (startIdx + 1)  # -1 is the last checkpoint
    if startIdx < 0 or startIdx > checkpoints[-1]:
        getLogger(__name__).error("Start index for HMC continuation is out of range: %d", startIdx)
        raise ValueError("Start index out of range")
    if startIdx not in checkpoints:
        getLogger(__name__).error("There is no checkpoint matching the given start index: %d", startIdx)
        raise ValueError("No checkpoint matching start index")
    return startIdx


def _loadCheckpoint(fname, startIdx, checkpoints, evManager, action, lattice):
    startIdx = _resolve_start(startIdx, checkpoints)
    rng, cfg, evolver = runfile.loadCheckpoint(fname, startIdx, evManager, action, lattice)
    return startIdx, rng, EvolutionStage.fromH5(cfg), evolver


def _ensureNoNewerConfigs(fname, checkpointIdx, checkpoints, configurations, overwrite):
    """Refuse (or, with ``overwrite``, delete) checkpoints and configurations newer than the
    start (hmc.py:262-288)."""
    for group, idxs in (("checkpoint", checkpoints), ("configuration", configurations)):
        if idxs and idxs[-1] > checkpointIdx:
            message = (f"Output file {fname} contains {group}s with greater index than HMC starting point.\n"
                       f"    Greatest index is {idxs[-1]}, start index is {checkpointIdx}.")
            if not overwrite:
                getLogger(__name__).error(message)
                raise RuntimeError("HMC start index is not latest")
            getLogger(__name__).warning(message + "\n    Overwriting")
            runfile.removeGreaterThan(fname, group, checkpointIdx)


def _iterTrajectories(ntr, maxNtr):
    if ntr is not None:
        yield from range(ntr)
        return
    count = 0
    while True:
        yield count
        count += 1
        if maxNtr is not None and count >= maxNtr:
            getLogger(__name__).warning("Reached maximum number of trajectories for 'unbounded' evolution.")
            return


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
        self._trajIdx = startIdx
        self._evManager = runfile.EvolverManager()

    @property
    def trajIdx(self):
        return self._trajIdx

    def __call__(self, phis, actVals, evolver, ntr, saveFreq, checkpointFreq):
        """Evolve the chains ([batch, nx*nt] device fields, their action values) ``ntr``
        trajectories; returns (phis, actVals, trajPoints of the last trajectory)."""
        if checkpointFreq != 0 and checkpointFreq % saveFreq != 0:
            raise ValueError("checkpointFreq must be a multiple of saveFreq.")
        if not isinstance(evolver, BatchedConstStepLeapfrog) or evolver.rngs != self.rngs:
            raise ValueError("the evolver must be a BatchedConstStepLeapfrog drawing from this driver's rngs")
        points = [1] * len(self.rngs)
        actVals = np.asarray(actVals, dtype=complex)
        for _ in range(ntr):
            phis, actVals, points = evolver.evolve(phis, actVals)
            self._trajIdx += 1
            if saveFreq != 0 and self._trajIdx % saveFreq == 0:
                host = phis.detach().cpu().numpy()
                ckpt = checkpointFreq != 0 and self._trajIdx % checkpointFreq == 0
                for b, fname in enumerate(self.outfnames):
                    stage = EvolutionStage(host[b], actVals[b], points[b])
                    name = runfile.writeTrajectory(fname, self._trajIdx, stage)
                    if ckpt:  # recorded as the ConstStepLeapfrog that would run this chain alone
                        single = ConstStepLeapfrog(self.action, evolver.length, evolver.nstep, self.rngs[b])
                        runfile.writeCheckpoint(fname, self._trajIdx, self.rngs[b], name, single, self._evManager)
        return phis, actVals, points


def newBatchedRun(lattice, params, rngs, makeAction, outfiles, overwrite):
    """Start one run directory per chain (same lattice, parameters and action)."""
    src = _source_of(makeAction)
    for fname in outfiles:
        runfile.initializeNewFile(fname, overwrite, lattice, params, src)
    action = makeAction(lattice, params) if callable(makeAction) else _function_from_source(makeAction)(lattice, params)
    return BatchedHMC(lattice, params, rngs, action, outfiles, 0)


def continueBatchedRun(infiles, startIdx, overwrite, makeAction=None, graph=False, split=None):
    """Resume every chain of a batched run from checkpoint ``startIdx`` (negative: from the
    last common checkpoint).  Returns (driver, phis [batch, nx*nt] numpy, actVals,
    evolver): the evolver is a BatchedConstStepLeapfrog over the restored generators."""
    lattice, params, makeActionSrc, _ = runfile.readMetadata(infiles[0])
    if makeAction is None:
        if not makeActionSrc:
            raise ValueError("the run stores no action source; pass makeAction")
        makeAction = _function_from_source(makeActionSrc)
    action = makeAction(lattice, params)
    common = sorted(set.intersection(*[set(runfile.indices(f, "checkpoint")) for f in infiles]))
    start = _resolve_start(startIdx, common)
    phis, acts, rngs, lengths = [], [], [], set()
    for fname in infiles:
        grp = runfile.readCheckpoint(fname, start)
        rngs.append(readStateH5(grp["rngState"]))
        ev = grp["evolver"]
        if str(ev["__name__"][()]) != "ConstStepLeapfrog":
            raise ValueError("batched runs resume ConstStepLeapfrog checkpoints only")
        lengths.add((float(ev["length"][()]), int(ev["nstep"][()])))
        stage = EvolutionStage.fromH5(runfile.readTrajectory(fname, start))
        phis.append(stage.phi)
        acts.append(stage.actVal)
        _ensureNoNewerConfigs(fname, start, runfile.indices(fname, "checkpoint"),
                              runfile.indices(fname, "configuration"), overwrite)
    if len(lengths) != 1:
        raise ValueError("the chains' checkpoints disagree on the trajectory length or step count")
    length, nstep = lengths.pop()
    evolver = BatchedConstStepLeapfrog(action, length, nstep, rngs, graph=graph, split=split)
    driver = BatchedHMC(lattice, params, rngs, action, infiles, start)
    return driver, np.stack(phis), np.array(acts), evolver
