"""isle_amd: the Hubbard fermion-action hot path of isle on AMD MI355X (gfx950).

Drop-in names for the reference's ``isle`` / ``isle.action`` surface on this path
(src/isle/__init__.py, src/isle/cpp_wrappers.py, src/isle/cpp/bind/*):

    Vector, CDVector, Species, Lattice, isBipartite, loadLattice,
    HubbardFermiMatrixExp, HubbardFermiMatrixDia (with P, Tplus, Tminus, Q), logdetM, solveM,
    logdetQ, solveQ, getQLU (the block LU of Q = M_p M_h^T), logdet, invert,
    leapfrog, rungeKutta4Flow, action.{Action, SumAction, HubbardGaugeAction, HubbardFermiAction,
    makeHubbardFermiAction, HFAHopping, HFABasis, HFAAlgorithm},
    evolver.{ConstStepLeapfrog, LinearStepLeapfrog, TwoPiJumps, UniformJump, BinarySelector,
    EvolutionStage},
    random.NumpyRNG, drivers.hmc.{newRun, continueRun, HMC} (run files with checkpoints,
    isle_amd.runfile), util.parameters, meas.{AllToAll, ChiralCondensate, Logdet}
    (measurements on solveM / logdetM, refined to the exact solve), applyM (M times rows),
    autograd.{action_value, logdetM_value} (the action and log det M as
    torch.autograd functions whose backward is the force sweep), and
    ensemble.{Chain, EnsembleLeapfrog} (many chains of uneven Nx and Nt advanced
    together, one HIP graph per trajectory).

Every routine of the fermion-action path runs in the hand-written HIP kernels of
libisle_hfa.so; applyM (the residual of the measurements' refinement) is one real GEMM
of the library kind (torch on the device), and the block products of the Q-matrix LU
(a small-size cross-check) are library complex GEMMs around the HIP LU.  There is no CPU
fallback.
"""
from . import action, autograd, drivers, ensemble, evolver, meas, runfile, util
from .core import (QLU, CDVector, HubbardFermiMatrixDia, HubbardFermiMatrixExp, Species, Vector, applyM, as_field,
                   getQLU, invert, logdet, logdetM, logdetQ, solveM, solveQ, to_first_log_branch)
from .integrator import leapfrog, leapfrog_batch, rungeKutta4Flow
from .lattice import LATTICES, Lattice, honeycomb, isBipartite, loadLattice

random = evolver  # isle.random.NumpyRNG lives in isle_amd.evolver

__all__ = [
    "CDVector", "HubbardFermiMatrixDia", "HubbardFermiMatrixExp", "LATTICES", "Lattice", "QLU", "Species", "Vector",
    "action",
    "autograd", "getQLU", "logdetQ", "solveQ",
    "applyM", "drivers", "ensemble", "meas", "runfile", "util",
    "as_field", "evolver", "honeycomb", "invert", "isBipartite", "leapfrog", "leapfrog_batch", "loadLattice",
    "logdet", "logdetM", "random", "rungeKutta4Flow", "solveM", "to_first_log_branch",
]
