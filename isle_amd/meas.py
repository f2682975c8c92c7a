"""Measurements built on solveM and logdetM, computed on the GPU
(src/isle/meas/propagator.py:12-64, src/isle/meas/chiralCondensate.py:11-58,
src/isle/meas/logdet.py:11-46).

* ``AllToAll(hfm, species, alpha=1)(stage, itr)`` - the all-to-all propagator M^-1 as a
  [nx, nt, nx, nt] device tensor indexed (sink site, sink time, source site, source
  time): solveM against every spacetime source (the identity), transposed so that the
  source is the right index.  alpha = 0 uses the spin-basis field -i phi and, as the
  reference's dense solve np.linalg.solve(M, I^T)^T does, returns (M^-1)^T in the
  spacetime indices: element [y, f, x, i] is M^-1[(i, x), (f, y)], sink and source
  exchanged relative to alpha = 1.  The last
  result is reused while (action value, trajectory point, itr) stay the same, as in the
  reference (the field itself is not compared).
* ``ChiralCondensate(seed, nsamples, hfm, species)(stage, itr)`` - records
  mean_s (r_s . M^-1 r_s) / (nx nt) (plain dot product, no conjugation) over ``nsamples``
  sources r_s = normal(0, 1, nx nt) + 0j drawn once from NumpyRNG(seed), in order, and
  reused for every configuration of the same size.

solveM's algorithm (unstabilised partial products) loses accuracy even where M is well
conditioned: 1e-7 relative at Nx=18, Nt=16 with DIA hopping, 3e-2 for the spin-basis
field.  Both solves therefore refine iteratively by default, x <- x + solveM(b - M x)
with the residual from ``core.applyM``, until the largest residual entry is below
1e-14 of the sources' or a step no longer lowers it (at most ``refine`` steps): the
result is then the exact solve to rounding.  ``refine=0`` returns solveM's result as the reference computes it.
* ``Logdet(hfm, alpha=1)(stage, itr)`` - records log det M of particles and holes.  With
  alpha = 0 (spin basis: the matrices of -i phi), where the reference calls a dense
  slogdet for stability, the log dets come from the pivoted block-cyclic elimination of
  the grouped products (precision "stable"; nx a multiple of 32) and else
  from the reference's product algorithm.

Results stay in memory (``values``: name -> list, one entry per call) and ``save(path)``
writes them as an .npz archive; the reference's buffered HDF5 measurement files are not
rebuilt.  The propagator is returned, not stored.
"""
import numpy as np
import torch

from . import _lib
from .action import HFABasis, HFAHopping, HubbardFermiAction
from .core import Species, applyM, as_field, logdetM, solveM
from .evolver import NumpyRNG


class _Measurement:
    NAMES = ()

    def __init__(self):
        self.values = {name: [] for name in self.NAMES}

    def save(self, path):
        np.savez(path, **{name: np.asarray(vals) for name, vals in self.values.items()})


def solve_refined(hfm, phi, species, rhss, refine=8, tol=1e-14, reuse_factor=False):
    """solveM followed by iterative refinement: at most ``refine`` steps, ending once the
    largest entry of the residual b - M x is below ``tol`` times the largest entry of b or
    when a step no longer lowers it (the rounding floor).  ``reuse_factor``: the previous
    solveM on ``hfm`` had this field and species (its factorization is reused)."""
    x = solveM(hfm, phi, species, rhss, reuse_factor=reuse_factor)
    if refine <= 0:
        return x
    goal = tol * rhss.abs().max().item()
    r = rhss - applyM(hfm, phi, species, x)
    res = r.abs().max().item()
    for _ in range(refine):
        if res <= goal:
            break
        x_new = x + solveM(hfm, phi, species, r, reuse_factor=True)  # same field: same factorization
        r_new = rhss - applyM(hfm, phi, species, x_new)
        res_new = r_new.abs().max().item()
        if not res_new < res:
            break
        x, r, res = x_new, r_new, res_new
    return x


def _nt(phi, nx):
    n = phi.shape[-1]
    if n % nx != 0:
        raise ValueError("Invalid field configuration: does not partition into spatial slices of size nx")
    return n // nx


class AllToAll:
    """All-to-all propagator of one species (propagator.py:12-64)."""

    # sources solved together: the solve and refinement hold about 12 buffers of
    # (sources x N) complex128, kept near 32 GiB
    WORKSPACE = 32 << 30

    def __init__(self, hfm, species, alpha=1, refine=8, chunk=None):
        if alpha not in (0, 1):
            raise ValueError("alpha must be 0 or 1")
        self.hfm = hfm
        self.nx = hfm.nx()
        self.species = species
        self._alpha = alpha
        self.refine = refine
        self.chunk = chunk
        self._key = None
        self._value = None

    def __call__(self, stage, itr):
        key = (complex(stage.actVal), int(stage.trajPoint), itr)
        if key == self._key:
            return self._value
        phi = as_field(stage.phi)
        nt = _nt(phi, self.nx)
        n = self.nx * nt
        if self._alpha == 0:
            phi = -1j * phi
        step = self.chunk or max(1, min(n, self.WORKSPACE // (12 * 16 * n)))
        res = torch.empty((n, n), dtype=torch.complex128, device=phi.device)
        for c0 in range(0, n, step):  # row i of res: M^-1 applied to source (unit vector) i
            c1 = min(n, c0 + step)
            sources = torch.zeros((c1 - c0, n), dtype=torch.complex128, device=phi.device)
            sources[torch.arange(c1 - c0), torch.arange(c0, c1)] = 1
            res[c0:c1] = solve_refined(self.hfm, phi, self.species, sources, self.refine, reuse_factor=c0 > 0)
            del sources
        # res[i, j] = (M^-1)[j, i].  alpha = 1: transpose so the source is the right index
        # (propagator.py:51-59).  alpha = 0: the reference solves densely,
        # np.linalg.solve(M, I^T)^T = (M^-1)^T (propagator.py:60-61), which is res as it is.
        if self._alpha == 1:
            res = res.transpose(0, 1)
        # split both spacetime indices (time-major) into (site, time)
        prop = res.reshape(nt, self.nx, nt, self.nx).permute(1, 0, 3, 2)
        self._key, self._value = key, prop
        return prop


class ChiralCondensate(_Measurement):
    """Stochastic chiral condensate of one species (chiralCondensate.py:11-58)."""

    NAMES = ("chiCon",)

    def __init__(self, seed, nsamples, hfm, species, refine=8):
        super().__init__()
        self.refine = refine
        self.nsamples = int(nsamples)
        self.rng = NumpyRNG(seed)
        self.hfm = hfm
        self.species = species
        self._rhss = None

    def _getRHSs(self, nt):
        n = nt * self.hfm.nx()
        if self._rhss is None or self._rhss.shape[1] != n:
            rows = [self.rng.normal(0, 1, n) + 0j for _ in range(self.nsamples)]
            self._rhss = torch.as_tensor(np.array(rows), device=_lib.device())
        return self._rhss

    def __call__(self, stage, itr):
        phi = as_field(stage.phi)
        nx = self.hfm.nx()
        nt = _nt(phi, nx)
        rhss = self._getRHSs(nt)
        res = solve_refined(self.hfm, phi, self.species, rhss, self.refine) / (nx * nt)
        value = complex((rhss * res).sum(dim=1).mean())
        self.values["chiCon"].append(value)
        return value


class Logdet(_Measurement):
    """log det M of particles and holes (logdet.py:11-46)."""

    NAMES = ("particles", "holes")

    def __init__(self, hfm, alpha=1):
        if alpha not in (0, 1):
            raise ValueError("alpha must be 0 or 1")
        super().__init__()
        self.hfm = hfm
        self.alpha = alpha
        self._spin = None

    def _spin_action(self):
        if self._spin is None:
            nx = self.hfm.nx()
            stable = nx % 32 == 0 and nx <= 16351
            self._spin = HubbardFermiAction(self.hfm.kappaTilde(), self.hfm.muTilde(), self.hfm.sigmaKappa(),
                                            HFAHopping(self.hfm.HOPPING),
                                            HFABasis.SPIN, precision="stable" if stable else "fp64")
        return self._spin

    def __call__(self, stage, itr):
        if self.alpha == 1:
            lp = logdetM(self.hfm, stage.phi, Species.PARTICLE)
            lh = logdetM(self.hfm, stage.phi, Species.HOLE)
        else:
            lp, lh = (complex(v) for v in self._spin_action().logdets(as_field(stage.phi))[0])
        self.values["particles"].append(lp)
        self.values["holes"].append(lh)
        return lp, lh
