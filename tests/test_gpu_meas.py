"""GPU measurements built on solveM / logdetM against the oracle and the dense solve.

The reference's solveM algorithm (partial products, z-recursion, LU of each partial
product) amplifies rounding even where M itself is well conditioned: at Nx=18, Nt=16 the
oracle's solveM is off the dense inverse by 7e-11 (EXP), 2e-8..2e-7 (DIA) and
2e-5..3e-2 for the spin-basis field -i phi (alpha 0, where the reference's propagator
calls a dense solve instead).  The measurements therefore refine solveM's result
iteratively (residual from core.applyM) by default, and are then required to equal the
dense solve with the oracle's M (numpy.linalg.inv / solve, the textbook routine) to
1e-10 of the largest entry, for EXP and DIA hopping, both species and alpha 0 / 1.  With
refine=0 (solveM as the reference computes it) they must be off the dense solve by at
most 10x the oracle's own solveM error (+1e-12): as accurate as the reference's
algorithm.  applyM itself equals the oracle's dense M times the rows to 1e-13.  The
explicit element map of the propagator (prop[y, f, x, i] = M^-1[f nx + y, i nx + x] for
alpha 1; M^-1[i nx + x, f nx + y] for alpha 0, where the reference's dense solve returns
(M^-1)^T) and the reference's own formulas are checked, so an exchanged site/time or
sink/source index fails.

Small case: honeycomb 3x3 torus (Nx=18), Nt=16, beta=4, U=2, mu=0.1 (N = 288).  Log dets:
the oracle's logdetM (alpha 1) and numpy's dense slogdet of M(-i phi) (alpha 0, the
reference's own method there), to 1e-10.  Config-2 case (honeycomb 6x6, Nx=72, Nt=64,
beta=6, U=3; N = 4608): the refined propagator equals numpy's dense inverse to 1e-10 and
M G = 1 to 1e-12; unrefined, it is within 10x of the oracle's solveM error.
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import meas  # noqa: E402
from isle_amd.evolver import EvolutionStage  # noqa: E402


def _setup(hop, l=3, nt=16, beta=4.0, U=2.0, mu=0.1, seed=5):
    kappa = ref.honeycomb(l)
    nx = kappa.shape[0]
    kt, mt = kappa * beta / nt, mu * beta / nt
    phi = np.random.RandomState(seed).normal(0, np.sqrt(U * beta / nt), nx * nt) + 0j
    if hop == "exp":
        return isle.HubbardFermiMatrixExp(kt, mt, -1), ref.HFMExp(kt, mt, -1), phi, nx, nt
    return isle.HubbardFermiMatrixDia(kt, mt, -1), ref.HFMDia(kt, mt, -1), phi, nx, nt


def _as_matrix(prop, nx, nt):
    """[nx, nt, nx, nt] (sink site, sink time, source site, source time) -> N x N."""
    return prop.permute(1, 0, 3, 2).reshape(nx * nt, nx * nt)


@pytest.mark.parametrize("hop", ["exp", "dia"])
@pytest.mark.parametrize("species", [0, 1])
def test_apply_m_matches_dense(hop, species):
    hfm, rh, phi, nx, nt = _setup(hop)
    x = np.random.RandomState(2).normal(size=(3, nx * nt)) + 1j * np.random.RandomState(3).normal(size=(3, nx * nt))
    for field in (phi, -1j * phi):
        got = isle.core.applyM(hfm, field, species, x).cpu().numpy()
        want = (rh.M(field, species) @ x.T).T
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("hop", ["exp", "dia"])
@pytest.mark.parametrize("species", [0, 1])
@pytest.mark.parametrize("alpha", [1, 0])
def test_all_to_all_is_dense_inverse(hop, species, alpha):
    print(f"all-to-all {hop} species {species} alpha {alpha}, Nx=18 Nt=16", flush=True)
    hfm, rh, phi, nx, nt = _setup(hop)
    stage = EvolutionStage(phi, 0.5, 1)
    prop = meas.AllToAll(hfm, isle.Species(species), alpha)(stage, 0)
    plain = meas.AllToAll(hfm, isle.Species(species), alpha, refine=0)(stage, 0)
    assert tuple(prop.shape) == (nx, nt, nx, nt)
    field = phi if alpha == 1 else -1j * phi
    dense = rh.M(field, species)
    # the reference's formulas (propagator.py:51-66): alpha 1 transposes solveM's rows,
    # alpha 0 solves densely and transposes, which leaves (M^-1)^T
    n = nx * nt
    res = np.linalg.inv(dense) if alpha == 1 else np.linalg.solve(dense, np.eye(n).T).T
    ref_prop = res.reshape(nt, nx, nt, nx).transpose(1, 0, 3, 2)
    want = np.linalg.inv(dense) if alpha == 1 else np.linalg.inv(dense).T
    scale = np.abs(want).max()
    got = prop.cpu().numpy()
    assert np.abs(got - ref_prop).max() < 1e-10 * scale
    y, f, x, i = 5, 3, 11, 9
    inv = np.linalg.inv(dense)
    sink, source = f * nx + y, i * nx + x
    assert abs(got[y, f, x, i] - (inv[sink, source] if alpha == 1 else inv[source, sink])) < 1e-10 * scale
    err = np.abs(_as_matrix(prop, nx, nt).cpu().numpy() - want).max() / scale
    err0 = np.abs(_as_matrix(plain, nx, nt).cpu().numpy() - want).max() / scale
    ref_sol = ref.solveM(rh, field, species, np.eye(nx * nt)).T
    ref_err = np.abs((ref_sol if alpha == 1 else ref_sol.T) - want).max() / scale
    print(f"  vs dense inverse: refined {err:.2e}, unrefined {err0:.2e} (oracle solveM {ref_err:.2e})", flush=True)
    assert err < 1e-10
    assert err0 <= 10 * ref_err + 1e-12


@pytest.mark.parametrize("hop", ["exp", "dia"])
@pytest.mark.parametrize("nrhs", [10, 100])
def test_solveM_paths_nx32(hop, nrhs):
    """Nx=32 (honeycomb 4x4), Nt=16: with 64 or more right-hand sides solveM runs its
    triangular solves as blocked GEMMs, with fewer in one kernel per matrix.  Either is the
    reference's algorithm: off the dense solve by at most 10x the oracle's solveM error."""
    hfm, rh, phi, nx, nt = _setup(hop, l=4)
    b = np.random.RandomState(7).normal(size=(nrhs, nx * nt)) + 1j * np.random.RandomState(8).normal(size=(nrhs, nx * nt))
    for species in (0, 1):
        got = isle.solveM(hfm, phi, species, b).cpu().numpy()
        exact = np.linalg.solve(rh.M(phi, species), b.T).T
        alg = ref.solveM(rh, phi, species, b)
        scale = np.abs(exact).max()
        err, ref_err = np.abs(got - exact).max() / scale, np.abs(alg - exact).max() / scale
        assert err <= 10 * ref_err + 1e-12, (species, err, ref_err)
    prop = meas.AllToAll(hfm, isle.Species.HOLE)(EvolutionStage(phi, 0.5, 1), 0)
    want = np.linalg.inv(rh.M(phi, 1))
    assert np.abs(_as_matrix(prop, nx, nt).cpu().numpy() - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("hop", ["exp", "dia"])
def test_solveM_reuses_factorization(hop):
    """solveM(reuse_factor=True) after a solveM with the same field and species gives the
    fresh solve bit for bit; asked to reuse across a change of species it refactorizes."""
    hfm, rh, phi, nx, nt = _setup(hop, l=4)
    rs = np.random.RandomState(3)
    b1, b2 = (torch.as_tensor(rs.normal(size=(100, nx * nt)) + 1j * rs.normal(size=(100, nx * nt)), device="cuda")
              for _ in range(2))
    isle.solveM(hfm, phi, 0, b1)
    reused = isle.solveM(hfm, phi, 0, b2, reuse_factor=True)
    fresh = isle.solveM(hfm, phi, 0, b2)
    assert torch.equal(reused, fresh)
    other = isle.solveM(hfm, phi, 1, b1, reuse_factor=True)  # species changed: refactorized
    assert torch.equal(other, isle.solveM(hfm, phi, 1, b1))
    exact = np.linalg.solve(rh.M(phi, 1), b1.cpu().numpy().T).T
    alg = ref.solveM(rh, phi, 1, b1.cpu().numpy())
    scale = np.abs(exact).max()
    assert np.abs(other.cpu().numpy() - exact).max() <= 10 * np.abs(alg - exact).max() + 1e-12 * scale


def test_all_to_all_chunked():
    """Sources solved in chunks of 100 (ragged last chunk: 288 = 2 x 100 + 88) give the
    propagator of one solve."""
    hfm, _, phi, nx, nt = _setup("dia")
    st = EvolutionStage(phi, 0.5, 1)
    whole = meas.AllToAll(hfm, isle.Species.HOLE)(st, 0)
    parts = meas.AllToAll(hfm, isle.Species.HOLE, chunk=100)(st, 0)
    assert (parts - whole).abs().max().item() < 1e-12 * whole.abs().max().item()


def test_all_to_all_memoized():
    hfm, _, phi, nx, nt = _setup("exp")
    a2a = meas.AllToAll(hfm, isle.Species.PARTICLE)
    st = EvolutionStage(phi, 0.5, 1)
    p1 = a2a(st, 3)
    assert a2a(st, 3) is p1                      # same (actVal, trajPoint, itr): cached
    assert a2a(EvolutionStage(phi, 0.5, 1), 4) is not p1


@pytest.mark.parametrize("hop", ["exp", "dia"])
def test_chiral_condensate(hop):
    hfm, rh, phi, nx, nt = _setup(hop)
    cc = meas.ChiralCondensate(1234, 5, hfm, isle.Species.HOLE)
    cc0 = meas.ChiralCondensate(1234, 5, hfm, isle.Species.HOLE, refine=0)
    fields = (phi, 2 * phi)
    got = [cc(EvolutionStage(f, 0.0, 1), k) for k, f in enumerate(fields)]
    got0 = [cc0(EvolutionStage(f, 0.0, 1), k) for k, f in enumerate(fields)]
    rng = np.random.RandomState(1234)  # the sources: drawn once, in order, from NumpyRNG(seed)
    rhss = np.array([rng.normal(0, 1, nx * nt) + 0j for _ in range(5)])
    for value, value0, field in zip(got, got0, fields):
        exact = np.linalg.solve(rh.M(field, 1), rhss.T).T / (nx * nt)
        want = np.mean([np.dot(r, x) for r, x in zip(rhss, exact)])
        res = ref.solveM(rh, field, 1, rhss) / (nx * nt)
        alg = np.mean([np.dot(r, x) for r, x in zip(rhss, res)])
        assert abs(value - want) < 1e-10 * abs(want)
        assert abs(value0 - want) <= 10 * abs(alg - want) + 1e-12 * abs(want)
    assert cc.values["chiCon"] == got


@pytest.mark.parametrize("hop", ["exp", "dia"])
def test_logdet_measurement(hop, tmp_path):
    hfm, rh, phi, nx, nt = _setup(hop)
    m1, m0 = meas.Logdet(hfm, 1), meas.Logdet(hfm, 0)
    lp, lh = m1(EvolutionStage(phi, 0.0, 1), 0)
    for species, val in ((0, lp), (1, lh)):
        want = ref.logdetM(rh, phi, species)
        assert abs(np.exp(1j * (val.imag - want.imag)) - 1) + abs(val.real - want.real) < 1e-10 * abs(want)
    lp0, lh0 = m0(EvolutionStage(phi, 0.0, 1), 0)
    for species, val in ((0, lp0), (1, lh0)):
        sign, logabs = np.linalg.slogdet(rh.M(-1j * phi, species))
        want = np.log(sign) + logabs
        assert abs(np.exp(1j * (val.imag - want.imag)) - 1) + abs(val.real - want.real) < 1e-10 * abs(want)
    m1.save(tmp_path / "ld.npz")
    with np.load(tmp_path / "ld.npz") as f:
        assert f["particles"][0] == lp and f["holes"][0] == lh


def test_all_to_all_config2():
    print("all-to-all exp, Nx=72 Nt=64 (N=4608)", flush=True)
    hfm, rh, phi, nx, nt = _setup("exp", l=6, nt=64, beta=6.0, U=3.0, mu=0.0, seed=16)
    stage = EvolutionStage(phi, 0.0, 1)
    g = _as_matrix(meas.AllToAll(hfm, isle.Species.PARTICLE)(stage, 0), nx, nt)
    g0 = _as_matrix(meas.AllToAll(hfm, isle.Species.PARTICLE, refine=0)(stage, 0), nx, nt)
    m = rh.M(phi, 0)
    inv = np.linalg.inv(m)
    scale = np.abs(inv).max()
    err = np.abs(g.cpu().numpy() - inv).max() / scale
    err0 = np.abs(g0.cpu().numpy() - inv).max() / scale
    ref_err = np.abs(ref.solveM(rh, phi, 0, np.eye(nx * nt)).T - inv).max() / scale
    md = torch.as_tensor(m, device="cuda")
    resid = (md @ g - torch.eye(nx * nt, dtype=torch.complex128, device="cuda")).abs().max().item()
    print(f"vs dense inverse: refined {err:.2e}, unrefined {err0:.2e} (oracle solveM {ref_err:.2e}); "
          f"max |MG - 1| {resid:.2e}", flush=True)
    assert err < 1e-10 and resid < 1e-12
    assert err0 <= 10 * ref_err + 1e-12
