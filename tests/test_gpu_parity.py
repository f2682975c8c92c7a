"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on identical inputs.

Tolerances (stated per test): logdet / eval / plain FP64 force 1e-10 relative
(imaginary parts compared modulo 2*pi); mixed-precision force 1e-5 relative (L2).
"""
import itertools
import os

import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.conftest import load_lattice
from tests.helpers import mod2pi_diff, random_phi, rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402

LATS = ("two_sites", "triangle", "tube_3-3_1", "one_site")


def _dev(x):
    return torch.as_tensor(x, device="cuda")


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_logdetM_vs_oracle(hop):
    rng = np.random.RandomState(8613)
    cls_g = isle.HubbardFermiMatrixExp if hop == ref.EXP else isle.HubbardFermiMatrixDia
    cls_r = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    for name in LATS:
        kappa = load_lattice(name)
        nx = kappa.shape[0]
        for nt, beta, mu, sk in itertools.product((2, 4, 8), (3,), (0.0, 0.5), (-1, 1)):
            kt, mt = kappa * beta / nt, mu * beta / nt
            g, r = cls_g(kt, mt, sk), cls_r(kt, mt, sk)
            for species, (re_, im_) in itertools.product((0, 1), ((True, False), (False, True), (True, True))):
                phi = random_phi(rng, nx * nt, re_, im_)
                a = isle.logdetM(g, phi, isle.Species(species))
                b = ref.logdetM(r, phi, species)
                assert mod2pi_diff(a, b) < 1e-10 * max(1, abs(b)), (name, nt, mu, sk, species, a, b)


@pytest.mark.parametrize("hop,basis", list(itertools.product((0, 1), (0, 1))))
def test_action_plain_vs_oracle(hop, basis):
    rng = np.random.RandomState(8613)
    for name in LATS:
        kappa = load_lattice(name)
        nx = kappa.shape[0]
        for nt, beta, sk in itertools.product((8, 16), (1, 4), (-1, 1)):
            kt = kappa * beta / nt
            g = act.HubbardFermiAction(kt, 0.0, sk, hop, basis)
            r = ref.HubbardFermiAction(kt, 0.0, sk, hop, basis)
            phi = random_phi(rng, nx * nt)
            assert mod2pi_diff(g.eval(phi), r.eval(phi)) < 1e-10 * max(1, abs(r.eval(phi))), (name, nt, beta, sk)
            fg = g.force(phi).cpu().numpy()
            fr = r.force(phi)
            assert np.max(np.abs(fg - fr)) < 1e-10 * max(1, np.max(np.abs(fr))), (name, nt, beta, sk)


@pytest.mark.parametrize("basis", [0, 1])
def test_shortcut_and_mu(basis):
    rng = np.random.RandomState(5)
    kappa = ref.honeycomb(3)
    nt, beta = 8, 3.0
    for mu, short in ((0.0, True), (0.3, False)):
        g = act.HubbardFermiAction(kappa * beta / nt, mu * beta / nt, 1, 1, basis, 0, short)
        r = ref.HubbardFermiAction(kappa * beta / nt, mu * beta / nt, 1, 1, basis, 0, short)
        phi = rng.normal(0, 0.4, 18 * nt) + 0j
        assert mod2pi_diff(g.eval(phi), r.eval(phi)) < 1e-10 * max(1, abs(r.eval(phi)))
        assert rel_l2(g.force(phi).cpu().numpy(), r.force(phi)) < 1e-10


def test_F_and_slice_matrices():
    kappa = ref.honeycomb(3) * 0.25
    phi = np.random.RandomState(2).normal(0, 0.3, 18 * 4) + 1j * np.random.RandomState(3).normal(0, 0.1, 18 * 4)
    for hop in (ref.EXP, ref.DIA):
        g = (isle.HubbardFermiMatrixExp if hop == ref.EXP else isle.HubbardFermiMatrixDia)(kappa, 0.1, -1)
        r = (ref.HFMExp if hop == ref.EXP else ref.HFMDia)(kappa, 0.1, -1)
        for tp, s, inv in itertools.product((0, 2), (0, 1), (False, True)):
            np.testing.assert_allclose(g.F(tp, phi, s, inv).cpu().numpy(), r.F(tp, phi, s, inv), rtol=1e-13, atol=1e-14)
        for s in (0, 1):
            np.testing.assert_allclose(g.K(s), r.K(s), rtol=1e-14, atol=1e-15)
        if hop == ref.EXP:
            for s, inv in itertools.product((0, 1), (False, True)):
                np.testing.assert_allclose(g.expKappa(s, inv), r.exp_kappa[(s, inv)], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(g.M(phi, 0).cpu().numpy(), r.M(phi, 0), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_solveM(hop):
    rng = np.random.RandomState(8613)
    kappa = load_lattice("tube_3-3_1")
    nx = kappa.shape[0]
    for nt, sk, species in itertools.product((4, 8), (-1, 1), (0, 1)):
        g = (isle.HubbardFermiMatrixExp if hop == ref.EXP else isle.HubbardFermiMatrixDia)(kappa / nt, 0.0, sk)
        r = (ref.HFMExp if hop == ref.EXP else ref.HFMDia)(kappa / nt, 0.0, sk)
        phi = random_phi(rng, nx * nt)
        rhss = np.array([random_phi(rng, nx * nt) for _ in range(10)])
        xg = isle.solveM(g, phi, species, rhss).cpu().numpy()
        np.testing.assert_allclose(xg, ref.solveM(r, phi, species, rhss), rtol=1e-10, atol=1e-12)
        m = r.M(phi, species)
        np.testing.assert_allclose(m @ xg.T, rhss.T, rtol=1e-10, atol=1e-12)


def _pivot_heavy(rng, n):
    """P L U with well-conditioned triangular factors (|re| + |im| of the entries of the unit
    lower L below 1) and a random row permutation P: partial pivoting then swaps almost
    every row, many of them with rows several panels further down."""
    s = 0.5 / np.sqrt(n)
    lo = np.tril(rng.uniform(-s, s, (n, n)) + 1j * rng.uniform(-s, s, (n, n)), -1) + np.eye(n)
    up = np.triu(rng.uniform(-s, s, (n, n)) + 1j * rng.uniform(-s, s, (n, n)), 1) + np.diag(
        rng.uniform(1, 2, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n)))
    return (lo @ up)[rng.permutation(n)]


@pytest.mark.parametrize("n", [1, 5, 16, 31, 32, 33, 63, 64, 65, 100, 288, 300, 384, 512])
def test_logdet_and_invert(n):
    """logdet (1e-11 relative, phase mod 2pi) and inverse (1e-10 relative Frobenius) against
    LAPACK through numpy/scipy, on Gaussian and pivot-heavy matrices; the sizes cover one
    partial panel, exact panel multiples, one row past a panel, and the largest order."""
    rng = np.random.RandomState(7 + n)
    mats = [rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)), _pivot_heavy(rng, n)]
    if n > 1:
        z = mats[0].copy()
        z[0, 0] = 0.0  # zero leading pivot
        mats.append(z)
    for a in mats:
        want = ref.ilogdet(a)
        assert mod2pi_diff(isle.logdet(a), want) < 1e-11 * max(1, abs(want))
        inv = np.linalg.inv(a)
        got = isle.invert(a).cpu().numpy()
        assert np.linalg.norm(got - inv) / np.linalg.norm(inv) < 1e-10
        assert np.linalg.norm(a @ got - np.eye(n)) / np.sqrt(n) < 1e-10


@pytest.mark.parametrize("cfg", [(3, 16, 4.0), (6, 64, 6.0)])
def test_mixed_force(cfg):
    """FP32 local chains + FP64 top level within 1e-5 (L2 relative) of the FP64 oracle."""
    l, nt, beta = cfg
    kappa = ref.honeycomb(l)
    nx = kappa.shape[0]
    u = 3.0 if l == 6 else 2.0
    rng = np.random.RandomState(42)
    kt = kappa * beta / nt
    r = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP)
    m = act.HubbardFermiAction(kt, 0.0, -1, 1, precision="mixed", group=8)
    for _ in range(2):
        phi = rng.normal(0, np.sqrt(u * beta / nt), nx * nt) + 0j
        err = rel_l2(m.force(phi).cpu().numpy(), r.force(phi))
        assert err < 1e-5, err
        assert mod2pi_diff(m.eval(phi), r.eval(phi)) < 1e-10 * abs(r.eval(phi))


def test_mixed_force_config3_hard_fields():
    """Nx=288, Nt=128, beta=8, U=3: the mixed force within 1e-5 (L2 relative) of the plain
    FP64 GPU force (itself pinned to the oracle) on start fields N(0, U beta/Nt) with seeds
    1000 and 1002, which lie near zeros of det M (|F| ~ 6x typical) where rounding in the
    group products would be amplified."""
    kappa = ref.honeycomb(12)
    nx, nt, beta, u = 288, 128, 8.0, 3.0
    kt, ut = kappa * beta / nt, u * beta / nt
    p = act.HubbardFermiAction(kt, 0.0, -1, 1)
    m = act.HubbardFermiAction(kt, 0.0, -1, 1, precision="mixed")
    phis = _dev(np.stack([np.random.RandomState(s).normal(0, np.sqrt(ut), nx * nt) + 0j for s in (1000, 1002)]))
    fp = p.force_batch(phis).cpu().numpy()
    fm = m.force_batch(phis).cpu().numpy()
    for b in range(2):
        assert rel_l2(fm[b], fp[b]) < 1e-5, (b, rel_l2(fm[b], fp[b]))


def test_batched_force_matches_single():
    kappa = ref.honeycomb(3)
    nt, beta = 16, 4.0
    rng = np.random.RandomState(3)
    phis = rng.normal(0, 0.7, (5, 18 * nt)) + 0j
    for prec in ("fp64", "mixed"):
        a = act.HubbardFermiAction(kappa * beta / nt, 0.0, -1, 1, precision=prec)
        fb = a.force_batch(_dev(phis)).cpu().numpy()
        eb = a.eval_batch(_dev(phis))
        for b in range(5):
            np.testing.assert_allclose(fb[b], a.force(phis[b]).cpu().numpy(), rtol=1e-12, atol=1e-13)
            assert abs(eb[b] - a.eval(phis[b])) < 1e-11 * abs(eb[b])


def test_batched_action_eval_matches_single():
    """Gauge + fermion action of a batch of 3 fields (the batched evaluation the evolvers
    use) equals the per-field evaluation and the oracle, 1e-12 relative."""
    from isle_amd.integrator import eval_batch
    kappa = ref.honeycomb(3)
    nx, nt, beta, u = 18, 16, 4.0, 2.0
    kt, ut = kappa * beta / nt, u * beta / nt
    full = act.SumAction(act.HubbardGaugeAction(ut), act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP))
    r = ref.SumAction(ref.HubbardGaugeAction(ut), ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP))
    rng = np.random.RandomState(12)
    phis = rng.normal(0, np.sqrt(ut), (3, nx * nt)) + 1j * rng.normal(0, 0.1, (3, nx * nt))
    got = eval_batch(full, _dev(phis))
    for b in range(3):
        want = r.eval(phis[b])
        assert mod2pi_diff(got[b], want) < 1e-12 * abs(want)
        assert mod2pi_diff(full.eval(_dev(phis[b])), want) < 1e-12 * abs(want)


def test_leapfrog_vs_oracle():
    kappa = ref.honeycomb(3)
    nt, beta, u = 16, 4.0, 2.0
    ut = u * beta / nt
    rng = np.random.RandomState(1075)
    phi = rng.normal(0, np.sqrt(ut), 18 * nt) + 0j
    pi = rng.normal(0, 1, 18 * nt) + 0j
    g = act.SumAction(act.HubbardGaugeAction(ut), act.HubbardFermiAction(kappa * beta / nt, 0.0, -1, 1))
    r = ref.SumAction(ref.HubbardGaugeAction(ut), ref.HubbardFermiAction(kappa * beta / nt, 0.0, -1, ref.EXP))
    p1, q1, s1 = isle.leapfrog(phi, pi, g, 1.0, 20)
    rp, rq, rs = ref.leapfrog(phi, pi, r, 1.0, 20)
    assert rel_l2(p1.cpu().numpy(), rp) < 1e-11
    assert rel_l2(q1.cpu().numpy(), rq) < 1e-11
    assert mod2pi_diff(s1, rs) < 1e-10 * abs(rs)
    # generic (unfused) path gives the same
    class Wrapped(act.Action):
        def __init__(self, a):
            self.a = a

        def eval(self, phi):
            return self.a.eval(phi)

        def force(self, phi):
            return self.a.force(phi)

    p2, q2, s2 = isle.leapfrog(phi, pi, Wrapped(g), 1.0, 20)
    assert rel_l2(p2.cpu().numpy(), rp) < 1e-11


def test_hmc_fixed_seed_matches_oracle():
    """Same seed, same momenta, same Metropolis draws: Delta H and decisions agree (config 1 shape)."""
    from isle_amd.evolver import ConstStepLeapfrog, EvolutionStage, NumpyRNG
    kappa = ref.honeycomb(3)
    nt, beta, u = 16, 4.0, 2.0
    ut = u * beta / nt
    g = act.SumAction(act.HubbardGaugeAction(ut), act.HubbardFermiAction(kappa * beta / nt, 0.0, -1, 1))
    r = ref.SumAction(ref.HubbardGaugeAction(ut), ref.HubbardFermiAction(kappa * beta / nt, 0.0, -1, ref.EXP))
    seed = 1075
    rng_g, rng_r = NumpyRNG(seed), np.random.RandomState(seed)
    phi0 = rng_r.normal(0, np.sqrt(ut), 18 * nt) + 0j
    rng_g.normal(0, np.sqrt(ut), 18 * nt)
    ev = ConstStepLeapfrog(g, 1.0, 20, rng_g)
    stage = EvolutionStage(phi0, g.eval(phi0))
    phi_r, act_r = phi0, r.eval(phi0)
    for _ in range(4):
        stage = ev.evolve(stage)
        phi_r, act_r, tp, e0, e1, _ = ref.const_step_leapfrog(rng_r, r, phi_r, act_r, 1.0, 20)
        assert tp == stage.trajPoint
        dh_g = np.real(ev.last_energies[1] - ev.last_energies[0])
        assert abs(dh_g - np.real(e1 - e0)) < 1e-8
    assert rel_l2(stage.phi.cpu().numpy(), phi_r) < 1e-10


def test_config2_gradient_check():
    """Nx=72, Nt=64, beta=6, U=3: the gradient of the action through torch.autograd (its
    backward is the force sweep) vs central FD (h=1e-6) on 64 entries, and the force vs
    the oracle force on 16 seeded fields.  The h = 1e-6 differences are taken twice: of the
    GPU's own FP64 action (round-off limited, see below) and, from the committed fixture
    tests/golden/fd_config2_h1e-6.npz, of the same action formula in extended precision
    (accurate to its O(h^2) truncation, ~1e-8): the force agrees with the latter within 3e-8
    absolute (measured 9.8e-9 on forces up to 13)."""
    kappa = ref.honeycomb(6)
    nx, nt, beta, u = 72, 64, 6.0, 3.0
    kt = kappa * beta / nt
    ut = u * beta / nt
    g = act.HubbardFermiAction(kt, 0.0, -1, 1)
    r = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP)
    rng = np.random.RandomState(2000)
    phi = rng.normal(0, np.sqrt(ut), nx * nt) + 0j
    # dS/dphi from autograd on a complex leaf: grad = conj(dS/dphi) for L = Re S
    leaf = torch.as_tensor(phi, device="cuda").requires_grad_(True)
    s_val = isle.autograd.action_value(g, leaf)
    s_val.real.backward()
    f = -leaf.grad.conj().cpu().numpy()  # force = -dS/dphi
    assert rel_l2(f, g.force(phi).cpu().numpy()) < 1e-14
    s0 = abs(g.eval(phi))
    idx = rng.choice(nx * nt, 64, replace=False)
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fd_config2_h1e-6.npz"))
    assert np.array_equal(golden["idx"], idx)
    err_golden = np.abs(f[idx] - golden["fd"]).max()
    print(f"config 2 FD (extended precision, h = 1e-6): max |FD - F| = {err_golden:.2e}", flush=True)
    assert err_golden < 3e-8
    # Central differences: truncation O(h^2) plus round-off |S| * eta / h with eta ~ 1e-13 the
    # relative accuracy of the FP64 action.  h = 1e-6 (the prescribed step) is round-off
    # limited: the bound 5 |S| 1e-13 / 1e-6 = 1.7e-4 at |S| = 350, measured 8.3e-5 (the action
    # carries ~2.4e-13 relative rounding).  The tight check extrapolates h = 2e-4 and 1e-4
    # (Richardson, truncation O(h^4)), leaving the round-off bound 5 |S| 1e-13 / 1e-4 (1.7e-6;
    # measured 1.1e-6; the truncation of the plain h = 1e-4 difference alone reaches 1.7e-6).
    def cdiff(i, h):
        e = np.zeros(nx * nt, dtype=complex)
        e[i] = h
        d = g.eval(phi + e) - g.eval(phi - e)
        return -complex(d.real, np.remainder(d.imag + np.pi, 2 * np.pi) - np.pi) / (2 * h)
    worst_h6 = worst_rich = 0.0
    for i in idx:
        assert abs(cdiff(i, 1e-6) - f[i]) < 5 * s0 * 1e-13 / 1e-6, (i, cdiff(i, 1e-6), f[i])
        rich = (4 * cdiff(i, 1e-4) - cdiff(i, 2e-4)) / 3
        assert abs(rich - f[i]) < 5 * s0 * 1e-13 / 1e-4, (i, rich, f[i])
        worst_h6 = max(worst_h6, abs(cdiff(i, 1e-6) - f[i]))
        worst_rich = max(worst_rich, abs(rich - f[i]))
    print(f"config 2 FD: |S| = {s0:.1f}, max |FD - F| at h = 1e-6: {worst_h6:.2e}, Richardson h = 1e-4: "
          f"{worst_rich:.2e}, max |F| = {np.abs(f[idx]).max():.2e}", flush=True)
    for s in range(2000, 2016):
        phi = np.random.RandomState(s).normal(0, np.sqrt(ut), nx * nt) + 0j
        assert rel_l2(g.force(phi).cpu().numpy(), r.force(phi)) < 1e-10


@pytest.mark.parametrize("precision,tol_phi,tol_pi,tol_s", [("fp64", 1e-7, 1e-6, 1e-8), ("mixed", 1e-6, 1e-5, 1e-7)])
def test_config3_leapfrog_start_matches_oracle(precision, tol_phi, tol_pi, tol_s):
    """Nx=288, Nt=128, beta=8, U=3 (the bench workload), chain 0's start field and momenta
    (seed 1000): the first 4 of the 40 leapfrog steps (eps = 1/40) against the oracle.
    Measured: FP64 phi 4e-9, pi 5e-8, S 1e-10; mixed phi 4e-8, pi 6e-7, S 1e-9.  (Near a zero
    of det M two FP64 evaluations of the reference algorithm with different summation orders
    already differ by ~1e-8 in the force; later in the trajectory the difference grows
    exponentially -- 3e-3 in pi after 8 steps, O(1) after 24 -- so whole trajectories from
    these starts are compared through their accept decisions, not their Delta H.)"""
    from isle_amd.integrator import leapfrog
    kappa = ref.honeycomb(12)
    nx, nt, beta, u = 288, 128, 8.0, 3.0
    kt, ut = kappa * beta / nt, u * beta / nt
    rng = np.random.RandomState(1000)
    phi = rng.normal(0, np.sqrt(ut), nx * nt) + 0j
    pi = rng.normal(0, 1, nx * nt) + 0j
    r = ref.SumAction(ref.HubbardGaugeAction(ut), ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP))
    rphi, rpi, rs = ref.leapfrog(phi, pi, r, 4 / 40.0, 4)
    g = act.SumAction(act.HubbardGaugeAction(ut),
                      act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision))
    gphi, gpi, gs = leapfrog(_dev(phi), _dev(pi), g, 4 / 40.0, 4)
    assert rel_l2(gphi.cpu().numpy(), rphi) < tol_phi
    assert rel_l2(gpi.cpu().numpy(), rpi) < tol_pi
    assert abs(gs - rs) < tol_s * abs(rs)


@pytest.mark.parametrize("hop,basis,mu", [(ref.EXP, ref.PARTICLE_HOLE, 0.0), (ref.EXP, ref.SPIN, 0.3),
                                          (ref.DIA, ref.PARTICLE_HOLE, 0.2), (ref.DIA, ref.SPIN, 0.5)])
def test_autograd_logdetM_and_action(hop, basis, mu):
    """torch.autograd through logdetM (per species) and the action, real and complex
    fields: gradients against the oracle's derivative rule and central differences of the
    GPU logdetM; 1e-9 relative."""
    kappa = ref.honeycomb(3)
    nx, nt, beta = 18, 16, 4.0
    kt, mt = kappa * beta / nt, mu * beta / nt
    g = act.HubbardFermiAction(kt, mt, -1, hop, basis)
    r = ref.HubbardFermiAction(kt, mt, -1, hop, basis)
    rng = np.random.RandomState(31)
    phi = rng.normal(0, 0.4, nx * nt) + 1j * rng.normal(0, 0.1, nx * nt)
    spin = basis == ref.SPIN
    arg = (lambda p: -1j * p) if spin else (lambda p: p)
    for species, k in ((0, r.kp), (1, r.kh)):
        leaf = torch.as_tensor(phi, device="cuda").requires_grad_(True)
        ld = isle.autograd.logdetM_value(g, leaf, species)
        assert mod2pi_diff(complex(ld.item()), ref.logdetM(r.hfm, arg(phi), species)) < 1e-10 * max(1, abs(ld.item()))
        ld.real.backward()
        d = leaf.grad.conj().cpu().numpy()  # d logdetM / dphi
        coef = (-1.0 if species == 0 else 1.0) if spin else (-1j if species == 0 else 1j)
        want = coef * (ref.force_direct_single_part(r.hfm, arg(phi), k, species) - 1.0)
        assert rel_l2(d, want) < 1e-9
        # central differences: truncation O(h^2) plus round-off |logdet| * 1e-13 / h
        h = 1e-4
        tol_fd = 1e-7 + 5 * abs(ld.item()) * 1e-13 / h
        for i in rng.choice(nx * nt, 4, replace=False):
            e = np.zeros(nx * nt, dtype=complex)
            e[i] = h
            fd = (isle.autograd.logdetM_value(g, _dev(phi + e), species).item()
                  - isle.autograd.logdetM_value(g, _dev(phi - e), species).item())
            fd = complex(fd.real, np.remainder(fd.imag + np.pi, 2 * np.pi) - np.pi) / (2 * h)
            assert abs(fd - d[i]) < tol_fd * max(1, abs(d[i])), (i, fd, d[i])
    # action on a real field (the HMC case): grad = -Re(force)
    x = torch.as_tensor(phi.real.copy(), device="cuda").requires_grad_(True)
    isle.autograd.action_value(g, x).real.backward()
    want = -np.real(r.force(phi.real + 0j))
    assert rel_l2(x.grad.cpu().numpy(), want) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_trajectory_replays(precision):
    """A leapfrog trajectory captured in a HIP graph and replayed three times, with eager
    work and synchronisations between the replays, equals the eager trajectory on the same
    inputs each time (1e-13 relative); the batched evolver gives the same accept/reject
    decisions and Delta H with and without the graph."""
    from isle_amd.evolver import BatchedConstStepLeapfrog, NumpyRNG
    from isle_amd.integrator import LeapfrogGraph, leapfrog_batch
    kappa = ref.honeycomb(3)
    nx, nt, beta, u = 18, 16, 4.0, 2.0
    kt, ut = kappa * beta / nt, u * beta / nt
    full = act.SumAction(act.HubbardGaugeAction(ut),
                         act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision))
    rng = np.random.RandomState(11)
    batch = 3
    lf = LeapfrogGraph(full, batch, nx * nt, 1.0, 6)
    for rep in range(3):
        phis = _dev(rng.normal(0, np.sqrt(ut), (batch, nx * nt)) + 0j)
        pis = _dev(rng.normal(0, 1, (batch, nx * nt)) + 0j)
        full[1].force_batch(_dev(rng.normal(0, 1, (2, nx * nt)) + 0j))  # unrelated eager work
        torch.cuda.synchronize()
        g_phi, g_pi = lf(phis, pis)
        torch.cuda.synchronize()
        e_phi, e_pi, _ = leapfrog_batch(phis, pis, full, 1.0, 6, evaluate=False)
        assert rel_l2(g_phi.cpu().numpy(), e_phi.cpu().numpy()) < 1e-13
        assert rel_l2(g_pi.cpu().numpy(), e_pi.cpu().numpy()) < 1e-13
    runs = []
    for graph in (False, True):
        rngs = [NumpyRNG(500 + b) for b in range(batch)]
        ev = BatchedConstStepLeapfrog(full, 1.0, 6, rngs, graph=graph)
        phis = _dev(np.stack([r.normal(0, np.sqrt(ut), nx * nt) + 0j for r in rngs]))
        acts = np.array([full.eval(phis[b]) for b in range(batch)])
        for _ in range(3):
            phis, acts, _ = ev.evolve(phis, acts)
        runs.append((ev.trajPoints, np.array(ev.deltaH)))
    assert runs[0][0] == runs[1][0]
    np.testing.assert_allclose(runs[0][1], runs[1][1], rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-4)])
def test_graph_split_over_streams(precision, tol):
    """A batch of 5 chains captured as 2 parts on streams of their own (each part with its
    own action and workspaces) equals the eager trajectory of the whole batch, on two
    replays with eager work between them (different plans, so agreement to the rounding of
    the precision, not bit-for-bit)."""
    from isle_amd.integrator import LeapfrogGraph, leapfrog_batch
    kappa = ref.honeycomb(3)
    nx, nt, beta, u = 18, 16, 4.0, 2.0
    kt, ut = kappa * beta / nt, u * beta / nt

    def make():
        return act.SumAction(act.HubbardGaugeAction(ut),
                             act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision))
    full = make()
    batch = 5
    lf = LeapfrogGraph(full, batch, nx * nt, 1.0, 6, split=[make(), make()])
    rng = np.random.RandomState(12)
    for rep in range(2):
        phis = _dev(rng.normal(0, np.sqrt(ut), (batch, nx * nt)) + 0j)
        pis = _dev(rng.normal(0, 1, (batch, nx * nt)) + 0j)
        full[1].force_batch(_dev(rng.normal(0, 1, (2, nx * nt)) + 0j))  # unrelated eager work
        torch.cuda.synchronize()
        g_phi, g_pi = lf(phis, pis)
        torch.cuda.synchronize()
        e_phi, e_pi, _ = leapfrog_batch(phis, pis, full, 1.0, 6, evaluate=False)
        assert rel_l2(g_phi.cpu().numpy(), e_phi.cpu().numpy()) < tol
        assert rel_l2(g_pi.cpu().numpy(), e_pi.cpu().numpy()) < tol


@pytest.mark.parametrize("merge", [True, False])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 5e-5)])
def test_ensemble_uneven_chains_match_single_runs(precision, tol, merge):
    """Chains of uneven (Nx, Nt) advanced together (one HIP graph per trajectory; with merge
    the two time extents of each lattice share their force launches on one stream, without
    it every bucket has its own stream) reproduce each chain run alone with the same seed:
    identical accept/reject decisions and Delta H within `tol` (absolute + relative) over 3
    trajectories; small sizes of the config-4 family (L in {3, 4}, Nt in {16, 32},
    beta = Nt/8, U = 2).  Batches of different size may run different tile shapes, so
    the mixed form's FP32 rounding differs between the runs (Delta H to ~1e-5)."""
    from isle_amd.ensemble import Chain, EnsembleLeapfrog
    from isle_amd.evolver import BatchedConstStepLeapfrog, NumpyRNG
    specs = [(3, 16), (4, 32), (3, 32), (4, 16), (3, 16), (4, 32)]
    def chains():
        return [Chain(ref.honeycomb(l), nt, nt / 8.0, 2.0, 0.0, -1, NumpyRNG(900 + i)) for i, (l, nt) in enumerate(specs)]
    print(f"ensemble vs single chains: {precision}, merge={merge}", flush=True)
    ens = EnsembleLeapfrog(chains(), 1.0, 8, precision=precision, group=32, graph=True, merge=merge)
    # merged: one stream per lattice, the Nx=32 lattice (87% of the work) split over two
    assert len(ens.groups) == (3 if merge else 4)
    for _ in range(3):
        ens.evolve()
    for i, c in enumerate(chains()):
        full = act.SumAction(act.HubbardGaugeAction(c.utilde),
                             act.HubbardFermiAction(c.kappa * c.beta / c.nt, 0.0, -1, act.HFAHopping.EXP,
                                                    precision=precision, group=32))
        ev = BatchedConstStepLeapfrog(full, 1.0, 8, [c.rng])
        phi = _dev(c.phi0.reshape(1, -1))
        a = np.array([full.eval(phi[0])])
        for _ in range(3):
            phi, a, _ = ev.evolve(phi, a)
        assert [p[0] for p in ev.trajPoints] == [p[i] for p in ens.trajPoints]
        np.testing.assert_allclose([d[0] for d in ev.deltaH], [d[i] for d in ens.deltaH], rtol=tol, atol=tol)
