"""Pins the oracle's actions, forces and leapfrog.

* DIRECT_SINGLE vs DIRECT_SQUARE (dense Q^{-1}) eval and force, all hopping x basis x sigmaKappa
  (the reference's tests/test_hubbardFermiAction.py:52-77,122-147);
* shortcut identities for real phi (tests/test_hubbardFermiAction.py:79-109,149-179);
* force == -dS/dphi by central finite differences (closes the sign/index of the force);
* SumAction of a dummy action and the gauge action (tests/test_sumaction.py:153-182);
* stabilised grouped force == plain force where the plain algorithm is accurate;
* leapfrog reversibility and O(eps^2) energy violation;
* a planted index shift in the force is caught by the finite-difference pin.
"""
import itertools

import numpy as np
import pytest

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, random_phi, rel_l2

LATTICES = ("tube_3-3_1", "one_site", "two_sites", "triangle")


def _act(kappa, nt, beta, mu, sk, hop, basis, alg, short=False):
    return ref.HubbardFermiAction(kappa * beta / nt, mu * beta / nt, sk, hop, basis, alg, short)


@pytest.mark.parametrize("hop,basis", list(itertools.product((ref.DIA, ref.EXP),
                                                              (ref.PARTICLE_HOLE, ref.SPIN))))
def test_single_vs_square(lattice, hop, basis):
    rng = np.random.RandomState(8613)
    for name in LATTICES:
        kappa = lattice(name)
        nx = kappa.shape[0]
        for nt, beta, sk in itertools.product((8,), (1, 4), (-1, 1)):
            a1 = _act(kappa, nt, beta, 0, sk, hop, basis, ref.DIRECT_SINGLE)
            a2 = _act(kappa, nt, beta, 0, sk, hop, basis, ref.DIRECT_SQUARE)
            for _ in range(2):
                phi = random_phi(rng, nx * nt)
                assert mod2pi_diff(a1.eval(phi), a2.eval(phi)) < 1e-9
                assert np.max(np.abs(a1.force(phi) - a2.force(phi))) < 1e-10


@pytest.mark.parametrize("hop,basis", list(itertools.product((ref.DIA, ref.EXP),
                                                              (ref.PARTICLE_HOLE, ref.SPIN))))
def test_shortcut(lattice, hop, basis):
    rng = np.random.RandomState(8613)
    for name in LATTICES:
        kappa = lattice(name)
        nx = kappa.shape[0]
        for nt, beta, sk in itertools.product((8, 16), (1, 4), (-1, 1)):
            a1 = _act(kappa, nt, beta, 0, sk, hop, basis, ref.DIRECT_SINGLE, False)
            a2 = _act(kappa, nt, beta, 0, sk, hop, basis, ref.DIRECT_SINGLE, True)
            phi = random_phi(rng, nx * nt, imag=False)
            assert mod2pi_diff(a1.eval(phi), a2.eval(phi)) < 1e-10
            assert np.max(np.abs(a1.force(phi) - a2.force(phi))) < 1e-10


@pytest.mark.parametrize("hop,basis,mu", [(ref.EXP, ref.PARTICLE_HOLE, 0.0), (ref.EXP, ref.SPIN, 0.3),
                                          (ref.DIA, ref.PARTICLE_HOLE, 0.2), (ref.DIA, ref.SPIN, 0.5)])
def test_force_is_minus_gradient(hop, basis, mu):
    kappa = ref.honeycomb(3)
    nx, nt, beta = 18, 8, 3.0
    act = _act(kappa, nt, beta, mu, -1, hop, basis, ref.DIRECT_SINGLE)
    rng = np.random.RandomState(3)
    phi = rng.normal(0, 0.4, nx * nt) + 0j
    f = act.force(phi)
    h = 1e-5
    for idx in rng.choice(nx * nt, 12, replace=False):
        e = np.zeros(nx * nt, dtype=complex)
        e[idx] = h
        s_p = act.eval(phi + e)
        s_m = act.eval(phi - e)
        d = complex(s_p) - complex(s_m)
        d = complex(d.real, np.remainder(d.imag + np.pi, 2 * np.pi) - np.pi)
        fd = -d / (2 * h)
        assert abs(fd - f[idx]) < 1e-7 * max(1, abs(f[idx])), (idx, fd, f[idx])


@pytest.mark.parametrize("hop,basis,mu", [(ref.EXP, ref.PARTICLE_HOLE, 0.0), (ref.EXP, ref.SPIN, 0.3),
                                          (ref.DIA, ref.PARTICLE_HOLE, 0.2), (ref.DIA, ref.SPIN, 0.5)])
def test_species_logdet_gradient(hop, basis, mu):
    """d logdetM_s/dphi = c_s (f_s - 1) with c = (-i, +i) in the particle/hole basis and
    (-1, +1) in the spin basis (f_s = forceDirectSinglePart at phi, resp. aux = -i phi; the
    -1 is the derivative of the explicit +-i sum(phi) term): the rule the autograd
    wrapper's backward uses, checked against central differences of logdetM."""
    kappa = ref.honeycomb(3)
    nx, nt, beta = 18, 8, 3.0
    act = _act(kappa, nt, beta, mu, -1, hop, basis, ref.DIRECT_SINGLE)
    rng = np.random.RandomState(5)
    phi = rng.normal(0, 0.4, nx * nt) + 1j * rng.normal(0, 0.1, nx * nt)
    spin = basis == ref.SPIN
    h = 1e-5
    for species, k in ((ref.PARTICLE, act.kp), (ref.HOLE, act.kh)):
        arg = (lambda p: -1j * p) if spin else (lambda p: p)
        coef = (-1.0 if species == ref.PARTICLE else 1.0) if spin else (-1j if species == ref.PARTICLE else 1j)
        grad = coef * (ref.force_direct_single_part(act.hfm, arg(phi), k, species) - 1.0)
        for idx in rng.choice(nx * nt, 6, replace=False):
            e = np.zeros(nx * nt, dtype=complex)
            e[idx] = h
            d = ref.logdetM(act.hfm, arg(phi + e), species) - ref.logdetM(act.hfm, arg(phi - e), species)
            d = complex(d.real, np.remainder(d.imag + np.pi, 2 * np.pi) - np.pi)
            assert abs(d / (2 * h) - grad[idx]) < 1e-7 * max(1, abs(grad[idx])), (species, idx)


def test_sum_action():
    rng = np.random.RandomState(8613)

    class Dummy:
        def eval(self, phi):
            return np.sum(phi)

        def force(self, phi):
            return phi

    for _ in range(5):
        s = ref.SumAction(Dummy(), ref.HubbardGaugeAction(1))
        phi = rng.normal(0, 0.2, 1000) + 1j * rng.normal(0, 0.2, 1000)
        assert abs(s.eval(phi) - (np.sum(phi) + np.dot(phi, phi) / 2)) < 1e-10
        assert np.linalg.norm(s.force(phi) - (phi - phi / 1)) < 1e-10


@pytest.mark.parametrize("hop,species", [(ref.EXP, ref.PARTICLE), (ref.EXP, ref.HOLE), (ref.DIA, ref.HOLE)])
def test_stable_force_matches_plain(hop, species):
    kappa = ref.honeycomb(3)
    nx, nt, beta = 18, 16, 4.0
    cls = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    hfm = cls(kappa * beta / nt, 0.1 * beta / nt, -1)
    phi = np.random.RandomState(9).normal(0, 0.5, nx * nt) + 0j
    k = hfm.K(species)
    plain = ref.force_direct_single_part(hfm, phi, k, species)
    for g in (1, 2, 4, 8):
        st = ref.force_part_stable(hfm, phi, species, g)
        assert rel_l2(st, plain) < 1e-11, g


def test_leapfrog_reversible_and_energy():
    kappa = ref.honeycomb(3)
    nx, nt, beta, u = 18, 8, 2.0, 2.0
    fa = _act(kappa, nt, beta, 0, -1, ref.EXP, ref.PARTICLE_HOLE, ref.DIRECT_SINGLE)
    act = ref.SumAction(ref.HubbardGaugeAction(u * beta / nt), fa)
    rng = np.random.RandomState(11)
    phi0 = rng.normal(0, np.sqrt(u * beta / nt), nx * nt) + 0j
    pi0 = rng.normal(0, 1, nx * nt) + 0j
    s0 = act.eval(phi0)
    phi1, pi1, s1 = ref.leapfrog(phi0, pi0, act, 1.0, 10)
    phi2, pi2, _ = ref.leapfrog(phi1, pi1, act, 1.0, 10, -1.0)
    assert np.linalg.norm(phi2 - phi0) / np.linalg.norm(phi0) < 1e-12
    assert np.linalg.norm(pi2 - pi0) / np.linalg.norm(pi0) < 1e-12
    dh = []
    for nst in (10, 20, 40):
        p1, q1, s1 = ref.leapfrog(phi0, pi0, act, 1.0, nst)
        dh.append(abs(np.real(s1 + np.linalg.norm(q1) ** 2 / 2 - s0 - np.linalg.norm(pi0) ** 2 / 2)))
    assert dh[1] < dh[0] / 3 and dh[2] < dh[1] / 3, dh


def test_planted_force_index_shift_is_caught(monkeypatch):
    kappa = ref.honeycomb(3)
    nx, nt, beta = 18, 8, 3.0
    act = _act(kappa, nt, beta, 0.0, -1, ref.EXP, ref.PARTICLE_HOLE, ref.DIRECT_SINGLE)
    phi = np.random.RandomState(4).normal(0, 0.4, nx * nt) + 0j
    good = act.force(phi)
    orig = ref.force_direct_single_part
    monkeypatch.setattr(ref, "force_direct_single_part",
                        lambda *a: np.roll(orig(*a), nx))
    bad = act.force(phi)
    monkeypatch.undo()
    idx = 5
    h = 1e-5
    e = np.zeros(nx * nt, dtype=complex)
    e[idx] = h
    fd = -(act.eval(phi + e) - act.eval(phi - e)) / (2 * h)
    assert abs(fd - good[idx]) < 1e-7
    assert abs(fd - bad[idx]) > 1e-4
