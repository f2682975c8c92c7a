"""Pins the oracle's fermion matrices, logdetM and solveM.

Independent checks (none retypes the oracle's formula):
* logdetM == log det of the densely assembled (NxNt)^2 M (the reference's own test,
  tests/test_hubbardFermiMatrix.py:125-146), for EXP and DIA, both species, real /
  imaginary / complex phi, several mu and sigmaKappa;
* closed forms at phi = 0 and at uniform phi (products of commuting factors);
* solveM residual M x^T = b^T (tests/test_hubbardFermiMatrix.py:178-198) and agreement
  with numpy.linalg.solve on the dense M;
* the stabilised grouped restatement agrees with the plain one, independently of the group size;
* a planted mistake in the oracle (wrong phase sign) is caught by the dense-M check.
"""
import itertools

import numpy as np
import pytest

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, random_phi, rel_mod2pi

LATTICES = ("c20", "tube_3-3_1", "one_site", "two_sites", "triangle")


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_logdetM_vs_dense(lattice, hop):
    rng = np.random.RandomState(8613)
    cls = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    for name in LATTICES:
        kappa = lattice(name)
        nx = kappa.shape[0]
        for nt, beta, mu, sk in itertools.product((2, 4, 8), (3, 6), (0, 0.1, 1.0), (-1, 1)):
            hfm = cls(kappa * beta / nt, mu * beta / nt, sk)
            for species, (re_, im_) in itertools.product((ref.PARTICLE, ref.HOLE),
                                                         ((True, False), (False, True), (True, True))):
                phi = random_phi(rng, nx * nt, re_, im_)
                plain = ref.ilogdet(hfm.M(phi, species))
                via = ref.logdetM(hfm, phi, species)
                assert mod2pi_diff(plain, via) < 1e-9 * max(1, abs(plain)), (name, nt, beta, mu, sk, species)


def test_closed_forms_exp():
    kappa = ref.honeycomb(3)
    eps = np.linalg.eigvalsh(kappa)
    nx = kappa.shape[0]
    for nt, beta, mu, sk in itertools.product((4, 16), (2.0, 5.0), (0.0, 0.3), (-1, 1)):
        kt, mt = kappa * beta / nt, mu * beta / nt
        et = eps * beta / nt
        hfm = ref.HFMExp(kt, mt, sk)
        for c in (0.0, 0.37, 0.2 - 0.1j):
            phi = np.full(nx * nt, c, dtype=complex)
            # particles: B = e^{i nt c} E^nt
            exp_p = np.sum(np.log(1 + np.exp(1j * nt * c + nt * (et - mt))))
            assert mod2pi_diff(ref.logdetM(hfm, phi, ref.PARTICLE), exp_p) < 1e-10 * abs(exp_p) + 1e-12
            # holes: det M_h = prod_t det(F_h(t)) det(1 + A^{-1}), F_h(t) = E_h e^{-i c}
            lam = sk * et + mt
            exp_h = np.sum(np.log(1 + np.exp(-1j * nt * c + nt * lam)))
            assert mod2pi_diff(ref.logdetM(hfm, phi, ref.HOLE), exp_h) < 1e-10 * abs(exp_h) + 1e-12


def test_closed_forms_dia():
    kappa = ref.honeycomb(3)
    eps = np.linalg.eigvalsh(kappa)
    nx = kappa.shape[0]
    for nt, beta, mu, sk in itertools.product((4, 16), (1.0, 3.0), (0.0, 0.5), (-1, 1)):
        kt, mt = kappa * beta / nt, mu * beta / nt
        et = eps * beta / nt
        hfm = ref.HFMDia(kt, mt, sk)
        for c in (0.0, 0.41):
            phi = np.full(nx * nt, c, dtype=complex)
            # particle: A^{-1} = e^{-i nt c} K_p^nt, logdetM = i nt nx c + log det(1 + A^{-1})
            exp_p = 1j * nt * nx * c + np.sum(np.log(1 + np.exp(-1j * nt * c) * (1 + mt - et) ** nt))
            assert mod2pi_diff(ref.logdetM(hfm, phi, ref.PARTICLE), exp_p) < 1e-10 * abs(exp_p) + 1e-12
            exp_h = -1j * nt * nx * c + np.sum(np.log(1 + np.exp(1j * nt * c) * (1 - mt - sk * et) ** nt + 0j))
            assert mod2pi_diff(ref.logdetM(hfm, phi, ref.HOLE), exp_h) < 1e-10 * abs(exp_h) + 1e-12


def test_one_site_scalar():
    # Nx = 1, kappa = 0: M_p is an nt x nt cyclic matrix, det = 1 + e^{-nt mu + i sum phi}
    for nt, mu in ((3, 0.1), (8, 0.0), (16, 0.25)):
        hfm = ref.HFMExp(np.zeros((1, 1)), mu, -1)
        phi = np.random.RandomState(nt).normal(0, 0.5, nt) + 0j
        exp_p = np.log(1 + np.exp(-nt * mu + 1j * np.sum(phi)))
        assert mod2pi_diff(ref.logdetM(hfm, phi, ref.PARTICLE), exp_p) < 1e-13


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_solveM_residual(lattice, hop):
    rng = np.random.RandomState(8613)
    cls = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    for name in ("tube_3-3_1", "two_sites", "triangle"):
        kappa = lattice(name)
        nx = kappa.shape[0]
        for nt, sk, species in itertools.product((4, 8), (-1, 1), (ref.PARTICLE, ref.HOLE)):
            hfm = cls(kappa / nt, 0.0, sk)
            phi = random_phi(rng, nx * nt)
            rhss = np.array([random_phi(rng, nx * nt) for _ in range(10)])
            res = ref.solveM(hfm, phi, species, rhss)
            m = hfm.M(phi, species)
            np.testing.assert_allclose(m @ res.T, rhss.T, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(res.T, np.linalg.solve(m, rhss.T), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_stable_logdet_matches_plain(hop):
    kappa = ref.honeycomb(3)
    nx = kappa.shape[0]
    cls = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    rng = np.random.RandomState(5)
    nt, beta, mu = 16, 4.0, 0.2
    hfm = cls(kappa * beta / nt, mu * beta / nt, -1)
    phi = rng.normal(0, 0.5, nx * nt) + 1j * rng.normal(0, 0.1, nx * nt)
    for species in (ref.PARTICLE, ref.HOLE):
        plain = ref.logdetM(hfm, phi, species)
        for g in (1, 2, 4, 8, 16):
            st = ref.logdetM_stable(hfm, phi, species, g)
            assert rel_mod2pi(st, plain) < 1e-12, (g, st, plain)


def test_planted_mistake_is_caught(monkeypatch):
    """Flip the phase sign of F in the oracle: the dense-M pin must fail."""
    kappa = ref.honeycomb(3) * 0.25
    hfm = ref.HFMExp(kappa, 0.0, -1)
    phi = np.random.RandomState(1).normal(0, 0.3, 18 * 4) + 0j
    dense = ref.ilogdet(hfm.M(phi, ref.PARTICLE))
    good = ref.logdetM(hfm, phi, ref.PARTICLE)
    assert mod2pi_diff(dense, good) < 1e-10
    orig = ref.HFMExp.F

    def bad_F(self, tp, phi, species, inv=False):
        return orig(self, tp, np.conj(phi) * -1, species, inv)

    monkeypatch.setattr(ref.HFMExp, "F", bad_F)
    bad = ref.logdetM(hfm, phi, ref.PARTICLE)
    monkeypatch.undo()
    assert mod2pi_diff(dense, bad) > 1e-3
