"""Pins the oracle's block LU of Q (qlu, qlu_reconstruct, solve_q, logdet_q) and Q itself.

Independent checks (none retypes the oracle's formula):
* Q assembled from P, T+ and T- equals M_p M_h^T, the product of the two fermion
  matrices that the dense-logdet tests pin (EXP and DIA, nt = 2, 3, 5, sigmaKappa = +-1,
  mu = 0 / 0.1);
* logdet_q of the block LU equals log det of the dense Q by LAPACK (the reference's own
  test, tests/test_hubbardFermiMatrix.py:148-166: nt 4, 8, 32, mu = 0, sigmaKappa = +-1),
  here also at nt = 1, 2, 3 (the special cases) and mu = 0.1;
* solve_q equals numpy.linalg.solve on the dense Q;
* the blocks multiplied back together (qlu_reconstruct) give Q;
* a sign flipped in any one of dinv, u, l, v, h breaks the reconstruction, so each block
  is pinned on its own.
"""
import itertools

import numpy as np
import pytest

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, random_phi

LATTICES = ("tube_3-3_1", "triangle", "two_sites", "one_site")


def _hfm(kappa, nt, hop, mu, sk):
    cls = ref.HFMExp if hop == ref.EXP else ref.HFMDia
    return cls(kappa * 3.0 / nt, mu * 3.0 / nt, sk)


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_q_is_mp_mh_transposed(lattice, hop):
    rng = np.random.RandomState(31)
    for name, nt, mu, sk in itertools.product(LATTICES, (2, 3, 5), (0.0, 0.1), (-1, 1)):
        kappa = lattice(name)
        hfm = _hfm(kappa, nt, hop, mu, sk)
        phi = random_phi(rng, kappa.shape[0] * nt)
        q = hfm.Q(phi)
        mm = hfm.M(phi, ref.PARTICLE) @ hfm.M(phi, ref.HOLE).T
        assert np.abs(q - mm).max() < 1e-12 * np.abs(mm).max(), (name, nt, mu, sk)


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_logdet_and_solve_vs_dense(lattice, hop):
    rng = np.random.RandomState(77)
    for name, nt, mu, sk in itertools.product(LATTICES, (1, 2, 3, 4, 8, 32), (0.0, 0.1), (-1, 1)):
        kappa = lattice(name)
        nx = kappa.shape[0]
        hfm = _hfm(kappa, nt, hop, mu, sk)
        phi = random_phi(rng, nx * nt)
        q = hfm.Q(phi)
        lu = ref.qlu(hfm, phi)
        assert len(lu.dinv) == nt and len(lu.u) == max(nt - 1, 0)
        dense = ref.ilogdet(q)
        assert mod2pi_diff(ref.logdet_q(lu), dense) < 1e-10 * max(1.0, abs(dense)), (name, nt, mu, sk)
        b = random_phi(rng, nx * nt)
        want = np.linalg.solve(q, b)
        assert np.abs(ref.solve_q(lu, b) - want).max() < 1e-10 * np.abs(want).max(), (name, nt, mu, sk)


@pytest.mark.parametrize("hop", [ref.EXP, ref.DIA])
def test_reconstruct_is_q(lattice, hop):
    rng = np.random.RandomState(5)
    for name, nt, sk in itertools.product(("tube_3-3_1", "triangle"), (3, 4, 8), (-1, 1)):
        kappa = lattice(name)
        hfm = _hfm(kappa, nt, hop, 0.1, sk)
        phi = random_phi(rng, kappa.shape[0] * nt)
        q = hfm.Q(phi)
        assert np.abs(ref.qlu_reconstruct(ref.qlu(hfm, phi)) - q).max() < 1e-11 * np.abs(q).max()


@pytest.mark.parametrize("part", ["dinv", "u", "l", "v", "h"])
def test_planted_block_sign_is_caught(part):
    kappa = ref.honeycomb(2)
    nt = 6
    hfm = ref.HFMExp(kappa * 3.0 / nt, 0.0, 1)
    phi = random_phi(np.random.RandomState(2), 8 * nt)
    lu = ref.qlu(hfm, phi)
    q = hfm.Q(phi)
    blocks = getattr(lu, part)
    blocks[1] = -blocks[1]
    assert np.abs(ref.qlu_reconstruct(lu) - q).max() > 1e-3 * np.abs(q).max()
