"""The oracle's force against central finite differences of its own action formula taken in
extended precision (tests/golden/fd_config2_h1e-6.npz, made by make_fd_config2.py beside
it): config 2 (honeycomb 6x6, Nx=72, Nt=64, beta=6, U=3, EXP, particle-hole, sigmaKappa=-1,
phi from RandomState(2000)), 64 entries, h = 1e-6.  In extended precision the difference's
rounding is ~1e-11, its truncation O(h^2) (~1e-8 here; the imaginary parts of the
differences, zero for the exact derivative on a real field, show its size): the oracle's force
agrees within 3e-8 absolute on forces up to ~13.  A planted mistake (the force of the
particles alone) fails."""
import os

import numpy as np

from oracle import hubbard_ref as ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fd_config2_h1e-6.npz")


def _config2():
    nx, nt, beta, u = 72, 64, 6.0, 3.0
    kt = ref.honeycomb(6) * beta / nt
    rng = np.random.RandomState(2000)
    phi = rng.normal(0, np.sqrt(u * beta / nt), nx * nt) + 0j
    idx = rng.choice(nx * nt, 64, replace=False)
    return kt, phi, idx


def test_oracle_force_matches_extended_precision_fd():
    kt, phi, idx = _config2()
    d = np.load(GOLDEN)
    assert np.array_equal(d["idx"], idx)
    f = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(phi)
    assert np.abs(f[idx] - d["fd"]).max() < 3e-8
    assert np.abs(d["fd"].imag).max() < 3e-8  # real field: real derivative, to truncation
    # planted mistake: the particle part alone is far off
    hfm = ref.HFMExp(kt, 0.0, -1)
    fp = ref.force_direct_single_part(hfm, phi, hfm.K(ref.PARTICLE), ref.PARTICLE)
    assert np.abs(-1j * fp[idx] - d["fd"]).max() > 1e-2
