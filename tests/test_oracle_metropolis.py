"""Pins the oracle's Metropolis step: binary_select and const_step_leapfrog.

binary_select (selector.py:20-30):
* a downhill proposal (dE < 0) is accepted without touching the generator;
* an uphill one (dE >= 0) consumes exactly one uniform, and its acceptance frequency
  is the Metropolis probability e^{-dE} (binomial bounds at e^{-dE} = 1/4: a flipped
  comparison gives 3/4);
* only the real part of the energy difference decides.

const_step_leapfrog (leapfrog.py:38-66) on the gauge action alone, S = sum phi^2/(2U):
the leapfrog map of a harmonic oscillator of frequency w = 1/sqrt(U) is linear, one
step is the 2x2 matrix [[1 - e^2 w^2/2, e], [-e w^2 (1 - e^2 w^2/4), 1 - e^2 w^2/2]]
per site, so the end point, both energies (S + |pi|^2/2) and the accept decision follow
in closed form from the momenta drawn by a replayed generator.  The generator must end
in the state of exactly one normal(0, 1, N) draw (+ one uniform when dE >= 0).

Each check is shown to catch a planted mistake in the oracle (uniform always drawn,
kinetic energy without the 1/2, a flipped comparison).
"""
import numpy as np
import pytest

from oracle import hubbard_ref as ref

U = 0.35


def _same_state(a, b):
    sa, sb = a.get_state(), b.get_state()
    return sa[0] == sb[0] and np.array_equal(sa[1], sb[1]) and sa[2:] == sb[2:]


def _harmonic_end(phi, pi, length, nstep):
    """(phi, pi) after nstep leapfrog steps of S = sum phi^2/(2U), from the step matrix."""
    e = length / nstep
    w2 = 1.0 / U
    step = np.array([[1 - e * e * w2 / 2, e], [-e * w2 * (1 - e * e * w2 / 4), 1 - e * e * w2 / 2]])
    m = np.linalg.matrix_power(step, nstep)
    return m[0, 0] * phi + m[0, 1] * pi, m[1, 0] * phi + m[1, 1] * pi


def _metropolis_checks(select):
    """Run the binary_select pins against ``select``; returns the list of failures."""
    bad = []
    # downhill: accepted, generator untouched
    for de in (-1e-12, -0.3, -40.0):
        rng, twin = np.random.RandomState(5), np.random.RandomState(5)
        if select(rng, 1.0, 1.0 + de) != 1 or not _same_state(rng, twin):
            bad.append(f"downhill {de}")
    # uphill: exactly one uniform, accepted iff u < e^{-dE}
    for de in (0.0, 1e-9, 0.7, 3.0):
        rng, twin = np.random.RandomState(11), np.random.RandomState(11)
        got = select(rng, 2.0, 2.0 + de)
        u = twin.uniform(0, 1)
        if not _same_state(rng, twin) or got != int(u < np.exp(-de)):
            bad.append(f"uphill {de}")
    # acceptance frequency e^{-dE} = 1/4 over 20000 proposals (sigma = 0.003)
    rng = np.random.RandomState(99)
    rate = np.mean([select(rng, 0.0, np.log(4.0)) for _ in range(20000)])
    if abs(rate - 0.25) > 0.015:
        bad.append(f"rate {rate}")
    # complex energies: only Re(E1 - E0) decides, not |E1 - E0| or the imaginary parts
    for e0, e1 in ((1.0 + 100j, 0.5 - 100j), (1.0 + 100j, 1.3 - 50j)):
        rng, twin = np.random.RandomState(3), np.random.RandomState(3)
        de = np.real(e1 - e0)
        want = 1 if de < 0 else int(twin.uniform(0, 1) < np.exp(-de))
        if select(rng, e0, e1) != want or not _same_state(rng, twin):
            bad.append("complex")
    return bad


def test_binary_select():
    assert _metropolis_checks(ref.binary_select) == []


def test_binary_select_planted_mistakes_are_caught():
    def always_draw(rng, e0, e1):
        u = rng.uniform(0, 1)
        de = np.real(e1 - e0)
        return 1 if de < 0 or np.exp(-de) > u else 0

    def flipped(rng, e0, e1):
        de = np.real(e1 - e0)
        return 1 if de < 0 or np.exp(-de) < rng.uniform(0, 1) else 0

    def modulus(rng, e0, e1):
        de = abs(e1 - e0)
        return 1 if np.real(e1 - e0) < 0 or np.exp(-de) > rng.uniform(0, 1) else 0

    assert any(b.startswith("downhill") for b in _metropolis_checks(always_draw))
    assert any(b.startswith("rate") for b in _metropolis_checks(flipped))
    assert "complex" in _metropolis_checks(modulus)


def _trajectory_checks(step_fn, seed, nx, length, nstep):
    """const_step_leapfrog pins on the gauge action; returns the list of failures."""
    bad = []
    g = ref.HubbardGaugeAction(U)
    phi0 = np.random.RandomState(seed + 1).normal(0, np.sqrt(U), nx) + 0j
    s0 = g.eval(phi0)
    rng, twin = np.random.RandomState(seed), np.random.RandomState(seed)
    phi_new, act_new, point, e0, e1 = step_fn(rng, g, phi0, s0, length, nstep)[:5]
    pi = twin.normal(0, 1, nx)
    q1, p1 = _harmonic_end(phi0.real, pi, length, nstep)
    h0 = np.sum(phi0.real ** 2) / (2 * U) + np.sum(pi ** 2) / 2
    h1 = np.sum(q1 ** 2) / (2 * U) + np.sum(p1 ** 2) / 2
    if abs(e0 - h0) > 1e-11 * h0 or abs(e1 - h1) > 1e-11 * h1:
        bad.append("energies")
    de = h1 - h0
    want = 1 if de < 0 else int(twin.uniform(0, 1) < np.exp(-de))
    if not _same_state(rng, twin):
        bad.append("draws")
    if point != want:
        bad.append("decision")
    end = q1 if want == 1 else phi0.real
    if np.abs(np.asarray(phi_new) - end).max() > 1e-11 or abs(act_new - (h1 - np.sum(p1 ** 2) / 2 if want else s0)) \
            > 1e-10 * abs(s0):
        bad.append("state")
    return bad, point


def test_const_step_leapfrog_harmonic():
    points = set()
    for nstep in (1, 3, 12):
        for seed in range(40):
            bad, point = _trajectory_checks(ref.const_step_leapfrog, 100 * seed + nstep, 24, 1.0, nstep)
            assert bad == [], (nstep, seed, bad)
            points.add(point)
    assert points == {0, 1}  # both branches of the decision were exercised


def test_const_step_leapfrog_planted_mistakes_are_caught(monkeypatch):
    def no_half(rng, action, phi, act_val, length, nstep):
        pi = rng.normal(0, 1, len(phi)) + 0j
        phi1, pi1, act1 = ref.leapfrog(phi, pi, action, length, nstep)
        e0 = act_val + np.linalg.norm(pi) ** 2
        e1 = act1 + np.linalg.norm(pi1) ** 2
        tp = ref.binary_select(rng, e0, e1)
        return (phi1, act1, 1, e0, e1) if tp else (np.asarray(phi), act_val, 0, e0, e1)

    found = set()
    for seed in range(10):
        found.update(_trajectory_checks(no_half, seed, 24, 1.0, 3)[0])
    assert "energies" in found
    # the uniform drawn on every trajectory (inside binary_select) breaks the draw count
    orig = ref.binary_select

    def eager(rng, e0, e1):
        u = rng.uniform(0, 1)
        de = np.real(e1 - e0)
        return 1 if de < 0 or np.exp(-de) > u else 0
    monkeypatch.setattr(ref, "binary_select", eager)
    found = set()
    for seed in range(40):
        found.update(_trajectory_checks(ref.const_step_leapfrog, seed, 24, 1.0, 3)[0])
    monkeypatch.setattr(ref, "binary_select", orig)
    assert "draws" in found
