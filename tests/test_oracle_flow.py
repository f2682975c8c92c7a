"""The oracle's adaptive RK4 holomorphic flow (oracle.rk4_flow), pinned by closed forms.

With only the gauge action S = sum phi^2 / (2U) the flow d phi / dt = conj(phi) / U is
linear: for phi = a + i b it is a' = a / U, b' = -b / U, and one step of any four-stage
fourth-order Runge-Kutta method multiplies a by T4(h) and b by T4(-h), h = +-eps / U,
T4(z) = 1 + z + z^2/2 + z^3/6 + z^4/24.  The textbook variant (classic, 3/8 rule) in
both directions must reproduce that: a wrong stage coefficient (such as 1/30 for 1/3) or
a stage without the direction fails it.

The reference variant (integrator.cpp:39-98: beta21 = 1/30 in the 3/8 rule, stage 4 with
-eps instead of -eps direction) is pinned by the stability function of its stage table
on the same linear flow: a step multiplies a by R = 1 + w^T Z (I - B Z)^{-1} 1 with B the
strictly lower stage matrix, w the weights and Z = diag(z, z, z, z4) the stages' own
factors (z4 = eps / U whatever the direction); b gets the same with -z, -z4.  It equals
T4 only for the classic rule going forward.  The step-size control is checked on its stated
rules: a last step shortened to end at the flow time, growth by 2 after error-free steps,
and a stop at time 0 when no step meets the tolerance.
"""
import numpy as np
import pytest

from oracle import hubbard_ref as ref

U = 0.7


def _t4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


# integrator.cpp:39-70, written as (weights, strictly lower stage matrix)
_TABLES = {0: ((1 / 6, 1 / 3, 1 / 3, 1 / 6), ((0, 0, 0, 0), (0.5, 0, 0, 0), (0, 0.5, 0, 0), (0, 0, 1, 0))),
           1: ((1 / 8, 3 / 8, 3 / 8, 1 / 8), ((0, 0, 0, 0), (1 / 30, 0, 0, 0), (-1 / 3, 1, 0, 0), (1, -1, 1, 0)))}


def _stability(n, zs):
    w, b = (np.array(x, dtype=float) for x in _TABLES[n])
    z = np.diag(zs)
    return 1 + w @ z @ np.linalg.solve(np.eye(4) - b @ z, np.ones(4))


@pytest.mark.parametrize("n", [0, 1])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_single_step_closed_form(n, direction):
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(1).normal(size=12) + 1j * np.random.RandomState(2).normal(size=12)
    eps = 0.05
    out, act, t = ref.rk4_flow(phi, g, eps, eps, n=n, direction=direction, tolerance=10.0, variant="textbook")
    h = direction * eps / U
    want = phi.real * _t4(h) + 1j * phi.imag * _t4(-h)
    assert t == eps
    assert np.abs(out - want).max() < 1e-14
    assert abs(act - g.eval(out)) < 1e-14


@pytest.mark.parametrize("n", [0, 1])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_reference_variant_stage_table(n, direction):
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(4).normal(size=10) + 1j * np.random.RandomState(5).normal(size=10)
    for eps in (0.02, 0.05, 0.2):
        out, act, t = ref.rk4_flow(phi, g, eps, eps, n=n, direction=direction, tolerance=10.0)
        z, z4 = direction * eps / U, eps / U
        want = phi.real * _stability(n, [z, z, z, z4]) + 1j * phi.imag * _stability(n, [-z, -z, -z, -z4])
        assert t == eps
        assert np.abs(out - want).max() < 1e-14
        # the classic rule going forward is the textbook method; everything else differs
        textbook = phi.real * _t4(z) + 1j * phi.imag * _t4(-z)
        assert (np.abs(out - textbook).max() < 1e-14) == (n == 0 and direction == 1.0)


def test_reference_variant_planted_mistake_is_caught(monkeypatch):
    """A stage table with beta31 and beta32 exchanged fails the stage-table pin."""
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(4).normal(size=10) + 0j
    w, b = ref._RK4["reference"][1]
    monkeypatch.setitem(ref._RK4["reference"], 1, (w, (b[0], b[2], b[1], b[3], b[4], b[5])))
    out, _, _ = ref.rk4_flow(phi, g, 0.2, 0.2, n=1, tolerance=10.0)
    want = phi.real * _stability(1, [0.2 / U] * 4)
    assert np.abs(out.real - want).max() > 1e-6


def test_step_control():
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(3).normal(size=8) + 0j  # real: Im S stays 0, every step error-free
    steps = []
    orig = ref._rk4_step

    def spy(p, a, eps, d, n, variant):
        steps.append(eps)
        return orig(p, a, eps, d, n, variant)
    ref._rk4_step = spy
    try:
        out, act, t = ref.rk4_flow(phi, g, 1.0, 0.1)
    finally:
        ref._rk4_step = orig
    assert steps == pytest.approx([0.1, 0.2, 0.4, 0.3])  # doubling, last step shortened
    assert t == pytest.approx(1.0)
    # no step can meet a tolerance of 1e-30 on a complex field: shrink to the minimum, stop
    cphi = phi + 0.5j
    out, act, t = ref.rk4_flow(cphi, g, 1.0, 0.1, tolerance=1e-30)
    assert t == 0.0 and np.array_equal(out, cphi)
