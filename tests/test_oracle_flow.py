"""The oracle's adaptive RK4 holomorphic flow (oracle.rk4_flow), pinned by closed forms.

With only the gauge action S = sum phi^2 / (2U) the flow d phi / dt = conj(phi) / U is
linear: for phi = a + i b it is a' = a / U, b' = -b / U, and one step of any four-stage
fourth-order Runge-Kutta method multiplies a by T4(h) and b by T4(-h), h = +-eps / U,
T4(z) = 1 + z + z^2/2 + z^3/6 + z^4/24.  Both variants (classic, 3/8 rule) and both
directions must reproduce that: a wrong stage coefficient (such as 1/30 for 1/3) or a
stage without the direction fails it.  The step-size control is checked on its stated
rules: a last step shortened to end at the flow time, growth by 2 after error-free steps,
and a stop at time 0 when no step meets the tolerance.
"""
import numpy as np
import pytest

from oracle import hubbard_ref as ref

U = 0.7


def _t4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


@pytest.mark.parametrize("n", [0, 1])
@pytest.mark.parametrize("direction", [1.0, -1.0])
def test_single_step_closed_form(n, direction):
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(1).normal(size=12) + 1j * np.random.RandomState(2).normal(size=12)
    eps = 0.05
    out, act, t = ref.rk4_flow(phi, g, eps, eps, n=n, direction=direction, tolerance=10.0)
    h = direction * eps / U
    want = phi.real * _t4(h) + 1j * phi.imag * _t4(-h)
    assert t == eps
    assert np.abs(out - want).max() < 1e-14
    assert abs(act - g.eval(out)) < 1e-14


def test_step_control():
    g = ref.HubbardGaugeAction(U)
    phi = np.random.RandomState(3).normal(size=8) + 0j  # real: Im S stays 0, every step error-free
    steps = []
    orig = ref._rk4_step

    def spy(p, a, eps, d, n):
        steps.append(eps)
        return orig(p, a, eps, d, n)
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
