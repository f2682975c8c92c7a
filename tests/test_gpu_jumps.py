"""GPU: the fermion-action invariance that TwoPiJumps / UniformJump rely on.

Honeycomb 3x3 torus (Nx=18), Nt=16, beta=4, U=2, mu=0.1 (DIA) / 0 (EXP), gauge + fermion
action.  In the particle-hole basis the action of every proposal (2 pi at random sites,
2 pi / Nt everywhere) computed by the shortcut (old value + change of the gauge action)
equals the full action evaluated by the GPU kernels at the new field to 1e-10.  In the
spin basis the field enters the fermion matrices as e^{+-phi}: the full action differs
from the shortcut formula there, and the evolvers evaluate it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd.action import HFABasis, HFAHopping, HubbardGaugeAction, makeHubbardFermiAction  # noqa: E402
from isle_amd.evolver import EvolutionStage, NumpyRNG, TwoPiJumps, UniformJump  # noqa: E402
from isle_amd.lattice import honeycomb  # noqa: E402


def _action(hop, basis):
    lat = honeycomb(3, nt=16)
    beta, U = 4.0, 2.0
    mu = 0.1 * beta / 16 if hop == HFAHopping.DIA else 0.0
    gauge = HubbardGaugeAction(U * beta / 16)
    return lat, gauge + makeHubbardFermiAction(lat, beta, mu, -1, hop, basis)


def _rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("hop", [HFAHopping.EXP, HFAHopping.DIA])
def test_shortcut_equals_full_action(hop):
    print(f"jump shortcut, {hop.name}, Nx=18 Nt=16", flush=True)
    lat, action = _action(hop, HFABasis.PARTICLE_HOLE)
    rng = np.random.RandomState(4)
    phi = rng.normal(0, 0.5, lat.lattSize()) + 0j
    act = complex(action.eval(phi))
    two_pi = TwoPiJumps(2, 3, action, lat, NumpyRNG(9))
    uni = UniformJump(action, lat, NumpyRNG(9))
    assert two_pi._action.shortcut and uni._action.shortcut
    for _ in range(4):
        sites = rng.choice(phi.size, 3, replace=False)
        shifts = rng.choice((-2 * np.pi, 2 * np.pi), 3)
        new = phi.copy()
        new[sites] += shifts
        assert _rel(two_pi._action.evalLocal(new, sites, shifts, act), complex(action.eval(new))) < 1e-10
    for shift in (2 * np.pi / 16, -2 * np.pi / 16):
        assert _rel(uni._action.evalGlobal(phi + shift, shift, act), complex(action.eval(phi + shift))) < 1e-10
    stage = EvolutionStage(phi, act, 1)
    for _ in range(3):
        stage = uni.evolve(two_pi.evolve(stage))
        assert _rel(stage.actVal, complex(action.eval(stage.phi))) < 1e-10


def test_spin_basis_is_evaluated():
    lat, action = _action(HFAHopping.EXP, HFABasis.SPIN)
    rng = np.random.RandomState(6)
    phi = rng.normal(0, 0.5, lat.lattSize()) + 0j
    act = complex(action.eval(phi))
    ev = TwoPiJumps(1, 2, action, lat, NumpyRNG(2))
    assert not ev._action.shortcut
    sites = np.array([3, 40])
    shifts = np.array([2 * np.pi, -2 * np.pi])
    new = phi.copy()
    new[sites] += shifts
    full = complex(action.eval(new))
    gauge_only = act + np.dot(2 * new[sites] - shifts, shifts) / (2 * (2.0 * 4.0 / 16))
    assert _rel(gauge_only, full) > 1e-3  # the fermion action is not invariant in the spin basis
    assert _rel(ev._action.evalLocal(new, sites, shifts, act), full) < 1e-12
