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


@pytest.mark.parametrize("basis", [HFABasis.PARTICLE_HOLE, HFABasis.SPIN])
def test_jump_chain_matches_oracle(basis):
    """TwoPiJumps then UniformJump, 6 sweeps, against oracle.two_pi_jumps / uniform_jump
    from the same seed with the oracle's gauge + fermion action evaluated in full (the
    product uses the shortcut in the particle-hole basis and the GPU action in the spin
    basis): same decisions, fields and action values."""
    from oracle import hubbard_ref as ref
    print(f"jump chain vs oracle, {basis.name}, Nx=18 Nt=16", flush=True)
    lat, _ = _action(HFAHopping.EXP, basis)
    ut = 20.0  # a wide gauge action: jumps are accepted and rejected
    action = HubbardGaugeAction(ut) + makeHubbardFermiAction(lat, 4.0, 0.0, -1, HFAHopping.EXP, basis)
    kt = lat.hopping() * 4.0 / 16
    oracle = ref.SumAction(ref.HubbardGaugeAction(ut),
                           ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP,
                                                  ref.PARTICLE_HOLE if basis == HFABasis.PARTICLE_HOLE else ref.SPIN))
    phi = np.random.RandomState(21).normal(0, 0.5, lat.lattSize()) + 0j
    rng, rs = NumpyRNG(31), np.random.RandomState(31)
    two_pi, uni = TwoPiJumps(3, 2, action, lat, rng), UniformJump(action, lat, rng)
    stage = EvolutionStage(phi, complex(action.eval(phi)), 1)
    want_phi, want_act = phi, oracle.eval(phi)
    decisions = []
    for sweep in range(6):
        stage = two_pi.evolve(stage)
        want_phi, want_act, batch = ref.two_pi_jumps(rs, oracle, want_phi, want_act, 3, 2)
        assert two_pi.acceptedPerBatch[-1] == batch, sweep
        stage = uni.evolve(stage)
        want_phi, want_act, took = ref.uniform_jump(rs, oracle, want_phi, want_act, 16)
        assert uni.accepted[-1] == took, sweep
        decisions += batch + [took]
        assert np.abs(np.asarray(stage.phi) - want_phi).max() < 1e-12
        assert _rel(complex(stage.actVal), want_act) < 1e-10
    assert set(decisions) == {0, 1}
