"""TwoPiJumps / UniformJump on the CPU (gauge actions only, so no GPU is touched).

With only gauge actions the new action value must be the gauge action of the new field,
sum phi^2 / (2 Utilde) summed over the gauge terms, computed here directly; the chain must
be the oracle's (oracle.two_pi_jumps / uniform_jump from a RandomState of the same seed,
evaluating the action in full): same draws, decisions, fields and action values,
including a sweep whose last batch is rejected after an earlier one was accepted (the
field of the last accepted batch is kept).  The fermion action's invariance, which
the shortcut relies on, is checked on the GPU in tests/test_gpu_jumps.py.
"""
import numpy as np
import pytest

from isle_amd import runfile
from isle_amd.action import HubbardGaugeAction, SumAction
from isle_amd.evolver import EvolutionStage, NumpyRNG, TwoPiJumps, UniformJump
from isle_amd.lattice import honeycomb
from oracle import hubbard_ref as ref

U1, U2 = 30.0, 120.0  # a 2 pi jump then costs O(1) in action: both outcomes occur


def _gauge(phi, us=(U1, U2)):
    phi = np.asarray(phi)
    return sum(np.dot(phi, phi) / (2 * u) for u in us)


def _start(seed=3, nt=4):
    lat = honeycomb(2, nt=nt)
    phi = np.random.RandomState(seed).normal(0, 1.0, lat.lattSize()) + 0.1j
    action = SumAction(HubbardGaugeAction(U1), HubbardGaugeAction(U2))
    return lat, phi, action


def _oracle_action(us=(U1, U2)):
    return ref.SumAction(*[ref.HubbardGaugeAction(u) for u in us])


def test_two_pi_jumps_values_and_draws():
    lat, phi, action = _start()
    ev = TwoPiJumps(4, 2, action, lat, NumpyRNG(11))
    assert ev._action.shortcut
    stage = EvolutionStage(phi, _gauge(phi), 1)
    rs, oracle = np.random.RandomState(11), _oracle_action()
    wphi, wact = phi, _gauge(phi)
    saw_rejected_last = False
    for k in range(25):
        stage = ev.evolve(stage)
        wphi, wact, wacc = ref.two_pi_jumps(rs, oracle, wphi, wact, 4, 2)
        assert ev.acceptedPerBatch[k] == wacc
        assert np.allclose(stage.phi, wphi, rtol=0, atol=1e-12)
        assert abs(stage.actVal - wact) < 1e-10 * abs(wact)
        assert abs(stage.actVal - _gauge(stage.phi)) < 1e-10 * abs(wact)  # the gauge action of the new field
        assert stage.trajPoint == (1 if any(wacc) else 0)
        saw_rejected_last |= any(wacc) and wacc[-1] == 0
    assert saw_rejected_last  # the case where the reference returns a rejected proposal


def test_uniform_jump_values_and_draws():
    lat, phi, action = _start(nt=8)
    ev = UniformJump(action, lat, NumpyRNG(5))
    stage = EvolutionStage(phi, _gauge(phi), 1)
    rs, oracle = np.random.RandomState(5), _oracle_action()
    cur, act = phi.copy(), _gauge(phi)
    for _ in range(40):
        stage = ev.evolve(stage)
        cur, act, took = ref.uniform_jump(rs, oracle, cur, act, 8)
        assert stage.trajPoint == took
        assert np.allclose(stage.phi, cur, rtol=0, atol=1e-12)
        assert abs(stage.actVal - act) < 1e-10 * abs(act)
    assert 0 < np.mean(ev.accepted) < 1


def test_jump_evolvers_checkpoint():
    lat, phi, action = _start()
    mgr = runfile.EvolverManager()
    for ev in (TwoPiJumps(2, 5, action, lat, NumpyRNG(1)), UniformJump(action, lat, NumpyRNG(1))):
        g = runfile.Group()
        mgr.save(ev, g)
        back = mgr.load(g, action, lat, NumpyRNG(1))
        assert type(back) is type(ev)
        if isinstance(ev, TwoPiJumps):
            assert (back.nBatches, back.batchSize) == (2, 5)
        else:
            assert back._absShift == pytest.approx(2 * np.pi / lat.nt())
