"""Pins the oracle's evolver restatements that go beyond one leapfrog trajectory.

Transforms in const_step_leapfrog (leapfrog.py:38-66, transform.py:73-101):
* a constant shift s moves the whole chain: with shift, from phi, the run equals the run
  without shift from phi - s, translated by s (same draws, decisions and action values),
  which fixes where backward / forward act and which action value is carried;
* the log det J weight of a transform that has one enters the energy and is carried by
  accepted stages (a transform with a constant log det J leaves every decision alone).

Jump evolvers (hubbard.py:15-184), both evaluating the action in full:
* with the particle-hole fermion action alone on a bipartite lattice, every 2 pi batch and
  every 2 pi / Nt uniform jump leaves the action unchanged, so every proposal is taken
  and the field moves by exact multiples of the shift;
* whatever the decisions, the returned action value is the action of the returned field
  (a restatement returning a rejected last proposal, as the reference does, fails this);
* the sites of a batch are distinct and the shifts are +-2 pi.
"""
import numpy as np

from oracle import hubbard_ref as ref

U = 0.4


def _gauge_plus_skew():
    """Gauge action plus a non-symmetric term, so that a translation is visible."""
    class Skew:
        def eval(self, phi):
            return 0.3 * np.sum(np.sin(phi + 0.2))

        def force(self, phi):
            return -0.3 * np.cos(phi + 0.2)
    return ref.SumAction(ref.HubbardGaugeAction(U), Skew())


def test_constant_shift_translates_the_chain():
    action = _gauge_plus_skew()
    n = 30
    shift = np.random.RandomState(7).normal(0, 0.5, n) + 0j
    phi = np.random.RandomState(8).normal(0, 1.0, n) + 0j
    tr = ref.ConstantShift(shift, action)
    rs_a, rs_b = np.random.RandomState(1), np.random.RandomState(1)
    a_phi, a_act, a_w = phi, action.eval(phi - shift), None
    b_phi, b_act = phi - shift, action.eval(phi - shift)
    points = set()
    for _ in range(12):
        a_phi, a_act, tp_a, e0a, e1a, a_w = ref.const_step_leapfrog(rs_a, action, a_phi, a_act, 1.5, 3, tr, a_w)
        b_phi, b_act, tp_b, e0b, e1b, _ = ref.const_step_leapfrog(rs_b, action, b_phi, b_act, 1.5, 3)
        assert tp_a == tp_b and abs(e0a - e0b) < 1e-12 and abs(e1a - e1b) < 1e-12
        assert np.abs(a_phi - (b_phi + shift)).max() < 1e-12 and abs(a_act - b_act) < 1e-12
        assert a_w == {"logdetJ": 0.0} or a_w == {"logdetJ": 0}
        points.add(tp_a)
    assert points == {0, 1}


def test_logdet_weight_enters_the_energy():
    action = _gauge_plus_skew()
    phi = np.random.RandomState(3).normal(0, 1.0, 20) + 0j

    class Scaled:
        """forward / backward are the identity, log det J = c everywhere."""

        def __init__(self, c):
            self.c = c

        def forward(self, p, a):
            return p, a, self.c

        def backward(self, p):
            return p, self.c
    for c in (0.0, 2.5):
        rs = np.random.RandomState(4)
        out = ref.const_step_leapfrog(rs, action, phi, action.eval(phi), 0.8, 5, Scaled(c))
        plain = ref.const_step_leapfrog(np.random.RandomState(4), action, phi, action.eval(phi), 0.8, 5)
        assert abs((out[3] - out[4]) - (plain[3] - plain[4])) < 1e-12  # same dE: the constant cancels
        assert abs(out[3] - plain[3] - c) < 1e-12
        assert out[2] == plain[2]
        assert out[5] == {"logdetJ": c}  # carried by the new stage, or kept by the old one


def _fermion_only():
    kappa = ref.honeycomb(2)
    nt, beta = 8, 3.0
    return ref.HubbardFermiAction(kappa * beta / nt, 0.0, -1, ref.EXP, ref.PARTICLE_HOLE), kappa.shape[0], nt


def test_jumps_leave_the_fermion_action_alone():
    action, nx, nt = _fermion_only()
    phi = np.random.RandomState(5).normal(0, 0.5, nx * nt) + 0j
    act = action.eval(phi)
    out, out_act, decisions = ref.two_pi_jumps(np.random.RandomState(6), action, phi, act, 6, 3)
    assert decisions == [1] * 6
    steps = (out - phi) / (2 * np.pi)
    assert np.abs(steps - np.round(steps)).max() < 1e-12 and np.abs(steps).sum() > 0
    assert abs(out_act - act) < 1e-9 * abs(act)
    for seed in range(4):
        out, out_act, took = ref.uniform_jump(np.random.RandomState(seed), action, phi, act, nt)
        assert took == 1 and abs(out_act - act) < 1e-9 * abs(act)
        assert np.allclose(np.abs(out - phi), 2 * np.pi / nt, rtol=0, atol=1e-12)


def test_jump_results_are_consistent():
    action = ref.SumAction(ref.HubbardGaugeAction(40.0), ref.HubbardGaugeAction(90.0))
    phi = np.random.RandomState(9).normal(0, 1.0, 32) + 0.1j
    rs = np.random.RandomState(10)
    act = action.eval(phi)
    seen = set()
    for _ in range(30):
        state = rs.get_state()
        phi1, act1, decisions = ref.two_pi_jumps(rs, action, phi, act, 3, 4)
        assert abs(act1 - action.eval(phi1)) < 1e-12 * abs(act1)
        seen.update(decisions)
        # the sites and signs of the first batch, read back from a replayed generator
        replay = np.random.RandomState()
        replay.set_state(state)
        sites = replay.choice(phi.size, 4, replace=False)
        signs = replay.choice((-2 * np.pi, 2 * np.pi), 4, replace=True)
        assert len(set(sites.tolist())) == 4 and set(np.abs(signs)) == {2 * np.pi}
        phi, act = phi1, act1
    assert seen == {0, 1}
    for _ in range(20):
        phi1, act1, took = ref.uniform_jump(rs, action, phi, act, 8)
        assert abs(act1 - action.eval(phi1)) < 1e-12 * abs(act1)
        phi, act = phi1, act1


def test_planted_last_proposal_mistake_is_caught():
    """The reference's defect (return the last proposal even if rejected) breaks the
    action-value consistency pin."""
    action = ref.SumAction(ref.HubbardGaugeAction(40.0), ref.HubbardGaugeAction(90.0))

    def last_proposal(rng, action, phi, act_val, n_batches, batch_size):
        phi = np.asarray(phi).copy()
        new = phi
        for _ in range(n_batches):
            sites = rng.choice(phi.size, batch_size, replace=False)
            shifts = rng.choice((-2 * np.pi, 2 * np.pi), batch_size, replace=True)
            new = phi.copy()
            new[sites] += shifts
            new_act = action.eval(new)
            if ref.binary_select(rng, act_val, new_act):
                phi, act_val = new, new_act
        return new, act_val, None
    rs = np.random.RandomState(10)
    phi = np.random.RandomState(9).normal(0, 1.0, 32) + 0.1j
    act = action.eval(phi)
    bad = 0
    for _ in range(30):
        phi1, act1, _ = last_proposal(rs, action, phi, act, 3, 4)
        bad += abs(act1 - action.eval(phi1)) > 1e-9 * abs(act1)
        phi, act = phi1, action.eval(phi1)
    assert bad > 0
