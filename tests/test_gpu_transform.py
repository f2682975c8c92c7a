"""GPU: leapfrog evolvers with transforms (src/isle/evolver/leapfrog.py:44-66,
src/isle/evolver/transform/identity.py, constantShift.py).

Honeycomb 3x3 torus (Nx=18), Nt=16, beta=4, U=2, mu=0, EXP hopping, particle-hole
basis, gauge + fermion action, trajectories of length 1 in 16 (identity) or 32 steps.
* The identity transform gives the chain without a transform, bit for bit, and its
  stages carry logdetJ = 0.
* A constant shift gives the oracle's chain (oracle.const_step_leapfrog with
  oracle.ConstantShift, pinned by the translation of the whole chain): same decisions,
  fields and action values within 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd.action import HFABasis, HFAHopping, HubbardGaugeAction, makeHubbardFermiAction  # noqa: E402
from isle_amd.evolver import ConstStepLeapfrog, EvolutionStage, NumpyRNG  # noqa: E402
from isle_amd.lattice import honeycomb  # noqa: E402
from isle_amd.transform import ConstantShift, Identity  # noqa: E402
from oracle import hubbard_ref as ref  # noqa: E402

UT = 2.0 * 4.0 / 16


def _actions():
    lat = honeycomb(3, nt=16)
    gpu = HubbardGaugeAction(UT) + makeHubbardFermiAction(lat, 4.0, 0.0, -1, HFAHopping.EXP, HFABasis.PARTICLE_HOLE)
    kt = lat.hopping() * 4.0 / 16
    cpu = ref.SumAction(ref.HubbardGaugeAction(UT), ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP))
    return lat, gpu, cpu


def test_identity_transform_is_no_transform():
    print("identity transform, Nx=18 Nt=16", flush=True)
    lat, gpu, _ = _actions()
    phi = np.random.RandomState(1075).normal(0, UT ** 0.5, lat.lattSize()) + 0j
    runs = []
    for tr in (None, Identity()):
        ev = ConstStepLeapfrog(gpu, 1.0, 16, NumpyRNG(5), transform=tr)
        stage = EvolutionStage(phi, complex(gpu.eval(phi)))
        chain = []
        for _ in range(4):
            stage = ev.evolve(stage)
            chain.append((np.asarray(stage.phi.cpu() if hasattr(stage.phi, "cpu") else stage.phi),
                          complex(stage.actVal), stage.trajPoint))
            if tr is not None:
                assert stage.logWeights.get("logdetJ") == 0
        runs.append(chain)
    for (p0, a0, t0), (p1, a1, t1) in zip(*runs):
        assert t0 == t1 and a0 == a1 and np.array_equal(p0, p1)


def test_constant_shift_matches_oracle():
    print("constant shift transform vs oracle, Nx=18 Nt=16", flush=True)
    lat, gpu, cpu = _actions()
    n = lat.lattSize()
    shift = np.random.RandomState(3).normal(0, 0.02, n) + 0j
    phi = np.random.RandomState(1075).normal(0, UT ** 0.5, n) + 0j
    ev = ConstStepLeapfrog(gpu, 1.0, 32, NumpyRNG(5), transform=ConstantShift(shift, gpu))
    stage = EvolutionStage(phi, complex(gpu.eval(phi - shift)))
    rs, tr = np.random.RandomState(5), ref.ConstantShift(shift, cpu)
    w_phi, w_act, w_weights = phi, cpu.eval(phi - shift), None
    points = []
    for k in range(5):
        stage = ev.evolve(stage)
        w_phi, w_act, tp, _, _, w_weights = ref.const_step_leapfrog(rs, cpu, w_phi, w_act, 1.0, 32, tr, w_weights)
        got = stage.phi.cpu().numpy() if hasattr(stage.phi, "cpu") else np.asarray(stage.phi)
        assert stage.trajPoint == tp, k
        assert np.abs(got - w_phi).max() < 1e-9 * np.abs(w_phi).max(), k
        assert abs(complex(stage.actVal) - w_act) < 1e-9 * abs(w_act), k
        points.append(tp)
    assert set(points) == {0, 1}
