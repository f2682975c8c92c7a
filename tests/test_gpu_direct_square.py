"""GPU: HFAAlgorithm.DIRECT_SQUARE (Q = M_p M_h^T, HIP LU) against the oracle.

The reference's own test compares DIRECT_SINGLE with DIRECT_SQUARE
(tests/test_hubbardFermiAction.py:52-77, 122-147); here the GPU DIRECT_SQUARE action and
force equal the oracle's DIRECT_SQUARE restatement (dense Q from its block assembly,
numpy inverse; pinned in tests/test_oracle_*.py) and the GPU DIRECT_SINGLE ones, to
1e-10 relative (action: imaginary part modulo 2 pi), on a honeycomb 3x3 torus (Nx=18),
Nt=16, beta=4 (N = 288) for EXP / DIA hopping, particle-hole / spin basis, sigmaKappa =
+-1 and mu = 0 / 0.1, with a complex field.  The action goes through the block LU of Q
(logdetQ, tests/test_gpu_qlu.py); graph capture is refused.  Beyond N = 512 the force's
dense Q^{-1} runs on the tall-panel LU: N = 576 (Nx=18, Nt=32) against the oracle.
"""
import numpy as np
import pytest

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as A  # noqa: E402


@pytest.mark.parametrize("hop", [A.HFAHopping.EXP, A.HFAHopping.DIA])
@pytest.mark.parametrize("basis", [A.HFABasis.PARTICLE_HOLE, A.HFABasis.SPIN])
@pytest.mark.parametrize("sk,mu", [(-1, 0.0), (1, 0.1)])
def test_direct_square_matches_oracle(hop, basis, sk, mu):
    print(f"direct square {hop.name} {basis.name} sk={sk} mu={mu}", flush=True)
    kappa = ref.honeycomb(3) * 4.0 / 16
    rs = np.random.RandomState(21)
    phi = rs.normal(0, 0.5, 18 * 16) + 0.1j * rs.normal(0, 1, 18 * 16)
    sq = A.HubbardFermiAction(kappa, mu, sk, hop, basis, A.HFAAlgorithm.DIRECT_SQUARE)
    single = A.HubbardFermiAction(kappa, mu, sk, hop, basis, A.HFAAlgorithm.DIRECT_SINGLE)
    r = ref.HubbardFermiAction(kappa, mu, sk, int(hop), int(basis), ref.DIRECT_SQUARE)
    s_gpu, s_ref, s_single = sq.eval(phi), r.eval(phi), single.eval(phi)
    assert mod2pi_diff(s_gpu, s_ref) < 1e-10 * abs(s_ref)
    assert mod2pi_diff(s_gpu, s_single) < 1e-10 * abs(s_ref)
    f_gpu = sq.force(phi).cpu().numpy()
    assert rel_l2(f_gpu, r.force(phi)) < 1e-10
    assert rel_l2(f_gpu, single.force(phi).cpu().numpy()) < 1e-10


def test_direct_square_force_beyond_512():
    print("direct square force, N = 576", flush=True)
    kappa = ref.honeycomb(3) * 4.0 / 32
    rs = np.random.RandomState(22)
    phi = rs.normal(0, 0.35, 18 * 32) + 0.1j * rs.normal(0, 1, 18 * 32)
    sq = A.HubbardFermiAction(kappa, 0.0, -1, A.HFAHopping.EXP, A.HFABasis.PARTICLE_HOLE, A.HFAAlgorithm.DIRECT_SQUARE)
    r = ref.HubbardFermiAction(kappa, 0.0, -1, ref.EXP, ref.PARTICLE_HOLE, ref.DIRECT_SQUARE)
    assert rel_l2(sq.force(phi).cpu().numpy(), r.force(phi)) < 1e-9


def test_direct_square_limits():
    kappa = ref.honeycomb(3) * 0.25
    sq = A.HubbardFermiAction(kappa, 0.0, -1, A.HFAHopping.EXP, A.HFABasis.PARTICLE_HOLE,
                              A.HFAAlgorithm.DIRECT_SQUARE)
    with pytest.raises(ValueError):
        sq.logdets(np.zeros(18 * 16, complex))
    with pytest.raises(ValueError):
        isle.integrator.LeapfrogGraph(A.HubbardGaugeAction(0.5) + sq, 1, 18 * 16, 1.0, 4)
    with pytest.raises(ValueError):
        A.HubbardFermiAction(kappa, 0.0, -1, algorithm=A.HFAAlgorithm.DIRECT_SQUARE, precision="mixed")
