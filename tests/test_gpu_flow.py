"""GPU: the adaptive RK4 holomorphic flow (rungeKutta4Flow) against the oracle's.

Honeycomb 3x3 torus (Nx=18), Nt=16, beta=4, U=2, mu=0.2 (a complex action), EXP hopping,
particle-hole basis, gauge + fermion action; the flow starts from N(0, 0.5) plus
0.05 i N(0, 1) and moves the imaginary part.  The oracle's RK4 is pinned by closed forms
in tests/test_oracle_flow.py.

* Fixed steps (a tolerance no step misses, no growth): five steps of 0.01; the GPU flow
  (force and action from the HIP kernels) and the oracle's end at the same field and
  action to 1e-9, for the classic and 3/8-rule steps in both directions, with the
  reference's stage coefficients (the default) and with the textbook ones.  (The
  force itself agrees to 1e-11 at the start field; the 3/8 rule's stages combine
  differences of stage slopes, which raises the rounding difference to ~1e-10.)
* Adaptive steps (tolerance 1e-3): the step sizes depend continuously on the measured
  error, so the two flows' step sequences agree to rounding, not exactly: the reached
  flow time and the field agree to 1e-6.  The flow conserves Im S to the step tolerance
  and does not lower Re S.
"""
import numpy as np
import pytest

from oracle import hubbard_ref as ref

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd.action import HFABasis, HFAHopping, HubbardGaugeAction, makeHubbardFermiAction  # noqa: E402
from isle_amd.lattice import honeycomb  # noqa: E402


def _actions():
    lat = honeycomb(3, nt=16)
    beta, U, nt = 4.0, 2.0, 16
    mt = 0.2 * beta / nt
    gpu = HubbardGaugeAction(U * beta / nt) + makeHubbardFermiAction(lat, beta, mt, -1, HFAHopping.EXP,
                                                                     HFABasis.PARTICLE_HOLE)
    kt = lat.hopping() * beta / nt
    cpu = ref.SumAction(ref.HubbardGaugeAction(U * beta / nt),
                        ref.HubbardFermiAction(kt, mt, -1, ref.EXP, ref.PARTICLE_HOLE))
    return lat, gpu, cpu


def _start(lat):
    rs = np.random.RandomState(12)
    return rs.normal(0, 0.5, lat.lattSize()) + 0.05j * rs.normal(0, 1, lat.lattSize())


@pytest.mark.parametrize("variant", ["reference", "textbook"])
@pytest.mark.parametrize("n,direction", [(0, 1.0), (1, 1.0), (0, -1.0), (1, -1.0)])
def test_fixed_step_flow_matches_oracle(n, direction, variant):
    print(f"rk4 flow, fixed steps, n={n} direction={direction} variant={variant}, Nx=18 Nt=16", flush=True)
    lat, gpu, cpu = _actions()
    phi = _start(lat)
    phi_g, act_g, t_g = isle.rungeKutta4Flow(phi, gpu, 0.05, 0.01, n=n, direction=direction, imActTolerance=10.0,
                                             adaptThreshold=0.0, variant=variant)
    phi_c, act_c, t_c = ref.rk4_flow(phi, cpu, 0.05, 0.01, n=n, direction=direction, tolerance=10.0, threshold=0.0,
                                     variant=variant)
    assert t_g == pytest.approx(0.05, abs=1e-15) and t_c == pytest.approx(0.05, abs=1e-15)
    phi_g = phi_g.cpu().numpy()
    assert np.abs(phi_g - phi_c).max() < 1e-9 * np.abs(phi_c).max()
    assert abs(act_g - act_c) < 1e-9 * abs(act_c)
    assert np.abs(phi_g.imag - phi.imag).max() > 1e-3  # the flow moved the imaginary part
    f_g, f_c = gpu.force(phi).cpu().numpy(), cpu.force(phi)
    assert np.abs(f_g - f_c).max() < 1e-11 * np.abs(f_c).max()


def test_adaptive_flow():
    lat, gpu, cpu = _actions()
    phi = _start(lat)
    phi_g, act_g, t_g = isle.rungeKutta4Flow(phi, gpu, 0.05, 0.01)
    phi_c, act_c, t_c = ref.rk4_flow(phi, cpu, 0.05, 0.01)
    assert t_g > 0.02 and abs(t_g - t_c) < 1e-6 * t_c
    phi_g = phi_g.cpu().numpy()
    assert np.abs(phi_g - phi_c).max() < 1e-6 * np.abs(phi_c).max()
    s0 = complex(gpu.eval(phi))
    assert abs(np.exp(1j * (act_g.imag - s0.imag)) - 1) < 1e-2  # Im S conserved to the step tolerance
    assert act_g.real >= s0.real  # d Re S / dt = |dS/dphi|^2 >= 0
