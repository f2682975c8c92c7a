"""GPU: the mixed force's memory schedules give the force of the default schedule.

With the diagonal fused into the FP32 right chains, a chain keeps only its last two
products; with recompute=k the FP64 suffix chains are kept every k slices and the rest is
recomputed by the same kernels.  Each action tunes its own tile shapes, so two schedules
may sum in different orders: the forces must agree to 2e-6 (FP32 rounding; a wrong
matrix in a schedule gives O(1) differences) and both be within 1e-5 of the oracle.  Workloads: honeycomb 6x6 torus (Nx=72), Nt=64, beta=6,
U=3 (groups of 16; k = 2, 4, 8) and the 12x12 torus (Nx=288), Nt=128, beta=8, U=3, two
fields (groups of 32; k = 8), both species, EXP hopping, particle-hole basis."""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402


@pytest.mark.parametrize("l,nt,beta,u,group,ks,batch", [(6, 64, 6.0, 3.0, 16, (2, 4, 8), 3),
                                                        (12, 128, 8.0, 3.0, 32, (8,), 2)])
def test_recompute_schedule_equals_default(l, nt, beta, u, group, ks, batch):
    nx = 2 * l * l
    print(f"recompute schedules, Nx={nx} Nt={nt} group {group} k {ks}", flush=True)
    kt = ref.honeycomb(l) * beta / nt
    rs = np.random.RandomState(31)
    phis = torch.as_tensor(rs.normal(0, np.sqrt(u * beta / nt), (batch, nx * nt)) + 0j, device="cuda")
    base = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", group=group)
    want = base.force_batch(phis).cpu().numpy()
    oracle = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(phis[0].cpu().numpy())
    assert np.linalg.norm(want[0] - oracle) < 1e-5 * np.linalg.norm(oracle)
    for k in ks:
        a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", group=group, recompute=k)
        got = a.force_batch(phis).cpu().numpy()
        diff = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"  k={k}: |F_k - F| / |F| = {diff:.2e}", flush=True)
        assert diff < 2e-6, k
