"""Force evaluations of several batches of fields with different time extents on one lattice,
merged into shared launches (HubbardFermiAction.force_parts_multi, isle_force_parts_multi),
against the same batches evaluated one at a time and against the oracle.

Shapes of the config-4 family: Nx=32 (L=4, the fused FP32 right chains) with Nt in
{16, 32, 64} and batches of 3, 2 and 4 fields; Nx=18 (L=3, the unfused path) with Nt in
{8, 16}.  beta = Nt/8, U = 2, EXP, particle-hole.  Tolerances (relative L2 per field): FP64
1e-12 between merged and separate (merged steps may run other tile shapes, which changes the
order of FP64 sums), 1e-8 against the oracle (the reference algorithm's own product of all
slices, summed in another order; 6.7e-10 at Nt=64 here); mixed 2e-6 between them (FP32
sums), 1e-5 against the oracle.
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402


@pytest.mark.parametrize("L,nts,batches", [(4, (16, 32, 64), (3, 2, 4)), (3, (8, 16), (2, 3))])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-6)])
def test_merged_force_matches_separate(L, nts, batches, precision, tol):
    print(f"merged force: Nx={2 * L * L} Nt={nts} batches={batches} {precision}", flush=True)
    kt = ref.honeycomb(L) / 8.0  # kappa beta / Nt with beta = Nt / 8
    ut = 2.0 / 8.0
    rs = np.random.RandomState(60 + L)
    fields = [rs.normal(0, np.sqrt(ut), (b, kt.shape[0] * nt)) + 0j for nt, b in zip(nts, batches)]
    dev = [torch.as_tensor(f, device="cuda") for f in fields]
    a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision, group=32)
    merged = [p.cpu().numpy() for p in a.force_parts_multi(dev)]
    again = [p.cpu().numpy() for p in a.force_parts_multi(dev)]  # the cached merged operation
    outs = [torch.empty_like(d) for d in dev]
    a.force_batch_multi(dev, outs)
    for f, d, m, m2, o in zip(fields, dev, merged, again, outs):
        sep = a.force_parts(d).cpu().numpy()
        assert np.array_equal(m, m2)
        for b in range(f.shape[0]):
            assert rel_l2(m[b], sep[b]) < tol, b
        full = a.force_batch(d).cpu().numpy()
        assert rel_l2(o.cpu().numpy(), full) < tol
        want = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(f[0])
        err = rel_l2(o.cpu().numpy()[0], want)
        print(f"  Nt={f.shape[1] // kt.shape[0]}: merged vs oracle {err:.2e}", flush=True)
        assert err < (1e-8 if precision == "fp64" else 1e-5)


def test_merged_force_rejects_repeated_shape():
    kt = ref.honeycomb(3) / 8.0
    a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed")
    d = torch.zeros((2, 18 * 16), dtype=torch.complex128, device="cuda")
    with pytest.raises(Exception):
        a.force_parts_multi([d, d.clone()])
