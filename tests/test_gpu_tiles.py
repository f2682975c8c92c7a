"""Every launch shape of the FP32 right-chain product of the mixed force (launch kind 30:
real x complex on k-contiguous slabs with the force diagonal fused), pinned one at a time
with HubbardFermiMatrix.pin_tile: tiles 0-3 (96 x 96 with 12 waves and 2 or 3 stages,
96 x 96 with 6 waves, 64 x 64); and the single-buffered register-staged tiles 25-27 of the
complex x complex products (kinds 3 and 18) in the mixed and the stable forms.

Shapes: the config-3 torus (Nx=288: full 96 x 96 tiles) and config-4 tori (Nx=128 and
Nx=32: partial tiles).  Every
pinned shape must reproduce the tuned force to 2e-6 (relative L2; the shapes differ only in
the order of FP32 sums) and, on one field (drawn from RandomState(1000) as in the config-3
parity test), the oracle's force to 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402

RCK_KIND = 30
RCK_TILES = range(4)


@pytest.mark.parametrize("L,nt,beta,U,batch", [(12, 128, 8.0, 3.0, 64), (8, 64, 8.0, 2.0, 12), (4, 32, 4.0, 2.0, 6)])
def test_rck_tiles_agree(L, nt, beta, U, batch):
    print(f"FP32 right-chain tiles: Nx={2 * L * L} Nt={nt} fields={batch}", flush=True)
    kt = isle.honeycomb(L).hopping() * beta / nt
    ut = U * beta / nt
    rs = np.random.RandomState(900 + L)
    phis = rs.normal(0, np.sqrt(ut), (batch, kt.shape[0] * nt)) + 0j
    # field 0 as in the config-3 parity test (seed 1000), where the mixed form is within 1e-5
    phis[0] = np.random.RandomState(1000).normal(0, np.sqrt(ut), kt.shape[0] * nt)
    dev = torch.as_tensor(phis, device="cuda")
    a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", group=32)
    tuned = a.force_batch(dev).cpu().numpy()
    want = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(phis[0])
    assert rel_l2(tuned[0], want) < 1e-5
    try:
        for tile in RCK_TILES:
            a.hfm.pin_tile(RCK_KIND, tile)
            got = a.force_batch(dev).cpu().numpy()
            errs = [rel_l2(got[b], tuned[b]) for b in range(batch)]
            err_oracle = rel_l2(got[0], want)
            print(f"  tile {tile}: max rel L2 vs tuned {max(errs):.2e}, vs oracle {err_oracle:.2e}", flush=True)
            assert max(errs) < 2e-6, tile
            assert err_oracle < 1e-5, tile
    finally:
        a.hfm.pin_tile(RCK_KIND, -1)


def test_pin_tile_rejects_bad_arguments():
    a = act.HubbardFermiAction(isle.honeycomb(2).hopping() * 0.25, 0.0, -1, act.HFAHopping.EXP, precision="mixed")
    with pytest.raises(Exception):
        a.hfm.pin_tile(RCK_KIND, 99)
    with pytest.raises(Exception):
        a.hfm.pin_tile(999, 0)


CC_KIND, CC_SUB_KIND = 3, 18  # complex x complex products: plain, and C -= A B
SB_TILES = (25, 26, 27)  # register-staged with one LDS buffer: 96x32 (K slab 16, 32), 64x64


@pytest.mark.parametrize("L,nt,batch", [(8, 64, 6), (12, 128, 4)])
def test_single_buffer_complex_tiles_agree(L, nt, batch):
    """The complex products of the mixed force's top level (kind 3) and of the stable form's
    elimination (kind 18) pinned to each single-buffered tile: the mixed force equals the tuned
    one to 2e-6 (its FP32 right chains start from the FP64 top level's factors, whose last-bit
    differences can move an FP32 rounding), the stable action to 1e-11 and the stable force to
    1e-9 (relative; FP64 throughout, the tiles differ only in the order of sums).  Nx=128 has
    partial 96-row tiles."""
    beta, U = 8.0, 2.0
    print(f"single-buffer complex tiles: Nx={2 * L * L} Nt={nt} fields={batch}", flush=True)
    kt = isle.honeycomb(L).hopping() * beta / nt
    ut = U * beta / nt
    rs = np.random.RandomState(700 + L)
    phis = rs.normal(0, np.sqrt(ut), (batch, kt.shape[0] * nt)) + 0j
    dev = torch.as_tensor(phis, device="cuda")
    mixed = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", group=32)
    stable = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="stable", group=32)
    f_mixed = mixed.force_batch(dev).cpu().numpy()
    f_stable = stable.force_batch(dev).cpu().numpy()
    s_stable = np.asarray(stable.eval_batch(dev))
    try:
        for tile in SB_TILES:
            for a in (mixed, stable):
                a.hfm.pin_tile(CC_KIND, tile)
                a.hfm.pin_tile(CC_SUB_KIND, tile)
            gm = mixed.force_batch(dev).cpu().numpy()
            gs = stable.force_batch(dev).cpu().numpy()
            es = np.asarray(stable.eval_batch(dev))
            em = max(rel_l2(gm[b], f_mixed[b]) for b in range(batch))
            ef = max(rel_l2(gs[b], f_stable[b]) for b in range(batch))
            ee = float(np.max(np.abs(es - s_stable) / np.maximum(1.0, np.abs(s_stable))))
            print(f"  tile {tile}: mixed force {em:.2e}, stable force {ef:.2e}, stable action {ee:.2e}", flush=True)
            assert em < 2e-6, tile
            assert ef < 1e-9, tile
            assert ee < 1e-11, tile
    finally:
        for a in (mixed, stable):
            a.hfm.pin_tile(CC_KIND, -1)
            a.hfm.pin_tile(CC_SUB_KIND, -1)
