"""Orders beyond one workgroup's registers: LU panels of more than 512 candidate rows run on
lu_panel_tall_kernel (panel in global memory), solves of more than 512 rows keep X in the
output array.  The reference has no size limit (blaze / LAPACK on the CPU).

* dense log det and inverse at n = 600, 1100, 2050 against numpy (slogdet, inv: textbook
  LAPACK routines): log det within 1e-11 relative, |A X - 1| / sqrt(n) within 1e-9; a
  pivot-heavy matrix (rows of a well-conditioned matrix shuffled so that the pivots come from
  far below the panel) at n = 800;
* logdetM (reference algorithm and grouped FP64) and the FP64 / mixed force at Nx = 576
  (honeycomb 12 x 24 torus, Nx % 32 == 0) and Nx = 578 (17 x 17, the unfused path), Nt = 8,
  against the oracle: log det 1e-10, FP64 force 1e-9, mixed force 1e-5;
* the stable form at Nx = 544 (16 x 17 cylinder, open), Nt = 64, groups of 16: log dets
  within 1e-10 of the oracle's stabilised restatement (its block-cyclic LU has 576 candidate
  rows per panel and a final 1088-row block);
* solveM at Nx = 578, Nt = 4 against the oracle's solveM (1e-9 relative).
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402


def _rand(rs, n):
    return rs.normal(size=(n, n)) + 1j * rs.normal(size=(n, n))


@pytest.mark.parametrize("n", [600, 1100, 2050])
def test_dense_logdet_and_inverse(n):
    print(f"dense LU beyond 512: n = {n}", flush=True)
    rs = np.random.RandomState(n)
    a = _rand(rs, n)
    sign, mag = np.linalg.slogdet(a)
    want = mag + 1j * np.angle(sign)
    got = isle.logdet(a)
    assert mod2pi_diff(got, want) < 1e-11 * abs(want), (got, want)
    x = isle.invert(a).cpu().numpy()
    assert np.linalg.norm(a @ x - np.eye(n)) / np.sqrt(n) < 1e-9


def test_dense_pivot_heavy():
    print("dense LU beyond 512: pivot-heavy n = 800", flush=True)
    n = 800
    rs = np.random.RandomState(3)
    base = _rand(rs, n) + 4 * np.sqrt(n) * np.eye(n)  # well conditioned
    perm = rs.permutation(n)
    a = base[perm]  # the pivot of every column sits in a shuffled row
    a[:40] *= 1e-3  # and the first rows are small: pivots come from far below the panel
    sign, mag = np.linalg.slogdet(a)
    want = mag + 1j * np.angle(sign)
    assert mod2pi_diff(isle.logdet(a), want) < 1e-11 * abs(want)
    x = isle.invert(a).cpu().numpy()
    assert np.linalg.norm(a @ x - np.eye(n)) / np.sqrt(n) < 1e-9


@pytest.mark.parametrize("li,lj", [(12, 24), (17, 17)])
def test_logdet_and_force_large_lattice(li, lj):
    nx, nt, beta, U = 2 * li * lj, 8, 1.0, 2.0
    print(f"logdetM and force at Nx={nx} Nt={nt}", flush=True)
    kt = ref.honeycomb(li, lj) * beta / nt
    phi = np.random.RandomState(nx).normal(0, np.sqrt(U * beta / nt), nx * nt) + 0j
    hfm = ref.HFMExp(kt, 0.0, -1)
    dev = torch.as_tensor(phi, device="cuda")
    for precision in ("fp64", "mixed"):
        a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision, group=4)
        got = a.logdets(dev)[0]
        for s in (0, 1):
            want = ref.logdetM(hfm, phi, s)
            assert mod2pi_diff(got[s], want) < 1e-10 * abs(want), (precision, s)
    f_want = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(phi)
    for precision, tol in (("fp64", 1e-9), ("mixed", 1e-5)):
        a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision, group=4)
        err = rel_l2(a.force(dev).cpu().numpy(), f_want)
        print(f"  {precision} force: {err:.2e}", flush=True)
        assert err < tol, precision


def test_stable_large_cylinder():
    kappa = ref.honeycomb(16, 17, periodic_j=False)
    nx, nt, beta, U, mu = kappa.shape[0], 64, 6.0, 2.0, 0.3
    assert nx == 544
    print(f"stable log det at Nx={nx} Nt={nt}", flush=True)
    kt, mt = kappa * beta / nt, mu * beta / nt
    phi = np.random.RandomState(9).normal(0, np.sqrt(U * beta / nt), nx * nt) + 0j
    hfm = ref.HFMDia(kt, mt, -1)
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=16)
    got = g.logdets(torch.as_tensor(phi, device="cuda"))[0]
    for s in (0, 1):
        want = ref.logdetM_stable(hfm, -1j * phi, s, g=16)
        assert mod2pi_diff(got[s], want) < 1e-10 * abs(want), s


def test_solveM_large_lattice():
    nx, nt = 578, 4
    print(f"solveM at Nx={nx} Nt={nt}", flush=True)
    kt = ref.honeycomb(17, 17) * 0.5 / nt
    rs = np.random.RandomState(5)
    phi = rs.normal(0, 0.3, nx * nt) + 0j
    rhs = rs.normal(size=(3, nx * nt)) + 1j * rs.normal(size=(3, nx * nt))
    hfm = isle.HubbardFermiMatrixExp(kt, 0.0, -1)
    got = isle.solveM(hfm, torch.as_tensor(phi, device="cuda"), isle.Species.PARTICLE, rhs)
    want = ref.solveM(ref.HFMExp(kt, 0.0, -1), phi, 0, rhs)
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    assert rel_l2(got, want) < 1e-9
