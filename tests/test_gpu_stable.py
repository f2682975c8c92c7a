"""GPU parity of precision="stable" (FP64 group products combined by the structured pivoted
elimination of their block-cyclic matrix) against the oracle.

The ill-conditioned case is a small version of the nanotube workload (honeycomb cylinder,
DIA hopping, spin basis, mu != 0, beta = 10): there the reference's own product of all
time slices loses the particle log-determinant completely (tests assert that it does), so
the comparison is with the oracle's stabilised restatement (``logdetM_stable``,
``force_part_stable``: dense GEPP / inverse of the same block-cyclic matrix).
Tolerances: logdet 1e-10 relative (imaginary parts modulo 2 pi); force parts 1e-9 relative
(L2) for groups of 16 slices and 1e-6 for groups of 32 (a group of 32 slices has condition
number ~1e6 here, which FP64 local chains carry with ~1e-10 relative error per group).
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402


def _cylinder(nt, seed=7, beta=10.0, U=2.5, mu=0.5):
    kappa = ref.honeycomb(4, 8, periodic_j=False)  # nx = 64
    nx = kappa.shape[0]
    kt, mt, ut = kappa * beta / nt, mu * beta / nt, U * beta / nt
    phi = np.random.RandomState(seed).normal(0, np.sqrt(ut), nx * nt) + 0j
    return kt, mt, phi


@pytest.mark.parametrize("nt,beta,group", [(128, 10.0, 16), (128, 10.0, 32), (64, 5.0, 32)])
def test_stable_logdet_ill_conditioned(nt, beta, group):
    kt, mt, phi = _cylinder(nt, beta=beta)
    hfm = ref.HFMDia(kt, mt, -1)
    aux = -1j * phi
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=group)
    assert g._group(nt) == group
    ld = g.logdets(torch.as_tensor(phi, device="cuda"))[0]
    for species in (0, 1):
        want = ref.logdetM_stable(hfm, aux, species, 8)
        assert mod2pi_diff(ld[species], want) < 1e-10 * abs(want), (species, ld[species], want)
    if beta == 10.0:  # the reference's own product is far off on this field (the test is not vacuous)
        plain = ref.logdetM(hfm, aux, 0)
        assert mod2pi_diff(plain, ref.logdetM_stable(hfm, aux, 0, 8)) > 1e-3 * abs(plain)


@pytest.mark.parametrize("nt,beta,group,tol", [(128, 10.0, 16, 1e-9), (128, 10.0, 32, 1e-6), (64, 5.0, 32, 1e-6)])
def test_stable_force_ill_conditioned(nt, beta, group, tol):
    kt, mt, phi = _cylinder(nt, beta=beta)
    hfm = ref.HFMDia(kt, mt, -1)
    aux = -1j * phi
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=group)
    parts = g.force_parts(torch.as_tensor(phi, device="cuda"))[0].cpu().numpy()
    for species in (0, 1):
        want = ref.force_part_stable(hfm, aux, species, 8)
        err = rel_l2(parts[species], want)
        assert err < tol, (species, err)
    # action-level force of the spin basis: f_h(aux) - f_p(aux)
    want = ref.force_part_stable(hfm, aux, 1, 8) - ref.force_part_stable(hfm, aux, 0, 8)
    assert rel_l2(g.force(phi).cpu().numpy(), want) < tol


def test_stable_matches_plain_where_plain_is_accurate():
    """Well-conditioned EXP torus (particle-hole basis, mu = 0): stable == reference algorithm."""
    kappa = ref.honeycomb(4, 8)  # nx = 64, periodic
    nt, beta = 64, 4.0
    kt = kappa * beta / nt
    phi = np.random.RandomState(11).normal(0, 0.3, 64 * nt) + 0j
    r = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP, ref.PARTICLE_HOLE)
    for group in (8, 16, 32):
        g = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="stable", group=group)
        assert mod2pi_diff(g.eval(phi), r.eval(phi)) < 1e-10 * abs(r.eval(phi)), group
        assert rel_l2(g.force(phi).cpu().numpy(), r.force(phi)) < 1e-9, group


def test_stable_batch_matches_single():
    kt, mt, phi0 = _cylinder(64, seed=3, beta=5.0)
    _, _, phi1 = _cylinder(64, seed=4, beta=5.0)
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=16)
    both = torch.as_tensor(np.stack([phi0, phi1]), device="cuda")
    ld = g.logdets(both)
    fp = g.force_parts(both).cpu().numpy()
    for i, phi in enumerate((phi0, phi1)):
        one = g.logdets(torch.as_tensor(phi, device="cuda"))[0]
        assert np.max(np.abs(ld[i] - one)) < 1e-12 * np.max(np.abs(one))
        f1 = g.force_parts(torch.as_tensor(phi, device="cuda"))[0].cpu().numpy()
        assert np.max(np.abs(fp[i] - f1)) < 1e-9 * np.max(np.abs(f1))


def test_stable_rejects_unsupported_nx():
    kappa = ref.honeycomb(3)  # nx = 18: not a multiple of 32
    g = act.HubbardFermiAction(kappa * 0.25, 0.0, -1, act.HFAHopping.EXP, precision="stable", group=4)
    with pytest.raises(Exception):
        g.logdets(torch.zeros(18 * 16, dtype=torch.complex128, device="cuda"))


@pytest.mark.parametrize("nt,beta,group", [(128, 10.0, 16), (128, 10.0, 32)])
def test_stable_fp32_right_chains(nt, beta, group):
    """right_chains="fp32": the stable force with FP32 right chains (fused diagonal) on the
    ill-conditioned cylinder, within 1e-5 of the stabilised oracle.  The FP32 chains keep
    the mixed form's conditioning cap (g * spread <= 12): a requested group of 32 slices is
    cut to 16 here (spread 0.56 per slice; at 32 slices the error is 1.8e-4)."""
    print(f"stable force, FP32 right chains, Nx=64 Nt={nt} group {group}", flush=True)
    kt, mt, phi = _cylinder(nt, beta=beta)
    hfm = ref.HFMDia(kt, mt, -1)
    aux = -1j * phi
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=group,
                               right_chains="fp32")
    assert g._group(nt) == 16
    parts = g.force_parts(torch.as_tensor(phi, device="cuda"))[0].cpu().numpy()
    for species in (0, 1):
        want = ref.force_part_stable(hfm, aux, species, 8)
        err = rel_l2(parts[species], want)
        print(f"  species {species}: L2 {err:.2e}", flush=True)
        assert err < 1e-5, (species, err)


def test_mixed_stable_logdet_chunks(monkeypatch):
    """precision="mixed" with logdet="stable" (the Metropolis log det through the stabilising
    block-cyclic LU, groups of 8 slices) on a config-4 torus (Nx=32, Nt=32, beta=4, U=2): the
    batch in one call, in chunks of two fields (the workspace bound lowered), and the oracle's
    log det of each field agree (1e-12 between the two GPU runs, 1e-10 with the oracle)."""
    print("mixed action, stable log det: Nx=32 Nt=32, 5 fields, whole and in chunks", flush=True)
    nt, beta, U = 32, 4.0, 2.0
    kt = ref.honeycomb(4) * beta / nt
    rs = np.random.RandomState(31)
    phis = rs.normal(0, np.sqrt(U * beta / nt), (5, 32 * nt)) + 0j
    a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", logdet="stable")
    assert a._logdet_group(nt) == 8
    dev = torch.as_tensor(phis, device="cuda")
    whole = a.logdets(dev)
    per_field = 136 * 32 * 32 * (nt // 8) * 2
    monkeypatch.setattr(act, "_STABLE_LOGDET_BYTES", 2 * per_field)
    chunked = a.logdets(dev)
    assert np.abs(chunked - whole).max() < 1e-12 * np.abs(whole).max()
    hfm = ref.HFMExp(kt, 0.0, -1)
    for b in range(5):
        for s in (0, 1):
            want = ref.logdetM(hfm, phis[b], s)
            assert mod2pi_diff(whole[b, s], want) < 1e-10 * abs(want), (b, s)


def test_mixed_stable_logdet_arguments():
    kt = ref.honeycomb(4) * 0.125
    with pytest.raises(ValueError):
        act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="fp64", logdet="stable")
    with pytest.raises(ValueError):
        act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed", logdet="lu")
