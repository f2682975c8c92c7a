"""GPU parity at the benchmark workloads' own shapes, against the oracle.

Config 3 (honeycomb 12x12 torus, Nx=288, Nt=128, beta=8, U=3, mu=0, EXP, particle-hole
basis, sigmaKappa=-1; fields N(0, U beta/Nt) from RandomState(1000..1003)):
* log det M of both species within 1e-10 (relative; imaginary parts modulo 2 pi) of
  oracle.logdetM, for the plain FP64 form (the reference's product) and for both FP64 forms
  the mixed action can evaluate for Metropolis: the grouped products multiplied together,
  and products of 8 slices combined through the stabilising block-cyclic LU;
* the mixed-precision force (FP32 right chains on MFMA) within 1e-5 of the oracle's force
  (relative L2 norm and max-norm ratio), the FP64 force within 1e-6.  The oracle's own
  force is the reference's unstabilised FP64 algorithm, 7e-8 from the stabilised
  restatement at seed 1000; its log dets 1e-11.

Config 4 (256 chains, L in {3,4,6,8} with Nx = 2L^2, Nt in {16,32,64}, beta = Nt/8, U=2,
mu=0, EXP, chain c seeded NumpyRNG(1000 + c), the 12 (L, Nt) buckets of the benchmark's
draw): the ensemble advances all 256 chains one trajectory (length 1, 20 steps; mixed
force, FP64 action); for the chains of every bucket, in order, the oracle's
const_step_leapfrog from the same seed gives the same accept decision, and Delta H within
5e-5 max(1, |Delta H|) until one chain per bucket has a regular trajectory (|Delta H| < 50;
2 of the 12 first chains and most random starts at Nx=128, Nt=64 pass near a zero of det M,
where Delta H ~ 1e3..1e4 differs between any two evaluation orders and only the decision is
compared).  The bucket's mixed force at
its first chain's start field is within 1e-5 and its FP64 action within 1e-10 of the
oracle's.

Config 5 (honeycomb cylinder, 12 cells around, 16 along, open: Nx=384; Nt=256, beta=10,
U=2.5, mu=0.5, DIA, spin basis; field from RandomState(7)): the stable log dets of both
species within 1e-10 of oracle.logdetM_stable, the stable force parts within 1e-7 of
oracle.force_part_stable with FP64 right chains and within 1e-5 with FP32 right chains on
MFMA (groups of 32 slices on both sides; the oracle's g=32 and g=16 agree to 8e-11 there).  The reference's own product is checked to be off there, so the
comparison is not vacuous.
"""
import numpy as np
import pytest
import torch

from oracle import hubbard_ref as ref
from tests.helpers import mod2pi_diff, rel_l2

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import action as act  # noqa: E402


def _max_rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("seed", [1000, 1001, 1002, 1003])
def test_config3_logdet_and_force(seed):
    print(f"config 3 vs oracle, seed {seed}: Nx=288 Nt=128", flush=True)
    kt = ref.honeycomb(12) * 8.0 / 128
    phi = np.random.RandomState(seed).normal(0, np.sqrt(3.0 * 8.0 / 128), 288 * 128) + 0j
    hfm = ref.HFMExp(kt, 0.0, -1)
    want = [ref.logdetM(hfm, phi, s) for s in (0, 1)]
    dev = torch.as_tensor(phi, device="cuda")
    for precision, form in (("fp64", "grouped"), ("mixed", "grouped"), ("mixed", "stable")):
        a = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision=precision, logdet=form)
        got = a.logdets(dev)[0]
        for s in (0, 1):
            assert mod2pi_diff(got[s], want[s]) < 1e-10 * abs(want[s]), (precision, form, s, got[s], want[s])
    f_want = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP).force(phi)
    f_mixed = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP, precision="mixed").force(dev).cpu().numpy()
    f_plain = act.HubbardFermiAction(kt, 0.0, -1, act.HFAHopping.EXP).force(dev).cpu().numpy()
    print(f"  mixed force: L2 {rel_l2(f_mixed, f_want):.2e} max {_max_rel(f_mixed, f_want):.2e}; "
          f"FP64: L2 {rel_l2(f_plain, f_want):.2e}", flush=True)
    assert rel_l2(f_mixed, f_want) < 1e-5 and _max_rel(f_mixed, f_want) < 1e-5
    assert rel_l2(f_plain, f_want) < 1e-6


def test_config4_buckets_against_oracle():
    from isle_amd.ensemble import Chain, EnsembleLeapfrog
    from isle_amd.evolver import NumpyRNG
    print("config 4 vs oracle: 256 chains, 12 buckets, one trajectory", flush=True)
    pick = np.random.RandomState(4)  # the benchmark's draw of the chains' shapes
    ls = pick.choice([3, 4, 6, 8], 256)
    nts = pick.choice([16, 32, 64], 256)
    lats = {l: isle.honeycomb(int(l)).hopping() for l in (3, 4, 6, 8)}
    chains = [Chain(lats[int(l)], int(nt), nt / 8.0, 2.0, 0.0, -1, NumpyRNG(1000 + c))
              for c, (l, nt) in enumerate(zip(ls, nts))]
    ens = EnsembleLeapfrog(chains, 1.0, 20, precision="mixed", group=32, graph=True)
    assert len(ens.buckets) == 12
    starts = [c.phi0.copy() for c in chains]
    ens.evolve()
    dh, points = ens.deltaH[-1], ens.trajPoints[-1]
    for idx, bucket in zip(ens.index, ens.buckets):
        regular = 0
        for c in idx[:8]:
            ch = chains[c]
            kt = ch.kappa * ch.beta / ch.nt
            oracle = ref.HubbardFermiAction(kt, 0.0, -1, ref.EXP)
            full = ref.SumAction(ref.HubbardGaugeAction(ch.utilde), oracle)
            rs = np.random.RandomState(1000 + c)
            phi0 = rs.normal(0, np.sqrt(ch.utilde), ch.nx * ch.nt) + 0j
            assert np.array_equal(phi0, starts[c])
            _, _, tp, e0, e1, _ = ref.const_step_leapfrog(rs, full, phi0, full.eval(phi0), 1.0, 20)
            want_dh = float(np.real(e1 - e0))
            print(f"  Nx={ch.nx} Nt={ch.nt} chain {c}: dH {dh[c]:.6f} oracle {want_dh:.6f}", flush=True)
            assert points[c] == tp, (ch.nx, ch.nt, c)
            if abs(want_dh) < 50.0:  # a trajectory that passes no near-zero of det M
                assert abs(dh[c] - want_dh) < 5e-5 * max(1.0, abs(want_dh)), (ch.nx, ch.nt, c, dh[c], want_dh)
                regular += 1
            if c == idx[0]:
                dev = torch.as_tensor(phi0, device="cuda")
                f = bucket.fermion.force(dev).cpu().numpy()
                f_want = oracle.force(phi0)
                assert rel_l2(f, f_want) < 1e-5 and _max_rel(f, f_want) < 1e-5, (ch.nx, ch.nt)
                s_want = oracle.eval(phi0)
                assert mod2pi_diff(complex(bucket.fermion.eval(dev)), s_want) < 1e-10 * abs(s_want), (ch.nx, ch.nt)
            if regular:
                break
        assert regular, (bucket.chains[0].nx, bucket.chains[0].nt)


def test_config5_stable_against_oracle():
    print("config 5 vs stabilised oracle: Nx=384 Nt=256 DIA spin mu=0.5", flush=True)
    kappa = ref.honeycomb(12, 16, periodic_j=False)
    nx, nt, beta, U, mu = 384, 256, 10.0, 2.5, 0.5
    kt, mt = kappa * beta / nt, mu * beta / nt
    phi = np.random.RandomState(7).normal(0, np.sqrt(U * beta / nt), nx * nt) + 0j
    aux = -1j * phi
    hfm = ref.HFMDia(kt, mt, -1)
    g = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=32)
    dev = torch.as_tensor(phi, device="cuda")
    got = g.logdets(dev)[0]
    for s in (0, 1):
        want = ref.logdetM_stable(hfm, aux, s, g=32)
        print(f"  species {s}: logdet {got[s]} oracle {want}", flush=True)
        assert mod2pi_diff(got[s], want) < 1e-10 * abs(want)
        if s == 0:  # the reference's own product of all slices is off here
            plain = ref.logdetM(hfm, aux, 0)
            assert mod2pi_diff(plain, want) > 1e-3 * abs(want)
    parts = g.force_parts(dev)[0].cpu().numpy()
    g32 = act.HubbardFermiAction(kt, mt, -1, act.HFAHopping.DIA, act.HFABasis.SPIN, precision="stable", group=32,
                                 right_chains="fp32")
    parts32 = g32.force_parts(dev)[0].cpu().numpy()
    for s in (0, 1):
        want = ref.force_part_stable(hfm, aux, s, g=32)
        err, err32 = rel_l2(parts[s], want), rel_l2(parts32[s], want)
        print(f"  species {s}: force part L2 {err:.2e} (FP64 chains), {err32:.2e} (FP32 chains)", flush=True)
        assert err < 1e-7 and err32 < 1e-5
