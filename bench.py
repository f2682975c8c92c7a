#!/usr/bin/env python3
"""HMC trajectories/s on the Nx=288, Nt=128 honeycomb torus: batched mixed precision
(FP32 MFMA force, FP64 action) against the plain FP64 form, with the force error.

Workload: 12x12-cell honeycomb torus (Nx=288), Nt=128, beta=8, U=3, mu=0, EXP hopping,
particle/hole basis, sigmaKappa=-1; HMC trajectories of length 1 with 40 leapfrog steps
(41 force evaluations and 1 action evaluation each); fields start from N(0, U*beta/Nt)
with per-chain seeds.  A "step" is one trajectory of every chain of the batch.

Prints one JSON line (see the fields below).  Only one GPU is used; --gpus > 1 exits.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "HMC traj/s at Nx=288 Nt=128: batched FP32-force vs plain FP64; force rel err"


def progress(msg):
    """One line on stderr as each part of the default bench ends (stdout carries the JSON)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
L, NT, BETA, U, NSTEP, LENGTH = 12, 128, 8.0, 3.0, 40, 1.0
THERM_LENGTH, THERM_NSTEP = 0.25, 100  # thermalisation trajectories: step 0.0025 instead of 0.025
REV_CHAINS = 8
PEAK = {"f32": 157.3, "f64": 78.6}  # dense MFMA TFLOP/s, MI355X
KIND_NAMES = ["gemm64_rc_pre", "gemm64_rc_row", "gemm64_cr_pre", "gemm64_cc", "gemm32_rc_row", "gemm32_rc_pre",
              "scale64", "scale32", "lu", "lu_solve", "diag64", "diag32", "copy64", "cvt32to64", "cvt64to32",
              "phase", "tcvt64to32", "diag64x32", "gemm64_cc_sub", "lu_shape", "block", "sigma", "gather",
              "tblock", "trsm", "gemm32_rc_pre_diag", "dreduce", "rgather", "gemm32_rc_diag", "tscale64to32",
              "gemm32_rck_diag", "colscale64to32", "lu_shape_defer",
              "gemm64_cc4", "gemm64_cc4_sub"]


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed trajectories (per chain) of the batched form")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chains", type=int, default=None,
                    help="independent chains advanced together (default 128; 16 for --config5)")
    ap.add_argument("--group", type=int, default=None,
                    help="time slices per group (default 32)")
    ap.add_argument("--plain-steps", type=int, default=3, help="timed trajectories of the plain FP64 forms")
    ap.add_argument("--therm", type=int, default=8, help="thermalisation trajectories (length 0.25, 100 steps) "
                    "before the timed ones")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--logdet", choices=["stable", "grouped"], default="stable",
                    help="the mixed form's Metropolis log det: groups of 8 slices through the stabilising "
                         "block-cyclic LU (default), or the force's group products multiplied together")
    ap.add_argument("--no-graph", action="store_true", help="launch the integration eagerly, not from a HIP graph")
    ap.add_argument("--no-merge", action="store_true",
                    help="config 4: every (Nx, Nt) bucket with its own launches, instead of the buckets of one "
                         "lattice sharing theirs")
    ap.add_argument("--streams", type=int, default=2,
                    help="parts of the batch integrated on streams of their own inside the graph")
    ap.add_argument("--config4", action="store_true",
                    help="the 256-chain ensemble of uneven lattices (L in {3,4,6,8}, Nt in {16,32,64}, beta=Nt/8, "
                         "U=2, 20 steps): prints chain-trajectories/s instead of the config-3 line")
    ap.add_argument("--right-chains", choices=("fp64", "fp32"), default="fp32",
                    help="config 5: precision of the stable force's right chains (FP32 on MFMA by default)")
    ap.add_argument("--config5", action="store_true",
                    help="the nanotube workload (12 x 16 cell honeycomb cylinder, Nx=384, Nt=256, beta=10, U=2.5, "
                         "mu=0.5, DIA, spin basis, 60 steps) with the stable block-cyclic form: prints "
                         "chain-trajectories/s for --chains chains instead of the config-3 line")
    ap.add_argument("--propagator", action="store_true",
                    help="the all-to-all propagator measurement (M^-1 of one species, every spacetime source) on "
                         "the Nx=288, Nt=128 torus: prints propagators/s instead of the HMC line")
    ap.add_argument("--logdetq", action="store_true",
                    help="log det Q through the block LU of Q = M_p M_h^T (DIRECT_SQUARE's action) on the "
                         "Nx=288, Nt=128 torus: prints log dets/s instead of the HMC line")
    ap.add_argument("--force-only", action="store_true",
                    help="profiling aid: run --steps batched mixed forces and one 1 GiB device copy "
                         "(the calibration kernel for HBM byte counters), print timings, no HMC")
    return ap.parse_args()


def config4(args):
    """256 independent chains, L drawn from {3,4,6,8} (Nx = 2L^2) and Nt from {16,32,64}
    with a seeded RNG, beta = Nt/8, U = 2, mu = 0, EXP, mixed force, HMC length 1 with 20
    steps; chain c starts from N(0, U beta/Nt) drawn from NumpyRNG(1000 + c).  One step =
    one trajectory of every chain (integration of all buckets replayed from one HIP graph)."""
    import torch
    import isle_amd as isle
    from isle_amd.ensemble import Chain, EnsembleLeapfrog
    from isle_amd.evolver import NumpyRNG
    pick = np.random.RandomState(4)
    n_chains = 256
    ls = pick.choice([3, 4, 6, 8], n_chains)
    nts = pick.choice([16, 32, 64], n_chains)
    lats = {l: isle.honeycomb(int(l)).hopping() for l in (3, 4, 6, 8)}
    chains = [Chain(lats[int(l)], int(nt), nt / 8.0, 2.0, 0.0, -1, NumpyRNG(1000 + c))
              for c, (l, nt) in enumerate(zip(ls, nts))]
    ens = EnsembleLeapfrog(chains, 1.0, 20, precision="mixed", group=32, graph=not args.no_graph,
                           merge=not args.no_merge)
    for _ in range(args.warmup):
        ens.evolve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ens.evolve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pts = np.array(ens.trajPoints[args.warmup:], dtype=float)
    dh = np.array(ens.deltaH[args.warmup:], dtype=float)
    launch = None
    if not args.no_graph:
        # after the timed region: the integration graph alone, and each stream's group of
        # buckets alone (one bucket per stream without merging)
        whole, alone = ens.time_groups()
        times = [t for _, t in alone]
        launch = {"graph_ms": whole, "streams": len(alone), "merged": not args.no_merge,
                  "longest_stream_alone_ms": max(times), "sum_streams_alone_ms": sum(times),
                  "overlap": sum(times) / whole, "streams_alone_ms": {lab: round(t, 3) for lab, t in alone}}
    # algorithmic work of one trajectory of every chain (per species: force
    # (2Nt-1) 4Nx^3 + 8Nx^3 + 8 Nt Nx^2, eval (Nt-1) 4Nx^3 + 8/3 Nx^3; 21 forces + 1 eval)
    def traj_flops(nx, nt):
        force = (2 * nt - 1) * 4 * nx ** 3 + 8 * nx ** 3 + 8 * nt * nx ** 2
        ev = (nt - 1) * 4 * nx ** 3 + 8 / 3 * nx ** 3
        return 2 * (21 * force + ev)
    work = sum(traj_flops(c.nx, c.nt) for c in chains)
    achieved = work * args.steps / dt / 1e12
    cpu = None
    if not args.no_cpu:
        # the oracle on the host: one force and one eval per (Nx, Nt) bucket on its first
        # chain's start field, scaled to 21 forces + 1 eval per chain of the bucket
        from oracle import hubbard_ref as ref
        total = 0.0
        for idx in ens.index:
            c = chains[idx[0]]
            fa = ref.HubbardFermiAction(c.kappa * c.beta / c.nt, 0.0, -1, ref.EXP)
            t0 = time.perf_counter()
            fa.force(c.phi0)
            tf = time.perf_counter() - t0
            t0 = time.perf_counter()
            fa.eval(c.phi0)
            te = time.perf_counter() - t0
            total += len(idx) * (21 * tf + te)
        cpu = {"value": n_chains / total, "unit": "chain-traj/s", "cores": os.cpu_count(), "kind": "port",
               "sample": "oracle (numpy complex128, OpenBLAS) force and eval once per (Nx, Nt) bucket, scaled to "
                         "21 forces + 1 eval per chain"}
    print(json.dumps({
        "metric": "chain-trajectories/s, 256 uneven chains (config 4)", "value": n_chains * args.steps / dt,
        "unit": "chain-traj/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 right chains + f64 suffix/top level/action", "data": "synthetic",
        "config": {"workload": "256 honeycomb tori, L in {3,4,6,8}, Nt in {16,32,64}, beta=Nt/8, U=2, mu=0, EXP, "
                               "particle-hole, sigmaKappa=-1, HMC length 1 with 20 steps",
                   "buckets": len(ens.buckets),
                   "parallelism": ("the buckets of one lattice share their force launches, one stream per lattice, "
                                   "one HIP graph" if not args.no_merge else "buckets on separate streams, one HIP graph")},
        "acceptance": float(pts.mean()) if pts.size else None,
        "median_dH": float(np.median(dh)) if dh.size else None,
        "launch_structure": launch,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK["f64"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK["f64"], "traffic": None,
                     "note": "all kernels together: the algorithmic flops of the chains' trajectories "
                             f"({work / 1e12:.3f} TFLOP per trajectory of every chain) per second, against the FP64 "
                             "MFMA peak (the FP32 right chains could run at twice that)"},
        "cpu_baseline": cpu}))


# the dominant kind's kernel as it appears in the rocprofv3 PMC summary of this workload
PMC_SUMMARY = "profiles/force_mixed_nx288_nt128_chains128_hbm_traffic.md"
PMC_CHAINS = 128
PMC_ROWS = {"gemm64_rc_row": ("gemm_rcd_kernel<double, 1, false, 96, 96, 16, 2, 2, 6>", "7077888")}


def _pmc_traffic(kind, chains):
    """HBM bytes per launch of the dominant kernel (calibrated FETCH + WRITE, bytes) from the
    committed rocprofv3 PMC summary of the 128-chain force, or None for other workloads
    (bench.py cannot run the counters itself: they need their own profiler passes)."""
    if chains != PMC_CHAINS or kind not in PMC_ROWS:
        return None
    name, grid = PMC_ROWS[kind]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), PMC_SUMMARY)
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip().strip("`") for c in line.split("|")]
                if len(cells) > 5 and cells[1] == name and cells[2] == grid:
                    return (float(cells[4]) + float(cells[5])) * 2 ** 20
    except (OSError, ValueError):
        return None
    return None


def propagator(args):
    """All-to-all propagator of the particles on the config-3 torus (Nx=288, Nt=128, beta=8,
    U=3, EXP) for a field N(0, U beta/Nt) from NumpyRNG(42): meas.AllToAll with iterative
    refinement (the exact solve to rounding) and without (solveM as the reference
    computes it).  One step = one propagator.  Algorithmic work per propagator: solveM
    against N = Nx Nt sources, Nt slice GEMMs of 8 Nx^2 N flops and Nt pairs of triangular
    solves of 8 Nx^2 N flops = 16 N^2 Nx real flops (refinement adds about twice that,
    not counted).  Accuracy: the largest entry of M G - 1 over the first 1024 sources
    (applyM).  CPU baseline: the oracle's solveM against 288 and 1152 of the sources, its
    fixed cost (partial products, LUs) and per-source cost extrapolated to all N."""
    import torch
    import isle_amd as isle
    from isle_amd import meas
    from isle_amd.evolver import EvolutionStage, NumpyRNG
    from oracle import hubbard_ref as ref
    kappa = isle.honeycomb(L).hopping()
    nx = kappa.shape[0]
    kt, ut = kappa * BETA / NT, U * BETA / NT
    n = nx * NT
    phi = NumpyRNG(42).normal(0, np.sqrt(ut), n) + 0j
    hfm = isle.HubbardFermiMatrixExp(kt, 0.0, -1)
    stage = EvolutionStage(phi, 0.0, 1)
    res = {}
    for refine in (8, 0):
        a2a = meas.AllToAll(hfm, isle.Species.PARTICLE, refine=refine)
        for w in range(args.warmup):
            a2a(stage, -1 - w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            g = a2a(stage, k)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        gm = g.permute(1, 0, 3, 2).reshape(n, n)
        cols = gm[:, :1024].T.contiguous()  # rows: M^-1 applied to sources 0..1023
        eye = torch.zeros_like(cols)
        eye[torch.arange(1024), torch.arange(1024)] = 1
        resid = (isle.applyM(hfm, phi, 0, cols) - eye).abs().max().item()
        res[refine] = (dt, resid)
        del g, gm, cols, a2a
    flops = 16.0 * n * n * nx
    dt, resid = res[8]
    line = {
        "metric": "all-to-all propagators/s at Nx=288 Nt=128 (M^-1, every spacetime source)",
        "value": 1.0 / dt, "unit": "propagators/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "honeycomb 12x12 torus Nx=288 Nt=128 beta=8 U=3 mu=0 EXP, particles, N = 36864 "
                               "sources in chunks", "refine": "until the residual stops falling (<= 8 steps)"},
        "max_abs_MG_minus_1_first_1024_sources": resid,
        "unrefined": {"ms_per_step": 1e3 * res[0][0], "max_abs_MG_minus_1_first_1024_sources": res[0][1]},
        "roofline": {"bound": "mfma", "achieved": flops / res[0][0] / 1e12, "peak": PEAK["f64"], "unit": "TFLOP/s",
                     "frac": flops / res[0][0] / 1e12 / PEAK["f64"], "traffic": None,
                     "kernel": "whole unrefined propagator (16 N^2 Nx flops)"},
    }
    if not args.no_cpu:
        # fixed cost (partial products, LUs) + per-source cost, from two sample sizes
        rh = ref.HFMExp(kt, 0.0, -1)
        times = {}
        for k in (288, 1152):
            sample = np.zeros((k, n))
            sample[np.arange(k), np.arange(k)] = 1
            t0 = time.perf_counter()
            ref.solveM(rh, phi, 0, sample)
            times[k] = time.perf_counter() - t0
        per = (times[1152] - times[288]) / (1152 - 288)
        t_cpu = max(times[288] - 288 * per, 0.0) + n * per
        line["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "propagators/s", "cores": os.cpu_count(),
                                "kind": "port",
                                "sample": f"oracle solveM (numpy/OpenBLAS, unrefined) against 288 and 1152 of the {n} "
                                          f"sources ({times[288]:.2f} s, {times[1152]:.2f} s); fixed + per-source "
                                          f"cost extrapolated to all sources: {t_cpu:.1f} s"}
    print(json.dumps(line))


def logdetq(args):
    """log det Q(phi) by the block LU of Q = M_p M_h^T (getQLU + logdetQ) on the config-3 torus
    (Nx=288, Nt=128, beta=8, U=3, EXP, sigmaKappa=-1) for a field N(0, U beta/Nt) from
    NumpyRNG(42).  One step = one log det Q.  Algorithmic work: the recursion's Nt block
    inversions and block products, 8 Nx^3 (7 Nt - 12) real flops (each complex Nx x Nx
    product or inverse counts 8 Nx^3).  Accuracy: against log det M_p + log det M_h from the
    DIRECT_SINGLE logdetM (det Q = det M_p det M_h).  CPU baseline: the oracle's QLU
    (numpy/OpenBLAS) on the same field."""
    import torch
    import isle_amd as isle
    from isle_amd.evolver import NumpyRNG
    from oracle import hubbard_ref as ref
    kappa = isle.honeycomb(L).hopping()
    nx = kappa.shape[0]
    kt, ut = kappa * BETA / NT, U * BETA / NT
    phi = NumpyRNG(42).normal(0, np.sqrt(ut), nx * NT) + 0j
    hfm = isle.HubbardFermiMatrixExp(kt, 0.0, -1)
    dphi = isle.as_field(phi)
    for _ in range(args.warmup):
        isle.logdetQ(hfm, dphi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got = isle.logdetQ(hfm, dphi)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    want = isle.logdetM(hfm, dphi, isle.Species.PARTICLE) + isle.logdetM(hfm, dphi, isle.Species.HOLE)
    d = complex(got) - complex(want)
    err = abs(complex(d.real, math.remainder(d.imag, 2 * math.pi))) / abs(want)
    flops = 8.0 * nx ** 3 * (7 * NT - 12)
    line = {
        "metric": "log det Q by the block LU of Q at Nx=288 Nt=128 (DIRECT_SQUARE action)",
        "value": 1.0 / dt, "unit": "logdets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "honeycomb 12x12 torus Nx=288 Nt=128 beta=8 U=3 mu=0 EXP sigmaKappa=-1"},
        "rel_err_vs_logdetM_p_plus_h": err,
        "roofline": {"bound": "mfma", "achieved": flops / dt / 1e12, "peak": PEAK["f64"], "unit": "TFLOP/s",
                     "frac": flops / dt / 1e12 / PEAK["f64"], "traffic": None,
                     "kernel": "whole log det Q (8 Nx^3 (7 Nt - 12) flops; a sequential recursion of Nt "
                               "block inversions)"},
    }
    if not args.no_cpu:
        rh = ref.HFMExp(kt, 0.0, -1)
        t0 = time.perf_counter()
        ref.logdet_q(ref.qlu(rh, phi))
        t_cpu = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "logdets/s", "cores": os.cpu_count(), "kind": "port",
                                "sample": f"oracle qlu + logdet_q (numpy/OpenBLAS) on the same field, once: "
                                          f"{t_cpu:.2f} s"}
    print(json.dumps(line))


def config5(args):
    """Nanotube workload: honeycomb cylinder of 12 cells around (periodic) and 16 along
    (open), Nx=384, Nt=256, beta=10, U=2.5, mu=0.5 (complex action), DIA hopping, spin basis,
    sigmaKappa=-1, HMC length 1 with 60 steps; chain c starts from N(0, U beta/Nt) drawn from
    NumpyRNG(7 + c).  The product of all 256 slices spans ~33 decades here, so eval and force
    use precision="stable" (FP64 groups of --group slices, default 32, combined by the pivoted
    block-cyclic elimination; groups of 16 double its sequential steps and run 2.2x slower).
    One step = one trajectory of every chain; 16 chains by default (the elimination's LU
    panels are one workgroup per matrix, so more chains per launch fill more of the GPU: 4
    chains gave 1.61, 16 gave 2.12, 32 gave 2.23 chain-trajectories/s in one build).  Checks: log det M of chain 0's start field against the oracle's stabilised
    restatement (CPU, groups of 16), and the force against FP64 chains over another grouping
    of the same elimination (32 slices, or 16 when the force uses 32)."""
    import torch
    import isle_amd as isle
    from isle_amd import action as A
    from isle_amd.evolver import BatchedConstStepLeapfrog, NumpyRNG
    from oracle import hubbard_ref as ref
    nt, beta, u, mu, nstep = 256, 10.0, 2.5, 0.5, 60
    chains = args.chains or 16
    kappa = isle.honeycomb(12, 16, periodic_j=False).hopping()
    nx = kappa.shape[0]
    kt, mt, ut = kappa * beta / nt, mu * beta / nt, u * beta / nt
    group = args.group or 32
    other = 16 if group == 32 else 32
    ferm = A.HubbardFermiAction(kt, mt, -1, A.HFAHopping.DIA, A.HFABasis.SPIN, precision="stable", group=group,
                                right_chains=args.right_chains)
    act = A.SumAction(A.HubbardGaugeAction(ut), ferm)
    rngs = [NumpyRNG(7 + c) for c in range(chains)]
    phi0 = np.stack([r.normal(0, np.sqrt(ut), nx * nt) + 0j for r in rngs])
    dev = torch.device("cuda")
    phis = torch.as_tensor(phi0, device=dev)
    # accuracy on chain 0's start field
    ld = ferm.logdets(phis[:1])[0]
    aux = -1j * phi0[0]
    hfm = ref.HFMDia(kt, mt, -1)
    t_cpu = time.perf_counter()
    want = [ref.logdetM_stable(hfm, aux, s, 16) for s in (0, 1)]
    t_cpu = time.perf_counter() - t_cpu
    ld_err = max(abs(complex(d.real, np.remainder(d.imag + np.pi, 2 * np.pi) - np.pi)) / abs(w)
                 for d, w in ((ld[s] - want[s], want[s]) for s in (0, 1)))
    plain_err = abs(ref.logdetM(hfm, aux, 0) - want[0]) / abs(want[0])
    f_used = ferm.force(phis[0]).cpu().numpy()
    f_other = A.HubbardFermiAction(kt, mt, -1, A.HFAHopping.DIA, A.HFABasis.SPIN, precision="stable",
                                   group=other).force(phis[0]).cpu().numpy()
    f_err = float(np.linalg.norm(f_used - f_other) / np.linalg.norm(f_other))
    # the stable force of one field, timed (both species)
    ferm.force(phis[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        ferm.force(phis[0])
    torch.cuda.synchronize()
    force_ms = 1e3 * (time.perf_counter() - t0) / 5
    acts = np.array([act.eval(phis[b]) for b in range(chains)])
    lf = BatchedConstStepLeapfrog(act, LENGTH, nstep, rngs, graph=not args.no_graph)
    for _ in range(args.warmup):
        phis, acts, _ = lf.evolve(phis, acts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        phis, acts, _ = lf.evolve(phis, acts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pts = np.array(lf.trajPoints[args.warmup:], dtype=float)
    dh = np.array(lf.deltaH[args.warmup:], dtype=float)
    print(json.dumps({
        "metric": "chain-trajectories/s, Nx=384 Nt=256 DIA spin-basis cylinder (config 5)",
        "value": chains * args.steps / dt, "unit": "chain-traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.right_chains == "fp64" else "f32 right chains + f64",
        "data": "synthetic",
        "config": {"workload": "honeycomb cylinder 12 (periodic) x 16 (open) cells, Nx=384, Nt=256, beta=10, U=2.5, "
                               "mu=0.5, DIA, spin basis, sigmaKappa=-1, HMC length 1 with 60 steps",
                   "chains": chains, "precision": "stable", "right_chains": args.right_chains,
                   "group": ferm._group(nt), "parallelism": "chains batched in every launch"},
        "force_ms_one_field": force_ms,
        "logdet_rel_err_vs_oracle": ld_err, "oracle_logdet_cpu_s": t_cpu,
        "reference_product_logdet_rel_err": float(plain_err),
        "force_rel_diff_vs_other_grouping": f_err,
        "force_rel_diff_note": f"the force used (groups of {ferm._group(nt)}, right chains as configured) against "
                               f"FP64 chains over groups of {other}",
        "acceptance": float(pts.mean()) if pts.size else None,
        "mean_dH": float(dh.mean()) if dh.size else None,
        "mean_exp_minus_dH": float(np.exp(-dh).mean()) if dh.size else None,
        "mean_action_phase_cos": float(np.mean(np.cos(np.imag(acts))))}))


def force_only(args):
    """Short workload for hardware-counter runs: the batched mixed force at Nx=288, Nt=128
    for all chains, then a device-to-device copy of 1 GiB (known HBM traffic: 1 GiB read
    and 1 GiB written) to scale the byte counters."""
    import torch
    import isle_amd as isle
    from isle_amd import _lib
    from isle_amd import action as A
    dev = _lib.device()
    lat = isle.honeycomb(L, nt=NT)
    nx = lat.nx()
    kt = lat.hopping() * BETA / NT
    ut = U * BETA / NT
    ferm = A.HubbardFermiAction(kt, 0.0, -1, A.HFAHopping.EXP, precision="mixed", group=args.group or 32)
    phis = torch.as_tensor(np.random.RandomState(5).normal(0, np.sqrt(ut), (args.chains or 128, nx * NT)) + 0j,
                           device=dev)
    for _ in range(args.warmup):
        ferm.force_batch(phis)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ferm.force_batch(phis)
    torch.cuda.synchronize()
    t_force = (time.perf_counter() - t0) / args.steps
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)  # 1 GiB
    torch.cuda.synchronize()
    dst = src.clone()
    torch.cuda.synchronize()
    print(json.dumps({"force_only": True, "chains": args.chains or 128, "force_ms": 1e3 * t_force,
                      "copy_bytes": int(dst.numel() * 8)}))


def cpu_baseline(kappa_tilde, ut, nx):
    """The oracle (numpy/OpenBLAS complex128 restatement of the reference algorithm) on
    the host, at the full size: with every host core, 2 forces + 2 evals after a warm-up
    force; with one BLAS thread (threadpoolctl), 1 force + 1 eval.  Each is scaled to
    41 forces + 1 eval per trajectory."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from threadpoolctl import threadpool_limits
    from oracle import hubbard_ref as ref
    act = ref.SumAction(ref.HubbardGaugeAction(ut), ref.HubbardFermiAction(kappa_tilde, 0.0, -1, ref.EXP))
    phi = np.random.RandomState(7).normal(0, np.sqrt(ut), nx * NT) + 0j

    def timed(reps):
        t0 = time.perf_counter()
        for _ in range(reps):
            act.force(phi)
        tf = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            act.eval(phi)
        return tf, (time.perf_counter() - t0) / reps

    act.force(phi)
    tf, te = timed(2)
    with threadpool_limits(limits=1):
        tf1, te1 = timed(1)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    traj, traj1 = (NSTEP + 1) * tf + te, (NSTEP + 1) * tf1 + te1
    return {"value": 1.0 / traj, "unit": "traj/s", "cores": cores, "kind": "port",
            "sample": f"oracle (numpy complex128, OpenBLAS) force {tf:.3f} s and eval {te:.3f} s at Nx=288 "
                      f"Nt=128, each timed twice; trajectory = {NSTEP + 1} forces + 1 eval",
            "force_s": tf, "eval_s": te,
            "one_thread": {"value": 1.0 / traj1, "unit": "traj/s", "cores": 1, "force_s": tf1, "eval_s": te1,
                           "sample": "the same with one BLAS thread, each timed once"}}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU (the chains of the batch share it); independent chains "
              "would shard across GPUs with no communication, which this benchmark does not launch",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    if args.force_only:
        force_only(args)
        return
    if args.config4:
        config4(args)
        return
    if args.config5:
        config5(args)
        return
    if args.propagator:
        propagator(args)
        return
    if args.logdetq:
        logdetq(args)
        return
    import torch
    import isle_amd as isle
    from isle_amd import _lib
    from isle_amd import action as A
    from isle_amd.evolver import BatchedConstStepLeapfrog, NumpyRNG

    dev = _lib.device()
    lat = isle.honeycomb(L, nt=NT)
    nx = lat.nx()
    kt = lat.hopping() * BETA / NT
    ut = U * BETA / NT
    B = args.chains or 128

    def make(prec):
        # the mixed form's Metropolis action: FP64, stabilising block-cyclic LU every 8 slices
        extra = {"logdet": args.logdet, "logdet_group": 8} if prec == "mixed" else {}
        ferm = A.HubbardFermiAction(kt, 0.0, -1, A.HFAHopping.EXP, precision=prec, group=args.group or 32, **extra)
        return A.SumAction(A.HubbardGaugeAction(ut), ferm), ferm

    def clone_rngs(rngs):
        out = []
        for r in rngs:
            c = NumpyRNG(0)
            c.set_state(r.get_state())
            out.append(c)
        return out

    def run(ev, phis, acts, n):
        for _ in range(n):
            phis, acts, _ = ev.evolve(phis, acts)
        return phis, acts

    def stats(ev, skip):
        dh = np.array(ev.deltaH[skip:], dtype=float).ravel()
        acc = np.array(ev.trajPoints[skip:], dtype=float).ravel()
        if not dh.size:
            return {}
        w = np.exp(-np.clip(dh, -700.0, None))
        p = acc.mean()
        return {"trajectories": int(dh.size), "acceptance": float(p),
                "acceptance_err": float(np.sqrt(max(p * (1 - p), 1e-300) / dh.size)),
                "mean_dH": float(dh.mean()), "median_dH": float(np.median(dh)),
                "mean_exp_minus_dH": float(w.mean()), "mean_exp_minus_dH_err": float(w.std(ddof=1) / np.sqrt(w.size))
                if w.size > 1 else None}

    # ---------------- start fields N(0, Utilde) per chain, then thermalisation: short
    #                  trajectories with small steps carry the fields away from the
    #                  random start (where length-1, 40-step trajectories are rejected)
    act_m, ferm_m = make("mixed")
    rngs = [NumpyRNG(1000 + b) for b in range(B)]
    phis = torch.as_tensor(np.stack([r.normal(0, np.sqrt(ut), nx * NT) + 0j for r in rngs]), device=dev)
    acts = np.array([act_m.eval(phis[b]) for b in range(B)])
    split = [make("mixed")[0] for _ in range(args.streams)] if args.streams > 1 else None
    t_th = time.perf_counter()
    if args.therm > 0:
        ev_th = BatchedConstStepLeapfrog(act_m, THERM_LENGTH, THERM_NSTEP, rngs, graph=False)
        phis, acts = run(ev_th, phis, acts, args.therm)
        therm_stats = stats(ev_th, 0)
    else:
        therm_stats = {}
    torch.cuda.synchronize()
    t_th = time.perf_counter() - t_th
    start_phis, start_acts, start_rngs = phis.clone(), acts.copy(), clone_rngs(rngs)

    progress("thermalisation done")
    # ---------------- force error of the mixed form against plain FP64 on the thermalised fields
    act_p, ferm_p = make("fp64")
    f_m = ferm_m.force_batch(phis).cpu().numpy()
    f_p = ferm_p.force_batch(phis).cpu().numpy()
    errs = np.linalg.norm(f_m - f_p, axis=1) / np.linalg.norm(f_p, axis=1)
    force_err = float(errs.max())
    force_err_max = float((np.abs(f_m - f_p).max(axis=1) / np.abs(f_p).max(axis=1)).max())
    del f_m, f_p

    progress("force error done")
    # ---------------- batched mixed precision: the timed trajectories
    ev = BatchedConstStepLeapfrog(act_m, LENGTH, NSTEP, rngs, graph=not args.no_graph, split=split)
    phis, acts = run(ev, phis, acts, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phis, acts = run(ev, phis, acts, args.steps)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    traj_per_s = B * args.steps / t_mixed
    mixed_stats = stats(ev, 0)

    progress("timed mixed trajectories done")
    # ---------------- batched plain FP64 from the same thermalised fields and generator states
    plain_batched, plain_stats = None, {}
    if not args.no_plain:
        ev_pb = BatchedConstStepLeapfrog(act_p, LENGTH, NSTEP, start_rngs, graph=not args.no_graph)
        pb_phis, pb_acts = run(ev_pb, start_phis.clone(), start_acts.copy(), 1)  # plan build, tuning, capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(ev_pb, pb_phis, pb_acts, args.plain_steps)
        torch.cuda.synchronize()
        plain_batched = B * args.plain_steps / (time.perf_counter() - t0)
        plain_stats = stats(ev_pb, 0)
        del ev_pb, pb_phis

    progress("batched plain FP64 trajectories done")
    # ---------------- reversibility: forward then reversed trajectory on thermalised fields
    #                  (the first REV chains), mixed and plain FP64 forces, the same momenta
    rev_n = min(REV_CHAINS, B)
    pi0 = torch.as_tensor(np.stack([NumpyRNG(77 + c).normal(0, 1, nx * NT) for c in range(rev_n)]) + 0j, device=dev)
    x0 = start_phis[:rev_n]
    rev = {}
    actions = (("mixed", make("mixed")[0]), ("fp64", make("fp64")[0]))
    for label, length, nstep in (("", LENGTH, NSTEP), ("_small_step", THERM_LENGTH, THERM_NSTEP)):
        for name, action in actions:
            p1, q1, a1 = isle.leapfrog_batch(x0, pi0, action, length, nstep, 1.0)
            p2, _, _ = isle.leapfrog_batch(p1, q1, action, length, nstep, -1.0, evaluate=False)
            err = (torch.linalg.norm(p2 - x0, dim=1) / torch.linalg.norm(x0, dim=1)).cpu().numpy()
            kin0 = (pi0.real ** 2).sum(dim=1).cpu().numpy() / 2
            kin1 = (q1.real ** 2 + q1.imag ** 2).sum(dim=1).cpu().numpy() / 2
            dh = np.real(np.asarray(a1) + kin1 - start_acts[:rev_n] - kin0)
            rev[name + label] = {"err": err, "dH": dh}
    regular = np.abs(rev["fp64"]["dH"]) < 10.0  # trajectories that passed no near-zero of det M
    reversibility = {
        "chains": rev_n, "regular_chains": int(regular.sum()),
        "mixed_median": float(np.median(rev["mixed"]["err"])),
        "fp64_median": float(np.median(rev["fp64"]["err"])),
        "mixed_max_regular": float(rev["mixed"]["err"][regular].max()) if regular.any() else None,
        "fp64_max_regular": float(rev["fp64"]["err"][regular].max()) if regular.any() else None,
        "dH_fp64": [round(float(x), 5) for x in rev["fp64"]["dH"]],
        "small_step": {"length": THERM_LENGTH, "nstep": THERM_NSTEP,
                       "mixed_max": float(rev["mixed_small_step"]["err"].max()),
                       "fp64_max": float(rev["fp64_small_step"]["err"].max()),
                       "dH_fp64_max_abs": float(np.abs(rev["fp64_small_step"]["dH"]).max()),
                       "dH_mixed_minus_fp64_max_abs": float(np.abs(rev["mixed_small_step"]["dH"]
                                                                   - rev["fp64_small_step"]["dH"]).max())},
    }

    progress("reversibility done")
    # ---------------- kernel roofline (mixed force of the whole chain batch)
    nk = len(KIND_NAMES)
    prof = np.zeros(4 * nk)
    g = ferm_m._group(NT)
    _lib.check(_lib.load().isle_profile_kernels(ferm_m.hfm.handle, _lib.ptr(phis), 1, NT, B, g, 2, 0, 3,
                                                _lib.dptr(prof), nk))
    prof = prof.reshape(nk, 4)
    dom = int(np.argmax(prof[:, 1]))
    launches, ms_op, flops_op = prof[dom, 0], prof[dom, 1], prof[dom, 2]
    avg_ms = ms_op / launches
    achieved = flops_op / launches / (avg_ms * 1e-3) / 1e12
    prec = "f32" if "32" in KIND_NAMES[dom] else "f64"
    traffic = _pmc_traffic(KIND_NAMES[dom], B)
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK[prec], "unit": "TFLOP/s",
                "frac": achieved / PEAK[prec], "traffic": traffic,
                "traffic_source": PMC_SUMMARY if traffic is not None else None, "kernel": KIND_NAMES[dom],
                "avg_launch_ms": avg_ms, "flops_per_launch": flops_op / launches,
                "force_ms": float(prof[:, 1].sum()),
                "breakdown_ms": {KIND_NAMES[k]: round(float(prof[k, 1]), 4) for k in range(nk) if prof[k, 0] > 0}}

    # the FP32 right-chain step, the chain kernel furthest below its peak, beside the dominant one
    k32 = KIND_NAMES.index("gemm32_rck_diag")
    if prof[k32, 0] > 0:
        a32 = prof[k32, 2] / prof[k32, 1] * 1e3 / 1e12
        roofline["fp32_right_chain"] = {"kernel": "gemm32_rck_diag", "achieved": a32, "peak": PEAK["f32"],
                                        "unit": "TFLOP/s", "frac": a32 / PEAK["f32"],
                                        "avg_launch_ms": float(prof[k32, 1] / prof[k32, 0]),
                                        "note": "epilogue-bound: the force diagonal reads the FP64 suffix "
                                                "entries (16 B per product element) after the K loop"}
    progress("roofline done")
    # ---------------- plain FP64, one chain, unbatched (from chain 0's thermalised field)
    plain = None
    if not args.no_plain:
        rng = clone_rngs(start_rngs[:1])
        phi1 = start_phis[0:1].clone()
        a1 = start_acts[0:1].copy()
        ev_p = BatchedConstStepLeapfrog(act_p, LENGTH, NSTEP, rng, graph=not args.no_graph)
        phi1, a1, _ = ev_p.evolve(phi1, a1)  # warm-up (plan build, tuning, graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            phi1, a1, _ = ev_p.evolve(phi1, a1)
        torch.cuda.synchronize()
        plain = args.plain_steps / (time.perf_counter() - t0)

    cpu = None if args.no_cpu else cpu_baseline(kt, ut, nx)
    out = {
        "metric": METRIC, "value": traj_per_s, "unit": "traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_mixed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 force (MFMA) + f64 action", "data": "synthetic",
        "config": {"workload": "honeycomb 12x12 torus Nx=288 Nt=128 beta=8 U=3 mu=0 EXP particle-hole "
                               "sigmaKappa=-1, HMC length 1 with 40 leapfrog steps",
                   "nx": nx, "nt": NT, "chains": B, "group": g, "streams": args.streams if not args.no_graph else 1,
                   "metropolis_logdet": ("FP64, stabilising block-cyclic LU every 8 slices" if args.logdet == "stable"
                                         else f"FP64 products of {g} slices multiplied together"),
                   "parallelism": "independent chains, 1 GPU"},
        "plain_fp64_traj_per_s": plain,
        "speedup_vs_plain_fp64": (traj_per_s / plain) if plain else None,
        "plain_fp64_batched_traj_per_s": plain_batched,
        "speedup_vs_plain_fp64_batched": (traj_per_s / plain_batched) if plain_batched else None,
        "plain_note": f"plain FP64 = the reference's algorithm in FP64 on the same kernels; single chain and "
                      f"{B} batched chains, {args.plain_steps} timed trajectories each after one warm-up",
        "force_rel_err": force_err, "force_rel_err_max": force_err_max,
        "force_rel_err_note": "max over the chains' thermalised fields of |F_mixed - F_fp64|_2 / |F_fp64|_2 "
                              "(and of the max-norm ratio)",
        "force_evals_per_s": traj_per_s * (NSTEP + 1),
        "thermalisation": {"trajectories": args.therm, "length": THERM_LENGTH, "nstep": THERM_NSTEP,
                           "seconds": t_th, **{k: v for k, v in therm_stats.items() if k in ("acceptance",
                                                                                             "median_dH")}},
        "hmc_mixed": mixed_stats, "hmc_fp64_batched": plain_stats,
        "acceptance": mixed_stats.get("acceptance"),
        "mean_dH": mixed_stats.get("mean_dH"),
        "median_dH": mixed_stats.get("median_dH"),
        "mean_exp_minus_dH": mixed_stats.get("mean_exp_minus_dH"),
        "reversibility": reversibility,
        "reversibility_err": reversibility["mixed_max_regular"],
        "reversibility_err_plain_fp64": reversibility["fp64_max_regular"],
        "reversibility_note": "|phi(fwd then reversed) - phi0| / |phi0| per chain on thermalised fields, same momenta "
                              "for both forces; at length 1 / 40 steps: max over the chains whose FP64 trajectory has "
                              "|dH| < 10 (the others pass near a zero of det M, where the trajectory itself is "
                              "chaotic), medians over all; small_step: the same at length 0.25 / 100 steps, where "
                              "every trajectory is regular",
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
