"""Many independent Markov chains of uneven lattice size and time extent on one GPU.

The reference runs one chain per process (``src/isle/drivers/hmc.py``) with
``ConstStepLeapfrog`` (``src/isle/evolver/leapfrog.py:38-66``) and ``BinarySelector``
(``src/isle/evolver/selector.py:20-30``).  Here chains are grouped into buckets of equal
(Nx, Nt); every bucket is a homogeneous batch with its own plan (so its launch shapes are
tuned per (Nx, Nt) by the plan's autotuner), all buckets run their leapfrog integration
on separate streams, and one trajectory of the whole ensemble is captured once in a HIP
graph (fork/join over the bucket streams) and replayed.  Momenta and the Metropolis
decision of every chain come from that chain's own NumpyRNG exactly as a lone
ConstStepLeapfrog would draw them, so each chain reproduces a single-chain run with the
same seed up to the arithmetic of the batched kernels.
"""
from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from .action import HFAHopping, HubbardFermiAction, HubbardGaugeAction, SumAction
from .evolver import BinarySelector
from .integrator import eval_batch, leapfrog_batch


class Chain:
    """One Markov chain: hopping matrix (nx x nx), nt, beta, U, mu, sigmaKappa, RNG."""

    def __init__(self, kappa, nt, beta, U, mu, sigma_kappa, rng, phi=None):
        self.kappa = np.asarray(kappa, dtype=np.float64)
        self.nx = self.kappa.shape[0]
        self.nt = int(nt)
        self.beta, self.U, self.mu, self.sigma_kappa = float(beta), float(U), float(mu), int(sigma_kappa)
        self.rng = rng
        self.utilde = self.U * self.beta / self.nt
        if phi is None:
            phi = rng.normal(0, np.sqrt(self.utilde), self.nx * self.nt) + 0j
        self.phi0 = np.asarray(phi, dtype=np.complex128)

    def key(self):
        return (self.nx, self.nt, self.beta, self.U, self.mu, self.sigma_kappa, self.kappa.tobytes())


class _Bucket:
    def __init__(self, chains, hopping, precision, group):
        c0 = chains[0]
        self.chains = chains
        kt = c0.kappa * c0.beta / c0.nt
        self.fermion = HubbardFermiAction(kt, c0.mu * c0.beta / c0.nt, c0.sigma_kappa, hopping,
                                          precision=precision, group=group)
        self.action = SumAction(HubbardGaugeAction(c0.utilde), self.fermion)
        self.n = c0.nx * c0.nt
        dev = _lib.device()
        self.phi = torch.as_tensor(np.stack([c.phi0 for c in chains]), device=dev)
        self.acts = np.asarray(eval_batch(self.action, self.phi), dtype=complex)
        self.selectors = [BinarySelector(c.rng) for c in chains]


class EnsembleLeapfrog:
    """Constant-step leapfrog HMC for an ensemble of chains of uneven (Nx, Nt).

    ``chains``: list of Chain.  ``evolve()`` advances every chain by one trajectory and
    returns the accept (1) / reject (0) decision of each chain in input order; ``deltaH``
    and ``trajPoints`` keep the history.  With ``graph=True`` the integration of all
    buckets is one replayed HIP graph."""

    def __init__(self, chains, length, nstep, hopping=HFAHopping.EXP, precision="mixed", group=32, graph=True):
        self.chains = list(chains)
        self.length, self.nstep = float(length), int(nstep)
        groups = OrderedDict()
        for i, c in enumerate(self.chains):
            groups.setdefault(c.key(), []).append(i)
        self.index = list(groups.values())  # chain indices of each bucket
        self.buckets = [_Bucket([self.chains[i] for i in idx], hopping, precision, group) for idx in self.index]
        self.graph = bool(graph)
        self._graph = None
        self._pool = ThreadPoolExecutor(max_workers=16)
        self.trajPoints, self.deltaH = [], []

    def phis(self):
        """Current field of every chain (input order), as host arrays."""
        out = [None] * len(self.chains)
        for idx, b in zip(self.index, self.buckets):
            host = b.phi.cpu().numpy()
            for j, i in enumerate(idx):
                out[i] = host[j]
        return out

    def _capture(self):
        main = torch.cuda.current_stream()
        self._streams = [torch.cuda.Stream() for _ in self.buckets]
        self._in = [(torch.zeros_like(b.phi), torch.zeros_like(b.phi)) for b in self.buckets]
        # warm-up on the bucket streams: builds and tunes every plan outside the capture
        for s, b, (p, q) in zip(self._streams, self.buckets, self._in):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                leapfrog_batch(p, q, b.action, self.length, self.nstep, evaluate=False)
            main.wait_stream(s)
        torch.cuda.synchronize()
        self._graph = torch.cuda.CUDAGraph()
        self._out = []
        with torch.cuda.graph(self._graph):
            cap = torch.cuda.current_stream()
            for s, b, (p, q) in zip(self._streams, self.buckets, self._in):
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    p1, q1, _ = leapfrog_batch(p, q, b.action, self.length, self.nstep, evaluate=False)
                self._out.append((p1, q1))
            for s in self._streams:
                cap.wait_stream(s)

    def time_buckets(self, reps=3):
        """Milliseconds of one trajectory's integration (the replayed work, no host part):
        ``(whole, alone)`` with ``whole`` the ensemble graph and ``alone[i]`` bucket i's
        integration captured in a graph of its own and replayed with nothing beside it, on
        the inputs of the last trajectory.  The longest ``alone`` is the ensemble's critical
        path; ``sum(alone) / whole`` is how much the buckets' streams overlap."""
        if self._graph is None:
            self._capture()

        def timed(replay):
            replay()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                replay()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps
        whole = timed(self._graph.replay)
        alone = []
        for b, (p, q) in zip(self.buckets, self._in):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                leapfrog_batch(p, q, b.action, self.length, self.nstep, evaluate=False)
            alone.append(timed(g.replay))
            del g
        return whole, alone

    def _momenta(self):
        """Every chain's momenta rng.normal(0, 1, n) (real parts), drawn on a thread pool: the
        chains' generators are independent and numpy releases the GIL while drawing."""
        real = [np.empty((len(b.chains), b.n)) for b in self.buckets]
        jobs = [(bi, j) for bi, b in enumerate(self.buckets) for j in range(len(b.chains))]

        def draw(job):
            bi, j = job
            real[bi][j] = self.buckets[bi].chains[j].rng.normal(0, 1, self.buckets[bi].n)
        list(self._pool.map(draw, jobs))
        return real

    def evolve(self):
        pis_host = self._momenta()
        if self.graph:
            if self._graph is None:
                self._capture()
            for b, (p, q), ph in zip(self.buckets, self._in, pis_host):
                p.copy_(b.phi)
                q.zero_()
                q.real.copy_(torch.from_numpy(ph))
            self._graph.replay()
            results = [(p1.clone(), q1.clone()) for p1, q1 in self._out]
        else:
            results = []
            for b, ph in zip(self.buckets, pis_host):
                q = torch.zeros_like(b.phi)
                q.real.copy_(torch.from_numpy(ph))
                p1, q1, _ = leapfrog_batch(b.phi, q, b.action, self.length, self.nstep, evaluate=False)
                results.append((p1, q1))
        points = [None] * len(self.chains)
        dhs = [None] * len(self.chains)
        for idx, b, (p1, q1), ph in zip(self.index, self.buckets, results, pis_host):
            acts1 = np.asarray(eval_batch(b.action, p1), dtype=complex)
            kin0 = np.einsum("ij,ij->i", ph, ph)
            kin1 = (q1.real.square() + q1.imag.square()).sum(dim=1).cpu().numpy()
            for j, i in enumerate(idx):
                e0 = b.acts[j] + kin0[j] / 2
                e1 = acts1[j] + kin1[j] / 2
                dhs[i] = float(np.real(e1 - e0))
                tp = b.selectors[j].selectTrajPoint(e0, e1)
                points[i] = tp
                if tp == 1:
                    b.phi[j] = p1[j]
                    b.acts[j] = acts1[j]
        self.trajPoints.append(points)
        self.deltaH.append(dhs)
        return points
