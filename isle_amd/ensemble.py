"""Many independent Markov chains of uneven lattice size and time extent on one GPU.

The reference runs one chain per process (``src/isle/drivers/hmc.py``) with
``ConstStepLeapfrog`` (``src/isle/evolver/leapfrog.py:38-66``) and ``BinarySelector``
(``src/isle/evolver/selector.py:20-30``).  Here chains are grouped into buckets of equal
(Nx, Nt) and parameters; every bucket is a homogeneous batch.  Buckets on one lattice (the
same hopping matrix and chemical potential per time slice, so the same slice matrices) share
one fermion action and one plan, and with ``merge=True`` (the default) they integrate
together: every kick evaluates the fermion force of all of them in shared launches (each
chain step one launch over the descriptors of all their chains), with each bucket's gauge
force added in its own kick.  Each such group runs on its own stream, and one trajectory of
the whole ensemble is captured once in a HIP graph (fork/join over the streams) and
replayed.  Launch shapes come from the plans' autotuner per (Nx, Nt) and per merged step.
Momenta and the Metropolis decision of every chain come from that chain's own NumpyRNG
exactly as a lone ConstStepLeapfrog would draw them, so each chain reproduces a single-chain
run with the same seed up to the arithmetic of the batched kernels.  A lattice that carries
more than half of the ensemble's work is split over two streams (config 4: Nx=128, 85% of
the flops).
"""
from collections import OrderedDict
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from .action import HFAHopping, HubbardFermiAction, HubbardGaugeAction, SumAction
from .evolver import BinarySelector
from .integrator import eval_batch, leapfrog_batch, leapfrog_multi


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

    def lattice_key(self):
        """Chains with equal keys have the same slice matrices (kappa beta/nt, mu beta/nt)."""
        kt = self.kappa * self.beta / self.nt
        return (kt.shape, kt.tobytes(), self.mu * self.beta / self.nt, self.sigma_kappa)


class _Bucket:
    def __init__(self, chains, fermion):
        c0 = chains[0]
        self.chains = chains
        self.fermion = fermion
        self.utilde = c0.utilde
        self.action = SumAction(HubbardGaugeAction(c0.utilde), fermion)
        self.n = c0.nx * c0.nt
        dev = _lib.device()
        self.phi = torch.as_tensor(np.stack([c.phi0 for c in chains]), device=dev)
        self.acts = np.asarray(eval_batch(self.action, self.phi), dtype=complex)
        self.selectors = [BinarySelector(c.rng) for c in chains]

    def label(self):
        c0 = self.chains[0]
        return f"Nx={c0.nx} Nt={c0.nt} x{len(self.chains)}"


class EnsembleLeapfrog:
    """Constant-step leapfrog HMC for an ensemble of chains of uneven (Nx, Nt).

    ``chains``: list of Chain.  ``evolve()`` advances every chain by one trajectory and
    returns the accept (1) / reject (0) decision of each chain in input order; ``deltaH``
    and ``trajPoints`` keep the history.  With ``graph=True`` the integration of all
    buckets is one replayed HIP graph; with ``merge=True`` the buckets of one lattice share
    their force launches (``groups`` lists the buckets integrated together)."""

    def __init__(self, chains, length, nstep, hopping=HFAHopping.EXP, precision="mixed", group=32, graph=True,
                 merge=True):
        self.chains = list(chains)
        self.length, self.nstep = float(length), int(nstep)
        keyed = OrderedDict()
        for i, c in enumerate(self.chains):
            keyed.setdefault(c.key(), []).append(i)
        self.index = list(keyed.values())  # chain indices of each bucket
        fermions = {}
        self.buckets = []
        self.groups = []  # bucket indices integrated together (one stream each)
        group_of = {}
        for idx in self.index:
            c0 = self.chains[idx[0]]
            lk = c0.lattice_key() if merge else len(self.buckets)
            if lk not in fermions:
                fermions[lk] = HubbardFermiAction(c0.kappa * c0.beta / c0.nt, c0.mu * c0.beta / c0.nt,
                                                  c0.sigma_kappa, hopping, precision=precision, group=group)
            bi = len(self.buckets)
            self.buckets.append(_Bucket([self.chains[i] for i in idx], fermions[lk]))
            # a group holds buckets of one lattice with distinct time extents
            for gi in group_of.get(lk, []):
                if all(self.buckets[b].chains[0].nt != c0.nt for b in self.groups[gi]):
                    self.groups[gi].append(bi)
                    break
            else:
                group_of.setdefault(lk, []).append(len(self.groups))
                self.groups.append([bi])
        if merge:
            self.groups = self._balance(self.groups)
        self.merge = bool(merge)
        self.graph = bool(graph)
        self._graph = None
        self._pool = ThreadPoolExecutor(max_workers=16)
        self.trajPoints, self.deltaH = [], []

    def _cost(self, b):
        """Work of one trajectory of bucket b, up to a constant: chains x Nt x Nx^3."""
        c0 = self.buckets[b].chains[0]
        return len(self.buckets[b].chains) * c0.nt * float(c0.nx) ** 3

    def _balance(self, groups):
        """A group with more than half of the ensemble's work and at least two buckets is
        the critical path of the graph: its buckets go to two streams, largest first onto the
        lighter one (each stream still shares its launches)."""
        total = sum(self._cost(b) for g in groups for b in g)
        out = []
        for g in groups:
            if len(g) >= 2 and sum(self._cost(b) for b in g) > 0.5 * total:
                parts, loads = [[], []], [0.0, 0.0]
                for b in sorted(g, key=self._cost, reverse=True):
                    i = 0 if loads[0] <= loads[1] else 1
                    parts[i].append(b)
                    loads[i] += self._cost(b)
                out.extend(sorted(p) for p in parts if p)
            else:
                out.append(g)
        return out

    def phis(self):
        """Current field of every chain (input order), as host arrays."""
        out = [None] * len(self.chains)
        for idx, b in zip(self.index, self.buckets):
            host = b.phi.cpu().numpy()
            for j, i in enumerate(idx):
                out[i] = host[j]
        return out

    def _integrate(self, gi, phis, pis):
        """One trajectory's integration of group gi from the given fields and momenta (one
        per bucket of the group); returns [(phi', pi')] in the group's bucket order."""
        members = [self.buckets[b] for b in self.groups[gi]]
        if not self.merge:
            b = members[0]
            p1, q1, _ = leapfrog_batch(phis[0], pis[0], b.action, self.length, self.nstep, evaluate=False)
            return [(p1, q1)]
        p1, q1 = leapfrog_multi(phis, pis, members[0].fermion, [b.utilde for b in members], self.length, self.nstep)
        return list(zip(p1, q1))

    def _capture(self):
        main = torch.cuda.current_stream()
        self._streams = [torch.cuda.Stream() for _ in self.groups]
        self._in = [(torch.zeros_like(b.phi), torch.zeros_like(b.phi)) for b in self.buckets]
        ins = [([self._in[b][0] for b in g], [self._in[b][1] for b in g]) for g in self.groups]
        # warm-up on the group streams: builds and tunes every plan outside the capture
        for gi, s in enumerate(self._streams):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                self._integrate(gi, *ins[gi])
            main.wait_stream(s)
        torch.cuda.synchronize()
        self._graph = torch.cuda.CUDAGraph()
        self._out = [None] * len(self.buckets)
        with torch.cuda.graph(self._graph):
            cap = torch.cuda.current_stream()
            for gi, s in enumerate(self._streams):
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    for b, res in zip(self.groups[gi], self._integrate(gi, *ins[gi])):
                        self._out[b] = res
            for s in self._streams:
                cap.wait_stream(s)

    def time_groups(self, reps=3):
        """Milliseconds of one trajectory's integration (the replayed work, no host part):
        ``(whole, alone)`` with ``whole`` the ensemble graph and ``alone`` a list of
        (label, ms): each group's integration captured in a graph of its own and replayed
        with nothing beside it, on the inputs of the last trajectory.  The longest is the
        ensemble's critical path; ``sum(alone) / whole`` is how much the streams overlap."""
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
        for gi, members in enumerate(self.groups):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._integrate(gi, [self._in[b][0] for b in members], [self._in[b][1] for b in members])
            label = ", ".join(self.buckets[b].label() for b in members)
            alone.append((label, timed(g.replay)))
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
            results = [None] * len(self.buckets)
            for gi, members in enumerate(self.groups):
                qs = []
                for b in members:
                    q = torch.zeros_like(self.buckets[b].phi)
                    q.real.copy_(torch.from_numpy(pis_host[b]))
                    qs.append(q)
                for b, res in zip(members, self._integrate(gi, [self.buckets[b].phi for b in members], qs)):
                    results[b] = res
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
