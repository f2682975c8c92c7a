"""The product's isBipartite (isle_amd.lattice) on the reference's truth table
(tests/test_lattice.py:13-77 of the reference: built-in lattices and the ill-labelled
(3,3) tube) and against the oracle's restatement on random graphs, including
disconnected ones, odd cycles and graphs whose last site decides (the reference's walk
does not continue from the last site)."""
import numpy as np

from isle_amd.lattice import Lattice, honeycomb, isBipartite
from oracle import hubbard_ref as ref
from tests.test_oracle_lattice import ILL_ADJ, TRUTH


def test_truth_table(lattice):
    for name, expected in TRUTH:
        h = lattice(name)
        assert isBipartite(h) == expected, name
        assert isBipartite(Lattice.from_hopping(h, nt=4)) == expected, name


def test_ill_labelled_tube():
    h = np.zeros((12, 12))
    for i, j in ILL_ADJ:
        h[i, j] = h[j, i] = 1.0
    assert not isBipartite(h)


def test_honeycomb_lattices():
    for l in (3, 4, 6, 8, 12):
        assert isBipartite(honeycomb(l, nt=8))
    assert isBipartite(honeycomb(12, 16, periodic_j=False, nt=8))


def test_random_graphs_agree_with_oracle():
    rng = np.random.RandomState(2024)
    seen = set()
    for trial in range(600):
        n = rng.randint(1, 13)
        h = np.zeros((n, n))
        kind = trial % 3
        for i in range(n):
            for j in range(i + 1, n):
                # kind 0: alternating-parity edges only (bipartite with good labels when
                # connected); 1: any edges; 2: sparse edges (often disconnected)
                p = {0: 0.5 * ((i + j) % 2), 1: 0.3, 2: 0.15}[kind]
                if rng.uniform() < p:
                    h[i, j] = h[j, i] = 1.0
        got, want = isBipartite(h), ref.is_bipartite(h)
        assert got == want, (trial, h)
        seen.add(got)
    assert seen == {True, False}


def test_large_lattice_needs_no_recursion():
    # a 3000-site ring with alternating labels: deeper than Python's default recursion limit
    n = 3000
    h = np.zeros((n, n))
    idx = np.arange(n)
    h[idx, (idx + 1) % n] = h[(idx + 1) % n, idx] = 1.0
    assert isBipartite(h)
