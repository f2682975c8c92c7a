"""Pins oracle.is_bipartite with the reference's own truth table (tests/test_lattice.py:13-77)."""
import numpy as np
import yaml

from oracle import hubbard_ref as ref

# tests/test_lattice.py:13-22 of the reference
TRUTH = [("one_site", True), ("two_sites", True), ("triangle", False), ("four_sites", True),
         ("tetrahedron", False), ("c20", False), ("c60_ipr", False), ("tube_3-3_1", True),
         ("tube_4-2_2", True), ("graphene_7_5", True)]

# adjacency of the ill-labelled (3,3) tube of tests/test_lattice.py:24-63 (sites 2 and 3 exchanged)
ILL_ADJ = [[0, 1], [0, 11], [0, 11], [1, 3], [1, 3], [3, 2], [2, 4], [2, 4], [4, 5], [5, 6],
           [5, 6], [6, 7], [7, 8], [7, 8], [8, 9], [9, 10], [9, 10], [10, 11]]


def test_truth_table(lattice):
    for name, expected in TRUTH:
        assert ref.is_bipartite(lattice(name)) == expected, name


def test_ill_labelled():
    h = np.zeros((12, 12))
    for i, j in ILL_ADJ:
        h[i, j] = h[j, i] = 1.0
    assert not ref.is_bipartite(h)


def test_honeycomb_configs_bipartite(lattice):
    # the synthetic honeycomb tori/cylinders have alternating A/B labels
    for l in (3, 4, 6):
        assert ref.is_bipartite(ref.honeycomb(l))
    assert ref.is_bipartite(ref.honeycomb(3, 4, periodic_j=False))
    # 18_sites.yml ("3x3 hex lattice") is isospectral with the 3x3 honeycomb torus
    e1 = np.linalg.eigvalsh(lattice("18_sites"))
    e2 = np.linalg.eigvalsh(ref.honeycomb(3))
    assert np.allclose(e1, e2, atol=1e-12)
    assert ref.is_bipartite(lattice("18_sites"))
