"""Lattices: hopping matrices, the bipartiteness test and lattice generators.

Mirrors src/isle/cpp/lattice.hpp:25-163 (Lattice), src/isle/cpp/lattice.cpp:16-84
(isBipartite) and src/isle/yamlio.py:16-50 (the ``!lattice`` YAML format).
"""

import json
import pathlib

import numpy as np
import yaml


class Lattice:
    """Spatial lattice with a symmetric hopping matrix and a number of time slices."""

    def __init__(self, nt, nx, name="", comment=""):
        self._nt = int(nt)
        self._hopping = np.zeros((int(nx), int(nx)))
        self._positions = np.zeros((int(nx), 3))
        self.name = name
        self.comment = comment

    def nx(self):
        return self._hopping.shape[0]

    def nt(self, nt=None):
        """Return nt, or set it when given (lattice.hpp: nt(std::size_t))."""
        if nt is not None:
            self._nt = int(nt)
        return self._nt

    def lattSize(self):
        return self.nx() * self._nt

    def setNeighbor(self, i, j, strength):
        """Set a symmetric hopping; zero strength removes the bond (lattice.hpp:80-95)."""
        self._hopping[i, j] = strength
        self._hopping[j, i] = strength

    def areNeighbors(self, i, j):
        return self._hopping[i, j] != 0

    def hopping(self):
        """Dense copy of the hopping matrix (nx x nx, float64)."""
        return self._hopping.copy()

    def getNeighbors(self, i):
        return [(int(j), float(self._hopping[i, j])) for j in np.nonzero(self._hopping[i])[0]]

    def position(self, i, x=None, y=None, z=0.0):
        """Position of site i as (x, y, z); sets it first when x and y are given
        (lattice.hpp: position(i) / position(i, x, y, z))."""
        if x is not None:
            self._positions[i] = (x, y, z)
        return tuple(float(v) for v in self._positions[i])

    @classmethod
    def from_hopping(cls, hopping, nt=0, name="", comment=""):
        h = np.asarray(hopping, dtype=np.float64)
        if h.shape[0] != h.shape[1] or not np.array_equal(h, h.T):
            raise ValueError("hopping matrix must be square and symmetric")
        lat = cls(nt, h.shape[0], name, comment)
        lat._hopping[...] = h
        return lat


def isBipartite(lat_or_hopping):
    """True iff the sites split into two sublattices with hopping only between them AND
    every even site is on the sublattice of site 0 and every odd site on the other
    (lattice.cpp:16-84).

    The sublattices come from a depth-first walk from site 0 that visits neighbours in
    increasing index order and does not walk on from the last site (it only labels that
    site's neighbours), as the reference does; a site left unreached fails the label test.
    The walk keeps its own stack of neighbour iterators, so lattice size is not bounded by
    Python's recursion limit.
    """
    h = lat_or_hopping.hopping() if isinstance(lat_or_hopping, Lattice) else np.asarray(lat_or_hopping)
    n = h.shape[0]
    if n == 0:
        return False
    neighbours = [np.flatnonzero(row) for row in h]
    side = np.zeros(n, dtype=np.int8)  # 0: not reached, +1: sublattice of site 0, -1: the other
    side[0] = 1
    walk = [(0, iter(neighbours[0]))]
    while walk:
        site, pending = walk[-1]
        for nxt in pending:
            if side[nxt] == 0:
                side[nxt] = -side[site]
                if site != n - 1:
                    walk.append((nxt, iter(neighbours[nxt])))
                    break
            elif side[nxt] == side[site]:
                return False
        else:
            walk.pop()
    return bool(np.array_equal(side, np.where(np.arange(n) % 2 == 0, 1, -1)))


class _LatticeLoader(yaml.SafeLoader):
    pass


def _construct_lattice(loader, node):
    d = loader.construct_mapping(node, deep=True)
    adjacency = d["adjacency"]
    hopping = d["hopping"]
    if not isinstance(hopping, list):
        hopping = [hopping] * len(adjacency)
    elif len(hopping) != len(adjacency):
        raise RuntimeError("Lengths of adjacency matrix and list of hopping strengths do not match")
    lat = Lattice(d.get("nt", 0), len(d["positions"]), d.get("name", ""), d.get("comment", ""))
    for (i, j), hop in zip(adjacency, hopping):
        lat.setNeighbor(i, j, hop)
    for i, pos in enumerate(d["positions"]):
        if len(pos) not in (2, 3):
            raise RuntimeError(f"Lattice site positions given with {len(pos)} coordinates; "
                               "only 2D and 3D positions are supported")
        lat.position(i, *pos)
    return lat


_LatticeLoader.add_constructor("!lattice", _construct_lattice)


class _LatticeDumper(yaml.SafeDumper):
    pass


def _represent_lattice(dumper, lat):
    adjacency, hopping = [], []
    for i in range(lat.nx()):
        for j, strength in lat.getNeighbors(i):
            if j > i:
                adjacency.append([i, j])
                hopping.append(strength)
    return dumper.represent_mapping("!lattice", {
        "name": lat.name, "comment": lat.comment, "nt": lat.nt(), "adjacency": adjacency,
        "hopping": hopping, "positions": [list(lat.position(i)) for i in range(lat.nx())]},
        flow_style=False)


_LatticeDumper.add_representer(Lattice, _represent_lattice)


def dumpLattice(lat):
    """The ``!lattice`` YAML document of a lattice (each bond once, i < j; the inverse of
    loadLattice; src/isle/yamlio.py:55-75)."""
    return yaml.dump(lat, Dumper=_LatticeDumper)


def loadLattice(text_or_file):
    """Parse a ``!lattice`` YAML document (string or open file) with a safe loader."""
    text = text_or_file.read() if hasattr(text_or_file, "read") else text_or_file
    return yaml.load(text, Loader=_LatticeLoader)


class _BuiltinLattices:
    """The lattices that ship with the package (src/isle/__init__.py:38-80): ``keys()``
    lists their names, ``[name]`` builds one (nt = 0) and ``loadExternal(fname)`` reads a
    ``!lattice`` YAML file.  The data is isle_amd/resources/lattices.json."""

    _DATA = pathlib.Path(__file__).resolve().parent / "resources" / "lattices.json"

    def __init__(self):
        self._table = None

    def _entries(self):
        if self._table is None:
            if not self._DATA.is_file():
                raise RuntimeError(f"the built-in lattice data {self._DATA} is missing from this installation")
            self._table = json.loads(self._DATA.read_text())
        return self._table

    def keys(self):
        return list(self._entries())

    def __contains__(self, name):
        return name in self._entries()

    def __getitem__(self, name):
        entry = self._entries().get(name)
        if entry is None:
            raise ValueError(f"no built-in lattice named {name!r}; known: {', '.join(self.keys())} "
                             "(LATTICES.loadExternal reads a lattice file)")
        lat = Lattice(0, len(entry["positions"]), entry["name"], entry["comment"])
        for (i, j), strength in zip(entry["adjacency"], entry["hopping"]):
            lat.setNeighbor(i, j, strength)
        for i, pos in enumerate(entry["positions"]):
            lat.position(i, *pos)
        return lat

    def loadExternal(self, fname):
        with open(fname) as f:
            return loadLattice(f)


LATTICES = _BuiltinLattices()


def honeycomb(l_i, l_j=None, periodic_j=True, kappa=1.0, nt=0):
    """Honeycomb lattice of l_i x l_j cells, sites A(i,j) = 2(i*l_j + j), B = A + 1.

    Bonds A(i,j)-B(i,j), A(i,j)-B(i-1,j), A(i,j)-B(i,j-1); periodic in i, and in j
    unless ``periodic_j`` is False (a cylinder).  Labels alternate A/B, so the
    torus is bipartite in the sense of isBipartite.
    """
    l_j = l_i if l_j is None else l_j
    nx = 2 * l_i * l_j
    lat = Lattice(nt, nx, f"honeycomb {l_i}x{l_j}{'' if periodic_j else ' (open j)'}")
    for i in range(l_i):
        for j in range(l_j):
            a = 2 * (i * l_j + j)
            for ii, jj in ((i, j), ((i - 1) % l_i, j), (i, j - 1)):
                if jj < 0:
                    if not periodic_j:
                        continue
                    jj %= l_j
                lat.setNeighbor(a, 2 * (ii * l_j + jj) + 1, kappa)
    return lat
