"""Run files of an HMC run: configurations, checkpoints and metadata.

The layout is the reference's configuration file (docs/pages/file_layout.md,
src/isle/h5io.py:63-181, src/isle/evolver/stage.py:74-117) with the same group and
dataset names.  h5py is not part of this stack, so a run "file" is a directory:

    <run>/meta.json                 lattice (``!lattice`` YAML), params (``!parameters``
                                    YAML), action (source of makeAction), version
    <run>/configuration/<i>.npz     phi, actVal, trajPoint [, phi_RT-P, logWeights/<k>,
                                    extra/<k>]                   (EvolutionStage.save)
    <run>/checkpoint/<i>.npz        cfg (name of the configuration group, the reference's
                                    soft link), rngState/{name, keys} with attributes pos,
                                    has_gauss, cached_gaussian, evolver/{__name__, state}

Inside an archive, groups are ``/``-separated name prefixes, an attribute ``a`` of group
``g`` is stored as ``g/.attrs/a`` and every group leaves a ``g/.group`` marker (so empty
groups survive).  Scalars are 0-d arrays, so ``group[name][()]`` reads a value as it does
with h5py, and code written against the reference's groups (evolver ``save`` /
``fromH5``) runs unchanged.  Archives hold plain numeric and string arrays only and are
read with ``allow_pickle=False``.

Each archive is written under a temporary name, flushed to disk and renamed into place,
so a run killed while writing leaves the complete archive or none; writing over an
existing trajectory or checkpoint is refused, as h5py refuses an existing name.
"""
import json
import os
import shutil
import sys
from pathlib import Path

import numpy as np

from .lattice import dumpLattice, loadLattice
from .util import dumpParameters, loadParameters

VERSION = "0.1"
META = "meta.json"
GROUPS = ("configuration", "checkpoint")


class Group(dict):
    """An HDF5-like group held in memory: items are sub-groups or datasets (numpy arrays),
    ``attrs`` a dict of scalars."""

    def __init__(self, name="/"):
        super().__init__()
        self.name = name
        self.attrs = {}

    def create_group(self, name):
        if name in self:
            raise ValueError(f"Unable to create group (name already exists): {name}")
        grp = Group(self.name.rstrip("/") + "/" + name)
        dict.__setitem__(self, name, grp)
        return grp

    def __setitem__(self, name, value):
        if name in self:
            raise ValueError(f"Unable to create dataset (name already exists): {name}")
        arr = np.asarray(value)
        if arr.dtype == object:
            raise TypeError(f"dataset {name}: only numeric and string data can be stored")
        dict.__setitem__(self, name, arr)


def _flatten(group, prefix, out):
    out[prefix + ".group"] = np.zeros(0)
    for key, val in group.attrs.items():
        out[prefix + ".attrs/" + key] = np.asarray(val)
    for key, val in group.items():
        if isinstance(val, Group):
            _flatten(val, prefix + key + "/", out)
        else:
            out[prefix + key] = val


def _unflatten(arrays, name):
    root = Group(name)
    for key in arrays.files:
        parts = key.split("/")
        is_attr = len(parts) >= 2 and parts[-2] == ".attrs"
        grp = root
        for part in parts[:-2] if is_attr else parts[:-1]:
            grp = grp[part] if part in grp else grp.create_group(part)
        leaf = parts[-1]
        if is_attr:
            grp.attrs[leaf] = arrays[key][()]
        elif leaf != ".group":
            dict.__setitem__(grp, leaf, arrays[key])
    return root


def _write_archive(path, group):
    """Write a group tree to ``path`` (which must not exist) atomically."""
    path = Path(path)
    if path.exists():
        raise ValueError(f"Unable to create group (name already exists): {path}")
    flat = {}
    _flatten(group, "", flat)
    tmp = path.with_name(f".{path.name}.tmp{os.getpid()}")
    with open(tmp, "wb") as f:
        np.savez(f, **flat)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def _read_archive(path, name):
    with np.load(path, allow_pickle=False) as arrays:
        return _unflatten(arrays, name)


def _is_run(path):
    return (Path(path) / META).is_file()


def writeMetadata(fname, lattice, params, makeActionSrc):
    """Store lattice, parameters and the action's source (h5io.py:63-79)."""
    meta = {
        "lattice": dumpLattice(lattice),
        "params": dumpParameters(params),
        "action": makeActionSrc,
        "version": {"isle_amd": VERSION, "python": sys.version.split()[0], "numpy": np.__version__},
    }
    path = Path(fname) / META
    tmp = path.with_name(f".{META}.tmp{os.getpid()}")
    tmp.write_text(json.dumps(meta, indent=1))
    os.replace(tmp, path)


def readMetadata(fname):
    """(lattice, params, makeActionSrc, versions) of a run (h5io.py:81-100)."""
    if isinstance(fname, (tuple, list)):
        fname = fname[0]
    path = Path(fname) / META
    if not path.is_file():
        raise KeyError(f"Cannot read metadata from {fname}: no {META}")
    meta = json.loads(path.read_text())
    return loadLattice(meta["lattice"]), loadParameters(meta["params"]), meta["action"], meta["version"]


def initializeNewFile(fname, overwrite, lattice, params, makeActionSrc, extraGroups=GROUPS):
    """Create a run directory with its metadata (h5io.py:102-121).  An existing run is
    removed first when ``overwrite``; anything at that path that is not a run is left
    alone and is an error."""
    fname = Path(fname)
    if fname.exists():
        if not overwrite:
            raise RuntimeError("Output file exists")
        if not _is_run(fname):
            raise RuntimeError(f"{fname} exists and is not a run directory; not removing it")
        shutil.rmtree(fname)
    fname.mkdir(parents=True)
    for group in extraGroups:
        (fname / group.strip("/")).mkdir()
    writeMetadata(fname, lattice, params, makeActionSrc)


def indices(fname, group):
    """Sorted trajectory indices stored under ``group`` ("configuration" or "checkpoint")."""
    d = Path(fname) / group
    if not d.is_dir():
        return []
    return sorted(int(p.stem) for p in d.glob("*.npz") if p.stem.lstrip("-").isdigit())


def writeTrajectory(fname, label, stage):
    """Store an EvolutionStage as configuration ``label`` (h5io.py:123-138); returns the
    name of the configuration group."""
    name = f"/configuration/{int(label)}"
    grp = Group(name)
    stage.save(grp)
    _write_archive(Path(fname) / "configuration" / f"{int(label)}.npz", grp)
    return name


def readTrajectory(fname, label):
    """The group of configuration ``label``."""
    return _read_archive(Path(fname) / "configuration" / f"{int(label)}.npz", f"/configuration/{int(label)}")


def writeCheckpoint(fname, label, rng, trajGrpName, evolver, evolverManager):
    """Store RNG state, the configuration's group name and the evolver as checkpoint
    ``label`` (h5io.py:140-162)."""
    grp = Group(f"/checkpoint/{int(label)}")
    rng.writeH5(grp.create_group("rngState"))
    grp["cfg"] = trajGrpName
    evolverManager.save(evolver, grp.create_group("evolver"))
    _write_archive(Path(fname) / "checkpoint" / f"{int(label)}.npz", grp)
    return grp


def readCheckpoint(fname, label):
    """The group of checkpoint ``label``."""
    return _read_archive(Path(fname) / "checkpoint" / f"{int(label)}.npz", f"/checkpoint/{int(label)}")


def loadCheckpoint(fname, label, evolverManager, action, lattice):
    """(rng, configuration group, evolver) of checkpoint ``label`` (h5io.py:164-181)."""
    from .evolver import readStateH5
    grp = readCheckpoint(fname, label)
    rng = readStateH5(grp["rngState"])
    cfg_name = str(grp["cfg"][()])
    cfg = _read_archive(Path(fname) / (cfg_name.strip("/") + ".npz"), cfg_name)
    evolver = evolverManager.load(grp["evolver"], action, lattice, rng)
    return rng, cfg, evolver


def loadConfiguration(fname, trajIdx=-1):
    """(phi, actVal) of configuration ``trajIdx``; negative indices count from the last
    stored configuration (h5io.py:183-203)."""
    if trajIdx < 0:
        trajIdx = indices(fname, "configuration")[trajIdx]
    grp = readTrajectory(fname, trajIdx)
    return grp["phi"][()], grp["actVal"][()]


def removeGreaterThan(fname, group, maxIdx):
    """Delete the trajectories or checkpoints of ``group`` with index > maxIdx."""
    for idx in indices(fname, group):
        if idx > maxIdx:
            (Path(fname) / group / f"{idx}.npz").unlink()


class EvolverManager:
    """Saves and restores evolvers and their transforms in checkpoints
    (src/isle/evolver/manager.py).

    The type is stored by name in ``__name__`` of its group (attribute
    ``__object__kind__`` = "Evolver" or "Transform"); the type is looked up among the
    evolvers of isle_amd.evolver, the transforms of isle_amd.transform and the classes in
    ``definitions``.  Source code of a custom evolver
    is never stored or executed: pass the class in ``definitions``."""

    def __init__(self, fname=None, definitions=None):
        self.fname = fname
        self.definitions = dict(definitions or {})

    def _types(self):
        from . import evolver as ev
        from . import transform as tr
        types = {cls.__name__: cls for cls in (ev.ConstStepLeapfrog, ev.LinearStepLeapfrog, ev.TwoPiJumps,
                                               ev.UniformJump, tr.Identity, tr.ConstantShift)}
        types.update(self.definitions)
        return types

    def save(self, evolver, group):
        """Store an evolver, or a transform of one, in ``group``."""
        from .transform import Transform
        name = type(evolver).__name__
        if name not in self._types():
            raise ValueError(f"cannot save an evolver of type {name}: not a built-in evolver and "
                             "not in definitions")
        group.attrs["__object__kind__"] = "Transform" if isinstance(evolver, Transform) else "Evolver"
        group["__name__"] = name
        evolver.save(group, self)

    def load(self, group, action, lattice, rng):
        name = str(group["__name__"][()])
        types = self._types()
        if name not in types:
            raise ValueError(f"checkpoint holds an evolver of type {name}; pass its class in definitions")
        return types[name].fromH5(group, self, action, lattice, rng)
