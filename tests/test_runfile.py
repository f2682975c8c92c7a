"""Run files and the HMC driver's checkpoint / resume, on the CPU.

The crash / resume test drives the real driver (isle_amd.drivers.hmc) with a numpy
evolver defined here (a Metropolis walk on a Gaussian action that consumes the run's RNG
as a leapfrog evolver does: normal draws, then a uniform when dE >= 0), so no GPU is
needed: a run that crashes after trajectory 7 and resumes from checkpoint 6 must write
configurations 7..12 bit-identical to those of an uninterrupted run.  The GPU test of the
same property with the fermion action is tests/test_gpu_hmc_resume.py.
"""
import numpy as np
import pytest

from isle_amd import runfile
from isle_amd.action import HFABasis, HFAHopping
from isle_amd.drivers import hmc
from isle_amd.evolver import BinarySelector, EvolutionStage, LinearStepLeapfrog, NumpyRNG, readStateH5
from isle_amd.lattice import dumpLattice, honeycomb, loadLattice
from isle_amd.util import dumpParameters, loadParameters, parameters


def make_gaussian_action(lattice, params):
    import numpy as np

    class GaussianAction:
        def eval(self, phi):
            phi = np.asarray(phi)
            return complex(0.5 * np.vdot(phi, phi).real / params.U)

    return GaussianAction()


class GaussianEvolver:
    """Metropolis random walk in numpy with the interface of the reference's evolvers."""

    def __init__(self, action, rng, step):
        self.action = action
        self.rng = rng
        self.step = step
        self.selector = BinarySelector(rng)
        self.trajPoints = []

    def evolve(self, stage):
        phi = np.asarray(stage.phi)
        prop = phi + self.step * self.rng.normal(0, 1, phi.size)
        act = self.action.eval(prop)
        tp = self.selector.selectTrajPoint(stage.sumLogWeights(), act)
        self.trajPoints.append(tp)
        return stage.accept(prop, act) if tp == 1 else stage.reject()

    def save(self, group, manager):
        group["step"] = self.step

    @classmethod
    def fromH5(cls, group, manager, action, lattice, rng):
        return cls(action, rng, float(group["step"][()]))

    def report(self):
        return f"GaussianEvolver step={self.step} acceptance={np.mean(self.trajPoints)}"


class Crash(RuntimeError):
    pass


class CrashingEvolver(GaussianEvolver):
    """Dies (like a lost node) when asked for trajectory ``crash_at``."""

    crash_at = None

    def evolve(self, stage):
        if len(self.trajPoints) + 1 == self.crash_at:
            raise Crash("node lost")
        return super().evolve(stage)


DEFS = {"GaussianEvolver": GaussianEvolver, "CrashingEvolver": GaussianEvolver}
PARAMS = parameters(beta=4.0, U=2.0, mu=0.0, sigmaKappa=-1, hopping=HFAHopping.EXP, basis=HFABasis.PARTICLE_HOLE)


def _lattice():
    lat = honeycomb(2, nt=4)
    for i in range(lat.nx()):
        lat.position(i, 0.5 * i, 0.25 * (i % 3), 0.0)
    return lat


def _start(seed, lat):
    rng = NumpyRNG(seed)
    phi = rng.normal(0, PARAMS.tilde("U", lat) ** 0.5, lat.lattSize()) + 0j
    return rng, phi


def _configs(run):
    out = {}
    for i in runfile.indices(run, "configuration"):
        g = runfile.readTrajectory(run, i)
        out[i] = (g["phi"][()], complex(g["actVal"][()]), int(g["trajPoint"][()]))
    return out


def _uninterrupted(tmp_path, ntr=12):
    lat = _lattice()
    rng, phi = _start(1075, lat)
    drv = hmc.newRun(lat, PARAMS, rng, make_gaussian_action, tmp_path / "a", False, DEFS)
    ev = GaussianEvolver(drv.action, rng, 0.4)
    drv(phi, ev, ntr, saveFreq=1, checkpointFreq=3)
    return tmp_path / "a"


def test_lattice_yaml_roundtrip():
    lat = _lattice()
    back = loadLattice(dumpLattice(lat))
    assert np.array_equal(back.hopping(), lat.hopping())
    assert back.nt() == lat.nt() and back.name == lat.name
    assert all(back.position(i) == lat.position(i) for i in range(lat.nx()))


def test_parameters_roundtrip_and_tilde():
    back = loadParameters(dumpParameters(PARAMS))
    assert back.asdict() == PARAMS.asdict()
    assert back.hopping is HFAHopping.EXP and back.basis is HFABasis.PARTICLE_HOLE
    assert back.tilde("U", 16) == 2.0 * 4.0 / 16 and back.tilde(3.0, _lattice()) == 3.0 * 4.0 / 4
    with pytest.raises(Exception):
        back.U = 1.0  # read-only, like the reference's parameters


def test_group_archive_roundtrip(tmp_path):
    g = runfile.Group("/x")
    g["phi"] = np.arange(6) * (1 + 2j)
    g["s"] = "np.MT19937"
    g["n"] = 7
    sub = g.create_group("sub")
    sub.attrs["pos"] = 3
    sub["z"] = 1.5 - 2j
    g.create_group("empty")
    with pytest.raises(ValueError):
        g["n"] = 8
    with pytest.raises(ValueError):
        g.create_group("sub")
    with pytest.raises(TypeError):
        g["bad"] = [object()]
    runfile._write_archive(tmp_path / "x.npz", g)
    with pytest.raises(ValueError):
        runfile._write_archive(tmp_path / "x.npz", g)
    back = runfile._read_archive(tmp_path / "x.npz", "/x")
    assert np.array_equal(back["phi"][()], g["phi"][()]) and back["phi"].dtype == np.complex128
    assert str(back["s"][()]) == "np.MT19937" and back["n"][()] == 7
    assert back["sub"].attrs["pos"] == 3 and back["sub"]["z"][()] == 1.5 - 2j
    assert "empty" in back and len(back["empty"]) == 0


@pytest.mark.parametrize("ndraw", [4, 5])  # even / odd: numpy's cached second Gaussian
def test_rng_state_roundtrip(ndraw):
    rng = NumpyRNG(3)
    rng.normal(0, 1, ndraw)
    rng.normal()
    g = runfile.Group()
    rng.writeH5(g)
    want = (rng.normal(0, 1, 9), rng.uniform(0, 1, 3))
    back = readStateH5(g)
    got = (back.normal(0, 1, 9), back.uniform(0, 1, 3))
    assert all(np.array_equal(a, b) for a, b in zip(want, got))


def test_stage_roundtrip():
    st = EvolutionStage(np.arange(4) + 1j, 2.5 - 1j, 0, {"logdetJ": 0.5 + 0.25j}, {"acc": np.arange(3)})
    g = runfile.Group()
    st.save(g)
    back = EvolutionStage.fromH5(g)
    assert np.array_equal(back.phi, st.phi) and back.actVal == st.actVal and back.trajPoint == 0
    assert back.logWeights == st.logWeights and np.array_equal(back.extra["acc"], np.arange(3))


def test_crash_resume_is_bit_identical(tmp_path):
    ref = _configs(_uninterrupted(tmp_path))
    assert sorted(ref) == list(range(1, 13))
    assert 0 < np.mean([v[2] for v in ref.values()]) < 1  # both accepts and rejects happen

    lat = _lattice()
    rng, phi = _start(1075, lat)
    drv = hmc.newRun(lat, PARAMS, rng, make_gaussian_action, tmp_path / "b", False, DEFS)
    ev = CrashingEvolver(drv.action, rng, 0.4)
    ev.crash_at = 8
    with pytest.raises(Crash):
        drv(phi, ev, 12, saveFreq=1, checkpointFreq=3)
    run = tmp_path / "b"
    assert runfile.indices(run, "configuration") == list(range(1, 8))
    assert runfile.indices(run, "checkpoint") == [3, 6]

    with pytest.raises(RuntimeError, match="after trajectory"):  # configuration 7 is newer than checkpoint 6
        hmc.continueRun(run, None, -1, False, DEFS)
    # the action is rebuilt from the source stored in the run (makeAction not passed)
    drv2, stage, ev2, save_freq, ckpt_freq = hmc.continueRun(run, None, -1, True, DEFS)
    assert (save_freq, ckpt_freq) == (1, 3) and drv2.trajIdx == 6 and isinstance(ev2, GaussianEvolver)
    assert runfile.indices(run, "configuration") == list(range(1, 7))
    drv2(stage, ev2, 6, saveFreq=1, checkpointFreq=3)
    got = _configs(run)
    assert sorted(got) == sorted(ref)
    for i in ref:
        assert np.array_equal(got[i][0], ref[i][0]) and got[i][1:] == ref[i][1:], i
    # the generators are in the same state at every common checkpoint
    for i in (9, 12):
        a = runfile.readCheckpoint(tmp_path / "a", i)["rngState"]
        b = runfile.readCheckpoint(run, i)["rngState"]
        assert np.array_equal(a["keys"][()], b["keys"][()]) and a.attrs == b.attrs


def test_continue_into_new_run(tmp_path):
    a = _uninterrupted(tmp_path)
    drv, stage, ev, _, _ = hmc.continueRun(a, tmp_path / "c", 6, False, DEFS, makeAction=make_gaussian_action)
    drv(stage, ev, 6, saveFreq=1, checkpointFreq=3)
    ref, got = _configs(a), _configs(tmp_path / "c")
    assert sorted(got) == list(range(7, 13))
    for i in got:
        assert np.array_equal(got[i][0], ref[i][0]) and got[i][1:] == ref[i][1:]


def test_driver_errors(tmp_path):
    a = _uninterrupted(tmp_path)
    with pytest.raises(ValueError, match="outside the checkpointed range"):
        hmc.continueRun(a, tmp_path / "d", 13, False, DEFS)
    with pytest.raises(ValueError, match="no checkpoint was written"):
        hmc.continueRun(a, tmp_path / "e", 4, False, DEFS)
    with pytest.raises(RuntimeError, match="exists"):
        hmc.newRun(_lattice(), PARAMS, NumpyRNG(1), make_gaussian_action, a, False, DEFS)
    other = tmp_path / "not_a_run"
    other.mkdir()
    (other / "keep.txt").write_text("x")
    with pytest.raises(RuntimeError, match="not a run"):
        hmc.newRun(_lattice(), PARAMS, NumpyRNG(1), make_gaussian_action, other, True, DEFS)
    assert (other / "keep.txt").exists()
    drv = hmc.newRun(_lattice(), PARAMS, NumpyRNG(1), make_gaussian_action, tmp_path / "f", False, DEFS)
    with pytest.raises(ValueError, match="configurations are written every"):
        drv(np.zeros(_lattice().lattSize(), complex), GaussianEvolver(drv.action, drv.rng, 0.1), 2, 2, 3)
    with pytest.raises(ValueError, match="definitions"):  # a checkpoint of an unknown evolver type
        hmc.continueRun(a, tmp_path / "g", 3, False, {})


def test_interrupted_write_leaves_nothing(tmp_path, monkeypatch):
    a = _uninterrupted(tmp_path, ntr=3)

    def dying_savez(f, **arrays):
        f.write(b"PK partial")
        raise Crash("killed while writing")

    monkeypatch.setattr(runfile.np, "savez", dying_savez)
    st = EvolutionStage(np.zeros(16, complex), 0.0, 1)
    with pytest.raises(Crash):
        runfile.writeTrajectory(a, 4, st)
    assert runfile.indices(a, "configuration") == [1, 2, 3]  # the temporary file is not a trajectory


def test_linear_step_schedule_resumes():
    """The leapfrog schedule: length and step count move linearly over ninterp
    trajectories, then stay; a checkpointed evolver continues where it stopped."""
    ev = LinearStepLeapfrog(None, (1.0, 0.5), (20, 5), 4, NumpyRNG(0))
    lengths = [next(ev._lengths) for _ in range(6)]
    nsteps = [int(next(ev._nsteps)) for _ in range(6)]
    assert lengths == pytest.approx([1.0, 0.875, 0.75, 0.625, 0.5, 0.5])
    assert nsteps == [20, 16, 12, 8, 5, 5]
    ev = LinearStepLeapfrog(None, (1.0, 0.5), (20, 5), 4, NumpyRNG(0), startPoint=2)
    ev._current += 1  # one trajectory run after resuming at 2
    next(ev._lengths)
    next(ev._nsteps)
    g = runfile.Group()
    ev.save(g, None)
    assert g["current"][()] == 3
    back = LinearStepLeapfrog.fromH5(g, None, None, None, NumpyRNG(0))
    assert next(back._lengths) == pytest.approx(0.625) and int(next(back._nsteps)) == 8


def test_evolver_with_transform_checkpoint():
    """A leapfrog evolver's transform goes into its checkpoint and comes back."""
    from isle_amd.transform import ConstantShift, Identity
    from isle_amd.evolver import ConstStepLeapfrog
    lat = honeycomb(2, nt=4)
    action = make_gaussian_action(lat, None)
    mgr = runfile.EvolverManager()
    shift = np.arange(lat.lattSize()) * 0.01 + 0j
    for tr in (None, Identity(), ConstantShift(shift, action)):
        g = runfile.Group()
        mgr.save(ConstStepLeapfrog(action, 0.7, 9, NumpyRNG(1), transform=tr), g)
        back = mgr.load(g, action, lat, NumpyRNG(1))
        assert (back.length, back.nstep) == (0.7, 9)
        if tr is None:
            assert back.transform is None
        else:
            assert type(back.transform) is type(tr)
            assert g["transform"].attrs["__object__kind__"] == "Transform"
        if isinstance(tr, ConstantShift):
            assert np.array_equal(back.transform.shift, shift)
