"""Output managers on the host (no GPU): the log-lin schedule against the
step sequences its definition fixes (linear, logarithmic with duplicated
block ends, log-lin, limited linear, linear with an interval that does not
divide the block), the skip-ahead used by the run loop, and the energies file
format with a stand-in sample."""
from __future__ import annotations

import gzip

import numpy as np
import pytest

from rumd_amd import output


class Recorder(output.LogLinOutputManager):
    def __init__(self, tmp_path):
        super().__init__(None, str(tmp_path / "out"), "rec")
        self.verbose = False
        self.writes = []
        self._t = None

    def Write(self):
        self.writes.append((self._t, self.logLin.block, self.logLin.index))

    def Update(self, t):
        self._t = t
        super().Update(t)


def _manager(tmp_path, bs, base, max_interval, user_max=-1, dup=None):
    m = Recorder(tmp_path)
    m.SetBlockSize(bs)
    m.SetLogLinParams(base, max_interval)
    m.SetUserMaxIndex(user_max)
    # what Sample.SetLogLinParameters decides, unless given
    m.SetDuplicateBlockEnds(not (base == max_interval or user_max >= 0) if dup is None else dup)
    m.Initialize(0, True)
    return m


def _run_every_step(m, nsteps):
    for t in range(nsteps):
        m.Update(t)
    return m.writes


def _run_skipping(m, nsteps):
    t = 0
    while t < nsteps:
        nxt = m.next_write_step(t)
        if nxt is None or nxt >= nsteps:
            break
        assert nxt >= t
        m.Update(nxt)
        t = nxt + 1
    return m.writes


def test_linear_schedule(tmp_path):
    m = _manager(tmp_path, 1024, 256, 256)
    w = _run_every_step(m, 2100)
    assert [t for t, _, _ in w] == list(range(0, 2100, 256))
    assert [(b, i) for t, b, i in w[:6]] == [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0), (1, 1)]


def test_logarithmic_schedule_duplicates_block_ends(tmp_path):
    m = _manager(tmp_path, 1024, 1, 0)
    w = _run_every_step(m, 2049)
    block0 = [t for t, b, _ in w if b == 0]
    assert block0 == [0] + [2 ** k for k in range(11)]  # 0, 1, 2, 4, ..., 1024
    # step 1024 closes block 0 (index 11) and opens block 1 (index 0)
    assert (1024, 0, 11) in w and (1024, 1, 0) in w
    block1 = [t for t, b, _ in w if b == 1]
    assert block1 == [1024] + [1024 + 2 ** k for k in range(11)]


def test_loglin_schedule(tmp_path):
    m = _manager(tmp_path, 1024, 1, 16)
    w = _run_every_step(m, 1025)
    block0 = [t for t, b, _ in w if b == 0]
    assert block0 == [0, 1, 2, 4, 8] + [16 * k for k in range(1, 65)]
    assert w[-1] == (1024, 1, 0)


def test_limited_linear_schedule(tmp_path):
    # limlin interval 5, numitems 3: the first three items of every block
    m = _manager(tmp_path, 1024, 5, 5, user_max=2)
    w = _run_every_step(m, 3000)
    assert [t for t, _, _ in w] == [0, 5, 10, 1024, 1029, 1034, 2048, 2053, 2058]


def test_uneven_linear_schedule(tmp_path):
    # an interval that does not divide the block: every multiple of it, the
    # phase carried across blocks
    m = _manager(tmp_path, 1024, 300, 300)
    w = _run_every_step(m, 5000)
    assert [t for t, _, _ in w] == list(range(0, 5000, 300))
    assert [b for t, b, _ in w if t in (900, 1200, 1800, 2100)] == [0, 1, 1, 2]


@pytest.mark.parametrize("par", [(1024, 256, 256, -1), (1024, 1, 0, -1), (1024, 1, 16, -1), (1024, 5, 5, 2),
                                 (1024, 300, 300, -1), (2048, 4, 64, -1)])
def test_skip_ahead_matches_every_step(tmp_path, par):
    bs, base, mi, um = par
    a = _run_every_step(_manager(tmp_path / "a", bs, base, mi, um), 5000)
    b = _run_skipping(_manager(tmp_path / "b", bs, base, mi, um), 5000)
    assert a == b and len(a) > 3


def test_schedule_errors(tmp_path):
    with pytest.raises(RuntimeError, match="blockSize has not been set"):
        _manager(tmp_path, 0, 1, 0)
    with pytest.raises(RuntimeError, match="not a multiple of base"):
        _manager(tmp_path, 1024, 3, 16)
    with pytest.raises(RuntimeError, match="greater than the blockSize"):
        _manager(tmp_path, 64, 128, 128)


def test_restart_manager_writes_each_block_start(tmp_path):
    # the restart manager: base = blockSize, no duplicated block ends
    m = _manager(tmp_path, 1000, 1000, 0, dup=False)
    w = _run_every_step(m, 3001)
    assert [t for t, _, _ in w] == [0, 1000, 2000, 3000]
    m2 = Recorder(tmp_path)
    m2.SetBlockSize(1000)
    m2.SetLogLinParams(1000, 0)
    m2.Initialize(2000, False)  # a restart at block 2: the directory must exist
    m2.Update(2000)
    assert m2.writes == [(2000, 2, 0)]
    m3 = Recorder(tmp_path / "missing")
    m3.SetBlockSize(1000)
    m3.SetLogLinParams(1000, 0)
    with pytest.raises(RuntimeError, match="Output directory does not exist"):
        m3.Initialize(2000, False)


def test_existing_directory_is_backed_up(tmp_path):
    d = tmp_path / "out"
    d.mkdir()
    (d / "old.txt").write_text("x")
    m = _manager(tmp_path, 1024, 256, 256)
    assert (tmp_path / "out.rumd_bak" / "old.txt").exists() and not (d / "old.txt").exists()
    assert m.outputDirectory == str(d)


# ------------------------------------------------------------------ energies file
class _Box:
    def GetVolume(self):
        return 1000.0


class _Itg:
    def GetTimeStep(self):
        return 0.0025

    def GetKineticEnergy(self, copy=True):
        return 675.0

    def GetThermostatState(self):
        return 0.125

    def GetTotalMomentum(self):
        return [2.0, -4.0, 1.0, 1000.0]


class _Pot:
    def GetID_String(self):
        return "potLJ_12_6"

    def GetPotentialEnergy(self):
        return -5900.0


class _Sample:
    def __init__(self):
        self.itg, self.pot = _Itg(), _Pot()

    def GetNumberOfParticles(self):
        return 1000

    def GetIntegrator(self):
        return self.itg

    def GetNumberOfDOFs(self):
        return 2997

    def GetSimulationBox(self):
        return _Box()

    def GetPotentialEnergy(self):
        return -6000.0

    def GetVirial(self):
        return 1500.0

    def GetPotentials(self):
        return [self.pot]


def _energies(tmp_path, sample, interval=10):
    m = output.EnergiesOutputManager(sample, str(tmp_path / "e"))
    m.SetVerbose(False)
    m.RegisterIntegrator(sample.GetIntegrator())
    m.RegisterPotential(sample.GetPotentials()[0])
    m.SetBlockSize(100)
    m.SetLogLinParams(interval, interval)
    m.SetDuplicateBlockEnds(False)
    return m


def test_energies_file_default_columns(tmp_path):
    s = _Sample()
    m = _energies(tmp_path, s)
    m.Initialize(0, True)
    for t in range(0, 120):
        m.Update(t)
    m.Terminate()
    with gzip.open(tmp_path / "e" / "energies0000.dat.gz", "rt") as fh:
        lines = fh.read().splitlines()
    # columns in the sorted order of their keys; Dt = dt * interval for linear saving
    assert lines[0] == "# ioformat=2 N=1000 Dt=0.025000 columns=ke,pe,p,T,Etot,W"
    assert len(lines) == 1 + 10
    vals = [float(x) for x in lines[1].split()]
    N, V, dof = 1000, 1000.0, 2997
    ke, pe, W = 675.0, -6000.0, 1500.0
    T = 2 * ke / dof
    assert vals == pytest.approx([ke / N, pe / N, (N * T + W) / V, T, (pe + ke) / N, W / N], rel=1e-6)
    assert (tmp_path / "e" / "energies0001.dat.gz").exists()
    assert (tmp_path / "e" / "LastComplete_energies.txt").read_text() == "0 100\n"


def test_energies_file_optional_columns(tmp_path):
    s = _Sample()
    m = _energies(tmp_path, s)
    for key in ("thermostat_Ps", "v_com_x", "density", "potLJ_12_6"):
        m.SetMetaData(key, True)
    m.SetMetaData("kineticEnergy", False)
    m.SetMetaData("precision", 9)
    with pytest.raises(RuntimeError, match="Unrecognized key"):
        m.SetMetaData("no_such_column", True)
    m.Initialize(0, True)
    m.Update(0)
    m.Terminate()
    with gzip.open(tmp_path / "e" / "energies0000.dat.gz", "rt") as fh:
        head, line = fh.read().splitlines()
    # byte order of the keys, as a std::map orders them ('L' < 't', '_' < 'i')
    assert head.endswith("columns=rho,potLJ_12_6,pe,p,T,Ps,Etot,v_com_x,W")
    vals = [float(x) for x in line.split()]
    assert vals[0] == pytest.approx(1.0) and vals[1] == pytest.approx(-5.9)
    assert vals[5] == pytest.approx(0.125) and vals[7] == pytest.approx(0.002)


def test_energies_log_schedule_header(tmp_path):
    s = _Sample()
    m = output.EnergiesOutputManager(s, str(tmp_path / "e"))
    m.SetVerbose(False)
    m.SetBlockSize(64)
    m.SetLogLinParams(1, 0)
    m.SetDuplicateBlockEnds(True)
    m.Initialize(0, True)
    for t in range(65):
        m.Update(t)
    m.Terminate()
    with gzip.open(tmp_path / "e" / "energies0000.dat.gz", "rt") as fh:
        lines = fh.read().splitlines()
    assert lines[0] == "# ioformat=2 N=1000 timeStepIndex=0 logLin=0,1,0,7,0 columns=ke,pe,p,T,Etot,W"
    assert len(lines) == 1 + 8  # 0, 1, 2, 4, ..., 64


def test_trajectory_writer_tokens(tmp_path):
    from rumd_amd import conf_io
    c = conf_io.Configuration(n=2, box=np.array([5.0, 5.0, 5.0]), types=np.array([0, 1], np.int32),
                              positions=np.array([[0.1, 0.2, 0.3], [-1.0, 2.0, -2.0]], np.float32),
                              masses=np.array([1.0, 2.0]), images=np.zeros((2, 3), np.float32),
                              pe=np.array([-1.5, 2.25]), vir=np.array([0.5, 0.75]))
    fn = tmp_path / "t.xyz.gz"
    conf_io.write_conf(fn, c, precision=4, integrator_info="IntegratorNVT,0.0025,0.45,0.2,0",
                       box_info="RectangularSimulationBox,5,5,5", loglin=(0, 0, 1, 0, 10, 0))
    c.positions = c.positions + 0.5
    conf_io.write_conf(fn, c, precision=4, box_info="RectangularSimulationBox,5,5,5", loglin=(1, 0, 1, 1, 10, 0),
                       mode="a")
    frames = conf_io.read_trajectory(fn)
    assert len(frames) == 2
    assert frames[0].meta["logLin"] == "0,1,0,10,0" and frames[1].meta["timeStepIndex"] == "1"
    assert frames[0].meta["columns"] == "type,x,y,z,imx,imy,imz,pe,vir"
    np.testing.assert_allclose(frames[1].positions, c.positions, atol=1e-4)
    with gzip.open(fn, "rt") as fh:
        head = fh.read().splitlines()[1]
    assert head.startswith("ioformat=2 timeStepIndex=0 logLin=0,1,0,10,0 numTypes=2 integrator=IntegratorNVT")
