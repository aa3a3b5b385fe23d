"""Output through Simulation.Run on the GPU: the energies, trajectory and
restart files RUMD's default managers write (src/EnergiesOutputManager.cu,
src/ConfigurationOutputManager.cu, src/Sample.cu:150-189), the step-0 energies
line against the FP64 oracle, output leaving the trajectory bit-identical, and
a restart from a restart file."""
from __future__ import annotations

import gzip
import shutil

import numpy as np
import pytest

import oracle
from capi import gpu_available
from helpers import ka_table, load_golden
from rumd_amd import configs, conf_io, rumd
from rumd_amd.Simulation import Simulation

pytestmark = pytest.mark.gpu

SF = 2
KA1000 = "KA_N1000_T0.466.xyz.gz"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _ka_sim(conf, outdir, ps0=0.0):
    sim = Simulation(conf, verbose=False)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    sim.SetPotential(pot)
    itg = rumd.IntegratorNVT(timeStep=0.004, targetTemperature=0.466, thermostatRelaxationTime=0.2)
    itg.SetThermostatState(ps0)
    sim.SetIntegrator(itg)
    sim.SetOutputDirectory(str(outdir))
    return sim, itg


def _energy_lines(path):
    with gzip.open(path, "rt") as fh:
        lines = fh.read().splitlines()
    return lines[0], np.array([[float(x) for x in ln.split()] for ln in lines[1:]])


def test_energies_trajectory_and_restart_files(tmp_path, golden_dir, monkeypatch):
    monkeypatch.chdir(tmp_path)
    c = load_golden(golden_dir / KA1000)
    ps0 = float(c.meta["Nose-Hoover-Ps"])
    out = tmp_path / "TrajectoryFiles"
    sim, itg = _ka_sim(c, out, ps0)
    sim.SetMomentumResetInterval(0)
    sim.SetOutputScheduling("energies", "linear", interval=64)
    sim.SetBlockSize(256)
    print("[case] output files KA_N1000 600 steps", flush=True)
    sim.Run(600)

    # energies: a file per block, a line every 64 steps, RUMD's default columns
    head, e0 = _energy_lines(out / "energies0000.dat.gz")
    assert head == "# ioformat=2 N=1000 Dt=0.256000 columns=ke,pe,p,T,Etot,W"
    assert e0.shape == (4, 6)
    assert _energy_lines(out / "energies0001.dat.gz")[1].shape == (4, 6)
    assert _energy_lines(out / "energies0002.dat.gz")[1].shape == (2, 6)  # steps 512, 576
    assert (out / "LastComplete_energies.txt").read_text() == "1 256\n"

    # the step-0 line against the oracle on the starting configuration
    pos4 = configs.pack_pos4(c.positions, c.types)
    vel4 = configs.pack_vel4(c.velocities, c.types, c.masses)
    f_ref, vir_ref = oracle.forces_brute(pos4, c.box, 2, ka_table(), SF)
    n, V, dof = c.n, float(np.prod(c.box)), 3 * c.n - 3
    ke = oracle.kinetic_energy(pos4, vel4, f_ref.astype(np.float32), 0.004, ps0, 0xFFFFFFFF)
    pe, W = f_ref[:, 3].sum(), vir_ref.sum() / 6.0
    T = 2.0 * ke / dof
    want = [ke / n, pe / n, (n * T + W) / V, T, (pe + ke) / n, W / n]
    np.testing.assert_allclose(e0[0], want, rtol=2e-5, atol=1e-6)
    # every line: Etot = ke + pe, p = (N T + N W) / V, T = 2 N ke / DOF
    for tab in (e0, _energy_lines(out / "energies0002.dat.gz")[1]):
        ke_, pe_, p_, T_, E_, W_ = tab.T
        np.testing.assert_allclose(E_, ke_ + pe_, rtol=0, atol=3e-5)
        np.testing.assert_allclose(p_, (n * T_ + n * W_) / V, rtol=3e-5, atol=1e-5)
        np.testing.assert_allclose(T_, 2 * n * ke_ / dof, rtol=3e-5)

    # trajectory: logarithmic frames, the block end duplicated into the next file
    frames = conf_io.read_trajectory(out / "trajectory0000.xyz.gz")
    assert [int(f.meta["timeStepIndex"]) for f in frames] == [0] + [2 ** k for k in range(9)]
    assert [int(f.meta["timeStepIndex"]) for f in conf_io.read_trajectory(out / "trajectory0002.xyz.gz")] == \
        [0, 1, 2, 4, 8, 16, 32, 64]
    assert frames[0].meta["columns"] == "type,x,y,z,imx,imy,imz,pe,vir"
    np.testing.assert_allclose(frames[0].positions, c.positions, atol=6e-5)
    np.testing.assert_allclose(frames[0].pe, f_ref[:, 3], atol=1e-4)
    b1 = conf_io.read_trajectory(out / "trajectory0001.xyz.gz")
    np.testing.assert_array_equal(b1[0].positions, frames[-1].positions)

    # restart files at each block start, full precision
    r0 = conf_io.read_conf(out / "restart0000.xyz.gz")
    np.testing.assert_allclose(r0.positions, c.positions, rtol=0, atol=1e-8)
    np.testing.assert_allclose(r0.velocities, c.velocities, rtol=0, atol=1e-8)
    assert r0.meta["integrator"].startswith("IntegratorNVT")
    assert (out / "restart0002.xyz.gz").exists()

    # timing file (Simulation.py:73-79, 600-606)
    timing = (tmp_path / "rumd_timing_info.dat").read_text().splitlines()
    assert timing[1].startswith("# nsteps") and timing[2].startswith("600 ")


def test_output_leaves_the_trajectory_unchanged(tmp_path, golden_dir, monkeypatch):
    """Output steps run eagerly, the rest in captured blocks; the final state
    is bit-identical to a run without output."""
    monkeypatch.chdir(tmp_path)
    c = load_golden(golden_dir / KA1000)
    ends = []
    for with_output in (True, False):
        sim, itg = _ka_sim(c, tmp_path / f"TF{int(with_output)}")
        sim.SetOutputScheduling("energies", "linear", interval=50)
        sim.SetOutputMetaData("energies", thermostat_Ps=True)
        print(f"[case] output on/off {with_output}", flush=True)
        sim.Run(700, suppressAllOutput=not with_output)
        ends.append((sim.sample.GetPositions(), sim.sample.GetVelocities(), itg.GetThermostatState()))
    np.testing.assert_array_equal(ends[0][0], ends[1][0])
    np.testing.assert_array_equal(ends[0][1], ends[1][1])
    assert ends[0][2] == ends[1][2]
    head, tab = _energy_lines(tmp_path / "TF1" / "energies0000.dat.gz")
    assert head.endswith("columns=ke,pe,p,T,Ps,Etot,W")
    assert tab.shape[0] == len(range(0, 700, 50))  # block 1024: every multiple of 50
    assert not (tmp_path / "TF0").exists()


def test_restart_continues_the_run(tmp_path, golden_dir, monkeypatch):
    monkeypatch.chdir(tmp_path)
    c = load_golden(golden_dir / KA1000)
    out = tmp_path / "TrajectoryFiles"
    sim, itg = _ka_sim(c, out, float(c.meta["Nose-Hoover-Ps"]))
    sim.SetOutputScheduling("energies", "linear", interval=32)
    sim.SetBlockSize(256)
    print("[case] restart: first run", flush=True)
    sim.Run(768)
    _, first = _energy_lines(out / "energies0002.dat.gz")
    shutil.copy(out / "energies0002.dat.gz", tmp_path / "first_block2.dat.gz")
    del sim
    sim2, itg2 = _ka_sim(c, out)
    sim2.SetOutputScheduling("energies", "linear", interval=32)
    sim2.SetBlockSize(256)
    print("[case] restart: from block 2", flush=True)
    sim2.Run(768, restartBlock=2)
    _, second = _energy_lines(out / "energies0002.dat.gz")
    assert second.shape == first.shape == (8, 6)
    # the restart file holds positions and velocities exactly and Ps to 11
    # digits; the continued run agrees to within chaotic growth over 256 steps
    np.testing.assert_allclose(second, first, rtol=1e-4, atol=1e-4)
