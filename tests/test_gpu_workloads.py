"""GPU parity on the benchmark workloads at their full sizes, against the CPU
oracle:

- Kob-Andersen 80:20, N = 1,048,576 (FCC 64^3 start, rho = 1.2, L = 95.6033,
  68^3 cell grid, 17^3 tiles), shifted-force cut-off, Nose-Hoover T = 0.45,
  FP32 and mixed positions: per-particle forces, PE and virial sums, and the
  neighbour sets of the cell-sorted build bit for bit, at the start and after
  200 graph-captured steps (the liquid, with its halo capacity and shifted
  frames);
- single-component LJ, N = 1000 (10^3 simple-cubic sites with jitter,
  rho = 0.8442, T = 0.72), shifted potential, NVE dt = 0.005 for 1000 steps
  (N2 lists): the short-time trajectory and the total energy along the run
  against the oracle's FP64 MD (Python/rumd/Simulation.py:579-597 is the
  reference's loop)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, lj_table, sorted_rows
from rumd_amd import configs, conf_io
import rumd_amd.rumd as rumd

pytestmark = pytest.mark.gpu
SP, SF = 1, 2


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _conf(sysm):
    return conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities, images=np.zeros_like(sysm.positions))


def _ka_1m_state_checks(s, pot, sysm, what, lists, nbar_range=(88.0, 97.0)):
    """Forces, PE and W against the FP64 oracle at the sample's current float
    positions; with `lists`, the sets of the current list (built at these
    positions) against the FP32 restatement of RUMD's cell build, and the mean
    list length within nbar_range."""
    pos4 = configs.pack_pos4(s.GetPositions(), sysm.types)
    table = ka_table()
    print(f"[case] KA N=1048576 {what}: oracle forces", flush=True)
    f_ref, vir_ref = oracle.forces_cells(pos4, sysm.box, 2, table, SF)
    e = force_errors(s.GetForces(), f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, (what, e)
    pe_ref, w_ref = float(f_ref[:, 3].sum()), float(vir_ref.sum()) / 6.0
    assert abs(s.GetPotentialEnergy() - pe_ref) <= 1e-5 * abs(pe_ref), what
    assert abs(s.GetVirial() - w_ref) <= 1e-5 * abs(w_ref), what
    if not lists:
        return
    print(f"[case] KA N=1048576 {what}: oracle lists", flush=True)
    b8 = configs.box8(sysm.box)
    nc = oracle.cell_grid(b8, 2.5, 0.3)
    assert list(nc) == [68, 68, 68]
    off_ref, idx_ref = oracle.nblist_cells_f32(pos4, b8, 2, configs.cut_skin2_table(table, 0.3), nc)
    off, ids = pot.GetNeighborListCSR(s)
    assert np.array_equal(off, off_ref), what
    assert np.array_equal(sorted_rows(off, ids), sorted_rows(off_ref, idx_ref)), what
    nbar = off[-1] / sysm.n
    assert nbar_range[0] < nbar < nbar_range[1], nbar


@pytest.mark.parametrize("mixed", [False, True])
def test_ka_1m_forces_sums_and_sets(mixed):
    sysm = configs.ka_random_lattice(1048576, density=1.2, temperature=0.45, seed=3001, lattice="fcc")
    assert abs(sysm.box[0] - 95.6033) < 1e-4
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetMixedPrecisionPositions(mixed)
    s.SetConfiguration(_conf(sysm))
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    s.SetPotential(pot)
    itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    s.SetIntegrator(itg)
    print(f"[case] KA N=1048576 mixed={mixed}: FCC start", flush=True)
    s.CalcF()
    # the lattice's neighbour shells give longer lists than the liquid's ~92
    _ka_1m_state_checks(s, pot, sysm, "FCC start", lists=True, nbar_range=(100.0, 112.0))
    print(f"[case] KA N=1048576 mixed={mixed}: 200 NVT steps", flush=True)
    s.RunSteps(0, 200, 10, 100)
    builds = pot.GetNumberOfNbListBuilds()
    assert builds >= 3  # the lattice melts: skin-triggered rebuilds inside captured blocks
    s.CalcF()  # forces from the list of the last skin-triggered build
    _ka_1m_state_checks(s, pot, sysm, "after 200 steps (list of the last build)", lists=False)
    s.SortParticles()  # rebuild at the current positions, then compare the sets
    s.CalcF()
    # a full sort; one more build when it changes the halo capacity
    assert pot.GetNumberOfNbListBuilds() in (builds + 1, builds + 2)
    _ka_1m_state_checks(s, pot, sysm, "after 200 steps (rebuilt)", lists=True)
    st = pot.GetNbTileStats()
    cap = pot.GetNbHaloCapacity()
    print(f"[case] tiles {st}, halo capacity {cap}", flush=True)
    assert st["tiles"] == 17 ** 3
    assert st["max_halo"] <= cap <= 2048  # every halo staged in LDS, u16 byte entries
    assert st["staged"] == st["tiles"] and st["fast"] == st["tiles"]
    # the first and last tile of each axis (of 17) have halos across the box;
    # at L = 95.6 all of them take the exact shifted frame
    assert st["wrapping"] == 17 ** 3 - 15 ** 3 and st["shifted"] == st["wrapping"]


def test_lj_1000_nve_against_oracle_md():
    """1000 NVE steps of the LJ liquid: positions follow the FP64 oracle over
    the first 100 steps (before chaos amplifies FP32 rounding), and the total
    energy per particle stays within the oracle run's own O(dt^2) band."""
    sysm = configs.lj_liquid(10, 0.8442, 0.72)
    dt, chunk, chunks = 0.005, 50, 20
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(_conf(sysm))
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.5)
    pot.SetNbListSkin(0.3)
    s.SetPotential(pot)
    itg = rumd.IntegratorNVE(timeStep=dt)
    s.SetIntegrator(itg)
    assert pot.GetNB_Method() == "default"
    n = sysm.n
    table = lj_table()
    pos = sysm.positions.astype(np.float64).copy()
    vel = sysm.velocities.astype(np.float64).copy()

    def oracle_energy():
        p4 = configs.pack_pos4(pos.astype(np.float32), sysm.types)
        f, _ = oracle.forces_brute(p4, sysm.box, 1, table, SP)
        vf = vel + f[:, :3] * dt
        return (f[:, 3].sum() + 0.25 * ((vel ** 2).sum() + (vf ** 2).sum())) / n

    s.CalcF()
    e_gpu = [(s.GetPotentialEnergy() + itg.GetKineticEnergy()) / n]
    e_orc = [oracle_energy()]
    assert abs(e_gpu[0] - e_orc[0]) < 1e-5
    for c in range(chunks):
        print(f"[case] LJ N=1000 NVE steps {c * chunk}-{(c + 1) * chunk}", flush=True)
        s.RunSteps(c * chunk, chunk, 10, 0)
        s.CalcF()
        e_gpu.append((s.GetPotentialEnergy() + itg.GetKineticEnergy()) / n)
        oracle.md_run(pos, vel, sysm.types, np.ones(n), sysm.box, 1, table, SP, 0.3, dt, chunk, False, 0.72, 0.0,
                      0.0)
        e_orc.append(oracle_energy())
        if c < 2:
            d = s.GetPositions().astype(np.float64) - pos
            d -= sysm.box * np.rint(d / sysm.box)
            assert np.abs(d).max() < (1e-4 if c == 0 else 1e-3), (c, np.abs(d).max())
    e_gpu, e_orc = np.array(e_gpu), np.array(e_orc)
    print(f"[case] energies gpu {e_gpu[[0, 1, -1]]} oracle {e_orc[[0, 1, -1]]}", flush=True)
    band = np.ptp(e_orc)
    assert band < 3e-3
    assert np.abs(e_gpu - e_orc).max() < max(band, 1e-3)
    assert abs((e_gpu[-1] - e_gpu[0]) - (e_orc[-1] - e_orc[0])) < max(band, 1e-3)
    assert pot.GetNumberOfNbListBuilds() >= 5
