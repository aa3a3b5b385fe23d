"""End-to-end parity through the C++ host layer (RUMD's Sample / Pot_LJ_12_6 /
IntegratorNVT API), which calls the HIP kernels through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, lj_table, load_golden
from rumd_amd import configs, conf_io
import rumd_amd.rumd as rumd
from rumd_amd.Simulation import Simulation

pytestmark = pytest.mark.gpu
SP, SF = 1, 2


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def ka_sample(conf, cutoff=rumd.ShiftedPotential, skin=0.3, method=None, mixed=False):
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    if mixed:
        s.SetMixedPrecisionPositions(True)
    s.SetConfiguration(conf)
    pot = rumd.Pot_LJ_12_6(cutoff_method=cutoff)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(skin)
    if method:
        pot.SetNB_Method(method)
    s.SetPotential(pot)
    return s, pot


def conf_from_system(sysm):
    return conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities,
                                 images=np.zeros_like(sysm.positions))


def test_calcf_pe_virial_golden(golden_dir):
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    s, pot = ka_sample(c)
    s.CalcF()
    f = s.GetForces()
    f_ref, vir_ref = oracle.forces_brute(configs.pack_pos4(c.positions, c.types), c.box, 2, ka_table(), SP)
    assert force_errors(f, f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
    assert s.GetPotentialEnergy() == pytest.approx(f_ref[:, 3].sum(), rel=1e-6)
    assert s.GetVirial() == pytest.approx(vir_ref.sum() / 6.0, rel=1e-5)
    assert pot.GetPotentialEnergy() == pytest.approx(f_ref[:, 3].sum(), rel=1e-6)
    # per-particle energies and virials against the file RUMD wrote
    fe = s.GetForcesAndEnergies()
    np.testing.assert_allclose(fe[:, 3], c.pe, atol=5e-5)
    np.testing.assert_allclose(s.GetVirials()[:, 3] / 2.0, c.vir, atol=1e-3)


def test_sorted_api_order_and_forces():
    sysm = configs.ka_random_lattice(20000, density=1.2, seed=123)
    conf = conf_from_system(sysm)
    s, pot = ka_sample(conf, cutoff=rumd.ShiftedForce)
    assert pot.GetNB_Method() == "default"
    s.CalcF()
    # host-side getters return original order even though the device sorted
    np.testing.assert_array_equal(s.GetPositions(), sysm.positions)
    np.testing.assert_array_equal(s.GetTypes(), sysm.types)
    f_ref, _ = oracle.forces_cells(configs.pack_pos4(sysm.positions, sysm.types), sysm.box, 2, ka_table(), SF)
    assert force_errors(s.GetForces(), f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
    # the first build sorts from scratch and settles the halo capacity, which
    # rebuilds once more when the capacity changes
    assert pot.GetNumberOfNbListBuilds() in (1, 2)
    assert 100 < pot.GetActualMaxNumNbrs() < 200
    # neighbour sets in original ids, bit-exact against the FP32 restatement
    off, ids = pot.GetNeighborListCSR(s)
    b8 = configs.box8(sysm.box)
    o2, i2 = oracle.nblist_cells_f32(configs.pack_pos4(sysm.positions, sysm.types), b8, 2,
                                     configs.cut_skin2_table(ka_table(), 0.3), oracle.cell_grid(b8, 2.5, 0.3))
    assert np.array_equal(off, o2)
    for i in range(0, sysm.n, 97):
        assert set(ids[off[i]:off[i + 1]].tolist()) == set(i2[o2[i]:o2[i + 1]].tolist())


def _run_api_loop(s, itg, n):
    for _ in range(n):
        s.CalcF()
        itg.Integrate()
    s.Synchronize()


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("mixed", [False, True])
def test_graph_blocks_equal_eager_loop(mixed, fused):
    """RunSteps (HIP-graph blocks of 10 steps) gives bit-identical results to
    eager steps, replay after replay: the CalcF/Integrate loop for the
    two-kernel step, RunSteps with blocks of one step for the fused one."""
    sysm = configs.ka_random_lattice(12000, density=1.2, temperature=0.45, seed=31)
    conf = conf_from_system(sysm)
    out = []
    for mode in ("eager", "graph"):
        s, pot = ka_sample(conf, cutoff=rumd.ShiftedForce, mixed=mixed)
        s.SetFusedSteps(fused)
        itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
        s.SetIntegrator(itg)
        if mode == "eager" and fused:
            s.RunSteps(0, 60, 1, 100)
        elif mode == "eager":
            for i in range(60):
                if i % 100 == 0:
                    itg.SetMomentumToZero()
                s.CalcF()
                itg.Integrate()
            s.Synchronize()
        else:
            s.RunSteps(0, 60, 10, 100)
        out.append((s.GetPositions(), s.GetVelocities(), s.GetImages(), itg.GetThermostatState(),
                    pot.GetNumberOfNbListBuilds()))
    (r0, v0, i0, p0, b0), (r1, v1, i1, p1, b1) = out
    assert b0 == b1 and b0 >= 1
    np.testing.assert_array_equal(r0, r1)
    np.testing.assert_array_equal(v0, v1)
    np.testing.assert_array_equal(i0, i1)
    assert p0 == p1


def test_nvt_trajectory_tracks_oracle(golden_dir):
    """20 NVT steps on the KA_N1000 fixture: GPU (FP32 forces) vs the FP64
    CPU MD of the oracle from the same start."""
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    s, pot = ka_sample(c)
    itg = rumd.IntegratorNVT(timeStep=0.004, targetTemperature=0.466, thermostatRelaxationTime=0.2)
    s.SetIntegrator(itg)
    ps0 = float(c.meta["Nose-Hoover-Ps"])
    itg.SetThermostatState(ps0)
    _run_api_loop(s, itg, 20)
    pos = c.positions.astype(np.float64).copy()
    vel = c.velocities.astype(np.float64).copy()
    omega2 = float(np.float32(4 * np.pi ** 2 / 0.2 ** 2))
    _, _, ps = oracle.md_run(pos, vel, c.types, np.ones(c.n), c.box, 2, ka_table(), SP, 0.3, 0.004, 20, True,
                             0.466, omega2, ps0)
    d = s.GetPositions().astype(np.float64) - pos
    d -= c.box * np.rint(d / c.box)
    assert np.abs(d).max() < 2e-4
    np.testing.assert_allclose(s.GetVelocities(), vel, atol=2e-3)
    assert itg.GetThermostatState() == pytest.approx(ps, abs=2e-4)


def test_nve_energy_drift_small(golden_dir):
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    s, pot = ka_sample(c, cutoff=rumd.ShiftedForce)
    itg = rumd.IntegratorNVE(timeStep=0.004)
    s.SetIntegrator(itg)
    s.CalcF()
    e0 = s.GetPotentialEnergy() + itg.GetKineticEnergy()
    s.RunSteps(0, 1000, 10, 0)
    s.CalcF()
    e1 = s.GetPotentialEnergy() + itg.GetKineticEnergy()
    assert abs(e1 - e0) / c.n < 2e-3


def test_simulation_driver_and_conf_round_trip(tmp_path, golden_dir, monkeypatch):
    monkeypatch.chdir(tmp_path)  # the timing file goes to the working directory
    sim = Simulation(golden_dir / "KA_N256_L6.0_T0.50.xyz.gz", verbose=False)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    sim.SetPotential(pot)
    itg = rumd.IntegratorNVT(timeStep=0.005, targetTemperature=0.5, thermostatRelaxationTime=0.2)
    sim.SetIntegrator(itg)
    sim.Run(200, suppressAllOutput=True)
    assert sim.last_mats > 0
    out = tmp_path / "end.xyz.gz"
    sim.WriteConf(out)
    d = conf_io.read_conf(out)
    np.testing.assert_allclose(d.positions, sim.sample.GetPositions(), atol=1e-8)
    np.testing.assert_allclose(d.forces, sim.sample.GetForces(), atol=1e-8)
    assert d.meta["integrator"].startswith("IntegratorNVT")
    # total momentum is removed at step 0 and every 100 steps
    p = np.asarray(itg.GetTotalMomentum())
    assert np.abs(p[:3]).max() < 1e-2


def test_overflow_grows_list_and_stays_exact():
    sysm = configs.ka_random_lattice(12000, density=1.2, seed=8)
    conf = conf_from_system(sysm)
    s, pot = ka_sample(conf, cutoff=rumd.ShiftedForce)
    pot.SetNbMaxNumNbrs(30)
    s.CalcF()
    f = s.GetForces()  # synchronises; overflow is handled here
    f_ref, _ = oracle.forces_cells(configs.pack_pos4(sysm.positions, sysm.types), sysm.box, 2, ka_table(), SF)
    assert force_errors(f, f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
    s.CalcF()
    f2 = s.GetForces()
    assert force_errors(f2, f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
    assert pot.GetActualMaxNumNbrs() > 30


def test_halo_capacity_follows_the_largest_halo():
    """The tile halo capacity grows when a halo exceeds it (those tiles are
    computed from global memory meanwhile) and shrinks when it is well above
    the largest halo; forces stay exact throughout."""
    sysm = configs.ka_random_lattice(40000, density=1.2, seed=21)
    conf = conf_from_system(sysm)
    s, pot = ka_sample(conf, cutoff=rumd.ShiftedForce, method="sort")
    f_ref, _ = oracle.forces_cells(configs.pack_pos4(sysm.positions, sysm.types), sysm.box, 2, ka_table(), SF)
    for cap in (64, 3072):
        print(f"[case] halo capacity {cap}", flush=True)
        pot.SetNbHaloCapacity(cap)
        s.CalcF()
        assert force_errors(s.GetForces(), f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
        adapted = pot.GetNbHaloCapacity()
        s.CalcF()
        assert force_errors(s.GetForces(), f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
        assert adapted == pot.GetNbHaloCapacity()  # settled after one adaptation
        if cap == 64:
            assert adapted > 64
            grown = adapted
        else:
            assert adapted == grown < cap


def test_error_paths():
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    with pytest.raises(RuntimeError):
        s.CalcF()  # no box
    sim_conf = conf_from_system(configs.ka_random_lattice(1000, seed=1))
    s.SetConfiguration(sim_conf)
    with pytest.raises(RuntimeError):
        s.RunSteps(0, 10, 10, 0)  # no integrator / potential
    with pytest.raises(ValueError):
        s.SetPositions(np.zeros((10, 3), np.float32))


def test_autotune_tunes_reuses_and_keeps_the_trajectory(tmp_path, monkeypatch):
    """Autotune times short runs over tp, list method and skin, saves the
    choice, reuses it for a matching simulation, and leaves the state as it
    was: the run after tuning equals a fresh run with those parameters."""
    from rumd_amd.Autotune import Autotune

    monkeypatch.chdir(tmp_path)  # the timing file goes to the working directory
    sysm = configs.ka_random_lattice(2000, density=1.2, temperature=0.45, seed=77)

    def make_sim():
        sim = Simulation(conf_from_system(sysm), verbose=False)
        pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
        pot.SetVerbose(False)
        for (i, j, sg, ep, rc) in configs.ka_params(2.5):
            pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
        pot.SetNbListSkin(0.3)
        sim.SetPotential(pot)
        sim.SetIntegrator(rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45,
                                             thermostatRelaxationTime=0.2))
        return sim, pot

    path = tmp_path / "autotune.json"
    sim, pot = make_sim()
    print("[case] autotune KA N=2000", flush=True)
    tech = Autotune(auto_filename=path, trial_seconds=0.02).Tune(sim)
    assert tech["tp"] in (1, 2, 4, 8) and tech["NB_method"] in ("none", "n2", "sort") and tech["skin"] > 0
    assert path.exists()
    assert sim.sample.GetThreadsPerParticle() == tech["tp"] and pot.GetNbListSkin() == pytest.approx(tech["skin"])
    at2 = Autotune(auto_filename=path)
    sim2, _ = make_sim()
    assert at2.Tune(sim2) == tech and at2.trials == []          # reused, nothing timed
    sim.Run(50)
    fresh, fpot = make_sim()
    fresh.SetPB_TP(tech["pb"], tech["tp"])
    fpot.SetNB_Method(tech["NB_method"])
    fpot.SetNbListSkin(tech["skin"])
    fresh.Run(50)
    np.testing.assert_array_equal(sim.sample.GetPositions(), fresh.sample.GetPositions())


def test_autotune_restores_fp64_positions_and_temperature_ramp(tmp_path, monkeypatch):
    """Tuning trials advance the state and may reallocate the particle arrays
    (SetPB_TP); afterwards the FP64 positions (hi + lo in mixed precision),
    velocities, Ps and a ramped target temperature are bit for bit what they
    were before Tune."""
    from rumd_amd.Autotune import Autotune

    monkeypatch.chdir(tmp_path)  # the run's output files go to the working directory
    sysm = configs.ka_random_lattice(3000, density=1.2, temperature=0.45, seed=78)
    sim = Simulation(conf_from_system(sysm), verbose=False, mixed_precision_positions=True, timing_filename=None)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    sim.SetPotential(pot)
    itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    sim.SetIntegrator(itg)
    itg.SetTemperatureChangePerTimeStep(1e-4)
    sim.Run(40)  # the low-order position parts and Ps are non-trivial now
    s = sim.sample
    pos64, vel = s.GetPositionsFP64(), s.GetVelocities()
    T0, ps0 = itg.GetTargetTemperature(), itg.GetThermostatState()
    assert np.abs(pos64 - s.GetPositions().astype(np.float64)).max() > 0.0 and ps0 != 0.0
    print("[case] autotune restore, mixed KA N=3000 with a T ramp", flush=True)
    Autotune(auto_filename=tmp_path / "autotune.json", trial_seconds=0.02).Tune(sim)
    np.testing.assert_array_equal(s.GetPositionsFP64(), pos64)
    np.testing.assert_array_equal(s.GetVelocities(), vel)
    assert itg.GetTargetTemperature() == T0
    assert itg.GetThermostatState() == ps0


def test_run_segments_start_at_the_first_step_and_end_at_resets():
    """RunSteps' blocks start at the run's first step and end early at each
    momentum-reset step (absolute multiples of the interval); PrepareRunSteps
    instantiates their graphs without running a step."""
    sysm = configs.ka_random_lattice(2000, density=1.2, temperature=0.45, seed=79)
    s, pot = ka_sample(conf_from_system(sysm), cutoff=rumd.ShiftedForce)
    s.SetFusedSteps(False)  # compared below with the CalcF/Integrate loop
    itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    s.SetIntegrator(itg)
    assert s.RunSegments(5, 20, 10, 100) == [(10, False), (10, False)]
    assert s.RunSegments(95, 20, 10, 100) == [(5, False), (10, True), (5, False)]
    assert s.RunSegments(0, 5, 10, 100) == [(5, True)]
    assert s.RunSegments(7, 3, 10, 0) == [(3, False)]
    assert s.RunSegments(99, 2, 10, 100) == [(1, False), (1, True)]
    s.CalcF()
    r0 = s.GetPositions()
    assert s.PrepareRunSteps(95, 20, 10, 100) == 2  # one graph per (length, reset) pair
    assert s.PrepareRunSteps(95, 20, 10, 100) == 0
    np.testing.assert_array_equal(s.GetPositions(), r0)  # nothing ran
    s.RunSteps(95, 20, 10, 100)
    ref, rpot = ka_sample(conf_from_system(sysm), cutoff=rumd.ShiftedForce)
    ritg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    ref.SetIntegrator(ritg)
    ref.CalcF()
    for i in range(95, 115):
        if i % 100 == 0:
            ritg.SetMomentumToZero()
        ref.CalcF()
        ritg.Integrate()
    np.testing.assert_array_equal(s.GetPositions(), ref.GetPositions())
    np.testing.assert_array_equal(s.GetVelocities(), ref.GetVelocities())


@pytest.mark.parametrize("mixed", [False, True])
def test_tile_resort_trajectory_equals_full_sort_trajectory(mixed):
    """Skin-triggered rebuilds take the tile re-sort inside captured blocks;
    since it reproduces the full sort's particle order bit for bit, 300 NVT
    steps give the same trajectory as with a full sort at every rebuild."""
    sysm = configs.ka_random_lattice(40000, density=1.2, temperature=0.8, seed=33)
    out = []
    for full in (False, True):
        s, pot = ka_sample(conf_from_system(sysm), cutoff=rumd.ShiftedForce, mixed=mixed)
        itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.8, thermostatRelaxationTime=0.2)
        s.SetIntegrator(itg)
        pot.SetNbForceFullSort(full)
        print(f"[case] 300 steps, full sort at every rebuild: {full}, mixed={mixed}", flush=True)
        s.RunSteps(0, 300, 10, 100)
        out.append((s.GetPositions(), s.GetVelocities(), s.GetImages(), itg.GetThermostatState(),
                    pot.GetNumberOfNbListBuilds()))
    (r0, v0, i0, p0, b0), (r1, v1, i1, p1, b1) = out
    assert b0 == b1 and b0 >= 4
    np.testing.assert_array_equal(r0, r1)
    np.testing.assert_array_equal(v0, v1)
    np.testing.assert_array_equal(i0, i1)
    assert p0 == p1
