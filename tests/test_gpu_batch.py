"""Batched replicas (ReplicaBatch over rumdx_batch_*): per-replica forces
against the FP64 oracle, neighbour sets bit-exact against the FP32
restatement, each replica independent of the others bit for bit, the
per-replica thermostat update, and exact forces through a list overflow."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, sorted_rows
from rumd_amd import configs

pytestmark = pytest.mark.gpu

SF = 2


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _rumd():
    import rumd_amd.rumd as rumd
    return rumd


def make_batch(systems, temps, tau=0.2, skin=0.3, dt=0.0025, max_nbrs=None):
    rumd = _rumd()
    b = rumd.ReplicaBatch(2)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    for s in systems:
        b.AddReplica(s.positions.astype(np.float32), s.types.astype(np.int32), s.velocities.astype(np.float32),
                     float(s.box[0]))
    b.SetPotential(pot, skin)
    b.SetTimeStep(dt)
    for r, t in enumerate(temps):
        b.SetThermostat(r, t, tau)
    if max_nbrs is not None:
        b.SetMaxNumNbrs(max_nbrs)
    return b


def _systems(sizes, seed0=500):
    return [configs.ka_random_lattice(n, density=1.2, temperature=0.45 + 0.3 * k, seed=seed0 + k)
            for k, n in enumerate(sizes)]


def test_batch_forces_and_lists_per_replica():
    sizes = [256, 700, 1500, 4000, 8192]
    systems = _systems(sizes)
    b = make_batch(systems, [0.5] * len(sizes))
    print(f"[case] batch forces sizes={sizes}", flush=True)
    b.CalcF()
    e = b.GetEnergies()
    cut2 = configs.cut_skin2_table(ka_table(), 0.3)
    for r, s in enumerate(systems):
        pos4 = configs.pack_pos4(s.positions, s.types)
        f = b.GetForcesAndEnergies(r)
        f_ref, vir_ref = (oracle.forces_brute if s.n <= 2000 else oracle.forces_cells)(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(f[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)
        assert e[r, 0] == pytest.approx(float(f_ref[:, 3].sum()), rel=1e-5, abs=1e-3)
        assert e[r, 1] == pytest.approx(float(vir_ref.sum()) / 6.0, rel=1e-5, abs=1e-3)
        b8 = configs.box8(s.box)
        off, ids = b.GetNeighborListCSR(r)
        o2, i2 = oracle.nblist_cells_f32(pos4, b8, 2, cut2, oracle.cell_grid(b8, 2.5, 0.3))
        assert np.array_equal(off, o2), r
        for i in range(s.n):
            assert set(ids[off[i]:off[i + 1]].tolist()) == set(i2[o2[i]:o2[i + 1]].tolist()), (r, i)


def test_batch_first_build_on_reused_memory():
    """The first build's counters start at zero even when the allocations
    reuse freed device memory that held other data (the batch runs on a
    non-blocking stream, so the zero fill must complete before it)."""
    import torch
    junk = torch.full((1 << 26,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    systems = _systems([700, 3000], seed0=650)
    b = make_batch(systems, [0.5, 0.5])
    print("[case] batch first build on reused memory", flush=True)
    b.CalcF()
    for r, s in enumerate(systems):
        pos4 = configs.pack_pos4(s.positions, s.types)
        f_ref, _ = oracle.forces_brute(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(b.GetForcesAndEnergies(r)[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)


def test_batch_replicas_are_independent():
    """A replica run inside a batch equals the same replica run alone, bit for
    bit (graph-captured blocks, rebuilds included)."""
    sizes = [300, 1200, 5000]
    temps = [0.45, 1.5, 0.8]
    systems = _systems(sizes, seed0=900)
    b = make_batch(systems, temps)
    print("[case] batch of 3, 60 steps", flush=True)
    b.RunSteps(60, 10)
    builds = b.GetNumberOfBuilds()
    assert all(k >= 2 for k in builds), builds
    for r, s in enumerate(systems):
        alone = make_batch([s], [temps[r]])
        alone.RunSteps(60, 10)
        print(f"[case] replica {r} alone", flush=True)
        assert np.array_equal(alone.GetPositions(0), b.GetPositions(r)), r
        assert np.array_equal(alone.GetVelocities(0), b.GetVelocities(r)), r
        assert alone.GetThermostatStates()[0] == b.GetThermostatStates()[r]


def test_batch_thermostat_update():
    """One step: Ps = dt omega^2 (sum m v^2 / (dof T) - 1), each replica with
    its own T and dof."""
    sizes = [400, 2000]
    temps = [0.6, 1.8]
    systems = _systems(sizes, seed0=300)
    dt, tau = 0.0025, 0.2
    b = make_batch(systems, temps, tau=tau, dt=dt)
    b.RunSteps(1, 1)
    e = b.GetEnergies()
    ps = b.GetThermostatStates()
    omega2 = float(np.float32(4 * np.pi ** 2 / tau ** 2))
    for r, s in enumerate(systems):
        dof = 3 * s.n - 3
        expect = dt * omega2 * (e[r, 3] / (dof * temps[r]) - 1.0)
        assert ps[r] == pytest.approx(expect, rel=1e-5), r


def test_batch_overflowed_lists_still_exact():
    systems = _systems([1000, 3000], seed0=700)
    b = make_batch(systems, [0.5, 0.5], max_nbrs=24)
    print("[case] batch overflow max_nbrs=24", flush=True)
    b.CalcF()
    for r, s in enumerate(systems):
        pos4 = configs.pack_pos4(s.positions, s.types)
        f_ref, _ = oracle.forces_brute(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(b.GetForcesAndEnergies(r)[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)
    b.Synchronize()
    assert b.GetMaxNumNbrs() > 24
    b.CalcF()
    f_ref, _ = oracle.forces_brute(configs.pack_pos4(systems[1].positions, systems[1].types), systems[1].box, 2,
                                   ka_table(), SF)
    err = force_errors(b.GetForcesAndEnergies(1)[:, :3], f_ref[:, :3])
    assert err["guarded_max"] <= FORCE_TOL_GUARDED, err


def test_batch_hot_replicas_rebuild_more_often():
    systems = _systems([2000, 2000], seed0=100)
    b = make_batch(systems, [0.4, 2.0])
    print("[case] batch hot/cold rebuild counts", flush=True)
    b.RunSteps(200, 10)
    builds = b.GetNumberOfBuilds()
    e = b.GetEnergies()
    t_kin = [2.0 * e[r, 2] / (3 * s.n - 3) for r, s in enumerate(systems)]
    assert t_kin[1] > t_kin[0]
    assert builds[1] > builds[0], builds


@pytest.mark.parametrize("tp", [2, 4, 8])
def test_batch_lanes_per_particle(tp):
    """Lists spread over 2, 4 or 8 lanes (a fixed butterfly combines them)
    give the oracle's forces; an overflowed list is rescanned by one lane."""
    systems = _systems([600, 3000], seed0=830)
    b = make_batch(systems, [0.5, 0.5])
    b.SetLanesPerParticle(tp)
    print(f"[case] batch lanes per particle {tp}", flush=True)
    b.CalcF()
    for r, s in enumerate(systems):
        pos4 = configs.pack_pos4(s.positions, s.types)
        f_ref, _ = oracle.forces_brute(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(b.GetForcesAndEnergies(r)[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)
    o = make_batch(systems[:1], [0.5], max_nbrs=24)
    o.SetLanesPerParticle(tp)
    o.CalcF()
    f_ref, _ = oracle.forces_brute(configs.pack_pos4(systems[0].positions, systems[0].types), systems[0].box, 2,
                                   ka_table(), SF)
    assert force_errors(o.GetForcesAndEnergies(0)[:, :3], f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED


@pytest.mark.parametrize("tp", [1, 4])
def test_batch_staged_forces_equal_gathered(tp):
    """The replica-staged force kernel (each replica copied whole into LDS,
    work items of at most 2048 slots, largest replicas first) gives the gather
    kernel's forces bit for bit, over 40 steps with rebuilds; also with an
    overflowed list."""
    sizes = [256, 1500, 2049, 5000, 8192]
    systems = _systems(sizes, seed0=870)
    temps = [0.5, 1.0, 2.0, 0.6, 0.9]
    out = []
    for staged in (False, True):
        b = make_batch(systems, temps)
        b.SetLanesPerParticle(tp)
        b.SetStagedForces(staged)
        print(f"[case] batch staged={staged} lanes={tp}", flush=True)
        b.RunSteps(40, 10)
        assert b.UsesStagedForces() == staged
        out.append([(b.GetPositions(r), b.GetForcesAndEnergies(r)) for r in range(len(sizes))])
    for r in range(len(sizes)):
        np.testing.assert_array_equal(out[0][r][0], out[1][r][0], err_msg=f"positions {r}")
        np.testing.assert_array_equal(out[0][r][1], out[1][r][1], err_msg=f"forces {r}")
    res = []
    for staged in (False, True):
        o = make_batch(systems[1:3], temps[1:3], max_nbrs=24)
        o.SetLanesPerParticle(tp)
        o.SetStagedForces(staged)
        o.CalcF()
        res.append([o.GetForcesAndEnergies(r) for r in range(2)])
    for r in range(2):
        np.testing.assert_array_equal(res[0][r], res[1][r])
    f_ref, _ = oracle.forces_brute(configs.pack_pos4(systems[1].positions, systems[1].types), systems[1].box, 2,
                                   ka_table(), SF)
    assert force_errors(res[1][0][:, :3], f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED


@pytest.mark.parametrize("tp", [1, 2])
def test_batch_fused_steps_equal_two_launches(tp):
    """Forces and the NVT update in one launch per step (whole-replica work
    items, rumdx_batch_forces_nvt) against the staged forces and the separate
    integrator (items of at most 2048 slots): the same bits for positions,
    velocities, forces and thermostat states over 60 steps with rebuilds,
    thermostatted and NVE replicas mixed."""
    sizes = [256, 1500, 2049, 5000, 8192]
    systems = _systems(sizes, seed0=950)
    temps = [2.0, 0.5, 1.0, 0.6, 0.9]
    out = []
    for fused in (False, True):
        b = make_batch(systems, temps)
        b.SetThermostat(2, temps[2], 0.0)  # NVE
        b.SetLanesPerParticle(tp)
        b.SetFusedSteps(fused)
        print(f"[case] batch fused steps={fused} lanes={tp}", flush=True)
        b.RunSteps(60, 10)
        b.RunSteps(3, 1)  # eager steps too
        assert b.UsesFusedSteps() == fused and b.UsesStagedForces()
        out.append(([b.GetPositions(r) for r in range(len(sizes))], [b.GetVelocities(r) for r in range(len(sizes))],
                    [b.GetForcesAndEnergies(r) for r in range(len(sizes))], b.GetThermostatStates(),
                    b.GetEnergies()))
    for k, name in enumerate(("positions", "velocities", "forces")):
        for r in range(len(sizes)):
            np.testing.assert_array_equal(out[0][k][r], out[1][k][r], err_msg=f"{name} {r}")
    np.testing.assert_array_equal(np.asarray(out[0][3]), np.asarray(out[1][3]), err_msg="thermostat states")
    np.testing.assert_array_equal(np.asarray(out[0][4]), np.asarray(out[1][4]), err_msg="energies")


def test_batch_workgroup_sort_equals_batch_sort():
    """Rebuild sorts by one workgroup per replica in LDS give the four-kernel
    sort's particle order, cell starts and lists: 60 steps with rebuilds of
    hot and cold replicas produce the same bits."""
    sizes = [256, 1500, 4100, 8192]
    systems = _systems(sizes, seed0=910)
    temps = [2.0, 0.5, 1.2, 0.8]
    out = []
    for wg in (False, True):
        b = make_batch(systems, temps)
        b.SetWorkgroupSort(wg)
        print(f"[case] batch workgroup sort={wg}", flush=True)
        b.RunSteps(60, 10)
        assert b.UsesWorkgroupSort() == wg
        st = []
        for r in range(len(sizes)):
            off, ids = b.GetNeighborListCSR(r)
            st.append((b.GetPositions(r), b.GetVelocities(r), b.GetForcesAndEnergies(r), off, ids))
        out.append(st)
    for r in range(len(sizes)):
        for k, name in enumerate(("positions", "velocities", "forces", "list offsets", "list ids")):
            np.testing.assert_array_equal(out[0][r][k], out[1][r][k], err_msg=f"{name} {r}")


def test_batch_save_restore_leaves_no_trace():
    """Trial steps between SaveState and RestoreState do not change the
    trajectory: the run continues bit for bit as if they never happened."""
    systems = _systems([500, 2500], seed0=870)
    a = make_batch(systems, [0.6, 1.2])
    a.RunSteps(40, 10)
    b = make_batch(systems, [0.6, 1.2])
    b.SaveState()
    b.SetLanesPerParticle(4)
    b.RunSteps(30, 10)
    b.RestoreState()
    b.SetLanesPerParticle(1)
    print("[case] batch save/restore", flush=True)
    b.RunSteps(40, 10)
    for r in range(2):
        assert np.array_equal(a.GetPositions(r), b.GetPositions(r)), r
        assert np.array_equal(a.GetVelocities(r), b.GetVelocities(r)), r
    assert a.GetThermostatStates() == b.GetThermostatStates()


def test_batch_autotune(tmp_path):
    from rumd_amd.Autotune import Autotune
    systems = _systems([400, 1600, 3000], seed0=910)
    b = make_batch(systems, [0.5, 0.9, 1.4])
    at = Autotune(auto_filename=tmp_path / "autotune.json", trial_seconds=0.02)
    print("[case] batch autotune", flush=True)
    tech = at.TuneBatch(b)
    assert tech["tp"] in (1, 2, 4, 8) and len(at.trials) == 4
    assert b.GetLanesPerParticle() == tech["tp"]
    # state untouched: the same run with the chosen lanes set directly
    ref = make_batch(systems, [0.5, 0.9, 1.4])
    ref.SetLanesPerParticle(tech["tp"])
    ref.RunSteps(20, 10)
    b.RunSteps(20, 10)
    for r in range(3):
        assert np.array_equal(ref.GetPositions(r), b.GetPositions(r)), r
    again = Autotune(auto_filename=tmp_path / "autotune.json")
    assert again.TuneBatch(b) == tech and again.trials == []


def test_batch_build_per_lane_path():
    """Lists longer than the build's LDS staging (capacity 300 here) take the
    one-lane-per-particle build; sets and forces still match the oracle."""
    sizes = [900, 2600]
    systems = _systems(sizes, seed0=960)
    b = make_batch(systems, [0.5, 0.5], max_nbrs=300)
    print("[case] batch build per-lane path", flush=True)
    b.CalcF()
    cut2 = configs.cut_skin2_table(ka_table(), 0.3)
    for r, s in enumerate(systems):
        pos4 = configs.pack_pos4(s.positions, s.types)
        f_ref, _ = oracle.forces_brute(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(b.GetForcesAndEnergies(r)[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)
        b8 = configs.box8(s.box)
        off, ids = b.GetNeighborListCSR(r)
        o2, i2 = oracle.nblist_cells_f32(pos4, b8, 2, cut2, oracle.cell_grid(b8, 2.5, 0.3))
        assert np.array_equal(off, o2), r
        for i in range(s.n):
            assert set(ids[off[i]:off[i + 1]].tolist()) == set(i2[o2[i]:o2[i + 1]].tolist()), (r, i)


def test_batch_full_replica_set_against_oracle():
    """The benchmark's own replica set (512 KA replicas, N ~ U[256, 8192],
    T ~ U[0.4, 2.0], 2.15 M particles) after 100 NVT steps in graph blocks
    (per-replica rebuilds of hot and cold replicas): the forces and energies
    of every replica, and the neighbour sets of a sample -- the smallest, the
    largest, the hottest, the coldest and every 64th -- against the oracle at
    their current positions."""
    systems, temps = configs.ka_replica_set()
    assert len(systems) == 512
    b = make_batch(systems, temps)
    print(f"[case] 512-replica set, {sum(s.n for s in systems)} particles: 100 steps", flush=True)
    b.RunSteps(100, 10)
    b.CalcF()
    ns = np.array([s.n for s in systems])
    sample = sorted({int(np.argmin(ns)), int(np.argmax(ns)), int(np.argmax(temps)), int(np.argmin(temps)),
                     *range(0, 512, 64)})
    builds = np.asarray(b.GetNumberOfBuilds())
    # hot replicas rebuild more often than cold ones
    assert builds[int(np.argmax(temps))] > builds[int(np.argmin(temps))]
    e = b.GetEnergies()
    cut2 = configs.cut_skin2_table(ka_table(), 0.3)
    for r in range(len(systems)):
        s = systems[r]
        pos = b.GetPositions(r)
        pos4 = configs.pack_pos4(pos, s.types)
        if r in sample:
            print(f"[case] replica {r}: N={s.n} T={temps[r]:.2f} builds={builds[r]}", flush=True)
        f = b.GetForcesAndEnergies(r)
        f_ref, vir_ref = (oracle.forces_brute if s.n <= 2000 else oracle.forces_cells)(pos4, s.box, 2, ka_table(), SF)
        err = force_errors(f[:, :3], f_ref[:, :3])
        assert err["guarded_max"] <= FORCE_TOL_GUARDED, (r, err)
        assert e[r, 0] == pytest.approx(float(f_ref[:, 3].sum()), rel=1e-5, abs=1e-3)
    # every list stale (RestoreState), rebuilt at the current positions: the
    # sampled replicas' sets are bit-exact against the FP32 restatement
    b.SaveState()
    b.RestoreState()
    b.CalcF()
    for r in sample:
        s = systems[r]
        pos4 = configs.pack_pos4(b.GetPositions(r), s.types)
        b8 = configs.box8(s.box)
        off, ids = b.GetNeighborListCSR(r)
        o2, i2 = oracle.nblist_cells_f32(pos4, b8, 2, cut2, oracle.cell_grid(b8, 2.5, 0.3))
        assert np.array_equal(off, o2), r
        assert np.array_equal(sorted_rows(off, ids), sorted_rows(o2, i2)), r
