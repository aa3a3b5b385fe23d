"""Intramolecular forces on the GPU (rumdx_bond_harmonic, rumdx_angle_cossq,
rumdx_angle_sq through BondHarmonic / AngleCosSq / AngleSq and MoleculeData) together with a pair
potential, against the FP64 oracle, on RUMD's pentamer melt fixture: every
list method (the sorted one permutes particles, so the bonded kernels read
them through the slot index), energies per potential, the virial, graph
capture, and energy conservation."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, load_golden
from rumd_amd import configs, rumd

pytestmark = pytest.mark.gpu

SP = 1
BOND_PARAMS = [(1.0, 100.0), (1.1, 150.0)]
ANGLE_PARAMS = [(2.0, 20.0), (1.8, 30.0)]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _topology(golden_dir):
    bonds = []
    for ln in open(golden_dir / "pentamer200_rho0.85_T0.5.top"):
        parts = ln.split()
        if ln.startswith("["):
            section = ln.strip()[1:-1].strip()
            continue
        if section == "bonds" and len(parts) == 4 and parts[0].isdigit():
            bonds.append([int(parts[1]), int(parts[2]), int(parts[3])])
    bonds = np.array(bonds, np.int32)
    # angles (i, i+1, i+2) inside each five-bead chain, type by chain parity
    angles = np.array([[5 * m + i, 5 * m + i + 1, 5 * m + i + 2, m % 2] for m in range(200) for i in range(3)], np.int32)
    return bonds, angles


def _sample(golden_dir, method, with_angles=True, exclude_bonds=False, angle_cls=None):
    c = load_golden(golden_dir / "pentamer200_rho0.85_T0.5.xyz.gz")
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(c)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.5)
    pot.SetNbListSkin(0.3)
    pot.SetNB_Method(method)
    s.SetPotential(pot)
    bonds, angles = _topology(golden_dir)
    s.ReadMoleculeData(str(golden_dir / "pentamer200_rho0.85_T0.5.top"))
    md = s.GetMoleculeData()
    assert md.GetNumberOfBonds() == len(bonds) and md.GetNumberOfMolecules() == 200
    bond = rumd.BondHarmonic()
    bond.SetVerbose(False)
    for t, (l0, k) in enumerate(BOND_PARAMS):
        bond.SetParams(bond_type=t, bond_length=l0, stiffness=k, exclude=exclude_bonds)
    s.AddPotential(bond)
    ang = None
    if with_angles:
        md.SetAngles(angles.reshape(-1).tolist())
        ang = (angle_cls or rumd.AngleCosSq)()
        ang.SetVerbose(False)
        for t, (th0, k) in enumerate(ANGLE_PARAMS):
            ang.SetParams(angle_type=t, theta0=th0, ktheta=k)
        s.AddPotential(ang)
    return c, s, pot, bond, ang, bonds, angles


@pytest.mark.parametrize("angle_form", ["cossq", "sq"])
@pytest.mark.parametrize("method", ["n2", "sort", "none"])
def test_bonded_with_pair_forces_match_oracle(golden_dir, method, angle_form):
    sq = angle_form == "sq"
    c, s, pot, bond, ang, bonds, angles = _sample(golden_dir, method, angle_cls=rumd.AngleSq if sq else None)
    print(f"[case] pentamer200 LJ + bonds + angles ({angle_form}), list {method}", flush=True)
    s.CalcF()
    f = s.GetForcesAndEnergies().astype(np.float64)
    pos4 = configs.pack_pos4(c.positions, c.types)
    table = configs.lj_param_table([(0, 0, 1.0, 1.0, 2.5)], 1)
    f_lj, vir_lj = oracle.forces_brute(pos4, c.box, 1, table, SP)
    f_b, vir_b = oracle.bonded(pos4, c.box, bonds, BOND_PARAMS, None, None)
    f_a, _ = oracle.bonded(pos4, c.box, None, None, angles, ANGLE_PARAMS, angle_sq=sq)
    f_ref = f_lj + f_b + f_a
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    np.testing.assert_allclose(f[:, 3], f_ref[:, 3], rtol=0, atol=5e-5)
    # each potential's own energy, and the total virial (angles add none)
    assert bond.GetPotentialEnergy() == pytest.approx(f_b[:, 3].sum(), rel=2e-6)
    assert ang.GetPotentialEnergy() == pytest.approx(f_a[:, 3].sum(), rel=2e-6)
    assert s.GetPotentialEnergy() == pytest.approx(f_ref[:, 3].sum(), rel=2e-6)
    assert s.GetVirial() == pytest.approx((vir_lj.sum() + vir_b.sum()) / 6.0, rel=1e-5)


def test_bonded_stress_and_initialize_order(golden_dir):
    """The pair potential runs first (its first list build sorts the
    particles), then angles and bonds add: the per-particle stress trace is
    minus the virial with the bonds included."""
    c, s, pot, bond, ang, bonds, angles = _sample(golden_dir, "sort")
    print("[case] pentamer200 stresses", flush=True)
    s.CalcF(True)
    w = s.GetVirials().astype(np.float64)
    sts = s.GetStresses().astype(np.float64)
    np.testing.assert_allclose(sts[:, 0] + sts[:, 1] + sts[:, 2], -w[:, 3], rtol=0, atol=2e-3)


def test_bonded_graph_blocks_equal_eager_and_conserve_energy(golden_dir):
    states = []
    for graph in (True, False):
        c, s, pot, bond, ang, bonds, angles = _sample(golden_dir, "sort")
        itg = rumd.IntegratorNVE(timeStep=0.002)
        s.SetIntegrator(itg)
        print(f"[case] pentamer200 NVE graph={graph}", flush=True)
        if graph:
            s.RunSteps(0, 200, 10, 0)
            assert not s.UsesFusedSteps()  # more than one force term: the two-kernel step
        else:
            for _ in range(200):
                s.CalcF()
                itg.Integrate()
        s.Synchronize()
        states.append((s.GetPositions(), s.GetVelocities()))
    np.testing.assert_array_equal(states[0][0], states[1][0])
    np.testing.assert_array_equal(states[0][1], states[1][1])
    # NVE drift over 2000 further steps
    c, s, pot, bond, ang, bonds, angles = _sample(golden_dir, "sort")
    itg = rumd.IntegratorNVE(timeStep=0.002)
    s.SetIntegrator(itg)
    s.CalcF()
    e0 = s.GetPotentialEnergy() + itg.GetKineticEnergy()
    s.RunSteps(0, 2000, 10, 0)
    s.CalcF()
    e1 = s.GetPotentialEnergy() + itg.GetKineticEnergy()
    print(f"  drift per particle {(e1 - e0) / c.n:.2e}", flush=True)
    assert abs(e1 - e0) / c.n < 2e-3


def _excluded(bonds, angles):
    ex = {(min(a, b), max(a, b)) for a, b, _ in bonds}
    ex |= {(min(a, c), max(a, c)) for a, _, c, _ in angles}
    return ex


def _oracle_lj_excluding(c, pos4, table, excl):
    """FP64 LJ over the rc + skin pairs minus the excluded ones (the oracle's
    list-driven sum), and those lists as sets."""
    b8 = configs.box8(c.box)
    off, idx = oracle.nblist_brute_f32(pos4, b8, 1, configs.cut_skin2_table(table, 0.3))
    sets = []
    off2, idx2 = [0], []
    for i in range(c.n):
        keep = [int(j) for j in idx[off[i]:off[i + 1]] if (min(i, j), max(i, j)) not in excl]
        sets.append(set(keep))
        idx2.extend(keep)
        off2.append(len(idx2))
    f, vir = oracle.forces_nl(pos4, c.box, 1, table, SP, np.array(off2, np.int64), np.array(idx2, np.int32))
    return f, vir, sets


@pytest.mark.parametrize("method,variant", [("n2", "plain"), ("sort", "plain"), ("none", "plain"),
                                            ("sort", "overflow"), ("sort", "small_halo"), ("n2", "overflow")])
def test_exclusions_leave_bonded_pairs_out(golden_dir, method, variant):
    """Bond pairs (exclude=True) and angle end pairs (excluded by default)
    leave the LJ lists: N2 build filter, sorted-list compaction, and the cell
    rescans of overflowed lists and unstaged tiles."""
    c, s, pot, bond, ang, bonds, angles = _sample(golden_dir, method, exclude_bonds=True)
    bond.SetExclusions(pot)
    ang.SetExclusions(pot)
    excl = _excluded(bonds, angles)
    assert pot.GetNumberOfExclusions() == len(excl) == 1400
    if variant == "overflow":
        pot.SetNbMaxNumNbrs(30)
    elif variant == "small_halo":
        pot.SetNbHaloCapacity(64)
    print(f"[case] exclusions list={method} {variant}", flush=True)
    s.CalcF()
    f = s.GetForcesAndEnergies().astype(np.float64)
    pos4 = configs.pack_pos4(c.positions, c.types)
    table = configs.lj_param_table([(0, 0, 1.0, 1.0, 2.5)], 1)
    f_lj, vir_lj, sets = _oracle_lj_excluding(c, pos4, table, excl)
    f_b, vir_b = oracle.bonded(pos4, c.box, bonds, BOND_PARAMS, None, None)
    f_a, _ = oracle.bonded(pos4, c.box, None, None, angles, ANGLE_PARAMS)
    f_ref = f_lj + f_b + f_a
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    np.testing.assert_allclose(f[:, 3], f_ref[:, 3], rtol=0, atol=5e-5)
    assert s.GetVirial() == pytest.approx((vir_lj.sum() + vir_b.sum()) / 6.0, rel=1e-5)
    if variant == "plain":
        off, ids = pot.GetNeighborListCSR(s)
        got = [set(ids[off[i]:off[i + 1]].tolist()) for i in range(c.n)]
        assert got == sets


def test_simulation_sets_exclusions_and_runs(golden_dir, tmp_path, monkeypatch):
    """Simulation.AddPotential wires the molecular potentials' exclusions into
    the pair potential in either order; a short NVT run stays finite."""
    from rumd_amd.Simulation import Simulation
    monkeypatch.chdir(tmp_path)
    sim = Simulation(str(golden_dir / "pentamer200_rho0.85_T0.5.xyz.gz"), verbose=False)
    sim.ReadMoleculeData(golden_dir / "pentamer200_rho0.85_T0.5.top")
    bond = rumd.BondHarmonic()
    bond.SetVerbose(False)
    bond.SetParams(bond_type=0, bond_length=1.0, stiffness=100.0, exclude=True)
    bond.SetParams(bond_type=1, bond_length=1.1, stiffness=150.0, exclude=True)
    sim.AddPotential(bond)  # before the pair potential
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.5)
    pot.SetNbListSkin(0.3)
    sim.AddPotential(pot)
    ang = rumd.AngleCosSq()
    ang.SetVerbose(False)
    ang.SetParams(angle_type=0, theta0=2.0, ktheta=20.0)
    ang.SetParams(angle_type=1, theta0=1.8, ktheta=30.0)
    sim.AddPotential(ang)  # after it; the topology file holds 9 angles
    assert pot.GetNumberOfExclusions() == 800 + 9
    with pytest.raises(RuntimeError):
        sim.SetPotential(bond)
    sim.SetIntegrator(rumd.IntegratorNVT(timeStep=0.002, targetTemperature=0.5, thermostatRelaxationTime=0.2))
    print("[case] Simulation pentamer200 with exclusions", flush=True)
    sim.Run(300, suppressAllOutput=True)
    assert np.isfinite(sim.sample.GetPositions()).all()
    assert np.isfinite(sim.sample.GetPotentialEnergy())


DIH_PARAMS = [(0.9, 2.1, -0.4, -3.0, 0.2, 0.1), (1.2, -1.0, 0.6, 0.9, -0.5, 0.3)]
PERIODIC_PARAMS = [(0.4, 2.0, -0.7, 0.9), (1.0, 1.5, 0.5, -0.4)]  # psi0, k1, k2, k3


def _dihedrals():
    return np.array([[5 * m + i, 5 * m + i + 1, 5 * m + i + 2, 5 * m + i + 3, (m + i) % 2]
                     for m in range(200) for i in range(2)], np.int32)


@pytest.mark.parametrize("dihedral_form", ["ryckaert", "periodic"])
@pytest.mark.parametrize("method", ["n2", "sort"])
def test_fene_and_harmonic_types_and_dihedrals(golden_dir, method, dihedral_form):
    """Bond type 0 through BondFENE, type 1 through BondHarmonic (a type belongs
    to the potential that set it), Ryckaert dihedrals with their end pairs
    excluded from the LJ list, all against the oracle."""
    c = load_golden(golden_dir / "pentamer200_rho0.85_T0.5.xyz.gz")
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(c)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.5)
    pot.SetNbListSkin(0.3)
    pot.SetNB_Method(method)
    s.SetPotential(pot)
    s.ReadMoleculeData(str(golden_dir / "pentamer200_rho0.85_T0.5.top"))
    md = s.GetMoleculeData()
    dih = _dihedrals()
    md.SetDihedrals(dih.reshape(-1).tolist())
    fene = rumd.BondFENE()
    fene.SetVerbose(False)
    fene.SetParams(bond_type=0, bond_length=1.5, stiffness=30.0)
    harm = rumd.BondHarmonic()
    harm.SetVerbose(False)
    harm.SetParams(bond_type=1, bond_length=1.1, stiffness=150.0)
    periodic = dihedral_form == "periodic"
    dparams = PERIODIC_PARAMS if periodic else DIH_PARAMS
    rb = rumd.PeriodicDihedral() if periodic else rumd.DihedralRyckaert()
    rb.SetVerbose(False)
    for t, cf in enumerate(dparams):
        rb.SetParams(dihedral_type=t, coeffs=list(cf))
    for q in (fene, harm, rb):
        s.AddPotential(q)
    rb.SetExclusions(pot)
    assert pot.GetNumberOfExclusions() == len(dih)
    print(f"[case] FENE + harmonic + dihedrals ({dihedral_form}), list {method}", flush=True)
    s.CalcF()
    f = s.GetForcesAndEnergies().astype(np.float64)
    bonds, _ = _topology(golden_dir)
    pos4 = configs.pack_pos4(c.positions, c.types)
    table = configs.lj_param_table([(0, 0, 1.0, 1.0, 2.5)], 1)
    excl = {(min(a, d), max(a, d)) for a, _, _, d, _ in dih}
    f_lj, vir_lj, _ = _oracle_lj_excluding(c, pos4, table, excl)
    f_f, vir_f = oracle.bonded(pos4, c.box, bonds[bonds[:, 2] == 0], [(1.5, 30.0), (0.0, 0.0)], fene=True)
    f_h, vir_h = oracle.bonded(pos4, c.box, bonds[bonds[:, 2] == 1], [(0.0, 0.0), (1.1, 150.0)])
    f_d, _ = oracle.bonded(pos4, c.box, dihedrals=dih, dihedral_params=dparams, periodic_dihedrals=periodic)
    f_ref = f_lj + f_f + f_h + f_d
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    np.testing.assert_allclose(f[:, 3], f_ref[:, 3], rtol=0, atol=5e-5)
    assert fene.GetPotentialEnergy() == pytest.approx(f_f[:, 3].sum(), rel=2e-6)
    assert harm.GetPotentialEnergy() == pytest.approx(f_h[:, 3].sum(), rel=2e-6)
    assert rb.GetPotentialEnergy() == pytest.approx(f_d[:, 3].sum(), rel=2e-5, abs=1e-4)
    assert s.GetVirial() == pytest.approx((vir_lj.sum() + vir_f.sum() + vir_h.sum()) / 6.0, rel=1e-5)
    # graph-captured steps with every term equal the eager loop
    itg = rumd.IntegratorNVE(timeStep=0.002)
    s.SetIntegrator(itg)
    s.RunSteps(0, 100, 10, 0)
    s.Synchronize()
    assert np.isfinite(s.GetPositions()).all()


def test_polymer_workload_small_against_oracle():
    """The polymer workload's force field (WCA without 1-2 pairs, stiff
    harmonic bonds, cos^2 angles) on a small version, after some NVT steps
    have bent the chains, against the oracle (1000 beads on a cubic lattice
    of spacing 1)."""
    from rumd_amd import conf_io
    sysm, bonds, angles = configs.polymer_melt(n_chains=100, beads=10, grid=(10, 10, 10))
    conf = conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities, images=np.zeros_like(sysm.positions))
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(conf)
    rc = float(np.float32(2.0 ** (1.0 / 6.0)))
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.0 ** (1.0 / 6.0))
    pot.SetNbListSkin(0.3)
    pot.SetNB_Method("sort")
    s.SetPotential(pot)
    md = s.CreateMoleculeData()
    md.SetBonds(bonds.reshape(-1).tolist())
    md.SetAngles(angles.reshape(-1).tolist())
    bond = rumd.BondHarmonic()
    bond.SetVerbose(False)
    bond.SetParams(bond_type=0, bond_length=1.0, stiffness=3000.0, exclude=True)
    ang = rumd.AngleCosSq()
    ang.SetVerbose(False)
    ang.SetParams(angle_type=0, theta0=float(np.pi), ktheta=10.0)
    s.AddPotential(bond)
    s.AddPotential(ang)
    bond.SetExclusions(pot)
    s.SetIntegrator(rumd.IntegratorNVT(timeStep=0.002, targetTemperature=1.0, thermostatRelaxationTime=0.2))
    print("[case] polymer 100x10", flush=True)
    s.RunSteps(0, 300, 10, 100)
    s.CalcF()
    f = s.GetForcesAndEnergies().astype(np.float64)
    pos = s.GetPositions()
    pos4 = configs.pack_pos4(pos, sysm.types)
    table = configs.lj_param_table([(0, 0, 1.0, 1.0, 2.0 ** (1.0 / 6.0))], 1)
    assert abs(table[0, 0] - rc) < 1e-6
    excl = {(int(a), int(b)) for a, b, _ in bonds}
    conf2 = conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=pos, masses=sysm.masses)
    f_lj, _, _ = _oracle_lj_excluding(conf2, pos4, table, excl)
    f_b, _ = oracle.bonded(pos4, sysm.box, bonds, [(1.0, 3000.0)], angles, [(float(np.pi), 10.0)])
    f_ref = f_lj + f_b
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
