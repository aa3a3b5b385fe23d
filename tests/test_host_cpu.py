"""Host-side logic of the C++ layer that runs without a GPU: module import,
parameter tables (Pot_LJ_12_6::SetParams + CopyParamsToGPU shifts), error
mapping, configuration IO and the workload generators."""
from __future__ import annotations

import numpy as np
import pytest

from rumd_amd import configs, conf_io
import rumd_amd.rumd as rumd


def test_module_surface_names():
    for name in ("Sample", "Pot_LJ_12_6", "Pot_IPL_12_6", "IntegratorNVT", "IntegratorNVE",
                 "RectangularSimulationBox", "NoShift", "ShiftedPotential", "ShiftedForce", "Device"):
        assert hasattr(rumd, name), name
    assert int(rumd.NoShift) == 0 and int(rumd.ShiftedPotential) == 1 and int(rumd.ShiftedForce) == 2
    assert rumd.NumberOfPairPotentialParameters() == 12


def test_lj_parameter_rows_match_reference_formulas():
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    for (i, j, s, e, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=s, Epsilon=e, Rcut=rc)
    ref = configs.lj_param_table(configs.ka_params(2.5), 2)
    for i in range(2):
        for j in range(2):
            row = ref[i * 2 + j]
            for k in (0, 1, 2, 10, 11):
                assert pot.GetPotentialParameter(i, j, k) == row[k], (i, j, k)
    # A12 = 4 eps sigma^12, A6 = -4 eps sigma^6, Rcut = rc sigma (AB pair)
    assert pot.GetPotentialParameter(0, 1, 1) == pytest.approx(4 * 1.5 * 0.8 ** 12, rel=1e-6)
    assert pot.GetPotentialParameter(0, 1, 2) == pytest.approx(-4 * 1.5 * 0.8 ** 6, rel=1e-6)
    assert pot.GetPotentialParameter(0, 1, 0) == np.float32(2.5) * np.float32(0.8)
    # shifted-force constants: u(rc) and s(rc) rc of LJ at rc = 2.5
    u = 4 * (2.5 ** -12 - 2.5 ** -6)
    src = 24 * (2 * 2.5 ** -13 - 2.5 ** -7)
    assert pot.GetPotentialParameter(0, 0, 11) == pytest.approx(u, rel=1e-6)
    assert pot.GetPotentialParameter(0, 0, 10) == pytest.approx(src, rel=1e-6)
    assert len(pot.GetParameterList(2)) == 4 * 10


def test_ipl_parameters_and_newton3():
    pot = rumd.Pot_IPL_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetParams(0, 1, A12=2.0, A6=-1.0, Rcut=2.0)
    assert pot.GetPotentialParameter(1, 0, 1) == 2.0
    pot2 = rumd.Pot_IPL_12_6(cutoff_method=rumd.ShiftedPotential)
    pot2.SetAssumeNewton3(False)
    pot2.SetParams(0, 1, A12=2.0, A6=-1.0, Rcut=2.0)
    assert pot2.GetPotentialParameter(1, 0, 1) == 0.0


def test_errors_map_to_python_exceptions():
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    with pytest.raises(RuntimeError):
        pot.SetNB_Method("bogus")
    with pytest.raises(RuntimeError):
        pot.SetNbListSkin(-1.0)
    with pytest.raises(RuntimeError):
        rumd.RectangularSimulationBox(0.0, 1.0, 1.0)
    itg = rumd.IntegratorNVT(timeStep=0.005, targetTemperature=1.0, thermostatRelaxationTime=0.2)
    with pytest.raises(RuntimeError):
        itg.SetThermostatOnParticleType(40, True)
    with pytest.raises(RuntimeError):
        itg.Integrate()  # no sample
    assert itg.GetInfoString(6).startswith("IntegratorNVT,0.005")
    nve = rumd.IntegratorNVE(timeStep=0.004)
    assert nve.GetInfoString(6) == "IntegratorNVE,0.004"


def test_box_and_info_string():
    b = rumd.RectangularSimulationBox(3.0, 4.0, 5.0)
    assert b.GetVolume() == pytest.approx(60.0)
    assert b.GetLength(1) == 4.0
    assert b.GetInfoString(6) == "RectangularSimulationBox,3,4,5"


def test_conf_io_round_trip(tmp_path, golden_dir):
    c = conf_io.read_conf(golden_dir / "KA_N1000_T0.466.xyz.gz")
    assert c.n == 1000 and c.num_types == 2 and c.forces.shape == (1000, 3)
    out = tmp_path / "x.xyz.gz"
    conf_io.write_conf(out, c, integrator_info="IntegratorNVT,0.004,0.466,0.2,0.1")
    d = conf_io.read_conf(out)
    assert np.array_equal(d.types, c.types)
    np.testing.assert_allclose(d.positions, c.positions, atol=1e-9)
    np.testing.assert_allclose(d.forces, c.forces, atol=1e-9)
    np.testing.assert_allclose(d.box, c.box)
    assert d.meta["integrator"].startswith("IntegratorNVT")


def test_workload_generators():
    s = configs.ka_random_lattice(65536, density=1.2, seed=2001)
    assert s.n == 65536 and (s.types == 0).sum() == 52429
    assert abs(s.box[0] - 37.94020) < 1e-4
    assert np.abs(s.positions).max() <= s.box[0] / 2 + 1e-4
    p = (s.velocities.astype(np.float64)).sum(axis=0)
    assert np.abs(p).max() < 1e-2
    f = configs.ka_random_lattice(4 * 16 ** 3, density=1.2, seed=3001, lattice="fcc")
    d = f.positions[:, None, :50].astype(np.float64) if False else None
    assert f.n == 16384
    lj = configs.lj_liquid(10, density=0.8442, temperature=0.72)
    assert lj.n == 1000 and abs(lj.box[0] - 10.58079) < 1e-4
    # identical seeds give identical configurations
    assert np.array_equal(configs.ka_random_lattice(1000, seed=5).positions,
                          configs.ka_random_lattice(1000, seed=5).positions)


def test_polymer_melt_generator():
    s, bonds, angles = configs.polymer_melt(n_chains=200, beads=10, grid=(10, 10, 20))
    assert s.n == 2000 and bonds.shape == (1800, 3) and angles.shape == (1600, 4)
    assert s.box[0] == pytest.approx(2000 ** (1 / 3))
    d = s.positions[bonds[:, 0]] - s.positions[bonds[:, 1]]
    d -= s.box * np.rint(d / s.box)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), s.box[0] / 10, rtol=1e-5)
    # chains never cross: bonded beads belong to the same chain of 10
    assert (bonds[:, 0] // 10 == bonds[:, 1] // 10).all()
    assert np.abs(s.velocities.mean(axis=0)).max() < 1e-6
    with pytest.raises(ValueError):
        configs.polymer_melt(n_chains=10, beads=10, grid=(7, 7, 2))


def test_bonded_class_ids_and_parameter_counts():
    """ID strings as the reference sets them, and DihedralPotential's parameter
    count check (six for Ryckaert, four for the periodic form)."""
    assert rumd.AngleCosSq().GetID_String() == "AngleCosSq"
    assert rumd.AngleSq().GetID_String() == "AngleSq"
    assert rumd.BondFENE().GetID_String() == "bondFene"
    assert rumd.PeriodicDihedral().GetID_String() == "PeriodicDihedral"
    with pytest.raises(RuntimeError, match="Wrong number of parameters"):
        rumd.PeriodicDihedral().SetParams(dihedral_type=0, coeffs=[0.1] * 6)
    with pytest.raises(RuntimeError, match="Wrong number of parameters"):
        rumd.DihedralRyckaert().SetParams(dihedral_type=0, coeffs=[0.1] * 4)
