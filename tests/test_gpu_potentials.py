"""Pair potential forms outside the 12-6 family on the GPU (rumdx_pair_form
through the tile, N2 and no-list kernels, and the host classes through
Sample.CalcF) against the FP64 oracle's restatement of each form."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import GpuSystem, gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, load_golden, unsort
from rumd_amd import configs, conf_io, rumd

pytestmark = pytest.mark.gpu

NS, SP, SF = 0, 1, 2
CM = {NS: rumd.NoShift, SP: rumd.ShiftedPotential, SF: rumd.ShiftedForce}
KA_ROWS = [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.8, 1.5, 2.5), (1, 1, 0.88, 0.5, 2.5)]
# per-particle energy, absolute: FP32 sums of ~60 terms of size up to ~10
PE_ATOL = 5e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _forms():
    """(name, form number, factory taking a cut-off method, SetParams rows)."""
    return [
        ("smooth_2_4", 1, lambda c: rumd.Pot_LJ_12_6_smooth_2_4(cutoff_method=c), KA_ROWS),
        ("gLJ_9_5", 2, lambda c: rumd.Pot_gLJ_m_n(set_m=9.0, set_n=5.0, cutoff_method=c), KA_ROWS),
        ("Gauss", 3, lambda c: rumd.Pot_Gauss(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.9, 1.2, 2.5), (1, 1, 0.8, 0.7, 2.5)]),
        ("Buckingham", 4, lambda c: rumd.Pot_Buckingham(cutoff_method=c),
         [(0, 0, 1.1, 14.0, 1.0, 2.5), (0, 1, 0.9, 14.0, 1.5, 2.3), (1, 1, 1.0, 13.0, 0.5, 2.4)]),
        ("IPL_18", 5, lambda c: rumd.Pot_IPL_n(n=18.0, cutoff_method=c), KA_ROWS),
        ("Yukawa", 6, lambda c: rumd.Pot_Yukawa(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.9, 1.2, 2.5), (1, 1, 0.8, 0.7, 2.5)]),
        ("Exp", 7, lambda c: rumd.Pot_Exp(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 2.5), (0, 1, 0.9, 1.2, 2.5), (1, 1, 0.8, 0.7, 2.5)]),
        ("LJ_Gauss", 8, lambda c: rumd.Pot_LJ_12_6_Gauss(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 2.5, 0.15, 1.5, 1.5), (0, 1, 0.8, 1.5, 2.5, 0.12, 1.5, 1.2),
          (1, 1, 0.88, 0.5, 2.5, 0.13, 1.5, 1.3)]),
        ("IPL_12", 9, lambda c: rumd.Pot_IPL_12(cutoff_method=c), KA_ROWS),
        ("IPL_9", 10, lambda c: rumd.Pot_IPL_9(cutoff_method=c), KA_ROWS),
        ("IPL_18_class", 11, lambda c: rumd.Pot_IPL_18(cutoff_method=c), KA_ROWS),
        ("IPL6_LN", 12, lambda c: rumd.Pot_IPL6_LN(cutoff_method=c),
         [(0, 0, 0.5, 1.0, 5.0), (0, 1, 0.4, 1.5, 5.0), (1, 1, 0.44, 0.5, 5.0)]),
        ("LJ_LN", 13, lambda c: rumd.Pot_LJ_LN(cutoff_method=c),
         [(0, 0, 1.0, 0.5, 1.0, 2.5), (0, 1, 0.8, 0.4, 1.5, 2.5), (1, 1, 0.88, 0.44, 0.5, 2.5)]),
        ("Girifalco", 14, lambda c: rumd.Pot_Girifalco(cutoff_method=c),
         [(0, 0, 0.5, 0.01, 0.5, 5.0), (0, 1, 0.6, 0.01, 0.4, 5.0), (1, 1, 0.4, 0.01, 0.44, 5.0)]),
        ("Dzugutov", 15, lambda c: rumd.Pot_Dzugutov(),
         [(0, 0, 1.87, 1.94, 1.1, 0.27, 5.82, 1.28), (0, 1, 1.5, 1.55, 1.1, 0.27, 0.5, 1.28),
          (1, 1, 1.65, 1.71, 1.1, 0.27, 0.5, 1.28)]),
        ("Fermi_Jagla", 16, lambda c: rumd.Pot_Fermi_Jagla(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 4.56, 28.88, 1.36, 1.0, 3.57, 2.36, 12.0, 2.5),
          (0, 1, 0.8, 1.5, 4.56, 28.88, 1.36, 1.0, 3.57, 2.36, 12.0, 2.5),
          (1, 1, 0.88, 0.5, 4.56, 28.88, 1.36, 1.0, 3.57, 2.36, 12.0, 2.5)]),
        ("Repulsive_Shoulder", 17, lambda c: rumd.Pot_Repulsive_Shoulder(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 10.0, 1.35, 2.5), (0, 1, 0.8, 1.5, 10.0, 1.08, 2.5), (1, 1, 0.88, 0.5, 10.0, 1.19, 2.5)]),
        ("SAAP", 18, lambda c: rumd.Pot_SAAP(cutoff_method=c),
         [(0, 0, 1.0, 1.0, 65214.64725, -9.452343340, -19.42488828, -1.958381959, -2.379111084, 1.051490962, -0.1, 2.5), (0, 1, 0.8, 1.5, 65214.64725, -9.452343340, -19.42488828, -1.958381959, -2.379111084, 1.051490962, -0.1, 2.5),
          (1, 1, 0.88, 0.5, 65214.64725, -9.452343340, -19.42488828, -1.958381959, -2.379111084, 1.051490962, -0.1, 2.5)]),
        ("SF_Coulomb", 19, lambda c: rumd.Pot_ShiftedForceCoulomb(),
         [(0, 0, 1.0, 2.5), (0, 1, -1.0, 2.5), (1, 1, 1.0, 2.5)]),
    ]


def _table(make, cutoff, rows, nt=2):
    pot = make(CM[cutoff])
    pot.SetVerbose(False)
    for r in rows:
        pot.SetParams(*r)
    assert pot.GetPotentialForm() != 0
    return np.asarray(pot.GetParamsTable(nt), dtype=np.float32)


@pytest.mark.parametrize("method", ["sorted", "n2", "none"])
@pytest.mark.parametrize("cutoff", [SP, SF])
@pytest.mark.parametrize("case", _forms(), ids=lambda c: c[0])
def test_form_forces_golden_ka1000(golden_dir, case, cutoff, method):
    name, form, make, rows = case
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    table = _table(make, cutoff, rows)
    pos4 = configs.pack_pos4(c.positions, c.types)
    g = GpuSystem(pos4, configs.pack_vel4(c.velocities, c.types, c.masses), configs.box8(c.box), 2, table,
                  method=method, with_sts=True)
    print(f"[case] form {name} cutoff={cutoff} method={method}", flush=True)
    g.calc_f(cutoff, form=form, stress=True)
    uidx = g.unsorted_index()
    f = unsort(g.get("f"), uidx)
    w = unsort(g.get("w"), uidx)
    f_ref, vir_ref = oracle.forces_form_brute(pos4, c.box, 2, table, form, cutoff)
    e = force_errors(f[:, :3], f_ref[:, :3])
    print(f"  guarded {e['guarded_max']:.2e} pe {np.abs(f[:, 3] - f_ref[:, 3]).max():.2e}", flush=True)
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    # FP32 sums: the bound grows with the largest per-particle energy
    pe_tol = max(PE_ATOL, 2e-6 * np.abs(f_ref[:, 3]).max())
    np.testing.assert_allclose(f[:, 3], f_ref[:, 3], rtol=0, atol=pe_tol)
    scale = max(1.0, np.abs(vir_ref).max())
    np.testing.assert_allclose(w[:, 3], vir_ref, rtol=0, atol=1e-5 * scale)
    # the stress branch: its trace is minus the virial
    sts = unsort(g.get("sts"), uidx)
    np.testing.assert_allclose(sts[:, 0] + sts[:, 1] + sts[:, 2], -w[:, 3], rtol=0, atol=1e-5 * scale)


@pytest.mark.parametrize("case", [_forms()[i] for i in (1, 5)], ids=lambda c: c[0])
def test_form_through_sample_sorted_tiles(case):
    """The host class through Sample.CalcF on a box with at least 16 cells per
    axis (shifted-frame tiles), sorted lists, shifted-force cut-off."""
    name, form, make, rows = case
    sysm = configs.ka_random_lattice(20000, density=1.2, temperature=0.45, seed=2301)
    conf = conf_io.Configuration(
        n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions, masses=sysm.masses,
        velocities=sysm.velocities, images=np.zeros_like(sysm.positions))
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(conf)
    pot = make(rumd.ShiftedForce)
    pot.SetVerbose(False)
    for r in rows:
        pot.SetParams(*r)
    pot.SetNbListSkin(0.3)
    s.SetPotential(pot)
    print(f"[case] Sample + {name} N=20000", flush=True)
    s.CalcF()
    f = s.GetForcesAndEnergies()
    table = np.asarray(pot.GetParamsTable(2), dtype=np.float32)
    f_ref, _ = oracle.forces_form_brute(configs.pack_pos4(sysm.positions, sysm.types), sysm.box, 2, table, form, SF)
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    assert s.GetPotentialEnergy() == pytest.approx(f_ref[:, 3].sum(), rel=1e-5, abs=1e-3)
    with pytest.raises(RuntimeError, match="12-6 potentials only"):
        pot.SetFP64Forces(True)
