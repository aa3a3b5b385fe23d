"""Stress path (the STR branch of fij, include/rumd/PairPotential.h:1232-1240,
and Sample::GetStress, src/Sample.cu:393-444) on the GPU against the FP64
oracle: per-particle sxx, syy, szz, syz in sts and sxz, sxy in w.y, w.z, through
every force kernel (tile lists, N2 lists, no list, the FP64 comparison path),
accumulation with initialize = false, and the summed tensor with its kinetic
part through the host API."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import GpuSystem, gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, load_golden, unsort
from rumd_amd import configs, rumd

pytestmark = pytest.mark.gpu

NS, SP, SF = 0, 1, 2
# per-particle stress components, absolute (the per-particle virial's bar:
# FP32 sums of ~90 terms of size up to ~1e2)
STRESS_ATOL = 5e-4
STRESS_ATOL_FP64 = 2e-5  # one FP32 rounding of each output component


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _gpu_stress(g):
    """(n, 6) stress in original order as (sxx, syy, szz, syz, sxz, sxy)."""
    uidx = g.unsorted_index()
    sts = unsort(g.get("sts"), uidx)
    w = unsort(g.get("w"), uidx)
    return np.concatenate([sts, w[:, 1:3]], axis=1).astype(np.float64), w, unsort(g.get("f"), uidx)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("cutoff", [SP, SF])
@pytest.mark.parametrize("method", ["n2", "sorted", "none"])
def test_stress_golden_ka1000(golden_dir, method, cutoff, fp64):
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    table = ka_table()
    pos4 = configs.pack_pos4(c.positions, c.types)
    g = GpuSystem(pos4, configs.pack_vel4(c.velocities, c.types, c.masses), configs.box8(c.box), 2, table,
                  method=method, with_sts=True)
    print(f"[case] stress KA_N1000 method={method} cutoff={cutoff} fp64={fp64}", flush=True)
    g.calc_f(cutoff, stress=True, fp64=fp64)
    s, w, f = _gpu_stress(g)
    s_ref = oracle.stress_brute(pos4, c.box, 2, table, cutoff)
    f_ref, vir_ref = oracle.forces_brute(pos4, c.box, 2, table, cutoff)
    np.testing.assert_allclose(s, s_ref, rtol=0, atol=STRESS_ATOL_FP64 if fp64 else STRESS_ATOL)
    # the stress branch leaves forces, energies and the virial as they were
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    np.testing.assert_allclose(w[:, 3], vir_ref, rtol=0, atol=5e-4)
    np.testing.assert_allclose(s[:, 0] + s[:, 1] + s[:, 2], -w[:, 3], rtol=0, atol=1e-3)


@pytest.mark.parametrize("cutoff", [SP, SF])
def test_stress_sorted_tiles_20000(cutoff):
    """A box with at least 16 cells per axis, so interior and wrapping tiles
    take the shifted-frame path; sums against the oracle to 1e-5."""
    sysm = configs.ka_random_lattice(20000, density=1.2, temperature=0.45, seed=2201)
    table = ka_table()
    pos4 = configs.pack_pos4(sysm.positions, sysm.types)
    b8 = configs.box8(sysm.box)
    g = GpuSystem(pos4, configs.pack_vel4(sysm.velocities, sysm.types, sysm.masses), b8, 2, table,
                  method="sorted", max_nbrs=160, with_sts=True)
    assert min(oracle.cell_grid(b8, 2.5, 0.3)) >= 16
    print(f"[case] stress KA_N20000 sorted cutoff={cutoff}", flush=True)
    g.calc_f(cutoff, stress=True)
    s, _, _ = _gpu_stress(g)
    s_ref = oracle.stress_brute(pos4, sysm.box, 2, table, cutoff)
    np.testing.assert_allclose(s, s_ref, rtol=0, atol=STRESS_ATOL)
    tot, tot_ref = s.sum(axis=0), s_ref.sum(axis=0)
    scale = np.abs(tot_ref).max()
    np.testing.assert_allclose(tot, tot_ref, rtol=0, atol=1e-5 * scale)


@pytest.mark.parametrize("fp64", [False, True])
def test_stress_accumulates_without_initialize(golden_dir, fp64):
    """initialize = false adds into sts and w.y, w.z (a second potential's
    contribution): two passes give twice one pass."""
    c = load_golden(golden_dir / "KA_N256_L6.0_T0.50.xyz.gz")
    table = ka_table()
    pos4 = configs.pack_pos4(c.positions, c.types)
    g = GpuSystem(pos4, configs.pack_vel4(c.velocities, c.types, c.masses), configs.box8(c.box), 2, table,
                  method="sorted", with_sts=True)
    print(f"[case] stress accumulate fp64={fp64}", flush=True)
    g.calc_f(SF, stress=True, fp64=fp64)
    one, _, _ = _gpu_stress(g)
    g.calc_f(SF, initialize=False, stress=True, fp64=fp64)
    two, _, _ = _gpu_stress(g)
    np.testing.assert_allclose(two, 2.0 * one, rtol=1e-6, atol=1e-5)


def test_host_get_stress_with_kinetic_part(golden_dir):
    """Sample.CalcF(True) then GetStress(): 1/(2V) (sum of per-particle stress
    - sum 2 m v_a v_b), in RUMD's (xx, yy, zz, yz, xz, xy) order."""
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(c)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    s.SetPotential(pot)
    print("[case] host GetStress KA_N1000 SF", flush=True)
    s.CalcF(True)
    got = np.array(s.GetStress())
    per = np.asarray(s.GetStresses(), dtype=np.float64)
    pos4 = configs.pack_pos4(c.positions, c.types)
    s_ref = oracle.stress_brute(pos4, c.box, 2, ka_table(), SF)
    np.testing.assert_allclose(per, s_ref[:, :4], rtol=0, atol=STRESS_ATOL)
    v = np.asarray(c.velocities, dtype=np.float64)
    m = np.asarray(c.masses, dtype=np.float64)[np.asarray(c.types)]
    kin = np.stack([m * v[:, 0] ** 2, m * v[:, 1] ** 2, m * v[:, 2] ** 2,
                    m * v[:, 1] * v[:, 2], m * v[:, 0] * v[:, 2], m * v[:, 0] * v[:, 1]], axis=1).sum(axis=0)
    V = float(np.prod(np.asarray(c.box, dtype=np.float64)))
    want = 0.5 * (s_ref.sum(axis=0) - 2.0 * kin) / V
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    # and the pressure relation: -trace/3 = (N T_kin + W) / V with W = sum vir / 6
    vir = np.asarray(s.GetVirials(), dtype=np.float64)[:, 3].sum() / 6.0
    p_from_stress = -got[:3].sum() / 3.0
    p_direct = (2.0 * kin[:3].sum() / 2.0 / 3.0 + vir) / V
    assert p_from_stress == pytest.approx(p_direct, rel=1e-5)
