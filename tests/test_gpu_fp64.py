"""All-FP64 force path (rumdx_pair_lj_fp64) against the FP64 oracle: the
comparison leg of the precision study.  Outputs are float4, so the bar is the
final FP32 rounding of each component (guarded error <= 2e-7), and the path
must be well below the FP32 kernel's error on the same input."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from capi import GpuSystem, gpu_available
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, load_golden, unsort
from rumd_amd import configs

pytestmark = pytest.mark.gpu

NS, SP, SF = 0, 1, 2
FP64_TOL_GUARDED = 2e-7  # one FP32 rounding of each output component


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


@pytest.mark.parametrize("method", ["n2", "sorted", "none"])
def test_fp64_golden_ka1000(golden_dir, method):
    c = load_golden(golden_dir / "KA_N1000_T0.466.xyz.gz")
    table = ka_table()
    pos4 = configs.pack_pos4(c.positions, c.types)
    g = GpuSystem(pos4, configs.pack_vel4(c.velocities, c.types, c.masses), configs.box8(c.box), 2, table,
                  method=method)
    print(f"[case] fp64 KA_N1000 method={method}", flush=True)
    g.calc_f(SP, fp64=True)
    uidx = g.unsorted_index()
    f = unsort(g.get("f"), uidx)
    f_ref, vir_ref = oracle.forces_brute(pos4, c.box, 2, table, SP)
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FP64_TOL_GUARDED, e
    pe = float(np.sum(f[:, 3], dtype=np.float64))
    assert pe == pytest.approx(float(f_ref[:, 3].sum()), rel=1e-7)
    w = unsort(g.get("w"), uidx)
    assert float(np.sum(w[:, 3], dtype=np.float64)) == pytest.approx(float(vir_ref.sum()), rel=1e-7)


@pytest.mark.parametrize("cutoff", [NS, SP, SF])
def test_fp64_ka_sorted_65536(cutoff):
    """Tile lists with wrapping halos; the FP64 path against the oracle and
    against the FP32 kernel's error on the same state."""
    sysm = configs.ka_random_lattice(65536, density=1.2, temperature=0.45, seed=2001)
    table = ka_table()
    pos4 = configs.pack_pos4(sysm.positions, sysm.types)
    g = GpuSystem(pos4, configs.pack_vel4(sysm.velocities, sysm.types, sysm.masses), configs.box8(sysm.box), 2,
                  table, method="sorted", max_nbrs=160)
    print(f"[case] fp64 KA_N65536 cutoff={cutoff}", flush=True)
    f_ref, _ = oracle.forces_cells(pos4, sysm.box, 2, table, cutoff)
    g.calc_f(cutoff, fp64=True)
    uidx = g.unsorted_index()
    e64 = force_errors(unsort(g.get("f"), uidx)[:, :3], f_ref[:, :3])
    g.calc_f(cutoff)
    e32 = force_errors(unsort(g.get("f"), uidx)[:, :3], f_ref[:, :3])
    assert e64["guarded_max"] <= FP64_TOL_GUARDED, e64
    assert e32["guarded_max"] <= FORCE_TOL_GUARDED, e32
    assert e64["guarded_max"] < e32["guarded_max"], (e64, e32)


def test_fp64_small_tile_halo_capacity():
    """Halos beyond the FP64 kernel's LDS capacity read positions from global
    memory; the result is the same."""
    sysm = configs.ka_random_lattice(20000, density=1.2, temperature=0.45, seed=7)
    table = ka_table()
    pos4 = configs.pack_pos4(sysm.positions, sysm.types)
    g = GpuSystem(pos4, configs.pack_vel4(sysm.velocities, sysm.types, sysm.masses), configs.box8(sysm.box), 2,
                  table, method="sorted", halo_cap=64)
    print("[case] fp64 KA_N20000 halo_cap=64", flush=True)
    g.calc_f(SF, fp64=True)
    f = unsort(g.get("f"), g.unsorted_index())
    f_ref, _ = oracle.forces_cells(pos4, sysm.box, 2, table, SF)
    e = force_errors(f[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FP64_TOL_GUARDED, e
