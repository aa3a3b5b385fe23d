"""The fused step (rumdx_pair_lj_nvt: 12-6 forces over a sorted list and the
NVT update in one kernel) against the two-call step (rumdx_pair_lj then
rumdx_nvt_step) and the oracle, through the C ABI and through
Sample.RunSteps.

The per-particle arithmetic of the two paths is the same, so positions,
velocities, images and forces agree bit for bit; only the FP64 kinetic-energy
sum is taken in another order (per cell tile instead of per 256 slots), so
the thermostat state agrees to rounding (relative 1e-12)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from capi import DeviceBuffer, GpuSystem, check, gpu_available, lib
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table
from rumd_amd import configs, conf_io
import rumd_amd.rumd as rumd

pytestmark = pytest.mark.gpu
NS, SP, SF = 0, 1, 2
RUMDX_E_INVALID = -1


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _system(n=12000, seed=5, rlo=False, method="sorted"):
    sysm = configs.ka_random_lattice(n, density=1.2, temperature=0.45, seed=seed)
    pos4 = configs.pack_pos4(sysm.positions, sysm.types)
    vel4 = configs.pack_vel4(sysm.velocities, sysm.types, sysm.masses)
    return sysm, GpuSystem(pos4, vel4, configs.box8(sysm.box), 2, ka_table(), method=method, rlo=rlo)


@pytest.mark.parametrize("rlo", [False, True])
@pytest.mark.parametrize("cutoff", [NS, SP, SF])
def test_capi_fused_step_equals_two_calls(cutoff, rlo):
    dt, tau, T = 0.0025, 0.2, 0.45
    omega2 = float(np.float32(4.0 * np.pi ** 2 / tau ** 2))
    on = 0xFFFFFFFF
    _, a = _system(rlo=rlo)
    _, b = _system(rlo=rlo)
    out = DeviceBuffer(16 * b.nvp)
    for g in (a, b):
        g.set_nh(ps=0.3, target_T=T)
    for step in range(4):
        print(f"[case] fused vs two calls cutoff={cutoff} rlo={rlo} step={step}", flush=True)
        a.calc_f(cutoff)
        a.nvt_step(dt, on_mask=on, omega2=omega2, fuse_skin=True)
        check(b.fused_step(cutoff, dt, on_mask=on, omega2=omega2, r_out=out))
        # the caller swaps the position arrays
        b.r, out = out, b.r
        b.p.r_dev = b.r.value
        for name in ("r", "v", "f", "w", "im") + (("rlo",) if rlo else ()):
            assert np.array_equal(a.get(name), b.get(name)), (name, step)
        assert np.array_equal(a.flags_host(), b.flags_host())
        na, nb = a.get_nh(), b.get_nh()
        assert nb.ke2_sum == pytest.approx(na.ke2_sum, rel=1e-12)
        assert nb.ps == pytest.approx(na.ps, rel=1e-12)
        assert nb.target_T == na.target_T


@pytest.mark.parametrize("thermo", [False, True])
def test_capi_fused_step_against_oracle(thermo):
    """Forces of the fused step against the oracle's cell-list forces; the
    update, given those forces, bit for bit against the oracle's leapfrog;
    the thermostat against the oracle's FP64 update."""
    dt, tau, T, ps0 = 0.0025, 0.2, 0.45, 0.25
    omega2 = float(np.float32(4.0 * np.pi ** 2 / tau ** 2))
    on = 0xFFFFFFFF if thermo else 0
    sysm, g = _system(n=20000, seed=11)
    out = DeviceBuffer(16 * g.nvp)
    g.set_nh(ps=ps0, target_T=T)
    g.update_nblist()
    r0, v0, im0 = g.get("r").copy(), g.get("v").copy(), g.get("im").copy()
    check(g.fused_step(SF, dt, on_mask=on, omega2=omega2, r_out=out))
    f = g.get("f")
    f_ref, _ = oracle.forces_cells(r0, sysm.box, 2, ka_table(), SF)
    assert force_errors(f[:, :3], f_ref[:, :3])["guarded_max"] <= FORCE_TOL_GUARDED
    r_h, v_h, im_h = r0.copy(), v0.copy(), im0.copy()
    ke2 = oracle.nvt_step(r_h, v_h, f, im_h, g.box8, dt, ps0, on)
    assert np.array_equal(out.download(np.float32, (g.nvp, 4))[: g.n], r_h)
    assert np.array_equal(g.get("v"), v_h)
    assert np.array_equal(g.get("im"), im_h)
    nh = g.get_nh()
    if thermo:
        ps_ref, _ = oracle.thermostat_update(ps0, T, 0.0, ke2, 3 * g.n - 3, dt, omega2)
        assert nh.ke2_sum == pytest.approx(ke2, rel=1e-12)
        assert nh.ps == pytest.approx(ps_ref, rel=1e-12, abs=1e-15)
    else:
        assert nh.ps == ps0


def test_capi_fused_step_rejects_bad_arguments():
    _, g = _system(n=4000)
    out = DeviceBuffer(16 * g.nvp)
    g.set_nh(ps=0.0, target_T=0.45)
    # the new positions may not overwrite the ones the forces read
    assert g.fused_step(SF, 0.0025, r_out=g.r) == RUMDX_E_INVALID
    assert g.fused_step(SF, 0.0025, r_out=None) == RUMDX_E_INVALID
    assert g.fused_step(7, 0.0025, r_out=out) == RUMDX_E_INVALID
    assert g.fused_step(SF, 0.0025, on_mask=1, omega2=1.0, dof=0, r_out=out) == RUMDX_E_INVALID
    assert b"rumdx_pair_lj_nvt" in lib().rumdx_last_error()
    # lists of the N2 method are not served
    _, h = _system(n=4000, method="n2")
    out_h = DeviceBuffer(16 * h.nvp)
    h.set_nh(ps=0.0, target_T=0.45)
    assert h.fused_step(SF, 0.0025, r_out=out_h) == RUMDX_E_INVALID


# ---------------------------------------------------------------- Sample level
def _conf(n, seed, temperature=0.45):
    sysm = configs.ka_random_lattice(n, density=1.2, temperature=temperature, seed=seed)
    return conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities,
                                 images=np.zeros_like(sysm.positions))


def _sample(conf, fused, thermo=True, mixed=False, cutoff=rumd.ShiftedForce):
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    if mixed:
        s.SetMixedPrecisionPositions(True)
    s.SetConfiguration(conf)
    s.SetFusedSteps(fused)
    pot = rumd.Pot_LJ_12_6(cutoff_method=cutoff)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    s.SetPotential(pot)
    if thermo:
        itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    else:
        itg = rumd.IntegratorNVE(timeStep=0.0025)
    s.SetIntegrator(itg)
    return s, pot, itg


def _state(s, pot, itg):
    return dict(r=s.GetPositions(), v=s.GetVelocities(), im=s.GetImages(), f=s.GetForces(),
                ps=itg.GetThermostatState(), builds=pot.GetNumberOfNbListBuilds())


@pytest.mark.parametrize("mixed", [False, True])
def test_sample_fused_nve_bit_identical_to_split(mixed):
    """NVE has no kinetic-energy feedback: 300 steps with list rebuilds give
    the same bits with and without the fused kernel."""
    conf = _conf(12000, seed=41, temperature=0.8)
    out = []
    for fused in (False, True):
        s, pot, itg = _sample(conf, fused, thermo=False, mixed=mixed)
        s.RunSteps(0, 300, 10, 0)
        assert s.UsesFusedSteps() == fused
        out.append(_state(s, pot, itg))
    a, b = out
    assert a["builds"] == b["builds"] and a["builds"] > 2
    for k in ("r", "v", "im", "f"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_sample_fused_nve_with_inner_lists_bit_identical_to_split():
    """The fused kernel reads the pruned inner lists and raises the prune
    request like the integrator's skin check."""
    conf = _conf(12000, seed=61, temperature=0.8)
    out = []
    for fused in (False, True):
        s, pot, itg = _sample(conf, fused, thermo=False)
        pot.SetNbListInnerSkin(0.1)
        s.RunSteps(0, 200, 10, 0)
        assert pot.HasNbInnerLists() and s.UsesFusedSteps() == fused
        out.append(_state(s, pot, itg))
    a, b = out
    assert a["builds"] == b["builds"]
    for k in ("r", "v", "im", "f"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_sample_fused_nvt_tracks_split():
    conf = _conf(12000, seed=43)
    out = []
    for fused in (False, True):
        s, pot, itg = _sample(conf, fused)
        s.RunSteps(0, 200, 10, 100)
        out.append(_state(s, pot, itg) | {"ke": itg.GetThermostatKineticEnergy()})
    a, b = out
    assert b["ps"] == pytest.approx(a["ps"], rel=1e-9)
    assert b["ke"] == pytest.approx(a["ke"], rel=1e-9)
    np.testing.assert_allclose(b["r"], a["r"], atol=1e-5)
    np.testing.assert_allclose(b["v"], a["v"], atol=1e-5)


def test_sample_fused_eager_graph_and_odd_blocks_agree():
    """Fused steps alternate two position arrays: eager steps (blocks of 1),
    graph blocks of 10, odd blocks of 7 (each graph copies back to the array
    it starts from) and runs that end on an odd step all give the same bits."""
    conf = _conf(12000, seed=47)
    runs = {"eager": [(0, 60, 1)], "graph10": [(0, 60, 10)], "odd7": [(0, 60, 7)],
            "split_calls": [(0, 13, 10), (13, 1, 10), (14, 33, 5), (47, 13, 10)]}
    out = {}
    for name, calls in runs.items():
        s, pot, itg = _sample(conf, True)
        for first, n, block in calls:
            print(f"[case] fused run {name} first={first} n={n} block={block}", flush=True)
            s.RunSteps(first, n, block, 100)
        assert s.UsesFusedSteps()
        out[name] = _state(s, pot, itg)
    ref = out["graph10"]
    for name, st in out.items():
        assert st["builds"] == ref["builds"], name
        for k in ("r", "v", "im", "f"):
            np.testing.assert_array_equal(st[k], ref[k], err_msg=f"{name} {k}")
        assert st["ps"] == ref["ps"], name


def test_sample_host_calls_after_odd_fused_steps():
    """After a run ending on an odd step the host getters, CalcF and the
    energies read the current positions."""
    conf = _conf(12000, seed=53)
    res = []
    for fused in (False, True):
        s, pot, itg = _sample(conf, fused, thermo=False)
        s.RunSteps(0, 5, 10, 0)
        assert s.UsesFusedSteps() == fused
        r = s.GetPositions()
        s.CalcF()
        res.append((r, s.GetForces(), s.GetPotentialEnergy(), s.GetVirial()))
    (r0, f0, e0, w0), (r1, f1, e1, w1) = res
    np.testing.assert_array_equal(r0, r1)
    np.testing.assert_array_equal(f0, f1)
    assert e0 == e1 and w0 == w1


def test_fused_steps_only_where_they_apply():
    conf = _conf(12000, seed=59)  # the default list method sorts from 10,000 particles
    s, pot, itg = _sample(conf, True)
    s.RunSteps(0, 2, 10, 0)
    assert s.UsesFusedSteps()
    pot.SetFP64Forces(True)
    assert not s.UsesFusedSteps()
    pot.SetFP64Forces(False)
    assert s.UsesFusedSteps()
    s.SetFusedSteps(False)
    assert not s.UsesFusedSteps()
    # N2 lists take the two-kernel step
    s2, pot2, _ = _sample(conf, True)
    pot2.SetNB_Method("n2")
    s2.RunSteps(0, 2, 10, 0)
    assert not s2.UsesFusedSteps()
    # (a sample with bonds or angles besides the pair potential takes the
    # two-kernel step too: tests/test_gpu_bonded.py checks it)
