"""Inner lists of the sorted method (rumdx_nblist_prune): the force kernels read
the tile-list entries within rc + inner_skin, pruned after every build and
whenever a particle has moved 0.49 inner_skin since the last prune.

Checked through the C ABI: each inner list is the ordered subsequence of its
tile list whose pairs lie within rc_ij + inner_skin (FP64 distances, with a
1e-5 band for the FP32 criterion), the forces over the inner lists equal the
FP64 oracle's, the prune flag follows the 0.49 inner_skin displacement rule,
and a trajectory of skin-triggered prunes keeps every force exact."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from capi import GpuSystem, check, gpu_available, lib
from helpers import FORCE_TOL_GUARDED, force_errors, ka_table, unsort
from rumd_amd import configs

pytestmark = pytest.mark.gpu

SF = 2
FLAG_REBUILD, FLAG_PRUNE = 0, 9
BAND = 1e-5


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _ka_system(n, seed, inner, skin=0.3):
    sysm = configs.ka_random_lattice(n, density=1.2, temperature=0.45, seed=seed)
    pos4 = configs.pack_pos4(sysm.positions, sysm.types)
    vel4 = configs.pack_vel4(sysm.velocities, sysm.types, sysm.masses)
    g = GpuSystem(pos4, vel4, configs.box8(sysm.box), 2, ka_table(), skin=skin, method="sorted")
    if inner:
        g.enable_inner(inner)
    return sysm, g


def _pair_d(pos, box, i, js):
    d = pos[js].astype(np.float64) - pos[i].astype(np.float64)
    d -= box * np.rint(d / box)
    return np.sqrt((d * d).sum(axis=1))


def _staged_slots(g):
    """Slots whose tile had its halo staged by the last build (the others,
    like overflowed lists, are computed by rescans and keep no inner list)."""
    info = g.tile_info.download(np.uint32, (g.ntiles, 4))
    cap = g.nb.halo_cap
    staged = ((info[:, 1] >> 16) == cap) & (info[:, 0] <= cap)
    out = np.zeros(g.nvp, dtype=bool)
    for t in np.nonzero(staged)[0]:
        out[info[t, 2]:info[t, 3]] = True
    return out[: g.n]


def _check_inner_lists(g, sysm, inner, rc_current=False):
    """Every staged particle's inner list is the ordered subsequence of its
    tile list whose pairs were within rc_ij + inner at the last prune (FP64
    distances at the recorded prune positions, 1e-5 band); with rc_current,
    it also holds every pair within rc_ij at the current positions (the
    validity the displacement rule guarantees).  Returns the entries dropped."""
    cnt_o, nbl_o = g.neighbor_lists()
    cnt_i, nbl_i = g.neighbor_lists(inner=True)
    r = g.get("r")
    pos_now = r[:, :3]
    pos_prune = g.prune_r.download(np.float32, (g.nvp, 4))[: g.n, :3]
    types = np.ascontiguousarray(r[:, 3]).view(np.int32)
    rc = np.array([[2.5, 2.0], [2.0, 2.2]])
    box = np.asarray(sysm.box, dtype=np.float64)
    staged = _staged_slots(g)
    assert staged.mean() > 0.5
    dropped = 0
    for s in np.nonzero(staged)[0]:
        if cnt_o[s] >= g.max_nbrs:
            continue
        outer = nbl_o[: cnt_o[s], s]
        inn = nbl_i[: cnt_i[s], s]
        lim = rc[types[s], types[outer]]
        d = _pair_d(pos_prune, box, s, outer)
        must = d < lim + inner - BAND
        may = d < lim + inner + BAND
        keep = np.isin(outer, inn)
        assert np.all(keep[must]), s
        assert not np.any(keep & ~may), s
        # the inner list is the kept entries of the tile list, in its order
        assert np.array_equal(outer[keep], inn), s
        if rc_current:
            assert np.all(keep[_pair_d(pos_now, box, s, outer) < lim]), s
        dropped += int(cnt_o[s] - cnt_i[s])
    return dropped


def test_inner_lists_are_the_tile_entries_within_rc_plus_inner_skin():
    sysm, g = _ka_system(20000, 71, 0.1)
    print("[case] KA N=20000 inner 0.1: build + prune", flush=True)
    g.update_nblist()
    assert g.flags_host()[FLAG_PRUNE] == 0
    dropped = _check_inner_lists(g, sysm, 0.1)
    # the skin shell between rc + 0.1 and rc + 0.3 holds about a fifth of the entries
    cnt_o, _ = g.neighbor_lists()
    assert 0.1 < dropped / cnt_o.sum() < 0.35


@pytest.mark.parametrize("inner", [0.05, 0.1])
def test_forces_over_inner_lists_match_the_oracle_and_the_tile_lists(inner):
    sysm, g_in = _ka_system(30000, 72, inner)
    _, g_out = _ka_system(30000, 72, None)
    print(f"[case] KA N=30000 SF inner {inner}", flush=True)
    g_in.calc_f(SF)
    g_out.calc_f(SF)
    f_in = unsort(g_in.get("f"), g_in.unsorted_index())
    f_out = unsort(g_out.get("f"), g_out.unsorted_index())
    f_ref, _ = oracle.forces_cells(configs.pack_pos4(sysm.positions, sysm.types), sysm.box, 2, ka_table(), SF)
    e = force_errors(f_in[:, :3], f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
    # same pairs, FP32 sums in possibly different lane groupings
    assert force_errors(f_in[:, :3], f_out[:, :3].astype(np.float64))["guarded_max"] <= 1e-6
    np.testing.assert_allclose(f_in[:, 3], f_out[:, 3], rtol=1e-5, atol=1e-5)


def test_prune_flag_follows_the_displacement_rule():
    sysm, g = _ka_system(8000, 73, 0.1)
    g.update_nblist()
    assert g.flags_host()[FLAG_PRUNE] == 0
    r = g.get("r").copy()
    step = np.float32(0.049 * 0.99)  # below 0.49 inner_skin: no prune, no rebuild
    r2 = r.copy()
    r2[5, 0] += step
    g.r.upload(np.vstack([r2, np.zeros((g.nvp - g.n, 4), np.float32)]))
    check(lib().rumdx_nblist_check(C.byref(g.p), C.byref(g.nb), g.box8, None))
    f = g.flags_host()
    assert f[FLAG_PRUNE] == 0 and f[FLAG_REBUILD] == 0
    r2[5, 0] = r[5, 0] + np.float32(0.049 * 1.02)  # above: prune, not rebuild
    g.r.upload(np.vstack([r2, np.zeros((g.nvp - g.n, 4), np.float32)]))
    check(lib().rumdx_nblist_check(C.byref(g.p), C.byref(g.nb), g.box8, None))
    f = g.flags_host()
    assert f[FLAG_PRUNE] == 1 and f[FLAG_REBUILD] == 0
    builds = f[2]
    check(lib().rumdx_nblist_prune(C.byref(g.p), C.byref(g.nb), g.box8, None))
    f = g.flags_host()
    assert f[FLAG_PRUNE] == 0 and f[2] == builds
    _check_inner_lists(g, sysm, 0.1)
    # the prune recorded the new positions: the same check passes now
    check(lib().rumdx_nblist_check(C.byref(g.p), C.byref(g.nb), g.box8, None))
    assert g.flags_host()[FLAG_PRUNE] == 0


def test_nvt_trajectory_with_prunes_keeps_forces_exact():
    """60 leapfrog steps with the fused skin and prune checks: prunes happen
    between builds, every inner list keeps the pairs now within rc, and the
    forces over the inner lists equal those over the tile lists and the
    oracle's."""
    sysm, g = _ka_system(16000, 74, 0.06)
    g.set_nh(0.0, 0.45, 0.0)
    g.calc_f(SF)
    prunes = 0
    for k in range(60):
        g.nvt_step(0.0025, on_mask=3, omega2=float((2 * np.pi / 0.2) ** 2), fuse_skin=True)
        fl = g.flags_host()
        if fl[FLAG_PRUNE] and not fl[FLAG_REBUILD]:
            prunes += 1
        g.calc_f(SF)
        if k % 20 == 19:
            _check_inner_lists(g, sysm, 0.06)
    print(f"[case] KA N=16000 60 NVT steps: {prunes} displacement-triggered prunes", flush=True)
    assert prunes >= 2
    # one more step without a prune: the lists are as the last prune left them
    g.nvt_step(0.0025, on_mask=3, omega2=float((2 * np.pi / 0.2) ** 2), fuse_skin=True)
    _check_inner_lists(g, sysm, 0.06, rc_current=not g.flags_host()[FLAG_PRUNE])
    uidx = g.unsorted_index()
    g.calc_f(SF)
    f_in = unsort(g.get("f"), uidx)
    saved = (g.nb.nbl16_in_dev, g.nb.num_in_dev)
    g.nb.nbl16_in_dev = g.nb.num_in_dev = None  # the tile lists themselves
    check(lib().rumdx_pair_lj(C.byref(g.p), C.byref(g.nb), g.d_params.value, g.nt, g.box8, SF, 1, 0, None))
    g.nb.nbl16_in_dev, g.nb.num_in_dev = saved
    f_out = unsort(g.get("f"), uidx)
    pos = unsort(g.get("r"), uidx)
    f_ref, _ = oracle.forces_cells(np.ascontiguousarray(pos), sysm.box, 2, ka_table(), SF)
    e_in, e_out = force_errors(f_in[:, :3], f_ref[:, :3]), force_errors(f_out[:, :3], f_ref[:, :3])
    d_io = force_errors(f_in[:, :3], f_out[:, :3].astype(np.float64))["guarded_max"]
    print(f"[case] after 61 steps: inner {e_in['guarded_max']:.2e}, tile lists {e_out['guarded_max']:.2e}, "
          f"inner vs tile {d_io:.2e}", flush=True)
    # the same pairs summed in other lane groupings: FP32 rounding apart (the
    # lattice start's compressed pairs make it larger than in a liquid)
    assert d_io <= max(2e-6, 2.0 * e_out["guarded_max"])
    assert e_in["guarded_max"] <= max(FORCE_TOL_GUARDED, 1.5 * e_out["guarded_max"]), (e_in, e_out)


def test_host_api_inner_lists_in_graph_blocks():
    """Pot.SetNbListInnerSkin through the host classes: graph blocks equal
    the eager loop bit for bit, and forces after 100 NVT steps (prunes and
    rebuilds inside captured blocks) match the oracle."""
    import rumd_amd.rumd as rumd
    from rumd_amd import conf_io
    sysm = configs.ka_random_lattice(12000, density=1.2, temperature=0.45, seed=75)
    conf = conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities, images=np.zeros_like(sysm.positions))
    out = []
    for mode in ("eager", "graph"):
        s = rumd.Sample(0, 0)
        s.SetVerbose(False)
        s.SetConfiguration(conf)
        pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
        pot.SetVerbose(False)
        for (i, j, sg, ep, rc) in configs.ka_params(2.5):
            pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
        pot.SetNbListSkin(0.3)
        pot.SetNbListInnerSkin(0.1)
        s.SetPotential(pot)
        s.SetFusedSteps(False)  # the graph runs the kernels of the CalcF/Integrate loop
        itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
        s.SetIntegrator(itg)
        print(f"[case] host API inner lists, {mode}", flush=True)
        if mode == "eager":
            for _ in range(100):
                s.CalcF()
                itg.Integrate()
            s.Synchronize()  # RunSteps ends with one (capacities may change there)
        else:
            s.RunSteps(1, 100, 10, 0)
        s.CalcF()
        assert pot.HasNbInnerLists() and abs(pot.GetNbListInnerSkin() - 0.1) < 1e-7
        out.append((s.GetPositions(), s.GetForces(), pot.GetNumberOfNbListBuilds()))
    (r0, f0, b0), (r1, f1, b1) = out
    assert b0 == b1 and np.array_equal(r0, r1) and np.array_equal(f0, f1)
    f_ref, _ = oracle.forces_cells(configs.pack_pos4(r1, sysm.types), sysm.box, 2, ka_table(), SF)
    e = force_errors(f1, f_ref[:, :3])
    assert e["guarded_max"] <= FORCE_TOL_GUARDED, e
