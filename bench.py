#!/usr/bin/env python3
"""Benchmark: atom-timesteps/s of the RUMD hot path on one MI355X.

Workload (one GPU): Kob-Andersen 80:20 binary LJ, N = 1,048,576 (FCC 64^3
start, density 1.2, box L = 95.6033), shifted-force cut-off rc = 2.5 sigma_ab,
Nose-Hoover NVT T = 0.45, tau = 0.2, dt = 0.0025, neighbour-list skin 0.3,
cell-sorted (Morton) list rebuilt by the skin criterion.  Precision: FP32 pair
forces; FP64 energy/virial/kinetic sums, thermostat state and position
bookkeeping ("mixed").

A step is CalcF + Integrate (with the device-side skin check and, when it
fires, the sort + list rebuild; the forces and the NVT update run as one
fused kernel, rumdx_pair_lj_nvt), run as HIP-graph-captured blocks of 10 steps
that start at the first timed step; momentum is reset every 100 steps (at
absolute step multiples of 100) as RUMD's Simulation.Run does.  Set-up runs
E equilibration steps from the lattice (the liquid the timed steps measure),
then W warm-up steps; the graphs of the timed steps are captured and
instantiated after the warm-up, before the timer starts; the K timed steps
are bracketed by stream synchronisation.

Also reported:
  roofline     the fused step kernel's algorithmic HBM bytes per launch
               (forces in RUMD's layout, 16 + 4 + 4 nbar + 16 + 16 bytes per
               particle, plus the update's 96: v, low position parts, last
               build positions, new positions) over its average duration,
               timed with HIP events on the sample's stream; the force kernel
               alone under roofline.force_only;
  cpu_baseline the FP64 CPU MD of the oracle (same algorithm: Verlet list with
               skin, cells, Nose-Hoover leapfrog) on the same start, on all host
               cores and on one core, over bounded samples.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def build_system(n_cells: int, seed: int):
    from rumd_amd import configs, conf_io
    s = configs.ka_random_lattice(4 * n_cells ** 3, density=1.2, temperature=0.45, seed=seed, lattice="fcc")
    conf = conf_io.Configuration(n=s.n, box=s.box, types=s.types, positions=s.positions, masses=s.masses,
                                 velocities=s.velocities, images=np.zeros_like(s.positions))
    return s, conf


def make_sample(conf, mixed: bool):
    import rumd_amd.rumd as rumd
    from rumd_amd import configs
    sample = rumd.Sample(0, 0)
    sample.SetVerbose(False)
    sample.SetMixedPrecisionPositions(mixed)
    sample.SetConfiguration(conf)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
    pot.SetVerbose(False)
    for (i, j, sg, ep, rc) in configs.ka_params(2.5):
        pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
    pot.SetNbListSkin(0.3)
    sample.SetPotential(pot)
    itg = rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=0.45, thermostatRelaxationTime=0.2)
    sample.SetIntegrator(itg)
    return sample, pot, itg


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(system, budget_s: float, max_steps: int):
    """FP64 CPU MD (oracle) on the same start; steps bounded by a time budget,
    on all host threads and on one thread."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from rumd_amd import configs
    table = configs.lj_param_table(configs.ka_params(2.5), 2)
    threads = oracle.num_threads()
    inv_m = np.ones(system.n)
    omega2 = float(np.float32(4 * np.pi ** 2 / 0.2 ** 2))

    def timed(budget, cap):
        pos = system.positions.astype(np.float64).copy()
        vel = system.velocities.astype(np.float64).copy()
        # one short run to size the sample (includes a list build)
        t0 = time.perf_counter()
        _, _, ps = oracle.md_run(pos, vel, system.types, inv_m, system.box, 2, table, 2, 0.3, 0.0025, 2, True,
                                 0.45, omega2, 0.0)
        t_probe = time.perf_counter() - t0
        steps = int(max(3, min(cap, budget / max(t_probe / 2.0, 1e-6))))
        t0 = time.perf_counter()
        builds, _, _ = oracle.md_run(pos, vel, system.types, inv_m, system.box, 2, table, 2, 0.3, 0.0025, steps,
                                     True, 0.45, omega2, ps)
        dt = time.perf_counter() - t0
        return system.n * steps / dt, steps, builds

    value, steps, builds = timed(budget_s, max_steps)
    oracle.set_num_threads(1)
    try:
        value1, steps1, builds1 = timed(budget_s / 2.0, max_steps)
    finally:
        oracle.set_num_threads(threads)
    return {
        "value": value,
        "unit": "atom-timesteps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"KA N={system.n}, {steps} NVT steps (FP64, Verlet list skin 0.3, {builds} list builds incl. "
                  f"the first), same start as the GPU run, {threads} OpenMP threads",
        "value_1core": value1,
        "sample_1core": f"{steps1} NVT steps ({builds1} list builds) on 1 thread",
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
    }


def read_traffic(profile_dir: Path, name: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE + WRITE_SIZE, each calibrated
    against kernels of known traffic); None when absent."""
    f = profile_dir / name
    if not f.exists():
        return None, None
    try:
        d = json.loads(f.read_text())
        return float(d["hbm_bytes_per_launch"]), str(d.get("kernel", ""))
    except (KeyError, ValueError):
        return None, None


def counter_calibration(n: int, reps: int) -> dict:
    """Kernels of known HBM traffic for scaling the rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters to bytes: k_reduce_f4 reads 16 n bytes, and
    k_momentum_apply reads and writes 16 n bytes (n float4 far larger than the
    caches).  Run under rocprofv3 --pmc together with the force kernel."""
    import ctypes as C
    from rumd_amd import _native
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    lib = _native.librumdx()

    class Particles(C.Structure):
        _fields_ = [(f, C.c_void_p) for f in ("r", "v", "f", "w", "im", "sts", "rlo", "idx")] + \
                   [("n", C.c_uint32), ("nvp", C.c_uint32)]

    lib.rumdx_workspace_bytes.restype = C.c_size_t
    lib.rumdx_workspace_bytes.argtypes = [C.c_uint32]
    lib.rumdx_reduce_f4.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.rumdx_zero_momentum.argtypes = [C.POINTER(Particles), C.c_void_p, C.c_void_p]
    arr, ws, out = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nbytes = 16 * n
    for ptr, size in ((arr, nbytes), (ws, lib.rumdx_workspace_bytes(n)), (out, 64)):
        if hip.hipMalloc(C.byref(ptr), size):
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(arr, 0, nbytes)
    p = Particles()
    p.v, p.n, p.nvp = arr.value, n, n
    for _ in range(reps):
        if lib.rumdx_reduce_f4(arr, n, 15, out, ws, None) or lib.rumdx_zero_momentum(C.byref(p), ws, None):
            raise RuntimeError("calibration kernel failed")
    hip.hipDeviceSynchronize()
    for ptr in (arr, ws, out):
        hip.hipFree(ptr)
    return {"n": n, "k_reduce_f4_read_bytes": nbytes, "k_momentum_apply_read_bytes": nbytes,
            "k_momentum_apply_write_bytes": nbytes}


def force_accuracy(sample, pot, system):
    """Per-particle force error of the FP32 kernel and of the all-FP64 kernel
    against the FP64 oracle (linked cells), all three on the same float
    positions (a sample without the FP64 low parts), so the errors come from
    the force arithmetic alone."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from rumd_amd import configs
    sample.Synchronize()
    table = configs.lj_param_table(configs.ka_params(2.5), 2)
    box = system.box.astype(np.float32).astype(np.float64)
    pos4 = configs.pack_pos4(sample.GetPositions(), system.types)
    f_ref = oracle.forces_cells(pos4, box, 2, table, 2)[0][:, :3]
    mag = np.linalg.norm(f_ref, axis=1)
    rms = float(np.sqrt(np.mean(mag ** 2)))
    out = {}
    for label, fp64 in (("fp32_kernel", False), ("fp64_kernel", True)):
        pot.SetFP64Forces(fp64)
        sample.CalcF()
        d = np.linalg.norm(sample.GetForces().astype(np.float64) - f_ref, axis=1)
        guarded = d / np.maximum(mag, rms)
        out[label] = {"guarded_max": float(guarded.max()), "guarded_p99": float(np.quantile(guarded, 0.99)),
                      "unguarded_median": float(np.median(d / np.maximum(mag, 1e-30)))}
    pot.SetFP64Forces(False)
    sample.CalcF()
    return out


def nve_drift(sample, nsteps: int, dt: float, nh_steps: int):
    """Total energy per particle over nsteps NVE steps from the current state,
    which the caller has equilibrated with nh_steps Nose-Hoover steps
    (momentum reset off): (E_end - E_start) / N per 1000 steps."""
    import rumd_amd.rumd as rumd
    nve = rumd.IntegratorNVE(timeStep=dt)
    sample.SetIntegrator(nve)
    n = sample.GetNumberOfParticles()
    sample.CalcF()
    e0 = (sample.GetPotentialEnergy() + nve.GetKineticEnergy()) / n
    sample.RunSteps(0, nsteps, 10, 0)
    sample.CalcF()
    e1 = (sample.GetPotentialEnergy() + nve.GetKineticEnergy()) / n
    return {"e_start": e0, "e_end": e1, "drift_per_1000_steps": (e1 - e0) * 1000.0 / nsteps, "steps": nsteps,
            "after_nh_steps": nh_steps}


def make_batch(systems, temps, dt: float, tau: float, skin: float):
    import rumd_amd.rumd as rumd
    from rumd_amd import configs
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
    return b


def batched_leg(nrep: int, warmup: int, steps: int, sample_reps: int, single_steps: int):
    """KA replicas (N ~ U[256, 8192], T ~ U[0.4, 2.0], rho 1.2, SF, NH tau 0.2,
    dt 0.0025, skin 0.3): all of them in one ReplicaBatch, against the same
    replicas run one at a time through the single-sample engine (a Sample
    each, RUMD's default list method), timed on an evenly spaced subset."""
    import rumd_amd.rumd as rumd
    from rumd_amd import configs, conf_io
    systems, temps = configs.ka_replica_set(nrep)
    b = make_batch(systems, temps, 0.0025, 0.2, 0.3)
    ntot = b.GetTotalParticles()
    b.RunSteps(warmup, 10)
    # lanes per particle of the batch force kernel, tuned on this state (the
    # trials run between a device-side save and restore of the batch state)
    import tempfile
    from rumd_amd.Autotune import Autotune
    with tempfile.TemporaryDirectory() as td:
        at = Autotune(auto_filename=Path(td) / "autotune.json", ignore_file=True, trial_seconds=0.15)
        tuned = at.TuneBatch(b, block=10)
    trials = ", ".join("%d: %.3f" % (t["params"]["tp"], t["ms_per_step"]) for t in at.trials)
    print(f"[bench] batch autotune: lanes per particle {tuned['tp']} ({trials} ms/step)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.RunSteps(steps, 10)
    el = time.perf_counter() - t0
    batched = ntot * steps / el
    print(f"[bench] batched {nrep} replicas, N_total={ntot}: {steps} steps in {el:.3f} s -> "
          f"{batched / 1e9:.3f} G atom-steps/s", file=sys.stderr, flush=True)
    # the batch force kernel against the HBM roofline: RUMD-layout bytes per
    # particle (16 + 4 + 4 nbar + 16 + 16) over its event-timed duration
    nbar_b, _ = b.GetNeighborListStats()
    fms = b.TimeForceKernel(20)
    bpp = 16 + 4 + 4 * nbar_b + 16 + 16
    ach = bpp * ntot / (fms * 1e-3) / 1e9
    batch_roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(ach / HBM_PEAK_GBS, 4),
                      "kernel": ("k_b_force_staged (batched LJ force, replicas staged in LDS)" if b.UsesStagedForces()
                                 else "k_b_force (batched LJ force, neighbours gathered through L2)"),
                      "kernel_ms": round(fms, 5), "bytes_per_atom": round(bpp, 2), "nbar": round(nbar_b, 2)}
    print(f"[bench] batch force kernel {fms * 1e3:.1f} us, {ach:.0f} GB/s", file=sys.stderr, flush=True)
    picks = list(range(0, nrep, max(1, nrep // sample_reps)))[:sample_reps]
    atom_steps, seconds = 0.0, 0.0
    for r in picks:
        s = systems[r]
        sample = rumd.Sample(0, 0)
        sample.SetVerbose(False)
        sample.SetMixedPrecisionPositions(False)
        sample.SetConfiguration(conf_io.Configuration(n=s.n, box=s.box, types=s.types, positions=s.positions,
                                                      masses=s.masses, velocities=s.velocities,
                                                      images=np.zeros_like(s.positions)))
        pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedForce)
        pot.SetVerbose(False)
        for (i, j, sg, ep, rc) in configs.ka_params(2.5):
            pot.SetParams(i, j, Sigma=sg, Epsilon=ep, Rcut=rc)
        pot.SetNbListSkin(0.3)
        sample.SetPotential(pot)
        sample.SetIntegrator(rumd.IntegratorNVT(timeStep=0.0025, targetTemperature=temps[r],
                                                thermostatRelaxationTime=0.2))
        sample.RunSteps(0, 20, 10, 0)
        t0 = time.perf_counter()
        sample.RunSteps(20, single_steps, 10, 0)
        seconds += time.perf_counter() - t0
        atom_steps += s.n * single_steps
    single = atom_steps / seconds
    print(f"[bench] one at a time ({len(picks)} replicas sampled): {single / 1e9:.4f} G atom-steps/s",
          file=sys.stderr, flush=True)
    return {
        "value": round(batched, 1),
        "unit": "atom-timesteps/s",
        "replicas": nrep,
        "particles": int(ntot),
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(el / steps * 1e3, 4),
        "one_at_a_time": {"value": round(single, 1), "replicas_timed": len(picks), "steps_each": single_steps,
                          "how": "a Sample per replica (default list method), graph blocks of 10, timed per replica "
                                 "and aggregated over an evenly spaced subset"},
        "speedup": round(batched / single, 2),
        "roofline": batch_roofline,
        "lanes_per_particle": tuned["tp"],
        "autotune_ms_per_step": {str(t["params"]["tp"]): round(t["ms_per_step"], 4) for t in at.trials},
        "workload": "KA replicas N~U[256,8192] (seed 4001), T~U[0.4,2.0] (seed 4002), rho=1.2, SF rc=2.5, "
                    "NH tau=0.2, dt=0.0025, skin=0.3, FP32 positions",
    }


def polymer_leg(warmup: int, steps: int, block: int) -> dict:
    """The polymer workload: 10^4 chains of 10 beads (N = 100,000, rho = 1.0),
    WCA pairs (LJ cut and shifted at 2^(1/6)) without the 1-2 pairs, harmonic
    bonds k = 3000, l0 = 1, cos^2 angles k = 10 with theta0 = pi (the workload
    does not fix theta0), NVT T = 1.0, tau = 0.2, dt = 0.002; chains start
    straight on a lattice."""
    import rumd_amd.rumd as rumd
    from rumd_amd import configs, conf_io
    sysm, bonds, angles = configs.polymer_melt()
    conf = conf_io.Configuration(n=sysm.n, box=sysm.box, types=sysm.types, positions=sysm.positions,
                                 masses=sysm.masses, velocities=sysm.velocities, images=np.zeros_like(sysm.positions))
    s = rumd.Sample(0, 0)
    s.SetVerbose(False)
    s.SetConfiguration(conf)
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetVerbose(False)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.0 ** (1.0 / 6.0))
    pot.SetNbListSkin(0.3)
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
    ang.SetExclude(False)
    s.AddPotential(bond)
    s.AddPotential(ang)
    bond.SetExclusions(pot)
    itg = rumd.IntegratorNVT(timeStep=0.002, targetTemperature=1.0, thermostatRelaxationTime=0.2)
    s.SetIntegrator(itg)
    s.RunSteps(0, warmup, block, 100)
    s.Synchronize()
    b0 = pot.GetNumberOfNbListBuilds()
    t0 = time.perf_counter()
    s.RunSteps(warmup, steps, block, 100)
    s.Synchronize()
    el = time.perf_counter() - t0
    n = sysm.n
    out = {
        "value": round(n * steps / el, 1), "unit": "atom-timesteps/s", "ms_per_step": round(el / steps * 1e3, 5),
        "workload": "10^4 chains x 10 beads N=100000 rho=1.0: WCA without 1-2 pairs, harmonic bonds k=3000 l0=1, "
                    "cos^2 angles k=10 theta0=pi, NVT T=1.0 tau=0.2 dt=0.002, sorted lists skin 0.3",
        "steps": steps, "warmup": warmup, "list_builds": pot.GetNumberOfNbListBuilds() - b0,
        "pe_per_particle": round(s.GetPotentialEnergy() / n, 6),
        "bond_energy_per_particle": round(bond.GetPotentialEnergy() / n, 6),
        "angle_energy_per_particle": round(ang.GetPotentialEnergy() / n, 6),
        "temperature": round(2.0 * itg.GetKineticEnergy() / (3 * n - 3), 4),
    }
    print(f"[bench] polymer: {out['value'] / 1e9:.3f} G atom-steps/s", file=sys.stderr, flush=True)
    return out


def timed_run(sample, pot, args, label: str) -> dict:
    """Equilibration and warm-up (untimed), the timed steps' graphs captured
    and instantiated, then K timed steps bracketed by stream synchronisation."""
    n = sample.GetNumberOfParticles()
    t0 = time.perf_counter()
    half = args.equilibrate // 2
    sample.RunSteps(0, half, args.block, 100)
    b_half = pot.GetNumberOfNbListBuilds()
    sample.RunSteps(half, args.equilibrate - half, args.block, 100)
    # rebuild interval of the liquid, over the second half of the equilibration
    eq_builds = pot.GetNumberOfNbListBuilds() - b_half
    kbar_eq = (args.equilibrate - half) / eq_builds if eq_builds else None
    print(f"[bench] {label}: equilibration {args.equilibrate} steps in {time.perf_counter() - t0:.2f} s "
          f"(second half: {eq_builds} list builds)", file=sys.stderr, flush=True)
    first = args.equilibrate
    t0 = time.perf_counter()
    sample.RunSteps(first, args.warmup, args.block, 100)
    print(f"[bench] {label}: warm-up {args.warmup} steps in {time.perf_counter() - t0:.2f} s", file=sys.stderr,
          flush=True)
    first += args.warmup
    # every graph the timed steps launch is captured and instantiated now
    t0 = time.perf_counter()
    made = sample.PrepareRunSteps(first, args.steps, args.block, 100)
    segments = sample.RunSegments(first, args.steps, args.block, 100)
    eager = sum(k for k, _ in segments if k < 2)
    captured = args.steps - eager
    print(f"[bench] {label}: {made} graphs captured before timing in {time.perf_counter() - t0:.3f} s; timed "
          f"steps: {captured} in {sum(1 for k, _ in segments if k >= 2)} graph launches, {eager} eager",
          file=sys.stderr, flush=True)
    builds0 = pot.GetNumberOfNbListBuilds()
    sample.Synchronize()
    t0 = time.perf_counter()
    sample.RunSteps(first, args.steps, args.block, 100)  # synchronises at the end
    elapsed = time.perf_counter() - t0
    builds = pot.GetNumberOfNbListBuilds() - builds0
    value = n * args.steps / elapsed
    print(f"[bench] {label}: timed {args.steps} steps in {elapsed:.4f} s -> {value / 1e9:.3f} G atom-steps/s "
          f"({builds} list builds)", file=sys.stderr, flush=True)
    return {"value": value, "elapsed": elapsed, "builds": builds, "captured": captured, "eager": eager,
            "kbar_equilibration": kbar_eq, "nh_steps": first + args.steps}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-counters", type=int, default=0, metavar="N",
                    help="first run kernels of known traffic over N float4 (for rocprofv3 --pmc scaling)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--equilibrate", type=int, default=1000,
                    help="NVT steps from the lattice before the warm-up (set-up: the timed steps run on the liquid)")
    ap.add_argument("--cells", type=int, default=64, help="FCC cells per axis (N = 4 cells^3)")
    ap.add_argument("--precision", choices=["mixed", "fp32"], default="mixed")
    ap.add_argument("--block", type=int, default=10)
    ap.add_argument("--seed", type=int, default=3001)
    ap.add_argument("--force-reps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the headline run (no accuracy, NVE, FP32, batch)")
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--single-replicas", type=int, default=32, help="replicas timed one at a time")
    ap.add_argument("--polymer-only", action="store_true",
                    help="run only the polymer workload and print its figures (for profiling)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched-replica workload and print its figures (for profiling)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(the single large system would shard by spatial domains, see the design notes).", file=sys.stderr)
        return 2

    if args.polymer_only:
        print(json.dumps({"polymer": polymer_leg(500, args.steps, args.block)}), flush=True)
        return 0

    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args.replicas, 100, args.batch_steps, args.single_replicas, 100)}),
              flush=True)
        return 0

    if args.calibrate_counters > 0:
        print(json.dumps({"calibration": counter_calibration(args.calibrate_counters, 3)}), file=sys.stderr, flush=True)

    system, conf = build_system(args.cells, args.seed)
    sample, pot, itg = make_sample(conf, args.precision == "mixed")
    n = sample.GetNumberOfParticles()
    print(f"[bench] KA N={n} L={system.box[0]:.5f} precision={args.precision}", file=sys.stderr, flush=True)

    run = timed_run(sample, pot, args, "headline")
    value, elapsed, builds = run["value"], run["elapsed"], run["builds"]
    captured, eager = run["captured"], run["eager"]

    nbar, nmax = pot.GetNeighborListStats()
    fused = sample.UsesFusedSteps()
    force_ms = pot.TimeForceKernel(args.force_reps)
    fused_ms = sample.TimeFusedStepKernel(args.force_reps) if fused else None
    build_ms = pot.TimeNbListBuild(10)
    # RUMD's layout: r_i, the list length, u32 entries, f_i and w_i written
    force_bytes = 16 + 4 + 4 * nbar + 16 + 16
    # the fused step's update: v and (mixed) the low position parts read and
    # written, the last build's positions read, the new positions written
    nvt_bytes = 16 * (2 + 2 + 1 + 1) if args.precision == "mixed" else 16 * (2 + 1 + 1)

    def hbm_roofline(bytes_per_atom, ms, pmc_file, kernel):
        achieved = bytes_per_atom * n / (ms * 1e-3) / 1e9
        traffic, _ = read_traffic(ROOT / "profiles", pmc_file)
        return {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": f"committed rocprofv3 PMC summary of this workload (profiles/{pmc_file}: "
                              "FETCH_SIZE + WRITE_SIZE per launch, calibrated); not measured in this run",
            "kernel": kernel,
            "kernel_ms": round(ms, 5),
            "bytes_per_atom": round(bytes_per_atom, 2),
        }

    force_roofline = hbm_roofline(force_bytes, force_ms, "force_kernel_ka1m_pmc.json",
                                  "k_pair_tiles<LjPair<SF>> (LJ force over the cell-tile lists)")
    if fused:
        # the step's dominant kernel: forces and the NVT update in one launch
        # (timed with the thermostat's kinetic-energy sum off)
        roofline = hbm_roofline(force_bytes + nvt_bytes, fused_ms, "fused_step_ka1m_pmc.json",
                                "k_pair_tiles<LjPair<SF>, fused NVT step> (forces and the leapfrog/Nose-Hoover "
                                "update over the cell-tile lists)")
        roofline["force_only"] = force_roofline
    else:
        roofline = force_roofline
    # steps per rebuild: over the timed steps, and over the second half of the
    # equilibration (a longer window; the byte model uses it)
    kbar = args.steps / builds if builds else None
    kbar_model = run["kbar_equilibration"] or kbar
    rebuild_bytes = 4 * nbar + 260
    if fused:
        step_bytes = force_bytes + nvt_bytes + (rebuild_bytes / kbar_model if kbar_model else 0.0)
        model_terms = (f"fused force+NVT 16+4+4nbar+16+16 + {nvt_bytes} (skin check included), "
                       "rebuild (4nbar+260)/kbar")
    else:
        step_bytes = force_bytes + (64 + 48 + 4) + 32 + (rebuild_bytes / kbar_model if kbar_model else 0.0)
        model_terms = "force 16+4+4nbar+16+16, NVT 64+48+4, skin check 32, rebuild (4nbar+260)/kbar"
    pe = sample.GetPotentialEnergy() / n
    ke = itg.GetKineticEnergy() / n

    extras = {}
    if not args.no_extras:
        print("[bench] NVE drift ...", file=sys.stderr, flush=True)
        drift = {args.precision: nve_drift(sample, 1000, 0.0025, run["nh_steps"])}
        if args.precision == "fp32":
            print("[bench] force accuracy vs the FP64 oracle ...", file=sys.stderr, flush=True)
            extras["force_error_vs_fp64_oracle"] = force_accuracy(sample, pot, system)
        if args.precision == "mixed":
            s32, p32, _ = make_sample(conf, False)
            r32 = timed_run(s32, p32, args, "FP32 positions")
            extras["fp32_positions"] = {"value": round(r32["value"], 1), "unit": "atom-timesteps/s",
                                        "ms_per_step": round(r32["elapsed"] / args.steps * 1e3, 5),
                                        "list_builds": r32["builds"]}
            print("[bench] force accuracy vs the FP64 oracle ...", file=sys.stderr, flush=True)
            extras["force_error_vs_fp64_oracle"] = force_accuracy(s32, p32, system)
            drift["fp32"] = nve_drift(s32, 1000, 0.0025, r32["nh_steps"])
            del s32, p32
        extras["nve_drift"] = drift
        print("[bench] batched replicas ...", file=sys.stderr, flush=True)
        extras["batched"] = batched_leg(args.replicas, 100, args.batch_steps, args.single_replicas, 100)
        print("[bench] polymer melt ...", file=sys.stderr, flush=True)
        extras["polymer"] = polymer_leg(500, 2000, args.block)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(system, args.cpu_budget, 400)

    line = {
        "metric": "atom-timesteps/s (KA LJ N=1M) mixed vs FP32; batched replica atom-steps/s",
        "value": round(value, 1),
        "unit": "atom-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces, f64 sums/thermostat/positions" if args.precision == "mixed" else "f32",
        "data": "synthetic (seeded FCC start, Maxwell velocities)",
        "config": {
            "workload": f"Kob-Andersen 80:20 LJ N={n} rho=1.2 shifted-force rc=2.5sigma NVT T=0.45 tau=0.2 "
                        f"dt=0.0025 skin=0.3, Morton cell-sorted list",
            "n": n,
            "precision": args.precision,
            "graph_block_steps": args.block,
            "nbar": round(nbar, 2),
            "nmax": int(nmax),
            "list_builds": builds,
            "kbar": round(kbar, 2) if kbar else None,
            "kbar_equilibration": round(run["kbar_equilibration"], 2) if run["kbar_equilibration"] else None,
            "halo_capacity": pot.GetNbHaloCapacity(),
            "list_rebuild_ms": round(build_ms, 4),
            "equilibration_steps": args.equilibrate,
            "timed_steps_captured": captured,
            "timed_steps_eager": eager,
            "model_bytes_per_atom_step": round(step_bytes, 1),
            "model_terms": model_terms,
            "fused_steps": fused,
            "model_hbm_fraction_step": round(value * step_bytes / 1e9 / HBM_PEAK_GBS, 4),
            "pe_per_particle": round(pe, 6),
            "ke_per_particle": round(ke, 6),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extras)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
