"""Autotune: picks the technical parameters of a Simulation by timing short
runs, and keeps the result in a file keyed by the user's parameters and the
device (the scheme of Python/rumd/Autotune.py:83-448).

Technical parameters, with their ranges:

    tp         [1, 2, 4, 8]  lanes sharing one particle's neighbour list in the
                              force kernels (RUMD's threads per particle)
    waves      [4, 5, 6]     waves per SIMD the cell-tile force kernel is
                              compiled for (sorted lists; the tile mapping's
                              occupancy: more waves hide more latency, fewer
                              leave more registers per lane)
    skin       non-negative float; the neighbour-list skin
    NB_method  ["none", "n2", "sort"]
    pb         kept for RUMD scripts; the kernels here use fixed 256-lane
               workgroups (one cell tile each for sorted lists), so pb only
               sets the particle padding

Typical use, as with RUMD:

    sim = Simulation(...); sim.SetPotential(pot); sim.SetIntegrator(itg)
    at = Autotune()            # or Autotune(skin=0.3) to fix a parameter
    at.Tune(sim)

Tuning runs advance a copy of the state: the configuration (positions in
full, with their FP64 low-order parts in mixed precision), the thermostat
state and the target temperature (which a temperature ramp advances every
step) are restored afterwards, so the user's trajectory is the one it would
have been with the tuned parameters set from the start.

TuneBatch(batch) does the same for a ReplicaBatch (batched replicas, which
RUMD does not have): it times lanes per particle of the batch force kernel
over graph blocks, keyed by the replicas' sizes, boxes, thermostats, time
step, skin and potential, with the batch state saved and restored on the
device around the trials.
"""
from __future__ import annotations

import json
import math
import time
from pathlib import Path

from . import rumd

_METHODS = ("none", "n2", "sort")
_TP = (1, 2, 4, 8)
_WAVES = (4, 5, 6)


class Autotune:
    def __init__(self, **kwargs):
        """Keyword arguments fix a parameter (pb=64, tp=1, skin=0.3,
        NB_method="sort"); ignore_file=True tunes even when the file has a
        match; auto_filename names the file (default autotune.json);
        trial_seconds sets the length of each timed trial."""
        self.auto_filename = Path(kwargs.pop("auto_filename", "autotune.json"))
        self.ignore_file = bool(kwargs.pop("ignore_file", False))
        self.trial_seconds = float(kwargs.pop("trial_seconds", 0.15))
        self.size_stop_no_nb = int(kwargs.pop("size_stop_no_nb", 5000))
        self.size_stop_n2 = int(kwargs.pop("size_stop_n2", 20000))
        self.verbose = bool(kwargs.pop("verbose", False))
        self.fixed = {}
        for key in ("pb", "tp", "skin", "NB_method", "waves"):
            if key in kwargs:
                self.fixed[key] = kwargs.pop(key)
        if kwargs:
            raise TypeError(f"unknown Autotune parameters: {sorted(kwargs)}")
        if "tp" in self.fixed and self.fixed["tp"] not in _TP:
            raise ValueError(f"tp must be one of {_TP}")
        if "NB_method" in self.fixed and self.fixed["NB_method"] not in _METHODS:
            raise ValueError(f"NB_method must be one of {_METHODS}")
        self.trials = []

    # ------------------------------------------------------------- RUMD API
    def Tune(self, sim) -> dict:
        """Use the saved parameters when the user parameters match, else time
        short runs over the ranges and save the best; then set them on sim."""
        user = self.GetUserParameters(sim)
        tech = None if self.ignore_file else self.CompareSavedParameters(user)
        if tech is None:
            tech = self.OptimizeTechnicalParameters(sim)
            self.WriteParametersToFile(user, tech)
        self.SetTechnicalParameters(sim, tech)
        return tech

    def GetUserParameters(self, sim) -> dict:
        """Device, N, composition, density, kinetic temperature, potentials and
        their parameters, integrator type and time step, position precision
        (Python/rumd/Autotune.py:256-314)."""
        s = sim.sample
        n = s.GetNumberOfParticles()
        nt = s.GetNumberOfTypes()
        itg_info = sim.itg.GetInfoString(8).split(",")
        return {
            "deviceName": rumd.Device.GetDevice().GetDeviceName(),
            "N": n,
            "numberEachType": [s.GetNumberThisType(t) for t in range(nt)],
            "rho": n / s.GetSimulationBox().GetVolume(),
            "T": sim.itg.GetKineticEnergy() / n * 2.0 / 3.0,
            "potentialTypes": [p.GetID_String() for p in sim.potentialList],
            "pot_parameters": [list(p.GetParameterList(nt)) if hasattr(p, "GetParameterList") else []
                               for p in sim.potentialList],
            "integratorType": itg_info[0],
            "timeStep": float(itg_info[1]),
            "mixedPrecision": bool(s.GetMixedPrecisionPositions()),
        }

    def GetCurrentTechnicalParameters(self, sim) -> dict:
        pot = sim.potentialList[0]
        return {"pb": sim.sample.GetParticlesPerBlock(), "tp": sim.sample.GetThreadsPerParticle(),
                "skin": pot.GetNbListSkin(), "NB_method": pot.GetNB_Method(),
                "waves": pot.GetNbTileWavesPerEU() or 5}

    def CompareSavedParameters(self, user: dict):
        """The saved technical parameters of an entry whose user parameters
        match (floats to 1e-6 relative, temperature to 10 %), or None."""
        for entry in self._read_file():
            if _same_user(entry.get("user", {}), user):
                return entry["tech"]
        return None

    def SetTechnicalParameters(self, sim, tech: dict) -> None:
        pot = sim.potentialList[0]
        if (sim.sample.GetParticlesPerBlock(), sim.sample.GetThreadsPerParticle()) != (tech["pb"], tech["tp"]):
            sim.SetPB_TP(int(tech["pb"]), int(tech["tp"]))
        pot.SetNB_Method(tech["NB_method"])
        pot.SetNbListSkin(float(tech["skin"]))
        pot.SetNbTileWavesPerEU(int(tech.get("waves", 5)))

    def WriteParametersToFile(self, user: dict, tech: dict) -> None:
        entries = [e for e in self._read_file() if not _same_user(e.get("user", {}), user)]
        entries.append({"user": user, "tech": tech, "trials": self.trials})
        self.auto_filename.write_text(json.dumps(entries, indent=1))

    # ------------------------------------------------------ batched replicas
    def TuneBatch(self, batch, block: int = 10) -> dict:
        """Lanes per particle of a ReplicaBatch's force kernel: the saved
        choice when the user parameters match, else the fastest of
        {1, 2, 4, 8} (or the fixed tp) over timed graph blocks."""
        user = self.GetBatchUserParameters(batch)
        tech = None if self.ignore_file else self.CompareSavedParameters(user)
        if tech is None:
            tps = [self.fixed["tp"]] if "tp" in self.fixed else list(_TP)
            self.trials = []
            batch.SaveState()
            try:
                nsteps = None
                best = None
                for tp in tps:
                    batch.SetLanesPerParticle(int(tp))
                    batch.RunSteps(2 * block, block)  # warm-up: list builds, graph capture
                    if nsteps is None:
                        t0 = time.perf_counter()
                        batch.RunSteps(2 * block, block)
                        per = (time.perf_counter() - t0) / (2 * block)
                        nsteps = max(2 * block, min(2000, block * math.ceil(self.trial_seconds / max(per, 1e-6) / block)))
                    t0 = time.perf_counter()
                    batch.RunSteps(nsteps, block)
                    ms = (time.perf_counter() - t0) * 1e3 / nsteps
                    batch.RestoreState()
                    self.trials.append({"params": {"tp": int(tp)}, "ms_per_step": ms})
                    if self.verbose:
                        print(f"[Autotune] batch tp={tp} -> {ms:.4f} ms/step")
                    if best is None or ms < best[0]:
                        best = (ms, int(tp))
            except BaseException:
                batch.RestoreState()
                raise
            tech = {"tp": best[1]}
            self.WriteParametersToFile(user, tech)
        batch.SetLanesPerParticle(int(tech["tp"]))
        return tech

    def GetBatchUserParameters(self, batch) -> dict:
        nrep = batch.GetNumberOfReplicas()
        return {
            "deviceName": rumd.Device.GetDevice().GetDeviceName(),
            "kind": "ReplicaBatch",
            "N": [batch.GetNumberOfParticles(r) for r in range(nrep)],
            "L": [float(batch.GetBoxLength(r)) for r in range(nrep)],
            "thermostats": [[float(x) for x in batch.GetThermostat(r)] for r in range(nrep)],
            "timeStep": float(batch.GetTimeStep()),
            "skin": float(batch.GetSkin()),
            "cutoffMethod": int(batch.GetCutoffMethod()),
            "params": [float(x) for x in batch.GetParamsTable()],
        }

    # --------------------------------------------------------- optimisation
    def OptimizeTechnicalParameters(self, sim) -> dict:
        """Coordinate search: the list method and tp at the current skin, a
        skin scan at the best of those, then tp once more.  Each trial times
        a run of graph-captured blocks after a warm-up that includes a list
        build; the state is restored after every trial."""
        s = sim.sample
        n = s.GetNumberOfParticles()
        saved = self._save(sim)
        current = self.GetCurrentTechnicalParameters(sim)
        best = dict(current)
        for key, val in self.fixed.items():
            best[key] = val
        if best["NB_method"] == "default":
            best["NB_method"] = "sort" if n > 8000 else "n2"
        methods = [self.fixed["NB_method"]] if "NB_method" in self.fixed else \
            [m for m in _METHODS if (m != "none" or n <= self.size_stop_no_nb) and (m != "n2" or n <= self.size_stop_n2)]
        tps = [self.fixed["tp"]] if "tp" in self.fixed else list(_TP)
        self.trials = []
        nsteps = self._trial_steps(sim, best, saved)

        def trial(cand):
            t = self._time(sim, cand, saved, nsteps)
            self.trials.append({"params": dict(cand), "ms_per_step": t})
            if self.verbose:
                print(f"[Autotune] {cand} -> {t:.4f} ms/step")
            return t

        scored = []
        for m in methods:
            for tp in tps:
                cand = dict(best, NB_method=m, tp=tp)
                scored.append((trial(cand), cand))
        best = min(scored, key=lambda x: x[0])[1]
        if "skin" not in self.fixed and best["NB_method"] != "none":
            s0 = float(best["skin"])
            skins = sorted({round(s0 * f, 4) for f in (0.6, 0.8, 1.0, 1.25, 1.5)})
            scored = [(trial(dict(best, skin=sk)), dict(best, skin=sk)) for sk in skins]
            best = min(scored, key=lambda x: x[0])[1]
            if "tp" not in self.fixed:
                scored = [(trial(dict(best, tp=tp)), dict(best, tp=tp)) for tp in tps]
                best = min(scored, key=lambda x: x[0])[1]
        if "waves" not in self.fixed and best["NB_method"] == "sort":
            scored = [(trial(dict(best, waves=wv)), dict(best, waves=wv)) for wv in _WAVES]
            best = min(scored, key=lambda x: x[0])[1]
        self._restore(sim, saved)
        return best

    # --------------------------------------------------------------- helpers
    def _read_file(self):
        if not self.auto_filename.exists():
            return []
        try:
            data = json.loads(self.auto_filename.read_text())
        except (OSError, ValueError):
            return []
        return data if isinstance(data, list) else []

    @staticmethod
    def _save(sim):
        """What tuning trials change besides the technical parameters: the
        configuration, the FP64 positions, Ps and the target temperature."""
        s = sim.sample
        itg = sim.itg
        return {
            "conf": s.GetConfiguration(),
            "pos64": s.GetPositionsFP64(),
            "ps": itg.GetThermostatState() if hasattr(itg, "GetThermostatState") else None,
            "T": itg.GetTargetTemperature() if hasattr(itg, "GetTargetTemperature") else None,
        }

    @staticmethod
    def _restore(sim, saved):
        sim.sample.SetConfiguration(saved["conf"])
        sim.sample.SetPositionsFP64(saved["pos64"])
        if saved["T"] is not None:
            sim.itg.SetTargetTemperature(saved["T"], False)
        if saved["ps"] is not None:
            sim.itg.SetThermostatState(saved["ps"])

    def _trial_steps(self, sim, params, saved) -> int:
        """Steps per trial for about trial_seconds, from one short probe."""
        probe = 20
        t = self._time(sim, params, saved, probe)  # ms per step
        steps = int(self.trial_seconds * 1e3 / max(t, 1e-3))
        return max(40, min(4000, 10 * math.ceil(steps / 10)))

    def _time(self, sim, params, saved, nsteps) -> float:
        self.SetTechnicalParameters(sim, params)
        self._restore(sim, saved)
        s = sim.sample
        s.RunSteps(0, 20, sim.graph_block, 0)  # warm-up: list build, graph capture
        s.Synchronize()
        t0 = time.perf_counter()
        s.RunSteps(20, nsteps, sim.graph_block, 0)
        return (time.perf_counter() - t0) * 1e3 / nsteps


def _same_user(a: dict, b: dict) -> bool:
    if set(a) != set(b):
        return False
    for k, va in a.items():
        vb = b[k]
        if k == "T":
            if abs(va - vb) > 0.1 * max(abs(vb), 1e-12):
                return False
        elif not _close(va, vb):
            return False
    return True


def _close(a, b) -> bool:
    if isinstance(a, (list, tuple)) and isinstance(b, (list, tuple)):
        return len(a) == len(b) and all(_close(x, y) for x, y in zip(a, b))
    if isinstance(a, float) or isinstance(b, float):
        try:
            return abs(float(a) - float(b)) <= 1e-6 * max(abs(float(a)), abs(float(b)), 1e-30)
        except (TypeError, ValueError):
            return False
    return a == b
