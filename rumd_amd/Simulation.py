"""Simulation driver with the run-loop semantics of RUMD's
Python/rumd/Simulation.py: per step, runtime actions (momentum reset every
100 steps by default), CalcF, Integrate; one more CalcF after the loop so the
forces match the final positions; timing reported as MATS (millions of
atom-timesteps per second).

When no Python runtime action other than the momentum reset is registered, the
loop runs as HIP-graph-captured blocks of steps (Sample.RunSteps); otherwise it
steps from Python exactly like the reference.
"""
from __future__ import annotations

import time
import weakref
from collections import OrderedDict

from . import conf_io
from . import rumd


class Simulation:
    def __init__(self, filename_or_conf, pb: int = 0, tp: int = 0, verbose: bool = True,
                 mixed_precision_positions: bool = False):
        self.sample = rumd.Sample(pb, tp)
        self.sample.SetVerbose(verbose)
        if mixed_precision_positions:
            self.sample.SetMixedPrecisionPositions(True)
        if isinstance(filename_or_conf, conf_io.Configuration):
            self.sample.SetConfiguration(filename_or_conf)
        else:
            self.sample.ReadConf(filename_or_conf)
        self.itg = None
        self.potentialList = []
        self.verbose = verbose
        self.block_size = 10
        self.runtime_actions = {}
        self.runtime_action_intervals = OrderedDict()
        self.runtime_action_intervals["momentum_reset"] = 100
        self.elapsed = 0.0
        self.last_mats = None

    # ------------------------------------------------------------ set-up
    def SetVerbose(self, vb: bool) -> None:
        self.verbose = vb
        self.sample.SetVerbose(vb)

    def SetPotential(self, potential) -> None:
        self.potentialList = [potential]
        self.sample.SetPotential(potential)

    def AddPotential(self, potential) -> None:
        self.potentialList.append(potential)
        self.sample.AddPotential(potential)

    def SetIntegrator(self, itg) -> None:
        self.itg = itg
        self.sample.SetIntegrator(itg)

    def SetPB_TP(self, pb: int, tp: int) -> None:
        self.sample.SetPB_TP(pb, tp)

    def SetBlockSize(self, block: int) -> None:
        """Steps per captured HIP graph in the fast path."""
        self.block_size = max(1, int(block))

    def SetMomentumResetInterval(self, interval: int) -> None:
        self.runtime_action_intervals["momentum_reset"] = int(interval)

    def SetRuntimeAction(self, name: str, method, interval: int) -> None:
        self.runtime_actions[name] = (weakref.ref(method.__self__), weakref.ref(method.__func__))
        self.runtime_action_intervals[name] = int(interval)

    def RemoveRuntimeAction(self, name: str) -> None:
        self.runtime_actions.pop(name, None)
        self.runtime_action_intervals.pop(name, None)

    def ResetMomentum(self) -> None:
        self.sample.GetIntegrator().SetMomentumToZero()

    def WriteConf(self, filename, mode: str = "w") -> None:
        self.sample.WriteConf(filename, mode)

    # ------------------------------------------------------------ run loop
    def ApplyRuntimeActions(self, i: int) -> None:
        for key, interval in self.runtime_action_intervals.items():
            if interval <= 0 or i % interval != 0:
                continue
            if key == "momentum_reset":
                self.ResetMomentum()
            else:
                obj, func = (r() for r in self.runtime_actions[key])
                func(obj)

    def Run(self, num_iter: int, initializeOutput: bool = True, restartBlock=None, suppressAllOutput: bool = False,
            force_timing: bool = False) -> float:
        if self.itg is None:
            raise ValueError("Integrator has not been set")
        if not self.potentialList:
            raise ValueError("Potential has not been set")
        if restartBlock is not None:
            raise ValueError("restart blocks need the output managers, which this build does not provide")
        t0 = time.perf_counter()
        python_actions = [k for k in self.runtime_action_intervals
                          if k != "momentum_reset" and self.runtime_action_intervals[k] > 0]
        if python_actions:
            for i in range(num_iter):
                self.ApplyRuntimeActions(i)
                self.sample.CalcF()
                self.itg.Integrate()
            self.sample.Synchronize()
        else:
            self.sample.RunSteps(0, int(num_iter), self.block_size,
                                 max(0, self.runtime_action_intervals.get("momentum_reset", 0)))
        self.elapsed = time.perf_counter() - t0
        self.last_mats = self.sample.GetNumberOfParticles() * num_iter / 1e6 / max(self.elapsed, 1e-12)
        if (not suppressAllOutput or force_timing) and self.verbose:
            print(f"{num_iter} steps, {self.elapsed:.3f} s, {self.last_mats:.3f} MATS")
        # forces consistent with the final positions (Simulation.py:610-614)
        self.sample.CalcF()
        return self.elapsed
