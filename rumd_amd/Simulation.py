"""Simulation driver with the run-loop semantics of RUMD's
Python/rumd/Simulation.py: per step, runtime actions (momentum reset every
100 steps by default), CalcF, the output managers, Integrate; one more CalcF
after the loop so the forces match the final positions; output scheduling
(energies linear every 256 steps, trajectory logarithmic, restart at every
block start) and the output block size chosen as the reference chooses it;
timing reported as MATS (millions of atom-timesteps per second).

Steps at which nothing happens on the host run as HIP-graph-captured blocks
(Sample.RunSteps, which applies the momentum reset itself).  A step at which an
output manager writes or a Python runtime action fires runs eagerly, in the
reference's order.
"""
from __future__ import annotations

import time
import weakref
from collections import OrderedDict

from . import conf_io
from . import rumd


class Simulation:
    def __init__(self, filename_or_conf, pb: int = 0, tp: int = 0, verbose: bool = True,
                 mixed_precision_positions: bool = False, timing_filename: str | None = "rumd_timing_info.dat"):
        self.sample = rumd.Sample(pb, tp)
        self.sample.SetVerbose(verbose)
        if mixed_precision_positions:
            self.sample.SetMixedPrecisionPositions(True)
        if isinstance(filename_or_conf, conf_io.Configuration):
            self.sample.SetConfiguration(filename_or_conf)
        else:
            self.sample.ReadConf(filename_or_conf)
        self.itg = None
        self.potentialList = []  # non-molecular potentials
        self.moleculePotentials = {}
        self.verbose = verbose
        self.graph_block = 10
        self.blockSize = None
        self.output_is_initialized = False
        # (Simulation.py:43-46)
        self.outputParameters = OrderedDict()
        self.outputParameters["energies"] = ("linear", 256)
        self.outputParameters["trajectory"] = ("logarithmic", 1)
        self.allowedManagers = ["energies", "trajectory"]
        self.outputDirectory = self.sample.GetOutputDirectory()
        self.nominalNumBlocks = 100
        self.runtime_actions = {}
        self.runtime_action_intervals = OrderedDict()
        self.runtime_action_intervals["momentum_reset"] = 100
        self.runtime_action_intervals["sort"] = 0  # Sample.SortParticles
        self.elapsed = 0.0
        self.last_mats = None
        # timing file written at construction and after each run (Simulation.py:73-79)
        self.timing_filename = timing_filename
        self.write_timing_info = timing_filename is not None
        if self.write_timing_info:
            with open(self.timing_filename, "w") as fh:
                fh.write("# %s \n" % rumd.Device.GetDevice().GetDeviceReport())
                fh.write("# nsteps, time(s), MATS\n")

    # ------------------------------------------------------------ set-up
    def SetVerbose(self, vb: bool) -> None:
        self.verbose = vb
        self.sample.SetVerbose(vb)

    def SetPotential(self, potential) -> None:
        """Replaces every potential (Simulation.py:160-176); molecular
        potentials must be added with AddPotential."""
        if potential.IsMolecularPotential():
            raise RuntimeError("Must use AddPotential for molecular potentials.")
        self.potentialList = [potential]
        self.moleculePotentials = {}
        self.sample.SetPotential(potential)

    def AddPotential(self, potential, clear: bool = False) -> None:
        """Adds a potential (Simulation.py:178-206): a molecular potential
        sets its exclusions on every non-bonded potential, a non-bonded one
        receives the exclusions of the molecular potentials present."""
        if potential.GetID_String() in self.moleculePotentials:
            raise RuntimeError("Simulation may only have one potential of type %s" % potential.GetID_String())
        if clear:
            self.sample.SetPotential(potential)
            self.moleculePotentials = {}
            self.potentialList = [potential]
        else:
            self.sample.AddPotential(potential)
        if not potential.IsMolecularPotential():
            if not clear:
                self.potentialList.append(potential)
            for pot_molecular in self.moleculePotentials.values():
                pot_molecular.SetExclusions(potential)
        else:
            self.moleculePotentials[potential.GetID_String()] = potential
            for pot_non_bond in self.potentialList:
                if not pot_non_bond.IsMolecularPotential() and hasattr(pot_non_bond, "SetExclusion"):
                    potential.SetExclusions(pot_non_bond)

    def SetIntegrator(self, itg) -> None:
        self.itg = itg
        self.sample.SetIntegrator(itg)

    def ReadMoleculeData(self, topologyFile) -> None:
        """Bonds and angles of the sample's molecules from a RUMD topology
        file (Simulation.py:113-120)."""
        self.sample.ReadMoleculeData(str(topologyFile))

    def SetPB_TP(self, pb: int, tp: int) -> None:
        self.sample.SetPB_TP(pb, tp)

    def SetBlockSize(self, blockSize) -> None:
        """Output block size (None = chosen at Run); a power of 2 for
        logarithmic scheduling (Simulation.py:222-227)."""
        self.blockSize = blockSize

    def SetGraphBlockSize(self, block: int) -> None:
        """Steps per captured HIP graph between host events."""
        self.graph_block = max(1, int(block))

    def SetOutputDirectory(self, outputDirectory: str) -> None:
        self.outputDirectory = outputDirectory
        self.sample.SetOutputDirectory(outputDirectory)

    def GetNumberOfParticles(self) -> int:
        return self.sample.GetNumberOfParticles()

    def GetVolume(self) -> float:
        return self.sample.GetSimulationBox().GetVolume()

    def SetOutputMetaData(self, manager_name: str, **kwargs) -> None:
        """e.g. SetOutputMetaData("energies", stress_xy=True, precision=9)
        (Simulation.py:252-271)."""
        for key, value in kwargs.items():
            self.sample.SetOutputManagerMetaData(manager_name, key, value)

    def SetOutputScheduling(self, manager_name: str, schedule: str, **kwargs) -> None:
        """schedule "none", "linear" (interval=), "logarithmic" (base=, default
        1), "loglin" (base=, maxInterval=) or "limlin" (interval=, numitems=)
        (Simulation.py:327-366)."""
        if manager_name not in self.allowedManagers:
            raise ValueError("SetOutputScheduling -- first argument must be one of " + str(self.allowedManagers))
        param_names = {"none": [], "linear": ["interval"], "logarithmic": ["base"],
                       "loglin": ["base", "maxInterval"], "limlin": ["interval", "numitems"]}
        if schedule not in param_names:
            raise ValueError("SetOutputScheduling -- second argument must be " + str(list(param_names)))
        if schedule == "logarithmic" and "base" not in kwargs:
            kwargs["base"] = 1
        self.outputParameters[manager_name] = tuple([schedule] + [kwargs[name] for name in param_names[schedule]])

    def SetMomentumResetInterval(self, interval: int) -> None:
        self.runtime_action_intervals["momentum_reset"] = int(interval)

    def SetSortInterval(self, sort_interval: int) -> None:
        self.runtime_action_intervals["sort"] = int(sort_interval)

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

    # ------------------------------------------------------------ output
    def InitializeOutput(self, num_iter: int) -> None:
        """Manager schedules and the output block size (Simulation.py:444-510):
        unless set, the smallest power of two >= 1024 that covers every linear
        interval and num_iter / 200."""
        num_active = 0
        min_block = 1
        for name, par in self.outputParameters.items():
            if par[0] == "none":
                self.sample.SetOutputManagerActive(name, False)
                continue
            num_active += 1
            self.sample.SetOutputManagerActive(name, True)
            schedule = par[0]
            if schedule in ("linear", "limlin"):
                interval = par[1]
                user_max = par[2] - 1 if schedule == "limlin" else -1
                self.sample.SetLogLinParameters(name, interval, interval, user_max)
                min_block = max(min_block, interval)
            elif schedule == "logarithmic":
                self.sample.SetLogLinParameters(name, par[1], 0)
            elif schedule == "loglin":
                self.sample.SetLogLinParameters(name, par[1], par[2])
            else:
                raise ValueError("Simulation.InitializeOutput: Unknown schedule %s" % schedule)
        if self.blockSize is None:
            if num_active > 0:
                min_block = max(min_block, num_iter // self.nominalNumBlocks // 2)
                block = 1024
                while block < min_block:
                    block *= 2
            else:
                block = num_iter // self.nominalNumBlocks
            self.blockSize = block
        if self.verbose:
            print("Output blockSize=%d, numBlocks=%d" % (self.blockSize, num_iter // self.blockSize + 1))
            print("Size of the last incomplete block", num_iter % self.blockSize)
        self.sample.SetOutputBlockSize(self.blockSize)
        self.output_is_initialized = True

    def CheckSortingClash(self) -> None:
        """A sort action is redundant with the sorted neighbour-list method,
        which sorts at every rebuild (Simulation.py:512-527)."""
        uses_sort = any(getattr(p, "GetNB_Method", lambda: "")() == "sort" for p in self.potentialList)
        if uses_sort and self.runtime_action_intervals.get("sort", 0) > 0:
            print("[Warning] sort_interval for ordinary sorting should be zero when using the sort-based NB method. "
                  "Setting to zero now.")
            self.SetSortInterval(0)

    # ------------------------------------------------------------ run loop
    def ApplyRuntimeActions(self, i: int) -> None:
        for key, interval in self.runtime_action_intervals.items():
            if interval <= 0 or i % interval != 0:
                continue
            if key == "momentum_reset":
                self.ResetMomentum()
            elif key == "sort":
                self.sample.SortParticles()
            else:
                obj, func = (r() for r in self.runtime_actions[key])
                func(obj)

    def _next_host_event(self, i: int, output: bool) -> int | None:
        """First step >= i that must run on the host: an output write or a
        Python runtime action (the momentum reset runs inside RunSteps)."""
        cand = []
        if output:
            t = self.sample.NextOutputStep(i)
            if t is not None:
                cand.append(t)
        for key, interval in self.runtime_action_intervals.items():
            if key != "momentum_reset" and interval > 0:
                cand.append(i + (-i) % interval)
        return min(cand) if cand else None

    def Run(self, num_iter: int, initializeOutput: bool = True, restartBlock=None, suppressAllOutput: bool = False,
            force_timing: bool = False) -> float:
        if self.itg is None:
            raise ValueError("Integrator has not been set")
        if not self.potentialList:
            raise ValueError("Potential has not been set")
        self.CheckSortingClash()
        first = 0
        output = not suppressAllOutput
        if output:
            if initializeOutput:
                self.InitializeOutput(num_iter)
            elif not self.output_is_initialized:
                raise ValueError("Cannot have initializeOutput=False if output has not previously been initialized")
            if restartBlock is not None:
                if not initializeOutput:
                    raise ValueError("When restarting, initializeOutput must be True")
                self.sample.ReadRestartConf(restartBlock)
                first = int(restartBlock) * self.blockSize
            self.sample.InitializeOutputManagers(first)
        elif restartBlock is not None:
            raise ValueError("restartBlock needs output (suppressAllOutput=False)")
        timing = output or force_timing
        t0 = time.perf_counter()
        mom = max(0, self.runtime_action_intervals.get("momentum_reset", 0))
        i = first
        while i < num_iter:
            nxt = self._next_host_event(i, output)
            stop = num_iter if nxt is None else min(nxt, num_iter)
            if stop > i:
                self.sample.RunSteps(i, stop - i, self.graph_block, mom)
                i = stop
                continue
            # a host step, in the reference's order (Simulation.py:581-597)
            self.ApplyRuntimeActions(i)
            self.sample.CalcF()
            if output:
                self.sample.NotifyOutputManagers(i)
            self.itg.Integrate()
            i += 1
        self.sample.Synchronize()
        self.elapsed = time.perf_counter() - t0
        nsteps = num_iter - first
        self.last_mats = self.sample.GetNumberOfParticles() * nsteps / 1e6 / max(self.elapsed, 1e-12)
        if timing:
            if self.verbose:
                print(f"{nsteps} steps, {self.elapsed:.3f} s, {self.last_mats:.3f} MATS")
            if self.write_timing_info:
                with open(self.timing_filename, "a") as fh:
                    fh.write("%d %.3f %.3f\n" % (nsteps, self.elapsed, self.last_mats))
        rumd.Device.GetDevice().CheckErrors()
        # forces consistent with the final positions (Simulation.py:610-614)
        self.sample.CalcF()
        if output:
            self.sample.TerminateOutputManagers()
        return self.elapsed
