"""Output managers: the energies file, the trajectory and the restart
configurations, each on a log-lin schedule.

Follows LogLinOutputManager (src/LogLinOutputManager.cu:28-232, the schedule
state of include/rumd/LogLin.h), EnergiesOutputManager
(src/EnergiesOutputManager.cu:11-215, columns of src/EnergiesMetaData.cc and
include/rumd/EnergiesMetaData.h:14-60), ConfigurationOutputManager
(src/ConfigurationOutputManager.cu:16-30) and the default managers that
Sample::CreateDefaultOutputManagers creates (src/Sample.cu:150-189).

The schedule is host bookkeeping; the quantities written come from the
sample's FP64 device sums (Sample.GetPotentialEnergy, GetVirial, GetStress,
Integrator.GetKineticEnergy).  ``next_write_step`` lets the run loop execute
graph-captured blocks of steps up to the next step at which a manager writes.
"""
from __future__ import annotations

import gzip
import os
import shutil
import sys
from dataclasses import dataclass

import numpy as np

from . import conf_io


class OutputError(RuntimeError):
    """RUMD_Error raised by an output manager (class, method, message)."""

    def __init__(self, cls: str, method: str, msg: str):
        super().__init__(f"{cls}::{method}: {msg}")


def _log2_floor_f32(x: int) -> int:
    # (unsigned int) log2(float(x)) as the reference computes it
    return int(np.log2(np.float32(x)))


@dataclass
class LogLin:
    """Schedule state (include/rumd/LogLin.h): the block, the next step to
    write (relative to the block start), the smallest interval, the item index
    in the block and its last value, the interval cap (0 = none; equal to base
    = linear) and the current interval."""
    block: int = 0
    nextTimeStep: int = 0
    base: int = 1
    index: int = 0
    maxIndex: int = 0
    maxInterval: int = 0
    increment: int = 0


class LogLinOutputManager:
    """Writes at steps chosen by a log-lin schedule inside blocks of
    ``blockSize`` steps (src/LogLinOutputManager.cu).  Subclasses implement
    ``Write``."""

    def __init__(self, sample, outputDirectory: str, baseFilename: str):
        self.sample = sample
        self.blockSize = 0
        self.user_maxIndex = -1
        self.outputDirectory = outputDirectory
        self.baseFilename = baseFilename
        self.backup_ext = ".rumd_bak"
        self.maxNumIoBlocks = 9999
        self.numFilenameDigits = 4
        self.logLin = LogLin()
        self.active = True
        self.duplicateBlockEnds = False
        self.verbose = True
        self.make_backup = True

    # ------------------------------------------------------------ settings
    def SetLogLinParams(self, base: int, maxInterval: int) -> None:
        self.logLin.base = int(base)
        self.logLin.maxInterval = int(maxInterval)

    def SetUserMaxIndex(self, user_maxIndex: int) -> None:
        self.user_maxIndex = int(user_maxIndex)

    def SetBlockSize(self, blockSize: int) -> None:
        self.blockSize = int(blockSize)

    def SetActive(self, active: bool) -> None:
        self.active = bool(active)

    def SetDuplicateBlockEnds(self, dup: bool) -> None:
        self.duplicateBlockEnds = bool(dup)

    def SetVerbose(self, vb: bool) -> None:
        self.verbose = bool(vb)

    def SetOutputDirectory(self, d: str) -> None:
        self.outputDirectory = d

    # ------------------------------------------------------------ schedule
    def CheckAndInitializeLogLin(self) -> None:
        """Validates base / maxInterval / blockSize and sets maxIndex
        (src/LogLinOutputManager.cu:33-103)."""
        L, bs, name = self.logLin, self.blockSize, self.baseFilename
        cls = "LogLinOutputManager"
        if bs == 0:
            raise OutputError(cls, "CheckAndInitializeLogLin", "blockSize has not been set")
        if L.base == 0:
            raise OutputError(cls, "CheckAndInitializeLogLin", f"Attempt to set base to zero in log-lin with baseFilename {name}")
        if L.base > bs:
            raise OutputError(cls, "CheckAndInitializeLogLin",
                              f"logLin.base is greater than the blockSize in log-lin with baseFilename {name}")
        if L.maxInterval % L.base != 0:
            raise OutputError(cls, "CheckAndInitializeLogLin",
                              f"maxInterval {L.maxInterval} is not a multiple of base {L.base} for log-lin with "
                              f"baseFilename {name}")
        if L.maxInterval != 0 and L.maxInterval != L.base and bs % L.maxInterval != 0:
            raise OutputError(cls, "CheckAndInitializeLogLin",
                              f"maxInterval {L.maxInterval} does not divide into blockSize {bs} for log-lin with "
                              f"baseFilename {name}, with non-linear saving")
        if bs % L.base != 0 and L.maxInterval != L.base:
            raise OutputError(cls, "CheckAndInitializeLogLin",
                              f"blockSize {bs} is not a multiple of base {L.base} for log-lin with baseFilename "
                              f"{name}, with non-linear saving")
        log2_block = _log2_floor_f32(bs // L.base)
        rounded = (1 << log2_block) * L.base
        if bs != rounded and L.maxInterval == 0:
            # pure logarithmic saving needs a power-of-two block: fall back to
            # one item per block
            print(f"WARNING (LogLinOutputManager::CheckAndInitializeLogLin, baseFilename {name}): Nominal block "
                  f"size {bs} is not a power of 2 times the base. This means effectively that the base of LogLin "
                  f"will be set to this value, i.e. no logarithmic saving", file=sys.stderr)
            L.base = bs
            L.maxIndex = 0
            self.duplicateBlockEnds = False
        else:
            L.maxIndex = log2_block + 1
            if L.maxInterval > 0:
                L.maxIndex = bs // L.maxInterval + _log2_floor_f32(L.maxInterval // L.base)
            uneven_linear = L.base == L.maxInterval and bs % L.base != 0
            if uneven_linear and self.duplicateBlockEnds:
                raise OutputError(cls, "CheckAndInitializeLogLin",
                                  "duplicateBlockEnds cannot be true when doing linear saving with an interval which "
                                  f"does not divide the block-size (baseFilename:{name})")
            if not self.duplicateBlockEnds and not uneven_linear:
                L.maxIndex -= 1
        if 0 <= self.user_maxIndex < L.maxIndex:
            L.maxIndex = self.user_maxIndex
        L.nextTimeStep = 0
        L.index = 0

    def Initialize(self, timeStepIndex: int, create_directory: bool) -> None:
        """Schedule and output directory at the start of a run or a restart
        (src/LogLinOutputManager.cu:107-151)."""
        if self.active:
            self.CheckAndInitializeLogLin()
        restart = timeStepIndex > 0
        exists = os.path.isdir(self.outputDirectory)
        may_create = create_directory and not restart
        if exists and may_create:
            self.MakeBackupDir()
            os.makedirs(self.outputDirectory)
        elif not exists and may_create:
            os.makedirs(self.outputDirectory)
        elif not exists and not may_create:
            raise OutputError("LogLinOutputManager", "Initialize",
                              "Output directory does not exist and either this manager does not have the right to "
                              "create it or this is a restart job")
        if not self.active:
            return
        L, bs = self.logLin, self.blockSize
        if timeStepIndex % bs != 0:
            raise OutputError("LogLinOutputManager", "Initialize", "a run must start at a block boundary")
        L.block = timeStepIndex // bs
        if L.base == L.maxInterval:
            # linear saving whose interval need not divide the block
            L.nextTimeStep = (L.maxInterval - timeStepIndex % L.maxInterval) % L.maxInterval
            if bs % L.maxInterval != 0:
                L.maxIndex = (bs - L.nextTimeStep) // L.maxInterval
        if 0 <= self.user_maxIndex < L.maxIndex:
            L.nextTimeStep = 0
            L.maxIndex = self.user_maxIndex

    def next_write_step(self, step: int) -> int | None:
        """The first step >= ``step`` at which ``Update`` writes (None when
        inactive): Update writes when step % blockSize equals nextTimeStep %
        blockSize, and nothing else changes the state in between."""
        if not self.active or self.blockSize <= 0:
            return None
        bs = self.blockSize
        return step + (self.logLin.nextTimeStep % bs - step % bs) % bs

    def Update(self, timeStepIndex: int) -> None:
        """Writes if this step is scheduled, then advances the schedule
        (src/LogLinOutputManager.cu:154-232)."""
        L, bs = self.logLin, self.blockSize
        if not (self.active and timeStepIndex % bs == L.nextTimeStep % bs):
            return
        if L.block == self.maxNumIoBlocks:
            raise OutputError("LogLinOutputManager", "Update", "Maximum number of blocks reached.")
        self.Write()
        cur, cur_max = L.index, L.maxIndex
        init_increments = False
        if cur == cur_max:
            with open(os.path.join(self.outputDirectory, f"LastComplete_{self.baseFilename}.txt"), "w") as fh:
                fh.write(f"{L.block} {bs}\n")
            if self.verbose and L.block % 10 == 0:
                print(f"Finished writing {self.baseFilename} nr. {L.block:{self.numFilenameDigits}d}")
            L.block += 1
            L.index = 0
            if bs % L.base != 0:
                # uneven linear saving: carry the phase into the next block
                L.nextTimeStep = (L.nextTimeStep + L.increment) % bs
                L.maxIndex = (bs - L.nextTimeStep - 1) // L.maxInterval
                if 0 <= self.user_maxIndex < L.maxIndex:
                    L.nextTimeStep = 0
                    L.maxIndex = self.user_maxIndex
            else:
                L.nextTimeStep = 0
            if self.duplicateBlockEnds:
                # the last item of a block is also the first of the next
                self.Write()
                init_increments = True
        elif cur > 0:
            L.nextTimeStep += L.increment
            L.increment *= 2
            if L.maxInterval > 0 and L.increment > L.maxInterval:
                L.increment = L.maxInterval
            L.index += 1
        else:
            init_increments = True
        if init_increments:
            L.increment = L.base
            L.nextTimeStep += L.increment
            L.index = 1

    def Terminate(self) -> None:
        """Closes open files (the run has ended)."""

    def Write(self) -> None:  # pragma: no cover - abstract
        raise NotImplementedError

    # ------------------------------------------------------------ files
    def MakeBackupDir(self) -> None:
        backup = self.outputDirectory + self.backup_ext
        if os.path.isdir(backup):
            shutil.rmtree(backup)
        if self.make_backup:
            if self.verbose:
                print(f"{self.outputDirectory} exists; renaming as {backup}")
            os.rename(self.outputDirectory, backup)
        else:
            if self.verbose:
                print(f"Removing existing output directory {self.outputDirectory} (back-up has been disabled)")
            shutil.rmtree(self.outputDirectory)

    def block_filename(self, ext: str) -> str:
        return os.path.join(self.outputDirectory,
                            f"{self.baseFilename}{self.logLin.block:0{self.numFilenameDigits}d}{ext}")


# (key, file string, on by default), include/rumd/EnergiesMetaData.h:20-55
_ENERGY_COLUMNS = [
    ("potentialEnergy", "pe", True), ("kineticEnergy", "ke", True), ("virial", "W", True),
    ("totalEnergy", "Etot", True), ("temperature", "T", True), ("pressure", "p", True),
    ("volume", "V", False), ("density", "rho", False), ("enthalpy", "H", False),
    ("potentialVirial", "pot_W", False),
    ("stress_xx", "sxx", False), ("stress_yy", "syy", False), ("stress_zz", "szz", False),
    ("stress_yz", "syz", False), ("stress_xz", "sxz", False), ("stress_xy", "sxy", False),
]
# IntegratorNVT::GetDataInfo (src/IntegratorNVT.cu:194-207)
_NVT_COLUMNS = [("thermostat_Ps", "Ps"), ("v_com_x", "v_com_x"), ("v_com_y", "v_com_y"), ("v_com_z", "v_com_z")]


class EnergiesMetaData:
    """Column switches and file strings; columns are written in the sorted
    order of their keys (the reference keeps them in a std::map)."""

    def __init__(self):
        self.bool_options = {k: on for k, _, on in _ENERGY_COLUMNS}
        self.fileStr = {k: s for k, s, _ in _ENERGY_COLUMNS}
        self.int_options = {"precision": 6}

    def Set(self, key: str, value) -> None:
        if isinstance(value, bool):
            if key not in self.bool_options:
                raise OutputError("EnergiesMetaData", "Set [bool]", f"Unrecognized key:{key}")
            self.bool_options[key] = value
        else:
            if key not in self.int_options:
                raise OutputError("EnergiesMetaData", "Set [int]", f"Unrecognized key:{key}")
            self.int_options[key] = int(value)

    def Get(self, key: str):
        if key in self.bool_options:
            return self.bool_options[key]
        if key in self.int_options:
            return self.int_options[key]
        raise OutputError("EnergiesMetaData", "Get", f"Unrecognized key:{key}")

    def active_keys(self) -> list[str]:
        return [k for k in sorted(self.bool_options) if self.bool_options[k]]


class EnergiesOutputManager(LogLinOutputManager):
    """energiesNNNN.dat.gz: one header line per block, one line per write
    (src/EnergiesOutputManager.cu:11-130)."""

    def __init__(self, sample, outputDirectory: str, baseFilename: str = "energies"):
        super().__init__(sample, outputDirectory, baseFilename)
        self.metaData = EnergiesMetaData()
        self.energies_ioformat = 2
        self._fh = None
        self._registered_itg = None

    def SetMetaData(self, key: str, value) -> None:
        self.metaData.Set(key, value)

    def RegisterPotential(self, pot) -> None:
        pid = pot.GetID_String()
        self.metaData.bool_options.setdefault(pid, False)
        self.metaData.fileStr[pid] = pid

    def RegisterIntegrator(self, itg) -> None:
        if self._registered_itg is itg:
            return
        for k, _ in _NVT_COLUMNS:
            self.metaData.bool_options.pop(k, None)
            self.metaData.fileStr.pop(k, None)
        if itg is not None and hasattr(itg, "GetThermostatState"):
            for k, s in _NVT_COLUMNS:
                self.metaData.bool_options[k] = False
                self.metaData.fileStr[k] = s
        self._registered_itg = itg

    def Write(self) -> None:
        L = self.logLin
        if L.index == 0:
            self._close()
            self._fh = gzip.open(self.block_filename(".dat.gz"), "wt")
            itg = self.sample.GetIntegrator()
            dt = float(np.float32(itg.GetTimeStep())) if itg is not None else 0.0
            head = f"# ioformat={self.energies_ioformat} N={self.sample.GetNumberOfParticles()}"
            if L.base != L.maxInterval:
                head += (f" timeStepIndex={L.nextTimeStep} logLin={L.block},{L.base},{L.index},{L.maxIndex},"
                         f"{L.maxInterval}")
            else:
                head += f" Dt={float(np.float32(dt * L.base)):f}"
            head += " columns=" + ",".join(self.metaData.fileStr[k] for k in self.metaData.active_keys())
            self._fh.write(head + "\n")
        elif self._fh is None:
            self._fh = gzip.open(self.block_filename(".dat.gz"), "at")
        self.WriteOneLine()

    def WriteOneLine(self) -> None:
        values = self.CalcEnergies()
        s, md = self.sample, self.metaData
        n = s.GetNumberOfParticles()
        for pot in s.GetPotentials():
            pid = pot.GetID_String()
            if md.bool_options.get(pid, False):
                values[pid] = pot.GetPotentialEnergy() / n
        itg = s.GetIntegrator()
        if itg is not None and md.bool_options.get("thermostat_Ps", False):
            values["thermostat_Ps"] = itg.GetThermostatState()
        if any(md.bool_options.get(k, False) for k in ("v_com_x", "v_com_y", "v_com_z")):
            p = itg.GetTotalMomentum()
            values["v_com_x"], values["v_com_y"], values["v_com_z"] = p[0] / p[3], p[1] / p[3], p[2] / p[3]
        prec = md.Get("precision")
        keys = md.active_keys()
        if keys:
            # the reference holds each value as a float before printing
            self._fh.write(" ".join(f"{float(np.float32(values[k])):.{prec}g}" for k in keys) + "\n")

    def CalcEnergies(self) -> dict:
        """EnergiesOutputManager::CalcEnergies (src/EnergiesOutputManager.cu:132-215):
        KE(t) averaged over the half steps, T = 2 KE / DOF, P = (N T + W) / V,
        E = PE + KE, H = E + P V; per-particle values divide by N."""
        s, md = self.sample, self.metaData
        itg = s.GetIntegrator()
        n = s.GetNumberOfParticles()
        dof = s.GetNumberOfDOFs()
        V = float(np.float32(s.GetSimulationBox().GetVolume()))
        need_stress = any(md.bool_options.get(k, False) for k in
                          ("stress_xx", "stress_yy", "stress_zz", "stress_xy", "stress_yz", "stress_xz"))
        need_ke = itg is not None and any(md.bool_options[k] for k in
                                          ("kineticEnergy", "temperature", "pressure", "totalEnergy"))
        need_pe = any(md.bool_options[k] for k in ("potentialEnergy", "pressure", "totalEnergy"))
        need_vir = md.bool_options["virial"] or md.bool_options["pressure"]
        if need_stress:
            s.CalcF(True)
        ke = itg.GetKineticEnergy() if need_ke else 0.0
        T = float(np.float32(2.0 * ke / dof))
        pe = s.GetPotentialEnergy() if need_pe else 0.0
        vir = s.GetVirial() if need_vir else 0.0
        E = float(np.float32(pe + ke))
        P = float(np.float32((n * T + vir) / V))
        H = float(np.float32(E + P * V))
        out = {
            "kineticEnergy": ke / n, "potentialEnergy": pe / n, "totalEnergy": E / n, "virial": vir / n,
            "potentialVirial": vir / n, "temperature": T, "pressure": P, "volume": V, "density": n / V,
            "enthalpy": H / n,
        }
        if need_stress:
            st = s.GetStress()
            for k, key in enumerate(("stress_xx", "stress_yy", "stress_zz", "stress_yz", "stress_xz", "stress_xy")):
                out[key] = st[k]
        return out

    def _close(self) -> None:
        if self._fh is not None:
            self._fh.close()
            self._fh = None

    def Terminate(self) -> None:
        self._close()


class ConfigurationMetaData:
    """Writer switches (include/rumd/ConfigurationMetaData.h:33-45)."""

    def __init__(self, **on):
        self.bool_options = {"images": False, "velocities": False, "forces": False, "pot_energies": False,
                             "virials": False, "logLin": False}
        self.int_options = {"precision": 6}
        for k, v in on.items():
            self.Set(k, v)

    def Set(self, key: str, value) -> None:
        if isinstance(value, bool):
            if key not in self.bool_options:
                raise OutputError("ConfigurationMetaData", "Set", f"Unrecognized key:{key}")
            self.bool_options[key] = value
        else:
            if key not in self.int_options:
                raise OutputError("ConfigurationMetaData", "Set", f"Unrecognized key:{key}")
            self.int_options[key] = int(value)

    def Get(self, key: str):
        return self.bool_options[key] if key in self.bool_options else self.int_options[key]


class ConfigurationOutputManager(LogLinOutputManager):
    """baseNNNN.xyz.gz: the first item of a block starts the file, later
    items append a frame (src/ConfigurationOutputManager.cu:16-30)."""

    def __init__(self, sample, outputDirectory: str, baseFilename: str, metaData: ConfigurationMetaData):
        super().__init__(sample, outputDirectory, baseFilename)
        self.metaData = metaData

    def SetMetaData(self, key: str, value) -> None:
        self.metaData.Set(key, value)

    def Write(self) -> None:
        md, L = self.metaData, self.logLin
        prec = md.Get("precision")
        conf = self.sample.GetConfiguration()
        if not md.Get("images"):
            conf.images = None
        if not md.Get("velocities"):
            conf.velocities = None
        if not md.Get("forces"):
            conf.forces = None
        if not md.Get("pot_energies"):
            conf.pe = None
        if not md.Get("virials"):
            conf.vir = None
        itg = self.sample.GetIntegrator()
        loglin = None
        if md.Get("logLin"):
            loglin = (L.nextTimeStep, L.block, L.base, L.index, L.maxIndex, L.maxInterval)
        conf_io.write_conf(self.block_filename(".xyz.gz"), conf, precision=prec,
                           integrator_info=itg.GetInfoString(prec + 2) if itg is not None else None,
                           box_info=self.sample.GetSimulationBox().GetInfoString(prec + 2),
                           loglin=loglin, mode="a" if L.index > 0 else "w")


def default_managers(sample, outputDirectory: str) -> dict:
    """restart, trajectory and energies as Sample::CreateDefaultOutputManagers
    sets them up (src/Sample.cu:150-189)."""
    restart = ConfigurationOutputManager(
        sample, outputDirectory, "restart",
        ConfigurationMetaData(precision=9, images=True, velocities=True, forces=True, pot_energies=True,
                              virials=True))
    trajectory = ConfigurationOutputManager(
        sample, outputDirectory, "trajectory",
        ConfigurationMetaData(precision=4, images=True, velocities=False, forces=False, pot_energies=True,
                              virials=True, logLin=True))
    trajectory.SetDuplicateBlockEnds(True)
    trajectory.SetLogLinParams(1, 0)
    energies = EnergiesOutputManager(sample, outputDirectory, "energies")
    energies.SetMetaData("precision", 6)
    # iteration order of the reference's std::map: energies, restart, trajectory
    return {"energies": energies, "restart": restart, "trajectory": trajectory}
