"""The ``rumd`` module surface: RUMD's SWIG names (Swig/rumd.i and the files it
includes) backed by the C++ host classes in the _rumd extension.

    import rumd_amd.rumd as rumd
    sample = rumd.Sample(0, 0)
    sample.ReadConf("start.xyz.gz")
    pot = rumd.Pot_LJ_12_6(cutoff_method=rumd.ShiftedPotential)
    pot.SetParams(0, 0, Sigma=1.0, Epsilon=1.0, Rcut=2.5)
    sample.SetPotential(pot)
"""
from __future__ import annotations

import numpy as np

from . import conf_io
from . import output as _output
from ._native import _rumd

Device = _rumd.Device
SimulationBox = _rumd.SimulationBox
RectangularSimulationBox = _rumd.RectangularSimulationBox
Potential = _rumd.Potential
PairPotential = _rumd.PairPotential
Pot_IPL_12_6 = _rumd.Pot_IPL_12_6
Pot_LJ_12_6 = _rumd.Pot_LJ_12_6
Pot_gLJ_m12_n6 = _rumd.Pot_gLJ_m12_n6
Pot_LJ_12_6_smooth_2_4 = _rumd.Pot_LJ_12_6_smooth_2_4
Pot_gLJ_m_n = _rumd.Pot_gLJ_m_n
Pot_Gauss = _rumd.Pot_Gauss
Pot_Buckingham = _rumd.Pot_Buckingham
Pot_IPL_n = _rumd.Pot_IPL_n
Pot_Yukawa = _rumd.Pot_Yukawa
Pot_Exp = _rumd.Pot_Exp
Pot_LJ_12_6_Gauss = _rumd.Pot_LJ_12_6_Gauss
Pot_IPL_12 = _rumd.Pot_IPL_12
Pot_IPL_9 = _rumd.Pot_IPL_9
Pot_IPL_18 = _rumd.Pot_IPL_18
Pot_IPL6_LN = _rumd.Pot_IPL6_LN
Pot_LJ_LN = _rumd.Pot_LJ_LN
Pot_Girifalco = _rumd.Pot_Girifalco
Pot_Dzugutov = _rumd.Pot_Dzugutov
Pot_Fermi_Jagla = _rumd.Pot_Fermi_Jagla
Pot_Repulsive_Shoulder = _rumd.Pot_Repulsive_Shoulder
Pot_SAAP = _rumd.Pot_SAAP
Pot_ShiftedForceCoulomb = _rumd.Pot_ShiftedForceCoulomb
BondHarmonic = _rumd.BondHarmonic
AngleCosSq = _rumd.AngleCosSq
AngleSq = _rumd.AngleSq
BondFENE = _rumd.BondFENE
DihedralRyckaert = _rumd.DihedralRyckaert
PeriodicDihedral = _rumd.PeriodicDihedral
MoleculeData = _rumd.MoleculeData
Integrator = _rumd.Integrator
IntegratorNVT = _rumd.IntegratorNVT
IntegratorNVE = _rumd.IntegratorNVE
NoShift = _rumd.NoShift
ShiftedPotential = _rumd.ShiftedPotential
ShiftedForce = _rumd.ShiftedForce
GetVersion = _rumd.GetVersion
NumberOfPairPotentialParameters = _rumd.NumberOfPairPotentialParameters
ReplicaBatch = _rumd.ReplicaBatch


class Sample(_rumd.Sample):
    """rumd.Sample (Swig/Sample.i) with configuration-file IO."""

    def __init__(self, pb: int = 0, tp: int = 0):
        super().__init__(pb, tp)
        self._box = None
        self._output_dir = "TrajectoryFiles"
        self._managers = _output.default_managers(self, self._output_dir)
        self._managers_initialized = False
        self._output_block_size = 0

    # ------------------------------------------------------------ potentials
    def SetPotential(self, potential) -> None:
        super().SetPotential(potential)
        self._managers["energies"].RegisterPotential(potential)
        self._managers_initialized = False

    def AddPotential(self, potential) -> None:
        super().AddPotential(potential)
        self._managers["energies"].RegisterPotential(potential)
        self._managers_initialized = False

    def SetIntegrator(self, newItg) -> None:
        super().SetIntegrator(newItg)
        self._managers["energies"].RegisterIntegrator(newItg)
        self._managers_initialized = False

    def SetVerbose(self, vb: bool) -> None:
        super().SetVerbose(vb)
        for m in self._managers.values():
            m.SetVerbose(vb)

    # ------------------------------------------------------------ output managers
    # Sample's output-manager interface (src/Sample.cu:482-575).
    def GetOutputDirectory(self) -> str:
        return self._output_dir

    def SetOutputDirectory(self, outputDirectory: str) -> None:
        self._output_dir = outputDirectory
        for m in self._managers.values():
            m.SetOutputDirectory(outputDirectory)
        self._managers_initialized = False

    def GetOutputManager(self, manager_name: str):
        return self._managers[manager_name]

    def AddOutputManager(self, manager_name: str, manager) -> None:
        self._managers[manager_name] = manager
        self._managers_initialized = False

    def SetOutputBlockSize(self, blockSize: int) -> None:
        if not blockSize:
            raise ValueError("Sample::SetOutputBlockSize: blockSize must be greater than zero")
        self._output_block_size = int(blockSize)
        for m in self._managers.values():
            m.SetBlockSize(blockSize)
        self._managers["restart"].SetLogLinParams(blockSize, 0)
        self._managers_initialized = False

    def GetOutputBlockSize(self) -> int:
        return self._output_block_size

    def SetOutputManagerActive(self, manager_name: str, active: bool) -> None:
        if manager_name not in self._managers:
            print(f"Warning [Sample::SetOutputManagerActive]: Manager {manager_name} not present")
            return
        self._managers[manager_name].SetActive(active)

    def SetOutputManagerMetaData(self, manager_name: str, key: str, value) -> None:
        if manager_name not in self._managers:
            print(f"Warning [Sample::SetOutputManagerMetaData]: Manager {manager_name} not present")
            return
        self._managers[manager_name].SetMetaData(key, value)

    def SetLogLinParameters(self, manager_name: str, base: int, maxInterval: int, user_maxIndex: int = -1) -> None:
        m = self._managers[manager_name]
        m.SetLogLinParams(base, maxInterval)
        m.SetUserMaxIndex(user_maxIndex)
        # linear, or logarithmic limited to the first items: no duplicate block ends
        m.SetDuplicateBlockEnds(not (base == maxInterval or user_maxIndex >= 0))
        self._managers_initialized = False

    def InitializeOutputManagers(self, timeStepIndex: int) -> None:
        energies = self._managers["energies"]
        for pot in self.GetPotentials():
            energies.RegisterPotential(pot)
        energies.RegisterIntegrator(self.GetIntegrator())
        create_dir = True
        for name in sorted(self._managers):
            self._managers[name].Initialize(int(timeStepIndex), create_dir)
            create_dir = False  # only the first one creates the directory
        self._managers_initialized = True

    def NotifyOutputManagers(self, timeStepIndex: int) -> None:
        if not self._managers_initialized:
            self.InitializeOutputManagers(timeStepIndex)
        for name in sorted(self._managers):
            self._managers[name].Update(int(timeStepIndex))

    def NextOutputStep(self, timeStepIndex: int):
        """The first step >= timeStepIndex at which a manager writes (None if
        no manager is active)."""
        steps = [m.next_write_step(int(timeStepIndex)) for m in self._managers.values()]
        steps = [t for t in steps if t is not None]
        return min(steps) if steps else None

    def TerminateOutputManagers(self) -> None:
        for m in self._managers.values():
            m.Terminate()

    def ReadRestartConf(self, restartBlock: int) -> None:
        """Reads restartNNNN.xyz.gz from the output directory, integrator
        state included (Sample::ReadRestartConf)."""
        m = self._managers["restart"]
        fn = f"{self._output_dir}/restart{int(restartBlock):0{m.numFilenameDigits}d}.xyz.gz"
        self.ReadConf(fn, init_itg=True)

    def SetConfiguration(self, conf: conf_io.Configuration) -> None:
        """Loads a configuration: box, particle count, types, masses,
        positions, velocities and images (what ReadConf stores,
        src/ConfigurationWriterReader.cu:225-343)."""
        box = RectangularSimulationBox(*[float(x) for x in conf.box])
        self._box = box
        self.SetSimulationBox(box)
        self.SetNumberOfParticles(int(conf.n))
        for t, m in enumerate(conf.masses):
            self.SetMass(t, float(m))
        self.SetTypes(np.ascontiguousarray(conf.types, dtype=np.int32))
        self.SetPositions(np.ascontiguousarray(conf.positions, dtype=np.float32))
        if conf.velocities is not None:
            self.SetVelocities(np.ascontiguousarray(conf.velocities, dtype=np.float32))
        if conf.images is not None:
            self.SetImages(np.ascontiguousarray(conf.images, dtype=np.float32))

    def ReadConf(self, filename, init_itg: bool = False) -> None:
        conf = conf_io.read_conf(filename)
        self.SetConfiguration(conf)
        if init_itg and self.GetIntegrator() is not None and "integrator" in conf.meta:
            self.GetIntegrator().InitializeFromInfoString(conf.meta["integrator"])

    def GetConfiguration(self) -> conf_io.Configuration:
        box = self.GetSimulationBox()
        nt = self.GetNumberOfTypes()
        f4 = self.GetForcesAndEnergies()
        w4 = self.GetVirials()
        return conf_io.Configuration(
            n=self.GetNumberOfParticles(),
            box=np.array([box.GetLength(k) for k in range(3)], dtype=np.float64),
            types=self.GetTypes(),
            positions=self.GetPositions(),
            masses=np.array([self.GetMass(t) for t in range(nt)], dtype=np.float64),
            images=self.GetImages(),
            velocities=self.GetVelocities(),
            forces=f4[:, :3].copy(),
            pe=f4[:, 3].astype(np.float64),
            vir=w4[:, 3].astype(np.float64) / 2.0,
        )

    def WriteConf(self, filename, mode: str = "w") -> None:
        if mode != "w":
            raise ValueError("only mode 'w' is supported")
        itg = self.GetIntegrator()
        info = itg.GetInfoString(9) if itg is not None else None
        conf_io.write_conf(filename, self.GetConfiguration(), integrator_info=info)


__all__ = [
    "Device", "SimulationBox", "RectangularSimulationBox", "Potential", "PairPotential", "Pot_IPL_12_6",
    "Pot_LJ_12_6", "Pot_gLJ_m12_n6", "Pot_LJ_12_6_smooth_2_4", "Pot_gLJ_m_n", "Pot_Gauss", "Pot_Buckingham",
    "Pot_IPL_n", "Pot_Yukawa", "Pot_Exp", "Pot_LJ_12_6_Gauss", "Pot_IPL_12", "Pot_IPL_9", "Pot_IPL_18", "Pot_IPL6_LN", "Pot_LJ_LN", "Pot_Girifalco", "Pot_Dzugutov", "Pot_Fermi_Jagla", "Pot_Repulsive_Shoulder", "Pot_SAAP", "Pot_ShiftedForceCoulomb",
    "BondHarmonic", "BondFENE", "AngleCosSq", "AngleSq", "DihedralRyckaert", "PeriodicDihedral", "MoleculeData", "Integrator", "IntegratorNVT", "IntegratorNVE", "NoShift", "ShiftedPotential",
    "ShiftedForce", "GetVersion", "NumberOfPairPotentialParameters", "Sample", "ReplicaBatch",
]
