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
from ._native import _rumd

Device = _rumd.Device
SimulationBox = _rumd.SimulationBox
RectangularSimulationBox = _rumd.RectangularSimulationBox
Potential = _rumd.Potential
PairPotential = _rumd.PairPotential
Pot_IPL_12_6 = _rumd.Pot_IPL_12_6
Pot_LJ_12_6 = _rumd.Pot_LJ_12_6
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
    "Pot_LJ_12_6", "Integrator", "IntegratorNVT", "IntegratorNVE", "NoShift", "ShiftedPotential",
    "ShiftedForce", "GetVersion", "NumberOfPairPotentialParameters", "Sample", "ReplicaBatch",
]
