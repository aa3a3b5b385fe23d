// Python surface of the host classes, with the names, keyword arguments and
// numpy shapes of RUMD's SWIG module (Swig/Sample.i, Swig/Potentials.i,
// Swig/Integrators.i, Swig/SimulationBox.i, Swig/rumd.i).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "replica_batch.h"
#include "rumd_host.h"

namespace py = pybind11;
using f32arr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using i32arr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

namespace {

// Nx3 float32 in original particle order (Sample.i GetPositions etc.)
py::array_t<float> get_xyz(Sample& s, const char* what) {
  const unsigned n = s.GetNumberOfParticles();
  if (!n) throw RUMD_Error("Sample", "Get", "Sample has no particles");
  s.Synchronize();
  std::vector<float> a(4ull * n);
  s.GetParticleData().Get(what, a.data(), s.GetStream());
  py::array_t<float> out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(3)});
  auto o = out.mutable_unchecked<2>();
  for (unsigned i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) o(i, k) = a[4ull * i + k];
  return out;
}

// Nx4 float32 (f: fx,fy,fz,u_i; w: w_mol,sxz,sxy,w_i; ...)
py::array_t<float> get_x4(Sample& s, const char* what) {
  const unsigned n = s.GetNumberOfParticles();
  if (!n) throw RUMD_Error("Sample", "Get", "Sample has no particles");
  s.Synchronize();
  py::array_t<float> out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(4)});
  s.GetParticleData().Get(what, out.mutable_data(), s.GetStream());
  return out;
}

const float* check_nx3(Sample& s, const f32arr& a) {
  if (a.ndim() != 2 || a.shape(0) != s.GetNumberOfParticles() || a.shape(1) != 3)
    throw py::value_error("Wrong array shape, must be Nx3");
  return a.data();
}

void reset_lists(Sample& s) {
  for (Potential* p : s.GetPotentials()) p->ResetInternalData();
}

}  // namespace

PYBIND11_MODULE(_rumd, m) {
  m.doc() = "C++ host classes of rumd_amd (RUMD's Sample/Potential/Integrator API) over the gfx950 C ABI";

  py::register_exception_translator([](std::exception_ptr p) {
    try {
      if (p) std::rethrow_exception(p);
    } catch (const RUMD_cudaResourceError& e) {
      PyErr_SetString(PyExc_ValueError, e.what());
    } catch (const RUMD_Error& e) {
      PyErr_SetString(PyExc_RuntimeError, e.what());
    }
  });

  m.attr("VERSION") = "3.5";
  m.def("GetVersion", []() { return std::string("3.5"); });
  m.def("NumberOfPairPotentialParameters", []() { return RUMDX_NUM_PARAM; });

  py::class_<Device, std::unique_ptr<Device, py::nodelete>>(m, "Device")
      .def_static("GetDevice", &Device::GetDevice, py::return_value_policy::reference)
      .def("GetDeviceName", &Device::GetDeviceName)
      .def("GetDeviceReport", &Device::GetDeviceReport)
      .def("GetDeviceMemory", &Device::GetDeviceMemory)
      .def("Synchronize", &Device::Synchronize)
      .def("Time", &Device::Time)
      .def("CheckErrors", &Device::CheckErrors);

  py::class_<SimulationBox>(m, "SimulationBox")
      .def("GetVolume", &SimulationBox::GetVolume)
      .def("GetLength", &SimulationBox::GetLength, py::arg("dir"))
      .def("GetInfoString", &SimulationBox::GetInfoString, py::arg("precision") = 9);
  py::class_<RectangularSimulationBox, SimulationBox>(m, "RectangularSimulationBox")
      .def(py::init<float, float, float>(), py::arg("Lx"), py::arg("Ly"), py::arg("Lz"))
      .def("SetLengths", &RectangularSimulationBox::SetLengths);

  py::class_<Potential>(m, "Potential")
      .def("SetVerbose", &Potential::SetVerbose, py::arg("vb"))
      .def("SetID_String", &Potential::SetID_String, py::arg("set_ID_String"))
      .def("GetID_String", &Potential::GetID_String)
      .def("ResetInternalData", &Potential::ResetInternalData)
      .def("GetPotentialEnergy", &Potential::GetPotentialEnergy)
      .def("EnergyIncludedInParticleSum", &Potential::EnergyIncludedInParticleSum)
      .def("IsMolecularPotential", &Potential::IsMolecularPotential);

  py::class_<PairPotential, Potential> pp(m, "PairPotential");
  py::enum_<PairPotential::CutoffMethod>(pp, "CutoffMethod")
      .value("NS", PairPotential::NS)
      .value("SP", PairPotential::SP)
      .value("SF", PairPotential::SF)
      .export_values();
  pp.def("GetNbListSkin", &PairPotential::GetNbListSkin)
      .def("GetMaxCutoff", &PairPotential::GetMaxCutoff)
      .def("GetNbListRebuildRequired", &PairPotential::GetNbListRebuildRequired)
      .def("GetNumberOfNbListBuilds", &PairPotential::GetNumberOfNbListBuilds)
      .def("SetNbListSkin", &PairPotential::SetNbListSkin, py::arg("skin"))
      .def("SetNB_Method", &PairPotential::SetNB_Method, py::arg("method"))
      .def("GetNB_Method", &PairPotential::GetNB_Method)
      .def("SetNbListAllocateMax", &PairPotential::SetNbListAllocateMax, py::arg("set_allocate_max"))
      .def("SetNbMaxNumNbrs", &PairPotential::SetNbMaxNumNbrs, py::arg("maxNumNbrs"))
      .def("SetAssumeNewton3", &PairPotential::SetAssumeNewton3, py::arg("assume_N3"))
      .def("GetPotentialParameter", &PairPotential::GetPotentialParameter, py::arg("i"), py::arg("j"), py::arg("index"))
      .def("GetParamsTable",
           [](const PairPotential& p, unsigned nt) {
             const std::vector<float> t = p.GetParamsTable(nt);
             py::array_t<float> out({static_cast<py::ssize_t>(nt * nt), static_cast<py::ssize_t>(RUMDX_NUM_PARAM)});
             std::copy(t.begin(), t.end(), out.mutable_data());
             return out;
           },
           py::arg("num_types"))
      .def("GetPotentialForm", &PairPotential::GetPotentialForm)
      .def("SetExclusion", [](PairPotential& p, unsigned i, unsigned j) { p.GetNeighborList().SetExclusion(i, j); },
           py::arg("particleI"), py::arg("particleJ"))
      .def("GetNumberOfExclusions", [](PairPotential& p) { return p.GetNeighborList().GetNumberOfExclusions(); })
      .def("GetParameterList", &PairPotential::GetPotentialParameters, py::arg("num_types"))
      .def("GetPotentialEnergy", &PairPotential::GetPotentialEnergy)
      .def("GetActualMaxNumNbrs", &PairPotential::GetActualMaxNumNbrs)
      .def("GetCutoffMethod", &PairPotential::GetCutoffMethod)
      .def("SetFP64Forces", &PairPotential::SetFP64Forces, py::arg("on"))
      .def("GetFP64Forces", &PairPotential::GetFP64Forces)
      .def("TimeForceKernel", &PairPotential::TimeForceKernel, py::arg("reps"),
           py::call_guard<py::gil_scoped_release>())
      .def("TimeNbListBuild", &PairPotential::TimeNbListBuild, py::arg("reps"), py::arg("full_sort") = false,
           py::call_guard<py::gil_scoped_release>())
      .def("GetNeighborListStats", &PairPotential::GetNeighborListStats)
      .def("SetNbHaloCapacity", [](PairPotential& p, unsigned cap) { p.GetNeighborList().SetHaloCapacity(cap); },
           py::arg("cap"))
      .def("GetNbHaloCapacity", [](PairPotential& p) { return p.GetNeighborList().GetHaloCapacity(); })
      .def("SetNbListInnerSkin", [](PairPotential& p, float s) { p.GetNeighborList().SetInnerSkin(s); },
           py::arg("inner_skin"))
      .def("GetNbListInnerSkin", [](PairPotential& p) { return p.GetNeighborList().GetInnerSkin(); })
      .def("HasNbInnerLists", [](PairPotential& p) { return p.GetNeighborList().HasInnerLists(); })
      .def("SetNbTileWavesPerEU", [](PairPotential& p, unsigned w) { p.GetNeighborList().SetTileWavesPerEU(w); },
           py::arg("waves"))
      .def("GetNbTileWavesPerEU", [](PairPotential& p) { return p.GetNeighborList().GetTileWavesPerEU(); })
      .def("SetNbForceFullSort", [](PairPotential& p, bool on) {
        p.GetNeighborList().SetForceFullSort(on);
        if (p.GetSample()) p.GetSample()->InvalidateGraphs();
      }, py::arg("on"))
      .def("GetNbTileStats", [](PairPotential& p) {
        const std::vector<uint32_t> t = p.GetNeighborList().TileStats();
        py::dict d;
        if (t.size() == 6) {
          d["tiles"] = t[0];
          d["staged"] = t[1];
          d["fast"] = t[2];
          d["wrapping"] = t[3];
          d["shifted"] = t[4];
          d["max_halo"] = t[5];
        }
        return d;
      })
      .def("GetNeighborListCSR", [](PairPotential& p, Sample& s) {
        // neighbour lists in original particle ids, CSR (offsets[n+1], ids)
        s.Synchronize();
        std::vector<uint32_t> cnt;
        const std::vector<uint32_t> lst = p.GetNeighborList().ExportList(cnt);
        const unsigned n = s.GetNumberOfParticles(), nvp = s.GetNumberOfVirtualParticles();
        const std::vector<uint32_t> idx = s.GetParticleData().UnsortedIndex(s.GetStream());
        std::vector<uint32_t> per(n, 0u);
        for (unsigned slot = 0; slot < n; ++slot) per[idx[slot]] = cnt[slot];
        py::array_t<int64_t> off(n + 1);
        auto o = off.mutable_unchecked<1>();
        o(0) = 0;
        for (unsigned i = 0; i < n; ++i) o(i + 1) = o(i) + per[i];
        py::array_t<int32_t> ids(static_cast<py::ssize_t>(o(n)));
        auto d = ids.mutable_unchecked<1>();
        for (unsigned slot = 0; slot < n; ++slot) {
          const uint32_t id = idx[slot];
          for (uint32_t k = 0; k < cnt[slot]; ++k) d(o(id) + k) = static_cast<int32_t>(idx[lst[static_cast<size_t>(k) * nvp + slot]]);
        }
        return py::make_tuple(off, ids);
      }, py::arg("sample"));

  py::class_<Pot_IPL_12_6, PairPotential>(m, "Pot_IPL_12_6")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_IPL_12_6::SetParams, py::arg("i"), py::arg("j"), py::arg("A12"), py::arg("A6"),
           py::arg("Rcut"));
  py::class_<Pot_LJ_12_6, Pot_IPL_12_6>(m, "Pot_LJ_12_6")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_LJ_12_6::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_gLJ_m12_n6, Pot_IPL_12_6>(m, "Pot_gLJ_m12_n6")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_gLJ_m12_n6::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<FormPairPotential, PairPotential>(m, "FormPairPotential");
  py::class_<Pot_LJ_12_6_smooth_2_4, FormPairPotential>(m, "Pot_LJ_12_6_smooth_2_4")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_LJ_12_6_smooth_2_4::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"),
           py::arg("Epsilon"), py::arg("Rcut"));
  py::class_<Pot_gLJ_m_n, FormPairPotential>(m, "Pot_gLJ_m_n")
      .def(py::init<float, float, PairPotential::CutoffMethod>(), py::arg("set_m"), py::arg("set_n"),
           py::arg("cutoff_method"))
      .def("SetParams", &Pot_gLJ_m_n::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_Gauss, FormPairPotential>(m, "Pot_Gauss")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Gauss::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_Buckingham, FormPairPotential>(m, "Pot_Buckingham")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Buckingham::SetParams, py::arg("i"), py::arg("j"), py::arg("rm"), py::arg("Alpha"),
           py::arg("Epsilon"), py::arg("Rcut"));
  py::class_<Pot_IPL_n, FormPairPotential>(m, "Pot_IPL_n")
      .def(py::init<float, PairPotential::CutoffMethod>(), py::arg("n"), py::arg("cutoff_method"))
      .def("SetParams", &Pot_IPL_n::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_Yukawa, FormPairPotential>(m, "Pot_Yukawa")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Yukawa::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_Exp, FormPairPotential>(m, "Pot_Exp")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Exp::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_LJ_12_6_Gauss, FormPairPotential>(m, "Pot_LJ_12_6_Gauss")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_LJ_12_6_Gauss::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"),
           py::arg("Epsilon"), py::arg("Rcut"), py::arg("Sigma0"), py::arg("Epsilon0"), py::arg("r0"));
  py::class_<Pot_IPL_fixed, FormPairPotential>(m, "Pot_IPL_fixed")
      .def("SetParams", &Pot_IPL_fixed::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_IPL_12, Pot_IPL_fixed>(m, "Pot_IPL_12")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"));
  py::class_<Pot_IPL_9, Pot_IPL_fixed>(m, "Pot_IPL_9")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"));
  py::class_<Pot_IPL_18, Pot_IPL_fixed>(m, "Pot_IPL_18")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"));
  py::class_<Pot_IPL6_LN, FormPairPotential>(m, "Pot_IPL6_LN")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_IPL6_LN::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<Pot_LJ_LN, FormPairPotential>(m, "Pot_LJ_LN")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_LJ_LN::SetParams, py::arg("i"), py::arg("j"), py::arg("Rm"), py::arg("R0"),
           py::arg("Epsilon"), py::arg("Rcut"));
  py::class_<Pot_Girifalco, FormPairPotential>(m, "Pot_Girifalco")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Girifalco::SetParams, py::arg("i"), py::arg("j"), py::arg("alpha"), py::arg("beta"),
           py::arg("R0"), py::arg("Rcut"));
  py::class_<Pot_Dzugutov, FormPairPotential>(m, "Pot_Dzugutov")
      .def(py::init<>())
      .def("SetParams", &Pot_Dzugutov::SetParams, py::arg("i"), py::arg("j"), py::arg("a"), py::arg("b"),
           py::arg("c"), py::arg("d"), py::arg("A"), py::arg("B"));
  py::class_<Pot_Fermi_Jagla, FormPairPotential>(m, "Pot_Fermi_Jagla")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Fermi_Jagla::SetParams, py::arg("i"), py::arg("j"), py::arg("a"), py::arg("epsilon0"),
           py::arg("A0"), py::arg("A1"), py::arg("A2"), py::arg("B0"), py::arg("B1"), py::arg("B2"), py::arg("n"),
           py::arg("Rcut"));
  py::class_<Pot_Repulsive_Shoulder, FormPairPotential>(m, "Pot_Repulsive_Shoulder")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_Repulsive_Shoulder::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"),
           py::arg("Epsilon"), py::arg("k0"), py::arg("Sigma1"), py::arg("Rcut"));
  py::class_<Pot_SAAP, FormPairPotential>(m, "Pot_SAAP")
      .def(py::init<PairPotential::CutoffMethod>(), py::arg("cutoff_method"))
      .def("SetParams", &Pot_SAAP::SetParams, py::arg("i"), py::arg("j"), py::arg("Sigma"), py::arg("Epsilon"),
           py::arg("a0"), py::arg("a1"), py::arg("a2"), py::arg("a3"), py::arg("a4"), py::arg("a5"), py::arg("a6"),
           py::arg("Rcut"));
  py::class_<Pot_ShiftedForceCoulomb, FormPairPotential>(m, "Pot_ShiftedForceCoulomb")
      .def(py::init<>())
      .def("SetParams", &Pot_ShiftedForceCoulomb::SetParams, py::arg("i"), py::arg("j"), py::arg("Epsilon"),
           py::arg("Rcut"));
  py::class_<BondHarmonic, Potential>(m, "BondHarmonic")
      .def(py::init<>())
      .def("SetParams", &BondHarmonic::SetParams, py::arg("bond_type"), py::arg("bond_length"), py::arg("stiffness"),
           py::arg("exclude") = false)
      .def("GetParams", &BondHarmonic::GetParams)
      .def("GetExcludedTypes", &BondHarmonic::GetExcludedTypes)
      .def("SetExclusions", &BondHarmonic::SetExclusions, py::arg("non_bond_pot"));
  py::class_<BondFENE, BondHarmonic>(m, "BondFENE").def(py::init<>());
  py::class_<DihedralRyckaert, Potential>(m, "DihedralRyckaert")
      .def(py::init<>())
      .def("SetParams", &DihedralRyckaert::SetParams, py::arg("dihedral_type"), py::arg("coeffs"))
      .def("SetExclude", &DihedralRyckaert::SetExclude, py::arg("exclude"))
      .def("SetExclusions", &DihedralRyckaert::SetExclusions, py::arg("non_bond_pot"));
  py::class_<PeriodicDihedral, DihedralRyckaert>(m, "PeriodicDihedral").def(py::init<>());
  py::class_<AngleCosSq, Potential>(m, "AngleCosSq")
      .def(py::init<>())
      .def("SetParams", &AngleCosSq::SetParams, py::arg("angle_type"), py::arg("theta0"), py::arg("ktheta"))
      .def("GetParams", &AngleCosSq::GetParams)
      .def("SetExclude", &AngleCosSq::SetExclude, py::arg("exclude"))
      .def("SetExclusions", &AngleCosSq::SetExclusions, py::arg("non_bond_pot"));
  py::class_<AngleSq, AngleCosSq>(m, "AngleSq").def(py::init<>());
  py::class_<MoleculeData>(m, "MoleculeData")
      .def("ReadTopology", &MoleculeData::ReadTopology, py::arg("filename"))
      .def("SetBonds", &MoleculeData::SetBonds, py::arg("abt"), py::arg("molecule") = std::vector<uint32_t>{})
      .def("SetAngles", &MoleculeData::SetAngles, py::arg("abct"))
      .def("SetDihedrals", &MoleculeData::SetDihedrals, py::arg("abcdt"))
      .def("GetNumberOfDihedrals", &MoleculeData::GetNumberOfDihedrals)
      .def("GetDihedrals", &MoleculeData::GetDihedrals)
      .def("GetNumberOfBonds", &MoleculeData::GetNumberOfBonds)
      .def("GetNumberOfAngles", &MoleculeData::GetNumberOfAngles)
      .def("GetNumberOfMolecules", &MoleculeData::GetNumberOfMolecules)
      .def("GetBonds", &MoleculeData::GetBonds)
      .def("GetAngles", &MoleculeData::GetAngles);
  m.attr("NoShift") = PairPotential::NS;
  m.attr("ShiftedPotential") = PairPotential::SP;
  m.attr("ShiftedForce") = PairPotential::SF;

  py::class_<Integrator>(m, "Integrator")
      .def("Integrate", &Integrator::Integrate)
      .def("SetTimeStep", &Integrator::SetTimeStep, py::arg("dt"))
      .def("GetTimeStep", &Integrator::GetTimeStep)
      .def("GetKineticEnergy", &Integrator::GetKineticEnergy, py::arg("copy") = true)
      .def("SetMomentumToZero", &Integrator::SetMomentumToZero)
      .def("GetInfoString", &Integrator::GetInfoString, py::arg("precision") = 9)
      .def("InitializeFromInfoString", &Integrator::InitializeFromInfoString, py::arg("infoStr"),
           py::arg("verbose") = true)
      .def("GetTotalMomentum", &Integrator::GetTotalMomentum);
  py::class_<IntegratorNVT, Integrator>(m, "IntegratorNVT")
      .def(py::init<float, double, double>(), py::arg("timeStep"), py::arg("targetTemperature"),
           py::arg("thermostatRelaxationTime") = 0.2)
      .def(py::init<float>(), py::arg("timeStep"))
      .def("SetRelaxationTime", &IntegratorNVT::SetRelaxationTime, py::arg("tau"))
      .def("GetRelaxationTime", &IntegratorNVT::GetRelaxationTime)
      .def("SetTargetTemperature", &IntegratorNVT::SetTargetTemperature, py::arg("T"), py::arg("reset") = true)
      .def("GetTargetTemperature", &IntegratorNVT::GetTargetTemperature)
      .def("SetThermostatState", &IntegratorNVT::SetThermostatState, py::arg("Ps"))
      .def("GetThermostatState", &IntegratorNVT::GetThermostatState)
      .def("GetThermostatKineticEnergy", &IntegratorNVT::GetThermostatKineticEnergy)
      .def("SetThermostatOn", &IntegratorNVT::SetThermostatOn, py::arg("on") = true)
      .def("SetThermostatOnParticleType", &IntegratorNVT::SetThermostatOnParticleType, py::arg("type"),
           py::arg("on"))
      .def("SetTemperatureChangePerTimeStep", &IntegratorNVT::SetTemperatureChangePerTimeStep,
           py::arg("temperatureChangePerTimeStep"));
  // %rename (IntegratorNVE) IntegratorNVT(float timeStep) (Swig/Integrators.i:16)
  m.def("IntegratorNVE", [](float timeStep) { return new IntegratorNVT(timeStep); }, py::arg("timeStep"),
        py::return_value_policy::take_ownership);

  py::class_<Sample>(m, "Sample")
      .def(py::init<unsigned, unsigned>(), py::arg("pb") = 0, py::arg("tp") = 0)
      .def("SetNumberOfParticles", &Sample::SetNumberOfParticles, py::arg("num_part"))
      .def("GetNumberOfParticles", &Sample::GetNumberOfParticles)
      .def("GetNumberOfVirtualParticles", &Sample::GetNumberOfVirtualParticles)
      .def("GetNumberOfTypes", &Sample::GetNumberOfTypes)
      .def("GetNumberThisType", &Sample::GetNumberThisType, py::arg("type"))
      .def("SetVerbose", &Sample::SetVerbose, py::arg("vb"))
      .def("SetSimulationBox", &Sample::SetSimulationBox, py::arg("simBox"), py::keep_alive<1, 2>())
      .def("GetSimulationBox", &Sample::GetSimulationBox, py::return_value_policy::reference_internal)
      .def("SetPotential", &Sample::SetPotential, py::arg("potential"), py::keep_alive<1, 2>())
      .def("AddPotential", &Sample::AddPotential, py::arg("potential"), py::keep_alive<1, 2>())
      .def("SetIntegrator", &Sample::SetIntegrator, py::arg("newItg"), py::keep_alive<1, 2>())
      .def("GetIntegrator", &Sample::GetIntegrator, py::return_value_policy::reference_internal)
      .def("GetPotentials", [](Sample& s) { return s.GetPotentials(); }, py::return_value_policy::reference_internal)
      .def("ReadMoleculeData", &Sample::ReadMoleculeData, py::arg("topology_file"))
      .def("GetMoleculeData", &Sample::GetMoleculeData, py::return_value_policy::reference_internal)
      .def("CreateMoleculeData", &Sample::CreateMoleculeData, py::return_value_policy::reference_internal)
      .def("SetPB_TP", &Sample::SetPB_TP, py::arg("pb"), py::arg("tp"))
      .def("GetParticlesPerBlock", &Sample::GetParticlesPerBlock)
      .def("GetThreadsPerParticle", &Sample::GetThreadsPerParticle)
      .def("SetMixedPrecisionPositions", &Sample::SetMixedPrecisionPositions, py::arg("on"))
      .def("GetMixedPrecisionPositions", &Sample::GetMixedPrecisionPositions)
      .def("SetFusedSteps", &Sample::SetFusedSteps, py::arg("on"))
      .def("GetFusedSteps", &Sample::GetFusedSteps)
      .def("UsesFusedSteps", [](Sample& s) { return s.FusedStepPotential() != nullptr; })
      .def("TimeFusedStepKernel", &Sample::TimeFusedStepKernel, py::arg("reps"),
           py::call_guard<py::gil_scoped_release>())
      .def("CalcF", &Sample::CalcF, py::arg("calc_stresses") = false)
      .def("GetPotentialEnergy", &Sample::GetPotentialEnergy)
      .def("GetVirial", &Sample::GetVirial)
      .def("GetStress", &Sample::GetStress)
      .def("GetNumberOfDOFs", &Sample::GetNumberOfDOFs)
      .def("SetNumberOfDOFs", &Sample::SetNumberOfDOFs, py::arg("DOFs"))
      .def("SetMass", &Sample::SetMass, py::arg("type"), py::arg("mass"))
      .def("GetMass", &Sample::GetMass, py::arg("type"))
      .def("GetMeanMass", &Sample::GetMeanMass)
      .def("SortParticles", &Sample::SortParticles)
      .def("Synchronize", &Sample::Synchronize)
      .def("RunSteps", &Sample::RunSteps, py::arg("first_step"), py::arg("nsteps"), py::arg("block") = 10,
           py::arg("momentum_interval") = 100, py::call_guard<py::gil_scoped_release>())
      .def("RunSegments", &Sample::RunSegments, py::arg("first_step"), py::arg("nsteps"), py::arg("block") = 10,
           py::arg("momentum_interval") = 100)
      .def("PrepareRunSteps", &Sample::PrepareRunSteps, py::arg("first_step"), py::arg("nsteps"),
           py::arg("block") = 10, py::arg("momentum_interval") = 100, py::call_guard<py::gil_scoped_release>())
      .def("GetPositions", [](Sample& s) { return get_xyz(s, "r"); })
      // Nx3 float64: r + rlo in mixed-precision mode, r widened otherwise
      .def("GetPositionsFP64", [](Sample& s) {
        const unsigned n = s.GetNumberOfParticles();
        if (!n) throw RUMD_Error("Sample", "GetPositionsFP64", "Sample has no particles");
        s.Synchronize();
        std::vector<float> hi(4ull * n), lo;
        s.GetParticleData().Get("r", hi.data(), s.GetStream());
        if (s.GetMixedPrecisionPositions()) {
          lo.resize(4ull * n);
          s.GetParticleData().Get("rlo", lo.data(), s.GetStream());
        }
        py::array_t<double> out({static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(3)});
        auto o = out.mutable_unchecked<2>();
        for (unsigned i = 0; i < n; ++i)
          for (int k = 0; k < 3; ++k)
            o(i, k) = static_cast<double>(hi[4ull * i + k]) + (lo.empty() ? 0.0 : static_cast<double>(lo[4ull * i + k]));
        return out;
      })
      .def("GetVelocities", [](Sample& s) { return get_xyz(s, "v"); })
      .def("GetForces", [](Sample& s) { return get_xyz(s, "f"); })
      .def("GetImages", [](Sample& s) { return get_xyz(s, "im"); })
      .def("GetForcesAndEnergies", [](Sample& s) { return get_x4(s, "f"); })
      .def("GetVirials", [](Sample& s) { return get_x4(s, "w"); })
      .def("GetStresses", [](Sample& s) { return get_x4(s, "sts"); })
      .def("GetTypes", [](Sample& s) {
        py::array_t<int32_t> out(s.GetNumberOfParticles());
        s.GetParticleData().GetTypes(out.mutable_data(), s.GetStream());
        return out;
      })
      .def("SetPositions", [](Sample& s, const f32arr& a) {
        s.Synchronize();
        s.GetParticleData().SetPositions(check_nx3(s, a), s.GetStream());
        reset_lists(s);
      }, py::arg("positions"))
      .def("SetPositionsFP64", [](Sample& s, const py::array_t<double, py::array::c_style | py::array::forcecast>& a) {
        if (a.ndim() != 2 || a.shape(0) != s.GetNumberOfParticles() || a.shape(1) != 3)
          throw py::value_error("Wrong array shape, must be Nx3");
        s.Synchronize();
        s.GetParticleData().SetPositionsFP64(a.data(), s.GetStream());
        reset_lists(s);
      }, py::arg("positions"))
      .def("SetVelocities", [](Sample& s, const f32arr& a) {
        s.Synchronize();
        s.GetParticleData().SetVelocities(check_nx3(s, a), s.GetStream());
      }, py::arg("velocities"))
      .def("SetImages", [](Sample& s, const f32arr& a) {
        s.Synchronize();
        s.GetParticleData().SetImages(check_nx3(s, a), s.GetStream());
      }, py::arg("images"))
      .def("SetTypes", [](Sample& s, const i32arr& t) {
        if (t.ndim() != 1 || t.shape(0) != s.GetNumberOfParticles())
          throw py::value_error("Wrong vector length, must be N (number of particles)");
        s.Synchronize();
        s.GetParticleData().SetTypes(t.data(), s.GetStream());
        for (Potential* p : s.GetPotentials()) p->SetSample(&s);
        s.InvalidateGraphs();
      }, py::arg("types"));

  // Batched replicas (no counterpart in RUMD's SWIG module).
  py::class_<ReplicaBatch>(m, "ReplicaBatch")
      .def(py::init<unsigned>(), py::arg("num_types"))
      .def("AddReplica", [](ReplicaBatch& b, const f32arr& pos, const i32arr& types, const f32arr& vel, float L) {
        if (pos.ndim() != 2 || pos.shape(1) != 3 || vel.ndim() != 2 || vel.shape(1) != 3 || types.ndim() != 1 ||
            pos.shape(0) != types.shape(0) || vel.shape(0) != types.shape(0))
          throw py::value_error("positions and velocities must be N x 3 and types N");
        const size_t n = static_cast<size_t>(types.shape(0));
        std::vector<float> p(pos.data(), pos.data() + 3 * n), v(vel.data(), vel.data() + 3 * n);
        std::vector<unsigned> t(n);
        for (size_t i = 0; i < n; ++i) {
          if (types.data()[i] < 0) throw py::value_error("negative type");
          t[i] = static_cast<unsigned>(types.data()[i]);
        }
        return b.AddReplica(p, t, v, L);
      }, py::arg("positions"), py::arg("types"), py::arg("velocities"), py::arg("box_length"))
      .def("SetMass", &ReplicaBatch::SetMass, py::arg("type"), py::arg("mass"))
      .def("SetPotential", &ReplicaBatch::SetPotential, py::arg("potential"), py::arg("skin"))
      .def("SetTimeStep", &ReplicaBatch::SetTimeStep, py::arg("dt"))
      .def("SetThermostat", &ReplicaBatch::SetThermostat, py::arg("replica"), py::arg("targetTemperature"),
           py::arg("thermostatRelaxationTime"))
      .def("SetMaxNumNbrs", &ReplicaBatch::SetMaxNumNbrs, py::arg("max_nbrs"))
      .def("GetMaxNumNbrs", &ReplicaBatch::GetMaxNumNbrs)
      .def("SetLanesPerParticle", &ReplicaBatch::SetLanesPerParticle, py::arg("tp"))
      .def("SetStagedForces", &ReplicaBatch::SetStagedForces, py::arg("on"))
      .def("GetStagedForces", &ReplicaBatch::GetStagedForces)
      .def("UsesStagedForces", &ReplicaBatch::UsesStagedForces)
      .def("SetWorkgroupSort", &ReplicaBatch::SetWorkgroupSort, py::arg("on"))
      .def("GetWorkgroupSort", &ReplicaBatch::GetWorkgroupSort)
      .def("UsesWorkgroupSort", &ReplicaBatch::UsesWorkgroupSort)
      .def("SetFusedSteps", &ReplicaBatch::SetFusedSteps, py::arg("on"))
      .def("GetFusedSteps", &ReplicaBatch::GetFusedSteps)
      .def("UsesFusedSteps", &ReplicaBatch::UsesFusedSteps)
      .def("GetLanesPerParticle", &ReplicaBatch::GetLanesPerParticle)
      .def("GetBoxLength", &ReplicaBatch::GetBoxLength, py::arg("replica"))
      .def("GetThermostat", &ReplicaBatch::GetThermostat, py::arg("replica"))
      .def("GetTimeStep", &ReplicaBatch::GetTimeStep)
      .def("GetSkin", &ReplicaBatch::GetSkin)
      .def("GetCutoffMethod", &ReplicaBatch::GetCutoffMethod)
      .def("GetParamsTable", &ReplicaBatch::GetParamsTable)
      .def("SaveState", &ReplicaBatch::SaveState)
      .def("RestoreState", &ReplicaBatch::RestoreState)
      .def("CalcF", &ReplicaBatch::CalcF)
      .def("TimeForceKernel", &ReplicaBatch::TimeForceKernel, py::arg("reps"),
           py::call_guard<py::gil_scoped_release>())
      .def("GetNeighborListStats", &ReplicaBatch::GetNeighborListStats)
      .def("RunSteps", &ReplicaBatch::RunSteps, py::arg("nsteps"), py::arg("block") = 10,
           py::call_guard<py::gil_scoped_release>())
      .def("Synchronize", &ReplicaBatch::Synchronize)
      .def("GetNumberOfReplicas", &ReplicaBatch::GetNumberOfReplicas)
      .def("GetNumberOfParticles", &ReplicaBatch::GetNumberOfParticles, py::arg("replica"))
      .def("GetTotalParticles", &ReplicaBatch::GetTotalParticles)
      .def("GetNumberOfSlots", &ReplicaBatch::GetNumberOfSlots)
      .def("GetPositions", [](ReplicaBatch& b, unsigned r) {
        auto v = b.GetPositions(r);
        py::array_t<float> out({static_cast<py::ssize_t>(v.size() / 3), static_cast<py::ssize_t>(3)});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
      }, py::arg("replica"))
      .def("GetVelocities", [](ReplicaBatch& b, unsigned r) {
        auto v = b.GetVelocities(r);
        py::array_t<float> out({static_cast<py::ssize_t>(v.size() / 3), static_cast<py::ssize_t>(3)});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
      }, py::arg("replica"))
      .def("GetForcesAndEnergies", [](ReplicaBatch& b, unsigned r) {
        auto v = b.GetForces(r);
        py::array_t<float> out({static_cast<py::ssize_t>(v.size() / 4), static_cast<py::ssize_t>(4)});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
      }, py::arg("replica"))
      .def("GetEnergies", [](ReplicaBatch& b) {
        auto v = b.GetEnergies();
        py::array_t<double> out({static_cast<py::ssize_t>(v.size() / 4), static_cast<py::ssize_t>(4)});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
      })
      .def("GetThermostatStates", &ReplicaBatch::GetThermostatStates)
      .def("GetNumberOfBuilds", &ReplicaBatch::GetNumberOfBuilds)
      .def("GetNeighborListCSR", [](ReplicaBatch& b, unsigned r) {
        std::vector<uint32_t> off, ids;
        b.GetNeighborList(r, off, ids);
        py::array_t<uint32_t> o(off.size()), d(ids.size());
        std::copy(off.begin(), off.end(), o.mutable_data());
        std::copy(ids.begin(), ids.end(), d.mutable_data());
        return py::make_tuple(o, d);
      }, py::arg("replica"));
}
