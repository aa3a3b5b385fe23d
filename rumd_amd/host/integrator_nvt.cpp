// IntegratorNVT host logic (include/rumd/IntegratorNVT.h, src/IntegratorNVT.cu).
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "rumd_host.h"

namespace {
const char* kNoSample = "There is no integrator associated with Sample or the latter has no particles";

// ParseInfoString (src/ParseInfoString.cc): "Name,p1,p2,..."
std::string parse_info(const std::string& info, std::vector<float>& params) {
  std::stringstream ss(info);
  std::string tok, name;
  std::getline(ss, name, ',');
  while (std::getline(ss, tok, ',')) params.push_back(std::stof(tok));
  return name;
}
}  // namespace

std::vector<double> Integrator::GetTotalMomentum() {
  if (!S || !S->GetNumberOfParticles()) throw RUMD_Error("Integrator", "GetTotalMomentum", kNoSample);
  const unsigned n = S->GetNumberOfParticles();
  std::vector<float> v(4ull * n);
  S->GetParticleData().Get("v", v.data(), S->GetStream());
  std::vector<double> p(4, 0.0);
  for (unsigned i = 0; i < n; ++i) {
    const double m = 1.0 / v[4ull * i + 3];
    for (int k = 0; k < 3; ++k) p[k] += m * v[4ull * i + k];
    p[3] += m;
  }
  return p;
}

IntegratorNVT::IntegratorNVT(float dt) {
  timeStep = dt;
  thermostat_on_ = 0u;
  relaxation_time_ = 0.2f;
  host_state_ = {0.0, 0.0, 0.0, 0.0};
}

IntegratorNVT::IntegratorNVT(float dt, double targetTemperature, double thermostatRelaxationTime) {
  timeStep = dt;
  thermostat_on_ = ~0u;
  relaxation_time_ = static_cast<float>(thermostatRelaxationTime);
  host_state_ = {0.0, targetTemperature, 0.0, 0.0};
}

void IntegratorNVT::SetSample(Sample* s) {
  S = s;
  if (d_state_.size() == 0) d_state_.resize(1);
  PushState();
  if (S) S->InvalidateGraphs();
}

void IntegratorNVT::PushState() {
  if (d_state_.size()) d_state_.upload(&host_state_, 1, S ? S->GetStream() : nullptr);
}

void IntegratorNVT::PullState() {
  if (d_state_.size()) d_state_.download(&host_state_, 1, S ? S->GetStream() : nullptr);
}

void IntegratorNVT::Integrate() {
  if (!S || !S->GetNumberOfParticles()) throw RUMD_Error("IntegratorNVT", "Integrate", kNoSample);
  if (!S->GetSimulationBox()) throw RUMD_Error("IntegratorNVT", "Integrate", "no simulation box");
  // omega^2 = 4 pi^2 / tau^2 formed in double, passed as float (src/IntegratorNVT.cu:88)
  const float omega2 = static_cast<float>(4. * M_PI * M_PI / (static_cast<double>(relaxation_time_) * relaxation_time_));
  NeighborList* nl = S->FusedCheckList();
  rumdx_check(rumdx_nvt_step(S->GetParticleData().Descriptor(), d_state_.get(), timeStep, thermostat_on_, omega2,
                             S->GetNumberOfDOFs(), S->GetSimulationBox()->GetBox8(), nl ? nl->Descriptor() : nullptr,
                             S->Workspace(), S->GetStream()),
              "IntegratorNVT", "Integrate");
  if (nl) nl->NoteFusedCheck();
}

void IntegratorNVT::IntegrateFused(PairPotential& pp) {
  if (!S || !S->GetNumberOfParticles()) throw RUMD_Error("IntegratorNVT", "Integrate", kNoSample);
  if (!S->GetSimulationBox()) throw RUMD_Error("IntegratorNVT", "Integrate", "no simulation box");
  const float omega2 = static_cast<float>(4. * M_PI * M_PI / (static_cast<double>(relaxation_time_) * relaxation_time_));
  pp.CalcFIntegrateNVT(d_state_.get(), timeStep, thermostat_on_, omega2, S->GetNumberOfDOFs());
  pp.GetNeighborList().NoteFusedCheck();
}

std::string IntegratorNVT::GetInfoString(unsigned precision) const {
  std::ostringstream s;
  if (thermostat_on_) {
    const_cast<IntegratorNVT*>(this)->PullState();
    s << "IntegratorNVT," << std::setprecision(precision) << timeStep << "," << std::setprecision(precision)
      << host_state_.target_T << "," << std::setprecision(precision) << relaxation_time_ << ","
      << std::setprecision(precision) << host_state_.ps;
  } else {
    s << "IntegratorNVE," << std::setprecision(precision) << timeStep;
  }
  return s.str();
}

void IntegratorNVT::InitializeFromInfoString(const std::string& info, bool verbose) {
  std::vector<float> p;
  const std::string name = parse_info(info, p);
  const bool nvt = name == "IntegratorNVT";
  if (!nvt && name != "IntegratorNVE")
    throw RUMD_Error("IntegratorNVT", "InitializeFromInfoString", "Expected IntegratorNVT or IntegratorNVE");
  if (p.size() != (nvt ? 4u : 1u))
    throw RUMD_Error("IntegratorNVT", "InitializeFromInfoString", "Wrong number of parameters in infoStr");
  if (verbose) std::cout << "[Info] Initializing " << name << ". time step:" << p[0] << std::endl;
  SetTimeStep(p[0]);
  if (nvt) {
    SetTargetTemperature(p[1]);
    SetRelaxationTime(p[2]);
    SetThermostatState(p[3]);
  }
}

void IntegratorNVT::SetMomentumToZero() {
  if (!S || !S->GetNumberOfParticles()) throw RUMD_Error("IntegratorNVT", "SetMomentumToZero", kNoSample);
  rumdx_check(rumdx_zero_momentum(S->GetParticleData().Descriptor(), S->Workspace(), S->GetStream()),
              "IntegratorNVT", "SetMomentumToZero");
}

double IntegratorNVT::GetKineticEnergy(bool) {
  if (!S || !S->GetNumberOfParticles()) throw RUMD_Error("IntegratorNVT", "GetKineticEnergy", "Sample has no particles");
  rumdx_check(rumdx_kinetic_energy(S->GetParticleData().Descriptor(), d_state_.get(), timeStep, thermostat_on_,
                                   S->ScalarOut(), S->Workspace(), S->GetStream()),
              "IntegratorNVT", "GetKineticEnergy");
  double ke = 0.0;
  hip_check(hipMemcpyAsync(&ke, S->ScalarOut(), sizeof(double), hipMemcpyDeviceToHost, S->GetStream()),
            "IntegratorNVT", "GetKineticEnergy");
  hip_check(hipStreamSynchronize(S->GetStream()), "IntegratorNVT", "GetKineticEnergy");
  return ke;
}

void IntegratorNVT::SetTimeStep(float dt) {
  timeStep = dt;
  if (S) S->InvalidateGraphs();
}

void IntegratorNVT::SetRelaxationTime(float tau) {
  if (!(tau > 0.f)) throw RUMD_Error("IntegratorNVT", "SetRelaxationTime", "relaxation time must be positive");
  relaxation_time_ = tau;
  SetThermostatState(0.0);
  if (S) S->InvalidateGraphs();
}

void IntegratorNVT::SetTargetTemperature(double T, bool reset) {
  PullState();
  host_state_.target_T = T;
  if (reset) host_state_.ps = 0.0;
  PushState();
}

double IntegratorNVT::GetTargetTemperature() {
  PullState();
  return host_state_.target_T;
}

void IntegratorNVT::SetThermostatState(double ps) {
  PullState();
  host_state_.ps = ps;
  PushState();
}

double IntegratorNVT::GetThermostatState() {
  if (!thermostat_on_) return 0.0;
  PullState();
  return host_state_.ps;
}

double IntegratorNVT::GetThermostatKineticEnergy() {
  PullState();
  return 0.5 * host_state_.ke2_sum;
}

void IntegratorNVT::SetThermostatOn(bool on) {
  thermostat_on_ = on ? ~0u : 0u;
  if (S) S->InvalidateGraphs();
}

void IntegratorNVT::SetThermostatOnParticleType(unsigned type, bool on) {
  if (type > 31) throw RUMD_Error("IntegratorNVT", "SetThermostatOnParticleType", "The particle type number is larger than 31");
  if (on) thermostat_on_ |= (1u << type); else thermostat_on_ &= ~(1u << type);
  if (S) S->InvalidateGraphs();
}

void IntegratorNVT::SetTemperatureChangePerTimeStep(double dT) {
  PullState();
  host_state_.dT_per_step = dT;
  PushState();
}
