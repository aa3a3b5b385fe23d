// Sample (include/rumd/Sample.h, src/Sample.cu): CalcF over the potential
// list, FP64 sums, and the captured step blocks of the run loop.
#include <algorithm>
#include <iostream>

#include "rumd_host.h"

namespace {
constexpr unsigned kPbThreshold = 10000;  // src/Sample.cu:36-39
}

Sample::Sample(unsigned pb, unsigned tp) : pb_(pb), tp_(tp) {
  Device::GetDevice().Init();
  hip_check(hipStreamCreate(&stream_), "Sample", "Sample");
  scalars_.resize(8);
}

Sample::~Sample() {
  InvalidateGraphs();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Sample::SetPB_TP(unsigned pb, unsigned tp) {
  if (pb * std::max(tp, 1u) > 1024)
    throw RUMD_cudaResourceError("Sample", "SetPB_TP", "Too many threads per block (pb*tp>maxthreads)");
  pb_ = pb;
  tp_ = tp;
  if (GetNumberOfParticles()) {
    // reallocate with the new padding, keeping the configuration
    const unsigned n = GetNumberOfParticles();
    std::vector<float> r(4ull * n), v(4ull * n), im(4ull * n);
    std::vector<int32_t> t(n);
    pd_.Get("r", r.data(), stream_);
    pd_.Get("v", v.data(), stream_);
    pd_.Get("im", im.data(), stream_);
    pd_.GetTypes(t.data(), stream_);
    const unsigned nt = pd_.GetNumberOfTypes();
    std::vector<float> masses(nt);
    for (unsigned k = 0; k < nt; ++k) masses[k] = pd_.GetMass(k);
    SetNumberOfParticles(n);
    for (unsigned k = 0; k < nt; ++k) pd_.SetMass(k, masses[k]);
    pd_.SetTypes(t.data(), stream_);
    std::vector<float> x3(3ull * n);
    for (unsigned i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) x3[3ull * i + k] = r[4ull * i + k];
    pd_.SetPositions(x3.data(), stream_);
    for (unsigned i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) x3[3ull * i + k] = v[4ull * i + k];
    pd_.SetVelocities(x3.data(), stream_);
    for (unsigned i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) x3[3ull * i + k] = im[4ull * i + k];
    pd_.SetImages(x3.data(), stream_);
  }
}

void Sample::SetMixedPrecisionPositions(bool on) {
  if (on == mixed_) return;
  mixed_ = on;
  if (GetNumberOfParticles()) SetPB_TP(pb_, tp_);
}

void Sample::SetNumberOfParticles(unsigned n) {
  if (n == 0) return;
  if (pb_ == 0) pb_ = n > kPbThreshold ? 64 : 32;
  if (tp_ == 0) tp_ = n > kPbThreshold ? 1 : 4;
  InvalidateGraphs();
  pd_.SetNumberOfParticles(n, pb_, mixed_);
  workspace_.resize(rumdx_workspace_bytes(n));
  if (verbose_) {
    std::cout << "Setting number of particles: N = " << n << ", pb = " << pb_ << ", tp = " << tp_
              << ", virtual particles = " << pd_.GetNumberOfVirtualParticles() << std::endl;
  }
  for (Potential* p : potentials_) p->SetSample(this);
  if (itg_) itg_->SetSample(this);
}

void Sample::SetSimulationBox(SimulationBox* box) {
  box_ = box;
  InvalidateGraphs();
  if (GetNumberOfParticles())
    for (Potential* p : potentials_) p->SetSample(this);
}

void Sample::SetPotential(Potential* p) {
  potentials_.clear();
  AddPotential(p);
}

void Sample::AddPotential(Potential* p) {
  p->SetVerbose(verbose_);
  // non-constraint potentials go to the front (src/Sample.cu:268-274), pair
  // potentials ahead of the others: the first force call of a step then is a
  // pair potential's, whose list build may sort the particles, before any
  // other term (bonds, angles) has been written in the old order
  auto pos = potentials_.begin();
  if (!dynamic_cast<PairPotential*>(p))
    while (pos != potentials_.end() && dynamic_cast<PairPotential*>(*pos)) ++pos;
  potentials_.insert(pos, p);
  if (GetNumberOfParticles()) p->SetSample(this);
  InvalidateGraphs();
}

void Sample::SetIntegrator(Integrator* itg) {
  itg_ = itg;
  if (GetNumberOfParticles()) itg_->SetSample(this);
  InvalidateGraphs();
}

unsigned Sample::GetNumberOfDOFs() const {
  if (user_dofs_) return user_dofs_;
  const unsigned n = GetNumberOfParticles();
  return n ? 3 * n - 3 : 0;  // src/Sample.cu:294-310
}

void Sample::SetMass(unsigned type, float mass) { pd_.SetMass(type, mass); }

float Sample::GetMeanMass() const {
  double m = 0.0;
  for (unsigned t = 0; t < GetNumberOfTypes(); ++t) m += GetNumberThisType(t) * GetMass(t);
  return static_cast<float>(m / GetNumberOfParticles());
}

NeighborList* Sample::FusedCheckList() {
  if (potentials_.empty()) return nullptr;
  auto* pp = dynamic_cast<PairPotential*>(potentials_.front());
  if (!pp || pp->GetNeighborList().GetResolvedMethod() == NeighborList::NONE) return nullptr;
  return &pp->GetNeighborList();
}

void Sample::CalcF(bool calc_stresses) {
  if (!box_) throw RUMD_Error("Sample", "CalcF", "No simulation box has been set");
  if (!GetNumberOfParticles()) throw RUMD_Error("Sample", "CalcF", "Sample has no particles");
  if (potentials_.empty()) std::cerr << "Warning [Sample::CalcF]: No potential has been set" << std::endl;
  BumpSlotEpoch();
  for (size_t k = 0; k < potentials_.size(); ++k) potentials_[k]->CalcF(k == 0, calc_stresses);
}

void Sample::Synchronize() {
  hip_check(hipStreamSynchronize(stream_), "Sample", "Synchronize");
  for (Potential* p : potentials_)
    if (auto* pp = dynamic_cast<PairPotential*>(p))
      if (pp->GetNeighborList().HandleOverflow()) InvalidateGraphs();
}

// FP64 device sums of the per-particle energies / virials (the host loops of
// Sample::GetPotentialEnergy / GetVirial, src/Sample.cu:353-391).
double Sample::GetPotentialEnergy() {
  Synchronize();
  rumdx_check(rumdx_reduce_f4(pd_.d_f.get(), GetNumberOfParticles(), 8u, scalars_.get(), workspace_.get(), stream_),
              "Sample", "GetPotentialEnergy");
  double out[4];
  scalars_.download(out, 4, stream_);
  double pe = out[3];
  for (Potential* p : potentials_)
    if (!p->EnergyIncludedInParticleSum()) pe += p->GetPotentialEnergy();
  return pe;
}

double Sample::GetVirial() {
  Synchronize();
  rumdx_check(rumdx_reduce_f4(pd_.d_w.get(), GetNumberOfParticles(), 8u, scalars_.get(), workspace_.get(), stream_),
              "Sample", "GetVirial");
  double out[4];
  scalars_.download(out, 4, stream_);
  double w = out[3] / 6.0;
  for (Potential* p : potentials_)
    if (!p->EnergyIncludedInParticleSum()) w += p->GetVirial();
  return w;
}

// Sample::GetStress (src/Sample.cu:393-444), kinetic part included.
std::vector<double> Sample::GetStress() {
  Synchronize();
  const unsigned n = GetNumberOfParticles();
  double s[6] = {0, 0, 0, 0, 0, 0};
  double out[4];
  rumdx_check(rumdx_reduce_f4(pd_.d_sts.get(), n, 15u, scalars_.get(), workspace_.get(), stream_), "Sample", "GetStress");
  scalars_.download(out, 4, stream_);
  for (int k = 0; k < 4; ++k) s[k] = out[k];
  rumdx_check(rumdx_reduce_f4(pd_.d_w.get(), n, 6u, scalars_.get(), workspace_.get(), stream_), "Sample", "GetStress");
  scalars_.download(out, 4, stream_);
  s[4] = out[1];
  s[5] = out[2];
  std::vector<float> v(4ull * n);
  pd_.Get("v", v.data(), stream_);
  for (unsigned i = 0; i < n; ++i) {
    const float m2 = 2.f / v[4ull * i + 3];
    const float vx = v[4ull * i], vy = v[4ull * i + 1], vz = v[4ull * i + 2];
    s[0] -= m2 * vx * vx;
    s[1] -= m2 * vy * vy;
    s[2] -= m2 * vz * vz;
    s[3] -= m2 * vy * vz;
    s[4] -= m2 * vx * vz;
    s[5] -= m2 * vx * vy;
  }
  const double V = box_->GetVolume();
  std::vector<double> r(6);
  for (int k = 0; k < 6; ++k) r[k] = 0.5 * s[k] / V;
  return r;
}

// Spatial sorting on demand (Sample::SortParticles): the sorted neighbour
// list reorders particles by cell at every rebuild, so a sort is a rebuild.
void Sample::SortParticles() {
  for (Potential* p : potentials_)
    if (auto* pp = dynamic_cast<PairPotential*>(p))
      if (pp->GetNeighborList().IsSorted()) pp->GetNeighborList().ResetNeighborList();
}

void Sample::InvalidateGraphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second.exec);
  graphs_.clear();
}

void Sample::SetFusedSteps(bool on) {
  if (on == fused_steps_) return;
  fused_steps_ = on;
  InvalidateGraphs();
}

float Sample::TimeFusedStepKernel(unsigned reps) {
  PrepareRun();
  PairPotential* pp = FusedStepPotential();
  if (pp == nullptr) throw RUMD_Error("Sample", "TimeFusedStepKernel", "the steps of this sample take no fused kernel");
  return pp->TimeFusedStepKernel(reps, itg_->GetTimeStep());
}

PairPotential* Sample::FusedStepPotential() {
  if (!fused_steps_ || potentials_.size() != 1 || dynamic_cast<IntegratorNVT*>(itg_) == nullptr) return nullptr;
  auto* pp = dynamic_cast<PairPotential*>(potentials_.front());
  return (pp != nullptr && pp->CanFuseStep()) ? pp : nullptr;
}

void Sample::PrepareRun() {
  if (!box_) throw RUMD_Error("Sample", "Run", "No simulation box has been set");
  if (!itg_) throw RUMD_Error("Sample", "Run", "Integrator has not been set");
  if (potentials_.empty()) throw RUMD_Error("Sample", "Run", "Potential has not been set");
  for (Potential* p : potentials_)
    if (auto* pp = dynamic_cast<PairPotential*>(p)) {
      if (pp->NumTypesInTable() != GetNumberOfTypes()) pp->Initialize();
      NeighborList& nl = pp->GetNeighborList();
      nl.EnsureAllocated();
      // a full sort is due (first build, or positions set by the host): run
      // it now, outside any graph, so that recorded steps hold the tile
      // re-sort of skin-triggered rebuilds (the build settles its capacities)
      if (nl.NeedsFullSort()) nl.UpdateNBlist();
    }
  if (FusedStepPotential() != nullptr) pd_.EnsureNextPositions();
}

void Sample::StepOnce(bool zero_momentum) {
  if (zero_momentum) itg_->SetMomentumToZero();
  PairPotential* pp = FusedStepPotential();
  if (pp != nullptr && pd_.HasNextPositions()) {
    BumpSlotEpoch();
    static_cast<IntegratorNVT*>(itg_)->IntegrateFused(*pp);
    return;
  }
  CalcF(false);
  itg_->Integrate();
}

std::vector<std::pair<unsigned, bool>> Sample::RunSegments(unsigned long first_step, unsigned nsteps,
                                                           unsigned block, unsigned momentum_interval) const {
  if (block == 0) block = 1;
  std::vector<std::pair<unsigned, bool>> seg;
  const unsigned long end = first_step + nsteps;
  for (unsigned long i = first_step; i < end;) {
    unsigned long stop = std::min<unsigned long>(i + block, end);
    if (momentum_interval > 0) {
      // the next reset step after i ends the block before it
      const unsigned long next_reset = (i / momentum_interval + 1) * momentum_interval;
      stop = std::min(stop, next_reset);
    }
    seg.emplace_back(static_cast<unsigned>(stop - i), momentum_interval > 0 && i % momentum_interval == 0);
    i = stop;
  }
  return seg;
}

// Records nsteps steps into a graph (nothing runs); the host-side state that
// recording changes (the fused skin check's pending mark) is restored, so the
// graph can be launched later as if recorded then.
Sample::Block Sample::CaptureBlock(unsigned nsteps, bool reset_first) {
  NeighborList* nl = FusedCheckList();
  float* home = pd_.d_r.get();
  hipGraph_t g = nullptr;
  hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "Sample", "RunSteps");
  try {
    // the first step re-checks the skin explicitly; later steps use the
    // check fused into the previous integrator step
    if (nl) nl->ResetFusedCheck();
    for (unsigned k = 0; k < nsteps; ++k) StepOnce(k == 0 && reset_first);
    // fused steps alternate the two position arrays: the graph ends in the
    // one it starts from (a copy after an odd number of them)
    pd_.MovePositionsTo(home, stream_);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    if (pd_.d_r.get() != home) pd_.SwapPositions();  // nothing ran
    throw;
  }
  hip_check(hipStreamEndCapture(stream_, &g), "Sample", "RunSteps");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  hip_check(e, "Sample", "RunSteps");
  return Block{exec, home};
}

unsigned Sample::PrepareRunSteps(unsigned long first_step, unsigned nsteps, unsigned block,
                                 unsigned momentum_interval) {
  PrepareRun();
  NeighborList* nl = FusedCheckList();
  const bool pending = nl != nullptr && nl->FusedCheckPending();
  unsigned made = 0;
  for (const auto& key : RunSegments(first_step, nsteps, block, momentum_interval)) {
    if (key.first < 2 || graphs_.count(key)) continue;
    graphs_.emplace(key, CaptureBlock(key.first, key.second));
    ++made;
  }
  if (nl) {
    if (pending) nl->NoteFusedCheck(); else nl->ResetFusedCheck();
  }
  return made;
}

void Sample::RunSteps(unsigned long first_step, unsigned nsteps, unsigned block, unsigned momentum_interval) {
  PrepareRun();
  for (const auto& key : RunSegments(first_step, nsteps, block, momentum_interval)) {
    if (key.first < 2) {
      StepOnce(key.second);
      continue;
    }
    auto it = graphs_.find(key);
    if (it == graphs_.end()) it = graphs_.emplace(key, CaptureBlock(key.first, key.second)).first;
    // the graph's first step checks the skin itself (see CaptureBlock)
    if (NeighborList* nl = FusedCheckList()) nl->ResetFusedCheck();
    pd_.MovePositionsTo(it->second.r_home, stream_);
    hip_check(hipGraphLaunch(it->second.exec, stream_), "Sample", "RunSteps");
    if (NeighborList* nl = FusedCheckList()) nl->NoteFusedCheck();
  }
  Synchronize();
}

// Sample::ReadMoleculeData (src/Sample.cu, MoleculeData::ReadTopology).
void Sample::ReadMoleculeData(const std::string& topology_file) {
  CreateMoleculeData()->ReadTopology(topology_file);
  InvalidateGraphs();
}

MoleculeData* Sample::GetMoleculeData() { return molecules_.get(); }

MoleculeData* Sample::CreateMoleculeData() {
  if (!GetNumberOfParticles()) throw RUMD_Error("Sample", "ReadMoleculeData", "Sample has no particles");
  molecules_ = std::make_unique<MoleculeData>(this);
  return molecules_.get();
}
