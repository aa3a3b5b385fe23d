// ReplicaBatch: host side of the batched-replica kernels (rumdx_batch_*).
#include "replica_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace {

uint32_t padded(uint32_t n) { return (n + RUMDX_BATCH_ALIGN - 1) / RUMDX_BATCH_ALIGN * RUMDX_BATCH_ALIGN; }

}  // namespace

ReplicaBatch::ReplicaBatch(unsigned num_types) : ntypes_(num_types), mass_(num_types, 1.f) {
  if (num_types == 0 || num_types > RUMDX_MAX_TYPES)
    throw RUMD_Error("ReplicaBatch", "ReplicaBatch", "number of types out of range");
  Device::GetDevice().Init();
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "ReplicaBatch", "ReplicaBatch");
}

ReplicaBatch::~ReplicaBatch() {
  InvalidateGraphs();
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Layout-changing calls keep the device state: positions and velocities are
// read back into the replica specs before the arrays are rebuilt.
void ReplicaBatch::Relayout() {
  if (allocated_) {
    for (unsigned r = 0; r < reps_.size() && r < slot0_.size(); ++r) {
      reps_[r].pos = GetPositions(r);
      reps_[r].vel = GetVelocities(r);
    }
  }
  allocated_ = false;
  saved_ = false;
  InvalidateGraphs();
}

unsigned ReplicaBatch::AddReplica(const std::vector<float>& positions, const std::vector<unsigned>& types,
                                  const std::vector<float>& velocities, float box_length) {
  const size_t n = types.size();
  if (n < 2 || positions.size() != 3 * n || velocities.size() != 3 * n)
    throw RUMD_Error("ReplicaBatch", "AddReplica", "positions/velocities must be n x 3 with n = len(types) >= 2");
  if (!(box_length > 0.f)) throw RUMD_Error("ReplicaBatch", "AddReplica", "box length must be positive");
  for (unsigned t : types)
    if (t >= ntypes_) throw RUMD_Error("ReplicaBatch", "AddReplica", "type index out of range");
  Relayout();
  Spec s;
  s.pos = positions;
  s.vel = velocities;
  s.types = types;
  s.L = box_length;
  reps_.push_back(std::move(s));
  return static_cast<unsigned>(reps_.size() - 1);
}

void ReplicaBatch::SetMass(unsigned type, float mass) {
  if (type >= ntypes_ || !(mass > 0.f)) throw RUMD_Error("ReplicaBatch", "SetMass", "bad type or mass");
  Relayout();
  mass_[type] = mass;
}

void ReplicaBatch::SetPotential(const PairPotential& pot, float skin) {
  if (!(skin > 0.f)) throw RUMD_Error("ReplicaBatch", "SetPotential", "skin must be positive");
  if (pot.GetPotentialForm() != 0)
    throw RUMD_Error("ReplicaBatch", "SetPotential", "batched replicas run the 12-6 potentials only");
  Relayout();
  table_ = pot.GetParamsTable(ntypes_);
  cutoff_method_ = pot.GetCutoffMethod();
  skin_ = skin;
  max_cutoff_ = 0.f;
  for (unsigned k = 0; k < ntypes_ * ntypes_; ++k) max_cutoff_ = std::max(max_cutoff_, table_[k * RUMDX_NUM_PARAM]);
  if (!(max_cutoff_ > 0.f)) throw RUMD_Error("ReplicaBatch", "SetPotential", "potential has no parameters set");
}

void ReplicaBatch::SetTimeStep(float dt) {
  dt_ = dt;
  InvalidateGraphs();
}

void ReplicaBatch::SetThermostat(unsigned replica, float target_temperature, float tau) {
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "SetThermostat", "replica index out of range");
  reps_[replica].T = target_temperature;
  reps_[replica].tau = tau;
  if (allocated_) {
    // in place: the replica's entry and its thermostat target
    Synchronize();
    const rumdx_replica rp = HostReplica(replica);
    hip_check(hipMemcpyAsync(rep_.get() + replica, &rp, sizeof(rp), hipMemcpyHostToDevice, stream_), "ReplicaBatch",
              "SetThermostat");
    rumdx_nh_state st;
    hip_check(hipMemcpyAsync(&st, nh_.get() + replica, sizeof(st), hipMemcpyDeviceToHost, stream_), "ReplicaBatch",
              "SetThermostat");
    hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "SetThermostat");
    st.target_T = target_temperature;
    hip_check(hipMemcpyAsync(nh_.get() + replica, &st, sizeof(st), hipMemcpyHostToDevice, stream_), "ReplicaBatch",
              "SetThermostat");
    hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "SetThermostat");
  }
}

void ReplicaBatch::SetMaxNumNbrs(unsigned m) {
  Relayout();
  user_max_nbrs_ = m;
  max_nbrs_ = 0;
}

void ReplicaBatch::SetStagedForces(bool on) {
  staged_forces_ = on;
  if (!allocated_) return;
  BuildForceWork();
  InvalidateGraphs();
}

// Staged forces: work items of one replica each, the largest replicas first,
// when the largest fits one workgroup's LDS: whole replicas for fused steps,
// else slot ranges of at most 2048.
void ReplicaBatch::BuildForceWork() {
  force_work_.release();
  desc_.force_work_dev = nullptr;
  desc_.force_work_n = 0;
  desc_.force_work_whole = 0;
  desc_.max_rep_n = max_rep_n_;
  int dev = 0, optin = 0;
  hip_check(hipGetDevice(&dev), "ReplicaBatch", "Allocate");
  hip_check(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev), "ReplicaBatch", "Allocate");
  const size_t staged_lds = 16ull * max_rep_n_ + 32ull * ntypes_ * ntypes_;
  if (!staged_forces_ || staged_lds > static_cast<size_t>(optin)) return;
  const bool whole = fused_steps_;
  const uint32_t chunk = whole ? 0xffffffffu : 2048u;
  const unsigned nrep = static_cast<unsigned>(reps_.size());
  std::vector<uint32_t> order(nrep);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return reps_[x].types.size() > reps_[y].types.size(); });
  std::vector<uint32_t> hw;
  for (uint32_t r : order) {
    const uint32_t n = static_cast<uint32_t>(reps_[r].types.size());
    const uint32_t len = std::max(1u, std::min(chunk, n));
    for (uint32_t c = 0; c < n; c += len) hw.insert(hw.end(), {r, slot0_[r] + c, slot0_[r] + std::min(n, c + len), 0u});
  }
  force_work_.resize(hw.size());
  force_work_.upload(hw.data(), hw.size(), stream_);
  desc_.force_work_dev = force_work_.get();
  desc_.force_work_n = static_cast<uint32_t>(hw.size() / 4);
  desc_.force_work_whole = whole ? 1u : 0u;
}

void ReplicaBatch::SetFusedSteps(bool on) {
  fused_steps_ = on;
  if (!allocated_) return;
  BuildForceWork();
  InvalidateGraphs();
}

void ReplicaBatch::SetWorkgroupSort(bool on) {
  workgroup_sort_ = on;
  if (!allocated_) return;
  desc_.max_rep_cells = on ? max_rep_cells_ : 0u;
  InvalidateGraphs();
}

void ReplicaBatch::SetLanesPerParticle(unsigned tp) {
  if (tp != 1 && tp != 2 && tp != 4 && tp != 8)
    throw RUMD_Error("ReplicaBatch", "SetLanesPerParticle", "lanes per particle must be 1, 2, 4 or 8");
  lanes_per_particle_ = tp;
  desc_.lanes_per_particle = tp;
  InvalidateGraphs();
}

namespace {
template <class T>
void copy_device(DeviceArray<T>& dst, const DeviceArray<T>& src, hipStream_t st) {
  if (dst.size() != src.size()) dst.resize(src.size());
  if (src.size())
    hip_check(hipMemcpyAsync(dst.get(), src.get(), src.size() * sizeof(T), hipMemcpyDeviceToDevice, st),
              "ReplicaBatch", "SaveState");
}
}  // namespace

void ReplicaBatch::SaveState() {
  if (!allocated_) Allocate();
  Synchronize();
  copy_device(save_r_, r_, stream_);
  copy_device(save_v_, v_, stream_);
  copy_device(save_f_, f_, stream_);
  copy_device(save_w_, w_, stream_);
  copy_device(save_im_, im_, stream_);
  copy_device(save_idx_, idx_, stream_);
  copy_device(save_nh_, nh_, stream_);
  copy_device(save_flags_, flags_, stream_);
  hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "SaveState");
  saved_ = true;
}

void ReplicaBatch::RestoreState() {
  if (!saved_ || !allocated_ || save_r_.size() != r_.size())
    throw RUMD_Error("ReplicaBatch", "RestoreState", "no saved state for the current layout");
  Synchronize();
  copy_device(r_, save_r_, stream_);
  copy_device(v_, save_v_, stream_);
  copy_device(f_, save_f_, stream_);
  copy_device(w_, save_w_, stream_);
  copy_device(im_, save_im_, stream_);
  copy_device(idx_, save_idx_, stream_);
  copy_device(nh_, save_nh_, stream_);
  copy_device(flags_, save_flags_, stream_);
  rumdx_check(rumdx_batch_request_rebuild(&desc_, stream_), "ReplicaBatch", "RestoreState");
  hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "RestoreState");
}

unsigned ReplicaBatch::GetNumberOfParticles(unsigned replica) const {
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "GetNumberOfParticles", "replica index out of range");
  return static_cast<unsigned>(reps_[replica].types.size());
}

float ReplicaBatch::GetBoxLength(unsigned replica) const {
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "GetBoxLength", "replica index out of range");
  return reps_[replica].L;
}

std::vector<float> ReplicaBatch::GetThermostat(unsigned replica) const {
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "GetThermostat", "replica index out of range");
  return {reps_[replica].T, reps_[replica].tau};
}

unsigned ReplicaBatch::GetTotalParticles() const {
  unsigned n = 0;
  for (const Spec& s : reps_) n += static_cast<unsigned>(s.types.size());
  return n;
}

rumdx_replica ReplicaBatch::HostReplica(unsigned r) const {
  const Spec& s = reps_[r];
  rumdx_replica rp{};
  rp.slot0 = slot0_[r];
  rp.n = static_cast<uint32_t>(s.types.size());
  rp.cell0 = cell0_[r];
  rp.nc = nc_[r];
  rp.L = s.L;
  rp.inv_L = 1.f / s.L;
  rp.omega2 = s.tau > 0.f ? static_cast<float>(4.0 * M_PI * M_PI / (static_cast<double>(s.tau) * s.tau)) : 0.f;
  rp.dof = 3 * rp.n - 3;
  return rp;
}

// Lays the replicas out (slot ranges padded to whole workgroups, cell grids
// one after another), allocates every array and uploads the replica specs
// (the state Relayout read back, or the configurations AddReplica was given).
// Thermostat states restart at Ps = 0.
void ReplicaBatch::Allocate() {
  if (reps_.empty()) throw RUMD_Error("ReplicaBatch", "Allocate", "no replicas");
  if (table_.empty()) throw RUMD_Error("ReplicaBatch", "Allocate", "no potential set");
  const unsigned nrep = static_cast<unsigned>(reps_.size());
  slot0_.assign(nrep, 0u);
  cell0_.assign(nrep, 0u);
  nc_.assign(nrep, 0);
  uint64_t slots = 0, cells = 0;
  double max_density = 0.0;
  unsigned max_n = 0;
  for (unsigned r = 0; r < nrep; ++r) {
    const Spec& s = reps_[r];
    slot0_[r] = static_cast<uint32_t>(slots);
    slots += padded(static_cast<uint32_t>(s.types.size()));
    const float box8[8] = {1.f, s.L, s.L, s.L, 1.f / s.L, 1.f / s.L, 1.f / s.L, 0.f};
    int32_t nc[3];
    rumdx_check(rumdx_cell_grid(box8, max_cutoff_, skin_, nc), "ReplicaBatch", "Allocate");
    nc_[r] = nc[0];
    cell0_[r] = static_cast<uint32_t>(cells);
    const uint64_t ncell = static_cast<uint64_t>(nc[0]) * nc[0] * nc[0];
    cells += (ncell + RUMDX_BATCH_CELL_BLOCK - 1) / RUMDX_BATCH_CELL_BLOCK * RUMDX_BATCH_CELL_BLOCK;
    max_density = std::max(max_density, s.types.size() / (static_cast<double>(s.L) * s.L * s.L));
    max_n = std::max(max_n, static_cast<unsigned>(s.types.size()));
  }
  if (slots >= (1ull << 31) || cells >= (1ull << 31))
    throw RUMD_cudaResourceError("ReplicaBatch", "Allocate", "batch too large");
  nslots_ = static_cast<uint32_t>(slots);
  ncells_ = static_cast<uint32_t>(cells);
  if (user_max_nbrs_) {
    max_nbrs_ = user_max_nbrs_;
  } else if (max_nbrs_ == 0) {
    // NeighborList's estimate (src/NeighborList.cu:595-610) at the densest replica
    const double rs = max_cutoff_ + skin_;
    max_nbrs_ = static_cast<unsigned>(std::max(50.0, 1.5 * 4.0 / 3.0 * M_PI * rs * rs * rs * max_density));
  }
  max_nbrs_ = std::max(1u, std::min(max_nbrs_, max_n));
  if (static_cast<uint64_t>(max_nbrs_) * nslots_ >= (1ull << 32))
    throw RUMD_cudaResourceError("ReplicaBatch", "Allocate", "neighbour lists exceed 2^32 entries");

  const size_t ns = nslots_;
  std::vector<float> hr(4 * ns, 0.f), hv(4 * ns, 0.f);
  std::vector<uint32_t> hb(ns / RUMDX_BATCH_ALIGN);
  std::vector<rumdx_replica> hrep(nrep);
  std::vector<rumdx_nh_state> hnh(nrep);
  std::vector<uint32_t> hflags(static_cast<size_t>(RUMDX_BATCH_FLAGS) * nrep, 0u);
  std::vector<uint32_t> hidx(ns, 0u);
  for (unsigned r = 0; r < nrep; ++r) {
    const Spec& s = reps_[r];
    hrep[r] = HostReplica(r);
    hnh[r] = rumdx_nh_state{0.0, static_cast<double>(s.T), 0.0, 0.0};
    hflags[RUMDX_BATCH_FLAGS * r] = 1u;  // first build
    const uint32_t n = static_cast<uint32_t>(s.types.size());
    for (uint32_t b = slot0_[r] / RUMDX_BATCH_ALIGN; b < (slot0_[r] + padded(n)) / RUMDX_BATCH_ALIGN; ++b) hb[b] = r;
    for (uint32_t i = 0; i < n; ++i) {
      hidx[slot0_[r] + i] = i;
      const size_t k = 4ull * (slot0_[r] + i);
      hr[k] = s.pos[3ull * i];
      hr[k + 1] = s.pos[3ull * i + 1];
      hr[k + 2] = s.pos[3ull * i + 2];
      uint32_t t = s.types[i];
      std::memcpy(&hr[k + 3], &t, sizeof(t));
      hv[k] = s.vel[3ull * i];
      hv[k + 1] = s.vel[3ull * i + 1];
      hv[k + 2] = s.vel[3ull * i + 2];
      hv[k + 3] = 1.f / mass_[t];
    }
  }
  InvalidateGraphs();
  r_.resize(4 * ns);
  v_.resize(4 * ns);
  f_.resize(4 * ns);
  w_.resize(4 * ns);
  im_.resize(4 * ns);
  last_r_.resize(4 * ns);
  r_.upload(hr.data(), hr.size(), stream_);
  v_.upload(hv.data(), hv.size(), stream_);
  block_rep_.resize(hb.size());
  block_rep_.upload(hb.data(), hb.size(), stream_);
  rep_.resize(nrep);
  rep_.upload(hrep.data(), nrep, stream_);
  nh_.resize(nrep);
  nh_.upload(hnh.data(), nrep, stream_);
  flags_.resize(hflags.size());
  flags_.upload(hflags.data(), hflags.size(), stream_);
  nbl_.resize(static_cast<size_t>(max_nbrs_) * ns);
  num_nbrs_.resize(ns);
  std::vector<uint32_t> hcb(ncells_ / RUMDX_BATCH_CELL_BLOCK);
  for (unsigned r = 0; r < nrep; ++r) {
    const uint64_t ncell = static_cast<uint64_t>(nc_[r]) * nc_[r] * nc_[r];
    const uint32_t b0 = cell0_[r] / RUMDX_BATCH_CELL_BLOCK;
    const uint32_t nb = static_cast<uint32_t>((ncell + RUMDX_BATCH_CELL_BLOCK - 1) / RUMDX_BATCH_CELL_BLOCK);
    for (uint32_t k = 0; k < nb; ++k) hcb[b0 + k] = r;
  }
  cellblock_rep_.resize(hcb.size());
  cellblock_rep_.upload(hcb.data(), hcb.size(), stream_);
  cell_count_.resize(ncells_);
  cell_start_.resize(ncells_);
  cell_n_.resize(ncells_);
  cell_tmp_.resize(ns);
  slot_cell_.resize(ns);
  idx_.resize(ns);
  idx_.upload(hidx.data(), hidx.size(), stream_);
  t_r_.resize(4 * ns);
  t_v_.resize(4 * ns);
  t_im_.resize(4 * ns);
  t_idx_.resize(ns);
  fixed_.resize(4 * ns);  // written by every build (all replicas rebuild first)
  params_.resize(table_.size());
  params_.upload(table_.data(), table_.size(), stream_);
  std::vector<float> c2(static_cast<size_t>(ntypes_) * ntypes_);
  for (size_t k = 0; k < c2.size(); ++k) {
    const float c = table_[k * RUMDX_NUM_PARAM] + skin_;  // squared in float (src/NeighborList.cu:1239-1240)
    c2[k] = c * c;
  }
  cut2_.resize(c2.size());
  cut2_.upload(c2.data(), c2.size(), stream_);
  workspace_.resize((rumdx_batch_workspace_bytes(nslots_) + sizeof(double) - 1) / sizeof(double));
  energies_.resize(4ull * nrep);

  std::memset(&desc_, 0, sizeof(desc_));
  desc_.r_dev = r_.get();
  desc_.v_dev = v_.get();
  desc_.f_dev = f_.get();
  desc_.w_dev = w_.get();
  desc_.im_dev = im_.get();
  desc_.nslots = nslots_;
  desc_.nrep = nrep;
  desc_.rep_dev = rep_.get();
  desc_.block_rep_dev = block_rep_.get();
  desc_.nh_dev = nh_.get();
  desc_.rep_flags_dev = flags_.get();
  desc_.nbl_dev = nbl_.get();
  desc_.num_nbrs_dev = num_nbrs_.get();
  desc_.last_r_dev = last_r_.get();
  desc_.cut_skin2_dev = cut2_.get();
  desc_.max_nbrs = max_nbrs_;
  desc_.ntypes = ntypes_;
  desc_.skin = skin_;
  desc_.ncells = ncells_;
  desc_.cellblock_rep_dev = cellblock_rep_.get();
  desc_.cell_count_dev = cell_count_.get();
  desc_.cell_start_dev = cell_start_.get();
  desc_.cell_n_dev = cell_n_.get();
  desc_.idx_dev = idx_.get();
  desc_.lanes_per_particle = lanes_per_particle_;
  desc_.fixed_dev = reinterpret_cast<int32_t*>(fixed_.get());
  desc_.t_r_dev = t_r_.get();
  desc_.t_v_dev = t_v_.get();
  desc_.t_im_dev = t_im_.get();
  desc_.t_idx_dev = t_idx_.get();
  desc_.cell_tmp_dev = cell_tmp_.get();
  desc_.slot_cell_dev = slot_cell_.get();
  max_rep_n_ = max_n;
  BuildForceWork();
  // the one-workgroup rebuild sort when a replica's counts, starts and slots
  // fit 64 KiB of LDS
  uint64_t max_cells = 0;
  for (unsigned r = 0; r < nrep; ++r)
    max_cells = std::max<uint64_t>(max_cells, static_cast<uint64_t>(nc_[r]) * nc_[r] * nc_[r]);
  max_rep_cells_ = (4ull * (max_n + 2 * max_cells) <= 65536ull) ? static_cast<uint32_t>(max_cells) : 0u;
  desc_.max_rep_cells = workgroup_sort_ ? max_rep_cells_ : 0u;
  allocated_ = true;
}

void ReplicaBatch::CalcF() {
  if (!allocated_) Allocate();
  rumdx_check(rumdx_batch_update_lists(&desc_, stream_), "ReplicaBatch", "CalcF");
  rumdx_check(rumdx_batch_forces(&desc_, params_.get(), cutoff_method_, stream_), "ReplicaBatch", "CalcF");
}

void ReplicaBatch::Step() {
  if (!allocated_) Allocate();
  if (UsesFusedSteps()) {
    rumdx_check(rumdx_batch_update_lists(&desc_, stream_), "ReplicaBatch", "RunSteps");
    rumdx_check(rumdx_batch_forces_nvt(&desc_, params_.get(), cutoff_method_, dt_, workspace_.get(), stream_),
                "ReplicaBatch", "RunSteps");
    return;
  }
  CalcF();
  rumdx_check(rumdx_batch_nvt_step(&desc_, dt_, workspace_.get(), stream_), "ReplicaBatch", "RunSteps");
}

float ReplicaBatch::TimeForceKernel(unsigned reps) {
  if (!allocated_) Allocate();
  if (reps == 0) reps = 1;
  rumdx_check(rumdx_batch_update_lists(&desc_, stream_), "ReplicaBatch", "TimeForceKernel");
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "ReplicaBatch", "TimeForceKernel");
  hip_check(hipEventCreate(&b), "ReplicaBatch", "TimeForceKernel");
  hip_check(hipEventRecord(a, stream_), "ReplicaBatch", "TimeForceKernel");
  for (unsigned k = 0; k < reps; ++k)
    rumdx_check(rumdx_batch_forces(&desc_, params_.get(), cutoff_method_, stream_), "ReplicaBatch", "TimeForceKernel");
  hip_check(hipEventRecord(b, stream_), "ReplicaBatch", "TimeForceKernel");
  hip_check(hipEventSynchronize(b), "ReplicaBatch", "TimeForceKernel");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "ReplicaBatch", "TimeForceKernel");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

std::vector<double> ReplicaBatch::GetNeighborListStats() {
  if (!allocated_) return {0.0, 0.0};
  std::vector<uint32_t> cnt(nslots_);
  num_nbrs_.download(cnt.data(), nslots_, stream_);
  double sum = 0.0, mx = 0.0, n = 0.0;
  for (size_t r = 0; r < reps_.size(); ++r) {
    const size_t nr = reps_[r].types.size();
    for (size_t k = 0; k < nr; ++k) {
      sum += cnt[slot0_[r] + k];
      mx = std::max(mx, static_cast<double>(cnt[slot0_[r] + k]));
    }
    n += static_cast<double>(nr);
  }
  return {n > 0.0 ? sum / n : 0.0, mx};
}

void ReplicaBatch::InvalidateGraphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

void ReplicaBatch::RunSteps(unsigned nsteps, unsigned block) {
  if (!allocated_) Allocate();
  if (block == 0) block = 1;
  unsigned done = 0;
  while (done < nsteps) {
    if (block == 1 || done + block > nsteps) {
      Step();
      ++done;
      continue;
    }
    auto it = graphs_.find(block);
    if (it == graphs_.end()) {
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "ReplicaBatch", "RunSteps");
      try {
        for (unsigned k = 0; k < block; ++k) Step();
      } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(stream_, &g), "ReplicaBatch", "RunSteps");
      hipGraphExec_t exec = nullptr;
      hip_check(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "ReplicaBatch", "RunSteps");
      (void)hipGraphDestroy(g);
      it = graphs_.emplace(block, exec).first;
    }
    hip_check(hipGraphLaunch(it->second, stream_), "ReplicaBatch", "RunSteps");
    done += block;
  }
  Synchronize();
}

// Grows the lists when some replica's overflowed (the forces were exact: the
// force kernel rescans the stencil for an overflowed list) and rebuilds all.
void ReplicaBatch::Synchronize() {
  hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "Synchronize");
  if (!allocated_) return;
  std::vector<uint32_t> fl(static_cast<size_t>(RUMDX_BATCH_FLAGS) * reps_.size());
  flags_.download(fl.data(), fl.size(), stream_);
  uint32_t longest = 0;
  bool overflow = false;
  for (size_t r = 0; r < reps_.size(); ++r) {
    overflow = overflow || fl[RUMDX_BATCH_FLAGS * r + 1] != 0u;
    longest = std::max(longest, fl[RUMDX_BATCH_FLAGS * r + 3]);
  }
  if (!overflow) return;
  unsigned max_n = 0;
  for (const Spec& s : reps_) max_n = std::max(max_n, static_cast<unsigned>(s.types.size()));
  max_nbrs_ = std::min(std::max(max_nbrs_ + max_nbrs_ / 2, longest + longest / 4 + 8), max_n);
  if (user_max_nbrs_) user_max_nbrs_ = max_nbrs_;
  InvalidateGraphs();
  nbl_.resize(static_cast<size_t>(max_nbrs_) * nslots_);
  desc_.nbl_dev = nbl_.get();
  desc_.max_nbrs = max_nbrs_;
  for (size_t r = 0; r < reps_.size(); ++r) {
    fl[RUMDX_BATCH_FLAGS * r] = 1u;  // rebuild
    fl[RUMDX_BATCH_FLAGS * r + 1] = 0u;
  }
  flags_.upload(fl.data(), fl.size(), stream_);
}

std::vector<uint32_t> ReplicaBatch::SlotIds(unsigned replica) {
  const unsigned n = GetNumberOfParticles(replica);
  std::vector<uint32_t> id(n);
  hip_check(hipMemcpyAsync(id.data(), idx_.get() + slot0_[replica], n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           stream_),
            "ReplicaBatch", "Get");
  hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "Get");
  for (uint32_t k : id)
    if (k >= n) throw RUMD_Error("ReplicaBatch", "Get", "slot ids are not a permutation");
  return id;
}

std::vector<float> ReplicaBatch::Slots(const DeviceArray<float>& a, unsigned replica, unsigned comps) {
  if (!allocated_) Allocate();
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "Get", "replica index out of range");
  Synchronize();
  const unsigned n = GetNumberOfParticles(replica);
  std::vector<float> all(4ull * n);
  hip_check(hipMemcpyAsync(all.data(), a.get() + 4ull * slot0_[replica], all.size() * sizeof(float),
                           hipMemcpyDeviceToHost, stream_),
            "ReplicaBatch", "Get");
  hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "Get");
  const std::vector<uint32_t> id = SlotIds(replica);
  std::vector<float> out(static_cast<size_t>(comps) * n);
  for (unsigned s = 0; s < n; ++s)
    for (unsigned c = 0; c < comps; ++c) out[static_cast<size_t>(comps) * id[s] + c] = all[4ull * s + c];
  return out;
}

std::vector<float> ReplicaBatch::GetPositions(unsigned replica) { return Slots(r_, replica, 3); }
std::vector<float> ReplicaBatch::GetVelocities(unsigned replica) { return Slots(v_, replica, 3); }
std::vector<float> ReplicaBatch::GetForces(unsigned replica) { return Slots(f_, replica, 4); }

std::vector<double> ReplicaBatch::GetEnergies() {
  if (!allocated_) Allocate();
  Synchronize();
  rumdx_check(rumdx_batch_energies(&desc_, dt_, energies_.get(), workspace_.get(), stream_), "ReplicaBatch",
              "GetEnergies");
  std::vector<double> e(4 * reps_.size());
  energies_.download(e.data(), e.size(), stream_);
  for (size_t r = 0; r < reps_.size(); ++r) e[4 * r + 1] /= 6.0;  // virial = sum w_i / 6 (src/Sample.cu:370-391)
  return e;
}

std::vector<double> ReplicaBatch::GetThermostatStates() {
  if (!allocated_) Allocate();
  Synchronize();
  std::vector<rumdx_nh_state> s(reps_.size());
  nh_.download(s.data(), s.size(), stream_);
  std::vector<double> ps(s.size());
  for (size_t r = 0; r < s.size(); ++r) ps[r] = s[r].ps;
  return ps;
}

std::vector<unsigned> ReplicaBatch::GetNumberOfBuilds() {
  if (!allocated_) Allocate();
  Synchronize();
  std::vector<uint32_t> fl(static_cast<size_t>(RUMDX_BATCH_FLAGS) * reps_.size());
  flags_.download(fl.data(), fl.size(), stream_);
  std::vector<unsigned> b(reps_.size());
  for (size_t r = 0; r < reps_.size(); ++r) b[r] = fl[RUMDX_BATCH_FLAGS * r + 2];
  return b;
}

void ReplicaBatch::GetNeighborList(unsigned replica, std::vector<uint32_t>& offsets, std::vector<uint32_t>& ids) {
  if (!allocated_) Allocate();
  if (replica >= reps_.size()) throw RUMD_Error("ReplicaBatch", "GetNeighborList", "replica index out of range");
  Synchronize();
  const unsigned n = GetNumberOfParticles(replica);
  const uint32_t s0 = slot0_[replica];
  std::vector<uint32_t> cnt(n);
  hip_check(hipMemcpyAsync(cnt.data(), num_nbrs_.get() + s0, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
            "ReplicaBatch", "GetNeighborList");
  const std::vector<uint32_t> id = SlotIds(replica);  // synchronises
  std::vector<uint32_t> col(n);
  // rows in particle order: row id[s] holds slot s's list, entries as particle ids
  offsets.assign(n + 1, 0u);
  for (unsigned s = 0; s < n; ++s) offsets[id[s] + 1] = cnt[s];
  for (unsigned i = 0; i < n; ++i) offsets[i + 1] += offsets[i];
  ids.assign(offsets[n], 0u);
  unsigned kmax = 0;
  for (uint32_t c : cnt) kmax = std::max(kmax, c);
  for (unsigned k = 0; k < kmax; ++k) {
    hip_check(hipMemcpyAsync(col.data(), nbl_.get() + static_cast<size_t>(k) * nslots_ + s0, n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, stream_),
              "ReplicaBatch", "GetNeighborList");
    hip_check(hipStreamSynchronize(stream_), "ReplicaBatch", "GetNeighborList");
    for (unsigned s = 0; s < n; ++s)
      if (k < cnt[s]) ids[offsets[id[s]] + k] = id[col[s] - s0];
  }
}
