// PairPotential / Pot_IPL_12_6 / Pot_LJ_12_6 and NeighborList host logic
// (include/rumd/PairPotential.h, src/PairPotential.cu, src/NeighborList.cu).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "rumd_host.h"
#include "../csrc/pot_forms.h"

namespace {
constexpr unsigned kNumParam = RUMDX_NUM_PARAM;
constexpr unsigned kMinSizeMemoryOpt = 10000;   // src/NeighborList.cu:27
constexpr unsigned kDefaultMaxNumNbrs = 50;     // src/NeighborList.cu:28
constexpr unsigned kSortThreshold = 8000;       // src/NeighborList.cu:31
constexpr uint32_t kHaloShrinkSlack = 128;     // halo capacity kept above the need before shrinking
}  // namespace

// ================================================================== NeighborList
NeighborList::NeighborList() = default;

void NeighborList::Initialize(Sample* s) {
  sample_ = s;
  allocated_ = false;
  fused_check_pending_ = false;
  sort_valid_ = false;
}

// The next sorted build must sort from scratch (positions were set, the
// arrays reallocated, or the build asked for by the host); graphs recorded
// with the tile re-sort are stale then.
void NeighborList::InvalidateSortOrder() {
  if (sort_valid_ && sample_) sample_->InvalidateGraphs();
  sort_valid_ = false;
}

void NeighborList::SetCutoffs(const PairPotential* pot) {
  if (!sample_) return;
  ntypes_ = sample_->GetNumberOfTypes();
  cutoffs_.assign(ntypes_ * ntypes_, 0.f);
  max_cutoff_ = 0.f;
  for (unsigned i = 0; i < ntypes_; ++i)
    for (unsigned j = 0; j < ntypes_; ++j) {
      const float rc = pot->GetPotentialParameter(i, j, 0);
      cutoffs_[i * ntypes_ + j] = rc;
      max_cutoff_ = std::max(max_cutoff_, rc);
    }
  allocated_ = false;  // cut-offs (and possibly the cell grid) changed
  halo_cap_ = 0;
}

void NeighborList::SetSkin(float skin) {
  if (!(skin >= 0.f)) throw RUMD_Error("NeighborList", "SetSkin", "skin must be non-negative");
  skin_ = skin;
  allocated_ = false;
  halo_cap_ = 0;
  if (sample_) sample_->InvalidateGraphs();
}

// Inner-list skin of the sorted method (RUMDX inner lists): zero (the
// default) switches the inner lists off, a negative value selects skin/3.
void NeighborList::SetInnerSkin(float inner) {
  if (!(inner == inner)) throw RUMD_Error("NeighborList", "SetInnerSkin", "inner skin is NaN");
  inner_skin_ = inner;
  allocated_ = false;
  if (sample_) sample_->InvalidateGraphs();
}

// Waves per SIMD of the 12-6 tile force kernel (0 = its default, or 4, 5, 6).
void NeighborList::SetTileWavesPerEU(unsigned w) {
  if (w != 0 && (w < 4 || w > 6))
    throw RUMD_Error("NeighborList", "SetTileWavesPerEU", "waves per SIMD must be 0 (default), 4, 5 or 6");
  tile_waves_ = w;
  desc_.tile_waves_per_eu = w;
  if (sample_) sample_->InvalidateGraphs();
}

float NeighborList::GetInnerSkin() const {
  if (inner_skin_ < 0.f) return skin_ / 3.f;
  return inner_skin_;
}

void NeighborList::SetNB_Method(const std::string& m) {
  if (m == "default") method_ = DEFAULT;
  else if (m == "none") method_ = NONE;
  else if (m == "n2") method_ = N2;
  else if (m == "sort") method_ = SORT;
  else throw RUMD_Error("NeighborList", "SetNB_Method", "unknown method '" + m + "' (default, none, n2, sort)");
  allocated_ = false;
  if (sample_) sample_->InvalidateGraphs();
}

std::string NeighborList::GetNB_Method() const {
  switch (method_) {
    case NONE: return "none";
    case N2: return "n2";
    case SORT: return "sort";
    default: return "default";
  }
}

NeighborList::Method NeighborList::GetResolvedMethod() const {
  // with exclusions "none" runs as N2 (NeighborList::SetExclusionType and
  // UpdateNBlist do the same switch in the reference)
  if (method_ == NONE && HasExclusions()) return N2;
  if (method_ != DEFAULT) return method_;
  const unsigned n = sample_ ? sample_->GetNumberOfParticles() : 0;
  return n > kSortThreshold ? SORT : N2;  // src/NeighborList.cu:586-587
}

void NeighborList::SetMaxNumNbrs(unsigned m) {
  user_max_nbrs_ = m;
  allocated_ = false;
  if (sample_) sample_->InvalidateGraphs();
}

void NeighborList::SetExclusion(unsigned i, unsigned j) { AddExclusions({i, j}); }

void NeighborList::AddExclusions(const std::vector<uint32_t>& pairs) {
  if (!sample_) throw RUMD_Error("NeighborList", "SetExclusion", "no sample");
  const unsigned n = sample_->GetNumberOfParticles();
  if (pairs.size() % 2) throw RUMD_Error("NeighborList", "SetExclusion", "exclusions come in pairs");
  if (excl_.size() != n) excl_.assign(n, {});
  for (size_t k = 0; k < pairs.size(); k += 2) {
    const uint32_t a = pairs[k], b = pairs[k + 1];
    if (a >= n || b >= n || a == b) throw RUMD_Error("NeighborList", "SetExclusion", "invalid particle pair");
    auto add = [&](uint32_t x, uint32_t y) {
      if (std::find(excl_[x].begin(), excl_[x].end(), y) == excl_[x].end()) {
        excl_[x].push_back(y);
        return true;
      }
      return false;
    };
    if (add(a, b)) ++num_exclusions_;
    add(b, a);
  }
  UploadExclusions();
}

void NeighborList::ClearExclusions() {
  excl_.clear();
  num_exclusions_ = 0;
  UploadExclusions();
}

// CSR over original ids, uploaded now (not under a graph capture); the next
// update rebuilds the list with them.
void NeighborList::UploadExclusions() {
  if (num_exclusions_ == 0) {
    desc_.excl_start_dev = nullptr;
    desc_.excl_ids_dev = nullptr;
  } else {
    const unsigned n = sample_->GetNumberOfParticles();
    std::vector<uint32_t> start(n + 1, 0u), ids;
    for (unsigned o = 0; o < n; ++o) {
      start[o] = static_cast<uint32_t>(ids.size());
      ids.insert(ids.end(), excl_[o].begin(), excl_[o].end());
    }
    start[n] = static_cast<uint32_t>(ids.size());
    const hipStream_t st = sample_->GetStream();
    excl_start_.resize(start.size());
    excl_start_.upload(start.data(), start.size(), st);
    excl_ids_.resize(ids.size());
    excl_ids_.upload(ids.data(), ids.size(), st);
    desc_.excl_start_dev = excl_start_.get();
    desc_.excl_ids_dev = excl_ids_.get();
  }
  allocated_ = false;  // the method may resolve differently; rebuild from scratch
  if (sample_) sample_->InvalidateGraphs();
}

void NeighborList::ResetNeighborList() {
  if (allocated_) rumdx_check(rumdx_nblist_request_rebuild(&desc_, sample_->GetStream()), "NeighborList", "ResetNeighborList");
  fused_check_pending_ = false;
  InvalidateSortOrder();
}

void NeighborList::Allocate() {
  const unsigned n = sample_->GetNumberOfParticles();
  const unsigned nvp = sample_->GetNumberOfVirtualParticles();
  if (ntypes_ == 0 || cutoffs_.empty()) throw RUMD_Error("NeighborList", "Allocate", "cut-offs not set");
  if (!sample_->GetSimulationBox()) throw RUMD_Error("NeighborList", "Allocate", "no simulation box");
  const Method method = GetResolvedMethod();
  // list capacity: NeighborList::UpdateNBlist (src/NeighborList.cu:595-610)
  if (user_max_nbrs_ > 0) {
    max_nbrs_ = std::min(user_max_nbrs_, n);
  } else if (grown_max_nbrs_ > 0) {
    max_nbrs_ = std::min(grown_max_nbrs_, n);
  } else {
    if (n < kMinSizeMemoryOpt || allocate_max_) {
      max_nbrs_ = n;
    } else {
      const double rs = static_cast<double>(max_cutoff_) + skin_;
      unsigned est = static_cast<unsigned>(1.5 * (4. / 3.) * M_PI * rs * rs * rs * n /
                                           sample_->GetSimulationBox()->GetVolume());
      max_nbrs_ = std::max(est, kDefaultMaxNumNbrs);
      max_nbrs_ = std::min(max_nbrs_, n);
    }
  }
  if (max_nbrs_ == 0) max_nbrs_ = 1;
  if (static_cast<uint64_t>(max_nbrs_) * nvp >= (1ull << 32))
    throw RUMD_cudaResourceError("NeighborList", "Allocate", "neighbour list exceeds 2^32 entries");
  const hipStream_t st = sample_->GetStream();
  if (method == SORT) {
    // entries are stored in groups of 8 (16 bytes) per slot
    nbl16_.resize(static_cast<size_t>((max_nbrs_ + 7u) & ~7u) * nvp);
    nbl_.release();
  } else {
    nbl_.resize(static_cast<size_t>(max_nbrs_) * nvp);
    nbl16_.release();
  }
  num_nbrs_.resize(nvp);
  last_r_.resize(4ull * nvp);
  flags_.resize(RUMDX_NB_FLAGS);
  std::vector<float> cs2(ntypes_ * ntypes_);
  for (size_t k = 0; k < cs2.size(); ++k) {
    const float c = cutoffs_[k] + skin_;  // squared in float (src/NeighborList.cu:1239-1240)
    cs2[k] = c * c;
  }
  cut_skin2_.resize(cs2.size());
  cut_skin2_.upload(cs2.data(), cs2.size(), st);
  std::memset(&desc_, 0, sizeof(desc_));
  if (num_exclusions_ > 0) {
    desc_.excl_start_dev = excl_start_.get();
    desc_.excl_ids_dev = excl_ids_.get();
  }
  desc_.nbl_dev = nbl_.get();
  desc_.num_nbrs_dev = num_nbrs_.get();
  desc_.last_r_dev = last_r_.get();
  desc_.flags_dev = flags_.get();
  desc_.cut_skin2_dev = cut_skin2_.get();
  desc_.max_nbrs = max_nbrs_;
  desc_.ntypes = ntypes_;
  desc_.skin = skin_;
  // the sample's threads per particle (Sample::SetPB_TP), as a power of two <= 8
  const unsigned tp = sample_->GetThreadsPerParticle();
  desc_.lanes_per_particle = tp >= 8 ? 8u : tp >= 4 ? 4u : tp >= 2 ? 2u : 1u;
  desc_.tile_waves_per_eu = tile_waves_;
  if (method == SORT) {
    int pair_pots = 0;
    for (Potential* p : sample_->GetPotentials())
      if (dynamic_cast<PairPotential*>(p)) ++pair_pots;
    if (pair_pots > 1)
      throw RUMD_Error("NeighborList", "Allocate",
                       "the sort method reorders particles and serves one pair potential per sample; "
                       "use SetNB_Method('n2') for the others");
    int32_t nc[3];
    rumdx_check(rumdx_cell_grid(sample_->GetSimulationBox()->GetBox8(), max_cutoff_, skin_, nc), "NeighborList",
                "Allocate");
    const size_t ncells = static_cast<size_t>(nc[0]) * nc[1] * nc[2];
    std::vector<uint32_t> rank(ncells);
    rumdx_check(rumdx_cell_rank_table(nc, rank.data()), "NeighborList", "Allocate");
    cell_rank_.resize(ncells);
    cell_rank_.upload(rank.data(), ncells, st);
    cell_start_.resize(ncells + 1);
    cell_count_.resize(ncells);
    slot_cell_.resize(nvp);
    lane_order_.resize(nvp);
    sort_scratch_.resize(rumdx_sort_scratch_bytes(nvp, static_cast<uint32_t>(ncells)));
    std::memcpy(desc_.nc, nc, sizeof(nc));
    desc_.cell_rank_dev = cell_rank_.get();
    desc_.cell_start_dev = cell_start_.get();
    desc_.cell_count_dev = cell_count_.get();
    desc_.slot_cell_dev = slot_cell_.get();
    desc_.sort_scratch_dev = sort_scratch_.get();
    desc_.lane_order_dev = lane_order_.get();
    // cell tiles of the LDS-staged build and force kernels
    const int ntiles = rumdx_tile_table(nc, rank.data(), nullptr, nullptr);
    if (ntiles <= 0) rumdx_check(ntiles == 0 ? RUMDX_E_INVALID : ntiles, "NeighborList", "Allocate");
    std::vector<uint32_t> first(ntiles + 1), origin(ntiles);
    const int nt2 = rumdx_tile_table(nc, rank.data(), first.data(), origin.data());
    if (nt2 != ntiles) rumdx_check(RUMDX_E_INVALID, "NeighborList", "Allocate");
    tile_first_rank_.resize(first.size());
    tile_first_rank_.upload(first.data(), first.size(), st);
    tile_origin_.resize(origin.size());
    tile_origin_.upload(origin.data(), origin.size(), st);
    if (halo_cap_ == 0) halo_cap_ = rumdx_halo_cap_estimate(nc, n);
    tile_info_.resize(4ull * ntiles);
    desc_.nbl16_dev = nbl16_.get();
    desc_.ntiles = static_cast<uint32_t>(ntiles);
    desc_.tile_first_rank_dev = tile_first_rank_.get();
    desc_.tile_origin_dev = tile_origin_.get();
    desc_.tile_info_dev = tile_info_.get();
    SizeTileHalo();
    // inner lists: the force kernels read the entries within rc + inner skin
    const float inner = GetInnerSkin();
    if (inner > 0.f && inner < skin_) {
      nbl16_in_.resize(nbl16_.size());
      num_in_.resize(nvp);
      prune_r_.resize(4ull * nvp);
      std::vector<float> ci2(ntypes_ * ntypes_);
      for (size_t k = 0; k < ci2.size(); ++k) {
        const float c = cutoffs_[k] + inner;
        ci2[k] = c * c;
      }
      cut_in2_.resize(ci2.size());
      cut_in2_.upload(ci2.data(), ci2.size(), st);
      desc_.nbl16_in_dev = nbl16_in_.get();
      desc_.num_in_dev = num_in_.get();
      desc_.prune_r_dev = prune_r_.get();
      desc_.cut_in2_dev = cut_in2_.get();
      desc_.inner_skin = inner;
    } else {
      nbl16_in_.release();
      num_in_.release();
      prune_r_.release();
      cut_in2_.release();
    }
  } else {
    nbl16_in_.release();
    num_in_.release();
    prune_r_.release();
    cut_in2_.release();
    cell_rank_.release();
    cell_start_.release();
    cell_count_.release();
    slot_cell_.release();
    lane_order_.release();
    sort_scratch_.release();
    tile_first_rank_.release();
    tile_origin_.release();
    tile_halo_.release();
    tile_info_.release();
  }
  allocated_nvp_ = nvp;
  allocated_ = true;
  fused_check_pending_ = false;
  InvalidateSortOrder();
  rumdx_check(rumdx_nblist_request_rebuild(&desc_, st), "NeighborList", "Allocate");
  sample_->InvalidateGraphs();
}

void NeighborList::EnsureAllocated() {
  if (!sample_ || GetResolvedMethod() == NONE) return;
  if (!allocated_ || allocated_nvp_ != sample_->GetNumberOfVirtualParticles()) Allocate();
}

void NeighborList::UpdateNBlist() {
  if (!sample_) throw RUMD_Error("NeighborList", "UpdateNBlist", "no sample");
  const Method method = GetResolvedMethod();
  if (method == NONE) return;
  EnsureAllocated();
  rumdx_particles* p = sample_->GetParticleData().Descriptor();
  const float* box = sample_->GetSimulationBox()->GetBox8();
  const hipStream_t st = sample_->GetStream();
  if (!fused_check_pending_) rumdx_check(rumdx_nblist_check(p, &desc_, box, st), "NeighborList", "UpdateNBlist");
  fused_check_pending_ = false;
  if (method == SORT) {
    // skin-triggered rebuilds re-sort tile by tile; the first build, and any
    // after the host changed positions or arrays, sorts from scratch
    const bool resort = UseTileResort() && !force_full_sort_;
    const int rc = resort ? rumdx_nblist_rebuild_sorted(p, &desc_, box, st) : rumdx_nblist_build_sorted(p, &desc_, box, st);
    rumdx_check(rc, "NeighborList", "UpdateNBlist");
    sort_valid_ = true;
    sample_->BumpSlotEpoch();  // the build may have moved particles between slots
    if (!resort && !force_full_sort_) Settle();
    Prune();
  }
  else
    rumdx_check(rumdx_nblist_build_n2(p, &desc_, box, st), "NeighborList", "UpdateNBlist");
}

// A full sort starts a new particle order.  Its halo and tile occupancy decide
// the halo capacity and whether later rebuilds re-sort tile by tile, so they
// are read back at once (outside a graph capture; a captured full sort is
// settled at the next synchronisation) and any rebuild they ask for runs now.
// Eager steps and graph blocks therefore see the same list history, whenever
// the host next synchronises.
// The inner lists' prune (rumdx_nblist_prune): on the device it runs only
// after a build or when a particle moved 0.49 inner skin since the last one.
void NeighborList::Prune() {
  if (desc_.nbl16_in_dev == nullptr) return;
  rumdx_check(rumdx_nblist_prune(sample_->GetParticleData().Descriptor(), &desc_,
                                 sample_->GetSimulationBox()->GetBox8(), sample_->GetStream()),
              "NeighborList", "UpdateNBlist");
}

void NeighborList::Settle() {
  const hipStream_t st = sample_->GetStream();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(st, &cs), "NeighborList", "UpdateNBlist");
  if (cs != hipStreamCaptureStatusNone) return;
  for (int round = 0; round < 4 && HandleOverflow(); ++round) {
    // capacities changed and a rebuild is pending
    rumdx_particles* p = sample_->GetParticleData().Descriptor();
    const float* box = sample_->GetSimulationBox()->GetBox8();
    const int rc = UseTileResort() ? rumdx_nblist_rebuild_sorted(p, &desc_, box, st)
                                   : rumdx_nblist_build_sorted(p, &desc_, box, st);
    rumdx_check(rc, "NeighborList", "UpdateNBlist");
    sort_valid_ = true;
    sample_->BumpSlotEpoch();
  }
}

void NeighborList::ReadFlags(uint32_t* f) {
  if (desc_.nbl16_dev != nullptr)
    rumdx_check(rumdx_nblist_build_stats(sample_->GetParticleData().Descriptor(), &desc_, sample_->GetStream()),
                "NeighborList", "UpdateNBlist");
  flags_.download(f, RUMDX_NB_FLAGS, sample_->GetStream());
}

bool NeighborList::HandleOverflow() {
  if (!allocated_) return false;
  uint32_t f[RUMDX_NB_FLAGS];
  ReadFlags(f);
  if (f[RUMDX_FLAG_ERROR]) {
    std::string why;
    if (f[RUMDX_FLAG_ERROR] & RUMDX_BUILD_ERR_HALO) why += " a cell tile's halo holds more than 65535 particles;";
    if (f[RUMDX_FLAG_ERROR] & RUMDX_BUILD_ERR_MOVED)
      why += " a particle moved more than one cell between list builds (the simulation is unstable);";
    if (f[RUMDX_FLAG_ERROR] & RUMDX_BUILD_ERR_WAIT) why += " a tile of the re-sort waited too long for the tiles before it;";
    if (f[RUMDX_FLAG_ERROR] & RUMDX_BUILD_ERR_TILE)
      why += " a cell tile holds more than " + std::to_string(RUMDX_RESORT_TILE_MAX) + " particles;";
    throw RUMD_Error("NeighborList", "UpdateNBlist", "the sorted neighbour list is unusable:" + why);
  }
  if (desc_.nbl16_dev != nullptr) {
    // a crowded tile (more than half the re-sort's capacity) sorts from scratch
    const bool was = UseTileResort();
    last_max_tile_ = f[RUMDX_FLAG_MAXTILE];
    if (was != UseTileResort()) InvalidateSortOrder();
  }
  if (desc_.nbl16_dev != nullptr && f[RUMDX_FLAG_MAXHALO] > 0u) {
    // Capacity for the largest halo of the last build plus 1/16.  Tiles beyond
    // the capacity computed exactly from global memory: grow for them.  The
    // force kernel's LDS is 16 bytes per unit of capacity, so a capacity well
    // above the need (a lattice start aliases with the cell grid, and its
    // halos are larger than the liquid's) costs occupancy: shrink then.
    const uint32_t need =
        std::min<uint32_t>(RUMDX_HALO_CAP_MAX, ((f[RUMDX_FLAG_MAXHALO] + f[RUMDX_FLAG_MAXHALO] / 16) + 63u) & ~63u);
    const bool grow = f[RUMDX_FLAG_MAXHALO] > halo_cap_ && halo_cap_ < RUMDX_HALO_CAP_MAX;
    const bool shrink = need + kHaloShrinkSlack <= halo_cap_;
    if (grow || shrink) {
      halo_cap_ = need;
      SizeTileHalo();
      sample_->InvalidateGraphs();
      if (!f[RUMDX_FLAG_OVERFLOW]) return true;
    }
  }
  if (!f[RUMDX_FLAG_OVERFLOW]) return false;
  const unsigned n = sample_->GetNumberOfParticles();
  // the build counts past capacity, so MAXLEN is the true longest list
  unsigned grown = std::min(std::max(max_nbrs_ + max_nbrs_ / 2, f[RUMDX_FLAG_MAXLEN] + f[RUMDX_FLAG_MAXLEN] / 4 + 8), n);
  if (grown <= max_nbrs_) grown = n;
  if (sample_->GetVerbose())
    fprintf(stderr, "[Info] NeighborList: list overflowed (longest %u); growing capacity %u -> %u\n",
            f[RUMDX_FLAG_MAXLEN], max_nbrs_, grown);
  if (user_max_nbrs_) user_max_nbrs_ = grown;
  grown_max_nbrs_ = grown;
  allocated_ = false;
  Allocate();
  return true;
}

void NeighborList::RequestRebuild() {
  if (!sample_) throw RUMD_Error("NeighborList", "RequestRebuild", "no sample");
  EnsureAllocated();
  if (allocated_) rumdx_check(rumdx_nblist_request_rebuild(&desc_, sample_->GetStream()), "NeighborList", "RequestRebuild");
}

void NeighborList::SetHaloCapacity(unsigned cap) {
  if (cap > RUMDX_HALO_CAP_MAX) throw RUMD_Error("NeighborList", "SetHaloCapacity", "capacity above RUMDX_HALO_CAP_MAX");
  halo_cap_ = cap;
  if (allocated_ && desc_.nbl16_dev != nullptr) SizeTileHalo();
  desc_.halo_cap = cap;
  if (sample_) sample_->InvalidateGraphs();
}

// The per-tile halo slot lists are laid out by the halo capacity: resize them
// and rebuild before the next force call (the build writes them).
void NeighborList::SizeTileHalo() {
  tile_halo_.resize(static_cast<size_t>(desc_.ntiles) * std::max(halo_cap_, 1u));
  desc_.tile_halo_dev = tile_halo_.get();
  desc_.halo_cap = halo_cap_;
  if (allocated_) rumdx_check(rumdx_nblist_request_rebuild(&desc_, sample_->GetStream()), "NeighborList", "SizeTileHalo");
}

std::vector<uint32_t> NeighborList::ExportList(std::vector<uint32_t>& counts) {
  const unsigned nvp = sample_->GetNumberOfVirtualParticles();
  counts.assign(nvp, 0u);
  if (!allocated_) return {};
  DeviceArray<uint32_t> out(static_cast<size_t>(max_nbrs_) * nvp);
  rumdx_check(rumdx_nblist_export(sample_->GetParticleData().Descriptor(), &desc_, out.get(), sample_->GetStream()),
              "NeighborList", "ExportList");
  std::vector<uint32_t> h(static_cast<size_t>(max_nbrs_) * nvp);
  out.download(h.data(), h.size(), sample_->GetStream());
  num_nbrs_.download(counts.data(), nvp, sample_->GetStream());
  return h;
}

void NeighborList::DownloadCounts(std::vector<uint32_t>& out) {
  if (allocated_) num_nbrs_.download(out.data(), std::min(out.size(), num_nbrs_.size()), sample_->GetStream());
}

unsigned NeighborList::GetActualMaxNumNbrs() {
  if (!allocated_) return 0;
  uint32_t f[RUMDX_NB_FLAGS];
  ReadFlags(f);
  return f[RUMDX_FLAG_MAXLEN];
}

int NeighborList::GetRebuildRequired() {
  if (!allocated_) return 1;
  uint32_t f[RUMDX_NB_FLAGS];
  flags_.download(f, RUMDX_NB_FLAGS, sample_->GetStream());
  return f[RUMDX_FLAG_REBUILD] ? 1 : 0;
}

unsigned NeighborList::GetNumberOfBuilds() {
  if (!allocated_) return 0;
  uint32_t f[RUMDX_NB_FLAGS];
  flags_.download(f, RUMDX_NB_FLAGS, sample_->GetStream());
  return f[RUMDX_FLAG_BUILDS];
}

std::vector<uint32_t> NeighborList::TileStats() {
  if (!allocated_ || desc_.nbl16_dev == nullptr) return {};
  const hipStream_t st = sample_->GetStream();
  std::vector<uint32_t> info(4ull * desc_.ntiles), f(RUMDX_NB_FLAGS);
  ReadFlags(f.data());
  tile_info_.download(info.data(), info.size(), st);
  // tile_info = {halo particles, flags | halo capacity << 16, slot0, slot1};
  // flag bit 0: fast, bits 1-6: per-axis shift below / above (RUMDX_TILE_*)
  uint32_t staged = 0, fast = 0, wrapping = 0, shifted = 0;
  for (uint32_t t = 0; t < desc_.ntiles; ++t) {
    const uint32_t npart = info[4ull * t], fl = info[4ull * t + 1];
    const bool s = (fl >> 16) == desc_.halo_cap && npart <= desc_.halo_cap;
    const bool fa = (fl & RUMDX_TILE_FAST) != 0u, wr = (fl & RUMDX_TILE_SHIFT_MASK) != 0u;
    staged += s ? 1u : 0u;
    fast += fa ? 1u : 0u;
    wrapping += wr ? 1u : 0u;
    shifted += (fa && wr) ? 1u : 0u;
  }
  return {desc_.ntiles, staged, fast, wrapping, shifted, f[RUMDX_FLAG_MAXHALO]};
}

// ================================================================== PairPotential
PairPotential::PairPotential(CutoffMethod m) : cutoff_method_(m) {
  if (m < NS || m > SF) throw RUMD_Error("PairPotential", "PairPotential", "Invalid cutoffMethod");
  SetID_String("base_pair_pot");
}

void PairPotential::SetParamsRow(unsigned i, unsigned j, const std::vector<float>& row) {
  if (i >= RUMDX_MAX_TYPES || j >= RUMDX_MAX_TYPES) throw RUMD_Error(GetID_String(), "SetParams", "type out of range");
  if (row.size() + 2 > kNumParam)
    throw RUMD_Error(GetID_String(), "CopyParamsToGPU", "more parameters than NumParam allows");
  std::vector<float> p(row);
  p.resize(kNumParam, 0.f);
  ComputeShifts(p);
  params_map_[{i, j}] = p;
  if (i != j && assume_newton3_) params_map_[{j, i}] = p;
  if (sample) {
    UploadParams();
    neighborList.SetCutoffs(this);
    sample->InvalidateGraphs();
  }
}

float PairPotential::GetPotentialParameter(unsigned i, unsigned j, unsigned idx) const {
  auto it = params_map_.find({i, j});
  if (it == params_map_.end() || idx >= it->second.size()) return 0.f;
  return it->second[idx];
}

std::vector<float> PairPotential::GetPotentialParameters(unsigned num_types) const {
  std::vector<float> v;
  for (unsigned i = 0; i < num_types; ++i)
    for (unsigned j = 0; j < num_types; ++j)
      for (unsigned p = 0; p < kNumParam - 2; ++p) v.push_back(GetPotentialParameter(i, j, p));
  return v;
}

std::vector<float> PairPotential::GetParamsTable(unsigned nt) const {
  std::vector<float> tab(static_cast<size_t>(nt) * nt * kNumParam, 0.f);
  for (unsigned i = 0; i < nt; ++i)
    for (unsigned j = 0; j < nt; ++j) {
      auto it = params_map_.find({i, j});
      if (it != params_map_.end())
        std::copy(it->second.begin(), it->second.end(), tab.begin() + (static_cast<size_t>(i) * nt + j) * kNumParam);
    }
  return tab;
}

void PairPotential::UploadParams() {
  const unsigned nt = sample->GetNumberOfTypes();
  const std::vector<float> tab = GetParamsTable(nt);
  if (d_params_.size() != tab.size()) d_params_.resize(tab.size());
  d_params_.upload(tab.data(), tab.size(), sample->GetStream());
  table_types_ = nt;
}

void PairPotential::Initialize() {
  UploadParams();
  neighborList.Initialize(sample);
  neighborList.SetCutoffs(this);
}

// The force kernel of this potential: FP32 pair arithmetic, or the all-FP64
// comparison path after SetFP64Forces(true).
void PairPotential::LaunchForce(const rumdx_particles* p, bool initialize, bool calc_stresses, hipStream_t st,
                                const char* method) {
  rumdx_nblist* nb = neighborList.GetResolvedMethod() == NeighborList::NONE ? nullptr : neighborList.Descriptor();
  const float* box = sample->GetSimulationBox()->GetBox8();
  const int form = GetPotentialForm();
  int rc;
  if (form != 0)
    rc = rumdx_pair_form(p, nb, d_params_.get(), table_types_, box, form, cutoff_method_, initialize ? 1 : 0,
                         calc_stresses ? 1 : 0, st);
  else if (fp64_forces_)
    rc = rumdx_pair_lj_fp64(p, nb, d_params_.get(), table_types_, box, cutoff_method_, initialize ? 1 : 0,
                            calc_stresses ? 1 : 0, st);
  else
    rc = rumdx_pair_lj(p, nb, d_params_.get(), table_types_, box, cutoff_method_, initialize ? 1 : 0,
                       calc_stresses ? 1 : 0, st);
  rumdx_check(rc, GetID_String().c_str(), method);
}

bool PairPotential::CanFuseStep() {
  if (!sample || GetPotentialForm() != 0 || fp64_forces_) return false;
  if (neighborList.GetResolvedMethod() != NeighborList::SORT) return false;
  const rumdx_nblist* nb = neighborList.Descriptor();
  if (nb->nbl16_dev == nullptr || nb->ntiles == 0) return false;
  // the workspace of rumdx_workspace_bytes(n) holds 4 partials per 256 particles
  const size_t n = sample->GetNumberOfParticles();
  return static_cast<size_t>(nb->ntiles) <= 4 * ((n + 255) / 256);
}

void PairPotential::CalcFIntegrateNVT(rumdx_nh_state* nh_dev, float dt, unsigned thermostat_on_mask, float omega2,
                                      unsigned dof) {
  if (!sample) throw RUMD_Error(GetID_String(), "CalcF", "no sample");
  if (table_types_ != sample->GetNumberOfTypes()) Initialize();
  neighborList.UpdateNBlist();
  ParticleData& pd = sample->GetParticleData();
  if (!pd.HasNextPositions())
    throw RUMD_Error(GetID_String(), "CalcF", "the fused step needs the sample's second position array");
  rumdx_check(rumdx_pair_lj_nvt(pd.Descriptor(), neighborList.Descriptor(), d_params_.get(), table_types_,
                                sample->GetSimulationBox()->GetBox8(), cutoff_method_, pd.NextPositions(), nh_dev, dt,
                                thermostat_on_mask, omega2, dof, sample->Workspace(), sample->GetStream()),
              GetID_String().c_str(), "CalcF");
  pd.SwapPositions();
}

void PairPotential::SetFP64Forces(bool on) {
  if (on && GetPotentialForm() != 0)
    throw RUMD_Error(GetID_String(), "SetFP64Forces", "the all-FP64 comparison path serves the 12-6 potentials only");
  fp64_forces_ = on;
  if (sample) sample->InvalidateGraphs();
}

void PairPotential::CalcF(bool initialize, bool calc_stresses) {
  if (!sample) throw RUMD_Error(GetID_String(), "CalcF", "no sample");
  if (table_types_ != sample->GetNumberOfTypes()) Initialize();
  neighborList.UpdateNBlist();
  LaunchForce(sample->GetParticleData().Descriptor(), initialize, calc_stresses, sample->GetStream(), "CalcF");
}

// PairPotential::GetPotentialEnergy (src/PairPotential.cu:252-267): this
// potential's energy alone, computed into private arrays, summed in FP64.
double PairPotential::GetPotentialEnergy() {
  if (!sample) throw RUMD_Error(GetID_String(), "GetPotentialEnergy", "no sample");
  if (table_types_ != sample->GetNumberOfTypes()) Initialize();
  neighborList.UpdateNBlist();
  const unsigned nvp = sample->GetNumberOfVirtualParticles();
  if (pe_f_.size() != 4ull * nvp) {
    pe_f_.resize(4ull * nvp);
    pe_w_.resize(4ull * nvp);
  }
  rumdx_particles p = *sample->GetParticleData().Descriptor();
  p.f_dev = pe_f_.get();
  p.w_dev = pe_w_.get();
  const hipStream_t st = sample->GetStream();
  LaunchForce(&p, true, false, st, "GetPotentialEnergy");
  rumdx_check(rumdx_reduce_f4(pe_f_.get(), p.n, 8u, sample->ScalarOut(), sample->Workspace(), st),
              GetID_String().c_str(), "GetPotentialEnergy");
  double out[4];
  hip_check(hipMemcpyAsync(out, sample->ScalarOut(), sizeof(out), hipMemcpyDeviceToHost, st), "PairPotential",
            "GetPotentialEnergy");
  hip_check(hipStreamSynchronize(st), "PairPotential", "GetPotentialEnergy");
  return out[3];
}

float PairPotential::TimeForceKernel(unsigned reps) {
  if (!sample) throw RUMD_Error(GetID_String(), "TimeForceKernel", "no sample");
  if (reps == 0) reps = 1;
  neighborList.UpdateNBlist();
  const hipStream_t st = sample->GetStream();
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "PairPotential", "TimeForceKernel");
  hip_check(hipEventCreate(&b), "PairPotential", "TimeForceKernel");
  hip_check(hipEventRecord(a, st), "PairPotential", "TimeForceKernel");
  for (unsigned k = 0; k < reps; ++k)
    LaunchForce(sample->GetParticleData().Descriptor(), true, false, st, "TimeForceKernel");
  hip_check(hipEventRecord(b, st), "PairPotential", "TimeForceKernel");
  hip_check(hipEventSynchronize(b), "PairPotential", "TimeForceKernel");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "PairPotential", "TimeForceKernel");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

float PairPotential::TimeFusedStepKernel(unsigned reps, float dt) {
  if (!sample) throw RUMD_Error(GetID_String(), "TimeFusedStepKernel", "no sample");
  if (!CanFuseStep()) throw RUMD_Error(GetID_String(), "TimeFusedStepKernel", "this potential takes no fused step");
  if (reps == 0) reps = 1;
  neighborList.UpdateNBlist();
  const hipStream_t st = sample->GetStream();
  ParticleData& pd = sample->GetParticleData();
  const size_t f4 = 4ull * sample->GetNumberOfVirtualParticles();
  DeviceArray<float> v(f4), im(f4), rlo, r_out(f4);
  DeviceArray<uint32_t> flags(RUMDX_NB_FLAGS);
  rumdx_particles p = *pd.Descriptor();
  rumdx_nblist nb = *neighborList.Descriptor();
  const auto copy = [&](void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st), "PairPotential", "TimeFusedStepKernel");
  };
  copy(v.get(), p.v_dev, sizeof(float) * f4);
  copy(im.get(), p.im_dev, sizeof(float) * f4);
  if (p.rlo_dev != nullptr) {
    rlo.resize(f4);
    copy(rlo.get(), p.rlo_dev, sizeof(float) * f4);
    p.rlo_dev = rlo.get();
  }
  copy(flags.get(), nb.flags_dev, sizeof(uint32_t) * RUMDX_NB_FLAGS);
  p.v_dev = v.get();
  p.im_dev = im.get();
  nb.flags_dev = flags.get();
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "PairPotential", "TimeFusedStepKernel");
  hip_check(hipEventCreate(&b), "PairPotential", "TimeFusedStepKernel");
  hip_check(hipEventRecord(a, st), "PairPotential", "TimeFusedStepKernel");
  for (unsigned k = 0; k < reps; ++k)
    rumdx_check(rumdx_pair_lj_nvt(&p, &nb, d_params_.get(), table_types_, sample->GetSimulationBox()->GetBox8(),
                                  cutoff_method_, r_out.get(), nullptr, dt, 0u, 0.f, 0u, nullptr, st),
                GetID_String().c_str(), "TimeFusedStepKernel");
  hip_check(hipEventRecord(b, st), "PairPotential", "TimeFusedStepKernel");
  hip_check(hipEventSynchronize(b), "PairPotential", "TimeFusedStepKernel");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "PairPotential", "TimeFusedStepKernel");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

float PairPotential::TimeNbListBuild(unsigned reps, bool full_sort) {
  if (!sample) throw RUMD_Error(GetID_String(), "TimeNbListBuild", "no sample");
  if (reps == 0) reps = 1;
  neighborList.UpdateNBlist();
  sample->Synchronize();  // settles list and halo capacities first
  neighborList.UpdateNBlist();
  neighborList.SetForceFullSort(full_sort);
  const hipStream_t st = sample->GetStream();
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "PairPotential", "TimeNbListBuild");
  hip_check(hipEventCreate(&b), "PairPotential", "TimeNbListBuild");
  hip_check(hipEventRecord(a, st), "PairPotential", "TimeNbListBuild");
  for (unsigned k = 0; k < reps; ++k) {
    neighborList.RequestRebuild();
    neighborList.UpdateNBlist();
  }
  hip_check(hipEventRecord(b, st), "PairPotential", "TimeNbListBuild");
  hip_check(hipEventSynchronize(b), "PairPotential", "TimeNbListBuild");
  neighborList.SetForceFullSort(false);
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "PairPotential", "TimeNbListBuild");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

std::vector<double> PairPotential::GetNeighborListStats() {
  if (!sample) throw RUMD_Error(GetID_String(), "GetNeighborListStats", "no sample");
  const unsigned n = sample->GetNumberOfParticles();
  std::vector<uint32_t> cnt(sample->GetNumberOfVirtualParticles());
  neighborList.DownloadCounts(cnt);
  double sum = 0.0, mx = 0.0;
  for (unsigned i = 0; i < n; ++i) {
    sum += cnt[i];
    mx = std::max(mx, static_cast<double>(cnt[i]));
  }
  return {sum / n, mx};
}

// ================================================================== Pot_IPL_12_6 / Pot_LJ_12_6
void Pot_IPL_12_6::SetParams(unsigned i, unsigned j, float A12, float A6, float Rcut) {
  if (verbose && sample && sample->GetVerbose())
    printf("%s: A12 = %g, A6 = %g, Rcut = %g.\n", GetID_String().c_str(), A12, A6, Rcut);
  SetParamsRow(i, j, {Rcut, A12, A6});
}

// u(rc) and s(rc) rc for the shifted potential / shifted force cut-offs
// (CopyParamsToGPU, src/PairPotential.cu:82-125), formed in FP64 from the
// stored float parameters.
void Pot_IPL_12_6::ComputeShifts(std::vector<float>& row) const {
  const double rc = row[0], a12 = row[1], a6 = row[2];
  double u = 0.0, s = 0.0;
  if (rc > 0.0) {
    const double i2 = 1.0 / (rc * rc), i6 = i2 * i2 * i2;
    u = i6 * (a12 * i6 + a6);
    s = i6 * i2 * (12.0 * a12 * i6 + 6.0 * a6);
  }
  row[kNumParam - 1] = static_cast<float>(u);
  row[kNumParam - 2] = static_cast<float>(s * rc);
}

void Pot_LJ_12_6::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  if (verbose && sample && sample->GetVerbose())
    printf("%s: Sigma = %g, Epsilon = %g, Rcut = %g.\n", GetID_String().c_str(), Sigma, Epsilon, Rcut);
  // include/rumd/PairPotential.h:202-207
  Pot_IPL_12_6::SetParams(i, j, static_cast<float>(4. * Epsilon * std::pow(static_cast<double>(Sigma), 12.)),
                          static_cast<float>(-4. * Epsilon * std::pow(static_cast<double>(Sigma), 6.)), Rcut * Sigma);
}

void Pot_gLJ_m12_n6::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  if (verbose && sample && sample->GetVerbose())
    printf("%s: Sigma = %g, Epsilon = %g, Rcut = %g.\n", GetID_String().c_str(), Sigma, Epsilon, Rcut);
  // include/rumd/PairPotential.h:322-327
  Pot_IPL_12_6::SetParams(i, j, static_cast<float>(Epsilon * std::pow(static_cast<double>(Sigma), 12.)),
                          static_cast<float>(-2. * Epsilon * std::pow(static_cast<double>(Sigma), 6.)), Rcut * Sigma);
}

// ================================================================== other potential forms
// u(rc) and s(rc) rc of a form (CopyParamsToGPU, src/PairPotential.cu:82-125),
// formed in FP64 from the stored float parameters.
void FormPairPotential::ComputeShifts(std::vector<float>& row) const {
  double p[10];
  for (int k = 0; k < 10; ++k) p[k] = row[k];
  const double rc = p[0];
  double u = 0.0, s = 0.0;
  if (rc > 0.0) s = rumdx::pot_form<double>(form_, rc * rc, p, u);
  row[kNumParam - 1] = static_cast<float>(u);
  row[kNumParam - 2] = static_cast<float>(s * rc);
}

void FormPairPotential::Report(const char* what) const {
  if (verbose && sample && sample->GetVerbose()) printf("%s: %s.\n", GetID_String().c_str(), what);
}

Pot_LJ_12_6_smooth_2_4::Pot_LJ_12_6_smooth_2_4(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_LJ_SMOOTH_2_4) {
  SetID_String("potLJ_12_6_smooth_2_4");
}

// Rcut in units of Sigma; c2, c4 make the force and its derivative vanish
// at Rcut (include/rumd/PairPotential.h:234-247).
void Pot_LJ_12_6_smooth_2_4::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  const float c2 = static_cast<float>(48. * std::pow(static_cast<double>(Rcut), -14.) -
                                      15. * std::pow(static_cast<double>(Rcut), -8.));
  const float c4 = static_cast<float>(6. * std::pow(static_cast<double>(Rcut), -10.) -
                                      21. * std::pow(static_cast<double>(Rcut), -16.));
  SetParamsRow(i, j, {Rcut * Sigma, 1.0f / (Sigma * Sigma), 4.0f * Epsilon, c2, c4});
  Report("smooth LJ parameters set");
}

static std::string glj_id(float m, float n) {
  std::ostringstream s;
  s << "potgLJ_" << m << "_" << n;
  return s.str();
}

Pot_gLJ_m_n::Pot_gLJ_m_n(float m_, float n_, CutoffMethod cm) : FormPairPotential(cm, RUMDX_POT_GLJ_M_N), m(m_), n(n_) {
  SetID_String(glj_id(m, n));
}

// include/rumd/PairPotential.h:280-292
void Pot_gLJ_m_n::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), m * n / (m - n) * Epsilon, m * 0.5f, n * 0.5f,
                      Epsilon / (m - n)});
  Report("generalised LJ parameters set");
}

Pot_Gauss::Pot_Gauss(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_GAUSS) { SetID_String("potGau"); }

// include/rumd/PairPotential.h:397-407
void Pot_Gauss::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), Epsilon});
  Report("Gaussian parameters set");
}

Pot_Buckingham::Pot_Buckingham(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_BUCKINGHAM) { SetID_String("potBuc"); }

// Rcut is absolute here (include/rumd/PairPotential.h:436-449)
void Pot_Buckingham::SetParams(unsigned i, unsigned j, float rm, float Alpha, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut, 1.f / (rm * rm), Alpha, 6.f * Epsilon / (Alpha - 6.f), Alpha * Epsilon / (Alpha - 6.f),
                      6.f * Epsilon * Alpha / (Alpha - 6)});
  Report("Buckingham parameters set");
}

static std::string ipl_id(float n) {
  std::ostringstream s;
  s << "potIPL_" << n;
  return s.str();
}

Pot_IPL_n::Pot_IPL_n(float n_, CutoffMethod cm) : FormPairPotential(cm, RUMDX_POT_IPL_N), n(n_) {
  SetID_String(ipl_id(n));
}

// include/rumd/PairPotential.h:614-625
void Pot_IPL_n::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), Epsilon, n});
  Report("IPL parameters set");
}

Pot_Yukawa::Pot_Yukawa(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_YUKAWA) { SetID_String("potYukawa"); }

// include/rumd/PairPotential.h:814-826
void Pot_Yukawa::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, Sigma, 1.f / Sigma, Epsilon});
  Report("Yukawa parameters set");
}

Pot_Exp::Pot_Exp(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_EXP) { SetID_String("potExp"); }

// include/rumd/PairPotential.h:1035-1044
void Pot_Exp::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / Sigma, Epsilon});
  Report("exponential parameters set");
}

Pot_LJ_12_6_Gauss::Pot_LJ_12_6_Gauss(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_LJ_GAUSS) {
  SetID_String("potLJG");
}

// include/rumd/PairPotential.h:350-365
void Pot_LJ_12_6_Gauss::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut, float Sigma0,
                                  float Epsilon0, float r0) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), 12.f * Epsilon, 2.0f * Sigma0 * Sigma0, r0, Epsilon0,
                      Epsilon});
  Report("LJ-Gauss parameters set");
}

Pot_IPL_fixed::Pot_IPL_fixed(CutoffMethod m, int form, float n, const char* id) : FormPairPotential(m, form), n_(n) {
  SetID_String(id);
}

// include/rumd/PairPotential.h:483-495 (and :526-538, :572-584)
void Pot_IPL_fixed::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), n_ * Epsilon, Epsilon});
  Report("IPL parameters set");
}

Pot_IPL6_LN::Pot_IPL6_LN(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_IPL6_LN) { SetID_String("potIPL6LN"); }

// include/rumd/PairPotential.h:658-669
void Pot_IPL6_LN::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, Epsilon, 1.f / (Sigma * Sigma), Epsilon / (Sigma * Sigma)});
  Report("IPL6/ln parameters set");
}

Pot_LJ_LN::Pot_LJ_LN(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_LJ_LN) { SetID_String("pot_LJ_LN"); }

// include/rumd/PairPotential.h:703-719: a = ln(Rm/R0) + 1/6, b = 2 ln(Rm/R0) + 1/6
void Pot_LJ_LN::SetParams(unsigned i, unsigned j, float Rm, float R0, float Epsilon, float Rcut) {
  const float l = std::log(Rm / R0);
  const float p1 = 1.f / (Rm * Rm);
  const float p3 = Epsilon * (l + 1.f / 6.f);
  SetParamsRow(i, j, {Rcut * Rm, p1, 1.f / (R0 * R0), p3, p1 * p3, (12.f * l + 1) / (6.f * l + 1)});
  Report("LJ/ln parameters set");
}

Pot_Girifalco::Pot_Girifalco(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_GIRIFALCO) {
  SetID_String("pot_Girifalco");
}

// include/rumd/PairPotential.h:757-768: Rcut in units of R0
void Pot_Girifalco::SetParams(unsigned i, unsigned j, float alpha, float beta, float R0, float Rcut) {
  SetParamsRow(i, j, {Rcut * R0, 1.f / (R0 * R0), alpha, beta});
  Report("Girifalco parameters set");
}

Pot_Dzugutov::Pot_Dzugutov() : FormPairPotential(NS, RUMDX_POT_DZUGUTOV) { SetID_String("potDzu"); }

// include/rumd/PairPotential.h:860-871: the cut-off sits just inside b
void Pot_Dzugutov::SetParams(unsigned i, unsigned j, float a, float b, float c, float d, float A, float B) {
  SetParamsRow(i, j, {b - 0.000001f, a, b, c, d, A, B});
  Report("Dzugutov parameters set");
}

Pot_Fermi_Jagla::Pot_Fermi_Jagla(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_FERMI_JAGLA) {
  SetID_String("potFermiJagla");
}

// include/rumd/PairPotential.h:924-941
void Pot_Fermi_Jagla::SetParams(unsigned i, unsigned j, float a, float epsilon0, float A0, float A1, float A2,
                                float B0, float B1, float B2, float n, float Rcut) {
  SetParamsRow(i, j, {Rcut * a, 1.f / (a * a), A0, A1 / A0, A2, B0, B1 / B0, B2, epsilon0, n});
  Report("Fermi-Jagla parameters set");
}

Pot_Repulsive_Shoulder::Pot_Repulsive_Shoulder(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_REP_SHOULDER) {
  SetID_String("potRepulsiveShoulder");
}

// include/rumd/PairPotential.h:983-994
void Pot_Repulsive_Shoulder::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float k0, float Sigma1,
                                       float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, 1.f / (Sigma * Sigma), 0.5f * Epsilon, k0, Sigma1});
  Report("repulsive shoulder parameters set");
}

Pot_SAAP::Pot_SAAP(CutoffMethod m) : FormPairPotential(m, RUMDX_POT_SAAP) { SetID_String("SAAP"); }

// include/rumd/PairPotential.h:1074-1090
void Pot_SAAP::SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float a0, float a1, float a2, float a3,
                         float a4, float a5, float a6, float Rcut) {
  SetParamsRow(i, j, {Rcut * Sigma, a0, a1, a2, a3, a4, a5, a6, 1.f / (Sigma * Sigma), Epsilon});
  Report("SAAP parameters set");
}

Pot_ShiftedForceCoulomb::Pot_ShiftedForceCoulomb() : FormPairPotential(NS, RUMDX_POT_SF_COULOMB) {
  SetID_String("potCoul");
}

// include/rumd/PairPotential.h:1165-1172: Rcut is absolute
void Pot_ShiftedForceCoulomb::SetParams(unsigned i, unsigned j, float Epsilon, float Rcut) {
  SetParamsRow(i, j, {Rcut, Epsilon, Epsilon / Rcut, Epsilon / (Rcut * Rcut)});
  Report("shifted-force Coulomb parameters set");
}
