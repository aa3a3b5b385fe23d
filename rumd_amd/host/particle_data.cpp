// ParticleData: device arrays in slot order, host access in original order
// (include/rumd/ParticleData.h, src/ParticleData.cu:86-473).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "rumd_host.h"

void ParticleData::SetNumberOfParticles(unsigned n, unsigned pb, bool mixed_positions) {
  if (n == 0) throw RUMD_Error("ParticleData", "SetNumberOfParticles", "zero particles");
  if (pb == 0) throw RUMD_Error("ParticleData", "SetNumberOfParticles", "zero particles per block");
  n_ = n;
  nvp_ = ((n + pb - 1) / pb) * pb;  // virtual particles (src/ParticleData.cu:86-108)
  mixed_ = mixed_positions;
  const size_t f4 = 4ull * nvp_;
  d_r.resize(f4);
  d_v.resize(f4);
  d_f.resize(f4);
  d_w.resize(f4);
  d_im.resize(f4);
  d_sts.resize(f4);
  d_r_next.release();
  if (mixed_) d_rlo.resize(f4); else d_rlo.release();
  d_unsorted_index.resize(nvp_);
  std::vector<uint32_t> id(nvp_);
  std::iota(id.begin(), id.end(), 0u);
  d_unsorted_index.upload(id.data(), nvp_, nullptr);
  types_.assign(n_, 0);
  if (masses_.empty()) masses_.assign(1, 1.f);
  // v.w = 1/m must be valid before the first step
  std::vector<float> v(f4, 0.f);
  for (unsigned s = 0; s < nvp_; ++s) v[4 * s + 3] = 1.f / masses_[0];
  d_v.upload(v.data(), f4, nullptr);
  desc_.r_dev = d_r.get();
  desc_.v_dev = d_v.get();
  desc_.f_dev = d_f.get();
  desc_.w_dev = d_w.get();
  desc_.im_dev = d_im.get();
  desc_.sts_dev = d_sts.get();
  desc_.rlo_dev = mixed_ ? d_rlo.get() : nullptr;
  desc_.unsorted_index_dev = d_unsorted_index.get();
  desc_.n = n_;
  desc_.nvp = nvp_;
}

void ParticleData::EnsureNextPositions() {
  if (!HasNextPositions()) d_r_next.resize(d_r.size());  // zero-filled like d_r's padding slots
}

void ParticleData::SwapPositions() {
  std::swap(d_r, d_r_next);
  desc_.r_dev = d_r.get();
}

void ParticleData::MovePositionsTo(float* home, hipStream_t st) {
  if (d_r.get() == home) return;
  if (home == nullptr || d_r_next.get() != home)
    throw RUMD_Error("ParticleData", "MovePositionsTo", "not a position array of this sample");
  hip_check(hipMemcpyAsync(home, d_r.get(), sizeof(float) * d_r.size(), hipMemcpyDeviceToDevice, st), "ParticleData",
            "MovePositionsTo");
  SwapPositions();
}

unsigned ParticleData::GetNumberThisType(unsigned t) const {
  return static_cast<unsigned>(std::count(types_.begin(), types_.end(), static_cast<int32_t>(t)));
}

void ParticleData::SetNumberOfTypes(unsigned nt) {
  if (nt == 0 || nt > RUMDX_MAX_TYPES) throw RUMD_Error("ParticleData", "SetNumberOfTypes", "number of types out of range");
  masses_.resize(nt, 1.f);
}

void ParticleData::SetMass(unsigned type, float mass) {
  if (!(mass > 0.f)) throw RUMD_Error("ParticleData", "SetMass", "mass must be positive");
  if (type >= masses_.size()) SetNumberOfTypes(type + 1);
  masses_[type] = mass;
  if (n_) UpdateMassesOnDevice(nullptr);
}

float ParticleData::GetMass(unsigned type) const {
  if (type >= masses_.size()) throw RUMD_Error("ParticleData", "GetMass", "no such type");
  return masses_[type];
}

std::vector<uint32_t> ParticleData::UnsortedIndex(hipStream_t st) {
  std::vector<uint32_t> idx(nvp_);
  d_unsorted_index.download(idx.data(), nvp_, st);
  return idx;
}

std::vector<float> ParticleData::GatherToSlots(const float* orig4, hipStream_t st) {
  const std::vector<uint32_t> idx = UnsortedIndex(st);
  std::vector<float> slot(4ull * nvp_, 0.f);
  for (unsigned s = 0; s < n_; ++s) std::memcpy(&slot[4ull * s], orig4 + 4ull * idx[s], 4 * sizeof(float));
  return slot;
}

void ParticleData::Scatter(const std::vector<float>& slot4, float* out4, hipStream_t st) {
  const std::vector<uint32_t> idx = UnsortedIndex(st);
  for (unsigned s = 0; s < n_; ++s) std::memcpy(out4 + 4ull * idx[s], &slot4[4ull * s], 4 * sizeof(float));
}

void ParticleData::SetPositions(const float* xyz, hipStream_t st) {
  std::vector<float> o(4ull * n_);
  for (unsigned i = 0; i < n_; ++i) {
    o[4ull * i] = xyz[3ull * i];
    o[4ull * i + 1] = xyz[3ull * i + 1];
    o[4ull * i + 2] = xyz[3ull * i + 2];
    std::memcpy(&o[4ull * i + 3], &types_[i], sizeof(int32_t));
  }
  const std::vector<float> slot = GatherToSlots(o.data(), st);
  d_r.upload(slot.data(), slot.size(), st);
  if (mixed_) {
    std::vector<float> z(4ull * nvp_, 0.f);
    d_rlo.upload(z.data(), z.size(), st);
  }
}

void ParticleData::SetPositionsFP64(const double* xyz, hipStream_t st) {
  std::vector<float> hi(3ull * n_);
  for (size_t k = 0; k < hi.size(); ++k) hi[k] = static_cast<float>(xyz[k]);
  SetPositions(hi.data(), st);
  if (!mixed_) return;
  std::vector<float> lo(4ull * n_, 0.f);
  for (unsigned i = 0; i < n_; ++i)
    for (int k = 0; k < 3; ++k)
      lo[4ull * i + k] = static_cast<float>(xyz[3ull * i + k] - static_cast<double>(hi[3ull * i + k]));
  const std::vector<float> slot = GatherToSlots(lo.data(), st);
  d_rlo.upload(slot.data(), slot.size(), st);
}

void ParticleData::SetTypes(const int32_t* types, hipStream_t st) {
  unsigned nt = static_cast<unsigned>(masses_.size());
  for (unsigned i = 0; i < n_; ++i) {
    if (types[i] < 0 || types[i] >= RUMDX_MAX_TYPES) throw RUMD_Error("ParticleData", "SetTypes", "type out of range");
    nt = std::max(nt, static_cast<unsigned>(types[i]) + 1);
  }
  std::vector<float> r(4ull * n_);
  Get("r", r.data(), st);
  types_.assign(types, types + n_);
  SetNumberOfTypes(nt);
  std::vector<float> xyz(3ull * n_);
  for (unsigned i = 0; i < n_; ++i)
    for (int k = 0; k < 3; ++k) xyz[3ull * i + k] = r[4ull * i + k];
  // keep the low-order parts across a type change
  std::vector<float> lo;
  if (mixed_) {
    lo.resize(4ull * nvp_);
    d_rlo.download(lo.data(), lo.size(), st);
  }
  SetPositions(xyz.data(), st);
  if (mixed_) d_rlo.upload(lo.data(), lo.size(), st);
  UpdateMassesOnDevice(st);
}

void ParticleData::UpdateMassesOnDevice(hipStream_t st) {
  std::vector<float> v(4ull * nvp_);
  d_v.download(v.data(), v.size(), st);
  const std::vector<uint32_t> idx = UnsortedIndex(st);
  for (unsigned s = 0; s < n_; ++s) v[4ull * s + 3] = 1.f / masses_[types_[idx[s]]];
  d_v.upload(v.data(), v.size(), st);
}

void ParticleData::SetVelocities(const float* xyz, hipStream_t st) {
  std::vector<float> o(4ull * n_);
  for (unsigned i = 0; i < n_; ++i) {
    for (int k = 0; k < 3; ++k) o[4ull * i + k] = xyz[3ull * i + k];
    o[4ull * i + 3] = 1.f / masses_[types_[i]];
  }
  std::vector<float> slot = GatherToSlots(o.data(), st);
  for (unsigned s = n_; s < nvp_; ++s) slot[4ull * s + 3] = 1.f / masses_[0];
  d_v.upload(slot.data(), slot.size(), st);
}

void ParticleData::SetImages(const float* xyz, hipStream_t st) {
  std::vector<float> o(4ull * n_, 0.f);
  for (unsigned i = 0; i < n_; ++i)
    for (int k = 0; k < 3; ++k) o[4ull * i + k] = xyz[3ull * i + k];
  const std::vector<float> slot = GatherToSlots(o.data(), st);
  d_im.upload(slot.data(), slot.size(), st);
}

void ParticleData::SetForces(const float* f4, hipStream_t st) {
  const std::vector<float> slot = GatherToSlots(f4, st);
  d_f.upload(slot.data(), slot.size(), st);
}

void ParticleData::Get(const char* what, float* out4, hipStream_t st) {
  const DeviceArray<float>* a = nullptr;
  const std::string w(what);
  if (w == "r") a = &d_r;
  else if (w == "v") a = &d_v;
  else if (w == "f") a = &d_f;
  else if (w == "w") a = &d_w;
  else if (w == "im") a = &d_im;
  else if (w == "sts") a = &d_sts;
  else if (w == "rlo") a = mixed_ ? &d_rlo : nullptr;
  if (a == nullptr) throw RUMD_Error("ParticleData", "Get", "unknown array " + w);
  std::vector<float> slot(4ull * nvp_);
  a->download(slot.data(), slot.size(), st);
  Scatter(slot, out4, st);
}

void ParticleData::GetTypes(int32_t* out, hipStream_t) { std::copy(types_.begin(), types_.end(), out); }

void ParticleData::ResetSorting(hipStream_t st) {
  // bring every array back to original order, then identity index
  std::vector<float> r(4ull * n_), v(4ull * n_), f(4ull * n_), w(4ull * n_), im(4ull * n_), lo;
  Get("r", r.data(), st);
  Get("v", v.data(), st);
  Get("f", f.data(), st);
  Get("w", w.data(), st);
  Get("im", im.data(), st);
  if (mixed_) {
    lo.resize(4ull * n_);
    Get("rlo", lo.data(), st);
  }
  std::vector<uint32_t> id(nvp_);
  std::iota(id.begin(), id.end(), 0u);
  d_unsorted_index.upload(id.data(), nvp_, st);
  auto put = [&](DeviceArray<float>& d, std::vector<float>& h) {
    h.resize(4ull * nvp_, 0.f);
    d.upload(h.data(), h.size(), st);
  };
  put(d_r, r);
  put(d_v, v);
  put(d_f, f);
  put(d_w, w);
  put(d_im, im);
  if (mixed_) put(d_rlo, lo);
}
