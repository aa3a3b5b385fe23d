// Integrator and FP64 reduction kernels (gfx950).
//
// Reference: IntegratorNVT::Integrate / integrateNVTAlgorithm
// (src/IntegratorNVT.cu:64-79, 364-418), UpdateThermostatState /
// updateNVTThermostatState (:82-92, 426-437), SetMomentumToZero with
// calculateMomentum / zeroTotalMomentum (:176-183, 444-472), GetKineticEnergy
// (:316-356); sumIdenticalArrays / reduce (src/rumd_algorithms.cu:162-250).
//
// The leapfrog update is one streaming pass with the thermostat's
// kinetic-energy sum fused in as an FP64 block reduction; a one-workgroup
// kernel then adds the block partials in a fixed order and advances Ps in FP64
// (bitwise reproducible).  The reference instead stores m v^2 per particle and
// reduces it in FP32 over ~4 launches.
#include <atomic>
#include <mutex>

#include "leapfrog.h"
#include "rumdx_common.h"

namespace rumdx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return RUMDX_OK;
  const int code = (e == hipErrorLaunchOutOfResources || e == hipErrorInvalidConfiguration) ? RUMDX_E_RESOURCE : RUMDX_E_HIP;
  return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

struct NvtArgs {
  float4* r;
  float4* v;
  const float4* f;
  float4* im;
  float4* rlo;
  rumdx_nh_state* nh;
  double* partials;
  const float4* last_r;  // skin check fused when non-null
  uint32_t* nb_flags;
  float skin2_over_4;
  const float4* prune_r;  // inner-list check fused when non-null
  float inner2;
  float dt;
  float omega2;
  uint32_t on_mask;
  uint32_t dof;
  uint32_t n;
};

template <bool THERMO, bool HILO, bool SKIN>
__global__ __launch_bounds__(kBlock) void k_nvt(NvtArgs a, Box bx) {
  __shared__ double s_red[kBlock / 64];
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = b * kBlock + threadIdx.x;
  float ke2 = 0.f;
  const float ps_all = THERMO ? static_cast<float>(a.nh->ps) : 0.f;
  if (i < a.n) {
    float4 r = a.r[i], v = a.v[i];
    const float4 f = a.f[i];
    float4 lo = HILO ? a.rlo[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool apply = THERMO && ((a.on_mask >> (__float_as_uint(r.w) & 31u)) & 1u);
    float3 shift;
    const float ke = nvt_particle<HILO>(r, v, f, shift, lo, apply ? ps_all : 0.f, a.dt, bx);
    if (apply) ke2 = ke;
    a.r[i] = r;
    a.v[i] = v;
    add_image_shift(a.im, i, shift);
    if (HILO) a.rlo[i] = lo;
    if (SKIN) {
      if (min_image_d2_f32(r, a.last_r[i], bx) > a.skin2_over_4) a.nb_flags[RUMDX_FLAG_REBUILD] = 1u;
      if (a.prune_r != nullptr && min_image_d2_f32(r, a.prune_r[i], bx) > a.inner2)
        a.nb_flags[RUMDX_FLAG_PRUNE] = 1u;
    }
  }
  if (THERMO) {
    const double s = block_sum(static_cast<double>(ke2), s_red);
    if (threadIdx.x == 0) a.partials[b] = s;
  }
}

// Fixed-order sums of `ncomp`-strided block partials into out[c] * scale.
__global__ __launch_bounds__(kBlock) void k_finish_sums(const double* __restrict__ partials, uint32_t np,
                                                        uint32_t ncomp, uint32_t mask, double scale,
                                                        double* __restrict__ out) {
  __shared__ double s_red[kBlock / 64];
  for (uint32_t c = 0; c < ncomp; ++c) {
    double s = 0.0;
    if ((mask >> c) & 1u) s = sum_partials(partials + c, np, ncomp, s_red);
    if (threadIdx.x == 0) out[c] = s * scale;
  }
}

// ---------------------------------------------------------------- momentum
__global__ __launch_bounds__(kBlock) void k_momentum_sum(const float4* __restrict__ v, double* partials, uint32_t n) {
  __shared__ double s_red[kBlock / 64];
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = b * kBlock + threadIdx.x;
  double p[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const float4 vi = v[i];
    const float m = 1.f / vi.w;  // calculateMomentum (src/IntegratorNVT.cu:462-470)
    p[0] = static_cast<double>(vi.x * m);
    p[1] = static_cast<double>(vi.y * m);
    p[2] = static_cast<double>(vi.z * m);
    p[3] = static_cast<double>(m);
  }
  for (int c = 0; c < 4; ++c) {
    const double s = block_sum(p[c], s_red);
    if (threadIdx.x == 0) partials[static_cast<size_t>(b) * 4 + c] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_momentum_apply(float4* __restrict__ v, const double* __restrict__ tot4,
                                                           uint32_t n) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const double inv_m = 1.0 / tot4[3];
  float4 vi = v[i];
  vi.x -= static_cast<float>(tot4[0] * inv_m);
  vi.y -= static_cast<float>(tot4[1] * inv_m);
  vi.z -= static_cast<float>(tot4[2] * inv_m);
  v[i] = vi;
}

// ---------------------------------------------------------------- kinetic energy
// IntegratorNVT::GetKineticEnergy (src/IntegratorNVT.cu:316-356), FP64 sum.
__global__ __launch_bounds__(kBlock) void k_kinetic(const float4* __restrict__ r, const float4* __restrict__ v,
                                                    const float4* __restrict__ f, const rumdx_nh_state* nh,
                                                    float dt, uint32_t on_mask, double* partials, uint32_t n) {
#pragma clang fp contract(off)
  __shared__ double s_red[kBlock / 64];
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = b * kBlock + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const float4 vi = v[i], fi = f[i];
    const double mass = static_cast<double>(1.f / vi.w);
    const float tsim = vi.w * dt;
    const bool apply = ((on_mask >> (__float_as_uint(r[i].w) & 31u)) & 1u) != 0u;
    const double ps = (on_mask != 0u && nh != nullptr) ? nh->ps : 0.0;
    const float factor = apply ? static_cast<float>(0.5f * dt * ps) : 0.f;
    const float plus = 1.f / (1.f + factor);
    const float minus = 1.f - factor;
    const float vfx = plus * (minus * vi.x + tsim * fi.x);
    const float vfy = plus * (minus * vi.y + tsim * fi.y);
    const float vfz = plus * (minus * vi.z + tsim * fi.z);
    const double vb = static_cast<double>(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
    const double vf = static_cast<double>(vfx * vfx + vfy * vfy + vfz * vfz);
    e = mass * ((vb + vf) / 2.0);
  }
  const double s = block_sum(e, s_red);
  if (threadIdx.x == 0) partials[b] = s;
}

// ---------------------------------------------------------------- float4 -> FP64 sums
__global__ __launch_bounds__(kBlock) void k_reduce_f4(const float4* __restrict__ a, uint32_t mask, double* partials,
                                                      uint32_t n) {
  __shared__ double s_red[kBlock / 64];
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = b * kBlock + threadIdx.x;
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i < n) x = a[i];
  const float c4[4] = {x.x, x.y, x.z, x.w};
  for (int c = 0; c < 4; ++c) {
    if (!((mask >> c) & 1u)) continue;
    const double s = block_sum(static_cast<double>(c4[c]), s_red);
    if (threadIdx.x == 0) partials[static_cast<size_t>(b) * 4 + c] = s;
  }
}

template <bool THERMO, bool HILO, bool SKIN>
void launch_nvt(const NvtArgs& a, const Box& bx, hipStream_t st) {
  const uint32_t nb = blocks_for(a.n);
  hipLaunchKernelGGL((k_nvt<THERMO, HILO, SKIN>), dim3(nb), dim3(kBlock), 0, st, a, bx);
  if (THERMO)
    hipLaunchKernelGGL(k_nvt_finish, dim3(1), dim3(kBlock), 0, st, a.partials, nb, a.nh, a.dt, a.omega2, a.dof);
}

}  // namespace
}  // namespace rumdx

using namespace rumdx;

extern "C" int rumdx_version(void) { return RUMDX_VERSION; }

extern "C" const char* rumdx_last_error(void) { return g_last_error.c_str(); }

extern "C" size_t rumdx_workspace_bytes(uint32_t n) {
  return Workspace::kPartials + sizeof(double) * 4 * static_cast<size_t>(blocks_for(n > 0 ? n : 1));
}

extern "C" int rumdx_nvt_step(rumdx_particles* p, rumdx_nh_state* nh_dev, float dt, uint32_t thermostat_on_mask,
                              float omega2, uint32_t dof, const float box[8], rumdx_nblist* nb, void* workspace_dev,
                              void* stream) {
  if (p == nullptr || box == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nvt_step: null argument");
  if (p->r_dev == nullptr || p->v_dev == nullptr || p->f_dev == nullptr || p->im_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_nvt_step: particle arrays not allocated");
  const bool thermo = thermostat_on_mask != 0u;
  if (thermo && (nh_dev == nullptr || workspace_dev == nullptr))
    return fail(RUMDX_E_INVALID, "rumdx_nvt_step: thermostat state and workspace required");
  if (thermo && dof == 0) return fail(RUMDX_E_INVALID, "rumdx_nvt_step: zero degrees of freedom");
  if (nb != nullptr && (nb->flags_dev == nullptr || nb->last_r_dev == nullptr))
    return fail(RUMDX_E_INVALID, "rumdx_nvt_step: neighbour list not allocated");
  if (p->n == 0) return RUMDX_OK;
  NvtArgs a{};
  a.r = reinterpret_cast<float4*>(p->r_dev);
  a.v = reinterpret_cast<float4*>(p->v_dev);
  a.f = reinterpret_cast<const float4*>(p->f_dev);
  a.im = reinterpret_cast<float4*>(p->im_dev);
  a.rlo = reinterpret_cast<float4*>(p->rlo_dev);
  a.nh = nh_dev;
  a.partials = workspace_dev ? Workspace::partials(workspace_dev) : nullptr;
  a.dt = dt;
  a.omega2 = omega2;
  a.on_mask = thermostat_on_mask;
  a.dof = dof;
  a.n = p->n;
  if (nb != nullptr) {
    a.last_r = reinterpret_cast<const float4*>(nb->last_r_dev);
    a.nb_flags = nb->flags_dev;
    a.skin2_over_4 = static_cast<float>(0.25 * static_cast<double>(nb->skin) * static_cast<double>(nb->skin));
    a.prune_r = reinterpret_cast<const float4*>(inner_prune_r(nb));
    a.inner2 = inner_check2(nb);
  }
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const bool hilo = p->rlo_dev != nullptr, skin = nb != nullptr;
  const int sel = (thermo ? 4 : 0) | (hilo ? 2 : 0) | (skin ? 1 : 0);
  switch (sel) {
    case 0: launch_nvt<false, false, false>(a, bx, st); break;
    case 1: launch_nvt<false, false, true>(a, bx, st); break;
    case 2: launch_nvt<false, true, false>(a, bx, st); break;
    case 3: launch_nvt<false, true, true>(a, bx, st); break;
    case 4: launch_nvt<true, false, false>(a, bx, st); break;
    case 5: launch_nvt<true, false, true>(a, bx, st); break;
    case 6: launch_nvt<true, true, false>(a, bx, st); break;
    default: launch_nvt<true, true, true>(a, bx, st); break;
  }
  return check_launch("rumdx_nvt_step");
}

extern "C" int rumdx_zero_momentum(rumdx_particles* p, void* workspace_dev, void* stream) {
  if (p == nullptr || p->v_dev == nullptr || workspace_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_zero_momentum: null argument");
  if (p->n == 0) return RUMDX_OK;
  const hipStream_t st = as_stream(stream);
  const uint32_t nb = blocks_for(p->n);
  double* tot = Workspace::scalars(workspace_dev);
  hipLaunchKernelGGL(k_momentum_sum, dim3(nb), dim3(kBlock), 0, st, reinterpret_cast<const float4*>(p->v_dev),
                     Workspace::partials(workspace_dev), p->n);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, st, Workspace::partials(workspace_dev), nb, 4u, 15u, 1.0,
                     tot);
  hipLaunchKernelGGL(k_momentum_apply, dim3(nb), dim3(kBlock), 0, st, reinterpret_cast<float4*>(p->v_dev), tot, p->n);
  return check_launch("rumdx_zero_momentum");
}

extern "C" int rumdx_kinetic_energy(const rumdx_particles* p, const rumdx_nh_state* nh_dev, float dt,
                                    uint32_t thermostat_on_mask, double* out_dev, void* workspace_dev, void* stream) {
  if (p == nullptr || out_dev == nullptr || workspace_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_kinetic_energy: null argument");
  if (thermostat_on_mask != 0u && nh_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_kinetic_energy: thermostat state required");
  if (p->n == 0) return fail(RUMDX_E_INVALID, "rumdx_kinetic_energy: sample has no particles");
  const uint32_t nb = blocks_for(p->n);
  hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(kBlock), 0, as_stream(stream), reinterpret_cast<const float4*>(p->r_dev),
                     reinterpret_cast<const float4*>(p->v_dev), reinterpret_cast<const float4*>(p->f_dev), nh_dev, dt,
                     thermostat_on_mask, Workspace::partials(workspace_dev), p->n);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, as_stream(stream), Workspace::partials(workspace_dev), nb,
                     1u, 1u, 0.5, out_dev);
  return check_launch("rumdx_kinetic_energy");
}

namespace rumdx {
namespace {

// Block partials of rumdx_reduce_f4_static: four doubles per kBlock elements
// for up to RUMDX_STATIC_REDUCE_MAX_N elements.  Each GPU holds its own copy
// at its own address; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (scratch_on_current_device).
constexpr size_t kStaticPartials = 4ull * (RUMDX_STATIC_REDUCE_MAX_N / kBlock);
static_assert(RUMDX_STATIC_REDUCE_MAX_N % kBlock == 0, "whole blocks");
__device__ __align__(256) double scratch[kStaticPartials];

constexpr int kMaxDevices = 64;
std::atomic<double*> g_scratch_addr[kMaxDevices];
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_on_current_device(const char* entry, double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(RUMDX_E_HIP, std::string(entry) + ": hipGetDevice: " + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(RUMDX_E_INVALID, std::string(entry) + ": device index out of range");
  double* ptr = g_scratch_addr[dev].load(std::memory_order_acquire);
  if (ptr == nullptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ptr = g_scratch_addr[dev].load(std::memory_order_relaxed);
    if (ptr == nullptr) {
      void* addr = nullptr;
      e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || addr == nullptr)
        return fail(RUMDX_E_HIP, std::string(entry) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
      ptr = static_cast<double*>(addr);
      g_scratch_addr[dev].store(ptr, std::memory_order_release);
    }
  }
  *out = ptr;
  return RUMDX_OK;
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_reduce_f4_static(const float* arr4_dev, uint32_t n, uint32_t comp_mask, double* out_dev,
                                      void* stream) {
  const char* entry = "rumdx_reduce_f4_static";
  if (arr4_dev == nullptr || out_dev == nullptr) return fail(RUMDX_E_INVALID, std::string(entry) + ": null argument");
  if (n == 0) return fail(RUMDX_E_INVALID, std::string(entry) + ": empty array");
  if (n > RUMDX_STATIC_REDUCE_MAX_N)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": n = " + std::to_string(n) + " exceeds the static scratch (" +
                                     std::to_string(RUMDX_STATIC_REDUCE_MAX_N) + " elements); use rumdx_reduce_f4");
  double* partials = nullptr;
  const int rc = scratch_on_current_device(entry, &partials);
  if (rc != RUMDX_OK) return rc;
  const uint32_t nb = blocks_for(n);
  hipLaunchKernelGGL(k_reduce_f4, dim3(nb), dim3(kBlock), 0, as_stream(stream), reinterpret_cast<const float4*>(arr4_dev),
                     comp_mask & 15u, partials, n);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, as_stream(stream), partials, nb, 4u, comp_mask & 15u,
                     1.0, out_dev);
  return check_launch(entry);
}

extern "C" int rumdx_reduce_f4(const float* arr4_dev, uint32_t n, uint32_t comp_mask, double* out_dev,
                               void* workspace_dev, void* stream) {
  if (arr4_dev == nullptr || out_dev == nullptr || workspace_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_reduce_f4: null argument");
  if (n == 0) return fail(RUMDX_E_INVALID, "rumdx_reduce_f4: empty array");
  const uint32_t nb = blocks_for(n);
  hipLaunchKernelGGL(k_reduce_f4, dim3(nb), dim3(kBlock), 0, as_stream(stream), reinterpret_cast<const float4*>(arr4_dev),
                     comp_mask & 15u, Workspace::partials(workspace_dev), n);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(kBlock), 0, as_stream(stream), Workspace::partials(workspace_dev), nb,
                     4u, comp_mask & 15u, 1.0, out_dev);
  return check_launch("rumdx_reduce_f4");
}
