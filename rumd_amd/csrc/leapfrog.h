// The leapfrog / Nose-Hoover update of one particle, shared by the
// single-sample and the batched integrators.
#pragma once

#include "rumdx_common.h"

namespace rumdx {

// One particle of integrateNVTAlgorithm, in source order without FMA
// contraction (reproducible bit for bit on the host).  With HILO the position
// is advanced and wrapped in FP64 from the hi/lo pair and stored back as a pair.
// The image shift of this step goes to `shift` (all zero unless the particle
// crossed the box), so the caller touches the image counts only then.
template <bool HILO>
__device__ __forceinline__ float nvt_particle(float4& r, float4& v, const float4 f, float3& shift, float4& lo,
                                              float ps, float dt, const Box& bx) {
#pragma clang fp contract(off)
  const float factor = 0.5f * ps * dt;
  const float plus = 1.f / (1.f + factor);
  const float minus = 1.f - factor;
  v.x = plus * (minus * v.x + v.w * f.x * dt);
  v.y = plus * (minus * v.y + v.w * f.y * dt);
  v.z = plus * (minus * v.z + v.w * f.z * dt);
  float ix, iy, iz;
  if (!HILO) {
    r.x += v.x * dt;
    r.y += v.y * dt;
    r.z += v.z * dt;
    ix = rintf(r.x * bx.iL[0]);
    iy = rintf(r.y * bx.iL[1]);
    iz = rintf(r.z * bx.iL[2]);
    r.x -= bx.L[0] * ix;
    r.y -= bx.L[1] * iy;
    r.z -= bx.L[2] * iz;
  } else {
    const double dtd = dt;
    double x = static_cast<double>(r.x) + static_cast<double>(lo.x) + static_cast<double>(v.x) * dtd;
    double y = static_cast<double>(r.y) + static_cast<double>(lo.y) + static_cast<double>(v.y) * dtd;
    double z = static_cast<double>(r.z) + static_cast<double>(lo.z) + static_cast<double>(v.z) * dtd;
    const double kx = rint(x * bx.iLd[0]), ky = rint(y * bx.iLd[1]), kz = rint(z * bx.iLd[2]);
    x -= bx.Ld[0] * kx;
    y -= bx.Ld[1] * ky;
    z -= bx.Ld[2] * kz;
    r.x = static_cast<float>(x);
    r.y = static_cast<float>(y);
    r.z = static_cast<float>(z);
    lo.x = static_cast<float>(x - static_cast<double>(r.x));
    lo.y = static_cast<float>(y - static_cast<double>(r.y));
    lo.z = static_cast<float>(z - static_cast<double>(r.z));
    ix = static_cast<float>(kx);
    iy = static_cast<float>(ky);
    iz = static_cast<float>(kz);
  }
  shift = make_float3(ix, iy, iz);
  return (v.x * v.x + v.y * v.y + v.z * v.z) / v.w;
}

// Adds a wrap's image shift to the particle's image counts (RUMD's image
// array); most steps move no particle across the box, and then nothing is read.
__device__ __forceinline__ void add_image_shift(float4* __restrict__ im, uint32_t i, float3 shift) {
  if (shift.x != 0.f || shift.y != 0.f || shift.z != 0.f) {
    float4 m = im[i];
    m.x += shift.x;
    m.y += shift.y;
    m.z += shift.z;
    im[i] = m;
  }
}

// The NVT step fused into a force kernel (rumdx_pair_lj_nvt): a tile's
// particles are advanced as soon as the tile's forces are complete.  Positions
// go to a second array (other workgroups still read the old ones);
// velocities, low position parts and image counts are the particle's own and
// are updated in place.
struct NvtFuse {
  float4* r_out;
  float4* v;
  float4* im;
  float4* rlo;            // null: plain FP32 positions
  const rumdx_nh_state* nh;  // null: thermostat off
  double* partials;       // one kinetic-energy partial per tile
  const float4* last_r;   // skin check against the last build
  uint32_t* nb_flags;
  float skin2_over_4;
  const float4* prune_r;  // inner-list check when non-null
  float inner2;
  float dt;
  uint32_t on_mask;
};

// The particle state the fused step reads besides the force; a kernel may
// load it early (nvt_fused_load) to overlap the reads with work.
struct NvtIn {
  float4 r, v, lo, last_r;
};

__device__ __forceinline__ NvtIn nvt_fused_load(uint32_t i, const float4* __restrict__ r, const NvtFuse& z) {
  NvtIn in;
  in.r = r[i];
  in.v = z.v[i];
  in.lo = z.rlo != nullptr ? z.rlo[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  in.last_r = z.last_r[i];
  return in;
}

// One particle of the fused step: the update of k_nvt for slot i from its
// force fi and state `in`; returns m v^2 when the particle is thermostatted,
// else 0.
__device__ __forceinline__ float nvt_fused(uint32_t i, float4 fi, NvtIn in, const NvtFuse& z, float ps,
                                           const Box& bx) {
  float4 ri = in.r;
  const bool apply = z.nh != nullptr && ((z.on_mask >> (__float_as_uint(ri.w) & 31u)) & 1u);
  float3 shift;
  float ke;
  if (z.rlo != nullptr) {
    ke = nvt_particle<true>(ri, in.v, fi, shift, in.lo, apply ? ps : 0.f, z.dt, bx);
    z.rlo[i] = in.lo;
  } else {
    ke = nvt_particle<false>(ri, in.v, fi, shift, in.lo, apply ? ps : 0.f, z.dt, bx);
  }
  z.r_out[i] = ri;
  z.v[i] = in.v;
  add_image_shift(z.im, i, shift);
  if (min_image_d2_f32(ri, in.last_r, bx) > z.skin2_over_4) z.nb_flags[RUMDX_FLAG_REBUILD] = 1u;
  if (z.prune_r != nullptr && min_image_d2_f32(ri, z.prune_r[i], bx) > z.inner2) z.nb_flags[RUMDX_FLAG_PRUNE] = 1u;
  return apply ? ke : 0.f;
}

namespace {

// UpdateThermostatState in FP64: fixed-order sum of the block partials of
// the integrator, then Ps += dt omega^2 min(KE2 / (dof T) - 1, 10) and the T
// ramp.  One workgroup; the kernel boundary makes the partials visible.
__global__ __launch_bounds__(kBlock) void k_nvt_finish(const double* __restrict__ partials, uint32_t np,
                                                       rumdx_nh_state* nh, float dt, float omega2, uint32_t dof) {
  __shared__ double s_red[kBlock / 64];
  const double tot = sum_partials(partials, np, 1, s_red);
  if (threadIdx.x == 0) {
    double dev = tot / (static_cast<double>(dof) * nh->target_T) - 1.0;
    if (dev > 10.0) dev = 10.0;  // runaway clamp (src/IntegratorNVT.cu:432-434)
    nh->ps += static_cast<double>(dt) * static_cast<double>(omega2) * dev;
    nh->ke2_sum = tot;
    nh->target_T += nh->dT_per_step;
  }
}

}  // namespace
}  // namespace rumdx
