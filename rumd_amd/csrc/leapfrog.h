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

}  // namespace rumdx
