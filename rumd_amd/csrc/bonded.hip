// Intramolecular forces (gfx950): harmonic bonds and cos^2 bond angles, and
// the slot index they read particles through.
//
// Reference: BondHarmonic::CalcF and the Harmonic kernel
// (src/BondPotential.cu:70-135, 281-357), AngleCosSq::CalcF and the CosSq
// kernel (src/AnglePotential.cu:69-98, 137-217).
//
// Design.  The reference runs a thread per bond (angle) and scatters the two
// (three) forces with float atomics, so its sums depend on timing.  Here a
// thread owns one particle and gathers every term it takes part in from a
// per-particle list (CSR over original ids): a bond is evaluated from both
// ends, an angle from all three atoms, each end keeping its own share.  No
// atomics, so the result is reproducible bit for bit and a particle's bonded
// sum is written once (stored or added, as initialize asks).  Particles are
// addressed through slot_of (original id -> current slot), which the spatial
// sort changes at every rebuild; rumdx_slot_index refreshes it on the device.
// Displacements are formed in FP64 and rounded, as in the pair kernels.
#include "rumdx_common.h"

namespace rumdx {
namespace {

// Minimum-image displacement a - b in FP64, rounded to FP32 (the reference's
// calculateDistance, include/rumd/SimulationBox.h:111-125).
__device__ __forceinline__ float3 displacement(float4 a, float4 b, const Box& bx) {
  double x = static_cast<double>(a.x) - static_cast<double>(b.x);
  double y = static_cast<double>(a.y) - static_cast<double>(b.y);
  double z = static_cast<double>(a.z) - static_cast<double>(b.z);
  x -= bx.Ld[0] * rint(x * bx.iLd[0]);
  y -= bx.Ld[1] * rint(y * bx.iLd[1]);
  z -= bx.Ld[2] * rint(z * bx.iLd[2]);
  return make_float3(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z));
}

__global__ __launch_bounds__(kBlock) void k_slot_index(const uint32_t* __restrict__ idx,
                                                       uint32_t* __restrict__ slot_of, uint32_t n) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) slot_of[idx[s]] = s;
}

struct BondSums {
  float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f, vir = 0.f;
  float sxx = 0.f, syy = 0.f, szz = 0.f, syz = 0.f, sxz = 0.f, sxy = 0.f;
};

template <bool STR>
__device__ __forceinline__ void store_sums(uint32_t s, const BondSums& a, bool init, float4* __restrict__ f,
                                           float4* __restrict__ w, float4* __restrict__ sts) {
  float4 fo = init ? make_float4(0.f, 0.f, 0.f, 0.f) : f[s];
  fo.x += a.fx;
  fo.y += a.fy;
  fo.z += a.fz;
  fo.w += a.u;
  f[s] = fo;
  float4 wo = init ? make_float4(0.f, 0.f, 0.f, 0.f) : w[s];
  wo.w += a.vir;
  if (STR) {
    wo.y += a.sxz;
    wo.z += a.sxy;
    float4 so = init ? make_float4(0.f, 0.f, 0.f, 0.f) : sts[s];
    so.x += a.sxx;
    so.y += a.syy;
    so.z += a.szz;
    so.w += a.syz;
    sts[s] = so;
  }
  w[s] = wo;
}

// Harmonic bonds, u = k/2 (r - l0)^2 per bond: each end takes F = ft d with
// ft = -k (1 - l0/r), half the bond energy and the virial ft r^2 (the
// reference's per-end bookkeeping).  params: (l0, k) per bond type.
template <bool STR>
__global__ __launch_bounds__(kBlock) void k_bond_harmonic(
    const float4* __restrict__ r, float4* __restrict__ f, float4* __restrict__ w, float4* __restrict__ sts,
    const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ start, const uint32_t* __restrict__ other,
    const uint32_t* __restrict__ btype, const float2* __restrict__ params, uint32_t ntypes, Box bx, uint32_t n,
    int init) {
  const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const uint32_t b0 = start[o], b1 = start[o + 1];
  if (b0 == b1 && !init) return;
  const uint32_t s = slot_of[o];
  const float4 ri = r[s];
  BondSums a;
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t t = btype[b];
    const float2 p = params[t < ntypes ? t : 0u];
    const float3 d = displacement(ri, r[slot_of[other[b]]], bx);
    const float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const float rij = sqrtf(d2);
    const float dl = rij - p.x;
    const float ft = -p.y * (1.0f - p.x / rij);
    a.fx += ft * d.x;
    a.fy += ft * d.y;
    a.fz += ft * d.z;
    a.u += 0.25f * p.y * dl * dl;  // half of k/2 (r - l0)^2
    a.vir += ft * d2;
    if (STR) {
      a.sxx -= ft * d.x * d.x;
      a.syy -= ft * d.y * d.y;
      a.szz -= ft * d.z * d.z;
      a.syz -= ft * d.y * d.z;
      a.sxz -= ft * d.x * d.z;
      a.sxy -= ft * d.x * d.y;
    }
  }
  store_sums<STR>(s, a, init != 0, f, w, sts);
}

// Bond angles, u = k/2 (cos theta - cos theta0)^2 for the angle at b of
// (a, b, c), written with the bond vectors d1 = r_b - r_a, d2 = r_c - r_b
// (so cos(pi - theta) appears): F_a = ff (c1 d1 - d2) / D, F_c = ff (d1 -
// c2 d2) / D, F_b = -F_a - F_c, ff = -k (cos(pi - theta) - cos(pi - theta0)),
// c1 = d1.d2/|d1|^2, c2 = d1.d2/|d2|^2, D = |d1||d2|.  The angle's energy goes
// to atom a, as in the reference; angles add no virial there.  Each list entry
// is (a, b, c, type) in original ids; params: (theta0, k) per angle type.
__global__ __launch_bounds__(kBlock) void k_angle_cossq(const float4* __restrict__ r, float4* __restrict__ f,
                                                        float4* __restrict__ w, float4* __restrict__ sts,
                                                        const uint32_t* __restrict__ slot_of,
                                                        const uint32_t* __restrict__ start,
                                                        const uint4* __restrict__ angles,
                                                        const float2* __restrict__ params, uint32_t ntypes, Box bx,
                                                        uint32_t n, int init, int str) {
  const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const uint32_t a0 = start[o], a1 = start[o + 1];
  if (a0 == a1 && !init) return;
  float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f;
  for (uint32_t q = a0; q < a1; ++q) {
    const uint4 an = angles[q];
    const float2 p = params[an.w < ntypes ? an.w : 0u];
    const float4 ra = r[slot_of[an.x]], rb = r[slot_of[an.y]], rc = r[slot_of[an.z]];
    const float3 d1 = displacement(rb, ra, bx);
    const float3 d2 = displacement(rc, rb, bx);
    const float c11 = d1.x * d1.x + d1.y * d1.y + d1.z * d1.z;
    const float c12 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    const float c22 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    const float icd = 1.f / sqrtf(c11 * c22);
    const float cc = c12 * icd;
    const float ccon = cosf(static_cast<float>(M_PI) - p.x);
    const float ff = -p.y * (cc - ccon);
    const float c1 = c12 / c11, c2 = c12 / c22;
    float gx, gy, gz;  // this atom's share
    if (o == an.x) {
      gx = ff * (c1 * d1.x - d2.x) * icd;
      gy = ff * (c1 * d1.y - d2.y) * icd;
      gz = ff * (c1 * d1.z - d2.z) * icd;
      u += 0.5f * p.y * (cc - ccon) * (cc - ccon);
    } else if (o == an.z) {
      gx = ff * (d1.x - c2 * d2.x) * icd;
      gy = ff * (d1.y - c2 * d2.y) * icd;
      gz = ff * (d1.z - c2 * d2.z) * icd;
    } else {
      gx = -ff * (c1 * d1.x - d2.x) * icd - ff * (d1.x - c2 * d2.x) * icd;
      gy = -ff * (c1 * d1.y - d2.y) * icd - ff * (d1.y - c2 * d2.y) * icd;
      gz = -ff * (c1 * d1.z - d2.z) * icd - ff * (d1.z - c2 * d2.z) * icd;
    }
    fx += gx;
    fy += gy;
    fz += gz;
  }
  const uint32_t s = slot_of[o];
  float4 fo = init ? make_float4(0.f, 0.f, 0.f, 0.f) : f[s];
  fo.x += fx;
  fo.y += fy;
  fo.z += fz;
  fo.w += u;
  f[s] = fo;
  if (init) {
    w[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (str) sts[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

int check_common(const char* entry, const rumdx_particles* p, const void* slot_of, const void* start,
                 const void* list, const void* params, const float* box) {
  if (p == nullptr || slot_of == nullptr || start == nullptr || list == nullptr || params == nullptr ||
      box == nullptr)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": null argument");
  if (p->r_dev == nullptr || p->f_dev == nullptr || p->w_dev == nullptr)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": particle arrays not allocated");
  return RUMDX_OK;
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_slot_index(const rumdx_particles* p, uint32_t* slot_of_dev, void* stream) {
  using namespace rumdx;
  if (p == nullptr || slot_of_dev == nullptr || p->unsorted_index_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_slot_index: null argument");
  if (p->n == 0) return RUMDX_OK;
  hipLaunchKernelGGL(k_slot_index, dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream), p->unsorted_index_dev,
                     slot_of_dev, p->n);
  return check_launch("rumdx_slot_index");
}

extern "C" int rumdx_bond_harmonic(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                   const uint32_t* bond_start_dev, const uint32_t* bond_other_dev,
                                   const uint32_t* bond_type_dev, const float* params_dev, uint32_t ntypes,
                                   const float box[8], int initialize, int calc_stresses, void* stream) {
  using namespace rumdx;
  const int rc = check_common("rumdx_bond_harmonic", p, slot_of_dev, bond_start_dev, bond_other_dev, params_dev, box);
  if (rc != RUMDX_OK) return rc;
  if (ntypes == 0) return fail(RUMDX_E_INVALID, "rumdx_bond_harmonic: no bond types");
  if (calc_stresses && p->sts_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_bond_harmonic: stresses requested but sts is null");
  if (p->n == 0) return RUMDX_OK;
  const Box bx = make_box(box);
  float4* r = reinterpret_cast<float4*>(p->r_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  const float2* par = reinterpret_cast<const float2*>(params_dev);
  if (calc_stresses)
    hipLaunchKernelGGL(k_bond_harmonic<true>, dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream), r, f, w,
                       sts, slot_of_dev, bond_start_dev, bond_other_dev, bond_type_dev, par, ntypes, bx, p->n,
                       initialize);
  else
    hipLaunchKernelGGL(k_bond_harmonic<false>, dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream), r, f, w,
                       sts, slot_of_dev, bond_start_dev, bond_other_dev, bond_type_dev, par, ntypes, bx, p->n,
                       initialize);
  return check_launch("rumdx_bond_harmonic");
}

extern "C" int rumdx_angle_cossq(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                 const uint32_t* angle_start_dev, const uint32_t* angle_atoms_dev,
                                 const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                                 int calc_stresses, void* stream) {
  using namespace rumdx;
  const int rc = check_common("rumdx_angle_cossq", p, slot_of_dev, angle_start_dev, angle_atoms_dev, params_dev, box);
  if (rc != RUMDX_OK) return rc;
  if (ntypes == 0) return fail(RUMDX_E_INVALID, "rumdx_angle_cossq: no angle types");
  if (calc_stresses && p->sts_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_angle_cossq: stresses requested but sts is null");
  if (p->n == 0) return RUMDX_OK;
  hipLaunchKernelGGL(k_angle_cossq, dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream),
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<float4*>(p->f_dev),
                     reinterpret_cast<float4*>(p->w_dev), reinterpret_cast<float4*>(p->sts_dev), slot_of_dev,
                     angle_start_dev,
                     reinterpret_cast<const uint4*>(angle_atoms_dev), reinterpret_cast<const float2*>(params_dev),
                     ntypes, make_box(box), p->n, initialize, calc_stresses);
  return check_launch("rumdx_angle_cossq");
}
