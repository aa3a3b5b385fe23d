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

// The same displacement kept in FP64 (for the dihedral algebra, whose
// differences of products cancel).
__device__ __forceinline__ double3 displacement_f64(float4 a, float4 b, const Box& bx) {
  double x = static_cast<double>(a.x) - static_cast<double>(b.x);
  double y = static_cast<double>(a.y) - static_cast<double>(b.y);
  double z = static_cast<double>(a.z) - static_cast<double>(b.z);
  x -= bx.Ld[0] * rint(x * bx.iLd[0]);
  y -= bx.Ld[1] * rint(y * bx.iLd[1]);
  z -= bx.Ld[2] * rint(z * bx.iLd[2]);
  return make_double3(x, y, z);
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

// Bonds: each end takes F = ft d, half the bond energy and the virial ft r^2
// (the reference's per-end bookkeeping).  params: (p0, k) per bond type.
//   harmonic (FORM 0): u = k/2 (r - l0)^2, ft = -k (1 - l0/r), p0 = l0
//   FENE     (FORM 1): u = -k R0^2/2 ln(1 - r^2/R0^2), ft = -k / (1 - r^2/R0^2),
//                      p0 = R0 (the FENE kernel, src/BondPotential.cu:358-430)
constexpr int kBondHarmonic = 0;
constexpr int kBondFene = 1;

template <bool STR, int FORM>
__global__ __launch_bounds__(kBlock) void k_bond(
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
    if (t >= ntypes) continue;  // a bond type this potential has no parameters for
    const float2 p = params[t];
    if (p.y == 0.f) continue;
    const float3 d = displacement(ri, r[slot_of[other[b]]], bx);
    const float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
    float ft, half_u;
    if (FORM == kBondFene) {
      const float rr = d2 / (p.x * p.x);
      ft = -p.y / (1.0f - rr);
      half_u = -0.25f * p.x * p.x * p.y * logf(1.0f - rr);
    } else {
      const float rij = sqrtf(d2);
      const float dl = rij - p.x;
      ft = -p.y * (1.0f - p.x / rij);
      half_u = 0.25f * p.y * dl * dl;  // half of k/2 (r - l0)^2
    }
    a.fx += ft * d.x;
    a.fy += ft * d.y;
    a.fz += ft * d.z;
    a.u += half_u;
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

// Bond angles for the angle theta at b of (a, b, c), written with the bond
// vectors d1 = r_b - r_a, d2 = r_c - r_b, whose cosine cc is cos(pi - theta):
// F_a = ff (c1 d1 - d2) / D, F_c = ff (d1 - c2 d2) / D, F_b = -F_a - F_c with
// c1 = d1.d2/|d1|^2, c2 = d1.d2/|d2|^2, D = |d1||d2| and ff = -du/dcc.
//   ANGLE_COSSQ (CosSq, src/AnglePotential.cu:137-217):
//     u = k/2 (cc - cos(pi - theta0))^2, ff = -k (cc - cos(pi - theta0));
//   ANGLE_SQ (SquaredAngle, src/AnglePotential.cu:289-373): theta =
//     pi - acos(cc), u = k/2 (theta - theta0)^2, ff = -k (theta - theta0) /
//     (1e-6 + sin theta), the reference's regularisation of the straight
//     angle.
// The angle's energy goes to atom a, as in the reference; angles add no
// virial there.  Each list entry is (a, b, c, type) in original ids; params:
// (theta0, k) per angle type.
constexpr int ANGLE_COSSQ = 0;
constexpr int ANGLE_SQ = 1;

template <int FORM>
__global__ __launch_bounds__(kBlock) void k_angle(const float4* __restrict__ r, float4* __restrict__ f,
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
    float ff, ua;
    if (FORM == ANGLE_SQ) {
      // theta = pi - acos(cc) taken as atan2(|d1 x d2|, -d1.d2) and sin theta
      // as |d1 x d2| / D: the same numbers, without acos's loss of digits
      // near the straight angle, where 1/sin theta amplifies them
      const float cx = d1.y * d2.z - d1.z * d2.y, cy = d1.z * d2.x - d1.x * d2.z, cz = d1.x * d2.y - d1.y * d2.x;
      const float cr = sqrtf(cx * cx + cy * cy + cz * cz);
      const float theta = atan2f(cr, -c12);
      ff = -p.y * (theta - p.x) / (1.0e-6f + cr * icd);
      ua = 0.5f * p.y * (theta - p.x) * (theta - p.x);
    } else {
      const float ccon = cosf(static_cast<float>(M_PI) - p.x);
      ff = -p.y * (cc - ccon);
      ua = 0.5f * p.y * (cc - ccon) * (cc - ccon);
    }
    const float c1 = c12 / c11, c2 = c12 / c22;
    float gx, gy, gz;  // this atom's share
    if (o == an.x) {
      gx = ff * (c1 * d1.x - d2.x) * icd;
      gy = ff * (c1 * d1.y - d2.y) * icd;
      gz = ff * (c1 * d1.z - d2.z) * icd;
      u += ua;
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

// Ryckaert-Bellemans dihedrals, u = sum_k c_k cos^k(psi) for (a, b, c, d)
// with psi the angle between the planes (a, b, c) and (b, c, d), zero for
// trans (the Ryckaert kernel, src/DihedralPotential.cu:150-260): with
// d1 = r_b - r_a, d2 = r_c - r_b, d3 = r_d - r_c, cos psi = cA / sqrt(cB1 cB2)
// (cA = c13 c22 - c12 c23, cB1 = c11 c22 - c12^2, cB2 = c22 c33 - c23^2),
// f_a = f1, f_b = -(1 + c12/c22) f1 + (c23/c22) f2, f_c = (c12/c22) f1 -
// (1 + c23/c22) f2, f_d = f2.  Nearly collinear bonds contribute nothing.  The
// energy goes to atom a.  The algebra runs in FP64: its differences of
// products (cA, cB1, cB2) cancel for nearly planar or nearly collinear
// geometries, which in FP32 costs several digits.  Each list entry is 5 u32:
// a, b, c, d (original ids) and the type; params: c_0..c_5 per type.
// DIH_PERIODIC (PeriodicDih, src/DihedralPotential.cu:337-458): psi =
// acos(cos psi) in [0, pi], params (psi0, k1, k2, k3, 0, 0), u = 1/2 (k1 (1 -
// cos x) + k2 (1 - cos 2x) + k3 (1 - cos 3x)) with x = psi - psi0, and the
// reference's force: -du/dpsi in the place where the Ryckaert form has
// -du/dcos psi, which is -sin psi times the gradient force.
constexpr int DIH_RYCKAERT = 0;
constexpr int DIH_PERIODIC = 1;

template <int FORM>
__global__ __launch_bounds__(kBlock) void k_dihedral(
    const float4* __restrict__ r, float4* __restrict__ f, float4* __restrict__ w, float4* __restrict__ sts,
    const uint32_t* __restrict__ slot_of, const uint32_t* __restrict__ start, const uint32_t* __restrict__ dih,
    const float* __restrict__ params, uint32_t ntypes, Box bx, uint32_t n, int init, int str) {
  const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const uint32_t q0 = start[o], q1 = start[o + 1];
  if (q0 == q1 && !init) return;
  double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
  for (uint32_t q = q0; q < q1; ++q) {
    const uint32_t* e = dih + 5ull * q;
    const uint32_t ia = e[0], ib = e[1], ic = e[2], id = e[3];
    const float* pf = params + 6u * (e[4] < ntypes ? e[4] : 0u);
    double p[6];
    for (int k = 0; k < 6; ++k) p[k] = pf[k];
    const float4 ra = r[slot_of[ia]], rb = r[slot_of[ib]], rc = r[slot_of[ic]], rd = r[slot_of[id]];
    const double3 d1 = displacement_f64(rb, ra, bx);
    const double3 d2 = displacement_f64(rc, rb, bx);
    const double3 d3 = displacement_f64(rd, rc, bx);
    const double c11 = d1.x * d1.x + d1.y * d1.y + d1.z * d1.z;
    const double c12 = d1.x * d2.x + d1.y * d2.y + d1.z * d2.z;
    const double c13 = d1.x * d3.x + d1.y * d3.y + d1.z * d3.z;
    const double c22 = d2.x * d2.x + d2.y * d2.y + d2.z * d2.z;
    const double c23 = d2.x * d3.x + d2.y * d3.y + d2.z * d3.z;
    const double c33 = d3.x * d3.x + d3.y * d3.y + d3.z * d3.z;
    const double cA = c13 * c22 - c12 * c23;
    const double cB1 = c11 * c22 - c12 * c12;
    const double cB2 = c22 * c33 - c23 * c23;
    constexpr double kCollinear = 1e-3;  // DIHEDRAL_EPS (include/rumd/DihedralPotential.h:9)
    if (fabs(c23 * c23 / (c22 * c33) - 1.0) < kCollinear || fabs(c12 * c12 / (c11 * c22) - 1.0) < kCollinear)
      continue;
    const double cD = sqrt(cB1 * cB2);
    const double cc = cA / cD;
    double ff, ud;
    if (FORM == DIH_PERIODIC) {
      const double ceq = cos(p[0]), seq = sin(p[0]), ss = sqrt(fabs(1.0 - cc * cc));
      const double cx = cc * ceq + ss * seq, sx = ss * ceq - cc * seq;  // cos and sin of psi - psi0
      ud = 0.5 * (p[1] * (1.0 - cx) + p[2] * (1.0 - cx * cx + sx * sx) + p[3] * (1.0 + 3.0 * cx - 4.0 * cx * cx * cx));
      ff = (-0.5 * p[1] * sx - 2.0 * p[2] * sx * cx - 1.5 * p[3] * (3.0 * sx - 4.0 * sx * sx * sx)) * c22;
    } else {
      ud = p[0] + (p[1] + (p[2] + (p[3] + (p[4] + p[5] * cc) * cc) * cc) * cc) * cc;
      ff = -(p[1] + (2.0 * p[2] + (3.0 * p[3] + (4.0 * p[4] + 5.0 * p[5] * cc) * cc) * cc) * cc) * c22;
    }
    const double t1 = cA, t2 = c11 * c23 - c12 * c13, t3 = -cB1;
    const double t4 = cB2, t5 = c13 * c23 - c12 * c33, t6 = -cA;
    const double cr1 = c12 / c22, cr2 = c23 / c22;
    const double g1 = ff / (cD * cB1), g2 = ff / (cD * cB2);
    const double3 f1 = make_double3(g1 * (t1 * d1.x + t2 * d2.x + t3 * d3.x), g1 * (t1 * d1.y + t2 * d2.y + t3 * d3.y),
                                    g1 * (t1 * d1.z + t2 * d2.z + t3 * d3.z));
    const double3 f2 = make_double3(g2 * (t4 * d1.x + t5 * d2.x + t6 * d3.x), g2 * (t4 * d1.y + t5 * d2.y + t6 * d3.y),
                                    g2 * (t4 * d1.z + t5 * d2.z + t6 * d3.z));
    double ka = 0.0, kb = 0.0;  // own force = ka f1 + kb f2
    if (o == ia) {
      ka = 1.0;
      u += ud;
    } else if (o == ib) {
      ka = -(1.0 + cr1);
      kb = cr2;
    } else if (o == ic) {
      ka = cr1;
      kb = -(1.0 + cr2);
    } else {
      kb = 1.0;
    }
    fx += ka * f1.x + kb * f2.x;
    fy += ka * f1.y + kb * f2.y;
    fz += ka * f1.z + kb * f2.z;
  }
  const uint32_t s = slot_of[o];
  float4 fo = init ? make_float4(0.f, 0.f, 0.f, 0.f) : f[s];
  fo.x += static_cast<float>(fx);
  fo.y += static_cast<float>(fy);
  fo.z += static_cast<float>(fz);
  fo.w += static_cast<float>(u);
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

namespace rumdx {
namespace {

template <int FORM>
int launch_bonds(const char* entry, const rumdx_particles* p, const uint32_t* slot_of_dev,
                 const uint32_t* bond_start_dev, const uint32_t* bond_other_dev, const uint32_t* bond_type_dev,
                 const float* params_dev, uint32_t ntypes, const float box[8], int initialize, int calc_stresses,
                 void* stream) {
  const std::string e(entry);
  const int rc = check_common(entry, p, slot_of_dev, bond_start_dev, bond_other_dev, params_dev, box);
  if (rc != RUMDX_OK) return rc;
  if (bond_type_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": null argument");
  if (ntypes == 0) return fail(RUMDX_E_INVALID, e + ": no bond types");
  if (calc_stresses && p->sts_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": stresses requested but sts is null");
  if (p->n == 0) return RUMDX_OK;
  const Box bx = make_box(box);
  float4* r = reinterpret_cast<float4*>(p->r_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  const float2* par = reinterpret_cast<const float2*>(params_dev);
  if (calc_stresses)
    hipLaunchKernelGGL((k_bond<true, FORM>), dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream), r, f, w, sts,
                       slot_of_dev, bond_start_dev, bond_other_dev, bond_type_dev, par, ntypes, bx, p->n, initialize);
  else
    hipLaunchKernelGGL((k_bond<false, FORM>), dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream), r, f, w, sts,
                       slot_of_dev, bond_start_dev, bond_other_dev, bond_type_dev, par, ntypes, bx, p->n, initialize);
  return check_launch(entry);
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_bond_harmonic(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                   const uint32_t* bond_start_dev, const uint32_t* bond_other_dev,
                                   const uint32_t* bond_type_dev, const float* params_dev, uint32_t ntypes,
                                   const float box[8], int initialize, int calc_stresses, void* stream) {
  return rumdx::launch_bonds<rumdx::kBondHarmonic>("rumdx_bond_harmonic", p, slot_of_dev, bond_start_dev,
                                                   bond_other_dev, bond_type_dev, params_dev, ntypes, box, initialize,
                                                   calc_stresses, stream);
}

extern "C" int rumdx_bond_fene(const rumdx_particles* p, const uint32_t* slot_of_dev, const uint32_t* bond_start_dev,
                               const uint32_t* bond_other_dev, const uint32_t* bond_type_dev, const float* params_dev,
                               uint32_t ntypes, const float box[8], int initialize, int calc_stresses, void* stream) {
  return rumdx::launch_bonds<rumdx::kBondFene>("rumdx_bond_fene", p, slot_of_dev, bond_start_dev, bond_other_dev,
                                               bond_type_dev, params_dev, ntypes, box, initialize, calc_stresses,
                                               stream);
}

namespace {
template <int FORM>
int launch_dihedrals(const char* entry, const rumdx_particles* p, const uint32_t* slot_of_dev,
                     const uint32_t* dihedral_start_dev, const uint32_t* dihedral_atoms_dev, const float* params_dev,
                     uint32_t ntypes, const float box[8], int initialize, int calc_stresses, void* stream) {
  using namespace rumdx;
  const std::string e(entry);
  const int rc = check_common(entry, p, slot_of_dev, dihedral_start_dev, dihedral_atoms_dev, params_dev, box);
  if (rc != RUMDX_OK) return rc;
  if (ntypes == 0) return fail(RUMDX_E_INVALID, e + ": no dihedral types");
  if (calc_stresses && p->sts_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": stresses requested but sts is null");
  if (p->n == 0) return RUMDX_OK;
  hipLaunchKernelGGL((k_dihedral<FORM>), dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream),
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<float4*>(p->f_dev),
                     reinterpret_cast<float4*>(p->w_dev), reinterpret_cast<float4*>(p->sts_dev), slot_of_dev,
                     dihedral_start_dev, dihedral_atoms_dev, params_dev, ntypes, make_box(box), p->n, initialize,
                     calc_stresses);
  return check_launch(entry);
}
}  // namespace

extern "C" int rumdx_dihedral_ryckaert(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                       const uint32_t* dihedral_start_dev, const uint32_t* dihedral_atoms_dev,
                                       const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                                       int calc_stresses, void* stream) {
  return launch_dihedrals<rumdx::DIH_RYCKAERT>("rumdx_dihedral_ryckaert", p, slot_of_dev, dihedral_start_dev,
                                              dihedral_atoms_dev, params_dev, ntypes, box, initialize, calc_stresses,
                                              stream);
}

extern "C" int rumdx_dihedral_periodic(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                       const uint32_t* dihedral_start_dev, const uint32_t* dihedral_atoms_dev,
                                       const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                                       int calc_stresses, void* stream) {
  return launch_dihedrals<rumdx::DIH_PERIODIC>("rumdx_dihedral_periodic", p, slot_of_dev, dihedral_start_dev,
                                              dihedral_atoms_dev, params_dev, ntypes, box, initialize, calc_stresses,
                                              stream);
}

namespace {
template <int FORM>
int launch_angles(const char* entry, const rumdx_particles* p, const uint32_t* slot_of_dev,
                  const uint32_t* angle_start_dev, const uint32_t* angle_atoms_dev, const float* params_dev,
                  uint32_t ntypes, const float box[8], int initialize, int calc_stresses, void* stream) {
  using namespace rumdx;
  const std::string e(entry);
  const int rc = check_common(entry, p, slot_of_dev, angle_start_dev, angle_atoms_dev, params_dev, box);
  if (rc != RUMDX_OK) return rc;
  if (ntypes == 0) return fail(RUMDX_E_INVALID, e + ": no angle types");
  if (calc_stresses && p->sts_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": stresses requested but sts is null");
  if (p->n == 0) return RUMDX_OK;
  hipLaunchKernelGGL((k_angle<FORM>), dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream),
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<float4*>(p->f_dev),
                     reinterpret_cast<float4*>(p->w_dev), reinterpret_cast<float4*>(p->sts_dev), slot_of_dev,
                     angle_start_dev,
                     reinterpret_cast<const uint4*>(angle_atoms_dev), reinterpret_cast<const float2*>(params_dev),
                     ntypes, make_box(box), p->n, initialize, calc_stresses);
  return check_launch(entry);
}
}  // namespace

extern "C" int rumdx_angle_cossq(const rumdx_particles* p, const uint32_t* slot_of_dev,
                                 const uint32_t* angle_start_dev, const uint32_t* angle_atoms_dev,
                                 const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                                 int calc_stresses, void* stream) {
  return launch_angles<rumdx::ANGLE_COSSQ>("rumdx_angle_cossq", p, slot_of_dev, angle_start_dev, angle_atoms_dev,
                                           params_dev, ntypes, box, initialize, calc_stresses, stream);
}

extern "C" int rumdx_angle_sq(const rumdx_particles* p, const uint32_t* slot_of_dev, const uint32_t* angle_start_dev,
                              const uint32_t* angle_atoms_dev, const float* params_dev, uint32_t ntypes,
                              const float box[8], int initialize, int calc_stresses, void* stream) {
  return launch_angles<rumdx::ANGLE_SQ>("rumdx_angle_sq", p, slot_of_dev, angle_start_dev, angle_atoms_dev,
                                        params_dev, ntypes, box, initialize, calc_stresses, stream);
}
