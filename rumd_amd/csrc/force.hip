// LJ 12-6 pair force / energy / virial (gfx950): the pair policy of
// Pot_IPL_12_6 / Pot_LJ_12_6 and its entry point rumdx_pair_lj.
//
// Reference: Pot_IPL_12_6::ComputeInteraction (include/rumd/PairPotential.h:161-168),
// fij (include/rumd/PairPotential.h:1198-1244).  The pair arithmetic is
// VALU-bound, so the LDS table holds pre-multiplied constants: per type pair
// {rc^2, A12, A6, s(rc) rc} and {12 A12, 6 A6, -u(rc), -(u(rc) + s(rc) rc^2)},
// the last two being the energy shifts of the shifted-potential and
// shifted-force cut-offs.
#include "force_kernels.h"

// Waves per SIMD of the 12-6 tile kernel (five fit a 2040-particle halo).
#ifndef RUMDX_TILE_WAVES_PER_EU
#define RUMDX_TILE_WAVES_PER_EU 5
#endif

namespace rumdx {
namespace {

template <int CUT>
struct LjPair {
  static constexpr uint32_t kRow = 2;
  static constexpr int kWavesPerEU = RUMDX_TILE_WAVES_PER_EU;
  static constexpr bool kTunedWaves = true;

  __device__ __forceinline__ void load_table(float4* s_tab, const float* __restrict__ params, uint32_t nt) const {
    for (uint32_t k = threadIdx.x; k < nt * nt; k += blockDim.x) {
      const float* q = params + static_cast<size_t>(k) * RUMDX_NUM_PARAM;
      const float rc = q[0], a12 = q[1], a6 = q[2];
      const float src = q[RUMDX_NUM_PARAM - 2], ucut = q[RUMDX_NUM_PARAM - 1];
      s_tab[2 * k] = make_float4(rc * rc, a12, a6, src);
      s_tab[2 * k + 1] = make_float4(12.f * a12, 6.f * a6, -ucut, -fmaf(src, rc, ucut));
    }
  }

  // The LJ term of one pair given its displacement (fij with
  // ComputeInteraction inlined): u = A12 r^-12 + A6 r^-6, s = -u'(r)/r,
  // shifted as the cut-off method asks (SP: u - u(rc); SF: s - s(rc) rc/r and
  // u - u(rc) + s(rc) rc (r - rc)).  Branch-free; pairs at or beyond rc are
  // masked.
  template <bool STR>
  __device__ __forceinline__ void core(float dx, float dy, float dz, float d2, uint32_t tj, const float4* prow,
                                       Acc& a) const {
    const float4 pa = prow[2 * tj];
    const float4 pb = prow[2 * tj + 1];
    // SF needs 1/r anyway: one rsq gives both 1/r and 1/r^2
    float rinv = 0.f, inv2;
    if (CUT == RUMDX_CUTOFF_SF) {
      rinv = __builtin_amdgcn_rsqf(d2);
      inv2 = rinv * rinv;
    } else {
      inv2 = __builtin_amdgcn_rcpf(d2);
    }
    const float inv6 = inv2 * inv2 * inv2;
    float s = (inv6 * inv2) * fmaf(pb.x, inv6, pb.y);
    float u;
    if (CUT == RUMDX_CUTOFF_SP) {
      u = fmaf(inv6, fmaf(pa.y, inv6, pa.z), pb.z);
    } else if (CUT == RUMDX_CUTOFF_SF) {
      u = fmaf(inv6, fmaf(pa.y, inv6, pa.z), pb.w);
      s = fmaf(-pa.w, rinv, s);
      u = fmaf(pa.w, d2 * rinv, u);
    } else {
      u = inv6 * fmaf(pa.y, inv6, pa.z);
    }
    const bool in = d2 < pa.x;
    s = in ? s : 0.f;
    u = in ? u : 0.f;
    accumulate<STR>(dx, dy, dz, d2, s, u, a);
  }
};

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_pair_lj(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                             uint32_t ntypes, const float box[8], int cutoff_method, int initialize,
                             int calc_stresses, void* stream) {
  using namespace rumdx;
  const char* entry = "rumdx_pair_lj";
  const int rc = check_pair_args(entry, p, nb, params_dev, ntypes, box, calc_stresses);
  if (rc != RUMDX_OK) return rc;
  if (p->n == 0) return RUMDX_OK;
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const bool init = initialize != 0, str = calc_stresses != 0;
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS:
      return dispatch_pair(LjPair<RUMDX_CUTOFF_NS>{}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SP:
      return dispatch_pair(LjPair<RUMDX_CUTOFF_SP>{}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SF:
      return dispatch_pair(LjPair<RUMDX_CUTOFF_SF>{}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    default: return fail(RUMDX_E_INVALID, "rumdx_pair_lj: invalid cutoff method");
  }
}

// One NVT step of a sample whose only force term is this 12-6 potential over a
// sorted list: the forces at the positions p->r_dev, and in the same kernel
// each particle's leapfrog / Nose-Hoover update with the skin check, the new
// positions going to r_out_dev; then the thermostat update from the per-tile
// kinetic-energy sums.  Per particle the arithmetic is that of rumdx_pair_lj
// followed by rumdx_nvt_step.
extern "C" int rumdx_pair_lj_nvt(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                                 uint32_t ntypes, const float box[8], int cutoff_method, float* r_out_dev,
                                 rumdx_nh_state* nh_dev, float dt, uint32_t thermostat_on_mask, float omega2,
                                 uint32_t dof, void* workspace_dev, void* stream) {
  using namespace rumdx;
  const char* entry = "rumdx_pair_lj_nvt";
  const int rc = check_pair_args(entry, p, nb, params_dev, ntypes, box, 0);
  if (rc != RUMDX_OK) return rc;
  if (nb == nullptr || nb->nbl16_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: needs a list of the sorted method");
  if (nb->flags_dev == nullptr || nb->last_r_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: neighbour list not allocated");
  if (p->v_dev == nullptr || p->im_dev == nullptr || r_out_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: particle arrays not allocated");
  if (r_out_dev == p->r_dev)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: r_out_dev must not alias the positions the forces read");
  if (cutoff_method != RUMDX_CUTOFF_NS && cutoff_method != RUMDX_CUTOFF_SP && cutoff_method != RUMDX_CUTOFF_SF)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: invalid cutoff method");
  const bool thermo = thermostat_on_mask != 0u;
  if (thermo && (nh_dev == nullptr || workspace_dev == nullptr))
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: thermostat state and workspace required");
  if (thermo && dof == 0) return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: zero degrees of freedom");
  // one kinetic-energy partial per tile in the workspace's partials
  if (thermo && static_cast<size_t>(nb->ntiles) > 4u * static_cast<size_t>(blocks_for(p->n > 0 ? p->n : 1)))
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_nvt: more tiles than the workspace holds partials (4 per 256 particles)");
  if (p->n == 0) return RUMDX_OK;
  NvtFuse fz{};
  fz.r_out = reinterpret_cast<float4*>(r_out_dev);
  fz.v = reinterpret_cast<float4*>(p->v_dev);
  fz.im = reinterpret_cast<float4*>(p->im_dev);
  fz.rlo = reinterpret_cast<float4*>(p->rlo_dev);
  fz.nh = thermo ? nh_dev : nullptr;
  fz.partials = thermo ? Workspace::partials(workspace_dev) : nullptr;
  fz.last_r = reinterpret_cast<const float4*>(nb->last_r_dev);
  fz.nb_flags = nb->flags_dev;
  fz.skin2_over_4 = static_cast<float>(0.25 * static_cast<double>(nb->skin) * static_cast<double>(nb->skin));
  fz.prune_r = reinterpret_cast<const float4*>(inner_prune_r(nb));
  fz.inner2 = inner_check2(nb);
  fz.dt = dt;
  fz.on_mask = thermostat_on_mask;
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const uint32_t lpp = nb->lanes_per_particle;
  const uint32_t lgp = lpp >= 8u ? 3u : lpp >= 4u ? 2u : lpp >= 2u ? 1u : 0u;
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS:
      launch_tiles<LjPair<RUMDX_CUTOFF_NS>, false, true, true>({}, p, nb, params_dev, ntypes, bx, lgp, fz, st);
      break;
    case RUMDX_CUTOFF_SP:
      launch_tiles<LjPair<RUMDX_CUTOFF_SP>, false, true, true>({}, p, nb, params_dev, ntypes, bx, lgp, fz, st);
      break;
    default:
      launch_tiles<LjPair<RUMDX_CUTOFF_SF>, false, true, true>({}, p, nb, params_dev, ntypes, bx, lgp, fz, st);
      break;
  }
  if (thermo)
    hipLaunchKernelGGL(k_nvt_finish, dim3(1), dim3(kBlock), 0, st, fz.partials, nb->ntiles, nh_dev, dt, omega2, dof);
  return check_launch(entry);
}
