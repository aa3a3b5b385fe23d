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
