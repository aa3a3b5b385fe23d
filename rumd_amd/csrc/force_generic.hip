// Pair force / energy / virial of the potential forms in pot_forms.h
// (gfx950): their pair policy and the entry point rumdx_pair_form.
//
// Reference: fij (include/rumd/PairPotential.h:1198-1244) around each
// class's ComputeInteraction; the reference generates one CalcF kernel per
// potential from a common skeleton (src/Generate_PP_FunctionBodies.py:3-127).
// Here the skeleton is the shared kernels of force_kernels.h and the form is a
// kernel argument: a wave-uniform switch, so one set of kernels serves every
// form.  Per type pair the LDS table holds three float4:
// {rc^2, p1, p2, p3}, {p4, p5, p6, p7}, {s(rc) rc, shift, p8, p9}, with the
// energy shift u(rc) for SP and s(rc) rc^2 + u(rc) for SF.
#include "force_kernels.h"
#include "pot_forms.h"

namespace rumdx {
namespace {

template <int CUT>
struct FormPair {
  int form;
  static constexpr uint32_t kRow = 3;
  static constexpr int kWavesPerEU = 3;
  static constexpr bool kTunedWaves = false;

  __device__ __forceinline__ void load_table(float4* s_tab, const float* __restrict__ params, uint32_t nt) const {
    for (uint32_t k = threadIdx.x; k < nt * nt; k += blockDim.x) {
      const float* q = params + static_cast<size_t>(k) * RUMDX_NUM_PARAM;
      const float rc = q[0], src = q[RUMDX_NUM_PARAM - 2], ucut = q[RUMDX_NUM_PARAM - 1];
      s_tab[3 * k] = make_float4(rc * rc, q[1], q[2], q[3]);
      s_tab[3 * k + 1] = make_float4(q[4], q[5], q[6], q[7]);
      s_tab[3 * k + 2] = make_float4(src, CUT == RUMDX_CUTOFF_SF ? fmaf(src, rc, ucut) : ucut, q[8], q[9]);
    }
  }

  // fij: the form's s and u inside rc, shifted as the cut-off method asks
  // (SP: u - u(rc); SF: s - s(rc) rc / r, u + s(rc) rc (r - rc) - u(rc),
  // the last two terms as s(rc) rc r - (s(rc) rc^2 + u(rc))).
  template <bool STR>
  __device__ __forceinline__ void core(float dx, float dy, float dz, float d2, uint32_t tj, const float4* prow,
                                       Acc& a) const {
    const float4 c0 = prow[3 * tj];
    float s = 0.f, u = 0.f;
    if (d2 < c0.x) {
      const float4 c1 = prow[3 * tj + 1];
      const float4 c2 = prow[3 * tj + 2];
      const float p[10] = {0.f, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.z, c2.w};
      s = pot_form<float>(form, d2, p, u);
      if (CUT == RUMDX_CUTOFF_SP) {
        u -= c2.y;
      } else if (CUT == RUMDX_CUTOFF_SF) {
        const float r = sqrtf(d2);
        s -= c2.x / r;
        u += fmaf(c2.x, r, -c2.y);
      }
    }
    accumulate<STR>(dx, dy, dz, d2, s, u, a);
  }
};

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_pair_form(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                               uint32_t ntypes, const float box[8], int form, int cutoff_method, int initialize,
                               int calc_stresses, void* stream) {
  using namespace rumdx;
  const char* entry = "rumdx_pair_form";
  const int rc = check_pair_args(entry, p, nb, params_dev, ntypes, box, calc_stresses);
  if (rc != RUMDX_OK) return rc;
  if (!pot_form_known(form)) return fail(RUMDX_E_INVALID, "rumdx_pair_form: unknown potential form");
  if (p->n == 0) return RUMDX_OK;
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const bool init = initialize != 0, str = calc_stresses != 0;
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS:
      return dispatch_pair(FormPair<RUMDX_CUTOFF_NS>{form}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SP:
      return dispatch_pair(FormPair<RUMDX_CUTOFF_SP>{form}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SF:
      return dispatch_pair(FormPair<RUMDX_CUTOFF_SF>{form}, entry, p, nb, params_dev, ntypes, bx, init, str, st);
    default: return fail(RUMDX_E_INVALID, "rumdx_pair_form: invalid cutoff method");
  }
}
