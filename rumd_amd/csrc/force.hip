// LJ 12-6 pair force / energy / virial kernels (gfx950).
//
// Reference: Calcf_NBL_tp_equal_one / Calcf_NBL_tp (src/PairPotential.cu:330-515),
// fij (include/rumd/PairPotential.h:1198-1244), Pot_IPL_12_6::ComputeInteraction
// (include/rumd/PairPotential.h:161-168).
//
// Design: one lane per particle, streaming the column-major neighbour list
// (a wave reads 64 consecutive u32 per list row), neighbour positions gathered
// through L2 (particles are cell-sorted in Morton order and blocks are mapped
// XCD-contiguous).  Pair arithmetic is FP32; the minimum-image displacement is
// formed in FP64 so pairs that cross the periodic boundary of a large box keep
// full FP32 accuracy.  The per-particle sums stay in registers (one lane owns a
// particle, no cross-lane traffic).
#include "rumdx_common.h"

namespace rumdx {
namespace {

// Per type pair, two float4 in LDS: {rc^2, A12, A6, s(rc) rc} and {u(rc), rc, -, -}.
struct PairTable {
  const float4* p;
  uint32_t nt;
};

struct Acc {
  float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f, vir = 0.f;
  float sxx = 0.f, syy = 0.f, szz = 0.f, syz = 0.f, sxz = 0.f, sxy = 0.f;
};

// Cell data used to rescan the stencil for a particle whose list overflowed.
struct CellFallback {
  const uint32_t* cell_rank;
  const uint32_t* cell_start;
  const uint32_t* slot_cell;
  int nc0, nc1, nc2;
};

__device__ __forceinline__ void load_table(float4* s_tab, const float* __restrict__ params, uint32_t nt) {
  for (uint32_t k = threadIdx.x; k < nt * nt; k += blockDim.x) {
    const float* q = params + static_cast<size_t>(k) * RUMDX_NUM_PARAM;
    const float rc = q[0];
    s_tab[2 * k] = make_float4(rc * rc, q[1], q[2], q[RUMDX_NUM_PARAM - 2]);
    s_tab[2 * k + 1] = make_float4(q[RUMDX_NUM_PARAM - 1], rc, 0.f, 0.f);
  }
}

// One pair term (fij with ComputeInteraction inlined).  The cut-off test and
// the shift follow the reference; evaluation is branch-free and masked.
template <int CUT, bool STR>
__device__ __forceinline__ void pair_term(double xi, double yi, double zi, float4 rj, const float4* prow,
                                          const Box& bx, Acc& a) {
  double x = xi - static_cast<double>(rj.x);
  double y = yi - static_cast<double>(rj.y);
  double z = zi - static_cast<double>(rj.z);
  x -= bx.Ld[0] * rint(x * bx.iLd[0]);
  y -= bx.Ld[1] * rint(y * bx.iLd[1]);
  z -= bx.Ld[2] * rint(z * bx.iLd[2]);
  const float dx = static_cast<float>(x), dy = static_cast<float>(y), dz = static_cast<float>(z);
  const float d2 = static_cast<float>(x * x + y * y + z * z);

  const uint32_t tj = __float_as_uint(rj.w);
  const float4 pa = prow[2 * tj];
  const float4 pb = prow[2 * tj + 1];

  const float inv2 = __builtin_amdgcn_rcpf(d2);
  const float inv6 = inv2 * inv2 * inv2;
  const float t = pa.y * inv6;
  float s = inv6 * inv2 * (12.f * t + 6.f * pa.z);
  float u = inv6 * (t + pa.z);
  if (CUT == RUMDX_CUTOFF_SP) {
    u -= pb.x;
  } else if (CUT == RUMDX_CUTOFF_SF) {
    const float rinv = __builtin_amdgcn_rsqf(d2);
    const float rr = d2 * rinv;
    s -= pa.w * rinv;
    u += pa.w * (rr - pb.y) - pb.x;
  }
  const bool in = d2 < pa.x;
  s = in ? s : 0.f;
  u = in ? u : 0.f;
  a.fx += dx * s;
  a.fy += dy * s;
  a.fz += dz * s;
  a.u += u;
  a.vir += d2 * s;
  if (STR) {
    a.sxx -= dx * dx * s;
    a.syy -= dy * dy * s;
    a.szz -= dz * dz * s;
    a.syz -= dy * dz * s;
    a.sxz -= dx * dz * s;
    a.sxy -= dx * dy * s;
  }
}

template <bool STR, bool INIT>
__device__ __forceinline__ void store_result(uint32_t i, Acc& a, float4* __restrict__ f, float4* __restrict__ w,
                                             float4* __restrict__ sts) {
  a.u *= 0.5f;  // each pair energy was counted from both ends
  if (INIT) {
    f[i] = make_float4(a.fx, a.fy, a.fz, a.u);
    w[i] = make_float4(0.f, STR ? a.sxz : 0.f, STR ? a.sxy : 0.f, a.vir);
    if (STR) sts[i] = make_float4(a.sxx, a.syy, a.szz, a.syz);
  } else {
    float4 fo = f[i];
    fo.x += a.fx;
    fo.y += a.fy;
    fo.z += a.fz;
    fo.w += a.u;
    f[i] = fo;
    float4 wo = w[i];
    wo.w += a.vir;
    if (STR) {
      wo.y += a.sxz;
      wo.z += a.sxy;
      float4 so = sts[i];
      so.x += a.sxx;
      so.y += a.syy;
      so.z += a.szz;
      so.w += a.syz;
      sts[i] = so;
    }
    w[i] = wo;
  }
}

// Neighbour-list force kernel.  fb.cell_rank == nullptr means the list was
// built by the N2 method, so an overflowed particle falls back to all pairs.
template <int CUT, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_lj_nbl(const float4* __restrict__ r, float4* __restrict__ f,
                                                   float4* __restrict__ w, float4* __restrict__ sts,
                                                   const uint32_t* __restrict__ nbl,
                                                   const uint32_t* __restrict__ num_nbrs,
                                                   const float* __restrict__ params, uint32_t nt, Box bx,
                                                   uint32_t n, uint32_t nvp, uint32_t max_nbrs, CellFallback fb) {
  extern __shared__ float4 s_tab[];
  load_table(s_tab, params, nt);
  __syncthreads();

  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = b * kBlock + threadIdx.x;
  if (i >= n) return;

  const float4 ri = r[i];
  const float4* prow = s_tab + 2 * __float_as_uint(ri.w) * nt;
  const double xi = ri.x, yi = ri.y, zi = ri.z;
  Acc a;
  const uint32_t nn = num_nbrs[i];
  if (nn < max_nbrs) {
    const uint32_t* col = nbl + i;
    uint32_t k = 0;
    for (; k + 4 <= nn; k += 4) {
      const uint32_t j0 = __builtin_nontemporal_load(col);
      const uint32_t j1 = __builtin_nontemporal_load(col + nvp);
      const uint32_t j2 = __builtin_nontemporal_load(col + 2 * static_cast<size_t>(nvp));
      const uint32_t j3 = __builtin_nontemporal_load(col + 3 * static_cast<size_t>(nvp));
      col += 4 * static_cast<size_t>(nvp);
      const float4 r0 = r[j0], r1 = r[j1], r2 = r[j2], r3 = r[j3];
      pair_term<CUT, STR>(xi, yi, zi, r0, prow, bx, a);
      pair_term<CUT, STR>(xi, yi, zi, r1, prow, bx, a);
      pair_term<CUT, STR>(xi, yi, zi, r2, prow, bx, a);
      pair_term<CUT, STR>(xi, yi, zi, r3, prow, bx, a);
    }
    for (; k < nn; ++k) {
      const uint32_t j = __builtin_nontemporal_load(col);
      col += nvp;
      pair_term<CUT, STR>(xi, yi, zi, r[j], prow, bx, a);
    }
  } else if (fb.cell_rank != nullptr) {
    // Overflowed list: rescan the 5x5x5 stencil around the cell the particle
    // was binned in at the last build.  Every pair now inside rc was inside
    // rc + skin then, so it is in that stencil.
    const uint32_t lin = fb.slot_cell[i];
    const int cx = static_cast<int>(lin % fb.nc0);
    const int cy = static_cast<int>((lin / fb.nc0) % fb.nc1);
    const int cz = static_cast<int>(lin / (static_cast<uint32_t>(fb.nc0) * fb.nc1));
    for (int dz = -2; dz <= 2; ++dz) {
      const int oz = (cz + dz + fb.nc2) % fb.nc2;
      for (int dy = -2; dy <= 2; ++dy) {
        const int oy = (cy + dy + fb.nc1) % fb.nc1;
        for (int dx = -2; dx <= 2; ++dx) {
          const int ox = (cx + dx + fb.nc0) % fb.nc0;
          const uint32_t rank = fb.cell_rank[ox + oy * fb.nc0 + oz * fb.nc0 * fb.nc1];
          const uint32_t s = fb.cell_start[rank], e = fb.cell_start[rank + 1];
          for (uint32_t j = s; j < e; ++j)
            if (j != i) pair_term<CUT, STR>(xi, yi, zi, r[j], prow, bx, a);
        }
      }
    }
  } else {
    for (uint32_t j = 0; j < n; ++j)
      if (j != i) pair_term<CUT, STR>(xi, yi, zi, r[j], prow, bx, a);
  }
  store_result<STR, INIT>(i, a, f, w, sts);
}

// All-pairs force kernel (neighbour-list method "none", the nbl == 0 branch of
// Calcf_NBL_tp, src/PairPotential.cu:431-441).  Positions are staged through
// LDS in kBlock tiles.
template <int CUT, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_lj_all(const float4* __restrict__ r, float4* __restrict__ f,
                                                   float4* __restrict__ w, float4* __restrict__ sts,
                                                   const float* __restrict__ params, uint32_t nt, Box bx,
                                                   uint32_t n) {
  extern __shared__ float4 s_dyn[];
  float4* s_tab = s_dyn;
  float4* s_r = s_dyn + 2 * nt * nt;
  load_table(s_tab, params, nt);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const float4 ri = live ? r[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double xi = ri.x, yi = ri.y, zi = ri.z;
  Acc a;
  __syncthreads();
  const float4* prow = s_tab + 2 * __float_as_uint(ri.w) * nt;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t jl = base + threadIdx.x;
    if (jl < n) s_r[threadIdx.x] = r[jl];
    __syncthreads();
    const uint32_t cnt = (n - base) < kBlock ? (n - base) : kBlock;
    if (live) {
      for (uint32_t k = 0; k < cnt; ++k)
        if (base + k != i) pair_term<CUT, STR>(xi, yi, zi, s_r[k], prow, bx, a);
    }
    __syncthreads();
  }
  if (live) store_result<STR, INIT>(i, a, f, w, sts);
}

template <int CUT, bool STR, bool INIT>
int launch_lj(const rumdx_particles* p, const rumdx_nblist* nb, const float* params, uint32_t nt, const Box& bx,
              hipStream_t st) {
  const uint32_t nblk = blocks_for(p->n);
  float4* r = reinterpret_cast<float4*>(p->r_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  const size_t tab_bytes = sizeof(float4) * 2 * nt * nt;
  if (nb != nullptr) {
    CellFallback fb{};
    if (nb->cell_rank_dev != nullptr) {
      fb.cell_rank = nb->cell_rank_dev;
      fb.cell_start = nb->cell_start_dev;
      fb.slot_cell = nb->slot_cell_dev;
      fb.nc0 = nb->nc[0];
      fb.nc1 = nb->nc[1];
      fb.nc2 = nb->nc[2];
    }
    hipLaunchKernelGGL((k_lj_nbl<CUT, STR, INIT>), dim3(nblk), dim3(kBlock), tab_bytes, st, r, f, w, sts,
                       nb->nbl_dev, nb->num_nbrs_dev, params, nt, bx, p->n, p->nvp, nb->max_nbrs, fb);
  } else {
    hipLaunchKernelGGL((k_lj_all<CUT, STR, INIT>), dim3(nblk), dim3(kBlock), tab_bytes + sizeof(float4) * kBlock,
                       st, r, f, w, sts, params, nt, bx, p->n);
  }
  return check_launch("rumdx_pair_lj");
}

template <int CUT>
int dispatch_lj(const rumdx_particles* p, const rumdx_nblist* nb, const float* params, uint32_t nt, const Box& bx,
                bool init, bool str, hipStream_t st) {
  if (str) return init ? launch_lj<CUT, true, true>(p, nb, params, nt, bx, st)
                       : launch_lj<CUT, true, false>(p, nb, params, nt, bx, st);
  return init ? launch_lj<CUT, false, true>(p, nb, params, nt, bx, st)
              : launch_lj<CUT, false, false>(p, nb, params, nt, bx, st);
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_pair_lj(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                             uint32_t ntypes, const float box[8], int cutoff_method, int initialize,
                             int calc_stresses, void* stream) {
  using namespace rumdx;
  if (p == nullptr || box == nullptr || params_dev == nullptr) return fail(RUMDX_E_INVALID, "rumdx_pair_lj: null argument");
  if (p->r_dev == nullptr || p->f_dev == nullptr || p->w_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj: particle arrays not allocated");
  if (calc_stresses && p->sts_dev == nullptr) return fail(RUMDX_E_INVALID, "rumdx_pair_lj: stresses requested but sts is null");
  if (ntypes == 0 || ntypes > RUMDX_MAX_TYPES) return fail(RUMDX_E_INVALID, "rumdx_pair_lj: ntypes out of range");
  if (p->n == 0) return RUMDX_OK;
  if (nb != nullptr && (nb->nbl_dev == nullptr || nb->num_nbrs_dev == nullptr || nb->max_nbrs == 0))
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj: neighbour list not allocated");
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const bool init = initialize != 0, str = calc_stresses != 0;
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS: return dispatch_lj<RUMDX_CUTOFF_NS>(p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SP: return dispatch_lj<RUMDX_CUTOFF_SP>(p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SF: return dispatch_lj<RUMDX_CUTOFF_SF>(p, nb, params_dev, ntypes, bx, init, str, st);
    default: return fail(RUMDX_E_INVALID, "rumdx_pair_lj: invalid cutoff method");
  }
}
