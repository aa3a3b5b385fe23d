// All-FP64 LJ 12-6 force / energy / virial kernels (gfx950): the comparison
// path for the FP32 kernels of force.hip.
//
// Reference arithmetic: fij (include/rumd/PairPotential.h:1198-1244) with
// Pot_IPL_12_6::ComputeInteraction (include/rumd/PairPotential.h:161-168),
// evaluated in double.  Positions are r (+ rlo in mixed mode) widened to
// double; the minimum image is formed in double; the cut-off shifts u(rc) and
// s(rc) rc are recomputed in double from the table's rc, A12 and A6 (RUMD
// computes them in float on the host, src/PairPotential.cu:82-125); the
// per-particle sums are double and are rounded once into the float4 outputs.
//
// The list formats are those of the FP32 kernels: tile lists (sorted method)
// are served by a tile kernel staging the halo as double positions in LDS;
// N2 lists and the no-list case by one lane per particle reading positions
// from global memory.
#include "rumdx_common.h"
#include "tiles.h"

namespace rumdx {
namespace {

// Per type pair, in LDS: {rc^2, A12, A6, 12 A12, 6 A6, -u(rc), s(rc) rc, rc}.
constexpr uint32_t kTabD = 8;

__device__ __forceinline__ void load_table_f64(double* s_tab, const float* __restrict__ params, uint32_t nt) {
  for (uint32_t k = threadIdx.x; k < nt * nt; k += blockDim.x) {
    const float* q = params + static_cast<size_t>(k) * RUMDX_NUM_PARAM;
    const double rc = q[0], a12 = q[1], a6 = q[2];
    const double i2 = 1.0 / (rc * rc), i6 = i2 * i2 * i2;
    const double ucut = i6 * (a12 * i6 + a6);
    const double src = i6 * i2 * (12.0 * a12 * i6 + 6.0 * a6) * rc;  // s(rc) rc = -u'(rc)
    double* t = s_tab + kTabD * k;
    t[0] = rc * rc;
    t[1] = a12;
    t[2] = a6;
    t[3] = 12.0 * a12;
    t[4] = 6.0 * a6;
    t[5] = -ucut;
    t[6] = src;
    t[7] = rc;
  }
}

struct AccD {
  double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0, vir = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, syz = 0.0, sxz = 0.0, sxy = 0.0;
};

struct PosD {
  double x, y, z;
  uint32_t t, pad;
};

__device__ __forceinline__ PosD pos_d(const float4* __restrict__ r, const float4* __restrict__ rlo, uint32_t j) {
  const float4 h = r[j];
  PosD p{static_cast<double>(h.x), static_cast<double>(h.y), static_cast<double>(h.z), __float_as_uint(h.w), 0u};
  if (rlo != nullptr) {
    const float4 l = rlo[j];
    p.x += static_cast<double>(l.x);
    p.y += static_cast<double>(l.y);
    p.z += static_cast<double>(l.z);
  }
  return p;
}

template <int CUT, bool STR>
__device__ __forceinline__ void pair_f64(const PosD& pi, const PosD& pj, const double* trow, const Box& bx,
                                         AccD& a) {
  double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
  dx -= bx.Ld[0] * rint(dx * bx.iLd[0]);
  dy -= bx.Ld[1] * rint(dy * bx.iLd[1]);
  dz -= bx.Ld[2] * rint(dz * bx.iLd[2]);
  const double d2 = dx * dx + dy * dy + dz * dz;
  const double* t = trow + kTabD * pj.t;
  if (!(d2 < t[0])) return;
  const double inv2 = 1.0 / d2, inv6 = inv2 * inv2 * inv2;
  double s = inv6 * inv2 * (t[3] * inv6 + t[4]);
  double u = inv6 * (t[1] * inv6 + t[2]);
  if (CUT == RUMDX_CUTOFF_SP) {
    u += t[5];
  } else if (CUT == RUMDX_CUTOFF_SF) {
    const double r = sqrt(d2);
    s -= t[6] / r;
    u += t[6] * (r - t[7]) + t[5];
  }
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
__device__ __forceinline__ void store_f64(uint32_t i, const AccD& a, float4* __restrict__ f, float4* __restrict__ w,
                                          float4* __restrict__ sts) {
  const double u = 0.5 * a.u;  // each pair energy was counted from both ends
  if (INIT) {
    f[i] = make_float4(static_cast<float>(a.fx), static_cast<float>(a.fy), static_cast<float>(a.fz),
                       static_cast<float>(u));
    w[i] = make_float4(0.f, STR ? static_cast<float>(a.sxz) : 0.f, STR ? static_cast<float>(a.sxy) : 0.f,
                       static_cast<float>(a.vir));
    if (STR)
      sts[i] = make_float4(static_cast<float>(a.sxx), static_cast<float>(a.syy), static_cast<float>(a.szz),
                           static_cast<float>(a.syz));
  } else {
    const float4 fo = f[i];
    f[i] = make_float4(static_cast<float>(fo.x + a.fx), static_cast<float>(fo.y + a.fy),
                       static_cast<float>(fo.z + a.fz), static_cast<float>(fo.w + u));
    float4 wo = w[i];
    wo.w = static_cast<float>(wo.w + a.vir);
    if (STR) {
      wo.y = static_cast<float>(wo.y + a.sxz);
      wo.z = static_cast<float>(wo.z + a.sxy);
      const float4 so = sts[i];
      sts[i] = make_float4(static_cast<float>(so.x + a.sxx), static_cast<float>(so.y + a.syy),
                           static_cast<float>(so.z + a.szz), static_cast<float>(so.w + a.syz));
    }
    w[i] = wo;
  }
}

// Tile lists of the sorted method: the halo as double positions in LDS when
// it fits, else positions from global memory (as is the stencil rescan of an
// overflowed list).
template <int CUT, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_lj_tiles_f64(const float4* __restrict__ r, const float4* __restrict__ rlo,
                                                         float4* __restrict__ f, float4* __restrict__ w,
                                                         float4* __restrict__ sts, const uint16_t* __restrict__ nbl16,
                                                         const uint32_t* __restrict__ num_nbrs,
                                                         const uint32_t* __restrict__ slot_cell,
                                                         const float* __restrict__ params, uint32_t nt, Box bx,
                                                         TileGeom tg, uint32_t cap, uint32_t nvp, uint32_t max_nbrs) {
  __shared__ HaloLds L;
  extern __shared__ double s_dd[];
  double* s_tab = s_dd;
  PosD* s_pd = reinterpret_cast<PosD*>(s_dd + kTabD * nt * nt);
  load_table_f64(s_tab, params, nt);
  TileGeom g = tg;
  g.halo_cap = 0;  // table only; the double copy follows
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  stage_halo(t, g, nullptr, nullptr, L);
  const Halo& h = L.h;
  const bool staged = h.npart <= cap;
  const bool bytes = h.npart <= kByteEntryHalo;
  if (staged) {
    for (uint32_t c = threadIdx.x; c < h.ncell; c += kBlock) {
      const uint32_t o0 = L.off[c], n = L.off[c + 1] - o0, g0 = L.start[c];
      for (uint32_t k = 0; k < n; ++k) s_pd[o0 + k] = pos_d(r, rlo, g0 + k);
    }
  }
  __syncthreads();
  for (uint32_t i = h.slot0 + threadIdx.x; i < h.slot1; i += kBlock) {
    const PosD pi = pos_d(r, rlo, i);
    const double* trow = s_tab + kTabD * nt * pi.t;
    AccD a;
    const uint32_t nn = num_nbrs[i];
    if (staged && nn < max_nbrs) {
      for (uint32_t k = 0; k < nn; ++k) {
        const uint32_t e = nbl16[(static_cast<size_t>(k >> 3) * nvp + i) * 8 + (k & 7u)];
        pair_f64<CUT, STR>(pi, s_pd[decode_entry(e, bytes)], trow, bx, a);
      }
    } else {
      const uint32_t lin = slot_cell[i];
      const int cx = static_cast<int>(lin % tg.nc0);
      const int cy = static_cast<int>((lin / tg.nc0) % tg.nc1);
      const int cz = static_cast<int>(lin / (static_cast<uint32_t>(tg.nc0) * tg.nc1));
      for (int dz = -2; dz <= 2; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
          for (int dx = -2; dx <= 2; ++dx) {
            const uint32_t hcell = halo_cell(h, tg, cx + dx, cy + dy, cz + dz);
            const uint32_t g0 = L.start[hcell], n = L.off[hcell + 1] - L.off[hcell];
            for (uint32_t j = g0; j < g0 + n; ++j)
              if (j != i) pair_f64<CUT, STR>(pi, pos_d(r, rlo, j), trow, bx, a);
          }
    }
    store_f64<STR, INIT>(i, a, f, w, sts);
  }
}

// N2 lists (u32 slots, column-major) or, with nbl == nullptr, all pairs.
template <int CUT, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_lj_slots_f64(const float4* __restrict__ r, const float4* __restrict__ rlo,
                                                         float4* __restrict__ f, float4* __restrict__ w,
                                                         float4* __restrict__ sts, const uint32_t* __restrict__ nbl,
                                                         const uint32_t* __restrict__ num_nbrs,
                                                         const float* __restrict__ params, uint32_t nt, Box bx,
                                                         uint32_t n, uint32_t nvp, uint32_t max_nbrs) {
  extern __shared__ double s_dd[];
  load_table_f64(s_dd, params, nt);
  __syncthreads();
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const PosD pi = pos_d(r, rlo, i);
  const double* trow = s_dd + kTabD * nt * pi.t;
  AccD a;
  const uint32_t nn = nbl != nullptr ? num_nbrs[i] : 0u;
  if (nbl != nullptr && nn < max_nbrs) {
    for (uint32_t k = 0; k < nn; ++k) pair_f64<CUT, STR>(pi, pos_d(r, rlo, nbl[static_cast<size_t>(k) * nvp + i]), trow, bx, a);
  } else {
    for (uint32_t j = 0; j < n; ++j)
      if (j != i) pair_f64<CUT, STR>(pi, pos_d(r, rlo, j), trow, bx, a);
  }
  store_f64<STR, INIT>(i, a, f, w, sts);
}

// LDS of the tile kernel: the table and the halo positions.
size_t tiles_f64_lds(uint32_t cap, uint32_t nt) { return sizeof(double) * kTabD * nt * nt + sizeof(PosD) * cap; }

// The halo capacity of the tile kernel: the list's, bounded by the device's
// LDS per workgroup (halos beyond it read positions from global memory).
uint32_t tiles_f64_cap(uint32_t cap, uint32_t nt) {
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds <= 0)
    lds = 64 * 1024;
  const size_t fixed = sizeof(HaloLds) + sizeof(double) * kTabD * nt * nt + 256;
  const size_t most = static_cast<size_t>(lds) > fixed ? (static_cast<size_t>(lds) - fixed) / sizeof(PosD) : 0;
  return cap < most ? cap : static_cast<uint32_t>(most);
}

template <int CUT, bool STR, bool INIT>
int launch_f64(const rumdx_particles* p, const rumdx_nblist* nb, const float* params, uint32_t nt, const Box& bx,
               hipStream_t st) {
  const float4* r = reinterpret_cast<const float4*>(p->r_dev);
  const float4* rlo = reinterpret_cast<const float4*>(p->rlo_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  if (nb != nullptr && nb->nbl16_dev != nullptr) {
    const TileGeom tg{nb->nc[0], nb->nc[1], nb->nc[2], nb->cell_rank_dev, nb->cell_start_dev,
                      nb->tile_first_rank_dev, nb->tile_origin_dev, nb->ntiles, nb->halo_cap, nullptr, nullptr};
    const uint32_t cap = tiles_f64_cap(nb->halo_cap, nt);
    hipLaunchKernelGGL((k_lj_tiles_f64<CUT, STR, INIT>), dim3(nb->ntiles), dim3(kBlock), tiles_f64_lds(cap, nt), st, r,
                       rlo, f, w, sts, nb->nbl16_dev, nb->num_nbrs_dev, nb->slot_cell_dev, params, nt, bx, tg, cap,
                       p->nvp, nb->max_nbrs);
  } else {
    hipLaunchKernelGGL((k_lj_slots_f64<CUT, STR, INIT>), dim3(blocks_for(p->n)), dim3(kBlock),
                       sizeof(double) * kTabD * nt * nt, st, r, rlo, f, w, sts, nb != nullptr ? nb->nbl_dev : nullptr,
                       nb != nullptr ? nb->num_nbrs_dev : nullptr, params, nt, bx, p->n, p->nvp,
                       nb != nullptr ? nb->max_nbrs : 0u);
  }
  return check_launch("rumdx_pair_lj_fp64");
}

template <int CUT>
int dispatch_f64(const rumdx_particles* p, const rumdx_nblist* nb, const float* params, uint32_t nt, const Box& bx,
                 bool init, bool str, hipStream_t st) {
  if (str) return init ? launch_f64<CUT, true, true>(p, nb, params, nt, bx, st)
                       : launch_f64<CUT, true, false>(p, nb, params, nt, bx, st);
  return init ? launch_f64<CUT, false, true>(p, nb, params, nt, bx, st)
              : launch_f64<CUT, false, false>(p, nb, params, nt, bx, st);
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_pair_lj_fp64(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                                  uint32_t ntypes, const float box[8], int cutoff_method, int initialize,
                                  int calc_stresses, void* stream) {
  using namespace rumdx;
  if (p == nullptr || box == nullptr || params_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: null argument");
  if (p->r_dev == nullptr || p->f_dev == nullptr || p->w_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: particle arrays not allocated");
  if (calc_stresses && p->sts_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: stresses requested but sts is null");
  if (nb != nullptr && nb->excl_start_dev != nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: the FP64 comparison path takes lists without exclusions");
  if (ntypes == 0 || ntypes > RUMDX_MAX_TYPES) return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: ntypes out of range");
  if (p->n == 0) return RUMDX_OK;
  if (nb != nullptr && ((nb->nbl_dev == nullptr && nb->nbl16_dev == nullptr) || nb->num_nbrs_dev == nullptr ||
                        nb->max_nbrs == 0))
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: neighbour list not allocated");
  if (nb != nullptr && nb->nbl16_dev != nullptr &&
      (nb->ntiles == 0 || nb->tile_first_rank_dev == nullptr || nb->tile_origin_dev == nullptr ||
       nb->slot_cell_dev == nullptr || nb->cell_rank_dev == nullptr || nb->cell_start_dev == nullptr))
    return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: tile arrays not allocated");
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  const bool init = initialize != 0, str = calc_stresses != 0;
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS: return dispatch_f64<RUMDX_CUTOFF_NS>(p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SP: return dispatch_f64<RUMDX_CUTOFF_SP>(p, nb, params_dev, ntypes, bx, init, str, st);
    case RUMDX_CUTOFF_SF: return dispatch_f64<RUMDX_CUTOFF_SF>(p, nb, params_dev, ntypes, bx, init, str, st);
    default: return fail(RUMDX_E_INVALID, "rumdx_pair_lj_fp64: invalid cutoff method");
  }
}
