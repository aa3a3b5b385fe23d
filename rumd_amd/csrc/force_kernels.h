// Pair-force kernels over every list format, parameterised by a pair policy.
//
// Reference: Calcf_NBL_tp_equal_one / Calcf_NBL_tp (src/PairPotential.cu:330-515),
// fij (include/rumd/PairPotential.h:1198-1244) and the per-potential CalcF
// bodies of src/Generate_PP_FunctionBodies.py:3-127.
//
// Design.  Lists of the sorted method are served by the tile kernel: one
// workgroup per cell tile (a brick of 4^3 cells, a contiguous slot range)
// stages the positions of the tile's halo in LDS, and list entries are LDS
// byte offsets in 16-byte groups of 8, so a pair costs one LDS read and no
// address arithmetic.  Lists of the N2 method (u32 slots, column-major like
// RUMD's) are served by a lane-per-particle kernel gathering positions
// through L2; the no-list method stages positions through LDS.  Pair
// arithmetic is FP32; where a pair may cross the periodic boundary the
// minimum-image displacement is formed in FP64, so large boxes keep full FP32
// accuracy.  Per-particle sums stay in registers; no atomics, so results are
// reproducible bit for bit.
//
// A pair policy PP supplies the potential: kRow float4 per type pair in LDS,
// kWavesPerEU (the tile kernel's occupancy target; kTunedWaves: the
// stepping variant also exists for 4 and 6 waves, rumdx_nblist.tile_waves_per_eu),
// load_table (RUMD's 12-float rows -> that LDS table) and core<STR> (the
// masked pair term for a displacement, accumulated into Acc).  The 12-6
// policy lives in force.hip, the other potential forms in force_generic.hip.
#pragma once

#include <string>

#include "leapfrog.h"
#include "rumdx_common.h"
#include "tiles.h"

namespace rumdx {
namespace {

// Per-particle sums of the pair terms.
struct Acc {
  float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f, vir = 0.f;
  float sxx = 0.f, syy = 0.f, szz = 0.f, syz = 0.f, sxz = 0.f, sxy = 0.f;
};

// Adds one pair's masked force factor s and energy u (fij, STR branch
// included: include/rumd/PairPotential.h:1225-1240).
template <bool STR>
__device__ __forceinline__ void accumulate(float dx, float dy, float dz, float d2, float s, float u, Acc& a) {
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

// Cell data used to rescan the stencil for a particle whose list overflowed.
struct CellFallback {
  const uint32_t* cell_rank;
  const uint32_t* cell_start;
  const uint32_t* slot_cell;
  int nc0, nc1, nc2;
  Exclusions ex;  // the list's exclusions, which a rescan must skip
};

// Pair term with the minimum-image displacement formed in FP64 and rounded
// to FP32 (any pair, including those across the periodic boundary).
template <bool STR, class PP>
__device__ __forceinline__ void pair_term(const PP& pp, double xi, double yi, double zi, float4 rj, const float4* prow,
                                          const Box& bx, Acc& a) {
  double x = xi - static_cast<double>(rj.x);
  double y = yi - static_cast<double>(rj.y);
  double z = zi - static_cast<double>(rj.z);
  x -= bx.Ld[0] * rint(x * bx.iLd[0]);
  y -= bx.Ld[1] * rint(y * bx.iLd[1]);
  z -= bx.Ld[2] * rint(z * bx.iLd[2]);
  pp.template core<STR>(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z),
                        static_cast<float>(x * x + y * y + z * z), __float_as_uint(rj.w), prow, a);
}

// Pair term for a pair that cannot wrap: the FP32 difference of two floats is
// correctly rounded, i.e. the same number the FP64 path produces.
template <bool STR, class PP>
__device__ __forceinline__ void pair_term_direct(const PP& pp, float4 ri, float4 rj, const float4* prow, Acc& a) {
  const float dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
  pp.template core<STR>(dx, dy, dz, dx * dx + dy * dy + dz * dz, __float_as_uint(rj.w), prow, a);
}

// Forces on particle i from the 5x5x5 cell stencil around the cell it was
// binned in at the last build, positions from global memory, displacements in
// FP64 (an overflowed list, or a tile whose halo was not staged).  Every pair
// now inside rc was inside rc + skin at the build, so it is in that stencil.
template <bool STR, class PP>
__device__ __forceinline__ void rescan_cells(const PP& pp, uint32_t i, double xi, double yi, double zi,
                                             const float4* __restrict__ r, const float4* prow, const Box& bx,
                                             const CellFallback& fb, Acc& a) {
  const ExclusionRange er = fb.ex.of(i);
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
          if (j != i && !fb.ex.excludes(er, j)) pair_term<STR>(pp, xi, yi, zi, r[j], prow, bx, a);
      }
    }
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

// Sums a particle's partial accumulators over the 2^lg lanes that share it
// (a fixed butterfly, so the result does not depend on timing).  Called by
// every lane of the wave.
template <bool STR>
__device__ __forceinline__ void combine_lanes(Acc& a, uint32_t lg) {
  for (uint32_t o = 1; o < (1u << lg); o <<= 1) {
    a.fx += __shfl_xor(a.fx, o, 64);
    a.fy += __shfl_xor(a.fy, o, 64);
    a.fz += __shfl_xor(a.fz, o, 64);
    a.u += __shfl_xor(a.u, o, 64);
    a.vir += __shfl_xor(a.vir, o, 64);
    if (STR) {
      a.sxx += __shfl_xor(a.sxx, o, 64);
      a.syy += __shfl_xor(a.syy, o, 64);
      a.szz += __shfl_xor(a.szz, o, 64);
      a.syz += __shfl_xor(a.syz, o, 64);
      a.sxz += __shfl_xor(a.sxz, o, 64);
      a.sxy += __shfl_xor(a.sxy, o, 64);
    }
  }
}

// Neighbour-list force kernel for u32 slot lists (N2 method).  2^lgp lanes
// share a particle (RUMD's threads per particle): lane part takes entries
// part, part + 2^lgp, ..., and a fixed butterfly combines them.  An
// overflowed list falls back to the cell stencil (fb.cell_rank != nullptr)
// or to all pairs.
template <class PP, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_pair_nbl(PP pp, const float4* __restrict__ r, float4* __restrict__ f,
                                                   float4* __restrict__ w, float4* __restrict__ sts,
                                                   const uint32_t* __restrict__ nbl,
                                                   const uint32_t* __restrict__ num_nbrs,
                                                   const float* __restrict__ params, uint32_t nt, Box bx,
                                                   uint32_t n, uint32_t nvp, uint32_t max_nbrs, CellFallback fb,
                                                   uint32_t lgp) {
  extern __shared__ float4 s_tab[];
  pp.load_table(s_tab, params, nt);
  __syncthreads();

  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t gid = b * kBlock + threadIdx.x;
  const uint32_t i = gid >> lgp, part = gid & ((1u << lgp) - 1u), step = 1u << lgp;
  const bool live = i < n;
  Acc a;
  if (live) {
    const float4 ri = r[i];
    const float4* prow = s_tab + PP::kRow * __float_as_uint(ri.w) * nt;
    const double xi = ri.x, yi = ri.y, zi = ri.z;
    const uint32_t nn = num_nbrs[i];
    if (nn < max_nbrs) {
      const size_t stride = static_cast<size_t>(nvp) << lgp;
      const uint32_t* col = nbl + static_cast<size_t>(part) * nvp + i;
      uint32_t k = part;
      for (; k + 3 * step < nn; k += 4 * step) {
        const uint32_t j0 = __builtin_nontemporal_load(col);
        const uint32_t j1 = __builtin_nontemporal_load(col + stride);
        const uint32_t j2 = __builtin_nontemporal_load(col + 2 * stride);
        const uint32_t j3 = __builtin_nontemporal_load(col + 3 * stride);
        col += 4 * stride;
        const float4 r0 = r[j0], r1 = r[j1], r2 = r[j2], r3 = r[j3];
        pair_term<STR>(pp, xi, yi, zi, r0, prow, bx, a);
        pair_term<STR>(pp, xi, yi, zi, r1, prow, bx, a);
        pair_term<STR>(pp, xi, yi, zi, r2, prow, bx, a);
        pair_term<STR>(pp, xi, yi, zi, r3, prow, bx, a);
      }
      for (; k < nn; k += step) {
        const uint32_t j = __builtin_nontemporal_load(col);
        col += stride;
        pair_term<STR>(pp, xi, yi, zi, r[j], prow, bx, a);
      }
    } else if (part == 0 && fb.cell_rank != nullptr) {
      rescan_cells<STR>(pp, i, xi, yi, zi, r, prow, bx, fb, a);
    } else if (part == 0) {
      const ExclusionRange er = fb.ex.of(i);
      for (uint32_t j = 0; j < n; ++j)
        if (j != i && !fb.ex.excludes(er, j)) pair_term<STR>(pp, xi, yi, zi, r[j], prow, bx, a);
    }
  }
  combine_lanes<STR>(a, lgp);
  if (live && part == 0) store_result<STR, INIT>(i, a, f, w, sts);
}

// All-pairs force kernel (neighbour-list method "none", the nbl == 0 branch of
// Calcf_NBL_tp, src/PairPotential.cu:431-441).  Positions are staged through
// LDS in kBlock tiles.
template <class PP, bool STR, bool INIT>
__global__ __launch_bounds__(kBlock) void k_pair_all(PP pp, const float4* __restrict__ r, float4* __restrict__ f,
                                                   float4* __restrict__ w, float4* __restrict__ sts,
                                                   const float* __restrict__ params, uint32_t nt, Box bx,
                                                   uint32_t n) {
  extern __shared__ float4 s_dyn[];
  float4* s_tab = s_dyn;
  float4* s_r = s_dyn + PP::kRow * nt * nt;
  pp.load_table(s_tab, params, nt);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const float4 ri = live ? r[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double xi = ri.x, yi = ri.y, zi = ri.z;
  Acc a;
  __syncthreads();
  const float4* prow = s_tab + PP::kRow * __float_as_uint(ri.w) * nt;
  for (uint32_t base = 0; base < n; base += kBlock) {
    const uint32_t jl = base + threadIdx.x;
    if (jl < n) s_r[threadIdx.x] = r[jl];
    __syncthreads();
    const uint32_t cnt = (n - base) < kBlock ? (n - base) : kBlock;
    if (live) {
      for (uint32_t k = 0; k < cnt; ++k)
        if (base + k != i) pair_term<STR>(pp, xi, yi, zi, s_r[k], prow, bx, a);
    }
    __syncthreads();
  }
  if (live) store_result<STR, INIT>(i, a, f, w, sts);
}

// Forces on the particles of one tile whose halo positions are in LDS.
// Waves take the tile's particles 64 >> lgp at a time in the build's lane
// order (descending list length, so a wave's lists have similar lengths);
// each list is spread over 2^lg lanes (groups part, part + 2^lg, ...; lg =
// lgp, more in a tile's last, partial chunk), combined by a butterfly.  FAST:
// the halo was staged in the shifted frame, so every pair takes the direct
// FP32 difference; otherwise the halo wraps and displacements are formed in
// FP64.  An overflowed list, or every list of a tile whose halo was not
// staged, rescans the particle's build-time cell stencil from global memory.
template <class PP, bool STR, bool INIT, bool FAST>
__device__ __forceinline__ void tile_particles(const PP& pp, uint32_t slot0, uint32_t slot1, bool staged,
                                               const uint32_t shifts,
                                               const float4* __restrict__ r, float4* __restrict__ f,
                                               float4* __restrict__ w, float4* __restrict__ sts,
                                               const uint16_t* __restrict__ nbl16,
                                               const uint32_t* __restrict__ num_nbrs,
                                               const uint16_t* __restrict__ lane_order, const float4* s_pos,
                                               const float4* s_tab, uint32_t nt, const Box& bx,
                                               const CellFallback& fb, uint32_t nvp, uint32_t max_nbrs,
                                               uint32_t lgp) {
  const uint32_t np = slot1 - slot0;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t per_wave = 64u >> lgp;  // particles per wave and chunk: 2^lgp lanes each
  for (uint32_t c0 = (threadIdx.x >> 6) * per_wave; c0 < np; c0 += (kBlock / 64) * per_wave) {
    const uint32_t cnt = (np - c0) < per_wave ? (np - c0) : per_wave;
    uint32_t lg = lgp;  // a partial chunk spreads its lists further, up to 8 lanes
    while (lg < 3u && (cnt << (lg + 1u)) <= 64u) ++lg;
    const uint32_t part = lane & ((1u << lg) - 1u);
    const uint32_t pl = lane >> lg;
    const bool live = pl < cnt;
    Acc a;
    uint32_t i = 0;
    if (live) {
      i = slot0 + lane_order[slot0 + c0 + pl];
      float4 ri = r[i];
      const float4* prow = s_tab + PP::kRow * __float_as_uint(ri.w) * nt;
      const uint32_t nn = num_nbrs[i];
      const uint32_t nfull = nn >> 3, rem = nn & 7u;
      const uint32_t step = 1u << lg;
      const size_t gstride = static_cast<size_t>(nvp) << lg;
      const u32x4* grp = reinterpret_cast<const u32x4*>(nbl16) + static_cast<size_t>(part) * nvp + i;
      const bool tail = rem != 0u && (nfull & (step - 1u)) == part;  // this lane owns group nfull
      if (staged && nn < max_nbrs) {
        if (FAST) {
          ri.x = shift_coord(ri.x, shifts, 0, bx.L[0]);
          ri.y = shift_coord(ri.y, shifts, 1, bx.L[1]);
          ri.z = shift_coord(ri.z, shifts, 2, bx.L[2]);
          // this lane's groups part, part + step, ... below ng, the last one
          // partial when it is the tail; each group's load is issued one
          // group ahead, so its latency overlaps the previous 8 pair terms
          const uint32_t ng = nfull + (tail ? 1u : 0u);
          u32x4 q = {0u, 0u, 0u, 0u};
          if (part < ng) q = __builtin_nontemporal_load(grp);
          for (uint32_t g = part; g < ng; g += step) {
            u32x4 qn = q;
            if (g + step < ng) qn = __builtin_nontemporal_load(grp + gstride);
            grp += gstride;
            if (g < nfull) {
              float4 rj[8];  // the group's positions first: eight LDS reads in flight
#pragma unroll
              for (int e = 0; e < 8; ++e) rj[e] = halo_pos(s_pos, group_entry(q, e));
#pragma unroll
              for (int e = 0; e < 8; ++e) pair_term_direct<STR>(pp, ri, rj[e], prow, a);
            } else {
#pragma unroll
              for (int e = 0; e < 7; ++e)
                if (static_cast<uint32_t>(e) < rem)
                  pair_term_direct<STR>(pp, ri, halo_pos(s_pos, group_entry(q, e)), prow, a);
            }
            q = qn;
          }
        } else {
          const double xi = ri.x, yi = ri.y, zi = ri.z;
          for (uint32_t g = part; g < nfull; g += step, grp += gstride) {
            const u32x4 q = __builtin_nontemporal_load(grp);
#pragma unroll
            for (int e = 0; e < 8; ++e)
              pair_term<STR>(pp, xi, yi, zi, halo_pos(s_pos, group_entry(q, e)), prow, bx, a);
          }
          if (tail) {
            const u32x4 q = __builtin_nontemporal_load(grp);
#pragma unroll
            for (int e = 0; e < 7; ++e)
              if (static_cast<uint32_t>(e) < rem)
                pair_term<STR>(pp, xi, yi, zi, halo_pos(s_pos, group_entry(q, e)), prow, bx, a);
          }
        }
      } else if (part == 0) {
        rescan_cells<STR>(pp, i, ri.x, ri.y, ri.z, r, prow, bx, fb, a);
      }
    }
    combine_lanes<STR>(a, lg);
    if (live && part == 0) store_result<STR, INIT>(i, a, f, w, sts);
  }
}

// Occupancy target of the tile force kernel: waves per SIMD, PP::kWavesPerEU.
// For the 12-6 policy five fit the LDS of a tile whose halo capacity is at
// most 2040 particles, and the compiler then keeps the kernel within 102
// VGPRs; policies with more live state ask for fewer waves so that nothing
// spills.

// Tile force kernel for lists of the sorted method: one workgroup per cell
// tile.  A tile whose halo fits the capacity stages it in LDS by one
// coalesced gather over the slot list the build recorded (in the shifted
// frame when the build marked the tile fast); list entries are LDS byte
// offsets, so neighbour positions come from LDS.  The kernel uses no static
// LDS: its dynamic LDS is the halo capacity and the parameter table, which
// sets how many workgroups share a CU.
//
// With FUSE (initialising, no stresses) the kernel also takes the NVT step:
// once the tile's forces are stored, its particles are advanced in slot order
// (coalesced; the force loop visits them by list length), the positions into
// fz.r_out, and the tile's sum of m v^2 goes to fz.partials[tile] for
// k_nvt_finish.
template <class PP, bool STR, bool INIT, int WPE = PP::kWavesPerEU, bool FUSE = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pair_tiles(
    PP pp, const float4* __restrict__ r, float4* __restrict__ f, float4* __restrict__ w, float4* __restrict__ sts,
    const uint16_t* __restrict__ nbl16, const uint32_t* __restrict__ num_nbrs,
    const uint16_t* __restrict__ lane_order, const float* __restrict__ params, uint32_t nt, Box bx,
    const uint4* __restrict__ tile_info, const uint32_t* __restrict__ halo_slots, uint32_t halo_cap,
    CellFallback fb, uint32_t nvp, uint32_t max_nbrs, uint32_t lgp, NvtFuse fz) {
  static_assert(!FUSE || (INIT && !STR), "the fused step initialises forces and takes no stresses");
  extern __shared__ float4 s_dyn[];
  float4* s_pos = s_dyn;
  float4* s_tab = s_dyn + halo_cap;
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  const uint4 info = tile_info[t];
  pp.load_table(s_tab, params, nt);
  const bool staged = (info.y >> 16) == halo_cap && info.x <= halo_cap;
  const bool fast = (info.y & kTileFast) != 0u;
  // the fused step's state of this thread's first particle, read now so that
  // the reads overlap the staging and the pair loop
  const uint32_t i_own = info.z + threadIdx.x;
  NvtIn own{};
  if (FUSE && i_own < info.w) own = nvt_fused_load(i_own, r, fz);
  if (staged) {
    // four halo particles per thread and pass: their slot and position loads
    // are issued together, so a pass waits for two round trips, not eight
    const uint32_t shifts = fast ? info.y : 0u;
    const uint32_t* hs = halo_slots + static_cast<size_t>(t) * halo_cap;
    const uint32_t nh = info.x;
    for (uint32_t q0 = threadIdx.x; q0 < nh; q0 += 4 * kBlock) {
      uint32_t slot[4];
      float4 p[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + u * kBlock;
        slot[u] = hs[q < nh ? q : q0];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) p[u] = r[slot[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + u * kBlock;
        if (q < nh) {
          p[u].x = shift_coord(p[u].x, shifts, 0, bx.L[0]);
          p[u].y = shift_coord(p[u].y, shifts, 1, bx.L[1]);
          p[u].z = shift_coord(p[u].z, shifts, 2, bx.L[2]);
          s_pos[q] = p[u];
        }
      }
    }
  }
  __syncthreads();
  if (staged && fast)
    tile_particles<PP, STR, INIT, true>(pp, info.z, info.w, true, info.y, r, f, w, sts, nbl16, num_nbrs, lane_order,
                                         s_pos, s_tab, nt, bx, fb, nvp, max_nbrs, lgp);
  else
    tile_particles<PP, STR, INIT, false>(pp, info.z, info.w, staged, 0u, r, f, w, sts, nbl16, num_nbrs, lane_order,
                                          s_pos, s_tab, nt, bx, fb, nvp, max_nbrs, lgp);
  if (FUSE) {
    // the tile's forces are stored (and read back from L2 in slot order);
    // the halo is no longer read, so its LDS holds the reduction
    __syncthreads();
    const float ps = fz.nh != nullptr ? static_cast<float>(fz.nh->ps) : 0.f;
    double ke2 = 0.0;
    for (uint32_t i = i_own; i < info.w; i += kBlock)
      ke2 += static_cast<double>(nvt_fused(i, f[i], i == i_own ? own : nvt_fused_load(i, r, fz), fz, ps, bx));
    if (fz.nh != nullptr) {
      const double sum = block_sum(ke2, reinterpret_cast<double*>(s_dyn));
      if (threadIdx.x == 0) fz.partials[t] = sum;
    }
  }
}

// Launches the tile kernel over lists of the sorted method: the pruned inner
// lists when the list has them (a subset in the same order and format: the
// pairs beyond them are outside rc), and for the stepping variants (no
// stresses, initialising) of a policy whose occupancy target is tunable, the
// build compiled for rumdx_nblist.tile_waves_per_eu.
template <class PP, bool STR, bool INIT, bool FUSE>
void launch_tiles(const PP& pp, const rumdx_particles* p, const rumdx_nblist* nb, const float* params, uint32_t nt,
                  const Box& bx, uint32_t lgp, const NvtFuse& fz, hipStream_t st) {
  float4* r = reinterpret_cast<float4*>(p->r_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  const size_t tab_bytes = sizeof(float4) * PP::kRow * nt * nt;
  const CellFallback fb{nb->cell_rank_dev, nb->cell_start_dev, nb->slot_cell_dev, nb->nc[0], nb->nc[1], nb->nc[2],
                        Exclusions{p->unsorted_index_dev, nb->excl_start_dev, nb->excl_ids_dev}};
  const bool inner = nb->nbl16_in_dev != nullptr && nb->num_in_dev != nullptr;
  const uint16_t* lists = inner ? nb->nbl16_in_dev : nb->nbl16_dev;
  const uint32_t* counts = inner ? nb->num_in_dev : nb->num_nbrs_dev;
  const auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(nb->ntiles), dim3(kBlock), tab_bytes + sizeof(float4) * nb->halo_cap, st, pp, r, f,
                       w, sts, lists, counts, nb->lane_order_dev, params, nt, bx,
                       reinterpret_cast<const uint4*>(nb->tile_info_dev), nb->tile_halo_dev, nb->halo_cap, fb, p->nvp,
                       nb->max_nbrs, lgp, fz);
  };
  if constexpr (PP::kTunedWaves && !STR && INIT) {
    if (nb->tile_waves_per_eu == 4u) return launch(k_pair_tiles<PP, STR, INIT, 4, FUSE>);
    // the fused step's update state does not fit six waves' registers (it
    // would spill), and the tile's LDS allows five workgroups per CU anyway
    if constexpr (!FUSE)
      if (nb->tile_waves_per_eu == 6u) return launch(k_pair_tiles<PP, STR, INIT, 6, FUSE>);
  }
  launch(k_pair_tiles<PP, STR, INIT, PP::kWavesPerEU, FUSE>);
}

template <class PP, bool STR, bool INIT>
int launch_pair(const PP& pp, const char* entry, const rumdx_particles* p, const rumdx_nblist* nb,
                const float* params, uint32_t nt, const Box& bx, hipStream_t st) {
  const uint32_t nblk = blocks_for(p->n);
  float4* r = reinterpret_cast<float4*>(p->r_dev);
  float4* f = reinterpret_cast<float4*>(p->f_dev);
  float4* w = reinterpret_cast<float4*>(p->w_dev);
  float4* sts = reinterpret_cast<float4*>(p->sts_dev);
  const size_t tab_bytes = sizeof(float4) * PP::kRow * nt * nt;
  const uint32_t lpp = nb != nullptr ? nb->lanes_per_particle : 1u;
  const uint32_t lgp = lpp >= 8u ? 3u : lpp >= 4u ? 2u : lpp >= 2u ? 1u : 0u;
  if (nb != nullptr && nb->nbl16_dev != nullptr) {
    launch_tiles<PP, STR, INIT, false>(pp, p, nb, params, nt, bx, lgp, NvtFuse{}, st);
  } else if (nb != nullptr) {
    CellFallback fb{};
    fb.ex = Exclusions{p->unsorted_index_dev, nb->excl_start_dev, nb->excl_ids_dev};
    if (nb->cell_rank_dev != nullptr) {
      fb.cell_rank = nb->cell_rank_dev;
      fb.cell_start = nb->cell_start_dev;
      fb.slot_cell = nb->slot_cell_dev;
      fb.nc0 = nb->nc[0];
      fb.nc1 = nb->nc[1];
      fb.nc2 = nb->nc[2];
    }
    hipLaunchKernelGGL((k_pair_nbl<PP, STR, INIT>), dim3(blocks_for(p->n << lgp)), dim3(kBlock), tab_bytes, st, pp,
                       r, f, w, sts, nb->nbl_dev, nb->num_nbrs_dev, params, nt, bx, p->n, p->nvp, nb->max_nbrs, fb, lgp);
  } else {
    hipLaunchKernelGGL((k_pair_all<PP, STR, INIT>), dim3(nblk), dim3(kBlock), tab_bytes + sizeof(float4) * kBlock,
                       st, pp, r, f, w, sts, params, nt, bx, p->n);
  }
  return check_launch(entry);
}

// Initialize and stress variants of one pair policy.
template <class PP>
int dispatch_pair(const PP& pp, const char* entry, const rumdx_particles* p, const rumdx_nblist* nb,
                  const float* params, uint32_t nt, const Box& bx, bool init, bool str, hipStream_t st) {
  if (str) return init ? launch_pair<PP, true, true>(pp, entry, p, nb, params, nt, bx, st)
                       : launch_pair<PP, true, false>(pp, entry, p, nb, params, nt, bx, st);
  return init ? launch_pair<PP, false, true>(pp, entry, p, nb, params, nt, bx, st)
              : launch_pair<PP, false, false>(pp, entry, p, nb, params, nt, bx, st);
}


// Argument checks shared by the pair-force entry points; RUMDX_OK or the
// failure code (with the message set).
inline int check_pair_args(const char* entry, const rumdx_particles* p, const rumdx_nblist* nb,
                           const float* params_dev, uint32_t ntypes, const float* box, int calc_stresses) {
  const std::string e(entry);
  if (p == nullptr || box == nullptr || params_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": null argument");
  if (p->r_dev == nullptr || p->f_dev == nullptr || p->w_dev == nullptr)
    return fail(RUMDX_E_INVALID, e + ": particle arrays not allocated");
  if (calc_stresses && p->sts_dev == nullptr) return fail(RUMDX_E_INVALID, e + ": stresses requested but sts is null");
  if (ntypes == 0 || ntypes > RUMDX_MAX_TYPES) return fail(RUMDX_E_INVALID, e + ": ntypes out of range");
  if (nb != nullptr && ((nb->nbl_dev == nullptr && nb->nbl16_dev == nullptr) || nb->num_nbrs_dev == nullptr ||
                        nb->max_nbrs == 0))
    return fail(RUMDX_E_INVALID, e + ": neighbour list not allocated");
  if (nb != nullptr && nb->nbl16_dev != nullptr &&
      (nb->ntiles == 0 || nb->tile_first_rank_dev == nullptr || nb->tile_origin_dev == nullptr ||
       nb->slot_cell_dev == nullptr || nb->lane_order_dev == nullptr || nb->tile_halo_dev == nullptr ||
       nb->tile_info_dev == nullptr || nb->cell_rank_dev == nullptr || nb->cell_start_dev == nullptr ||
       nb->halo_cap > RUMDX_HALO_CAP_MAX))
    return fail(RUMDX_E_INVALID, e + ": tile arrays not allocated");
  if (nb != nullptr && nb->excl_start_dev != nullptr && (nb->excl_ids_dev == nullptr || p->unsorted_index_dev == nullptr))
    return fail(RUMDX_E_INVALID, e + ": exclusions need the exclusion ids and the particles' unsorted index");
  if (nb != nullptr) {
    const uint32_t lpp = nb->lanes_per_particle;
    if (lpp > 8u || (lpp & (lpp - 1u)) != 0u || static_cast<uint64_t>(p->n) * (lpp ? lpp : 1u) >= (1ull << 32))
      return fail(RUMDX_E_INVALID, e + ": lanes_per_particle must be 0, 1, 2, 4 or 8");
  }
  return RUMDX_OK;
}

}  // namespace
}  // namespace rumdx
