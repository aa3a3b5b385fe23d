// Cell tiles: the unit of work of the sorted neighbour-list build and of the
// tile force kernel.
//
// Cells are stored in Morton order of their coordinates, so an aligned
// kTileEdge^3 brick of cells is a contiguous run of ranks and therefore of
// particle slots.  One workgroup owns one brick ("tile") and stages its halo
// -- the bricks's cells dilated by the 2-cell stencil, at most 8^3 cells -- into
// LDS: a per-halo-cell table (global first slot, LDS offset) and, when it fits,
// the halo particles' positions.  List entries of the tile format are LDS
// offsets into that halo (16 bit), so the force kernel reads neighbour
// positions from LDS instead of gathering them through L2.
#pragma once

#include "rumdx_common.h"

namespace rumdx {

constexpr int kTileEdge = 4;                            // cells per tile edge (a power of two)
constexpr int kHaloEdgeMax = kTileEdge + 4;             // 2-cell stencil on each side
constexpr uint32_t kHaloCellsMax = kHaloEdgeMax * kHaloEdgeMax * kHaloEdgeMax;  // 512

struct TileGeom {
  int nc0, nc1, nc2;
  const uint32_t* cell_rank;   // linear cell -> rank
  const uint32_t* cell_start;  // rank -> first slot (ncells + 1)
  const uint32_t* first_rank;  // tile -> first rank (ntiles + 1)
  const uint32_t* origin;      // tile -> cx0 | cy0 << 10 | cz0 << 20
  uint32_t ntiles;
  uint32_t halo_cap;           // particles the LDS position buffer holds
  uint32_t* halo_slots;        // tile t's halo slots at [t * halo_cap, ...) (build output)
  uint4* tile_info;            // per tile: {halo particles, kTileFast | shifts | cap << 16, slot0, slot1}
};

// tile_info flags.  kTileFast: the build staged the halo (npart <= halo_cap)
// and every axis on which the halo wraps allows the shifted frame below.  Per axis a, (kTileShiftDown << 2a): the
// halo wraps below the tile (its far cells sit at x >= 0, near +L/2) and
// x >= 0 becomes x - L; (kTileShiftUp << 2a): it wraps above and x < 0
// becomes x + L.  Positions are in [-L/2, L/2).
constexpr uint32_t kTileFast = 1u;
constexpr uint32_t kTileShiftDown = 2u;
constexpr uint32_t kTileShiftUp = 4u;
static_assert(kTileFast == RUMDX_TILE_FAST, "tile flag bits of the header");
static_assert((((kTileShiftDown | kTileShiftUp) << 0) | ((kTileShiftDown | kTileShiftUp) << 2) |
               ((kTileShiftDown | kTileShiftUp) << 4)) == RUMDX_TILE_SHIFT_MASK,
              "tile shift bits of the header");

// The shifted frame puts a wrapping halo on one side of the periodic
// boundary, so plain FP32 differences there are the correctly rounded
// minimum-image displacements.  It needs at least kShiftMinCells cells on the
// axis (the tile's own side of the halo then stays clear of x = 0) and an
// exact shift: the moved coordinates lie within 2 cells of +-L/2 and move to
// within 2 cells beyond it, and x -+ L is exact when all of
// [L/2 - 2a, L/2 + 2a] lies in one binade (x is a multiple of that binade's
// ulp and L of twice it).
constexpr int kShiftMinCells = 16;

__device__ __forceinline__ bool shift_exact(float len, int nc) {
  const float half = 0.5f * len, reach = 2.05f * len / static_cast<float>(nc);
  return ilogbf(half - reach) == ilogbf(half + reach);
}

__device__ __forceinline__ float shift_coord(float x, uint32_t fl, int axis, float len) {
  const uint32_t f = fl >> (2 * axis);
  if ((f & kTileShiftDown) && x >= 0.f) return x - len;
  if ((f & kTileShiftUp) && x < 0.f) return x + len;
  return x;
}

// Halo of one tile, in LDS.
struct Halo {
  int base[3];   // first halo cell coordinate (may be negative: periodic)
  int ext[3];    // halo cells per axis (<= 8, <= nc)
  int own[3];    // the tile's own cells per axis (<= 4); they sit at halo index 2..2+own-1
  bool wraps;    // some halo cell lies across the periodic boundary
  uint32_t ncell;
  uint32_t npart;  // particles in the halo
  uint32_t slot0, slot1;  // the tile's own slots
};

struct HaloLds {
  uint32_t start[kHaloCellsMax];    // global first slot of each halo cell
  uint32_t off[kHaloCellsMax + 1];  // LDS offset of each halo cell's first particle
  uint32_t wave_sum[kMaxWorkgroup / 64];
  Halo h;
};

// Tile-format list entries come in groups of 8 u16 (16 bytes): entry k of
// slot i is u16 number k%8 of group (k/8)*nvp + i.  An entry is the LDS byte
// offset 16*h of halo particle h when the halo holds at most kByteEntryHalo
// particles (so the force kernel adds nothing to form the address), else h.
constexpr uint32_t kByteEntryHalo = 4096;
static_assert(RUMDX_HALO_CAP_MAX <= kByteEntryHalo, "a staged halo must be addressable by byte entries");

// Tiles of at most this many particles get a lane order (by list length).
constexpr uint32_t kOrderMax = 1024;

__device__ __forceinline__ uint32_t encode_entry(uint32_t h, bool bytes) { return bytes ? (h << 4) : h; }
__device__ __forceinline__ uint32_t decode_entry(uint32_t e, bool bytes) { return bytes ? (e >> 4) : e; }
// Halo position of a byte entry (s_pos: the staged halo positions).
__device__ __forceinline__ float4 halo_pos(const float4* s_pos, uint32_t entry) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_pos) + entry);
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t group_entry(const u32x4& q, int e) {
  const uint32_t w = (e >> 1) == 0 ? q.x : (e >> 1) == 1 ? q.y : (e >> 1) == 2 ? q.z : q.w;
  return (e & 1) ? (w >> 16) : (w & 0xffffu);
}

// Appends halo offset o as entry `cnt` of a list being built: packs it into
// the 16-byte group held in w, and stores the group once it is full.
__device__ __forceinline__ void group_append(u32x4& w, uint32_t cnt, uint32_t o, u32x4* __restrict__ grp,
                                             uint32_t nvp, uint32_t i) {
  const uint32_t e = cnt & 7u;
  const uint32_t v = (e & 1u) ? (o << 16) : o;
  const uint32_t sel = e >> 1;
  w.x |= sel == 0 ? v : 0u;
  w.y |= sel == 1 ? v : 0u;
  w.z |= sel == 2 ? v : 0u;
  w.w |= sel == 3 ? v : 0u;
  if (e == 7u) {
    grp[static_cast<size_t>(cnt >> 3) * nvp + i] = w;
    w = u32x4{0u, 0u, 0u, 0u};
  }
}

// Halo index of cell coordinate c along one axis (c inside the halo).
__device__ __forceinline__ int halo_axis(int c, int base, int nc) {
  int d = c - base;
  d += (d < 0) ? nc : 0;
  d -= (d >= nc) ? nc : 0;
  return d;
}

// Fills the halo table of tile t and, if it fits, copies the halo particles'
// positions from pos into s_pos.  Must be called by all threads of the block
// (at least 256, a multiple of 64, at most kMaxWorkgroup).  With cellof (LDS
// scratch of halo_cap u16) the copy is flat: each halo particle learns its
// cell from cellof and every thread issues up to four independent loads, and
// the halo's slots also go to slots_out (halo order) when that is given.
__device__ __forceinline__ void stage_halo(uint32_t t, const TileGeom& g, const float4* __restrict__ pos,
                                           float4* s_pos, HaloLds& L, uint16_t* cellof = nullptr,
                                           uint32_t* __restrict__ slots_out = nullptr) {
  if (threadIdx.x == 0) {
    const uint32_t o = g.origin[t];
    const int c0[3] = {static_cast<int>(o & 1023u), static_cast<int>((o >> 10) & 1023u),
                       static_cast<int>((o >> 20) & 1023u)};
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    bool wraps = false;
    uint32_t ncell = 1;
    for (int a = 0; a < 3; ++a) {
      const int bw = (nc[a] - c0[a]) < kTileEdge ? (nc[a] - c0[a]) : kTileEdge;
      const int ext = (bw + 4) < nc[a] ? (bw + 4) : nc[a];
      L.h.base[a] = c0[a] - 2;
      L.h.ext[a] = ext;
      L.h.own[a] = bw;
      wraps = wraps || (c0[a] - 2 < 0) || (c0[a] - 2 + ext > nc[a]);
      ncell *= static_cast<uint32_t>(ext);
    }
    L.h.wraps = wraps;
    L.h.ncell = ncell;
    L.h.slot0 = g.cell_start[g.first_rank[t]];
    L.h.slot1 = g.cell_start[g.first_rank[t + 1]];
  }
  __syncthreads();
  const Halo& h = L.h;
  // per-cell counts; two cells per thread cover 512
  uint32_t cnt[2] = {0u, 0u};
  for (uint32_t r = 0; r < 2; ++r) {
    const uint32_t c = threadIdx.x * 2 + r;
    if (c < h.ncell) {
      const int hx = static_cast<int>(c % h.ext[0]);
      const int hy = static_cast<int>((c / h.ext[0]) % h.ext[1]);
      const int hz = static_cast<int>(c / (h.ext[0] * h.ext[1]));
      int cx = h.base[0] + hx, cy = h.base[1] + hy, cz = h.base[2] + hz;
      cx += (cx < 0) ? g.nc0 : 0;
      cx -= (cx >= g.nc0) ? g.nc0 : 0;
      cy += (cy < 0) ? g.nc1 : 0;
      cy -= (cy >= g.nc1) ? g.nc1 : 0;
      cz += (cz < 0) ? g.nc2 : 0;
      cz -= (cz >= g.nc2) ? g.nc2 : 0;
      const uint32_t rank = g.cell_rank[cx + cy * g.nc0 + cz * g.nc0 * g.nc1];
      const uint32_t s = g.cell_start[rank];
      L.start[c] = s;
      cnt[r] = g.cell_start[rank + 1] - s;
    }
  }
  // exclusive scan of the counts (thread-pair sums, wave scan, wave totals)
  const uint32_t v = cnt[0] + cnt[1];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= static_cast<uint32_t>(d)) x += y;
  }
  if (lane == 63) L.wave_sum[wave] = x;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
  for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
    woff += (k < wave) ? L.wave_sum[k] : 0u;
    tot += L.wave_sum[k];
  }
  const uint32_t ex = woff + x - v;
  if (threadIdx.x * 2 < h.ncell) L.off[threadIdx.x * 2] = ex;
  if (threadIdx.x * 2 + 1 < h.ncell) L.off[threadIdx.x * 2 + 1] = ex + cnt[0];
  if (threadIdx.x == 0) {
    L.off[h.ncell] = tot;
    L.h.npart = tot;
  }
  __syncthreads();
  if (tot <= g.halo_cap && cellof != nullptr) {
    for (uint32_t c = threadIdx.x; c < h.ncell; c += blockDim.x) {
      const uint32_t o1 = L.off[c + 1];
      for (uint32_t q = L.off[c]; q < o1; ++q) cellof[q] = static_cast<uint16_t>(c);
    }
    __syncthreads();
    for (uint32_t q0 = 0; q0 < tot; q0 += 4 * blockDim.x) {
      uint32_t src[4];
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + u * blockDim.x + threadIdx.x;
        const uint32_t c = q < tot ? cellof[q] : 0u;
        // past the halo, slot 0 stands in: every load is unconditional, so
        // the four stay in registers and in flight together
        src[u] = q < tot ? L.start[c] + (q - L.off[c]) : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = pos[src[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + u * blockDim.x + threadIdx.x;
        if (q < tot) {
          s_pos[q] = v[u];
          if (slots_out != nullptr) slots_out[q] = src[u];
        }
      }
    }
  } else if (tot <= g.halo_cap) {
    // one thread per halo cell copies that cell's (contiguous) particles
    for (uint32_t c = threadIdx.x; c < h.ncell; c += blockDim.x) {
      const uint32_t s = L.start[c], o0 = L.off[c], n = L.off[c + 1] - o0;
      for (uint32_t k = 0; k < n; ++k) s_pos[o0 + k] = pos[s + k];
    }
  }
  __syncthreads();
}

// Halo cell index of the stencil cell (cx+dx, cy+dy, cz+dz).
__device__ __forceinline__ uint32_t halo_cell(const Halo& h, const TileGeom& g, int cx, int cy, int cz) {
  const int hx = halo_axis(cx, h.base[0], g.nc0);
  const int hy = halo_axis(cy, h.base[1], g.nc1);
  const int hz = halo_axis(cz, h.base[2], g.nc2);
  return static_cast<uint32_t>(hx + h.ext[0] * (hy + h.ext[1] * hz));
}

}  // namespace rumdx
