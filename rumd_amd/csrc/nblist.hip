// Neighbour-list kernels (gfx950): skin check, N2 build and the sorted
// (cell-binned, Morton-ordered) build.
//
// Reference: NeighborList::UpdateNBlist / CheckRebuildRequired
// (src/NeighborList.cu:569-690), displacement_since_nb_build (:970-997),
// calculateNBL_N2 (:1001-1099), CalculateNBL_Sorting (:717-781) with
// calculateCellIndices / calculateCellsSorted / calculateNBL_CellsSorted
// (:1142-1319), Sample::SortParticlesByKey (include/rumd/Sample.h:174-188) and
// ParticleData::UpdateAfterSorting (src/ParticleData.cu:493-528).
//
// Every build kernel first reads flags[REBUILD] and exits when it is clear, so a
// step can launch the whole chain unconditionally (no host round trip, capturable
// in a graph).  A one-workgroup kernel after the tile build clears the flag and
// records the build (the N2 build's last workgroup does it there).
//
// The sorted build is a deterministic counting sort: histogram of cell ranks,
// two-pass exclusive scan, scatter, then a gather that puts each cell's
// particles in ascending old-slot order (what a stable key sort gives).  The gather writes
// into scratch; the list kernel reads the sorted positions from there and copies
// every array back to its home buffer for its own slot, so buffer addresses
// never change (captured graphs stay valid).
#include <algorithm>
#include <utility>
#include <vector>

#include "rumdx_common.h"
#include "tiles.h"

namespace rumdx {
namespace {

struct SortScratch {
  uint32_t* key_lin;    // nvp: linear cell of each old slot
  uint32_t* perm;       // nvp: old slots grouped by cell (scatter output)
  float4* t_r;          // nvp each
  float4* t_v;
  float4* t_f;
  float4* t_w;
  float4* t_im;
  float4* t_rlo;
  uint32_t* t_idx;      // nvp
  uint32_t* chunk_sum;  // one per kScanChunk cells
  uint32_t* start_next; // ncells + 1: cell starts of the build in progress (the build copies them home)
  uint32_t* tile_max;   // per tile: longest list of the last build
  unsigned long long* lookback;  // per tile: the re-sort's published counts
  uint32_t* arrive;     // the tile build's arrival counters: kArriveGroups, then the total, one 128-byte line each
};

// The tile build counts its blocks' arrivals on kArriveGroups counters, block
// b on counter b % kArriveGroups, each on its own 128-byte line, and the last
// arrival of each group on a total: thousands of blocks on one address would
// serialise at the memory side (a short build of many small tiles waited on
// it).  The counters live in the zero-filled sort scratch and are left at zero.
constexpr uint32_t kArriveGroups = 8;
constexpr uint32_t kArriveStride = 32;  // u32 per 128-byte line

// Counts the calling block's arrival on the grouped counters `arrive`
// (thread 0 only); true for the block that arrives last in the launch.  The
// counters are left at zero for the next launch that uses them.
__device__ __forceinline__ bool grouped_arrival_last(uint32_t* arrive) {
  const uint32_t g = blockIdx.x % kArriveGroups;
  const uint32_t members = (gridDim.x - g + kArriveGroups - 1) / kArriveGroups;
  const uint32_t groups = gridDim.x < kArriveGroups ? gridDim.x : kArriveGroups;
  uint32_t* ctr = arrive + g * kArriveStride;
  if (__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != members - 1) return false;
  __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t* total = arrive + kArriveGroups * kArriveStride;
  if (__hip_atomic_fetch_add(total, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != groups - 1) return false;
  __hip_atomic_store(total, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Grid cap of the particle-parallel rebuild kernels (they loop over the rest).
constexpr uint32_t kChainGridCap = 1024;

// Cells per workgroup of the two-pass count scan.
constexpr uint32_t kScanPerThread = 16;
constexpr uint32_t kScanChunk = kBlock * kScanPerThread;

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

uint32_t scan_chunks(uint32_t ncells) { return (ncells + kScanChunk - 1) / kScanChunk; }

SortScratch carve(void* base, uint32_t nvp, uint32_t ncells) {
  char* p = static_cast<char*>(base);
  SortScratch s;
  const size_t u32n = align256(sizeof(uint32_t) * nvp);
  const size_t f4n = align256(sizeof(float4) * nvp);
  s.key_lin = reinterpret_cast<uint32_t*>(p); p += u32n;
  s.perm = reinterpret_cast<uint32_t*>(p); p += u32n;
  s.t_r = reinterpret_cast<float4*>(p); p += f4n;
  s.t_v = reinterpret_cast<float4*>(p); p += f4n;
  s.t_f = reinterpret_cast<float4*>(p); p += f4n;
  s.t_w = reinterpret_cast<float4*>(p); p += f4n;
  s.t_im = reinterpret_cast<float4*>(p); p += f4n;
  s.t_rlo = reinterpret_cast<float4*>(p); p += f4n;
  s.t_idx = reinterpret_cast<uint32_t*>(p); p += u32n;
  s.chunk_sum = reinterpret_cast<uint32_t*>(p); p += align256(sizeof(uint32_t) * scan_chunks(ncells));
  s.start_next = reinterpret_cast<uint32_t*>(p); p += align256(sizeof(uint32_t) * (ncells + 1ull));
  s.tile_max = reinterpret_cast<uint32_t*>(p); p += align256(sizeof(uint32_t) * ncells);  // ntiles <= ncells
  s.lookback = reinterpret_cast<unsigned long long*>(p); p += align256(sizeof(unsigned long long) * ncells);
  s.arrive = reinterpret_cast<uint32_t*>(p);
  return s;
}

size_t scratch_bytes(uint32_t nvp, uint32_t ncells) {
  const size_t u32n = align256(sizeof(uint32_t) * nvp);
  const size_t f4n = align256(sizeof(float4) * nvp);
  return 3 * u32n + 6 * f4n + align256(sizeof(uint32_t) * scan_chunks(ncells)) +
         align256(sizeof(uint32_t) * (ncells + 1ull)) + align256(sizeof(uint32_t) * ncells) +
         align256(sizeof(unsigned long long) * ncells) + align256(sizeof(uint32_t) * kArriveStride * (kArriveGroups + 1));
}

__device__ __forceinline__ bool rebuild_pending(const uint32_t* flags) {
  return __hip_atomic_load(flags + RUMDX_FLAG_REBUILD, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

// Called by every block of the final build kernel after its work: the block
// that finishes last clears the rebuild request and counts the build.
__device__ __forceinline__ void finish_build(uint32_t* flags) {
  if (last_to_arrive(flags + RUMDX_FLAG_DONE, gridDim.x) && threadIdx.x == 0) {
    flags[RUMDX_FLAG_REBUILD] = 0u;
    flags[RUMDX_FLAG_BUILDS] += 1u;
    flags[RUMDX_FLAG_EXCL_PENDING] = 1u;
    flags[RUMDX_FLAG_PRUNE] = 1u;
  }
}

// ------------------------------------------------------------------ skin check
__global__ __launch_bounds__(kBlock) void k_skin_check(const float4* __restrict__ r, const float4* __restrict__ last_r,
                                                       uint32_t* __restrict__ flags, float skin2_over_4, Box bx,
                                                       uint32_t n, const float4* __restrict__ prune_r, float inner2) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 ri = r[i];
  const float d2 = min_image_d2_f32(ri, last_r[i], bx);
  if (d2 > skin2_over_4) flags[RUMDX_FLAG_REBUILD] = 1u;
  if (prune_r != nullptr && min_image_d2_f32(ri, prune_r[i], bx) > inner2) flags[RUMDX_FLAG_PRUNE] = 1u;
}

__global__ void k_request_rebuild(uint32_t* flags) {
  if (threadIdx.x == 0) flags[RUMDX_FLAG_REBUILD] = 1u;
}

// ------------------------------------------------------------------ N2 build
// All pairs in slot order, positions staged through LDS.  List entries come out
// in ascending j (RUMD's atomicInc order is unspecified; only sets compare).
__global__ __launch_bounds__(kBlock) void k_build_n2(const float4* __restrict__ r, float4* __restrict__ last_r,
                                                     uint32_t* __restrict__ nbl, uint32_t* __restrict__ num_nbrs,
                                                     uint32_t* __restrict__ flags, const float* __restrict__ cut_skin2,
                                                     uint32_t nt, Box bx, uint32_t n, uint32_t nvp,
                                                     uint32_t max_nbrs, Exclusions ex) {
  if (!rebuild_pending(flags)) return;
  extern __shared__ float4 s_dyn[];
  float4* s_r = s_dyn;
  float* s_cut = reinterpret_cast<float*>(s_dyn + kBlock);
  for (uint32_t k = threadIdx.x; k < nt * nt; k += kBlock) s_cut[k] = cut_skin2[k];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const float4 ri = live ? r[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t ti = __float_as_uint(ri.w);
  const ExclusionRange er = live ? ex.of(i) : ExclusionRange{};
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < n; base += kBlock) {
    __syncthreads();
    if (base + threadIdx.x < n) s_r[threadIdx.x] = r[base + threadIdx.x];
    __syncthreads();
    const uint32_t m = (n - base) < kBlock ? (n - base) : kBlock;
    if (live) {
      for (uint32_t k = 0; k < m; ++k) {
        const uint32_t j = base + k;
        if (j == i) continue;
        const float4 rj = s_r[k];
        const float d2 = min_image_d2_f32(ri, rj, bx);
        if (d2 < s_cut[ti * nt + __float_as_uint(rj.w)] && !ex.excludes(er, j)) {
          if (cnt < max_nbrs) nbl[static_cast<size_t>(cnt) * nvp + i] = j;
          ++cnt;  // keeps counting past capacity: the true length sizes the regrowth
        }
      }
    }
  }
  if (live) {
    last_r[i] = ri;
    num_nbrs[i] = cnt < max_nbrs ? cnt : max_nbrs;  // == max_nbrs marks an overflowed list
    if (cnt >= max_nbrs) flags[RUMDX_FLAG_OVERFLOW] = 1u;
  }
  __shared__ uint32_t s_max[kBlock / 64];
  const uint32_t longest = block_max_u32(live ? cnt : 0u, s_max);
  if (threadIdx.x == 0) atomicMax(flags + RUMDX_FLAG_MAXLEN, longest);
  finish_build(flags);
}

// ------------------------------------------------------------------ sorted build
struct CellGeom {
  int nc0, nc1, nc2;
  uint32_t ncells;
};

__device__ __forceinline__ uint32_t cell_of(float4 rr, const Box& bx, const CellGeom& g) {
  const int cx = cell_coord(rr.x, bx.iL[0], g.nc0);
  const int cy = cell_coord(rr.y, bx.iL[1], g.nc1);
  const int cz = cell_coord(rr.z, bx.iL[2], g.nc2);
  return static_cast<uint32_t>(cx + cy * g.nc0 + cz * g.nc0 * g.nc1);
}

// Runs of equal keys among the live lanes of a wave (consecutive lanes with
// the same key).  Particles stay in cell order between builds, so a wave's
// 64 particles fall into a few cells: one atomic per run instead of one per
// particle, and no lanes of one instruction on the same address.  Returns the
// lane's index inside its run; `len` is the run's length and `head` its
// first lane.  Called by all lanes of the wave.
__device__ __forceinline__ uint32_t key_run(uint32_t key, bool live, uint32_t& len, uint32_t& head) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool prev_live = __shfl_up(live ? 1 : 0, 1, 64) != 0;
  const bool start = live && (lane == 0 || !prev_live || prev != key);
  const uint64_t starts = __builtin_amdgcn_ballot_w64(start);
  const uint64_t bounds = starts | __builtin_amdgcn_ballot_w64(!live);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
  head = 63u - static_cast<uint32_t>(__builtin_clzll((starts & upto) | 1ull));
  const uint64_t above = bounds & ~upto;
  const uint32_t end = above != 0ull ? static_cast<uint32_t>(__builtin_ctzll(above)) : 64u;
  len = end - head;
  return lane - head;
}

// calculateCellIndices (src/NeighborList.cu:1142-1172) plus the histogram.
__global__ __launch_bounds__(kBlock) void k_cell_hist(const float4* __restrict__ r, uint32_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ cell_rank,
                                                      uint32_t* __restrict__ key_lin,
                                                      uint32_t* __restrict__ cell_count, Box bx, CellGeom g,
                                                      uint32_t n) {
  if (!rebuild_pending(flags)) return;
  for (uint32_t b0 = blockIdx.x * kBlock; b0 < n; b0 += gridDim.x * kBlock) {
    const uint32_t i = b0 + threadIdx.x;
    const bool live = i < n;
    uint32_t rank = 0;
    if (live) {
      const uint32_t lin = cell_of(r[i], bx, g);
      key_lin[i] = lin;
      rank = cell_rank[lin];
    }
    uint32_t len, head;
    const uint32_t pos = key_run(rank, live, len, head);
    if (live && pos == 0) atomicAdd(cell_count + rank, len);
  }
}

// Total of v over a kBlock-thread workgroup, returned in every thread.
__device__ __forceinline__ uint32_t block_total_u32(uint32_t v, uint32_t* lds4) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63u) == 0) lds4[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t k = 0; k < kBlock / 64; ++k) t += lds4[k];
  __syncthreads();
  return t;
}

// Exclusive scan of the cell counts in two passes over kScanChunk-cell
// chunks: chunk totals, then each chunk adds the totals before it and scans
// itself.  cell_start[c] = sum of counts before c, cell_start[ncells] = n; the
// counts stay, as the scatter's per-cell cursors.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ cell_count,
                                                        const uint32_t* __restrict__ flags,
                                                        uint32_t* __restrict__ chunk_sum, uint32_t ncells) {
  if (!rebuild_pending(flags)) return;
  __shared__ uint32_t s_w[kBlock / 64];
  const uint32_t base = blockIdx.x * kScanChunk;
  uint32_t v = 0;
  for (uint32_t k = threadIdx.x; k < kScanChunk; k += kBlock) {
    const uint32_t c = base + k;
    v += c < ncells ? cell_count[c] : 0u;
  }
  v = block_total_u32(v, s_w);
  if (threadIdx.x == 0) chunk_sum[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ cell_count,
                                                       const uint32_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ chunk_sum,
                                                       uint32_t* __restrict__ cell_start, uint32_t ncells) {
  if (!rebuild_pending(flags)) return;
  __shared__ uint32_t s_w[kBlock / 64];
  uint32_t carry = 0;
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += kBlock) carry += chunk_sum[k];
  carry = block_total_u32(carry, s_w);

  const uint32_t c0 = blockIdx.x * kScanChunk + threadIdx.x * kScanPerThread;
  uint32_t c[kScanPerThread];
  const bool full = c0 + kScanPerThread <= ncells;
  if (full) {
    const uint4* src = reinterpret_cast<const uint4*>(cell_count + c0);
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread / 4; ++q) {
      const uint4 x = src[q];
      c[4 * q] = x.x;
      c[4 * q + 1] = x.y;
      c[4 * q + 2] = x.z;
      c[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanPerThread; ++k) c[k] = c0 + k < ncells ? cell_count[c0 + k] : 0u;
  }
  uint32_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) v += c[k];
  // exclusive scan of the per-thread sums: inside each wave, then wave totals
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= static_cast<uint32_t>(o)) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t off = carry, tot = 0;
  for (uint32_t k = 0; k < kBlock / 64; ++k) {
    off += (k < wave) ? s_w[k] : 0u;
    tot += s_w[k];
  }
  off += x - v;
  uint32_t st[kScanPerThread];
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    st[k] = off;
    off += c[k];
  }
  if (full) {
    uint4* ds = reinterpret_cast<uint4*>(cell_start + c0);
#pragma unroll
    for (uint32_t q = 0; q < kScanPerThread / 4; ++q)
      ds[q] = make_uint4(st[4 * q], st[4 * q + 1], st[4 * q + 2], st[4 * q + 3]);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kScanPerThread; ++k)
      if (c0 + k < ncells) cell_start[c0 + k] = st[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) cell_start[ncells] = carry + tot;
}

// Places every old slot in its cell's slot range, counting the cell's cursor
// down so the counts are zero again for the next build.  The order inside a
// cell is whatever the atomics give; the gather restores it.
__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t* __restrict__ key_lin,
                                                    const uint32_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ cell_rank,
                                                    const uint32_t* __restrict__ cell_start,
                                                    uint32_t* __restrict__ cell_count, uint32_t* __restrict__ perm,
                                                    uint32_t n) {
  if (!rebuild_pending(flags)) return;
  for (uint32_t b0 = blockIdx.x * kBlock; b0 < n; b0 += gridDim.x * kBlock) {
    const uint32_t i = b0 + threadIdx.x;
    const bool live = i < n;
    const uint32_t rank = live ? cell_rank[key_lin[i]] : 0u;
    uint32_t len, head;
    const uint32_t pos = key_run(rank, live, len, head);
    // the run's first lane takes len cursor values at once
    uint32_t top = 0;
    if (live && pos == 0) top = atomicSub(cell_count + rank, len);
    top = __shfl(top, static_cast<int>(head), 64);
    const uint32_t slot = cell_start[rank] + top - len + pos;
    if (live && slot < n) perm[slot] = i;  // always in range when hist and scatter saw the same keys
  }
}

// Gathers the particle arrays into scratch in sorted order.  Thread k takes
// the old slot x the scatter put at k; its new slot is the cell's first slot
// plus the number of cell members with a smaller old slot, i.e. cells keep
// ascending old-slot order (what a stable key sort gives).
__global__ __launch_bounds__(kBlock) void k_gather(const uint32_t* __restrict__ flags,
                                                   const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ key_lin,
                                                   const uint32_t* __restrict__ cell_rank,
                                                   const uint32_t* __restrict__ cell_start,
                                                   const float4* __restrict__ r,
                                                   const float4* __restrict__ v, const float4* __restrict__ f,
                                                   const float4* __restrict__ w, const float4* __restrict__ im,
                                                   const float4* __restrict__ rlo, const uint32_t* __restrict__ idx,
                                                   SortScratch s, uint32_t n) {
  if (!rebuild_pending(flags)) return;
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    const uint32_t x = perm[k];
    const uint32_t rank = cell_rank[key_lin[x]];
    const uint32_t b0 = cell_start[rank], e0 = cell_start[rank + 1];
    uint32_t d = b0;
    for (uint32_t j = b0; j < e0; ++j) d += perm[j] < x ? 1u : 0u;
    s.t_r[d] = r[x];
    s.t_v[d] = v[x];
    s.t_f[d] = f[x];
    s.t_w[d] = w[x];
    s.t_im[d] = im[x];
    if (rlo != nullptr) s.t_rlo[d] = rlo[x];
    s.t_idx[d] = idx[x];
  }
}

struct HomeArrays {
  float4* r;
  float4* v;
  float4* f;
  float4* w;
  float4* im;
  float4* rlo;
  uint32_t* idx;
  float4* last_r;
  uint32_t* slot_cell;
  uint16_t* lane_order;
  uint32_t* cell_start;  // the build copies its cells' starts here from start_next
};

// ------------------------------------------------------------------ incremental re-sort
// Between skin-triggered rebuilds particles stay in the cell order of the last
// build and move less than a cell, so the particles of a tile's cells after
// the move all sat in those cells or their face/edge/corner neighbours before
// it.  One workgroup per tile gathers those candidates, keeps the ones whose
// new cell is its own, orders them by (cell, old slot) -- the order the
// global counting sort gives -- and places them after the particles of all
// tiles before it.  That offset comes from a decoupled look-back in tile
// (= slot) order: each tile publishes its count as soon as it has it, then
// adds its predecessors' counts until it meets one that already holds the
// running total.  Published words are relaxed agent-scope atomics carrying
// their own payload, so no fences are needed; the data the tiles write is
// read by the next kernel.
//
// Look-back word of a tile: count (bits 0-31), build epoch (bits 32-61, so
// words of earlier builds are stale and the array is never cleared), state
// (bits 62-63: 1 = the tile's own count, 2 = the count of every tile up to
// and including it).
constexpr unsigned long long kLbOwn = 1ull, kLbInclusive = 2ull;
constexpr uint32_t kLbEpochMask = 0x3fffffffu;
constexpr uint32_t kLookbackSpin = 1u << 20;  // polls of one predecessor's word before giving up
constexpr uint32_t kResortCap = RUMDX_RESORT_TILE_MAX;
constexpr uint32_t kTileCells = static_cast<uint32_t>(kTileEdge * kTileEdge * kTileEdge);
constexpr uint32_t kDilCellsMax = static_cast<uint32_t>((kTileEdge + 2) * (kTileEdge + 2) * (kTileEdge + 2));
static_assert(kDilCellsMax <= kBlock && kTileCells <= 64, "one thread per dilated cell, one lane per tile cell");
static_assert(kResortCap <= 65536, "member indices are 16 bit");

// Sum of a u32 over the 64 lanes of a wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ unsigned long long lb_pack(unsigned long long state, uint32_t epoch, uint32_t count) {
  return (state << 62) | (static_cast<unsigned long long>(epoch & kLbEpochMask) << 32) | count;
}

struct ParticleArrays {
  const float4* r;
  const float4* v;
  const float4* f;
  const float4* w;
  const float4* im;
  const float4* rlo;
  const uint32_t* idx;
};

struct ResortLds {
  uint32_t start[kDilCellsMax];    // old first slot of each cell of the dilated tile
  uint32_t off[kDilCellsMax + 1];  // first candidate of each of those cells
  uint32_t wave_sum[kBlock / 64];
  uint32_t cnt[kTileCells];        // members per cell of the tile (tile-local rank)
  uint32_t coff[kTileCells + 1];   // their exclusive prefix
  uint32_t cur[kTileCells];
  uint32_t nmem, prefix, rank0, ncell_t;
  int c0[3], own[3], ext[3];
  uint32_t slot[kResortCap];       // old slot of each member
  uint8_t lr[kResortCap];          // tile-local cell of each member
  uint16_t order[kResortCap];      // members grouped by cell
};

__global__ __launch_bounds__(kBlock) void k_tile_resort(uint32_t* __restrict__ flags, SortScratch s, TileGeom tg,
                                                        Box bx, ParticleArrays o, uint32_t n) {
  if (!rebuild_pending(flags)) return;
  __shared__ ResortLds L;
  // tiles in slot order, in dispatch order: a tile waits only on earlier ones
  const uint32_t t = blockIdx.x;
  const uint32_t epoch = flags[RUMDX_FLAG_BUILDS];  // unchanged during the launch
  const uint32_t tid = threadIdx.x;
  const int nc[3] = {tg.nc0, tg.nc1, tg.nc2};
  if (tid == 0) {
    const uint32_t org = tg.origin[t];
    const int c0[3] = {static_cast<int>(org & 1023u), static_cast<int>((org >> 10) & 1023u),
                       static_cast<int>((org >> 20) & 1023u)};
    for (int a = 0; a < 3; ++a) {
      const int own = (nc[a] - c0[a]) < kTileEdge ? (nc[a] - c0[a]) : kTileEdge;
      L.c0[a] = c0[a];
      L.own[a] = own;
      L.ext[a] = (own + 2) < nc[a] ? (own + 2) : nc[a];
    }
    L.rank0 = tg.first_rank[t];
    L.ncell_t = tg.first_rank[t + 1] - L.rank0;
    L.nmem = 0u;
  }
  if (tid < kTileCells) {
    L.cnt[tid] = 0u;
    L.cur[tid] = 0u;
  }
  __syncthreads();
  // old slot ranges of the tile's cells dilated by one cell (each cell once)
  const uint32_t ndil = static_cast<uint32_t>(L.ext[0] * L.ext[1] * L.ext[2]);
  uint32_t cnt = 0;
  if (tid < ndil) {
    const int h[3] = {static_cast<int>(tid % L.ext[0]), static_cast<int>((tid / L.ext[0]) % L.ext[1]),
                      static_cast<int>(tid / (L.ext[0] * L.ext[1]))};
    int c[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = L.c0[a] - 1 + h[a];
      c[a] += c[a] < 0 ? nc[a] : 0;
      c[a] -= c[a] >= nc[a] ? nc[a] : 0;
    }
    const uint32_t rank = tg.cell_rank[c[0] + c[1] * nc[0] + c[2] * nc[0] * nc[1]];
    const uint32_t st = tg.cell_start[rank];
    L.start[tid] = st;
    cnt = tg.cell_start[rank + 1] - st;
  }
  {
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint32_t x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= static_cast<uint32_t>(d)) x += y;
    }
    if (lane == 63) L.wave_sum[wave] = x;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t k = 0; k < wave; ++k) woff += L.wave_sum[k];
    if (tid < ndil) L.off[tid] = woff + x - cnt;
    if (tid == ndil - 1) L.off[ndil] = woff + x;
  }
  __syncthreads();
  // candidates: members are those whose new cell is one of the tile's
  const uint32_t tot = L.off[ndil];
  for (uint32_t q = tid; q < tot; q += kBlock) {
    uint32_t lo = 0, hi = ndil;  // last dilated cell with off <= q
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (L.off[mid] <= q) lo = mid; else hi = mid;
    }
    const uint32_t slot = L.start[lo] + (q - L.off[lo]);
    if (slot >= n) continue;  // never, for cell starts left by a sorted build of these particles
    const float4 rr = o.r[slot];
    const int cx = cell_coord(rr.x, bx.iL[0], nc[0]);
    const int cy = cell_coord(rr.y, bx.iL[1], nc[1]);
    const int cz = cell_coord(rr.z, bx.iL[2], nc[2]);
    int dx = cx - L.c0[0], dy = cy - L.c0[1], dz = cz - L.c0[2];
    dx += dx < 0 ? nc[0] : 0;
    dy += dy < 0 ? nc[1] : 0;
    dz += dz < 0 ? nc[2] : 0;
    if (dx < L.own[0] && dy < L.own[1] && dz < L.own[2]) {
      const uint32_t lc = tg.cell_rank[cx + cy * nc[0] + cz * nc[0] * nc[1]] - L.rank0;
      if (lc >= L.ncell_t) continue;  // never: the tile's cells are its rank range
      const uint32_t k = atomicAdd(&L.nmem, 1u);
      if (k < kResortCap) {
        L.slot[k] = slot;
        L.lr[k] = static_cast<uint8_t>(lc);
      }
      atomicAdd(&L.cnt[lc], 1u);
    }
  }
  __syncthreads();
  const uint32_t nmem = L.nmem;
  const bool fits = nmem <= kResortCap;
  if (tid == 0) __hip_atomic_store(s.lookback + t, lb_pack(kLbOwn, epoch, nmem), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (tid < 64) {  // exclusive prefix of the per-cell counts (one wave)
    const uint32_t v = tid < L.ncell_t ? L.cnt[tid] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (tid >= static_cast<uint32_t>(d)) x += y;
    }
    if (tid <= kTileCells) L.coff[tid] = x - v;
    if (tid == 63) L.coff[kTileCells] = x;
  }
  __syncthreads();
  if (fits)
    for (uint32_t k = tid; k < nmem; k += kBlock) {
      const uint32_t c = L.lr[k];
      L.order[L.coff[c] + atomicAdd(&L.cur[c], 1u)] = static_cast<uint16_t>(k);
    }
  if (tid < 64) {
    // this tile's first slot: the particles of every tile before it.  The
    // first wave reads 64 predecessors' words at a time, nearest first, until
    // one holds an inclusive count; the sum runs up to that lane.
    const uint32_t lane = tid;
    uint32_t acc = 0;
    bool ok = true;
    for (int32_t base = static_cast<int32_t>(t); base > 0; base -= 64) {
      const int32_t j = base - 1 - static_cast<int32_t>(lane);
      unsigned long long w = 0;
      bool ready = j < 0;
      for (uint32_t spin = 0;; ++spin) {
        if (!ready) {
          w = __hip_atomic_load(s.lookback + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ready = (w >> 62) != 0ull && (static_cast<uint32_t>(w >> 32) & kLbEpochMask) == (epoch & kLbEpochMask);
        }
        if (__builtin_amdgcn_ballot_w64(!ready) == 0ull) break;
        if (spin == kLookbackSpin) {
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (!ok) break;
      const uint32_t count = j >= 0 ? static_cast<uint32_t>(w) : 0u;
      const uint64_t incl = __builtin_amdgcn_ballot_w64(j >= 0 && (w >> 62) == kLbInclusive);
      const uint32_t upto = incl != 0ull ? static_cast<uint32_t>(__builtin_ctzll(incl)) : 63u;
      acc += wave_sum_u32(lane <= upto ? count : 0u);
      if (incl != 0ull) break;
    }
    if (lane == 0) {
      if (!ok) atomicOr(flags + RUMDX_FLAG_ERROR, RUMDX_BUILD_ERR_WAIT);
      __hip_atomic_store(s.lookback + t, lb_pack(kLbInclusive, epoch, acc + nmem), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      L.prefix = acc;
      if (!fits) atomicOr(flags + RUMDX_FLAG_ERROR, RUMDX_BUILD_ERR_TILE);
      // every particle is a member of exactly one tile unless one moved more than a cell
      if (t == tg.ntiles - 1 && acc + nmem != n) atomicOr(flags + RUMDX_FLAG_ERROR, RUMDX_BUILD_ERR_MOVED);
    }
  }
  __syncthreads();
  const uint32_t S = L.prefix;
  for (uint32_t c = tid; c < L.ncell_t; c += kBlock) s.start_next[L.rank0 + c] = S + L.coff[c];
  if (t == tg.ntiles - 1 && tid == 0) s.start_next[L.rank0 + L.ncell_t] = S + nmem;
  if (!fits) return;
  // new slot: the cell's first slot plus the members of the cell with a smaller old slot
  for (uint32_t j = tid; j < nmem; j += kBlock) {
    const uint32_t k = L.order[j];
    const uint32_t c = L.lr[k], my = L.slot[k];
    const uint32_t b = L.coff[c], e = L.coff[c + 1];
    uint32_t rk = 0;
    for (uint32_t jj = b; jj < e; ++jj) rk += L.slot[L.order[jj]] < my ? 1u : 0u;
    const uint32_t d = S + b + rk;
    if (d < n) {  // in range unless particles were lost (then the build is flagged)
      s.t_r[d] = o.r[my];
      s.t_v[d] = o.v[my];
      s.t_f[d] = o.f[my];
      s.t_w[d] = o.w[my];
      s.t_im[d] = o.im[my];
      if (o.rlo != nullptr) s.t_rlo[d] = o.rlo[my];
      s.t_idx[d] = o.idx[my];
    }
  }
}

// The statistics of the last sorted build from its per-tile records: longest
// list (kept as the maximum with the old value), largest halo and largest
// tile.  One workgroup, launched on request between steps rather than after
// every build, so captured steps carry no launch for it.
__global__ __launch_bounds__(kBlock) void k_build_stats(uint32_t* __restrict__ flags,
                                                        const uint4* __restrict__ tile_info,
                                                        const uint32_t* __restrict__ tile_max, uint32_t ntiles) {
  if (flags[RUMDX_FLAG_STATS] == 0u) return;
  __shared__ uint32_t s_max[kBlock / 64];
  uint32_t ml = 0, mh = 0, mt = 0;
  for (uint32_t t = threadIdx.x; t < ntiles; t += kBlock) {
    const uint4 info = tile_info[t];
    ml = tile_max[t] > ml ? tile_max[t] : ml;
    mh = info.x > mh ? info.x : mh;
    mt = (info.w - info.z) > mt ? (info.w - info.z) : mt;
  }
  ml = block_max_u32(ml, s_max);
  mh = block_max_u32(mh, s_max);
  mt = block_max_u32(mt, s_max);
  if (threadIdx.x == 0) {
    flags[RUMDX_FLAG_MAXLEN] = ml > flags[RUMDX_FLAG_MAXLEN] ? ml : flags[RUMDX_FLAG_MAXLEN];
    flags[RUMDX_FLAG_MAXHALO] = mh;
    flags[RUMDX_FLAG_MAXTILE] = mt;
    flags[RUMDX_FLAG_STATS] = 0u;
  }
}

// Per-wave LDS of the cooperative list build: the stencil segments of the
// cell being built, a window of its candidates (LDS offsets, in stencil
// order) and the staged lists of the two particles tested together.  A staged
// list holds as many entries as the window, so appends need no bounds check.
constexpr uint32_t kCandWindow = 512;
constexpr uint32_t kStageMax = 256;  // longest list the staged path writes (max_nbrs bound)
struct WaveBuild {
  uint16_t cand[kCandWindow];
  uint16_t seg_start[128];  // LDS offset of the segment's first particle
  uint16_t seg_pre[128];    // candidates before the segment
  uint16_t seg_len[128];
  __attribute__((aligned(16))) uint16_t list[2][kCandWindow];
};

// Orders the LDS accesses of one wave (lanes read what other lanes wrote).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Appends the accepted candidates of one 64-lane chunk to the list of slot i,
// which holds cnt entries so far; returns the new count (uniform).
__device__ __forceinline__ uint32_t append_chunk(bool acc, uint32_t entry, uint32_t cnt, uint16_t* __restrict__ nbl16,
                                                 uint32_t i, uint32_t nvp, uint32_t max_nbrs) {
  const uint64_t mask = __ballot(acc);
  const uint32_t k = cnt + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
  if (acc && k < max_nbrs) nbl16[(static_cast<size_t>(k >> 3) * nvp + i) * 8 + (k & 7u)] = static_cast<uint16_t>(entry);
  return cnt + static_cast<uint32_t>(__popcll(mask));
}

// One particle's test of one chunk of candidates: the accepted ones go to its
// staged list in candidate order (ballot + lane count); returns the new count.
// The loop is short of scalar issue slots as much as of vector ones: the two
// compares give their lane masks directly (one scalar AND, no vector select)
// and the append is a plain exec-masked store.
template <bool WRAPS>
__device__ __forceinline__ uint32_t test_chunk(float4 pi, float4 rj, uint32_t e, float cut, uint32_t self,
                                               const Box& bx, uint16_t* list, uint32_t cnt) {
  const float d2 = WRAPS ? min_image_d2_f32(pi, rj, bx) : plain_d2_f32(pi, rj);
  const bool near = d2 < cut, other = e != self;
  const uint64_t mask = __builtin_amdgcn_ballot_w64(near) & __builtin_amdgcn_ballot_w64(other);
  const uint32_t k = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                               __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
  uint16_t* at = list + cnt;  // cnt is uniform: the base stays scalar
  if (near && other) at[k] = static_cast<uint16_t>(e);
  return cnt + static_cast<uint32_t>(__builtin_popcountll(mask));
}

// Tests N chunks of 64 window candidates (from q0) against particle a and,
// when TWO, particle b, appending the accepted ones to their staged lists.
// The chunks' window and position reads are issued together and shared by
// both particles.
template <bool WRAPS, bool TWO, int N>
__device__ __forceinline__ void scan_chunks(const WaveBuild& wb, uint16_t* list_a, uint16_t* list_b,
                                            const float4* s_pos, float4 pa, float4 pb, const float* crow_a,
                                            const float* crow_b, uint32_t self_a, uint32_t self_b, const Box& bx,
                                            uint32_t q0, uint32_t& cnt_a, uint32_t& cnt_b) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t e[N];
  float4 rj[N];
  float cut_a[N], cut_b[N];
#pragma unroll
  for (int u = 0; u < N; ++u) e[u] = wb.cand[q0 + 64u * u + lane];
#pragma unroll
  for (int u = 0; u < N; ++u) rj[u] = halo_pos(s_pos, e[u]);
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const uint32_t tj = __float_as_uint(rj[u].w);
    cut_a[u] = crow_a[tj];
    if (TWO) cut_b[u] = crow_b[tj];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) {
    cnt_a = test_chunk<WRAPS>(pa, rj[u], e[u], cut_a[u], self_a, bx, list_a, cnt_a);
    if (TWO) cnt_b = test_chunk<WRAPS>(pb, rj[u], e[u], cut_b[u], self_b, bx, list_b, cnt_b);
  }
}

// Scans the whole padded window for particle a (and b when TWO).
template <bool WRAPS, bool TWO>
__device__ __forceinline__ void scan_window(const WaveBuild& wb, uint16_t* list_a, uint16_t* list_b,
                                            const float4* s_pos, float4 pa, float4 pb, const float* crow_a,
                                            const float* crow_b, uint32_t self_a, uint32_t self_b, const Box& bx,
                                            uint32_t padded, uint32_t& cnt_a, uint32_t& cnt_b) {
  const uint32_t full = padded & ~255u;  // passes of four chunks, then the rest
  uint32_t q0 = 0;
  for (; q0 < full; q0 += 256)
    scan_chunks<WRAPS, TWO, 4>(wb, list_a, list_b, s_pos, pa, pb, crow_a, crow_b, self_a, self_b, bx, q0, cnt_a, cnt_b);
  if (padded - q0 >= 128) {
    scan_chunks<WRAPS, TWO, 2>(wb, list_a, list_b, s_pos, pa, pb, crow_a, crow_b, self_a, self_b, bx, q0, cnt_a, cnt_b);
    q0 += 128;
  }
  if (padded - q0 >= 64)
    scan_chunks<WRAPS, TWO, 1>(wb, list_a, list_b, s_pos, pa, pb, crow_a, crow_b, self_a, self_b, bx, q0, cnt_a, cnt_b);
}

// Writes one staged list to slot i as whole 16-byte groups (group g at
// nbl16[(g nvp + i) 8]); entries of the last group past the count are never
// read.
__device__ __forceinline__ void store_staged(const uint16_t* list, uint32_t cnt, uint32_t max_nbrs,
                                             uint16_t* __restrict__ nbl16, uint32_t nvp, uint32_t i, uint32_t lane) {
  const uint32_t stored = cnt < max_nbrs ? cnt : max_nbrs;
  if (lane * 8u < stored) {
    const u32x4 g = *reinterpret_cast<const u32x4*>(list + lane * 8u);
    reinterpret_cast<u32x4*>(nbl16)[static_cast<size_t>(lane) * nvp + i] = g;
  }
}

// The cooperative build of one cell when its whole stencil fits one window
// and a list fits the wave's staging buffer: the window holds the
// candidates' LDS byte offsets (the list entries themselves), padded to a
// multiple of 64 with the halo's sentinel (a position at +inf, never
// accepted), so chunks read without bounds checks.  The cell's particles are
// tested two at a time against each chunk; each list is staged in LDS and
// written as whole 16-byte groups.
template <bool WRAPS>
__device__ __forceinline__ void build_cell_staged(const float4* s_pos, const float* s_cut,
                                                  uint32_t nt, const Box& bx, uint32_t coff, uint32_t m,
                                                  uint32_t gstart, uint32_t nseg, uint32_t total, uint32_t sentinel,
                                                  WaveBuild& wb, uint32_t* s_cnt, uint16_t* __restrict__ nbl16,
                                                  uint32_t nvp, uint32_t max_nbrs) {
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t sg = lane; sg < nseg; sg += 64) {
    const uint32_t pre = wb.seg_pre[sg], len = wb.seg_len[sg], st = wb.seg_start[sg];
    for (uint32_t q = 0; q < len; ++q) wb.cand[pre + q] = static_cast<uint16_t>((st + q) << 4);
  }
  const uint32_t padded = (total + 63u) & ~63u;
  if (total + lane < padded) wb.cand[total + lane] = static_cast<uint16_t>(sentinel);
  wave_lds_sync();
  for (uint32_t p = 0; p < m; p += 2) {
    const bool two = p + 1 < m;
    const uint32_t pa_off = coff + p, pb_off = two ? pa_off + 1 : pa_off;
    const float4 pa = s_pos[pa_off], pb = s_pos[pb_off];
    const float* crow_a = s_cut + __float_as_uint(pa.w) * nt;
    const float* crow_b = s_cut + __float_as_uint(pb.w) * nt;
    uint32_t cnt_a = 0, cnt_b = 0;
    if (two)
      scan_window<WRAPS, true>(wb, wb.list[0], wb.list[1], s_pos, pa, pb, crow_a, crow_b, pa_off << 4, pb_off << 4, bx,
                               padded, cnt_a, cnt_b);
    else
      scan_window<WRAPS, false>(wb, wb.list[0], wb.list[1], s_pos, pa, pb, crow_a, crow_b, pa_off << 4, pb_off << 4,
                                bx, padded, cnt_a, cnt_b);
    wave_lds_sync();
    store_staged(wb.list[0], cnt_a, max_nbrs, nbl16, nvp, gstart + p, lane);
    if (two) store_staged(wb.list[1], cnt_b, max_nbrs, nbl16, nvp, gstart + p + 1, lane);
    if (lane == 0) {
      s_cnt[p] = cnt_a;
      if (two) s_cnt[p + 1] = cnt_b;
    }
    wave_lds_sync();  // the group stores have read the lists
  }
}


// Lists of one cell of a staged tile, built by one wave: lanes enumerate the
// cell's stencil candidates (the 125 cells, or 25 rows of 5 consecutive halo
// cells, in RUMD's (dZ, dY, dX) order, ascending slot inside each cell) and
// test them against one particle at a time; a ballot places the accepted
// entries in candidate order.  Writes the lists and leaves each particle's
// full count (which may exceed max_nbrs) in s_cnt[slot - slot0].
template <bool WRAPS>
__device__ __forceinline__ void build_cell_wave(const Halo& h, const HaloLds& L, const TileGeom& tg,
                                                const float4* s_pos, const float* s_cut, uint32_t nt, const Box& bx,
                                                int cx, int cy, int cz, WaveBuild& wb, uint32_t* s_cnt,
                                                uint16_t* __restrict__ nbl16, uint32_t nvp, uint32_t max_nbrs,
                                                bool bytes) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t hc = halo_cell(h, tg, cx, cy, cz);
  // wave-uniform values, made scalar so the loops below stay on full waves
  const uint32_t coff = __builtin_amdgcn_readfirstlane(L.off[hc]);
  const uint32_t m = __builtin_amdgcn_readfirstlane(L.off[hc + 1]) - coff;
  const uint32_t gstart = __builtin_amdgcn_readfirstlane(L.start[hc]);
  if (m == 0) return;
  wave_lds_sync();  // the previous cell's tables have been read
  const int hx = halo_axis(cx, h.base[0], tg.nc0);
  const bool rows = hx >= 2 && hx + 2 < h.ext[0];
  const uint32_t nseg = rows ? 25u : 125u;
  // segment table and its exclusive prefix (two 64-lane passes)
  uint32_t carry = 0;
  for (uint32_t s0 = 0; s0 < nseg; s0 += 64) {
    const uint32_t sg = s0 + lane;
    uint32_t st = 0, len = 0;
    if (sg < nseg) {
      uint32_t hcell;
      if (rows) {
        const int hz = halo_axis(cz + static_cast<int>(sg / 5) - 2, h.base[2], tg.nc2);
        const int hy = halo_axis(cy + static_cast<int>(sg % 5) - 2, h.base[1], tg.nc1);
        hcell = static_cast<uint32_t>(hx - 2 + h.ext[0] * (hy + h.ext[1] * hz));
        st = L.off[hcell];
        len = L.off[hcell + 5] - st;
      } else {
        hcell = halo_cell(h, tg, cx + static_cast<int>(sg % 5) - 2, cy + static_cast<int>((sg / 5) % 5) - 2,
                          cz + static_cast<int>(sg / 25) - 2);
        st = L.off[hcell];
        len = L.off[hcell + 1] - st;
      }
    }
    uint32_t x = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= static_cast<uint32_t>(o)) x += y;
    }
    if (sg < nseg) {
      wb.seg_start[sg] = static_cast<uint16_t>(st);
      wb.seg_pre[sg] = static_cast<uint16_t>(carry + x - len);
      wb.seg_len[sg] = static_cast<uint16_t>(len);
    }
    carry += __shfl(x, 63, 64);
  }
  const uint32_t total = __builtin_amdgcn_readfirstlane(carry);
  wave_lds_sync();
  if (bytes && total <= kCandWindow && max_nbrs <= kStageMax) {
    build_cell_staged<WRAPS>(s_pos, s_cut, nt, bx, coff, m, gstart, nseg, total, tg.halo_cap << 4, wb,
                             s_cnt + (gstart - h.slot0), nbl16, nvp, max_nbrs);
    return;
  }
  for (uint32_t pg = 0; pg < m; pg += 64) {
    const uint32_t mg = (m - pg) < 64u ? (m - pg) : 64u;
    uint32_t cnt_v = 0;  // lane l: count of particle pg + l
    for (uint32_t w0 = 0; w0 < total; w0 += kCandWindow) {
      const uint32_t wn = (total - w0) < kCandWindow ? (total - w0) : kCandWindow;
      wave_lds_sync();  // the previous window has been read
      for (uint32_t sg = lane; sg < nseg; sg += 64) {
        const uint32_t pre = wb.seg_pre[sg], len = wb.seg_len[sg], st = wb.seg_start[sg];
        const uint32_t lo = pre > w0 ? pre : w0;
        const uint32_t hi = (pre + len) < (w0 + wn) ? (pre + len) : (w0 + wn);
        for (uint32_t q = lo; q < hi; ++q) wb.cand[q - w0] = static_cast<uint16_t>(st + (q - pre));
      }
      wave_lds_sync();
      for (uint32_t p = 0; p < mg; ++p) {
        const uint32_t po = coff + pg + p;
        const float4 pi = s_pos[po];
        const float* crow = s_cut + __float_as_uint(pi.w) * nt;
        const uint32_t i = gstart + pg + p;
        uint32_t cnt = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cnt_v), static_cast<int>(p)));
        // four 64-candidate chunks per pass: their LDS reads are issued
        // together, so one round trip serves 256 candidates
        for (uint32_t q0 = 0; q0 < wn; q0 += 4 * 64) {
          uint32_t oo[4];
          float4 rj[4];
          float cut[4];
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t q = q0 + 64 * u + lane;
            oo[u] = q < wn ? static_cast<uint32_t>(wb.cand[q]) : po;
          }
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u) rj[u] = s_pos[oo[u]];
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u) cut[u] = crow[__float_as_uint(rj[u].w)];
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u) {
            const float d2 = WRAPS ? min_image_d2_f32(pi, rj[u], bx) : plain_d2_f32(pi, rj[u]);
            const bool acc = oo[u] != po && d2 < cut[u];
            cnt = append_chunk(acc, encode_entry(oo[u], bytes), cnt, nbl16, i, nvp, max_nbrs);
          }
        }
        cnt_v = lane == p ? cnt : cnt_v;
      }
    }
    if (lane < mg) s_cnt[gstart + pg + lane - h.slot0] = cnt_v;
  }
}

// Lists of one particle read from global memory (tiles whose halo does not
// fit in LDS): the same stencil order, one lane per particle.
__device__ __forceinline__ uint32_t build_particle_lane(uint32_t i, float4 ri, int cx, int cy, int cz, const Halo& h,
                                                        const HaloLds& L, const TileGeom& tg, const float4* t_r,
                                                        const float* crow, const Box& bx, u32x4* grp, uint32_t nvp,
                                                        uint32_t max_nbrs, bool bytes) {
  const uint32_t hc = halo_cell(h, tg, cx, cy, cz);
  const uint32_t my_off = L.off[hc] + (i - L.start[hc]);
  u32x4 w = {0u, 0u, 0u, 0u};
  uint32_t cnt = 0;
  for (int dz = -2; dz <= 2; ++dz)
    for (int dy = -2; dy <= 2; ++dy)
      for (int dx = -2; dx <= 2; ++dx) {
        const uint32_t hcell = halo_cell(h, tg, cx + dx, cy + dy, cz + dz);
        const uint32_t o0 = L.off[hcell], o1 = L.off[hcell + 1], g0 = L.start[hcell];
        for (uint32_t oo = o0; oo < o1; ++oo) {
          if (oo == my_off) continue;
          const float4 rj = t_r[g0 + (oo - o0)];
          const float d2 = h.wraps ? min_image_d2_f32(ri, rj, bx) : plain_d2_f32(ri, rj);
          if (d2 < crow[__float_as_uint(rj.w)]) {
            if (cnt < max_nbrs) group_append(w, cnt, encode_entry(oo, bytes), grp, nvp, i);
            ++cnt;  // keeps counting past capacity: the true length sizes the regrowth
          }
        }
      }
  const uint32_t stored = cnt < max_nbrs ? cnt : max_nbrs;
  if (stored & 7u) grp[static_cast<size_t>(stored >> 3) * nvp + i] = w;  // partial last group
  return cnt;
}

// Workgroup of the tile build: one tile's halo positions are staged in LDS
// once and shared by all its waves, so a larger workgroup has more waves in
// flight per byte of LDS (the candidate loop is bound by LDS latency).
#ifndef RUMDX_BUILD_BLOCK
#define RUMDX_BUILD_BLOCK 512
#endif
constexpr uint32_t kBuildBlock = RUMDX_BUILD_BLOCK;
static_assert(kBuildBlock % 64 == 0 && kBuildBlock >= kBlock && kBuildBlock <= kMaxWorkgroup, "build workgroup size");

// calculateNBL_CellsSorted (src/NeighborList.cu:1204-1319), one workgroup per
// cell tile with the tile's halo staged in LDS: for each particle, scan the
// 5x5x5 cell stencil in RUMD's (dZ, dY, dX) order, ascending slot inside each
// cell, and accept j != i with d^2 < (rc_ij + skin)^2 (FP32, the reference's
// minimum-image arithmetic; a tile whose halo does not wrap needs no image
// correction, and the plain difference is then the same number).  Entries are
// halo positions.  Each particle's sorted data is copied back to its home slot.
__global__ __launch_bounds__(kBuildBlock) void k_build_tiles(uint32_t* __restrict__ flags, SortScratch s,
                                                             TileGeom tg, const float* __restrict__ cut_skin2,
                                                             uint32_t nt, Box bx, HomeArrays o,
                                                             uint16_t* __restrict__ nbl16,
                                                             uint32_t* __restrict__ num_nbrs, uint32_t nvp,
                                                             uint32_t max_nbrs) {
  if (!rebuild_pending(flags)) return;
  // Arrival at the start (grouped counters, kArriveGroups): its round trip
  // overlaps the halo loads, and the block that arrives last (every block has
  // read the request by then) completes the build at its end.
  const bool last = threadIdx.x == 0 && grouped_arrival_last(s.arrive);
  __shared__ HaloLds L;
  __shared__ float s_cut[RUMDX_MAX_TYPES * RUMDX_MAX_TYPES];
  __shared__ __align__(16) uint32_t s_key[kOrderMax];  // list lengths, then sort keys (same index, same thread)
  __shared__ WaveBuild s_wave[kBuildBlock / 64];
  extern __shared__ float4 s_pos[];
  for (uint32_t k = threadIdx.x; k < nt * nt; k += kBuildBlock) s_cut[k] = cut_skin2[k];
  // the sentinel after the halo capacity: a position no distance test accepts
  if (threadIdx.x == 0) s_pos[tg.halo_cap] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  // the halo in halo order also goes to the force kernel's slot list; the
  // waves' build buffers are free during staging and hold its cell table
  static_assert(sizeof(s_wave) >= sizeof(uint16_t) * RUMDX_HALO_CAP_MAX, "cell table scratch");
  stage_halo(t, tg, s.t_r, s_pos, L, reinterpret_cast<uint16_t*>(s_wave),
             tg.halo_slots + static_cast<size_t>(t) * tg.halo_cap);
  const Halo& h = L.h;
  const bool staged = h.npart <= tg.halo_cap;
  const bool bytes = h.npart <= kByteEntryHalo;
  const uint32_t np = h.slot1 - h.slot0;
  uint32_t longest = 0;  // this thread's longest list
  // wave-cooperative lists for staged tiles; one lane per particle otherwise
  const bool coop = staged && np <= kOrderMax;
  if (coop) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0x = h.base[0] + 2, c0y = h.base[1] + 2, c0z = h.base[2] + 2;
    const uint32_t nown = static_cast<uint32_t>(h.own[0] * h.own[1] * h.own[2]);
    for (uint32_t cc = wave; cc < nown; cc += kBuildBlock / 64) {
      const int a = static_cast<int>(cc % h.own[0]);
      const int b = static_cast<int>((cc / h.own[0]) % h.own[1]);
      const int c = static_cast<int>(cc / (h.own[0] * h.own[1]));
      int cx = c0x + a, cy = c0y + b, cz = c0z + c;
      cx -= cx >= tg.nc0 ? tg.nc0 : 0;
      cy -= cy >= tg.nc1 ? tg.nc1 : 0;
      cz -= cz >= tg.nc2 ? tg.nc2 : 0;
      // a cell whose stencil stays inside the box on every axis (and cannot
      // reach L/2: at least 7 cells per axis) needs no image correction
      const bool cell_wraps =
          h.wraps && (stencil_wraps(cx, tg.nc0) || stencil_wraps(cy, tg.nc1) || stencil_wraps(cz, tg.nc2));
      if (cell_wraps)
        build_cell_wave<true>(h, L, tg, s_pos, s_cut, nt, bx, cx, cy, cz, s_wave[wave], s_key, nbl16, nvp, max_nbrs,
                              bytes);
      else
        build_cell_wave<false>(h, L, tg, s_pos, s_cut, nt, bx, cx, cy, cz, s_wave[wave], s_key, nbl16, nvp, max_nbrs,
                               bytes);
    }
  }
  __syncthreads();
  for (uint32_t i = h.slot0 + threadIdx.x; i < h.slot1; i += kBuildBlock) {
    const float4 ri = s.t_r[i];
    const int cx = cell_coord(ri.x, bx.iL[0], tg.nc0);
    const int cy = cell_coord(ri.y, bx.iL[1], tg.nc1);
    const int cz = cell_coord(ri.z, bx.iL[2], tg.nc2);
    uint32_t cnt;
    if (coop) {
      cnt = s_key[i - h.slot0];
    } else {
      cnt = build_particle_lane(i, ri, cx, cy, cz, h, L, tg, s.t_r, s_cut + __float_as_uint(ri.w) * nt, bx,
                                reinterpret_cast<u32x4*>(nbl16), nvp, max_nbrs, bytes);
    }
    const uint32_t stored = cnt < max_nbrs ? cnt : max_nbrs;
    num_nbrs[i] = stored;  // == max_nbrs marks an overflowed list
    if (np <= kOrderMax) s_key[i - h.slot0] = ((0xffffu - min(stored, 0xffffu)) << 16) | (i - h.slot0);
    if (cnt >= max_nbrs) flags[RUMDX_FLAG_OVERFLOW] = 1u;
    longest = cnt > longest ? cnt : longest;
    o.slot_cell[i] = static_cast<uint32_t>(cx + cy * tg.nc0 + cz * tg.nc0 * tg.nc1);
    o.r[i] = ri;
    o.last_r[i] = ri;
    o.v[i] = s.t_v[i];
    o.f[i] = s.t_f[i];
    o.w[i] = s.t_w[i];
    o.im[i] = s.t_im[i];
    if (o.rlo != nullptr) o.rlo[i] = s.t_rlo[i];
    o.idx[i] = s.t_idx[i];
  }
  // Lane order of the force kernel: the tile's particles by descending list
  // length, ties by slot (a rank count over unique keys; deterministic).
  // Keys past np are padded with a value above every key, so the count reads
  // four keys per LDS load.  block_max_u32 synchronises the workgroup, so
  // every key is written.
  const uint32_t np4 = (np + 3u) & ~3u;
  if (np <= kOrderMax && threadIdx.x < np4 - np) s_key[np + threadIdx.x] = 0xffffffffu;
  __shared__ uint32_t s_max[kBuildBlock / 64];
  longest = block_max_u32(longest, s_max);
  for (uint32_t k = threadIdx.x; k < np; k += kBuildBlock) {
    uint32_t rk = k;
    if (np <= kOrderMax) {
      const uint32_t key = s_key[k];
      const uint4* k4 = reinterpret_cast<const uint4*>(s_key);
      uint32_t r0 = 0, r1 = 0;
      uint32_t j = 0;
      for (; j + 8 <= np4; j += 8) {
        const uint4 a = k4[j >> 2], b = k4[(j >> 2) + 1];
        r0 += (a.x < key ? 1u : 0u) + (a.y < key ? 1u : 0u) + (a.z < key ? 1u : 0u) + (a.w < key ? 1u : 0u);
        r1 += (b.x < key ? 1u : 0u) + (b.y < key ? 1u : 0u) + (b.z < key ? 1u : 0u) + (b.w < key ? 1u : 0u);
      }
      if (j < np4) {
        const uint4 a = k4[j >> 2];
        r0 += (a.x < key ? 1u : 0u) + (a.y < key ? 1u : 0u) + (a.z < key ? 1u : 0u) + (a.w < key ? 1u : 0u);
      }
      rk = r0 + r1;
    }
    o.lane_order[h.slot0 + rk] = static_cast<uint16_t>(k);
  }
  if (threadIdx.x == 0) {
    // per-tile records (longest list here, halo and slot range in tile_info),
    // reduced by k_build_stats when the host asks; plain stores, so no block
    // waits on a same-address atomic
    s.tile_max[t] = longest;
    if (last) {
      flags[RUMDX_FLAG_REBUILD] = 0u;
      flags[RUMDX_FLAG_BUILDS] += 1u;
      flags[RUMDX_FLAG_EXCL_PENDING] = 1u;
      flags[RUMDX_FLAG_PRUNE] = 1u;
      flags[RUMDX_FLAG_STATS] = 1u;
    }
    if (h.npart > 65535u) atomicOr(flags + RUMDX_FLAG_ERROR, RUMDX_BUILD_ERR_HALO);
    // wrapping axes, and whether the shifted frame serves them
    const int nc[3] = {tg.nc0, tg.nc1, tg.nc2};
    uint32_t fl = 0;
    bool shiftable = true;
    for (int a = 0; a < 3; ++a) {
      const bool below = h.base[a] < 0, above = h.base[a] + h.ext[a] > nc[a];
      if (below) fl |= kTileShiftDown << (2 * a);
      if (above) fl |= kTileShiftUp << (2 * a);
      shiftable = shiftable && ((!below && !above) || (nc[a] >= kShiftMinCells && h.ext[a] < nc[a] &&
                                                       !(below && above) && shift_exact(bx.L[a], nc[a])));
    }
    if (staged && shiftable) fl |= kTileFast;
    tg.tile_info[t] = make_uint4(h.npart, fl | (tg.halo_cap << 16), h.slot0, h.slot1);
  }
  // the cell starts of this build, home for the force kernels' rescans and
  // the next build (the other tiles of this launch read start_next)
  const uint32_t r0 = tg.first_rank[t], r1 = tg.first_rank[t + 1];
  for (uint32_t c = r0 + threadIdx.x; c < r1; c += kBuildBlock) o.cell_start[c] = tg.cell_start[c];
  if (t == tg.ntiles - 1 && threadIdx.x == 0) o.cell_start[r1] = tg.cell_start[r1];
}

// Exclusions on the lists of the sorted method (apply_exclusion_list,
// src/NeighborList.cu:1323-1432), after a rebuild only: each list loses the
// entries of the particles its owner must not interact with, in place and in
// order.  Entries are decoded through the tile's halo slot list, so tiles whose
// halo was not staged (their force kernel rescans cells and skips exclusions
// there) and overflowed lists are left alone.  A completed build raises
// flags[EXCL_PENDING]; without it every block returns at once (no atomics on
// the many idle steps).  With it, every block counts its arrival on the
// build's grouped counters and the last one clears the flag, so no block can
// miss it.
__global__ __launch_bounds__(kBlock) void k_apply_exclusions(uint32_t* __restrict__ flags,
                                                             const uint4* __restrict__ tile_info,
                                                             const uint32_t* __restrict__ tile_halo, uint32_t halo_cap,
                                                             Exclusions ex, uint16_t* __restrict__ nbl16,
                                                             uint32_t* __restrict__ num_nbrs, uint32_t nvp,
                                                             uint32_t max_nbrs, uint32_t* __restrict__ arrive) {
  if (__hip_atomic_load(flags + RUMDX_FLAG_EXCL_PENDING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  // arrival at the start, on the grouped counters (the build left them at zero)
  const bool last = threadIdx.x == 0 && grouped_arrival_last(arrive);
  {
    const uint32_t t = blockIdx.x;
    const uint4 info = tile_info[t];
    const bool staged = (info.y >> 16) == halo_cap && info.x <= halo_cap;
    const bool bytes = info.x <= kByteEntryHalo;
    const uint32_t* hs = tile_halo + static_cast<size_t>(t) * halo_cap;
    for (uint32_t i = info.z + threadIdx.x; staged && i < info.w; i += kBlock) {
      const ExclusionRange er = ex.of(i);
      if (er.e0 == er.e1) continue;
      const uint32_t nn = num_nbrs[i];
      if (nn >= max_nbrs) continue;
      uint32_t kept = 0;
      for (uint32_t k = 0; k < nn; ++k) {
        const size_t at = (static_cast<size_t>(k >> 3) * nvp + i) * 8u + (k & 7u);
        const uint16_t e = nbl16[at];
        const uint32_t h = decode_entry(e, bytes);
        if (h < info.x && ex.excludes(er, hs[h])) continue;
        nbl16[(static_cast<size_t>(kept >> 3) * nvp + i) * 8u + (kept & 7u)] = e;
        ++kept;
      }
      num_nbrs[i] = kept;
    }
  }
  if (last) flags[RUMDX_FLAG_EXCL_PENDING] = 0u;
}

// Converts tile-format entries (halo offsets) back to slots.
__global__ __launch_bounds__(kBlock) void k_export_tiles(TileGeom tg, const uint16_t* __restrict__ nbl16,
                                                         const uint32_t* __restrict__ num_nbrs,
                                                         uint32_t* __restrict__ out, uint32_t nvp) {
  __shared__ HaloLds L;
  TileGeom g = tg;
  g.halo_cap = 0;  // table only, no positions
  stage_halo(blockIdx.x, g, nullptr, nullptr, L);
  const Halo& h = L.h;
  const bool bytes = h.npart <= kByteEntryHalo;
  for (uint32_t i = h.slot0 + threadIdx.x; i < h.slot1; i += kBlock) {
    const uint32_t nn = num_nbrs[i];
    for (uint32_t k = 0; k < nn; ++k) {
      const uint32_t oo = decode_entry(nbl16[(static_cast<size_t>(k >> 3) * nvp + i) * 8 + (k & 7u)], bytes);
      uint32_t lo = 0, hi = h.ncell;  // last cell with off <= oo
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (L.off[mid] <= oo) lo = mid; else hi = mid;
      }
      out[static_cast<size_t>(k) * nvp + i] = L.start[lo] + (oo - L.off[lo]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_export_slots(const uint32_t* __restrict__ nbl,
                                                         const uint32_t* __restrict__ num_nbrs,
                                                         uint32_t* __restrict__ out, uint32_t n, uint32_t nvp) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t nn = num_nbrs[i];
  for (uint32_t k = 0; k < nn; ++k) out[static_cast<size_t>(k) * nvp + i] = nbl[static_cast<size_t>(k) * nvp + i];
}

int check_nb(const rumdx_particles* p, const rumdx_nblist* nb, const char* who) {
  if (p == nullptr || nb == nullptr) return fail(RUMDX_E_INVALID, std::string(who) + ": null argument");
  if (nb->flags_dev == nullptr || nb->num_nbrs_dev == nullptr || nb->last_r_dev == nullptr ||
      nb->cut_skin2_dev == nullptr)
    return fail(RUMDX_E_INVALID, std::string(who) + ": neighbour list not allocated");
  if (nb->max_nbrs == 0) return fail(RUMDX_E_INVALID, std::string(who) + ": max_nbrs is zero");
  if (static_cast<uint64_t>(nb->max_nbrs) * p->nvp >= (1ull << 32))
    return fail(RUMDX_E_INVALID, std::string(who) + ": max_nbrs * nvp exceeds 2^32 entries");
  if (nb->ntypes == 0 || nb->ntypes > RUMDX_MAX_TYPES) return fail(RUMDX_E_INVALID, std::string(who) + ": ntypes out of range");
  return RUMDX_OK;
}

}  // namespace
}  // namespace rumdx

using namespace rumdx;

extern "C" size_t rumdx_sort_scratch_bytes(uint32_t nvp, uint32_t ncells) { return scratch_bytes(nvp, ncells); }

extern "C" int rumdx_cell_grid(const float box[8], float max_cutoff, float skin, int32_t nc_out[3]) {
  if (box == nullptr || nc_out == nullptr) return fail(RUMDX_E_INVALID, "rumdx_cell_grid: null argument");
  if (!(max_cutoff + skin > 0.f)) return fail(RUMDX_E_INVALID, "rumdx_cell_grid: cutoff + skin must be positive");
  const Box bx = make_box(box);
  for (int a = 0; a < 3; ++a) {
    // (int)floor(n_Cells * L / (maxCutoff + skin)) with n_Cells = 2, all in float
    const float q = 2.0f * bx.L[a] / (max_cutoff + skin);
    int nc = static_cast<int>(floorf(q));
    if (nc < 5) nc = 5;
    nc_out[a] = nc;
  }
  return RUMDX_OK;
}

extern "C" int rumdx_cell_rank_table(const int32_t nc[3], uint32_t* rank_out) {
  if (nc == nullptr || rank_out == nullptr) return fail(RUMDX_E_INVALID, "rumdx_cell_rank_table: null argument");
  for (int a = 0; a < 3; ++a)
    if (nc[a] < 1 || nc[a] > 2048) return fail(RUMDX_E_INVALID, "rumdx_cell_rank_table: cells per axis out of range");
  // Rank = position of the cell's Morton code (x in bit 0, y in bit 1, z in
  // bit 2 of each triple) among all cells of the grid.
  const uint32_t n0 = nc[0], n1 = nc[1], n2 = nc[2];
  const uint64_t ncells = static_cast<uint64_t>(n0) * n1 * n2;
  std::vector<std::pair<uint64_t, uint32_t>> code(ncells);
  for (uint32_t z = 0; z < n2; ++z)
    for (uint32_t y = 0; y < n1; ++y)
      for (uint32_t x = 0; x < n0; ++x) {
        uint64_t m = 0;
        for (uint32_t b = 0; b < 11; ++b)
          m |= (static_cast<uint64_t>((x >> b) & 1u) << (3 * b)) | (static_cast<uint64_t>((y >> b) & 1u) << (3 * b + 1)) |
               (static_cast<uint64_t>((z >> b) & 1u) << (3 * b + 2));
        const uint32_t lin = x + y * n0 + z * n0 * n1;
        code[lin] = {m, lin};
      }
  std::sort(code.begin(), code.end());
  for (uint64_t k = 0; k < ncells; ++k) rank_out[code[k].second] = static_cast<uint32_t>(k);
  return RUMDX_OK;
}

extern "C" int rumdx_nblist_check(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream) {
  if (p == nullptr || nb == nullptr || box == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nblist_check: null argument");
  if (nb->flags_dev == nullptr || nb->last_r_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_nblist_check: neighbour list not allocated");
  if (p->n == 0) return RUMDX_OK;
  // skin2_over_4 = 0.25*skin*skin formed in double, passed as float (src/NeighborList.cu:578-579)
  const float s24 = static_cast<float>(0.25 * static_cast<double>(nb->skin) * static_cast<double>(nb->skin));
  hipLaunchKernelGGL(k_skin_check, dim3(blocks_for(p->n)), dim3(kBlock), 0, as_stream(stream),
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<const float4*>(nb->last_r_dev),
                     nb->flags_dev, s24, make_box(box), p->n, reinterpret_cast<const float4*>(inner_prune_r(nb)),
                     inner_check2(nb));
  return check_launch("rumdx_nblist_check");
}

extern "C" int rumdx_nblist_request_rebuild(rumdx_nblist* nb, void* stream) {
  if (nb == nullptr || nb->flags_dev == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nblist_request_rebuild: null argument");
  hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(64), 0, as_stream(stream), nb->flags_dev);
  return check_launch("rumdx_nblist_request_rebuild");
}

extern "C" int rumdx_nblist_build_n2(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream) {
  int rc = check_nb(p, nb, "rumdx_nblist_build_n2");
  if (rc != RUMDX_OK) return rc;
  if (box == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nblist_build_n2: null box");
  if (nb->nbl_dev == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nblist_build_n2: slot list not allocated");
  if (nb->excl_start_dev != nullptr && (nb->excl_ids_dev == nullptr || p->unsorted_index_dev == nullptr))
    return fail(RUMDX_E_INVALID, "rumdx_nblist_build_n2: exclusions need the exclusion ids and the unsorted index");
  if (p->n == 0) return RUMDX_OK;
  const size_t lds = sizeof(float4) * kBlock + sizeof(float) * nb->ntypes * nb->ntypes;
  hipLaunchKernelGGL(k_build_n2, dim3(blocks_for(p->n)), dim3(kBlock), lds, as_stream(stream),
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<float4*>(nb->last_r_dev),
                     nb->nbl_dev, nb->num_nbrs_dev, nb->flags_dev, nb->cut_skin2_dev, nb->ntypes, make_box(box), p->n,
                     p->nvp, nb->max_nbrs, Exclusions{p->unsorted_index_dev, nb->excl_start_dev, nb->excl_ids_dev});
  return check_launch("rumdx_nblist_build_n2");
}

// ------------------------------------------------------------------ inner lists
// Prune of the tile lists (rumdx_nblist_prune).  Workgroups loop over the
// tiles (a capped grid, so an idle launch costs little); each stages its
// tile's halo in LDS as the force kernel does and gives one lane per
// particle, in the build's lane order.  A lane walks its particle's list,
// keeps the entries within rc + inner_skin and packs them into 16-byte
// groups through a 4-word shift register (one funnel shift per word).
namespace rumdx {
namespace {

constexpr uint32_t kPruneGrid = 1024;

// Appends entry e (16 bits) to the group being packed in w: every entry moves
// down one slot and e enters slot 7, so after eight appends slot k holds the
// k-th entry.
__device__ __forceinline__ void shift_in(u32x4& w, uint32_t e) {
  w.x = __builtin_amdgcn_alignbit(w.y, w.x, 16);
  w.y = __builtin_amdgcn_alignbit(w.z, w.y, 16);
  w.z = __builtin_amdgcn_alignbit(w.w, w.z, 16);
  w.w = __builtin_amdgcn_alignbit(e, w.w, 16);
}

template <bool FAST>
__device__ __forceinline__ uint32_t prune_particle(uint32_t i, float4 ri, uint32_t nn, const float* crow,
                                                   const float4* s_pos, const Box& bx,
                                                   const uint16_t* __restrict__ nbl16, uint16_t* __restrict__ nbl16_in,
                                                   uint32_t nvp) {
  const u32x4* grp = reinterpret_cast<const u32x4*>(nbl16) + i;
  u32x4* out = reinterpret_cast<u32x4*>(nbl16_in) + i;
  u32x4 w = {0u, 0u, 0u, 0u};
  uint32_t kept = 0;
  const uint32_t ng = (nn + 7u) >> 3;
  for (uint32_t g = 0; g < ng; ++g) {
    const u32x4 q = __builtin_nontemporal_load(grp + static_cast<size_t>(g) * nvp);
    const uint32_t ne = (nn - 8u * g) < 8u ? (nn - 8u * g) : 8u;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if (static_cast<uint32_t>(e) < ne) {
        const uint32_t ent = group_entry(q, e);
        const float4 rj = halo_pos(s_pos, ent);
        const float d2 = FAST ? plain_d2_f32(ri, rj) : min_image_d2_f32(ri, rj, bx);
        if (d2 < crow[__float_as_uint(rj.w)]) {
          shift_in(w, ent);
          ++kept;
          if ((kept & 7u) == 0u) out[static_cast<size_t>((kept >> 3) - 1u) * nvp] = w;
        }
      }
    }
  }
  if (kept & 7u) {  // the partial last group: move its entries down to slots 0..
    for (uint32_t k = kept & 7u; k < 8u; ++k) shift_in(w, 0u);
    out[static_cast<size_t>(kept >> 3) * nvp] = w;
  }
  return kept;
}

__global__ __launch_bounds__(kBlock) void k_prune_tiles(const uint32_t* __restrict__ flags,
                                                        const float4* __restrict__ r, float4* __restrict__ prune_r,
                                                        const uint16_t* __restrict__ nbl16,
                                                        const uint32_t* __restrict__ num_nbrs,
                                                        const uint16_t* __restrict__ lane_order,
                                                        uint16_t* __restrict__ nbl16_in, uint32_t* __restrict__ num_in,
                                                        const float* __restrict__ cut_in2, uint32_t nt, Box bx,
                                                        const uint4* __restrict__ tile_info,
                                                        const uint32_t* __restrict__ halo_slots, uint32_t halo_cap,
                                                        uint32_t ntiles, uint32_t nvp, uint32_t max_nbrs) {
  if (__hip_atomic_load(flags + RUMDX_FLAG_PRUNE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  extern __shared__ float4 s_dyn[];
  float4* s_pos = s_dyn;
  float* s_cut = reinterpret_cast<float*>(s_dyn + halo_cap);
  for (uint32_t k = threadIdx.x; k < nt * nt; k += kBlock) s_cut[k] = cut_in2[k];
  for (uint32_t t = xcd_block(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
    const uint4 info = tile_info[t];
    const bool staged = (info.y >> 16) == halo_cap && info.x <= halo_cap;
    const bool fast = (info.y & kTileFast) != 0u;
    const uint32_t shifts = fast ? info.y : 0u;
    __syncthreads();  // the previous tile's halo has been read (and s_cut written)
    if (staged) {
      const uint32_t* hs = halo_slots + static_cast<size_t>(t) * halo_cap;
      for (uint32_t q = threadIdx.x; q < info.x; q += kBlock) {
        float4 pq = r[hs[q]];
        pq.x = shift_coord(pq.x, shifts, 0, bx.L[0]);
        pq.y = shift_coord(pq.y, shifts, 1, bx.L[1]);
        pq.z = shift_coord(pq.z, shifts, 2, bx.L[2]);
        s_pos[q] = pq;
      }
    }
    __syncthreads();
    const uint32_t np = info.w - info.z;
    for (uint32_t k = threadIdx.x; k < np; k += kBlock) {
      const uint32_t i = info.z + lane_order[info.z + k];
      float4 ri = r[i];
      prune_r[i] = ri;
      const uint32_t nn = num_nbrs[i];
      if (!staged || nn >= max_nbrs) {  // the force kernels rescan these particles
        num_in[i] = nn;
        continue;
      }
      const float* crow = s_cut + __float_as_uint(ri.w) * nt;
      if (fast) {
        ri.x = shift_coord(ri.x, shifts, 0, bx.L[0]);
        ri.y = shift_coord(ri.y, shifts, 1, bx.L[1]);
        ri.z = shift_coord(ri.z, shifts, 2, bx.L[2]);
        num_in[i] = prune_particle<true>(i, ri, nn, crow, s_pos, bx, nbl16, nbl16_in, nvp);
      } else {
        num_in[i] = prune_particle<false>(i, ri, nn, crow, s_pos, bx, nbl16, nbl16_in, nvp);
      }
    }
  }
}

__global__ void k_prune_done(uint32_t* flags) {
  if (threadIdx.x == 0) flags[RUMDX_FLAG_PRUNE] = 0u;
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_nblist_prune(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream) {
  using namespace rumdx;
  const char* entry = "rumdx_nblist_prune";
  if (p == nullptr || nb == nullptr || box == nullptr) return fail(RUMDX_E_INVALID, std::string(entry) + ": null argument");
  if (nb->nbl16_dev == nullptr || nb->tile_info_dev == nullptr || nb->tile_halo_dev == nullptr ||
      nb->lane_order_dev == nullptr || nb->num_nbrs_dev == nullptr || nb->flags_dev == nullptr || nb->ntiles == 0)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": the sorted (tile) lists are not allocated");
  if (nb->nbl16_in_dev == nullptr || nb->num_in_dev == nullptr || nb->prune_r_dev == nullptr ||
      nb->cut_in2_dev == nullptr)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": inner lists not allocated");
  if (!(nb->inner_skin > 0.f)) return fail(RUMDX_E_INVALID, std::string(entry) + ": inner_skin must be positive");
  if (nb->halo_cap > RUMDX_HALO_CAP_MAX || nb->ntypes == 0 || nb->ntypes > RUMDX_MAX_TYPES)
    return fail(RUMDX_E_INVALID, std::string(entry) + ": halo capacity or type count out of range");
  if (p->r_dev == nullptr) return fail(RUMDX_E_INVALID, std::string(entry) + ": particle arrays not allocated");
  if (p->n == 0) return RUMDX_OK;
  const hipStream_t st = as_stream(stream);
  const uint32_t grid = nb->ntiles < kPruneGrid ? nb->ntiles : kPruneGrid;
  const size_t lds = sizeof(float4) * nb->halo_cap + sizeof(float) * nb->ntypes * nb->ntypes;
  hipLaunchKernelGGL(k_prune_tiles, dim3(grid), dim3(kBlock), lds, st, nb->flags_dev,
                     reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<float4*>(nb->prune_r_dev),
                     nb->nbl16_dev, nb->num_nbrs_dev, nb->lane_order_dev, nb->nbl16_in_dev, nb->num_in_dev,
                     nb->cut_in2_dev, nb->ntypes, make_box(box), reinterpret_cast<const uint4*>(nb->tile_info_dev),
                     nb->tile_halo_dev, nb->halo_cap, nb->ntiles, p->nvp, nb->max_nbrs);
  hipLaunchKernelGGL(k_prune_done, dim3(1), dim3(64), 0, st, nb->flags_dev);
  return check_launch(entry);
}

namespace rumdx {
namespace {

int check_sorted_args(const char* entry, const rumdx_particles* p, const rumdx_nblist* nb, const float* box) {
  const std::string e(entry);
  int rc = check_nb(p, nb, entry);
  if (rc != RUMDX_OK) return rc;
  if (box == nullptr) return fail(RUMDX_E_INVALID, e + ": null box");
  if (nb->cell_rank_dev == nullptr || nb->cell_start_dev == nullptr || nb->cell_count_dev == nullptr ||
      nb->slot_cell_dev == nullptr || nb->sort_scratch_dev == nullptr || nb->nbl16_dev == nullptr ||
      nb->lane_order_dev == nullptr || nb->tile_halo_dev == nullptr || nb->tile_info_dev == nullptr ||
      nb->tile_first_rank_dev == nullptr || nb->tile_origin_dev == nullptr || nb->ntiles == 0)
    return fail(RUMDX_E_INVALID, e + ": cell or tile arrays not allocated");
  if (nb->halo_cap > RUMDX_HALO_CAP_MAX) return fail(RUMDX_E_INVALID, e + ": halo_cap too large");
  if (p->v_dev == nullptr || p->im_dev == nullptr || p->unsorted_index_dev == nullptr || p->f_dev == nullptr ||
      p->w_dev == nullptr)
    return fail(RUMDX_E_INVALID, e + ": particle arrays not allocated");
  if (nb->excl_start_dev != nullptr && nb->excl_ids_dev == nullptr)
    return fail(RUMDX_E_INVALID, e + ": exclusion ids missing");
  for (int a = 0; a < 3; ++a)
    if (nb->nc[a] < 5) return fail(RUMDX_E_INVALID, e + ": fewer than 5 cells per axis");
  return RUMDX_OK;
}

// The tile build, its finishing kernel and the exclusion pass: the common
// end of both sorted builds (the particles are staged in sorted order in s,
// their cell starts in s.start_next).
void launch_tile_build(rumdx_particles* p, rumdx_nblist* nb, const Box& bx, const SortScratch& s, hipStream_t st) {
  const TileGeom tg{nb->nc[0],           nb->nc[1],
                    nb->nc[2],           nb->cell_rank_dev,
                    s.start_next,        nb->tile_first_rank_dev,
                    nb->tile_origin_dev, nb->ntiles,
                    nb->halo_cap,        nb->tile_halo_dev,
                    reinterpret_cast<uint4*>(nb->tile_info_dev)};
  const HomeArrays home{reinterpret_cast<float4*>(p->r_dev), reinterpret_cast<float4*>(p->v_dev),
                        reinterpret_cast<float4*>(p->f_dev), reinterpret_cast<float4*>(p->w_dev),
                        reinterpret_cast<float4*>(p->im_dev), reinterpret_cast<float4*>(p->rlo_dev),
                        p->unsorted_index_dev, reinterpret_cast<float4*>(nb->last_r_dev), nb->slot_cell_dev,
                        nb->lane_order_dev, nb->cell_start_dev};
  hipLaunchKernelGGL(k_build_tiles, dim3(nb->ntiles), dim3(kBuildBlock), sizeof(float4) * (nb->halo_cap + 1u), st,
                     nb->flags_dev, s, tg, nb->cut_skin2_dev, nb->ntypes, bx, home, nb->nbl16_dev, nb->num_nbrs_dev,
                     p->nvp, nb->max_nbrs);
  if (nb->excl_start_dev != nullptr)
    hipLaunchKernelGGL(k_apply_exclusions, dim3(nb->ntiles), dim3(kBlock), 0, st, nb->flags_dev,
                       reinterpret_cast<const uint4*>(nb->tile_info_dev), nb->tile_halo_dev, nb->halo_cap,
                       Exclusions{p->unsorted_index_dev, nb->excl_start_dev, nb->excl_ids_dev}, nb->nbl16_dev,
                       nb->num_nbrs_dev, p->nvp, nb->max_nbrs, s.arrive);
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_nblist_build_sorted(rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream) {
  int rc = check_sorted_args("rumdx_nblist_build_sorted", p, nb, box);
  if (rc != RUMDX_OK) return rc;
  if (p->n == 0) return RUMDX_OK;
  const CellGeom g{nb->nc[0], nb->nc[1], nb->nc[2],
                   static_cast<uint32_t>(nb->nc[0]) * static_cast<uint32_t>(nb->nc[1]) * static_cast<uint32_t>(nb->nc[2])};
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  SortScratch s = carve(nb->sort_scratch_dev, p->nvp, g.ncells);
  const uint32_t nblk = blocks_for(p->n);
  const float4* r = reinterpret_cast<const float4*>(p->r_dev);

  // The chain runs every step and exits at once unless a rebuild is pending,
  // so the particle-parallel kernels loop over a capped grid: an idle launch
  // then costs little more than its dispatch.
  const uint32_t ngrid = std::min(nblk, kChainGridCap);
  hipLaunchKernelGGL(k_cell_hist, dim3(ngrid), dim3(kBlock), 0, st, r, nb->flags_dev, nb->cell_rank_dev, s.key_lin,
                     nb->cell_count_dev, bx, g, p->n);
  const uint32_t nchunk = scan_chunks(g.ncells);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nchunk), dim3(kBlock), 0, st, nb->cell_count_dev, nb->flags_dev, s.chunk_sum,
                     g.ncells);
  hipLaunchKernelGGL(k_scan_apply, dim3(nchunk), dim3(kBlock), 0, st, nb->cell_count_dev, nb->flags_dev, s.chunk_sum,
                     s.start_next, g.ncells);
  hipLaunchKernelGGL(k_scatter, dim3(ngrid), dim3(kBlock), 0, st, s.key_lin, nb->flags_dev, nb->cell_rank_dev,
                     s.start_next, nb->cell_count_dev, s.perm, p->n);
  hipLaunchKernelGGL(k_gather, dim3(ngrid), dim3(kBlock), 0, st, nb->flags_dev, s.perm, s.key_lin, nb->cell_rank_dev,
                     s.start_next, r, reinterpret_cast<const float4*>(p->v_dev),
                     reinterpret_cast<const float4*>(p->f_dev), reinterpret_cast<const float4*>(p->w_dev),
                     reinterpret_cast<const float4*>(p->im_dev), reinterpret_cast<const float4*>(p->rlo_dev),
                     p->unsorted_index_dev, s, p->n);
  launch_tile_build(p, nb, bx, s, st);
  return check_launch("rumdx_nblist_build_sorted");
}

extern "C" int rumdx_nblist_rebuild_sorted(rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream) {
  int rc = check_sorted_args("rumdx_nblist_rebuild_sorted", p, nb, box);
  if (rc != RUMDX_OK) return rc;
  if (p->n == 0) return RUMDX_OK;
  const uint32_t ncells =
      static_cast<uint32_t>(nb->nc[0]) * static_cast<uint32_t>(nb->nc[1]) * static_cast<uint32_t>(nb->nc[2]);
  const Box bx = make_box(box);
  const hipStream_t st = as_stream(stream);
  SortScratch s = carve(nb->sort_scratch_dev, p->nvp, ncells);
  // the previous build's cell starts (home) are the old cells of the re-sort
  const TileGeom tg{nb->nc[0], nb->nc[1], nb->nc[2], nb->cell_rank_dev, nb->cell_start_dev,
                    nb->tile_first_rank_dev, nb->tile_origin_dev, nb->ntiles, nb->halo_cap, nullptr, nullptr};
  const ParticleArrays old{reinterpret_cast<const float4*>(p->r_dev), reinterpret_cast<const float4*>(p->v_dev),
                           reinterpret_cast<const float4*>(p->f_dev), reinterpret_cast<const float4*>(p->w_dev),
                           reinterpret_cast<const float4*>(p->im_dev), reinterpret_cast<const float4*>(p->rlo_dev),
                           p->unsorted_index_dev};
  hipLaunchKernelGGL(k_tile_resort, dim3(nb->ntiles), dim3(kBlock), 0, st, nb->flags_dev, s, tg, bx, old, p->n);
  launch_tile_build(p, nb, bx, s, st);
  return check_launch("rumdx_nblist_rebuild_sorted");
}

extern "C" int rumdx_nblist_build_stats(const rumdx_particles* p, rumdx_nblist* nb, void* stream) {
  if (p == nullptr || nb == nullptr || nb->flags_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_nblist_build_stats: null argument");
  if (nb->nbl16_dev == nullptr || nb->tile_info_dev == nullptr || nb->sort_scratch_dev == nullptr || nb->ntiles == 0)
    return RUMDX_OK;  // no sorted build: the N2 build keeps the flags itself
  const uint32_t ncells =
      static_cast<uint32_t>(nb->nc[0]) * static_cast<uint32_t>(nb->nc[1]) * static_cast<uint32_t>(nb->nc[2]);
  const SortScratch s = carve(nb->sort_scratch_dev, p->nvp, ncells);
  hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(kBlock), 0, as_stream(stream), nb->flags_dev,
                     reinterpret_cast<const uint4*>(nb->tile_info_dev), s.tile_max, nb->ntiles);
  return check_launch("rumdx_nblist_build_stats");
}

extern "C" int rumdx_tile_table(const int32_t nc[3], const uint32_t* rank, uint32_t* first_rank_out,
                                uint32_t* origin_out) {
  if (nc == nullptr) return fail(RUMDX_E_INVALID, "rumdx_tile_table: null argument");
  int nb[3];
  for (int a = 0; a < 3; ++a) {
    if (nc[a] < 1 || nc[a] > 1023) return fail(RUMDX_E_INVALID, "rumdx_tile_table: cells per axis out of range");
    nb[a] = (nc[a] + kTileEdge - 1) / kTileEdge;
  }
  const uint64_t ntiles = static_cast<uint64_t>(nb[0]) * nb[1] * nb[2];
  if (first_rank_out == nullptr && origin_out == nullptr) return static_cast<int>(ntiles);
  if (rank == nullptr || first_rank_out == nullptr || origin_out == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_tile_table: null output");
  // The origin cell of an aligned brick has the smallest Morton code (hence
  // rank) of the brick, and bricks occupy disjoint code ranges.
  std::vector<std::pair<uint32_t, uint32_t>> t;
  t.reserve(ntiles);
  for (int z = 0; z < nb[2]; ++z)
    for (int y = 0; y < nb[1]; ++y)
      for (int x = 0; x < nb[0]; ++x) {
        const uint32_t cx = x * kTileEdge, cy = y * kTileEdge, cz = z * kTileEdge;
        const uint32_t lin = cx + cy * nc[0] + cz * static_cast<uint32_t>(nc[0]) * nc[1];
        t.emplace_back(rank[lin], cx | (cy << 10) | (cz << 20));
      }
  std::sort(t.begin(), t.end());
  for (uint64_t k = 0; k < ntiles; ++k) {
    first_rank_out[k] = t[k].first;
    origin_out[k] = t[k].second;
  }
  first_rank_out[ntiles] = static_cast<uint32_t>(static_cast<uint64_t>(nc[0]) * nc[1] * nc[2]);
  return static_cast<int>(ntiles);
}

extern "C" uint32_t rumdx_halo_cap_estimate(const int32_t nc[3], uint32_t n) {
  if (nc == nullptr) return 0;
  const double ncells = static_cast<double>(nc[0]) * nc[1] * nc[2];
  const double halo = static_cast<double>(kHaloCellsMax) * n / ncells;
  uint32_t cap = static_cast<uint32_t>(1.08 * halo + 64.0);
  cap = (cap + 63u) & ~63u;
  return cap > RUMDX_HALO_CAP_MAX ? RUMDX_HALO_CAP_MAX : cap;
}

extern "C" int rumdx_nblist_export(const rumdx_particles* p, const rumdx_nblist* nb, uint32_t* out_dev,
                                   void* stream) {
  if (p == nullptr || nb == nullptr || out_dev == nullptr || nb->num_nbrs_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_nblist_export: null argument");
  if (p->n == 0) return RUMDX_OK;
  const hipStream_t st = as_stream(stream);
  if (nb->nbl16_dev != nullptr) {
    if (nb->tile_first_rank_dev == nullptr || nb->tile_origin_dev == nullptr || nb->cell_rank_dev == nullptr ||
        nb->cell_start_dev == nullptr)
      return fail(RUMDX_E_INVALID, "rumdx_nblist_export: tile arrays not allocated");
    const TileGeom tg{nb->nc[0], nb->nc[1], nb->nc[2], nb->cell_rank_dev, nb->cell_start_dev,
                      nb->tile_first_rank_dev, nb->tile_origin_dev, nb->ntiles, 0u, nullptr, nullptr};
    hipLaunchKernelGGL(k_export_tiles, dim3(nb->ntiles), dim3(kBlock), 0, st, tg, nb->nbl16_dev, nb->num_nbrs_dev,
                       out_dev, p->nvp);
  } else {
    if (nb->nbl_dev == nullptr) return fail(RUMDX_E_INVALID, "rumdx_nblist_export: list not allocated");
    hipLaunchKernelGGL(k_export_slots, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, nb->nbl_dev, nb->num_nbrs_dev,
                       out_dev, p->n, p->nvp);
  }
  return check_launch("rumdx_nblist_export");
}
