// Batched replicas (gfx950): hundreds of independent samples of uneven size
// and temperature advanced by one launch per kernel.
//
// Each replica follows the single-sample algorithms: RUMD's cell grid and 5^3
// stencil for the list (calculateNBL_CellsSorted, src/NeighborList.cu:
// 1204-1319, with the FP32 criterion of :1239-1240), the skin criterion
// (displacement_since_nb_build, :970-997), the LJ pair term of fij
// (include/rumd/PairPotential.h:1198-1244) and the Nose-Hoover leapfrog with
// its thermostat update (src/IntegratorNVT.cu:64-92, 364-437).
//
// Layout: replica r owns slots [slot0, slot0 + n) of the shared float4
// arrays, slot0 a multiple of RUMDX_BATCH_ALIGN (= the workgroup size), so a
// workgroup never spans two replicas.  block_rep maps a workgroup to its
// replica; per-replica sums are block partials added in a fixed order by one
// workgroup per replica, so every replica's trajectory is independent of the
// others and reproducible bit for bit.  A rebuild sorts the replica's
// particles by cell inside its slot range (a stable counting sort staged
// through t_*, with idx following the particles), so a cell is a contiguous
// slot range and a particle's neighbours sit in a few contiguous runs of
// memory.  Lists are u32 slots, column-major over the whole batch (entry k
// of slot i at k*nslots + i).  Only replicas whose rebuild flag is set run
// the build kernels; the others exit at once.
#include <algorithm>

#include "leapfrog.h"
#include "rumdx_common.h"

namespace rumdx {
namespace {

static_assert(RUMDX_BATCH_ALIGN == kBlock, "one workgroup per aligned slot block");

enum RepFlag : uint32_t { kRebuild = 0, kOverflow = 1, kBuilds = 2, kMaxLen = 3, kDone = 4 };
constexpr uint32_t kRepFlags = RUMDX_BATCH_FLAGS;

__device__ __forceinline__ Box rep_box(const rumdx_replica& rp) {
  Box b;
  for (int a = 0; a < 3; ++a) {
    b.L[a] = rp.L;
    b.iL[a] = rp.inv_L;
    b.Ld[a] = static_cast<double>(rp.L);
    b.iLd[a] = 1.0 / b.Ld[a];
  }
  return b;
}

__device__ __forceinline__ uint32_t rep_cell(float4 p, const Box& bx, int nc) {
  const int cx = cell_coord(p.x, bx.iL[0], nc);
  const int cy = cell_coord(p.y, bx.iL[1], nc);
  const int cz = cell_coord(p.z, bx.iL[2], nc);
  return static_cast<uint32_t>(cx + nc * (cy + nc * cz));
}

// Fixed-point positions for the pair forces: per axis round(x/L 2^32) as a
// 32-bit integer (x in [-L/2, L/2) maps into int32; the conversion wraps
// modulo 2^32), and the type bits.  The difference of two such integers,
// taken modulo 2^32, is the minimum-image displacement in units of L 2^-32,
// exactly (no float rounding of a coordinate near L/2, unlike the FP32 minimum
// image); scaled in FP64 and rounded once it costs four VALU per axis, where
// the FP64 coordinate route costs about three times as much.  The
// quantisation, L 2^-33, is far below the positions' float ulp.
__device__ __forceinline__ int32_t fixed_axis(float x, double scale) {
  return static_cast<int32_t>(static_cast<uint32_t>(static_cast<long long>(rint(static_cast<double>(x) * scale))));
}
__device__ __forceinline__ int4 to_fixed(float4 r, double scale) {
  return make_int4(fixed_axis(r.x, scale), fixed_axis(r.y, scale), fixed_axis(r.z, scale), __float_as_int(r.w));
}
__device__ __forceinline__ double fixed_scale(const rumdx_replica& rp) {
  return 4294967296.0 / static_cast<double>(rp.L);
}

// Orders the LDS accesses of one wave (lanes read what other lanes wrote).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bool rebuild_pending(const uint32_t* flags, uint32_t rep) {
  return __hip_atomic_load(flags + kRepFlags * rep + kRebuild, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

struct BatchArgs {
  float4* r;
  float4* v;
  float4* f;
  float4* w;
  float4* im;
  int4* fixed;
  const rumdx_replica* rep;
  const uint32_t* block_rep;
  uint32_t* flags;
  uint32_t nslots;
};

BatchArgs args_of(const rumdx_batch* b) {
  return BatchArgs{reinterpret_cast<float4*>(b->r_dev), reinterpret_cast<float4*>(b->v_dev),
                   reinterpret_cast<float4*>(b->f_dev), reinterpret_cast<float4*>(b->w_dev),
                   reinterpret_cast<float4*>(b->im_dev), reinterpret_cast<int4*>(b->fixed_dev), b->rep_dev,
                   b->block_rep_dev, b->rep_flags_dev, b->nslots};
}

// ------------------------------------------------------------------ lists
// Cell of every slot of a rebuilding replica, and the per-cell counts.
__global__ __launch_bounds__(kBlock) void k_b_hist(BatchArgs a, uint32_t* __restrict__ slot_cell,
                                                   uint32_t* __restrict__ cell_count) {
  const uint32_t rep = a.block_rep[blockIdx.x];
  if (!rebuild_pending(a.flags, rep)) return;
  const rumdx_replica rp = a.rep[rep];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i - rp.slot0 >= rp.n) return;
  const uint32_t c = rp.cell0 + rep_cell(a.r[i], rep_box(rp), rp.nc);
  slot_cell[i] = c;
  atomicAdd(cell_count + c, 1u);
}

// One workgroup per replica: exclusive scan of its cell counts into slot
// offsets (cell_start), the counts kept in cell_n and zeroed for the scatter.
__global__ __launch_bounds__(kBlock) void k_b_scan(BatchArgs a, uint32_t* __restrict__ cell_count,
                                                   uint32_t* __restrict__ cell_start, uint32_t* __restrict__ cell_n) {
  const uint32_t rep = blockIdx.x;
  if (!rebuild_pending(a.flags, rep)) return;
  __shared__ uint32_t s_w[kBlock / 64];
  const rumdx_replica rp = a.rep[rep];
  const uint32_t ncell = static_cast<uint32_t>(rp.nc) * rp.nc * rp.nc;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t carry = rp.slot0;
  for (uint32_t base = 0; base < ncell; base += kBlock) {
    const uint32_t c = base + threadIdx.x;
    const uint32_t v = c < ncell ? cell_count[rp.cell0 + c] : 0u;
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
      off += k < wave ? s_w[k] : 0u;
      tot += s_w[k];
    }
    if (c < ncell) {
      cell_start[rp.cell0 + c] = off + x - v;
      cell_n[rp.cell0 + c] = v;
      cell_count[rp.cell0 + c] = 0u;
    }
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_b_scatter(BatchArgs a, const uint32_t* __restrict__ slot_cell,
                                                      const uint32_t* __restrict__ cell_start,
                                                      uint32_t* __restrict__ cell_count, uint32_t* __restrict__ cell_tmp) {
  const uint32_t rep = a.block_rep[blockIdx.x];
  if (!rebuild_pending(a.flags, rep)) return;
  const rumdx_replica rp = a.rep[rep];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i - rp.slot0 >= rp.n) return;
  const uint32_t c = slot_cell[i];
  cell_tmp[cell_start[c] + atomicAdd(cell_count + c, 1u)] = i;
}

// Sort staging: the particle arrays of rebuilding replicas in cell order.
struct SortStage {
  float4* r;
  float4* v;
  float4* im;
  uint32_t* idx;
};

// The same sort by one 1024-thread workgroup per rebuilding replica, in LDS:
// cell counts (then running cursors), their exclusive scan, the particles
// placed by cell, and each particle's rank among its cell's members by old
// slot (stable).  cell_start, cell_n, slot_cell and the staging arrays come
// out as from k_b_hist, k_b_scan, k_b_scatter and k_b_gather; cell_count is
// not touched (it stays zero).
constexpr uint32_t kSortBlock = 1024;
__global__ __launch_bounds__(kSortBlock) void k_b_sort(BatchArgs a, uint32_t* __restrict__ slot_cell,
                                                       uint32_t* __restrict__ cell_start,
                                                       uint32_t* __restrict__ cell_n,
                                                       const uint32_t* __restrict__ idx, SortStage st,
                                                       uint32_t max_cells) {
  const uint32_t rep = blockIdx.x;
  if (!rebuild_pending(a.flags, rep)) return;
  extern __shared__ uint32_t s_sort[];
  __shared__ uint32_t s_w[kSortBlock / 64];
  const rumdx_replica rp = a.rep[rep];
  const uint32_t ncell = static_cast<uint32_t>(rp.nc) * rp.nc * rp.nc;
  uint32_t* s_cnt = s_sort;                  // counts, then cursors (max_cells)
  uint32_t* s_start = s_sort + max_cells;    // replica-local cell starts (max_cells)
  uint32_t* s_tmp = s_sort + 2 * max_cells;  // old slots placed by cell (n)
  for (uint32_t c = threadIdx.x; c < ncell; c += kSortBlock) s_cnt[c] = 0u;
  __syncthreads();
  const Box bx = rep_box(rp);
  for (uint32_t k = threadIdx.x; k < rp.n; k += kSortBlock) {
    const uint32_t i = rp.slot0 + k;
    const uint32_t lc = rep_cell(a.r[i], bx, rp.nc);
    slot_cell[i] = rp.cell0 + lc;
    atomicAdd(s_cnt + lc, 1u);
  }
  __syncthreads();
  // exclusive scan of the counts, kSortBlock cells at a time
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < ncell; base += kSortBlock) {
    const uint32_t c = base + threadIdx.x;
    const uint32_t v = c < ncell ? s_cnt[c] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= static_cast<uint32_t>(o)) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t off = carry, tot = 0;
    for (uint32_t k = 0; k < kSortBlock / 64; ++k) {
      off += k < wave ? s_w[k] : 0u;
      tot += s_w[k];
    }
    if (c < ncell) {
      const uint32_t start = off + x - v;
      s_start[c] = start;
      cell_start[rp.cell0 + c] = rp.slot0 + start;
      cell_n[rp.cell0 + c] = v;
      s_cnt[c] = start;  // the cell's placement cursor
    }
    carry += tot;
    __syncthreads();
  }
  // place the old slots by cell (any order inside a cell), then rank
  for (uint32_t k = threadIdx.x; k < rp.n; k += kSortBlock) {
    const uint32_t i = rp.slot0 + k;
    const uint32_t lc = slot_cell[i] - rp.cell0;
    s_tmp[atomicAdd(s_cnt + lc, 1u)] = i;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < rp.n; k += kSortBlock) {
    const uint32_t x = s_tmp[k];
    const uint32_t lc = slot_cell[x] - rp.cell0;
    const uint32_t b0 = s_start[lc], m = (lc + 1 < ncell ? s_start[lc + 1] : rp.n) - b0;
    uint32_t d = b0;
    for (uint32_t j = b0; j < b0 + m; ++j) d += s_tmp[j] < x ? 1u : 0u;
    d += rp.slot0;
    st.r[d] = a.r[x];
    st.v[d] = a.v[x];
    st.im[d] = a.im[x];
    st.idx[d] = idx[x];
  }
}

// Gathers every particle of a rebuilding replica into its sorted slot: the
// cell's first slot plus the number of the cell's particles with a smaller
// old slot (a stable counting sort), and zeroes the counts for the next
// build.
__global__ __launch_bounds__(kBlock) void k_b_gather(BatchArgs a, const uint32_t* __restrict__ slot_cell,
                                                     const uint32_t* __restrict__ cell_start,
                                                     const uint32_t* __restrict__ cell_n,
                                                     uint32_t* __restrict__ cell_count,
                                                     const uint32_t* __restrict__ cell_tmp,
                                                     const uint32_t* __restrict__ idx, SortStage st) {
  const uint32_t rep = a.block_rep[blockIdx.x];
  if (!rebuild_pending(a.flags, rep)) return;
  const rumdx_replica rp = a.rep[rep];
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k - rp.slot0 >= rp.n) return;
  const uint32_t x = cell_tmp[k];
  const uint32_t c = slot_cell[x];
  const uint32_t b0 = cell_start[c], m = cell_n[c];
  uint32_t d = b0;
  for (uint32_t j = b0; j < b0 + m; ++j) d += cell_tmp[j] < x ? 1u : 0u;
  st.r[d] = a.r[x];
  st.v[d] = a.v[x];
  st.im[d] = a.im[x];
  st.idx[d] = idx[x];
  cell_count[c] = 0u;
}

// Stencil candidates of one cell in RUMD's (dZ, dY, dX) order, ascending slot
// inside each cell: 25 rows of five cells (each a contiguous range of sorted
// slots) when the cell's row does not wrap around the box, else 125 cells.
// seg gives segment sg's first slot and length; returns the count.
__device__ __forceinline__ uint32_t stencil_segment(const rumdx_replica& rp, const uint32_t* __restrict__ cell_start,
                                                    const uint32_t* __restrict__ cell_n, int cx, int cy, int cz,
                                                    bool rows, uint32_t sg, uint32_t& first) {
  const int nc = rp.nc;
  if (rows) {
    const int oz = (cz + static_cast<int>(sg / 5) - 2 + nc) % nc, oy = (cy + static_cast<int>(sg % 5) - 2 + nc) % nc;
    const uint32_t c_row = rp.cell0 + static_cast<uint32_t>(nc * (oy + nc * oz));
    const uint32_t lo = c_row + static_cast<uint32_t>(cx - 2), hi = c_row + static_cast<uint32_t>(cx + 2);
    first = cell_start[lo];
    return cell_start[hi] + cell_n[hi] - first;
  }
  const int oz = (cz + static_cast<int>(sg / 25) - 2 + nc) % nc;
  const int oy = (cy + static_cast<int>((sg / 5) % 5) - 2 + nc) % nc;
  const int ox = (cx + static_cast<int>(sg % 5) - 2 + nc) % nc;
  const uint32_t c = rp.cell0 + static_cast<uint32_t>(ox + nc * (oy + nc * oz));
  first = cell_start[c];
  return cell_n[c];
}

// Per-wave LDS of the batch build: the cell's stencil as a window of
// replica-local slots, the segment table, and one particle's list.
constexpr uint32_t kBatchWindow = 1024;
constexpr uint32_t kBatchStage = 256;
constexpr uint32_t kBatchCellsPerWave = RUMDX_BATCH_CELL_BLOCK / (kBlock / 64);
struct BatchWave {
  uint16_t cand[kBatchWindow];  // slot - slot0
  uint32_t seg_first[128];
  uint16_t seg_pre[128];
  uint16_t seg_len[128];
  uint32_t list[2][kBatchStage];  // the staged lists of the two particles tested together
};
static_assert(RUMDX_BATCH_CELL_BLOCK % (kBlock / 64) == 0, "whole cells per wave");

// Appends the accepted candidates of one 64-lane chunk to a staged list of
// cnt entries (ballot order = candidate order); returns the new count.
__device__ __forceinline__ uint32_t b_append(bool acc, uint32_t j, uint32_t cnt, uint32_t* list, uint32_t max_nbrs) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(acc);
  const uint32_t k = cnt + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
  if (acc && k < max_nbrs) list[k] = j;
  return cnt + static_cast<uint32_t>(__builtin_popcountll(mask));
}

// Tests the cell's window against particle a (slot ia) and, when TWO,
// particle b: each pass reads four 64-candidate chunks' slots and positions
// once for both.  Without WRAPS (a stencil that does not wrap, nc >= 7) the
// plain FP32 difference is the minimum-image one.
template <bool WRAPS, bool TWO>
__device__ __forceinline__ void b_scan_window(const BatchWave& wb, uint32_t* list_a, uint32_t* list_b,
                                              const float4* __restrict__ sr, uint32_t slot0, uint32_t total,
                                              float4 ra, float4 rb, const float* crow_a, const float* crow_b,
                                              uint32_t ia, uint32_t ib, const Box& bx, uint32_t max_nbrs,
                                              uint32_t& cnt_a, uint32_t& cnt_b) {
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t q0 = 0; q0 < total; q0 += 4 * 64) {
    uint32_t j[4];
    bool in[4];  // lanes past the window end hold slot ia and test nothing
    float4 rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t q = q0 + 64u * u + lane;
      in[u] = q < total;
      j[u] = in[u] ? slot0 + wb.cand[q] : ia;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) rj[u] = sr[j[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t tj = __float_as_uint(rj[u].w);
      const float da = WRAPS ? min_image_d2_f32(ra, rj[u], bx) : plain_d2_f32(ra, rj[u]);
      cnt_a = b_append(j[u] != ia && da < crow_a[tj], j[u], cnt_a, list_a, max_nbrs);
      if (TWO) {
        const float db = WRAPS ? min_image_d2_f32(rb, rj[u], bx) : plain_d2_f32(rb, rj[u]);
        cnt_b = b_append(in[u] && j[u] != ib && db < crow_b[tj], j[u], cnt_b, list_b, max_nbrs);
      }
    }
  }
}

// Window chunks a lane holds in registers while the cell's particles are
// tested (windows of up to 512 candidates): the candidates' slots and
// positions are read once per cell, not once per pair of particles.
constexpr uint32_t kRegChunks = 8;

// The test of particles a (and b when TWO) against a window held in
// registers (chunk u: candidates 64u + lane, `nch` chunks, `total`
// candidates); the same order and criterion as b_scan_window.
template <bool WRAPS, bool TWO>
__device__ __forceinline__ void b_test_held(const uint32_t (&j)[kRegChunks], const float4 (&rj)[kRegChunks],
                                            uint32_t nch, uint32_t total, uint32_t* list_a, uint32_t* list_b,
                                            float4 ra, float4 rb, const float* crow_a, const float* crow_b,
                                            uint32_t ia, uint32_t ib, const Box& bx, uint32_t max_nbrs,
                                            uint32_t& cnt_a, uint32_t& cnt_b) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (uint32_t u = 0; u < kRegChunks; ++u) {
    if (u < nch) {
      const bool in = 64u * u + lane < total;
      const uint32_t tj = __float_as_uint(rj[u].w);
      const float da = WRAPS ? min_image_d2_f32(ra, rj[u], bx) : plain_d2_f32(ra, rj[u]);
      cnt_a = b_append(in && j[u] != ia && da < crow_a[tj], j[u], cnt_a, list_a, max_nbrs);
      if (TWO) {
        const float db = WRAPS ? min_image_d2_f32(rb, rj[u], bx) : plain_d2_f32(rb, rj[u]);
        cnt_b = b_append(in && j[u] != ib && db < crow_b[tj], j[u], cnt_b, list_b, max_nbrs);
      }
    }
  }
}

// The lists of a rebuilding replica from its sorted staging, one wave per
// cell (RUMDX_BATCH_CELL_BLOCK cells per workgroup, all of one replica): the
// cell's stencil goes into an LDS window once, and each of its particles
// tests it 64 candidates at a time (FP32 minimum-image criterion), a ballot
// placing the accepted ones in candidate order.  A cell whose stencil does
// not fit the window, or lists longer than the staging, is built one lane
// per particle.  Then the cell's particles move to their sorted slots.  The
// last workgroup of the replica clears its rebuild request.
__global__ __launch_bounds__(kBlock) void k_b_build(BatchArgs a, const uint32_t* __restrict__ cellblock_rep,
                                                    uint32_t* __restrict__ slot_cell,
                                                    const uint32_t* __restrict__ cell_start,
                                                    const uint32_t* __restrict__ cell_n, SortStage st,
                                                    uint32_t* __restrict__ idx, const float* __restrict__ cut_skin2,
                                                    uint32_t nt, uint32_t* __restrict__ nbl,
                                                    uint32_t* __restrict__ num_nbrs, float4* __restrict__ last_r,
                                                    uint32_t max_nbrs) {
  const uint32_t rep = cellblock_rep[blockIdx.x];
  if (!rebuild_pending(a.flags, rep)) return;
  __shared__ float s_cut[RUMDX_MAX_TYPES * RUMDX_MAX_TYPES];
  __shared__ BatchWave s_wave[kBlock / 64];
  __shared__ uint32_t s_max[kBlock / 64];
  for (uint32_t k = threadIdx.x; k < nt * nt; k += kBlock) s_cut[k] = cut_skin2[k];
  __syncthreads();
  const rumdx_replica rp = a.rep[rep];
  uint32_t* fl = a.flags + kRepFlags * rep;
  const Box bx = rep_box(rp);
  const int nc = rp.nc;
  const uint32_t ncell = static_cast<uint32_t>(nc) * nc * nc;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  BatchWave& wb = s_wave[wave];
  uint32_t longest = 0;
  for (uint32_t cc = 0; cc < kBatchCellsPerWave; ++cc) {
    const uint32_t local = blockIdx.x * RUMDX_BATCH_CELL_BLOCK - rp.cell0 + wave * kBatchCellsPerWave + cc;
    if (local >= ncell) break;
    const uint32_t c = rp.cell0 + local;
    const uint32_t m = __builtin_amdgcn_readfirstlane(cell_n[c]);
    const uint32_t cs = __builtin_amdgcn_readfirstlane(cell_start[c]);
    if (m == 0) continue;
    const int cx = static_cast<int>(local % nc), cy = static_cast<int>((local / nc) % nc),
              cz = static_cast<int>(local / (nc * nc));
    const bool rows = cx >= 2 && cx + 2 < nc;
    const uint32_t nseg = rows ? 25u : 125u;
    // segment table and its exclusive prefix
    uint32_t carry = 0;
    for (uint32_t s0 = 0; s0 < nseg; s0 += 64) {
      const uint32_t sg = s0 + lane;
      uint32_t first = 0, len = 0;
      if (sg < nseg) len = stencil_segment(rp, cell_start, cell_n, cx, cy, cz, rows, sg, first);
      uint32_t x = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= static_cast<uint32_t>(o)) x += y;
      }
      if (sg < nseg) {
        wb.seg_first[sg] = first;
        wb.seg_pre[sg] = static_cast<uint16_t>(carry + x - len < 0xffffu ? carry + x - len : 0xffffu);
        wb.seg_len[sg] = static_cast<uint16_t>(len < 0xffffu ? len : 0xffffu);
      }
      carry += __shfl(x, 63, 64);
    }
    const uint32_t total = __builtin_amdgcn_readfirstlane(carry);
    if (total <= kBatchWindow && max_nbrs <= kBatchStage && rp.n <= 65536u) {  // u16 window entries
      wave_sync_lds();
      for (uint32_t sg = lane; sg < nseg; sg += 64) {
        const uint32_t pre = wb.seg_pre[sg], len = wb.seg_len[sg], first = wb.seg_first[sg];
        for (uint32_t q = 0; q < len; ++q) wb.cand[pre + q] = static_cast<uint16_t>(first + q - rp.slot0);
      }
      wave_sync_lds();
      const bool wraps = stencil_wraps(cx, nc) || stencil_wraps(cy, nc) || stencil_wraps(cz, nc);
      // a window of at most kRegChunks chunks stays in registers for all of
      // the cell's particles (lanes past its end hold the cell's first slot)
      const bool held = total <= 64u * kRegChunks;
      const uint32_t nch = (total + 63u) >> 6;
      uint32_t hj[kRegChunks];
      float4 hr[kRegChunks];
#pragma unroll
      for (uint32_t u = 0; u < kRegChunks; ++u) {
        hj[u] = cs;
        if (held && u < nch) {
          const uint32_t q = 64u * u + lane;
          hj[u] = q < total ? rp.slot0 + wb.cand[q] : cs;
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kRegChunks; ++u) hr[u] = (held && u < nch) ? st.r[hj[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
      for (uint32_t p = 0; p < m; p += 2) {
        const bool two = p + 1 < m;
        const uint32_t ia = cs + p, ib = two ? ia + 1 : ia;
        const float4 ra = st.r[ia], rb = st.r[ib];
        const float* crow_a = s_cut + __float_as_uint(ra.w) * nt;
        const float* crow_b = s_cut + __float_as_uint(rb.w) * nt;
        uint32_t cnt_a = 0, cnt_b = 0;
        if (held) {
          if (wraps) {
            if (two)
              b_test_held<true, true>(hj, hr, nch, total, wb.list[0], wb.list[1], ra, rb, crow_a, crow_b, ia, ib, bx,
                                      max_nbrs, cnt_a, cnt_b);
            else
              b_test_held<true, false>(hj, hr, nch, total, wb.list[0], wb.list[1], ra, rb, crow_a, crow_b, ia, ib, bx,
                                       max_nbrs, cnt_a, cnt_b);
          } else {
            if (two)
              b_test_held<false, true>(hj, hr, nch, total, wb.list[0], wb.list[1], ra, rb, crow_a, crow_b, ia, ib, bx,
                                       max_nbrs, cnt_a, cnt_b);
            else
              b_test_held<false, false>(hj, hr, nch, total, wb.list[0], wb.list[1], ra, rb, crow_a, crow_b, ia, ib,
                                        bx, max_nbrs, cnt_a, cnt_b);
          }
        } else if (wraps) {
          if (two)
            b_scan_window<true, true>(wb, wb.list[0], wb.list[1], st.r, rp.slot0, total, ra, rb, crow_a, crow_b, ia,
                                      ib, bx, max_nbrs, cnt_a, cnt_b);
          else
            b_scan_window<true, false>(wb, wb.list[0], wb.list[1], st.r, rp.slot0, total, ra, rb, crow_a, crow_b, ia,
                                       ib, bx, max_nbrs, cnt_a, cnt_b);
        } else {
          if (two)
            b_scan_window<false, true>(wb, wb.list[0], wb.list[1], st.r, rp.slot0, total, ra, rb, crow_a, crow_b, ia,
                                       ib, bx, max_nbrs, cnt_a, cnt_b);
          else
            b_scan_window<false, false>(wb, wb.list[0], wb.list[1], st.r, rp.slot0, total, ra, rb, crow_a, crow_b,
                                        ia, ib, bx, max_nbrs, cnt_a, cnt_b);
        }
        wave_sync_lds();
        for (int t = 0; t < (two ? 2 : 1); ++t) {
          const uint32_t i = t ? ib : ia, cnt = t ? cnt_b : cnt_a;
          const uint32_t stored = cnt < max_nbrs ? cnt : max_nbrs;
          for (uint32_t k = lane; k < stored; k += 64) nbl[static_cast<size_t>(k) * a.nslots + i] = wb.list[t][k];
          if (lane == 0) {
            num_nbrs[i] = stored;
            if (cnt >= max_nbrs) fl[kOverflow] = 1u;
          }
          longest = cnt > longest ? cnt : longest;
        }
        wave_sync_lds();  // the flushes have read the lists
      }
    } else {
      // one lane per particle over the segments (any stencil size)
      for (uint32_t p = lane; p < m; p += 64) {
        const uint32_t i = cs + p;
        const float4 ri = st.r[i];
        const float* crow = s_cut + __float_as_uint(ri.w) * nt;
        uint32_t cnt = 0;
        for (uint32_t sg = 0; sg < nseg; ++sg) {
          uint32_t first;
          const uint32_t len = stencil_segment(rp, cell_start, cell_n, cx, cy, cz, rows, sg, first);
          for (uint32_t j = first; j < first + len; ++j) {
            const float4 rj = st.r[j];
            if (j != i && min_image_d2_f32(ri, rj, bx) < crow[__float_as_uint(rj.w)]) {
              if (cnt < max_nbrs) nbl[static_cast<size_t>(cnt) * a.nslots + i] = j;
              ++cnt;  // keeps counting past capacity: the true length sizes the regrowth
            }
          }
        }
        num_nbrs[i] = cnt < max_nbrs ? cnt : max_nbrs;
        if (cnt >= max_nbrs) fl[kOverflow] = 1u;
        longest = cnt > longest ? cnt : longest;
      }
    }
    // the cell's particles move to their sorted slots
    for (uint32_t p = lane; p < m; p += 64) {
      const uint32_t i = cs + p;
      const float4 ri = st.r[i];
      a.r[i] = ri;
      a.fixed[i] = to_fixed(ri, fixed_scale(rp));
      a.v[i] = st.v[i];
      a.im[i] = st.im[i];
      idx[i] = st.idx[i];
      slot_cell[i] = c;
      last_r[i] = ri;
    }
  }
  longest = block_max_u32(longest, s_max);
  if (threadIdx.x == 0) atomicMax(fl + kMaxLen, longest);
  const uint32_t nblk = (ncell + RUMDX_BATCH_CELL_BLOCK - 1) / RUMDX_BATCH_CELL_BLOCK;
  if (last_to_arrive(fl + kDone, nblk) && threadIdx.x == 0) {
    fl[kRebuild] = 0u;
    fl[kBuilds] += 1u;
  }
}

// ------------------------------------------------------------------ forces
// FP32 pair arithmetic with the minimum-image displacement taken exactly on
// the fixed-point positions (to_fixed above) and rounded once to FP32.
struct Acc {
  float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f, vir = 0.f;
};

// One axis of the displacement: the wrapping difference of the fixed-point
// coordinates (the minimum image, exactly) scaled by L 2^-32 in FP64 (exact:
// a 32-bit integer times a float over a power of two) and rounded once to FP32,
// so the displacement is the correctly rounded one.
__device__ __forceinline__ float fixed_delta(int32_t a, int32_t b, double unit) {
  return static_cast<float>(
      static_cast<double>(static_cast<int32_t>(static_cast<uint32_t>(a) - static_cast<uint32_t>(b))) * unit);
}

template <int CUT>
__device__ __forceinline__ void pair_b(int4 fi, int4 fj, double unit, const float4* prow, Acc& a) {
  const float dx = fixed_delta(fi.x, fj.x, unit);
  const float dy = fixed_delta(fi.y, fj.y, unit);
  const float dz = fixed_delta(fi.z, fj.z, unit);
  const float d2 = dx * dx + dy * dy + dz * dz;
  const float4 pa = prow[2 * fj.w];
  const float4 pb = prow[2 * fj.w + 1];
  // 1/r^2 from its own reciprocal (1 ulp) rather than the square of 1/r
  // (2.5 ulp): the r^-14 force term multiplies that error by seven.  The
  // batch kernel waits on gathers more than on VALU, so the second
  // transcendental costs little here.
  float rinv = 0.f;
  const float inv2 = __builtin_amdgcn_rcpf(d2);
  if (CUT == RUMDX_CUTOFF_SF) rinv = __builtin_amdgcn_rsqf(d2);
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
  a.fx += dx * s;
  a.fy += dy * s;
  a.fz += dz * s;
  a.u += u;
  a.vir += d2 * s;
}

struct CellArrays {
  const uint32_t* slot_cell;
  const uint32_t* cell_start;
  const uint32_t* cell_n;
};

// The LDS table of the LJ pair term: per type pair {rc^2, A12, A6, s(rc) rc}
// and {12 A12, 6 A6, -u(rc), -(u(rc) + s(rc) rc^2)}.
__device__ __forceinline__ void load_lj_table(float4* s_tab, const float* __restrict__ params, uint32_t nt) {
  for (uint32_t k = threadIdx.x; k < nt * nt; k += blockDim.x) {
    const float* q = params + static_cast<size_t>(k) * RUMDX_NUM_PARAM;
    const float rc = q[0], a12 = q[1], a6 = q[2];
    const float src = q[RUMDX_NUM_PARAM - 2], ucut = q[RUMDX_NUM_PARAM - 1];
    s_tab[2 * k] = make_float4(rc * rc, a12, a6, src);
    s_tab[2 * k + 1] = make_float4(12.f * a12, 6.f * a6, -ucut, -fmaf(src, rc, ucut));
  }
}

// Neighbour positions read from the batch's fixed-point array.
struct GlobalFixed {
  const int4* fixed;
  __device__ __forceinline__ int4 operator()(uint32_t j) const { return fixed[j]; }
};

// Neighbour positions read from a replica staged in LDS (slot j at s[j - slot0]).
struct StagedFixed {
  const int4* s;
  uint32_t slot0;
  __device__ __forceinline__ int4 operator()(uint32_t j) const { return s[j - slot0]; }
};

// One lane's share of slot i's pair terms: entries part, part + 2^lgp, ... of
// its list, four loads in flight, positions through `pos`.  An overflowed
// list rescans the build-time stencil (from global memory) in its first lane.
template <int CUT, class Pos>
__device__ __forceinline__ void slot_terms(const BatchArgs& a, const rumdx_replica& rp, uint32_t i, int4 fi,
                                           uint32_t part, uint32_t lgp, const uint32_t* __restrict__ nbl,
                                           const uint32_t* __restrict__ num_nbrs, uint32_t max_nbrs,
                                           const CellArrays& cells, const float4* prow, const Pos& pos, Acc& acc) {
  const uint32_t step = 1u << lgp;
  const double unit = static_cast<double>(rp.L) * 2.3283064365386963e-10;  // L 2^-32, exact
  const uint32_t nn = num_nbrs[i];
  if (nn >= max_nbrs) {
    if (part == 0) {
      // overflowed list: rescan the build-time stencil (every pair now
      // inside rc was inside rc + skin then)
      const int nc = rp.nc;
      const uint32_t lc = cells.slot_cell[i] - rp.cell0;
      const int cx = static_cast<int>(lc % nc), cy = static_cast<int>((lc / nc) % nc),
                cz = static_cast<int>(lc / (nc * nc));
      for (int dz = -2; dz <= 2; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
          for (int dx = -2; dx <= 2; ++dx) {
            const uint32_t c = rp.cell0 + static_cast<uint32_t>((cx + dx + nc) % nc +
                                                                nc * ((cy + dy + nc) % nc + nc * ((cz + dz + nc) % nc)));
            const uint32_t b0 = cells.cell_start[c], m = cells.cell_n[c];
            for (uint32_t j = b0; j < b0 + m; ++j)
              if (j != i) pair_b<CUT>(fi, a.fixed[j], unit, prow, acc);
          }
    }
    return;
  }
  // eight list entries (and their positions) in flight per pass: the list
  // reads come from HBM, and fewer waves than that would leave their latency
  // exposed at one staged workgroup per CU
  const size_t stride = static_cast<size_t>(a.nslots) << lgp;
  const uint32_t* col = nbl + static_cast<size_t>(part) * a.nslots + i;
  uint32_t k = part;
  for (; k + 7 * step < nn; k += 8 * step) {
    uint32_t j[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) j[u] = col[u * stride];
    col += 8 * stride;
    int4 fj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) fj[u] = pos(j[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u) pair_b<CUT>(fi, fj[u], unit, prow, acc);
  }
  for (; k < nn; k += step, col += stride) pair_b<CUT>(fi, pos(*col), unit, prow, acc);
}

// Sums the 2^lgp lanes of each slot (a fixed butterfly); every lane of the
// wave takes part.
__device__ __forceinline__ void combine_slot_lanes(Acc& acc, uint32_t lgp) {
  for (uint32_t o = 1; o < (1u << lgp); o <<= 1) {
    acc.fx += __shfl_xor(acc.fx, o, 64);
    acc.fy += __shfl_xor(acc.fy, o, 64);
    acc.fz += __shfl_xor(acc.fz, o, 64);
    acc.u += __shfl_xor(acc.u, o, 64);
    acc.vir += __shfl_xor(acc.vir, o, 64);
  }
}

__device__ __forceinline__ void store_slot(const BatchArgs& a, uint32_t i, const Acc& acc) {
  a.f[i] = make_float4(acc.fx, acc.fy, acc.fz, 0.5f * acc.u);
  a.w[i] = make_float4(0.f, 0.f, 0.f, acc.vir);
}

// One particle's list over 2^lgp lanes (lane part takes entries part,
// part + 2^lgp, ...), combined by a fixed butterfly: the workgroup covers
// kBlock >> lgp slots, all of one replica (its 256-slot block); neighbour
// positions are gathered through L2.
template <int CUT>
__global__ __launch_bounds__(kBlock) void k_b_force(BatchArgs a, const uint32_t* __restrict__ nbl,
                                                    const uint32_t* __restrict__ num_nbrs, uint32_t max_nbrs,
                                                    CellArrays cells, const float* __restrict__ params,
                                                    uint32_t nt, uint32_t lgp) {
  extern __shared__ float4 s_tab[];
  load_lj_table(s_tab, params, nt);
  __syncthreads();
  const uint32_t step = 1u << lgp;
  const uint32_t i = (blockIdx.x * kBlock + threadIdx.x) >> lgp, part = threadIdx.x & (step - 1u);
  const uint32_t rep = a.block_rep[i / kBlock];
  const rumdx_replica rp = a.rep[rep];
  const bool live = i - rp.slot0 < rp.n;
  Acc acc;
  if (live) {
    const int4 fi = a.fixed[i];
    slot_terms<CUT>(a, rp, i, fi, part, lgp, nbl, num_nbrs, max_nbrs, cells, s_tab + 2 * fi.w * nt,
                    GlobalFixed{a.fixed}, acc);
  }
  combine_slot_lanes(acc, lgp);
  if (live && part == 0) store_slot(a, i, acc);
}

// The same pair terms with the replica staged: a workgroup takes one work
// item (replica, first slot, end slot), copies the replica's fixed-point
// positions into LDS, and walks the item's slots kStagedBlock >> lgp at a
// time, reading every neighbour from LDS.  Lane mapping and summation order
// are those of k_b_force, so the forces are the same bits.
//
// With FUSE every work item is a whole replica, and the workgroup then also
// takes the replica's NVT step (k_b_nvt's per-particle update and its
// per-256-slot kinetic-energy partials, in the same order): nothing else
// reads this replica's positions during the launch, so they are updated in
// place.
constexpr uint32_t kStagedBlock = 1024;
static_assert(kStagedBlock <= kMaxWorkgroup, "staged force workgroup");
static_assert(kStagedBlock % kBlock == 0, "whole integrator blocks per pass");

struct BatchNvt {
  const rumdx_nh_state* nh;
  const float4* last_r;
  double* partials;  // one per 256-slot block, as k_b_nvt writes them
  float dt;
  float skin2_over_4;
};

template <int CUT, bool FUSE = false>
__global__ __launch_bounds__(kStagedBlock) void k_b_force_staged(BatchArgs a, const uint32_t* __restrict__ nbl,
                                                                 const uint32_t* __restrict__ num_nbrs,
                                                                 uint32_t max_nbrs, CellArrays cells,
                                                                 const float* __restrict__ params, uint32_t nt,
                                                                 uint32_t lgp, const uint4* __restrict__ work,
                                                                 BatchNvt nv) {
  extern __shared__ float4 s_dyn[];
  float4* s_tab = s_dyn;
  int4* s_fix = reinterpret_cast<int4*>(s_dyn + 2 * nt * nt);
  load_lj_table(s_tab, params, nt);
  const uint4 wk = work[blockIdx.x];
  const rumdx_replica rp = a.rep[wk.x];
  for (uint32_t k = threadIdx.x; k < rp.n; k += kStagedBlock) s_fix[k] = a.fixed[rp.slot0 + k];
  __syncthreads();
  const StagedFixed pos{s_fix, rp.slot0};
  const uint32_t step = 1u << lgp, part = threadIdx.x & (step - 1u);
  for (uint32_t base = wk.y; base < wk.z; base += kStagedBlock >> lgp) {
    const uint32_t i = base + (threadIdx.x >> lgp);
    const bool live = i < wk.z;
    Acc acc;
    if (live) {
      const int4 fi = pos(i);
      slot_terms<CUT>(a, rp, i, fi, part, lgp, nbl, num_nbrs, max_nbrs, cells, s_tab + 2 * fi.w * nt, pos, acc);
    }
    combine_slot_lanes(acc, lgp);
    if (live && part == 0) store_slot(a, i, acc);
  }
  if (FUSE) {
    // the replica's forces are stored; its slots in 256-slot blocks, each
    // block's m v^2 summed as k_b_nvt's block_sum does (its four waves' sums
    // in order)
    __syncthreads();
    double* s_w = reinterpret_cast<double*>(s_dyn);  // the staging is no longer read
    const Box bx = rep_box(rp);
    const float ps = rp.omega2 > 0.f ? static_cast<float>(nv.nh[wk.x].ps) : 0.f;
    const uint32_t nb = (rp.n + kBlock - 1) / kBlock;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < nb; b0 += kStagedBlock / kBlock) {
      const uint32_t i = rp.slot0 + b0 * kBlock + threadIdx.x;
      float ke2 = 0.f;
      if (i - rp.slot0 < rp.n) {
        float4 r = a.r[i], v = a.v[i];
        float4 lo = make_float4(0.f, 0.f, 0.f, 0.f);
        float3 shift;
        ke2 = nvt_particle<false>(r, v, a.f[i], shift, lo, ps, nv.dt, bx);
        a.r[i] = r;
        a.fixed[i] = to_fixed(r, fixed_scale(rp));
        a.v[i] = v;
        add_image_shift(a.im, i, shift);
        if (min_image_d2_f32(r, nv.last_r[i], bx) > nv.skin2_over_4) a.flags[kRepFlags * wk.x + kRebuild] = 1u;
      }
      const double wsum = wave_sum(static_cast<double>(ke2));
      if (lane == 0) s_w[wave] = wsum;
      __syncthreads();
      const uint32_t q = threadIdx.x;  // block b0 + q, q < 4
      if (q < kStagedBlock / kBlock && b0 + q < nb) {
        double sum = 0.0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) sum += s_w[q * (kBlock / 64) + k];
        nv.partials[rp.slot0 / kBlock + b0 + q] = sum;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ integrator
// Leapfrog + Nose-Hoover per replica, the skin check fused in, and one FP64
// partial of sum m v^2 per workgroup.
__global__ __launch_bounds__(kBlock) void k_b_nvt(BatchArgs a, const rumdx_nh_state* __restrict__ nh,
                                                  const float4* __restrict__ last_r, float dt, float skin2_over_4,
                                                  double* __restrict__ partials) {
  __shared__ double s_red[kBlock / 64];
  const uint32_t rep = a.block_rep[blockIdx.x];
  const rumdx_replica rp = a.rep[rep];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  float ke2 = 0.f;
  if (i - rp.slot0 < rp.n) {
    const Box bx = rep_box(rp);
    const float ps = rp.omega2 > 0.f ? static_cast<float>(nh[rep].ps) : 0.f;
    float4 r = a.r[i], v = a.v[i];
    float4 lo = make_float4(0.f, 0.f, 0.f, 0.f);
    float3 shift;
    ke2 = nvt_particle<false>(r, v, a.f[i], shift, lo, ps, dt, bx);
    a.r[i] = r;
    a.fixed[i] = to_fixed(r, fixed_scale(rp));
    a.v[i] = v;
    add_image_shift(a.im, i, shift);
    if (min_image_d2_f32(r, last_r[i], bx) > skin2_over_4) a.flags[kRepFlags * rep + kRebuild] = 1u;
  }
  const double s = block_sum(static_cast<double>(ke2), s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// One workgroup per replica: its block partials in order, then
// Ps += dt omega^2 min(KE2 / (dof T) - 1, 10) in FP64.
__global__ __launch_bounds__(kBlock) void k_b_nvt_finish(const rumdx_replica* __restrict__ reps,
                                                         const double* __restrict__ partials,
                                                         rumdx_nh_state* __restrict__ nh, float dt) {
  __shared__ double s_red[kBlock / 64];
  const rumdx_replica rp = reps[blockIdx.x];
  const uint32_t b0 = rp.slot0 / kBlock, nb = (rp.n + kBlock - 1) / kBlock;
  const double tot = sum_partials(partials + b0, nb, 1, s_red);
  if (threadIdx.x == 0) {
    rumdx_nh_state& s = nh[blockIdx.x];
    s.ke2_sum = tot;
    if (rp.omega2 > 0.f) {
      double dev = tot / (static_cast<double>(rp.dof) * s.target_T) - 1.0;
      if (dev > 10.0) dev = 10.0;  // runaway clamp (src/IntegratorNVT.cu:432-434)
      s.ps += static_cast<double>(dt) * static_cast<double>(rp.omega2) * dev;
      s.target_T += s.dT_per_step;
    }
  }
}

// ------------------------------------------------------------------ energies
// Per-workgroup partials of {sum u_i, sum w_i, leapfrog kinetic energy
// (IntegratorNVT::GetKineticEnergy, src/IntegratorNVT.cu:316-356), sum m v^2}.
__global__ __launch_bounds__(kBlock) void k_b_energy(BatchArgs a, const rumdx_nh_state* __restrict__ nh, float dt,
                                                     double* __restrict__ partials) {
  __shared__ double s_red[kBlock / 64];
  const uint32_t rep = a.block_rep[blockIdx.x];
  const rumdx_replica rp = a.rep[rep];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  double c[4] = {0.0, 0.0, 0.0, 0.0};
  if (i - rp.slot0 < rp.n) {
    const float4 f = a.f[i], w = a.w[i], v = a.v[i];
    const double m = 1.0 / static_cast<double>(v.w);
    const double ps = rp.omega2 > 0.f ? nh[rep].ps : 0.0;
    // v(t + h/2) predicted from the current force, as the reference does
    const double factor = 0.5 * ps * dt, plus = 1.0 / (1.0 + factor), minus = 1.0 - factor;
    const double vx = plus * (minus * v.x + v.w * f.x * dt), vy = plus * (minus * v.y + v.w * f.y * dt),
                 vz = plus * (minus * v.z + v.w * f.z * dt);
    const double vm2 = static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y + static_cast<double>(v.z) * v.z;
    const double vp2 = vx * vx + vy * vy + vz * vz;
    c[0] = f.w;
    c[1] = w.w;
    c[2] = 0.25 * m * (vm2 + vp2);
    c[3] = m * vm2;
  }
  for (int k = 0; k < 4; ++k) {
    const double s = block_sum(c[k], s_red);
    if (threadIdx.x == 0) partials[4 * static_cast<size_t>(blockIdx.x) + k] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_b_energy_finish(const rumdx_replica* __restrict__ reps,
                                                            const double* __restrict__ partials,
                                                            double* __restrict__ out) {
  __shared__ double s_red[kBlock / 64];
  const rumdx_replica rp = reps[blockIdx.x];
  const uint32_t b0 = rp.slot0 / kBlock, nb = (rp.n + kBlock - 1) / kBlock;
  for (int k = 0; k < 4; ++k) {
    const double s = sum_partials(partials + 4 * static_cast<size_t>(b0) + k, nb, 4, s_red);
    if (threadIdx.x == 0) out[4 * static_cast<size_t>(blockIdx.x) + k] = s;
  }
}

__global__ void k_b_request(uint32_t* flags, uint32_t nrep) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrep; r += gridDim.x * blockDim.x)
    flags[kRepFlags * r + kRebuild] = 1u;
}

int check_batch(const rumdx_batch* b, const char* who) {
  if (b == nullptr) return fail(RUMDX_E_INVALID, std::string(who) + ": null batch");
  if (b->r_dev == nullptr || b->v_dev == nullptr || b->f_dev == nullptr || b->w_dev == nullptr ||
      b->im_dev == nullptr || b->rep_dev == nullptr || b->block_rep_dev == nullptr || b->rep_flags_dev == nullptr ||
      b->fixed_dev == nullptr)
    return fail(RUMDX_E_INVALID, std::string(who) + ": batch arrays not allocated");
  if (b->nslots == 0 || b->nslots % RUMDX_BATCH_ALIGN != 0 || b->nrep == 0)
    return fail(RUMDX_E_INVALID, std::string(who) + ": nslots must be a positive multiple of RUMDX_BATCH_ALIGN");
  return RUMDX_OK;
}

}  // namespace
}  // namespace rumdx

using namespace rumdx;

extern "C" size_t rumdx_batch_workspace_bytes(uint32_t nslots) {
  return sizeof(double) * 4 * (static_cast<size_t>(nslots) / RUMDX_BATCH_ALIGN + 1);
}

extern "C" int rumdx_batch_request_rebuild(rumdx_batch* b, void* stream) {
  int rc = check_batch(b, "rumdx_batch_request_rebuild");
  if (rc != RUMDX_OK) return rc;
  hipLaunchKernelGGL(k_b_request, dim3(blocks_for(b->nrep)), dim3(kBlock), 0, as_stream(stream), b->rep_flags_dev,
                     b->nrep);
  return check_launch("rumdx_batch_request_rebuild");
}

extern "C" int rumdx_batch_update_lists(rumdx_batch* b, void* stream) {
  int rc = check_batch(b, "rumdx_batch_update_lists");
  if (rc != RUMDX_OK) return rc;
  if (b->nbl_dev == nullptr || b->num_nbrs_dev == nullptr || b->last_r_dev == nullptr || b->cut_skin2_dev == nullptr ||
      b->cell_count_dev == nullptr || b->cell_start_dev == nullptr || b->cell_n_dev == nullptr ||
      b->cell_tmp_dev == nullptr || b->slot_cell_dev == nullptr || b->idx_dev == nullptr || b->t_r_dev == nullptr ||
      b->t_v_dev == nullptr || b->t_im_dev == nullptr || b->t_idx_dev == nullptr || b->fixed_dev == nullptr ||
      b->cellblock_rep_dev == nullptr || b->ncells == 0 || b->ncells % RUMDX_BATCH_CELL_BLOCK != 0)
    return fail(RUMDX_E_INVALID, "rumdx_batch_update_lists: list arrays not allocated");
  if (b->max_nbrs == 0 || b->ntypes == 0 || b->ntypes > RUMDX_MAX_TYPES)
    return fail(RUMDX_E_INVALID, "rumdx_batch_update_lists: max_nbrs or ntypes out of range");
  if (static_cast<uint64_t>(b->max_nbrs) * b->nslots >= (1ull << 32))
    return fail(RUMDX_E_INVALID, "rumdx_batch_update_lists: max_nbrs * nslots exceeds 2^32 entries");
  const BatchArgs a = args_of(b);
  const hipStream_t st = as_stream(stream);
  const uint32_t nblk = b->nslots / kBlock;
  const SortStage stage{reinterpret_cast<float4*>(b->t_r_dev), reinterpret_cast<float4*>(b->t_v_dev),
                        reinterpret_cast<float4*>(b->t_im_dev), b->t_idx_dev};
  const size_t sort_lds = sizeof(uint32_t) * (static_cast<size_t>(b->max_rep_n) + 2ull * b->max_rep_cells);
  if (b->max_rep_cells != 0) {
    if (b->max_rep_n == 0 || sort_lds > 65536u)
      return fail(RUMDX_E_INVALID, "rumdx_batch_update_lists: the one-workgroup sort needs at most 64 KiB of LDS");
    hipLaunchKernelGGL(k_b_sort, dim3(b->nrep), dim3(kSortBlock), sort_lds, st, a, b->slot_cell_dev,
                       b->cell_start_dev, b->cell_n_dev, b->idx_dev, stage, b->max_rep_cells);
  } else {
    hipLaunchKernelGGL(k_b_hist, dim3(nblk), dim3(kBlock), 0, st, a, b->slot_cell_dev, b->cell_count_dev);
    hipLaunchKernelGGL(k_b_scan, dim3(b->nrep), dim3(kBlock), 0, st, a, b->cell_count_dev, b->cell_start_dev,
                       b->cell_n_dev);
    hipLaunchKernelGGL(k_b_scatter, dim3(nblk), dim3(kBlock), 0, st, a, b->slot_cell_dev, b->cell_start_dev,
                       b->cell_count_dev, b->cell_tmp_dev);
    hipLaunchKernelGGL(k_b_gather, dim3(nblk), dim3(kBlock), 0, st, a, b->slot_cell_dev, b->cell_start_dev,
                       b->cell_n_dev, b->cell_count_dev, b->cell_tmp_dev, b->idx_dev, stage);
  }
  hipLaunchKernelGGL(k_b_build, dim3(b->ncells / RUMDX_BATCH_CELL_BLOCK), dim3(kBlock), 0, st, a,
                     b->cellblock_rep_dev, b->slot_cell_dev, b->cell_start_dev, b->cell_n_dev, stage, b->idx_dev,
                     b->cut_skin2_dev, b->ntypes, b->nbl_dev, b->num_nbrs_dev,
                     reinterpret_cast<float4*>(b->last_r_dev), b->max_nbrs);
  return check_launch("rumdx_batch_update_lists");
}

namespace rumdx {
namespace {

// Argument checks shared by the batch force entry points; RUMDX_OK or the
// failure code (with the message set).
int check_batch_forces(const rumdx_batch* b, const float* params_dev, const char* entry) {
  const std::string e(entry);
  int rc = check_batch(b, entry);
  if (rc != RUMDX_OK) return rc;
  if (params_dev == nullptr || b->nbl_dev == nullptr || b->num_nbrs_dev == nullptr || b->fixed_dev == nullptr)
    return fail(RUMDX_E_INVALID, e + ": null params, list or fixed-point positions");
  if (b->ntypes == 0 || b->ntypes > RUMDX_MAX_TYPES) return fail(RUMDX_E_INVALID, e + ": ntypes out of range");
  if (b->slot_cell_dev == nullptr || b->cell_start_dev == nullptr || b->cell_n_dev == nullptr || b->max_nbrs == 0)
    return fail(RUMDX_E_INVALID, e + ": cell arrays not allocated");
  const uint32_t lpp = b->lanes_per_particle;
  if (lpp > 8u || (lpp & (lpp - 1u)) != 0u)
    return fail(RUMDX_E_INVALID, e + ": lanes_per_particle must be 0, 1, 2, 4 or 8");
  if (b->force_work_dev != nullptr && (b->force_work_n == 0 || b->max_rep_n == 0))
    return fail(RUMDX_E_INVALID, e + ": empty staged work list");
  return RUMDX_OK;
}

// The replica-staged force kernel over the batch's work list (with FUSE, the
// NVT update too).
template <bool FUSE>
int launch_staged(const rumdx_batch* b, const float* params_dev, int cutoff_method, const BatchNvt& nv,
                  hipStream_t st, const char* entry) {
  const uint32_t lpp = b->lanes_per_particle;
  const uint32_t lgp = lpp >= 8u ? 3u : lpp >= 4u ? 2u : lpp >= 2u ? 1u : 0u;
  const BatchArgs a = args_of(b);
  const CellArrays cells{b->slot_cell_dev, b->cell_start_dev, b->cell_n_dev};
  size_t lds_staged = sizeof(float4) * 2 * b->ntypes * b->ntypes + sizeof(int4) * b->max_rep_n;
  if (FUSE) lds_staged = std::max(lds_staged, sizeof(double) * (kStagedBlock / 64));  // the update's wave sums
  const auto* work = reinterpret_cast<const uint4*>(b->force_work_dev);
  const auto launch = [&](auto kernel) -> int {
    // above 64 KiB a workgroup's dynamic LDS must be opted into
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_staged));
    if (e != hipSuccess)
      return fail(RUMDX_E_INVALID, std::string(entry) + ": staged LDS of " + std::to_string(lds_staged) +
                                       " bytes refused: " + hipGetErrorString(e));
    hipLaunchKernelGGL(kernel, dim3(b->force_work_n), dim3(kStagedBlock), lds_staged, st, a, b->nbl_dev,
                       b->num_nbrs_dev, b->max_nbrs, cells, params_dev, b->ntypes, lgp, work, nv);
    return check_launch(entry);
  };
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS: return launch(k_b_force_staged<RUMDX_CUTOFF_NS, FUSE>);
    case RUMDX_CUTOFF_SP: return launch(k_b_force_staged<RUMDX_CUTOFF_SP, FUSE>);
    case RUMDX_CUTOFF_SF: return launch(k_b_force_staged<RUMDX_CUTOFF_SF, FUSE>);
    default: return fail(RUMDX_E_INVALID, std::string(entry) + ": invalid cutoff method");
  }
}

}  // namespace
}  // namespace rumdx

extern "C" int rumdx_batch_forces(rumdx_batch* b, const float* params_dev, int cutoff_method, void* stream) {
  int rc = check_batch_forces(b, params_dev, "rumdx_batch_forces");
  if (rc != RUMDX_OK) return rc;
  const uint32_t lpp = b->lanes_per_particle;
  const uint32_t lgp = lpp >= 8u ? 3u : lpp >= 4u ? 2u : lpp >= 2u ? 1u : 0u;
  const BatchArgs a = args_of(b);
  const hipStream_t st = as_stream(stream);
  const uint32_t nblk = (b->nslots / kBlock) << lgp;
  const size_t lds = sizeof(float4) * 2 * b->ntypes * b->ntypes;
  const CellArrays cells{b->slot_cell_dev, b->cell_start_dev, b->cell_n_dev};
  if (b->force_work_dev != nullptr)
    return launch_staged<false>(b, params_dev, cutoff_method, BatchNvt{}, st, "rumdx_batch_forces");
  switch (cutoff_method) {
    case RUMDX_CUTOFF_NS:
      hipLaunchKernelGGL(k_b_force<RUMDX_CUTOFF_NS>, dim3(nblk), dim3(kBlock), lds, st, a, b->nbl_dev, b->num_nbrs_dev,
                         b->max_nbrs, cells, params_dev, b->ntypes, lgp);
      break;
    case RUMDX_CUTOFF_SP:
      hipLaunchKernelGGL(k_b_force<RUMDX_CUTOFF_SP>, dim3(nblk), dim3(kBlock), lds, st, a, b->nbl_dev, b->num_nbrs_dev,
                         b->max_nbrs, cells, params_dev, b->ntypes, lgp);
      break;
    case RUMDX_CUTOFF_SF:
      hipLaunchKernelGGL(k_b_force<RUMDX_CUTOFF_SF>, dim3(nblk), dim3(kBlock), lds, st, a, b->nbl_dev, b->num_nbrs_dev,
                         b->max_nbrs, cells, params_dev, b->ntypes, lgp);
      break;
    default: return fail(RUMDX_E_INVALID, "rumdx_batch_forces: invalid cutoff method");
  }
  return check_launch("rumdx_batch_forces");
}

extern "C" int rumdx_batch_nvt_step(rumdx_batch* b, float dt, void* workspace_dev, void* stream) {
  int rc = check_batch(b, "rumdx_batch_nvt_step");
  if (rc != RUMDX_OK) return rc;
  if (workspace_dev == nullptr || b->nh_dev == nullptr || b->last_r_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_batch_nvt_step: null workspace, thermostat state or last_r");
  const BatchArgs a = args_of(b);
  const hipStream_t st = as_stream(stream);
  double* partials = static_cast<double*>(workspace_dev);
  hipLaunchKernelGGL(k_b_nvt, dim3(b->nslots / kBlock), dim3(kBlock), 0, st, a, b->nh_dev,
                     reinterpret_cast<const float4*>(b->last_r_dev), dt, 0.25f * b->skin * b->skin, partials);
  hipLaunchKernelGGL(k_b_nvt_finish, dim3(b->nrep), dim3(kBlock), 0, st, b->rep_dev, partials, b->nh_dev, dt);
  return check_launch("rumdx_batch_nvt_step");
}

extern "C" int rumdx_batch_forces_nvt(rumdx_batch* b, const float* params_dev, int cutoff_method, float dt,
                                      void* workspace_dev, void* stream) {
  int rc = check_batch_forces(b, params_dev, "rumdx_batch_forces_nvt");
  if (rc != RUMDX_OK) return rc;
  if (b->force_work_dev == nullptr || b->force_work_whole == 0)
    return fail(RUMDX_E_INVALID, "rumdx_batch_forces_nvt: needs a staged work list of whole replicas");
  if (workspace_dev == nullptr || b->nh_dev == nullptr || b->last_r_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_batch_forces_nvt: null workspace, thermostat state or last_r");
  const hipStream_t st = as_stream(stream);
  double* partials = static_cast<double*>(workspace_dev);
  const BatchNvt nv{b->nh_dev, reinterpret_cast<const float4*>(b->last_r_dev), partials, dt,
                    0.25f * b->skin * b->skin};
  rc = launch_staged<true>(b, params_dev, cutoff_method, nv, st, "rumdx_batch_forces_nvt");
  if (rc != RUMDX_OK) return rc;
  hipLaunchKernelGGL(k_b_nvt_finish, dim3(b->nrep), dim3(kBlock), 0, st, b->rep_dev, partials, b->nh_dev, dt);
  return check_launch("rumdx_batch_forces_nvt");
}

extern "C" int rumdx_batch_energies(rumdx_batch* b, float dt, double* out_dev, void* workspace_dev, void* stream) {
  int rc = check_batch(b, "rumdx_batch_energies");
  if (rc != RUMDX_OK) return rc;
  if (workspace_dev == nullptr || out_dev == nullptr || b->nh_dev == nullptr)
    return fail(RUMDX_E_INVALID, "rumdx_batch_energies: null workspace, output or thermostat state");
  const BatchArgs a = args_of(b);
  const hipStream_t st = as_stream(stream);
  double* partials = static_cast<double*>(workspace_dev);
  hipLaunchKernelGGL(k_b_energy, dim3(b->nslots / kBlock), dim3(kBlock), 0, st, a, b->nh_dev, dt, partials);
  hipLaunchKernelGGL(k_b_energy_finish, dim3(b->nrep), dim3(kBlock), 0, st, b->rep_dev, partials, out_dev);
  return check_launch("rumdx_batch_energies");
}
