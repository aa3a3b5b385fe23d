// Device-side helpers shared by the rumdx kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "rumdx.h"

namespace rumdx {

// Threads per workgroup for the per-particle kernels: four wave64s.
constexpr uint32_t kBlock = 256;
constexpr uint32_t kMaxWorkgroup = 1024;  // largest workgroup of any kernel here
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.
constexpr uint32_t kNumXcd = 8;

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);

// ------------------------------------------------------------------ launch geometry
inline uint32_t blocks_for(uint32_t n, uint32_t per = kBlock) { return (n + per - 1) / per; }

// Maps the hardware workgroup id to a logical block so that each XCD receives
// one contiguous range of logical blocks (hence of cell-sorted particles): the
// neighbour positions a block gathers then sit in that XCD's own L2.  Bijective
// on [0, nb) for any nb.
__device__ __forceinline__ uint32_t xcd_block(uint32_t hw, uint32_t nb) {
  const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
  const uint32_t x = hw % kNumXcd, k = hw / kNumXcd;
  const uint32_t base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + k;
}

// ------------------------------------------------------------------ box
// RUMD's float[8] box, expanded on the host (loadMemory semantics,
// include/rumd/SimulationBox.h:165-178): a cubic box uses L[0] and 1/L[0]
// for all three axes.
struct Box {
  float L[3];
  float iL[3];
  double Ld[3];
  double iLd[3];
};

inline Box make_box(const float box[8]) {
  Box b;
  const bool cubic = static_cast<unsigned>(box[0]) != 0u;
  for (int a = 0; a < 3; ++a) {
    b.L[a] = cubic ? box[1] : box[1 + a];
    b.iL[a] = cubic ? box[4] : box[4 + a];
    b.Ld[a] = static_cast<double>(b.L[a]);
    b.iLd[a] = 1.0 / b.Ld[a];
  }
  return b;
}

// ------------------------------------------------------------------ exclusions
// A slot's excluded partners: [e0, e1) of the exclusion list, owner's id.
struct ExclusionRange {
  uint32_t e0 = 0, e1 = 0;
};

// Exclusion lists of a neighbour list (CSR over original ids, see
// rumdx_nblist.excl_*); start == nullptr means none.  Slots map to original
// ids through idx.
struct Exclusions {
  const uint32_t* idx;
  const uint32_t* start;
  const uint32_t* ids;

  __device__ __forceinline__ ExclusionRange of(uint32_t slot) const {
    ExclusionRange r;
    if (start != nullptr) {
      const uint32_t o = idx[slot];
      r.e0 = start[o];
      r.e1 = start[o + 1];
    }
    return r;
  }
  // Whether the particle in slot j is excluded for the owner of er.
  __device__ __forceinline__ bool excludes(const ExclusionRange& er, uint32_t j) const {
    if (er.e0 == er.e1) return false;
    const uint32_t oj = idx[j];
    for (uint32_t q = er.e0; q < er.e1; ++q)
      if (ids[q] == oj) return true;
    return false;
  }
};

// FP32 squared distance without the minimum-image correction: identical to
// min_image_d2_f32 when the pair cannot wrap (|dx| < L/2 so rint gives 0),
// which holds for every pair of a tile whose halo does not wrap and for the
// stencil of a cell that does not wrap (stencil_wraps).
__device__ __forceinline__ float plain_d2_f32(float4 a, float4 b) {
#pragma clang fp contract(off)
  const float dx = a.x - b.x;
  const float dy = a.y - b.y;
  const float dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

// Whether the 5-cell stencil of cell c wraps around the box on an axis of nc
// cells.  When it does not and nc >= 7, every candidate lies within 3 cells
// (< L/2) of the cell, so RUMD's minimum-image correction rounds to zero and
// the plain FP32 difference is the same number.
__device__ __forceinline__ bool stencil_wraps(int c, int nc) { return nc < 7 || c < 2 || c + 2 >= nc; }

// RUMD's FP32 minimum-image squared distance (RectangularSimulationBox::
// calculateDistance, include/rumd/SimulationBox.h:111-125), evaluated in source
// order without FMA contraction so that it is reproducible bit for bit on the
// host.  Used by every neighbour-list criterion and by the skin check.
__device__ __forceinline__ float min_image_d2_f32(float4 a, float4 b, const Box& bx) {
#pragma clang fp contract(off)
  float dx = a.x - b.x;
  float dy = a.y - b.y;
  float dz = a.z - b.z;
  dx -= bx.L[0] * rintf(dx * bx.iL[0]);
  dy -= bx.L[1] * rintf(dy * bx.iL[1]);
  dz -= bx.L[2] * rintf(dz * bx.iL[2]);
  return dx * dx + dy * dy + dz * dz;
}

// Cell coordinate of one axis (calculateCellCoordinates,
// src/NeighborList.cu:1106-1133), uncontracted; out-of-range values clamp.
__device__ __forceinline__ int cell_coord(float x, float iL, int nc) {
#pragma clang fp contract(off)
  const float s = x * iL + 0.5f;
  int c = static_cast<int>(static_cast<float>(nc) * s);
  c = c >= nc ? nc - 1 : c;
  c = c < 0 ? 0 : c;
  return c;
}

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Fixed-order sum over a kBlock-thread workgroup: lane tree inside each wave,
// then the four wave totals in wave order.  Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4) {
  v = wave_sum(v);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (uint32_t k = 0; k < kBlock / 64; ++k) s += lds4[k];
  }
  __syncthreads();
  return s;
}

// Maximum of v over the workgroup (lds4: blockDim.x / 64 entries), valid in thread 0.  Flags
// that collect a maximum (the longest list) take one atomic per workgroup:
// one per particle on a single address serialises at the memory side.
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t* lds4) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t y = __shfl_xor(v, o, 64);
    v = y > v ? y : v;
  }
  if ((threadIdx.x & 63u) == 0) lds4[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t m = 0;
  if (threadIdx.x == 0)
    for (uint32_t k = 0; k < blockDim.x / 64; ++k) m = lds4[k] > m ? lds4[k] : m;
  __syncthreads();
  return m;
}

// Returns true (in every thread) for the block that arrives last at `counter`,
// which is left at zero for the next launch.  No memory fence is issued: this
// orders nothing but the arrivals themselves, which is enough for a block to
// clear a flag every other block has already read.  (An agent-scope fence on
// gfx950 writes back the whole L2 of the XCD, so data handed between blocks
// goes through a separate finishing kernel instead.)
__device__ __forceinline__ bool last_to_arrive(uint32_t* counter, uint32_t nblocks) {
  __shared__ uint32_t s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == nblocks - 1) ? 1u : 0u;
    if (s_last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return s_last != 0u;
}

// Fixed-order sum of partials[0..np) by one workgroup; valid in thread 0.
// Each thread adds its partials k, k + kBlock, ... in that order; the loads of
// eight of them are issued together, so one workgroup does not wait for one
// round trip per partial.
__device__ __forceinline__ double sum_partials(const double* partials, uint32_t np, uint32_t stride,
                                               double* lds4) {
  double v = 0.0;
  uint32_t k = threadIdx.x;
  for (; k + 7u * kBlock < np; k += 8u * kBlock) {
    double a[8];
#pragma unroll
    for (uint32_t u = 0; u < 8u; ++u) a[u] = partials[static_cast<size_t>(k + u * kBlock) * stride];
#pragma unroll
    for (uint32_t u = 0; u < 8u; ++u) v += a[u];
  }
  for (; k < np; k += kBlock) v += partials[static_cast<size_t>(k) * stride];
  return block_sum(v, lds4);
}

// ------------------------------------------------------------------ workspace
// Layout of the caller's reduction workspace.
struct Workspace {
  static constexpr size_t kCounters = 0;    // uint32[16]
  static constexpr size_t kScalars = 64;    // double[8] scratch results
  static constexpr size_t kPartials = 256;  // double[4 * nblocks]
  static uint32_t* counters(void* ws) { return reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + kCounters); }
  static double* scalars(void* ws) { return reinterpret_cast<double*>(static_cast<char*>(ws) + kScalars); }
  static double* partials(void* ws) { return reinterpret_cast<double*>(static_cast<char*>(ws) + kPartials); }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Inner lists (sorted method): the positions of the last prune when the list
// has inner lists (else null), and the squared displacement since that prune
// above which the skin check asks for a new one: (0.49 inner_skin)^2, a 2 %
// margin under the inner_skin/2 that keeps every pair within rc listed.
inline const float* inner_prune_r(const rumdx_nblist* nb) {
  return (nb->nbl16_in_dev != nullptr && nb->num_in_dev != nullptr) ? nb->prune_r_dev : nullptr;
}
inline float inner_check2(const rumdx_nblist* nb) {
  const double h = 0.49 * static_cast<double>(nb->inner_skin);
  return static_cast<float>(h * h);
}

}  // namespace rumdx
