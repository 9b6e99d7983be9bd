// pbf_kernels.hip -- gfx950 kernels of the PBF step and their launchers.
//
// One step (the reference's Particles::timeStep, src/particles.cpp:250-297):
//   predict        gravity, x* = x, clamp_response, cell hash, histogram
//   scan           wave64 exclusive scan of the cell histogram
//   scatter/rank   counting sort by cell, deterministic in-cell rank by id,
//                  gather into sorted order
//   nbr_lambda     27-cell search -> chunked ELL neighbour lists (fixed per
//                  step), then the first density / lambda pass in the same
//                  thread
//   lambda         density / lambda pass            (S - 1 more)
//   deltap         delta-p + clamp pass              (x S; the last one
//                  writes the (x*, v) records)
//   xsph           XSPH viscosity, vorticity, density
//   vort_commit    vorticity confinement, x <- x*; between the steps of one
//                  call fused with the next step's predict
// Particles stay in the order (cell, id), so the lists of a particle are in
// ascending sorted index, i.e. the same order a CPU visits the 27 cells in.
// The list passes come with one lane, one quad (4 lanes) or one wave per
// particle, chosen per pass by state size (PassRule in pbf_internal.h); each
// adds its terms in list order from +0, so all three give the same bits.
//
// Every pass is a per-particle device function.  The flat schedule runs one
// launch per pass over all particles of all scenes; the persistent schedule
// runs whole scenes inside one workgroup (all passes, all steps, workgroup
// barriers between passes) in a single launch.  Both call the same
// functions, so both give the same bits.
#include <hip/hip_runtime.h>

#include "pbf_internal.h"

namespace pbf {

namespace {

constexpr int kBlock = 256;
// Neighbour loads in flight per group (build-time tunables).
#ifndef PBF_UNROLL
#define PBF_UNROLL 4
#endif
constexpr int kUnroll = PBF_UNROLL;

// Device pointers and sizes of one context, passed by value to every kernel.
struct Dev {
  float4 *posA, *posB, *velA, *velB, *predP, *predQ, *vort, *vxsph, *xv;
  // x* as 12-byte records, written by the delta-p passes before the last:
  // the λ passes after the first gather x* only, and 10.7 records share a
  // 128-byte line instead of 8.
  float3* xs3;
  uint32_t *cell, *arrival, *cellSorted, *tmpId, *cellCount, *cellStart, *nbr, *ncnt, *flags;
  const SceneGrid* grids;
  const uint32_t* slotScene;
  MeshView mesh;  // triangle-mesh collider (ntri 0: the analytic box walls)
  // Dense split (kDenseList in pbf_internal.h): the queue (NULL: off), its
  // threshold, and the per-lane workgroups of a split launch (0: not split).
  uint32_t* dense;
  uint32_t dense_list;
  uint32_t nlane;
  // Build split: the queue of particles in crowded cells (NULL: off) and its
  // threshold on the cell's population.
  uint32_t* dense_b;
  uint32_t dense_cell;
  // Nonzero: the prediction stores the advanced v in velA and the sort moves
  // it to velB (read only by debug stops and the adjoint).
  uint32_t move_v;
  uint32_t n;
  int K4;
};

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Whether x queues its dense particles: a single scene of at most
// kDenseMaxN particles whose list build runs one lane per particle.
static bool dense_split(const Ctx& x) {
  const PassRule& b = x.rules[kPassBuild];
  const bool lane_build = x.n > b.wave_max && !(x.n <= b.quad_max && x.n < (int64_t{1} << 30));
  return x.nscenes == 1 && x.d.dense && x.dense_list > 0 && x.n <= kDenseMaxN && lane_build;
}

Dev dev_of(const Ctx& x) {
  const DevState& s = x.d;
  Dev d;
  d.posA = s.posA; d.posB = s.posB; d.velA = s.velA; d.velB = s.velB;
  d.predP = s.predP; d.predQ = s.predQ; d.vort = s.vort; d.vxsph = s.vxsph; d.xv = s.xv;
  d.xs3 = reinterpret_cast<float3*>(s.xs3);
  d.cell = s.cell; d.arrival = s.arrival; d.cellSorted = s.cellSorted; d.tmpId = s.tmpId;
  d.cellCount = s.cellCount; d.cellStart = s.cellStart; d.nbr = s.nbr; d.ncnt = s.ncnt;
  d.flags = s.flags; d.grids = s.grids; d.slotScene = s.slotScene;
  d.mesh.tri = s.meshTri;
  d.mesh.node = s.meshNode;
  d.mesh.ntri = x.mesh_ntri;
  d.mesh.nnode = x.mesh_nnode;
  d.mesh.skin = x.mesh_skin;
  d.mesh.flag = s.meshFlag;
  d.mesh.margin = x.mesh_margin;
  d.dense = dense_split(x) ? s.dense : nullptr;
  d.dense_list = x.dense_list;
  d.dense_b = dense_split(x) && x.dense_cell > 0 ? s.dense_b : nullptr;
  d.dense_cell = x.dense_cell;
  d.move_v = x.sorted_v ? 1u : 0u;
  d.nlane = 0;
  d.n = (uint32_t)x.n;
  d.K4 = k4_of(x);
  return d;
}

__device__ __forceinline__ uint32_t id_of(const float4& p) { return __float_as_uint(p.w); }

// ---------------------------------------------------------------- scans --
// Inclusive scan across the 64 lanes of a wave with DPP row shifts and row
// broadcasts (rows of 16 lanes, then rows 1/3 from row 0/2, then rows 2/3
// from lane 31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)x;
}

constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

// Block-local exclusive prefix of a 1024-element tile; tile totals to
// partials.  With clear, the counts are set to zero once read.
__global__ __launch_bounds__(kBlock) void k_scan_block(uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, uint32_t n,
                                                       uint32_t* __restrict__ partials, bool clear) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  if (base + kScanItems <= n) {
    uint4 q = *reinterpret_cast<const uint4*>(in + base);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    if (clear) *reinterpret_cast<uint4*>(in + base) = make_uint4(0u, 0u, 0u, 0u);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++) v[k] = (base + k < n) ? in[base + k] : 0u;
    if (clear)
      for (int k = 0; k < kScanItems; k++)
        if (base + k < n) in[base + k] = 0u;
  }
  uint32_t t = (v[0] + v[1]) + (v[2] + v[3]);
  uint32_t incl = wave_incl_scan(t);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t woff = 0;
  for (int w = 0; w < wid; w++) woff += wsum[w];
  uint32_t e = woff + incl - t;
  uint32_t o[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    o[k] = e;
    e += v[k];
  }
  if (base + kScanItems <= n) {
    *reinterpret_cast<uint4*>(out + base) = make_uint4(o[0], o[1], o[2], o[3]);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
      if (base + k < n) out[base + k] = o[k];
  }
  if (threadIdx.x == kBlock - 1) partials[blockIdx.x] = woff + incl;
}

// Exclusive scan of the tile totals in place (one workgroup); total to out[n].
__global__ __launch_bounds__(1024) void k_scan_partials(uint32_t* __restrict__ partials,
                                                        uint32_t nb, uint32_t* __restrict__ out,
                                                        uint32_t n) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (uint32_t b = 0; b < nb; b += 1024) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < nb ? partials[i] : 0u;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wid; w++) woff += wsum[w];
    uint32_t c0 = carry;
    if (i < nb) partials[i] = c0 + woff + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c0 + woff + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* __restrict__ out, uint32_t n,
                                                     const uint32_t* __restrict__ partials) {
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e < n) out[e] += partials[e / kScanTile];
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, size_t n,
                                                     uint32_t v) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

// Agent-scope loads and stores of the cell histogram inside the persistent
// kernel: the histogram is built with device-scope atomics, so it is read and
// cleared with atomic accesses rather than through the per-CU cache.
__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier between two passes of the persistent kernel.  A scene belongs
// to one workgroup for all its steps and no other workgroup reads its
// arrays, so a workgroup barrier (which makes the workgroup's global writes
// visible to all its threads; its waves share one CU) is enough.  The
// histogram, which the prediction fills with atomics performed in L2, is
// zeroed and read with agent-scope accesses (store_agent, load_agent).
__device__ __forceinline__ void phase_barrier() { __syncthreads(); }

// Exclusive scan of in[0..n) into out[0..n) inside one workgroup of BLOCK
// threads, every value offset by base; out[n] = base + total.  All threads
// of the block must call it.
template <int BLOCK>
__device__ void block_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                           uint32_t base, uint32_t* wsum /* BLOCK/64 + 1 shared */) {
  constexpr int W = BLOCK / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t carry = base;
  for (uint32_t b = 0; b < n; b += BLOCK * kScanItems) {
    const uint32_t e0 = b + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) v[k] = (e0 + k < n) ? load_agent(in + e0 + k) : 0u;
    const uint32_t t = (v[0] + v[1]) + (v[2] + v[3]);
    const uint32_t incl = wave_incl_scan(t);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0, total = 0;
    for (int w = 0; w < W; w++) {
      if (w < wid) woff += wsum[w];
      total += wsum[w];
    }
    uint32_t e = carry + woff + incl - t;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      if (e0 + k < n) out[e0 + k] = e;
      e += v[k];
    }
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

// ----------------------------------------------------- per-particle passes --
__device__ __forceinline__ uint32_t cell_lin(const SceneGrid& g, float3 p, int* cx = nullptr) {
  int ix = axis_cell(p.x, g.lo[0], g.inv_cell, g.dims[0]);
  int iy = axis_cell(p.y, g.lo[1], g.inv_cell, g.dims[1]);
  int iz = axis_cell(p.z, g.lo[2], g.inv_cell, g.dims[2]);
  if (cx) {
    cx[0] = ix;
    cx[1] = iy;
    cx[2] = iz;
  }
  return g.cell_base + ((uint32_t)iz * (uint32_t)g.dims[1] + (uint32_t)iy) * (uint32_t)g.dims[0] +
         (uint32_t)ix;
}

// applyForceVelocity (src/particles.cpp:175-183) + cell hash + histogram.
// With MOVE=false x* = x (used by the load-time density pass).
// Prediction and cell of particle i from its committed x and v; returns
// the cell.
template <bool MOVE>
__device__ __forceinline__ uint32_t predict_from(uint32_t i, const Dev& d, const SimConst& c,
                                                 const SceneGrid& g, float4 x, float4 v,
                                                 float4* __restrict__ pred) {
  float3 p = make_float3(x.x, x.y, x.z);
  if (MOVE) {
    v.y = v.y - c.gdt;
    p = clamp_response(g, c, d.mesh, p, make_float3(v.x, v.y, v.z));
    if (d.move_v) d.velA[i] = v;
  }
  pred[i] = make_float4(p.x, p.y, p.z, 0.0f);
  const uint32_t cl = cell_lin(g, p);
  d.cell[i] = cl;
  return cl;
}

template <bool MOVE>
__device__ __forceinline__ uint32_t predict_cell(uint32_t i, const Dev& d, const SimConst& c,
                                                 const SceneGrid& g) {
  return predict_from<MOVE>(i, d, c, g, d.posA[i], MOVE ? d.velA[i] : make_float4(0, 0, 0, 0),
                           d.predP);
}

template <bool MOVE>
__device__ __forceinline__ void predict_one(uint32_t i, const Dev& d, const SimConst& c,
                                            const SceneGrid& g) {
  const uint32_t cl = predict_cell<MOVE>(i, d, c, g);
  d.arrival[i] = atomicAdd(&d.cellCount[cl], 1u);
}

// Histogram slot of each lane's cell with one atomic per run of equal cells
// in consecutive lanes (the state is sorted by the previous step's cells,
// so a wave holds a few runs of ~8).  Every lane of the wave must call it;
// lanes with valid = false take no slot.  The slots of one cell are still
// 0..count-1; their order within the cell is irrelevant (the sort ranks by
// id).
__device__ __forceinline__ uint32_t wave_run_slot(uint32_t cl, bool valid,
                                                  uint32_t* __restrict__ count) {
  const uint32_t lane = __lane_id();
  const uint32_t prev = __shfl_up(cl, 1);
  const bool head = lane == 0 || prev != cl;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all for lane 63)
  const uint32_t start = 63u - (uint32_t)__clzll((long long)(heads & upto));
  const unsigned long long after = heads & ~upto;
  const uint32_t end = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
  uint32_t base = 0;
  if (head && valid) base = atomicAdd(count + cl, end - start);
  base = __shfl(base, (int)start);
  return base + (lane - start);
}

__device__ __forceinline__ void scatter_one(uint32_t i, const Dev& d) {
  const uint32_t slot = d.cellStart[d.cell[i]] + d.arrival[i];
  if (slot < d.n) d.tmpId[slot] = id_of(d.posA[i]);
}

// Rank of a particle inside its cell = number of cell members with a smaller
// id; the gather target is cellStart + rank, which makes the order (cell, id)
// independent of atomic arrival order.
// Moves x, x* and (with d.move_v) v to the sorted slot.
// Everything the particle moves is loaded before the count, and the cell's
// ids are read four at a time with independent loads, so the loads of one
// particle overlap instead of waiting on each other.
__device__ __forceinline__ void rank_one(uint32_t i, const Dev& d,
                                         const float4* __restrict__ pred) {
  const float4 x = d.posA[i];
  const float4 v = d.move_v ? d.velA[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 q = pred[i];
  const uint32_t c = d.cell[i];
  const uint32_t s = d.cellStart[c];
  const uint32_t e = min(d.cellStart[c + 1], d.n);
  const uint32_t id = id_of(x);
  const uint32_t* __restrict__ ids = d.tmpId;
  uint32_t rank = 0;
  const uint32_t last = e > s ? e - 1 : s;
  for (uint32_t t = s; t < e; t += 4) {
    // Slots past the end re-read the last member and do not count.
    const uint32_t a0 = ids[t];
    const uint32_t a1 = ids[min(t + 1, last)];
    const uint32_t a2 = ids[min(t + 2, last)];
    const uint32_t a3 = ids[min(t + 3, last)];
    rank += (uint32_t)(a0 < id) + (uint32_t)(t + 1 < e && a1 < id) +
            (uint32_t)(t + 2 < e && a2 < id) + (uint32_t)(t + 3 < e && a3 < id);
  }
  const uint32_t dst = s + rank;
  if (dst >= d.n) return;
  d.posB[dst] = x;
  if (d.move_v) d.velB[dst] = v;
  d.predQ[dst] = q;
  d.cellSorted[dst] = c;
  // Build split: a particle of a crowded cell gets its list from a quad.
  if (d.dense_b && e - s > d.dense_cell) {
    const uint32_t slot = atomicAdd(&d.flags[kDenseBuildSlot], 1u);
    if (slot < d.n) d.dense_b[slot] = dst;  // each particle once per step: slot < n
  }
}

// Distance from coordinate x to the near face of cell k+side of the axis
// (side -1: lower face of cell k, +1: upper face), never negative.
__device__ __forceinline__ float face_gap(float x, float lo, float cell, int k, int side) {
  if (side == 0) return 0.0f;
  const float f = lo + (float)(side < 0 ? k : k + 1) * cell;
  // The face is within a few ulps of the coordinates involved of where the
  // cell assignment puts it; tol bounds that generously.
  const float tol = (fabsf(x) + fabsf(lo) + fabsf(f)) * 0x1p-20f;
  const float gap = (side < 0 ? x - f : f - x) - tol;
  return gap > 0.0f ? gap : 0.0f;
}

// Neighbour sets of src/particles.cpp:258-265 (|x*_i - x*_j|^2 <= H2, j != i),
// found in the 27 cells around i.  For each (dz, dy) the cells dx = -1..1 are
// contiguous in sorted order, so the search walks at most 9 index ranges in
// ascending order.  Rows and end cells whose box lies farther than 1.01 h
// from the particle are skipped (they hold no neighbour; the 1% margin covers
// the rounding of the cell faces), which leaves the list and its order
// unchanged.  Candidates are tested 32 at a time, branch-free, into a hit
// mask (self cleared afterwards); the hits are then appended in order.
__device__ __forceinline__ void nbr_one(uint32_t i, const Dev& d, float h2, const SceneGrid& g) {
  const float4 xi = d.predQ[i];
  int cc[3];
  cell_lin(g, make_float3(xi.x, xi.y, xi.z), cc);
  const float cell = 1.0f / g.inv_cell;
  const float hm2 = h2 * 1.0201f;  // (1.01 h)^2
  const float4* __restrict__ pred = d.predQ;
  const uint32_t cap = (uint32_t)d.K4 * kChunk;
  uint32_t cnt = 0;
  // A 4-entry shift register: the latest hit enters at b3, and after every
  // fourth hit b0..b3 hold the chunk in order.
  uint32_t b0 = 0u, b1 = 0u, b2 = 0u, b3 = 0u;
  auto push = [&](uint32_t j) {
    b0 = b1;
    b1 = b2;
    b2 = b3;
    b3 = j;
    cnt++;
    if ((cnt & 3u) == 0u) {
      const int chunk = (int)(cnt >> 2) - 1;
      if (chunk < d.K4)
        *reinterpret_cast<uint4*>(d.nbr + nbr_chunk(i, chunk, d.K4)) = make_uint4(b0, b1, b2, b3);
    }
  };
  auto hit = [&](const float4& xj) -> uint32_t {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float d2 = rx * rx + ry * ry + rz * rz;
    return d2 <= h2 ? 1u : 0u;
  };
  const float gxl = face_gap(xi.x, g.lo[0], cell, cc[0], -1);
  const float gxr = face_gap(xi.x, g.lo[0], cell, cc[0], 1);
  for (int dz = -1; dz <= 1; dz++) {
    const int z = cc[2] + dz;
    if (z < 0 || z >= g.dims[2]) continue;
    const float gz = face_gap(xi.z, g.lo[2], cell, cc[2], dz);
    for (int dy = -1; dy <= 1; dy++) {
      const int y = cc[1] + dy;
      if (y < 0 || y >= g.dims[1]) continue;
      const float gy = face_gap(xi.y, g.lo[1], cell, cc[1], dy);
      const float yz2 = gy * gy + gz * gz;
      if (yz2 > hm2) continue;
      const int x0 = (cc[0] > 0 && gxl * gxl + yz2 <= hm2) ? cc[0] - 1 : cc[0];
      const int x1 = (cc[0] + 1 < g.dims[0] && gxr * gxr + yz2 <= hm2) ? cc[0] + 1 : cc[0];
      const uint32_t row =
          g.cell_base + ((uint32_t)z * (uint32_t)g.dims[1] + (uint32_t)y) * (uint32_t)g.dims[0];
      const uint32_t s = d.cellStart[row + (uint32_t)x0];
      const uint32_t e = min(d.cellStart[row + (uint32_t)x1 + 1u], d.n);
      for (uint32_t jb = s; jb < e; jb += 32) {
        const uint32_t len = min(e - jb, 32u);
        uint32_t mask = 0;
        // Groups of 4 may read up to 3 candidates past the range (the x*
        // arrays are padded by 64 entries); their bits are masked off.
        for (uint32_t o = 0; o < len; o += 4) {
          const uint32_t j = jb + o;
          const float4 a = at(pred, j), b = at(pred, j + 1u), c4 = at(pred, j + 2u),
                       q = at(pred, j + 3u);
          mask |= (hit(a) | (hit(b) << 1) | (hit(c4) << 2) | (hit(q) << 3)) << o;
        }
        if (len < 32u) mask &= (1u << len) - 1u;
        const uint32_t self = i - jb;
        if (self < 32u) mask &= ~(1u << self);
        while (mask) {
          const uint32_t u = (uint32_t)__builtin_ctz(mask);
          mask &= mask - 1u;
          push(jb + u);
        }
      }
    }
  }
  // Partial last chunk: its r = cnt % 4 entries are b(4-r)..b3.  The slots
  // past them are filled with a repeat of the first, so every slot of the
  // chunks in use holds a particle index.
  const uint32_t r = cnt & 3u;
  if (r) {
    const int chunk = (int)(cnt >> 2);
    const uint32_t e0 = r == 1u ? b3 : (r == 2u ? b2 : b1);
    const uint32_t e1 = r == 1u ? b3 : (r == 2u ? b3 : b2);
    const uint32_t e2 = r == 3u ? b3 : e0;
    if (chunk < d.K4)
      *reinterpret_cast<uint4*>(d.nbr + nbr_chunk(i, chunk, d.K4)) = make_uint4(e0, e1, e2, e0);
  }
  if (cnt > cap) {
    atomicMax(&d.flags[0], cnt);
    cnt = cap;
  }
  d.ncnt[i] = cnt;
}

// ---- per-neighbour terms of the list passes ----------------------------------
// Each pass adds, for every neighbour j in list order, a few terms to sums
// that start at +0.  The terms are written once here and used by both the
// per-lane loops and the wide (wave-per-particle) passes, so both add the
// same values in the same order.

// newtonStepCalculateLambda (src/particles.cpp:189-197): W, g/rho0 (3),
// |g/rho0|^2.
__device__ __forceinline__ void lambda_terms(const SimConst& c, const float3& xi,
                                             const float3& xj, float* t) {
  const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float w = poly6_sel(c, r2);
  float3 gr = spiky_sel(c, rx, ry, rz, r2);
  gr.x = gr.x * c.inv_rho0;
  gr.y = gr.y * c.inv_rho0;
  gr.z = gr.z * c.inv_rho0;
  t[0] = w;
  t[1] = gr.x;
  t[2] = gr.y;
  t[3] = gr.z;
  t[4] = gr.x * gr.x + gr.y * gr.y + gr.z * gr.z;
}

// newtonStepUpdatePosition (src/particles.cpp:208-210) with s_corr (:195):
// (lambda_i + lambda_j + s_ij) g_ij (3).  x .w carries lambda.
template <int NT>
__device__ __forceinline__ void deltap_terms(const SimConst& c, const float4& xi,
                                             const float4& xj, float* t) {
  const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float w = poly6_sel(c, r2);
  const float s = c.negk * ipow_n<NT>(w * c.S, c.n_tensile);
  const float3 gr = spiky_sel(c, rx, ry, rz, r2);
  const float sc = (xi.w + xj.w) + s;
  t[0] = sc * gr.x;
  t[1] = sc * gr.y;
  t[2] = sc * gr.z;
}

// calculateVorticityApplyXSPHViscosity (src/particles.cpp:221-231): the
// vorticity (v_j - v_i) x g (3), the viscosity (v_j - v_i) W (3), W.
__device__ __forceinline__ void xsph_terms(const SimConst& c, const float3& xi, const float3& vi,
                                           const float3& xj, const float3& vj, float* t) {
  const float u0 = vj.x - vi.x, u1 = vj.y - vi.y, u2 = vj.z - vi.z;
  const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float3 gr = spiky_sel(c, rx, ry, rz, r2);
  t[0] = u1 * gr.z - u2 * gr.y;
  t[1] = u2 * gr.x - u0 * gr.z;
  t[2] = u0 * gr.y - u1 * gr.x;
  const float w = poly6_sel(c, r2);
  t[3] = u0 * w;
  t[4] = u1 * w;
  t[5] = u2 * w;
  t[6] = w;
}

// applyVorticity (src/particles.cpp:238-240): |omega_j| g_ij (3); x .w
// carries |omega|.
__device__ __forceinline__ void vort_terms(const SimConst& c, const float4& xi, const float4& xj,
                                           float* t) {
  const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float3 gr = spiky_sel(c, rx, ry, rz, r2);
  t[0] = xj.w * gr.x;
  t[1] = xj.w * gr.y;
  t[2] = xj.w * gr.z;
}

// lambda from the sums of newtonStepCalculateLambda (src/particles.cpp:
// 198-204): C = rho/rho0 - 1, D = sum |g/rho0|^2 + |G|^2, lambda = -C/(D+eps).
__device__ __forceinline__ float lambda_of(const SimConst& c, float dens, float denom, float Gx,
                                           float Gy, float Gz) {
  const float ci = dens / c.rho0 - 1.0f;
  denom = denom + (Gx * Gx + Gy * Gy + Gz * Gz);
  return -ci / (denom + c.eps);
}

// newtonStepCalculateLambda (src/particles.cpp:185-204).  out = (x*, lambda).
// x* comes from `in` (after the sort), or with PACKED from the 12-byte
// records that the previous delta-p pass wrote (then `in` is not read).
template <bool PACKED>
__device__ __forceinline__ void lambda_one(uint32_t i, const Dev& d, const SimConst& c,
                                           const float4* __restrict__ in,
                                           float4* __restrict__ out) {
  const float3 xi = PACKED ? d.xs3[i] : make_float3(in[i].x, in[i].y, in[i].z);
  const int cnt = (int)d.ncnt[i];
  float dens = 0.0f, denom = 0.0f, Gx = 0.0f, Gy = 0.0f, Gz = 0.0f;
  const float3* __restrict__ xs3 = d.xs3;
  visit_neighbors<kUnroll>(
      i, cnt, d.K4, d.nbr,
      [&](uint32_t j) {
        if (PACKED) return at(xs3, j);
        const float4 q = at(in, j);
        return make_float3(q.x, q.y, q.z);
      },
      [&](uint32_t, const float3& xj) {
        float t[5];
        lambda_terms(c, xi, xj, t);
        dens = dens + t[0];
        Gx = Gx + t[1];
        Gy = Gy + t[2];
        Gz = Gz + t[3];
        denom = denom + t[4];
      });
  out[i] = make_float4(xi.x, xi.y, xi.z, lambda_of(c, dens, denom, Gx, Gy, Gz));
}

// (x*, v) record of sorted particle i after the solve: 6 consecutive floats,
// so XSPH gathers both with two adjacent 12-byte loads.
__device__ __forceinline__ float3* xv_record(const Dev& d, uint32_t i) {
  return reinterpret_cast<float3*>(d.xv) + 2 * (size_t)i;
}

template <bool LAST>
__device__ __forceinline__ void deltap_finish(uint32_t i, const Dev& d, const SimConst& c,
                                              const SceneGrid& g, const float4& xi, float dx,
                                              float dy, float dz);

// newtonStepUpdatePosition (src/particles.cpp:206-213) with s_corr (:195)
// and clamp (:51-84).  in = (x*, lambda), out = (x*', 0).  The last
// iteration also forms v = (x* - x) / dt (updateVelocity, :215-217).  NT >= 0
// is the tensile exponent N known at compile time (the scenes' 4): the
// same products, unrolled; NT < 0 reads it from c.
template <bool LAST, int NT = -1>
__device__ __forceinline__ void deltap_one(uint32_t i, const Dev& d, const SimConst& c,
                                           const SceneGrid& g, const float4* __restrict__ in,
                                           float4* __restrict__ out) {
  const float4 xi = in[i];
  const int cnt = (int)d.ncnt[i];
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  visit_neighbors<kUnroll>(
      i, cnt, d.K4, d.nbr, [&](uint32_t j) { return at(in, j); },
      [&](uint32_t, const float4& xj) {
        float t[3];
        deltap_terms<NT>(c, xi, xj, t);
        dx = dx + t[0];
        dy = dy + t[1];
        dz = dz + t[2];
      });
  deltap_finish<LAST>(i, d, c, g, xi, dx, dy, dz);
}

// The end of newtonStepUpdatePosition from the summed correction: scale by
// 1/rho0, clamp (src/particles.cpp:51-84), and store x* (and, in the last
// iteration, v = (x* - x)/dt).
template <bool LAST>
__device__ __forceinline__ void deltap_finish(uint32_t i, const Dev& d, const SimConst& c,
                                              const SceneGrid& g, const float4& xi, float dx,
                                              float dy, float dz) {
  dx = dx * c.inv_rho0;
  dy = dy * c.inv_rho0;
  dz = dz * c.inv_rho0;
  const float3 p =
      clamp_dp(g, c, d.mesh, make_float3(xi.x, xi.y, xi.z), make_float3(dx, dy, dz));
  if (!LAST) {
    // The next λ pass reads x* from the 12-byte records only (this pass
    // gathers from its input x*).
    d.xs3[i] = p;
  } else {
    // The last iteration writes only the (x*, v) record that XSPH gathers.
    const float4 x = d.posB[i];
    const float3 v = make_float3(c.inv_dt * (p.x - x.x), c.inv_dt * (p.y - x.y),
                                 c.inv_dt * (p.z - x.z));
    float3* r = xv_record(d, i);
    r[0] = make_float3(p.x, p.y, p.z);
    r[1] = v;
  }
}

// The stores of calculateVorticityApplyXSPHViscosity from its sums:
// out = (x*, |omega|), vxsph = (v + C*visc, density), vort = (omega, 0).
__device__ __forceinline__ void xsph_finish(uint32_t i, const Dev& d, const SimConst& c,
                                            const float3& xi, const float3& vi, float ox,
                                            float oy, float oz, float sx, float sy, float sz,
                                            float dens, float4* __restrict__ out) {
  const float wn = sqrtf(ox * ox + oy * oy + oz * oz);
  out[i] = make_float4(xi.x, xi.y, xi.z, wn);
  d.vxsph[i] = make_float4(vi.x + c.xc * sx, vi.y + c.xc * sy, vi.z + c.xc * sz, dens);
  d.vort[i] = make_float4(ox, oy, oz, 0.0f);
}

// calculateVorticityApplyXSPHViscosity (src/particles.cpp:219-234), Jacobi:
// every neighbour velocity is the pre-XSPH v.  out = (x*, |omega|),
// vxsph = (v + C*visc, density), vort = (omega, 0).
__device__ __forceinline__ void xsph_one(uint32_t i, const Dev& d, const SimConst& c,
                                         float4* __restrict__ out) {
  const float3* __restrict__ xv = reinterpret_cast<const float3*>(d.xv);
  const float3 xi = xv[2 * (size_t)i], vi = xv[2 * (size_t)i + 1];
  const int cnt = (int)d.ncnt[i];
  float ox = 0.0f, oy = 0.0f, oz = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, dens = 0.0f;
  struct XV {
    float3 x, v;
  };
  visit_neighbors<kUnroll>(
      i, cnt, d.K4, d.nbr, [&](uint32_t j) { return XV{at(xv, 2 * j), at(xv, 2 * j + 1)}; },
      [&](uint32_t, const XV& q) {
        float t[7];
        xsph_terms(c, xi, vi, q.x, q.v, t);
        ox = ox + t[0];
        oy = oy + t[1];
        oz = oz + t[2];
        sx = sx + t[3];
        sy = sy + t[4];
        sz = sz + t[5];
        dens = dens + t[6];
      });
  xsph_finish(i, d, c, xi, vi, ox, oy, oz, sx, sy, sz, dens, out);
}

__device__ __forceinline__ void vort_finish(uint32_t i, const Dev& d, const SimConst& c,
                                            const float4& xi, float ex, float ey, float ez,
                                            float4* x_out, float4* v_out, bool commit);

// applyVorticity + updatePosition (src/particles.cpp:236-248).
// in = (x*, |omega|).  Commits posA = (x*, id), velA = (v, density).
// Returns the committed x (xyz, id) and v of particle i; writes them unless
// the caller goes on with the next step's prediction (commit = false).
__device__ __forceinline__ void vort_one(uint32_t i, const Dev& d, const SimConst& c,
                                         const float4* __restrict__ in, float4* x_out = nullptr,
                                         float4* v_out = nullptr, bool commit = true) {
  const float4 xi = in[i];
  const int cnt = (int)d.ncnt[i];
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  visit_neighbors<kUnroll>(
      i, cnt, d.K4, d.nbr, [&](uint32_t j) { return at(in, j); },
      [&](uint32_t, const float4& xj) {
        float t[3];
        vort_terms(c, xi, xj, t);
        ex = ex + t[0];
        ey = ey + t[1];
        ez = ez + t[2];
      });
  vort_finish(i, d, c, xi, ex, ey, ez, x_out, v_out, commit);
}

// applyVorticity + updatePosition from eta (src/particles.cpp:238-248):
// if |eta| > EPS_D, v += dt*eps_v*(eta/|eta| x omega_i); x <- x*.
__device__ __forceinline__ void vort_finish(uint32_t i, const Dev& d, const SimConst& c,
                                            const float4& xi, float ex, float ey, float ez,
                                            float4* x_out, float4* v_out, bool commit) {
  float4 v = d.vxsph[i];
  const float e2 = ex * ex + ey * ey + ez * ez;
  if (sqrtf(e2) > c.eps_d) {
    const float rn = 1.0f / sqrtf(e2);
    const float u0 = rn * ex, u1 = rn * ey, u2 = rn * ez;
    const float4 w = d.vort[i];
    const float c0 = u1 * w.z - u2 * w.y;
    const float c1 = u2 * w.x - u0 * w.z;
    const float c2 = u0 * w.y - u1 * w.x;
    v.x = v.x + c.dtv * c0;
    v.y = v.y + c.dtv * c1;
    v.z = v.z + c.dtv * c2;
  }
  const float4 x = make_float4(xi.x, xi.y, xi.z, d.posB[i].w);
  d.posA[i] = x;
  if (commit) d.velA[i] = v;
  if (x_out) *x_out = x;
  if (v_out) *v_out = v;
}

// Density of the committed particles at point p (estimateDensityAt,
// src/particles.cpp:446-453): the particles of the 27 cells around p (the
// state sorted by cell, x in predQ), in (cell, id) order, r = x_j - p.  A
// point outside the grid takes the clamped boundary cell, whose
// neighbourhood still holds every particle within h of it.
__device__ __forceinline__ float density_point(float3 p, const Dev& d, const SimConst& c,
                                               const SceneGrid& g) {
  const float4* __restrict__ pred = d.predQ;
  int cc[3];
  cell_lin(g, p, cc);
  const int x0 = cc[0] > 0 ? cc[0] - 1 : 0;
  const int x1 = cc[0] + 1 < g.dims[0] ? cc[0] + 1 : g.dims[0] - 1;
  float dens = 0.0f;
  for (int dz = -1; dz <= 1; dz++) {
    const int z = cc[2] + dz;
    if (z < 0 || z >= g.dims[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int y = cc[1] + dy;
      if (y < 0 || y >= g.dims[1]) continue;
      const uint32_t row =
          g.cell_base + ((uint32_t)z * (uint32_t)g.dims[1] + (uint32_t)y) * (uint32_t)g.dims[0];
      const uint32_t s = d.cellStart[row + (uint32_t)x0];
      const uint32_t e = min(d.cellStart[row + (uint32_t)x1 + 1u], d.n);
      for (uint32_t j = s; j < e; j++) {
        const float4 xj = pred[j];
        const float rx = xj.x - p.x, ry = xj.y - p.y, rz = xj.z - p.z;
        dens = dens + poly6_br(c, rx * rx + ry * ry + rz * rz);
      }
    }
  }
  return dens;
}

// Load-time density (src/particles.cpp:440-444 -> :158-163): the density at
// the particle's own position, self included.  Written by id.
__device__ __forceinline__ void density_one(uint32_t i, const Dev& d, const SimConst& c,
                                            const SceneGrid& g, float* __restrict__ out) {
  const float4 xi = d.predQ[i];
  out[id_of(d.posB[i])] = density_point(make_float3(xi.x, xi.y, xi.z), d, c, g);
}

// ----------------------------------------- one particle on one quad --
// The list build and the lambda, delta-p and XSPH passes of particle i on
// the four lanes of a quad; every lane of the quad calls them.  The quad
// kernels and the dense split's quad workgroups share them.
__device__ __forceinline__ uint32_t quad_index() {
  return (xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) >> 2;
}

template <bool PACKED>
__device__ __forceinline__ void lambda_quad_one(uint32_t i, const Dev& d, const SimConst& c,
                                                const float4* in, float4* out) {
  const float3 xi = PACKED ? d.xs3[i] : make_float3(in[i].x, in[i].y, in[i].z);
  float s[5];
  quad_sums<5>(
      i, (int)d.ncnt[i], d.K4, d.nbr,
      [&](uint32_t j) {
        if (PACKED) return at(d.xs3, j);
        const float4 q = at(in, j);
        return make_float3(q.x, q.y, q.z);
      },
      [&](uint32_t, const float3& xj, float* t) { lambda_terms(c, xi, xj, t); }, s);
  if ((threadIdx.x & 3u) == 0)
    out[i] = make_float4(xi.x, xi.y, xi.z, lambda_of(c, s[0], s[4], s[1], s[2], s[3]));
}

template <bool LAST, int NT>
__device__ __forceinline__ void deltap_quad_one(uint32_t i, const Dev& d, const SimConst& c,
                                                const SceneGrid& g, const float4* in) {
  const float4 xi = in[i];
  float s[3];
  quad_sums<3>(
      i, (int)d.ncnt[i], d.K4, d.nbr, [&](uint32_t j) { return at(in, j); },
      [&](uint32_t, const float4& xj, float* t) { deltap_terms<NT>(c, xi, xj, t); }, s);
  if ((threadIdx.x & 3u) == 0) deltap_finish<LAST>(i, d, c, g, xi, s[0], s[1], s[2]);
}

struct XVRec {
  float3 x, v;
};

__device__ __forceinline__ void xsph_quad_one(uint32_t i, const Dev& d, const SimConst& c,
                                              float4* out) {
  const float3* __restrict__ xv = reinterpret_cast<const float3*>(d.xv);
  const float3 xi = xv[2 * (size_t)i], vi = xv[2 * (size_t)i + 1];
  float s[7];
  quad_sums<7>(
      i, (int)d.ncnt[i], d.K4, d.nbr,
      [&](uint32_t j) { return XVRec{at(xv, 2 * j), at(xv, 2 * j + 1)}; },
      [&](uint32_t, const XVRec& q, float* t) { xsph_terms(c, xi, vi, q.x, q.v, t); }, s);
  if ((threadIdx.x & 3u) == 0)
    xsph_finish(i, d, c, xi, vi, s[0], s[1], s[2], s[3], s[4], s[5], s[6], out);
}

// The neighbour build with a quad per particle (single scene; nbr_one's
// rows, pruning and order).  The quad walks each row R candidates at a
// time: lane k tests candidates k, k + 4, k + 8, ..., and four DPP
// broadcasts give every lane the quad's R-bit hit mask in candidate order.
// A hit lane stores its entry at slot cnt + (hits before it in the mask),
// so the list is nbr_one's.  The slots past the end of the last chunk hold
// i (a particle index, as the readers' gathers need; their terms are not
// added).  The first lambda pass follows on the same quad (the build
// split) or as its own quad launch.
__device__ __forceinline__ uint32_t quad_or(uint32_t m) {
  return (uint32_t)__float_as_int(quad_lane(__int_as_float((int)m), 0)) |
         (uint32_t)__float_as_int(quad_lane(__int_as_float((int)m), 1)) |
         (uint32_t)__float_as_int(quad_lane(__int_as_float((int)m), 2)) |
         (uint32_t)__float_as_int(quad_lane(__int_as_float((int)m), 3));
}

// R candidates per round (16 or 32: four or eight per lane).  Both uses
// take 16: 32 cost mid-size quad builds up to 3%, and in the build split
// (the dam break 0.927 -> 0.920 ms per step) it raised the per-lane build
// kernel that carries it from 72 to 107 VGPRs, which slowed the 4M run's
// build (4.08 -> 4.13 ms per step).
template <int R>
__device__ __forceinline__ void nbr_quad_one(uint32_t i, const Dev& d, float h2,
                                             const SceneGrid& g) {
  static_assert(R == 16 || R == 32, "16 or 32 candidates per round");
  const uint32_t k = threadIdx.x & 3u;
  const float4 xi = d.predQ[i];
  int cc[3];
  cell_lin(g, make_float3(xi.x, xi.y, xi.z), cc);
  const float cell = 1.0f / g.inv_cell;
  const float hm2 = h2 * 1.0201f;  // (1.01 h)^2
  const float4* __restrict__ pred = d.predQ;
  const uint32_t cap = (uint32_t)d.K4 * kChunk;
  auto hit = [&](const float4& xj) -> uint32_t {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float d2 = rx * rx + ry * ry + rz * rz;
    return d2 <= h2 ? 1u : 0u;
  };
  auto put = [&](uint32_t slot, uint32_t j) {
    if (slot < cap) d.nbr[nbr_chunk(i, (int)(slot >> 2), d.K4) + (slot & 3u)] = j;
  };
  uint32_t cnt = 0;
  const float gxl = face_gap(xi.x, g.lo[0], cell, cc[0], -1);
  const float gxr = face_gap(xi.x, g.lo[0], cell, cc[0], 1);
  for (int dz = -1; dz <= 1; dz++) {
    const int z = cc[2] + dz;
    if (z < 0 || z >= g.dims[2]) continue;
    const float gz = face_gap(xi.z, g.lo[2], cell, cc[2], dz);
    for (int dy = -1; dy <= 1; dy++) {
      const int y = cc[1] + dy;
      if (y < 0 || y >= g.dims[1]) continue;
      const float gy = face_gap(xi.y, g.lo[1], cell, cc[1], dy);
      const float yz2 = gy * gy + gz * gz;
      if (yz2 > hm2) continue;
      const int x0 = (cc[0] > 0 && gxl * gxl + yz2 <= hm2) ? cc[0] - 1 : cc[0];
      const int x1 = (cc[0] + 1 < g.dims[0] && gxr * gxr + yz2 <= hm2) ? cc[0] + 1 : cc[0];
      const uint32_t row =
          g.cell_base + ((uint32_t)z * (uint32_t)g.dims[1] + (uint32_t)y) * (uint32_t)g.dims[0];
      const uint32_t s = d.cellStart[row + (uint32_t)x0];
      const uint32_t e = min(d.cellStart[row + (uint32_t)x1 + 1u], d.n);
      for (uint32_t jb = s; jb < e; jb += R) {
        // Up to R - 1 candidates past the range are read (the x* arrays are
        // padded by 64 entries); their bits are dropped.
        uint32_t mine = 0;
#pragma unroll
        for (uint32_t u = 0; u < R / 4; u++) {
          const uint32_t j = jb + k + 4u * u;
          mine |= (hit(at(pred, j)) & (uint32_t)(j < e && j != i)) << (4u * u + k);
        }
        const uint32_t m = quad_or(mine);
#pragma unroll
        for (uint32_t u = 0; u < R / 4; u++) {
          const uint32_t t = 4u * u + k;
          if ((m >> t) & 1u) put(cnt + (uint32_t)__builtin_popcount(m & ((1u << t) - 1u)), jb + t);
        }
        cnt += (uint32_t)__builtin_popcount(m);
      }
    }
  }
  // Slots cnt .. end of its chunk: lane k fills slot cnt + k if that is in
  // the same chunk.
  const uint32_t pad = cnt + k;
  if ((cnt & 3u) && (pad >> 2) == (cnt >> 2)) put(pad, i);
  if (k == 0) {
    if (cnt > cap) atomicMax(&d.flags[0], cnt);
    d.ncnt[i] = min(cnt, cap);
  }
}

// Dense split.  In a split launch the first d.nlane workgroups run the
// per-lane pass (lane_particle: ~0 for queued particles, which they skip)
// and the rest take the queued particles, a quad each, in queue order (a
// particle's result does not depend on which workgroup computes it).
__device__ __forceinline__ bool dense_quads_part(const Dev& d) {
  return d.nlane && blockIdx.x >= d.nlane;
}
__device__ __forceinline__ uint32_t lane_particle(const Dev& d) {
  if (!d.nlane) return xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const uint32_t i = xcd_block(blockIdx.x, d.nlane) * kBlock + threadIdx.x;
  return i < d.n && d.ncnt[i] > d.dense_list ? 0xffffffffu : i;
}
template <class F>
__device__ __forceinline__ void for_dense_quads(const Dev& d, F f) {
  const uint32_t qn = min(d.flags[kDenseSlot], d.n);
  const uint32_t stride = (gridDim.x - d.nlane) * (kBlock / 4);
  for (uint32_t q = (blockIdx.x - d.nlane) * (kBlock / 4) + (threadIdx.x >> 2); q < qn;
       q += stride)
    f(d.dense[q]);  // a quad's lanes share q
}

// Whether sorted particle i's cell holds more than dense_cell particles
// (the test the sort's queue used).
__device__ __forceinline__ bool crowded(const Dev& d, uint32_t i) {
  const uint32_t c = d.cellSorted[i];
  return min(d.cellStart[c + 1], d.n) - d.cellStart[c] > d.dense_cell;
}

// Queues particle i (its list longer than the threshold) for the dense split.
__device__ __forceinline__ void queue_dense(const Dev& d, uint32_t i, uint32_t cnt) {
  if (cnt <= d.dense_list) return;
  const uint32_t slot = atomicAdd(&d.flags[kDenseSlot], 1u);
  if (slot < d.n) d.dense[slot] = i;  // each particle once per step: slot < n
}

// ------------------------------------------------------- flat schedule --
template <bool MULTI>
__device__ __forceinline__ SceneGrid grid_of(uint32_t i, const SceneGrid& g0, const Dev& d) {
  if (MULTI) return d.grids[d.slotScene[i]];
  return g0;
}

__device__ __forceinline__ uint32_t flat_index() { return blockIdx.x * kBlock + threadIdx.x; }
// Gather-heavy passes take the XCD-aware block order.
__device__ __forceinline__ uint32_t gather_index() {
  return xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
}

template <bool MULTI, bool MOVE>
__global__ __launch_bounds__(kBlock) void k_predict(Dev d, SimConst c, SceneGrid g0) {
  const uint32_t i = flat_index();
  const bool valid = i < d.n;
  uint32_t cl = 0xffffffffu;
  if (valid) cl = predict_cell<MOVE>(i, d, c, grid_of<MULTI>(i, g0, d));
  const uint32_t slot = wave_run_slot(cl, valid, d.cellCount);
  if (valid) d.arrival[i] = slot;
}

// The prediction of k_predict<MULTI, true> from the committed state, written
// to `out` (the array where a prediction fused into the previous step's last
// pass would be) instead of predP.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_predict_into(Dev d, SimConst c, SceneGrid g0,
                                                         float4* out) {
  const uint32_t i = flat_index();
  const bool valid = i < d.n;
  uint32_t cl = 0xffffffffu;
  if (valid)
    cl = predict_from<true>(i, d, c, grid_of<MULTI>(i, g0, d), d.posA[i], d.velA[i], out);
  const uint32_t slot = wave_run_slot(cl, valid, d.cellCount);
  if (valid) d.arrival[i] = slot;
}

// The dense split's queue counters start at zero in every step: the build
// queue's before the rank pass fills it (here), the list queue's before the
// build fills it (in k_rank).
__global__ __launch_bounds__(kBlock) void k_scatter(Dev d) {
  const uint32_t i = flat_index();
  if (i < d.n) scatter_one(i, d);
  if (i == 0 && d.dense_b) d.flags[kDenseBuildSlot] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_rank(Dev d, const float4* pred) {
  const uint32_t i = flat_index();
  if (i < d.n) rank_one(i, d, pred);
  if (i == 0 && d.dense) d.flags[kDenseSlot] = 0u;
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_nbr_build(Dev d, float h2, SceneGrid g0) {
  const uint32_t i = gather_index();
  if (i >= d.n) return;
  nbr_one(i, d, h2, grid_of<MULTI>(i, g0, d));
  if (!MULTI && d.dense) queue_dense(d, i, d.ncnt[i]);
}

// The neighbour build followed, in the same thread, by the first density/λ
// pass over the list it has just written (its own stores, read back in
// program order).  Waves in the memory-bound build overlap waves in the
// VALU-bound λ pass.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_nbr_lambda(Dev d, SimConst c, SceneGrid g0,
                                                       float4* out) {
  if (!MULTI && dense_quads_part(d)) {
    // Build split: the particles of crowded cells, a quad each.
    const uint32_t qn = min(d.flags[kDenseBuildSlot], d.n);
    const uint32_t stride = (gridDim.x - d.nlane) * (kBlock / 4);
    for (uint32_t q = (blockIdx.x - d.nlane) * (kBlock / 4) + (threadIdx.x >> 2); q < qn;
         q += stride) {
      const uint32_t i = d.dense_b[q];  // a quad's lanes share q
      nbr_quad_one<16>(i, d, c.h2, g0);
      __threadfence_block();  // the quad's list stores before its loads
      lambda_quad_one<false>(i, d, c, d.predQ, out);
      if ((threadIdx.x & 3u) == 0 && d.dense) queue_dense(d, i, d.ncnt[i]);
    }
    return;
  }
  uint32_t i = gather_index();
  if (!MULTI && d.nlane) {
    i = xcd_block(blockIdx.x, d.nlane) * kBlock + threadIdx.x;
    if (i < d.n && crowded(d, i)) return;  // built by a quad
  }
  if (i >= d.n) return;
  nbr_one(i, d, c.h2, grid_of<MULTI>(i, g0, d));
  lambda_one<false>(i, d, c, d.predQ, out);
  if (!MULTI && d.dense) queue_dense(d, i, d.ncnt[i]);
}

template <bool PACKED>
__global__ __launch_bounds__(kBlock) void k_lambda(Dev d, SimConst c, const float4* in,
                                                   float4* out) {
  if (dense_quads_part(d)) {
    for_dense_quads(d, [&](uint32_t i) { lambda_quad_one<PACKED>(i, d, c, in, out); });
    return;
  }
  const uint32_t i = lane_particle(d);
  if (i < d.n) lambda_one<PACKED>(i, d, c, in, out);
}

template <bool MULTI, bool LAST, int NT>
__global__ __launch_bounds__(kBlock) void k_deltap(Dev d, SimConst c, SceneGrid g0,
                                                   const float4* in, float4* out) {
  if (!MULTI && dense_quads_part(d)) {
    for_dense_quads(d, [&](uint32_t i) { deltap_quad_one<LAST, NT>(i, d, c, g0, in); });
    return;
  }
  const uint32_t i = lane_particle(d);
  if (i < d.n) deltap_one<LAST, NT>(i, d, c, grid_of<MULTI>(i, g0, d), in, out);
}

__global__ __launch_bounds__(kBlock) void k_xsph(Dev d, SimConst c, float4* out) {
  if (dense_quads_part(d)) {
    for_dense_quads(d, [&](uint32_t i) { xsph_quad_one(i, d, c, out); });
    return;
  }
  const uint32_t i = lane_particle(d);
  if (i < d.n) xsph_one(i, d, c, out);
}

// Copies the (x*, v) records of the last iteration into float4 arrays (for
// debug stops and the adjoint, which read x* and v there).
__global__ __launch_bounds__(kBlock) void k_unpack_record(Dev d, float4* __restrict__ pos,
                                                          float4* __restrict__ vel) {
  const uint32_t i = flat_index();
  if (i >= d.n) return;
  const float3* r = xv_record(d, i);
  const float3 x = r[0], v = r[1];
  pos[i] = make_float4(x.x, x.y, x.z, 0.0f);
  vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// Copies the 12-byte x* records into a float4 array (w = 0), for debug
// stops and the adjoint's tape after a delta-p pass before the last.
__global__ __launch_bounds__(kBlock) void k_unpack_xs3(Dev d, float4* __restrict__ pos) {
  const uint32_t i = flat_index();
  if (i >= d.n) return;
  const float3 x = d.xs3[i];
  pos[i] = make_float4(x.x, x.y, x.z, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_vort_commit(Dev d, SimConst c, const float4* in) {
  const uint32_t i = gather_index();
  if (i < d.n) vort_one(i, d, c, in);
}

// Vorticity confinement and commit, then, from the committed x and v still
// in registers, the next step's prediction, cell and histogram slot (as
// k_predict; cellCount must be zero).  Used between the steps of one call.
// The prediction goes to the vort array (this particle's omega, already
// read): predP is `in`, still being gathered by other threads.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_vort_commit_predict(Dev d, SimConst c, SceneGrid g0,
                                                                const float4* in) {
  const uint32_t i = gather_index();
  const bool valid = i < d.n;
  uint32_t cl = 0xffffffffu;
  if (valid) {
    float4 x, v;
    vort_one(i, d, c, in, &x, &v, false);
    cl = predict_from<true>(i, d, c, grid_of<MULTI>(i, g0, d), x, v, d.vort);
  }
  const uint32_t slot = wave_run_slot(cl, valid, d.cellCount);
  if (valid) d.arrival[i] = slot;
}

template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_densities(Dev d, SimConst c, SceneGrid g0,
                                                      float* out) {
  const uint32_t i = flat_index();
  if (i < d.n) density_one(i, d, c, grid_of<MULTI>(i, g0, d), out);
}

// Density field at m points of one scene: pts[3q..3q+2], or with LATTICE the
// point (min(hi, lo + ix*step), ...) of q = (iz*ny + iy)*nx + ix.
template <bool LATTICE>
__global__ __launch_bounds__(kBlock) void k_density_field(Dev d, SimConst c, SceneGrid g,
                                                          const float* __restrict__ pts,
                                                          DensityLattice L, uint32_t m,
                                                          float* __restrict__ out) {
  const uint32_t q = flat_index();
  if (q >= m) return;
  float3 p;
  if (LATTICE) {
    const uint32_t ix = q % L.n[0], iy = (q / L.n[0]) % L.n[1], iz = q / (L.n[0] * L.n[1]);
    p = make_float3(smin(L.hi[0], L.lo[0] + (float)ix * L.step),
                    smin(L.hi[1], L.lo[1] + (float)iy * L.step),
                    smin(L.hi[2], L.lo[2] + (float)iz * L.step));
  } else {
    p = make_float3(pts[3 * (size_t)q], pts[3 * (size_t)q + 1], pts[3 * (size_t)q + 2]);
  }
  out[q] = density_point(p, d, c, g);
}

// ------------------------------------------------- persistent schedule --
// One launch advances every scene by nsteps steps.  Workgroups take scenes
// from a queue in the host's order (largest first); a workgroup runs all
// passes of all steps of its scene, with workgroup barriers between passes.
// A scene's particles and cells are contiguous ranges of the global arrays,
// so every pass writes exactly what the flat schedule writes.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_persistent(Dev d, SimConst c, const SceneRun* runs,
                                                      uint32_t nruns, uint32_t* queue,
                                                      int iterations, int nsteps) {
  __shared__ uint32_t wsum[BLOCK / 64 + 1];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();
    if (item >= nruns) break;
    const SceneRun r = runs[item];
    const SceneGrid g = d.grids[r.scene];
    const uint32_t b = r.offset, e = r.offset + r.count;
    const uint32_t c0 = g.cell_base, nc = r.ncells;
    for (int step = 0; step < nsteps; step++) {
      for (uint32_t k = threadIdx.x; k < nc; k += BLOCK) store_agent(d.cellCount + c0 + k, 0u);
      phase_barrier();
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) predict_one<true>(i, d, c, g);
      phase_barrier();
      block_scan<BLOCK>(d.cellCount + c0, d.cellStart + c0, nc, b, wsum);
      phase_barrier();
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) scatter_one(i, d);
      phase_barrier();
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) rank_one(i, d, d.predP);
      phase_barrier();
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) nbr_one(i, d, c.h2, g);
      phase_barrier();
      for (int it = 0; it < iterations; it++) {
        if (it == 0) {
          for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK)
            lambda_one<false>(i, d, c, d.predQ, d.predP);
        } else {
          for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK)
            lambda_one<true>(i, d, c, d.predQ, d.predP);
        }
        phase_barrier();
        if (it == iterations - 1) {
          for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK)
            deltap_one<true>(i, d, c, g, d.predP, d.predQ);
        } else {
          for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK)
            deltap_one<false>(i, d, c, g, d.predP, d.predQ);
        }
        phase_barrier();
      }
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) xsph_one(i, d, c, d.predP);
      phase_barrier();
      for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) vort_one(i, d, c, d.predP);
      phase_barrier();
    }
  }
}

// ------------------------------------------ small states: wide passes --
// With a few thousand particles a list pass lasts as long as one lane's walk
// through its own list (about eight dependent chunk gathers for ~31
// neighbours) while most of the chip idles.  The wide passes give each
// particle a wave of its own: lane k evaluates the terms of neighbour k
// with the term functions above, the terms go to LDS, and one lane per sum
// adds them in list order from +0 -- the additions of the per-lane loops in
// the same order, hence the same bits.  Lists longer than 64 are taken 64
// entries at a time, each block added to the running sums before the next.
constexpr int kWide = 64;            // one wave per particle and workgroup
constexpr int kWideRow = kWide + 1;  // padded rows: the sum lanes hit distinct banks

// Adds the first m terms of each of the NS series in T to the running sums
// (lane s < NS holds series s), in order.  Uniform control flow: every lane
// of the wave calls it.
template <int NS>
__device__ __forceinline__ void wide_add(float (*T)[kWideRow], int m, float& acc) {
  const int lane = (int)threadIdx.x;
  __syncthreads();
  if (lane < NS) {
    int q = 0;
    for (; q + 4 <= m; q += 4) {
      const float a0 = T[lane][q], a1 = T[lane][q + 1], a2 = T[lane][q + 2],
                  a3 = T[lane][q + 3];
      acc = acc + a0;
      acc = acc + a1;
      acc = acc + a2;
      acc = acc + a3;
    }
    for (; q < m; q++) acc = acc + T[lane][q];
  }
  __syncthreads();
}

// The NS sums, on every lane.
template <int NS>
__device__ __forceinline__ void wide_result(float (*T)[kWideRow], float acc, float* out) {
  if ((int)threadIdx.x < NS) T[threadIdx.x][0] = acc;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NS; q++) out[q] = T[q][0];
}

// Sums over the cnt list entries of particle i; term(j, t) gives the NS
// terms of neighbour j.
template <int NS, class Term>
__device__ __forceinline__ void wide_list_sums(uint32_t i, int cnt, const Dev& d, Term term,
                                               float (*T)[kWideRow], float* out) {
  const int lane = (int)threadIdx.x;
  float acc = 0.0f;
  for (int base = 0; base < cnt; base += kWide) {
    const int k = base + lane;
    if (k < cnt) {
      float t[NS];
      term(d.nbr[nbr_slot(i, k, d.K4)], t);
#pragma unroll
      for (int q = 0; q < NS; q++) T[q][lane] = t[q];
    }
    wide_add<NS>(T, min(kWide, cnt - base), acc);
  }
  wide_result<NS>(T, acc, out);
}

// nbr_one and the first lambda of particle blockIdx.x: the lanes test 64
// candidates at a time, the hits enter the list in candidate order (a ballot
// and the count of hits in lower lanes give each its slot), and the lambda
// terms of the hits, whose x* the testing lanes hold, are added block by
// block.  The same rows, end cells, candidate ranges and test as nbr_one,
// so the same list; the partial last chunk is padded with its first entry
// as nbr_one pads it.
template <bool MULTI>
__global__ __launch_bounds__(kWide) void k_nbr_lambda_wide(Dev d, SimConst c, SceneGrid g0,
                                                           float4* out) {
  __shared__ float T[5][kWideRow];
  const uint32_t i = blockIdx.x;
  if (i >= d.n) return;
  const SceneGrid g = grid_of<MULTI>(i, g0, d);
  const uint32_t lane = threadIdx.x;
  const float4 x4 = d.predQ[i];
  const float3 xi = make_float3(x4.x, x4.y, x4.z);
  int cc[3];
  cell_lin(g, xi, cc);
  const float cell = 1.0f / g.inv_cell;
  const float h2 = c.h2, hm2 = h2 * 1.0201f;  // (1.01 h)^2
  const float4* __restrict__ pred = d.predQ;
  const uint32_t cap = (uint32_t)d.K4 * kChunk;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t cnt = 0;
  float acc = 0.0f;
  // The nine (dz, dy) rows, each a contiguous candidate range in sorted
  // order, found by lanes 0..8 at once (the culling of nbr_one), then walked
  // as one sequence in row order, 64 candidates per round.
  __shared__ uint32_t row_s[9], row_at[10];
  if (lane < 9u) {
    const int dz = (int)lane / 3 - 1, dy = (int)lane % 3 - 1;
    const int z = cc[2] + dz, y = cc[1] + dy;
    uint32_t s = 0, len = 0;
    if (z >= 0 && z < g.dims[2] && y >= 0 && y < g.dims[1]) {
      const float gxl = face_gap(xi.x, g.lo[0], cell, cc[0], -1);
      const float gxr = face_gap(xi.x, g.lo[0], cell, cc[0], 1);
      const float gz = face_gap(xi.z, g.lo[2], cell, cc[2], dz);
      const float gy = face_gap(xi.y, g.lo[1], cell, cc[1], dy);
      const float yz2 = gy * gy + gz * gz;
      if (!(yz2 > hm2)) {
        const int x0 = (cc[0] > 0 && gxl * gxl + yz2 <= hm2) ? cc[0] - 1 : cc[0];
        const int x1 = (cc[0] + 1 < g.dims[0] && gxr * gxr + yz2 <= hm2) ? cc[0] + 1 : cc[0];
        const uint32_t row = g.cell_base +
                             ((uint32_t)z * (uint32_t)g.dims[1] + (uint32_t)y) * (uint32_t)g.dims[0];
        s = d.cellStart[row + (uint32_t)x0];
        const uint32_t e = min(d.cellStart[row + (uint32_t)x1 + 1u], d.n);
        len = e > s ? e - s : 0u;
      }
    }
    row_s[lane] = s;
    row_at[lane + 1] = len;
  }
  __syncthreads();
  if (lane == 0) {
    row_at[0] = 0;
    for (int r = 0; r < 9; r++) row_at[r + 1] += row_at[r];
  }
  __syncthreads();
  const uint32_t total = row_at[9];
  for (uint32_t tb = 0; tb < total; tb += kWide) {
    const uint32_t t = tb + lane;
    bool hit = false;
    float3 xj = xi;
    uint32_t j = 0;
    if (t < total) {
      int r = 0;
      while (row_at[r + 1] <= t) r++;
      j = row_s[r] + (t - row_at[r]);
      if (j != i) {
        const float4 q = pred[j];
        xj = make_float3(q.x, q.y, q.z);
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        hit = rx * rx + ry * ry + rz * rz <= h2;
      }
    }
    const unsigned long long mask = __ballot(hit);
    const uint32_t before = (uint32_t)__popcll(mask & below);
    const uint32_t pos = cnt + before;
    if (hit && pos < cap) {
      d.nbr[nbr_slot(i, (int)pos, d.K4)] = j;
      float tm[5];
      lambda_terms(c, xi, xj, tm);
#pragma unroll
      for (int q = 0; q < 5; q++) T[q][before] = tm[q];
    }
    const uint32_t m = (uint32_t)__popcll(mask);
    const uint32_t room = cnt < cap ? cap - cnt : 0u;
    wide_add<5>(T, (int)min(m, room), acc);
    cnt += m;
  }
  const uint32_t nc = min(cnt, cap);
  // Pad the partial last chunk (its first entry repeated), as nbr_one does;
  // the barrier in wide_result orders the list stores before the reads.
  float sum[5];
  wide_result<5>(T, acc, sum);
  if (lane == 0) {
    const uint32_t r = nc & 3u;
    if (r) {
      const int first = (int)(nc - r);
      const uint32_t e0 = d.nbr[nbr_slot(i, first, d.K4)];
      for (uint32_t q = r; q < 4u; q++) d.nbr[nbr_slot(i, first + (int)q, d.K4)] = e0;
    }
    if (cnt > cap) atomicMax(&d.flags[0], cnt);
    d.ncnt[i] = nc;
    out[i] = make_float4(xi.x, xi.y, xi.z, lambda_of(c, sum[0], sum[4], sum[1], sum[2], sum[3]));
  }
}

// lambda_one, wide.
template <bool PACKED>
__global__ __launch_bounds__(kWide) void k_lambda_wide(Dev d, SimConst c, const float4* in,
                                                       float4* out) {
  __shared__ float T[5][kWideRow];
  const uint32_t i = blockIdx.x;
  if (i >= d.n) return;
  const float3 xi = PACKED ? d.xs3[i] : make_float3(in[i].x, in[i].y, in[i].z);
  float sum[5];
  wide_list_sums<5>(
      i, (int)d.ncnt[i], d,
      [&](uint32_t j, float* t) {
        float3 xj;
        if (PACKED) {
          xj = at(d.xs3, j);
        } else {
          const float4 q = at(in, j);
          xj = make_float3(q.x, q.y, q.z);
        }
        lambda_terms(c, xi, xj, t);
      },
      T, sum);
  if (threadIdx.x == 0)
    out[i] = make_float4(xi.x, xi.y, xi.z, lambda_of(c, sum[0], sum[4], sum[1], sum[2], sum[3]));
}

// deltap_one, wide.
template <bool MULTI, bool LAST, int NT>
__global__ __launch_bounds__(kWide) void k_deltap_wide(Dev d, SimConst c, SceneGrid g0,
                                                       const float4* in) {
  __shared__ float T[3][kWideRow];
  const uint32_t i = blockIdx.x;
  if (i >= d.n) return;
  const float4 xi = in[i];
  float sum[3];
  wide_list_sums<3>(
      i, (int)d.ncnt[i], d,
      [&](uint32_t j, float* t) { deltap_terms<NT>(c, xi, at(in, j), t); }, T, sum);
  if (threadIdx.x == 0)
    deltap_finish<LAST>(i, d, c, grid_of<MULTI>(i, g0, d), xi, sum[0], sum[1], sum[2]);
}

// xsph_one, wide.
__global__ __launch_bounds__(kWide) void k_xsph_wide(Dev d, SimConst c, float4* out) {
  __shared__ float T[7][kWideRow];
  const uint32_t i = blockIdx.x;
  if (i >= d.n) return;
  const float3* __restrict__ xv = reinterpret_cast<const float3*>(d.xv);
  const float3 xi = xv[2 * (size_t)i], vi = xv[2 * (size_t)i + 1];
  float sum[7];
  wide_list_sums<7>(
      i, (int)d.ncnt[i], d,
      [&](uint32_t j, float* t) { xsph_terms(c, xi, vi, at(xv, 2 * j), at(xv, 2 * j + 1), t); },
      T, sum);
  if (threadIdx.x == 0)
    xsph_finish(i, d, c, xi, vi, sum[0], sum[1], sum[2], sum[3], sum[4], sum[5], sum[6], out);
}

// vort_one, wide; with PREDICT the next step's prediction, cell and
// histogram slot follow as in k_vort_commit_predict (one atomic per
// particle: the slot order inside a cell does not matter, the rank orders
// by id).
template <bool PREDICT, bool MULTI>
__global__ __launch_bounds__(kWide) void k_vort_wide(Dev d, SimConst c, SceneGrid g0,
                                                     const float4* in) {
  __shared__ float T[3][kWideRow];
  const uint32_t i = blockIdx.x;
  if (i >= d.n) return;
  const float4 xi = in[i];
  float sum[3];
  wide_list_sums<3>(
      i, (int)d.ncnt[i], d, [&](uint32_t j, float* t) { vort_terms(c, xi, at(in, j), t); }, T,
      sum);
  if (threadIdx.x != 0) return;
  if (!PREDICT) {
    vort_finish(i, d, c, xi, sum[0], sum[1], sum[2], nullptr, nullptr, true);
  } else {
    float4 x, v;
    vort_finish(i, d, c, xi, sum[0], sum[1], sum[2], &x, &v, false);
    const uint32_t cl = predict_from<true>(i, d, c, grid_of<MULTI>(i, g0, d), x, v, d.vort);
    d.arrival[i] = atomicAdd(&d.cellCount[cl], 1u);
  }
}

// ---- quad list passes: four lanes per particle ------------------------------
// With one lane per particle a state of some ten thousand particles leaves
// the chip with few waves, each walking its particles' lists one entry
// after another; with one wave per particle every 64-entry block of a list
// costs a sequential sum in one lane.  With a quad per particle (quad_sums)
// there are four times the waves of the lane-per-particle passes and each
// walks a quarter of the list; the sums are the same bits.
template <bool PACKED>
__global__ __launch_bounds__(kBlock) void k_lambda_quad(Dev d, SimConst c, const float4* in,
                                                        float4* out) {
  const uint32_t i = quad_index();
  if (i < d.n) lambda_quad_one<PACKED>(i, d, c, in, out);  // a quad's lanes leave together
}

template <bool LAST, int NT>
__global__ __launch_bounds__(kBlock) void k_deltap_quad(Dev d, SimConst c, SceneGrid g,
                                                        const float4* in) {
  const uint32_t i = quad_index();
  if (i < d.n) deltap_quad_one<LAST, NT>(i, d, c, g, in);
}

__global__ __launch_bounds__(kBlock) void k_nbr_quad(Dev d, float h2, SceneGrid g) {
  const uint32_t i = quad_index();
  if (i < d.n) nbr_quad_one<16>(i, d, h2, g);  // a quad's lanes leave together
}

// The vorticity confinement and commit (no next prediction) with a quad per
// particle.
// With PREDICT, the next step's prediction from the committed state too,
// and its histogram slot (one atomic per particle: the sort ranks a cell's
// particles by id, whatever their slots).
template <bool PREDICT>
__global__ __launch_bounds__(kBlock) void k_vort_quad(Dev d, SimConst c, SceneGrid g,
                                                      const float4* in) {
  const uint32_t i = quad_index();
  if (i >= d.n) return;
  const float4 xi = in[i];
  float s[3];
  quad_sums<3>(
      i, (int)d.ncnt[i], d.K4, d.nbr, [&](uint32_t j) { return at(in, j); },
      [&](uint32_t, const float4& xj, float* t) { vort_terms(c, xi, xj, t); }, s);
  if ((threadIdx.x & 3u) != 0) return;
  if (PREDICT) {
    float4 x, v;
    vort_finish(i, d, c, xi, s[0], s[1], s[2], &x, &v, false);
    const uint32_t cl = predict_from<true>(i, d, c, g, x, v, d.vort);
    d.arrival[i] = atomicAdd(&d.cellCount[cl], 1u);
  } else {
    vort_finish(i, d, c, xi, s[0], s[1], s[2], nullptr, nullptr, true);
  }
}

__global__ __launch_bounds__(kBlock) void k_xsph_quad(Dev d, SimConst c, float4* out) {
  const uint32_t i = quad_index();
  if (i < d.n) xsph_quad_one(i, d, c, out);
}

// Scan, scatter and rank of a small state in one workgroup: the three
// passes that otherwise take four or five launches, with workgroup barriers
// between them (every thread of the workgroup shares one CU's cache, so a
// barrier makes the phase's global writes visible).  The same per-particle
// code as the flat launches: the same bits.
constexpr int kSortSmallBlock = 1024;
__global__ __launch_bounds__(kSortSmallBlock) void k_sort_small(Dev d, const float4* pred,
                                                                uint32_t ncell) {
  __shared__ uint32_t wsum[kSortSmallBlock / 64 + 1];
  // The dense split's queue counters, as k_scatter and k_rank reset them
  // (visible to the rank phase's atomics after the barriers below).
  if (threadIdx.x == 0) {
    if (d.dense_b) d.flags[kDenseBuildSlot] = 0u;
    if (d.dense) d.flags[kDenseSlot] = 0u;
  }
  block_scan<kSortSmallBlock>(d.cellCount, d.cellStart, ncell, 0u, wsum);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < d.n; i += kSortSmallBlock) scatter_one(i, d);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < d.n; i += kSortSmallBlock) rank_one(i, d, pred);
  // The histogram is spent (the scan above read it): zero it for the next
  // step's prediction, which then needs no fill launch of its own.
  for (uint32_t k = threadIdx.x; k < ncell; k += kSortSmallBlock) d.cellCount[k] = 0u;
}

// ------------------------------------------------------------ read-back --
__global__ __launch_bounds__(kBlock) void k_unpermute(uint32_t n, const float4* __restrict__ ids,
                                                      const float4* __restrict__ src, int first,
                                                      int count, float* __restrict__ dst) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 v = src[i];
  const float c[4] = {v.x, v.y, v.z, v.w};
  const size_t o = (size_t)id_of(ids[i]) * (size_t)count;
  for (int a = 0; a < count; a++) dst[o + a] = c[first + a];
}

__global__ __launch_bounds__(kBlock) void k_unpermute_u32(uint32_t n,
                                                          const float4* __restrict__ ids,
                                                          const uint32_t* __restrict__ src,
                                                          uint32_t* __restrict__ dst) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dst[id_of(ids[i])] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_nbr_export(Dev d, uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ out, int stride) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= d.n) return;
  const uint32_t id = id_of(d.posB[i]);
  const int cnt = (int)d.ncnt[i];
  counts[id] = (uint32_t)cnt;
  const int m = cnt < stride ? cnt : stride;
  for (int k = 0; k < m; k++)
    out[(size_t)id * stride + k] = id_of(d.posB[d.nbr[nbr_slot(i, k, d.K4)]]);
}

// Sum and maximum of the neighbour counts (64-bit sum, maximum).
__global__ __launch_bounds__(kBlock) void k_nbr_stats(uint32_t n, const uint32_t* __restrict__ ncnt,
                                                      unsigned long long* __restrict__ sum,
                                                      uint32_t* __restrict__ mx) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  unsigned long long s = 0;
  uint32_t m = 0;
  for (; i < n; i += gridDim.x * kBlock) {
    const uint32_t c = ncnt[i];
    s += c;
    m = c > m ? c : m;
  }
  atomicAdd(sum, s);
  atomicMax(mx, m);
}

// Exhaustive check of the short division, reciprocal and square root
// sequences against the IEEE operations: thread m takes significand m and
// runs it through every binade of the checked ranges.  bad[0..2] count the
// mismatches of the division by b (0 and [2^div_e_lo, 2^div_e_hi)), the
// reciprocal ([2^-100, 2^100)) and the square root ([2^-96, 2^128); below
// 2^-96, including 0 and the subnormals, a result neither NaN nor under
// 2^-47 counts as a mismatch).
__global__ __launch_bounds__(kBlock) void k_verify_math(float b, float rb, int div_e_lo,
                                                        int div_e_hi, uint32_t* __restrict__ bad) {
  const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= (1u << 23)) return;
  auto differ = [](float u, float v) { return __float_as_uint(u) != __float_as_uint(v) ? 1u : 0u; };
  uint32_t nd = 0, nr = 0, ns = 0;
  if (m == 0) nd += differ(div_seq(0.0f, b, rb), 0.0f / b);
  for (int e = div_e_lo; e < div_e_hi; e++) {
    const float a = __uint_as_float(((uint32_t)(e + 127) << 23) | m);
    nd += differ(div_seq(a, b, rb), a / b);
  }
  for (int e = -100; e < 100; e++) {
    const float z = __uint_as_float(((uint32_t)(e + 127) << 23) | m);
    nr += differ(rcp_seq(z), 1.0f / z);
  }
  for (int e = -96; e < 128; e++) {
    const float x = __uint_as_float(((uint32_t)(e + 127) << 23) | m);
    ns += differ(sqrt_rn(x), sqrtf(x));
  }
  // Below 2^-96 (0, subnormals, binades -126..-97): NaN or under 2^-47.
  auto tiny_ok = [](float r) { return (r != r || r < 0x1p-47f) ? 0u : 1u; };
  for (int e = -127; e < -96; e++) {
    const float x = __uint_as_float(((uint32_t)(e + 127) << 23) | m);
    ns += tiny_ok(sqrt_rn(x));
  }
  if (nd) atomicAdd(bad, nd);
  if (nr) atomicAdd(bad + 1, nr);
  if (ns) atomicAdd(bad + 2, ns);
}


}  // namespace

hipError_t launch_verify_math(float b, float rb, int div_e_lo, int div_e_hi, uint32_t* bad,
                              hipStream_t s) {
  k_verify_math<<<(1u << 23) / kBlock, kBlock, 0, s>>>(b, rb, div_e_lo, div_e_hi, bad);
  return hipGetLastError();
}

// ------------------------------------------------------------ launchers --
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + kBlock - 1) / kBlock;
  unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
  k_fill_u32<<<grid, kBlock, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_scan_u32(uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials,
                           hipStream_t s, bool clear_in) {
  const uint32_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb > 0) k_scan_block<<<nb, kBlock, 0, s>>>(in, out, n, partials, clear_in);
  k_scan_partials<<<1, 1024, 0, s>>>(partials, nb, out, n);
  if (nb > 1) k_scan_add<<<blocks_for(n), kBlock, 0, s>>>(out, n, partials);
  return hipGetLastError();
}

static bool multi(const Ctx& x) { return x.nscenes > 1; }

hipError_t launch_predict(Ctx& x, const SimConst& c) {
  hipError_t e = launch_fill_u32(x.d.cellCount, x.ncell, 0u, x.stream);
  if (e != hipSuccess) return e;
  const Dev d = dev_of(x);
  if (multi(x))
    k_predict<true, true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0]);
  else
    k_predict<false, true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0]);
  return hipGetLastError();
}

hipError_t launch_scan(Ctx& x) {
  // The histogram is spent once scanned: the tile pass clears it for the
  // next prediction, which then needs no fill launch of its own.
  return launch_scan_u32(x.d.cellCount, x.d.cellStart, x.ncell, x.d.blockSums, x.stream, true);
}

bool sort_in_one_workgroup(const Ctx& x) {
  return x.n <= (int64_t)kSortSmallMax && x.ncell <= kSortSmallCells;
}

hipError_t launch_sort_small(Ctx& x, const float4* pred) {
  const Dev d = dev_of(x);
  k_sort_small<<<1, kSortSmallBlock, 0, x.stream>>>(d, pred ? pred : d.predP, x.ncell);
  return hipGetLastError();
}

hipError_t launch_scatter_rank(Ctx& x, const float4* pred) {
  const Dev d = dev_of(x);
  k_scatter<<<blocks_for(x.n), kBlock, 0, x.stream>>>(d);
  k_rank<<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, pred ? pred : d.predP);
  return hipGetLastError();
}

// Lanes per particle of list pass p: 64 (a wave: k_*_wide), 4 (a quad:
// k_*_quad, single scenes) or 1.
static int lanes(const Ctx& x, ListPass p) {
  const PassRule& r = x.rules[p];
  if (x.n <= r.wave_max) return 64;
  // Quad kernels index 4n threads in 32 bits.
  if (x.nscenes == 1 && x.n <= r.quad_max && x.n < (int64_t{1} << 30)) return 4;
  return 1;
}
static bool wide(const Ctx& x, ListPass p) { return lanes(x, p) == 64; }
static bool quad(const Ctx& x, ListPass p) { return lanes(x, p) == 4; }
static unsigned wide_grid(const Ctx& x) { return (unsigned)x.n; }
static unsigned quad_grid(const Ctx& x) { return blocks_for(x.n * 4); }

// A per-lane pass over the particles, split with the dense queue when x
// has one: the per-lane workgroups, then as many quad workgroups.
static Dev split_dev(const Ctx& x, ListPass p, unsigned* grid) {
  Dev d = dev_of(x);
  *grid = blocks_for(x.n);
  if (d.dense && lanes(x, p) == 1) {
    d.nlane = *grid;
    *grid *= 2u;
  }
  return d;
}

hipError_t launch_nbr_build(Ctx& x, const SimConst& c) {
  const Dev d = dev_of(x);
  if (quad(x, kPassBuild)) {
    k_nbr_quad<<<quad_grid(x), kBlock, 0, x.stream>>>(d, c.h2, x.grids[0]);
    return hipGetLastError();
  }
  if (multi(x))
    k_nbr_build<true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c.h2, x.grids[0]);
  else
    k_nbr_build<false><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c.h2, x.grids[0]);
  return hipGetLastError();
}

hipError_t launch_nbr_lambda(Ctx& x, const SimConst& c, float4* out) {
  const Dev d = dev_of(x);
  if (quad(x, kPassBuild)) {
    // The quad build, then the first lambda pass over the finished lists.
    k_nbr_quad<<<quad_grid(x), kBlock, 0, x.stream>>>(d, c.h2, x.grids[0]);
    k_lambda_quad<false><<<quad_grid(x), kBlock, 0, x.stream>>>(d, c, x.d.predQ, out);
    return hipGetLastError();
  }
  if (wide(x, kPassBuild)) {
    if (multi(x))
      k_nbr_lambda_wide<true><<<wide_grid(x), kWide, 0, x.stream>>>(d, c, x.grids[0], out);
    else
      k_nbr_lambda_wide<false><<<wide_grid(x), kWide, 0, x.stream>>>(d, c, x.grids[0], out);
    return hipGetLastError();
  }
  if (multi(x)) {
    k_nbr_lambda<true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], out);
  } else if (d.dense_b) {
    Dev ds = d;  // build split: the per-lane workgroups, then as many quad workgroups
    ds.nlane = blocks_for(x.n);
    k_nbr_lambda<false><<<2u * ds.nlane, kBlock, 0, x.stream>>>(ds, c, x.grids[0], out);
  } else {
    k_nbr_lambda<false><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], out);
  }
  return hipGetLastError();
}

hipError_t launch_lambda(Ctx& x, const SimConst& c, const float4* in, float4* out, bool packed) {
  if (quad(x, kPassLambda)) {
    if (packed)
      k_lambda_quad<true><<<quad_grid(x), kBlock, 0, x.stream>>>(dev_of(x), c, in, out);
    else
      k_lambda_quad<false><<<quad_grid(x), kBlock, 0, x.stream>>>(dev_of(x), c, in, out);
    return hipGetLastError();
  }
  if (wide(x, kPassLambda)) {
    if (packed)
      k_lambda_wide<true><<<wide_grid(x), kWide, 0, x.stream>>>(dev_of(x), c, in, out);
    else
      k_lambda_wide<false><<<wide_grid(x), kWide, 0, x.stream>>>(dev_of(x), c, in, out);
    return hipGetLastError();
  }
  unsigned ng = 0;
  const Dev d = split_dev(x, kPassLambda, &ng);
  if (packed)
    k_lambda<true><<<ng, kBlock, 0, x.stream>>>(d, c, in, out);
  else
    k_lambda<false><<<ng, kBlock, 0, x.stream>>>(d, c, in, out);
  return hipGetLastError();
}

template <int NT>
static void launch_deltap_nt(const Ctx& x, const Dev& d, const SimConst& c, const float4* in,
                             float4* out, bool last) {
  const unsigned nb = blocks_for(x.n);
  if (quad(x, kPassDeltap)) {
    if (last)
      k_deltap_quad<true, NT><<<quad_grid(x), kBlock, 0, x.stream>>>(d, c, x.grids[0], in);
    else
      k_deltap_quad<false, NT><<<quad_grid(x), kBlock, 0, x.stream>>>(d, c, x.grids[0], in);
    return;
  }
  if (wide(x, kPassDeltap)) {
    const unsigned ng = wide_grid(x);
    if (multi(x)) {
      if (last)
        k_deltap_wide<true, true, NT><<<ng, kWide, 0, x.stream>>>(d, c, x.grids[0], in);
      else
        k_deltap_wide<true, false, NT><<<ng, kWide, 0, x.stream>>>(d, c, x.grids[0], in);
    } else {
      if (last)
        k_deltap_wide<false, true, NT><<<ng, kWide, 0, x.stream>>>(d, c, x.grids[0], in);
      else
        k_deltap_wide<false, false, NT><<<ng, kWide, 0, x.stream>>>(d, c, x.grids[0], in);
    }
    return;
  }
  if (multi(x)) {
    if (last)
      k_deltap<true, true, NT><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0], in, out);
    else
      k_deltap<true, false, NT><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0], in, out);
  } else {
    unsigned ng = 0;
    const Dev ds = split_dev(x, kPassDeltap, &ng);
    if (last)
      k_deltap<false, true, NT><<<ng, kBlock, 0, x.stream>>>(ds, c, x.grids[0], in, out);
    else
      k_deltap<false, false, NT><<<ng, kBlock, 0, x.stream>>>(ds, c, x.grids[0], in, out);
  }
}

hipError_t launch_deltap(Ctx& x, const SimConst& c, const float4* in, float4* out, bool last) {
  const Dev d = dev_of(x);
  if (c.n_tensile == 4)
    launch_deltap_nt<4>(x, d, c, in, out, last);
  else
    launch_deltap_nt<-1>(x, d, c, in, out, last);
  return hipGetLastError();
}

hipError_t launch_xsph(Ctx& x, const SimConst& c, float4* out) {
  if (quad(x, kPassXsph)) {
    k_xsph_quad<<<quad_grid(x), kBlock, 0, x.stream>>>(dev_of(x), c, out);
    return hipGetLastError();
  }
  if (wide(x, kPassXsph)) {
    k_xsph_wide<<<wide_grid(x), kWide, 0, x.stream>>>(dev_of(x), c, out);
    return hipGetLastError();
  }
  unsigned ng = 0;
  const Dev d = split_dev(x, kPassXsph, &ng);
  k_xsph<<<ng, kBlock, 0, x.stream>>>(d, c, out);
  return hipGetLastError();
}

hipError_t launch_unpack_xs3(Ctx& x, float4* pos) {
  k_unpack_xs3<<<blocks_for(x.n), kBlock, 0, x.stream>>>(dev_of(x), pos);
  return hipGetLastError();
}

hipError_t launch_unpack_record(Ctx& x, float4* pos, float4* vel) {
  k_unpack_record<<<blocks_for(x.n), kBlock, 0, x.stream>>>(dev_of(x), pos, vel);
  return hipGetLastError();
}

hipError_t launch_vort_commit_predict(Ctx& x, const SimConst& c, const float4* in) {
  // Both sorts leave the histogram zero (the single-workgroup sort and the
  // general scan clear it once read), so the next prediction needs no fill.
  const bool small = x.small_sort && sort_in_one_workgroup(x);
  const Dev d = dev_of(x);
  if (wide(x, kPassVort) && !small) {
    // Beyond the smallest states the wide pass commits and the prediction
    // follows as its own launch (the serial tail of one lane per particle,
    // with one histogram atomic each, costs more than the launch).  It is
    // written where the fused prediction would be.
    k_vort_wide<false, false><<<wide_grid(x), kWide, 0, x.stream>>>(d, c, x.grids[0], in);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (multi(x))
      k_predict_into<true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], d.vort);
    else
      k_predict_into<false><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], d.vort);
    return hipGetLastError();
  }
  if (quad(x, kPassVort) && x.n <= kVortPredictQuadMax) {
    k_vort_quad<true><<<quad_grid(x), kBlock, 0, x.stream>>>(d, c, x.grids[0], in);
    return hipGetLastError();
  }
  if (wide(x, kPassVort)) {
    if (multi(x))
      k_vort_wide<true, true><<<wide_grid(x), kWide, 0, x.stream>>>(d, c, x.grids[0], in);
    else
      k_vort_wide<true, false><<<wide_grid(x), kWide, 0, x.stream>>>(d, c, x.grids[0], in);
    return hipGetLastError();
  }
  if (multi(x))
    k_vort_commit_predict<true><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], in);
  else
    k_vort_commit_predict<false><<<blocks_for(x.n), kBlock, 0, x.stream>>>(d, c, x.grids[0], in);
  return hipGetLastError();
}

hipError_t launch_vort_commit(Ctx& x, const SimConst& c, const float4* in) {
  if (quad(x, kPassVort)) {
    k_vort_quad<false><<<quad_grid(x), kBlock, 0, x.stream>>>(dev_of(x), c, x.grids[0], in);
    return hipGetLastError();
  }
  if (wide(x, kPassVort)) {
    k_vort_wide<false, false><<<wide_grid(x), kWide, 0, x.stream>>>(dev_of(x), c, x.grids[0], in);
    return hipGetLastError();
  }
  k_vort_commit<<<blocks_for(x.n), kBlock, 0, x.stream>>>(dev_of(x), c, in);
  return hipGetLastError();
}

hipError_t launch_persistent(Ctx& x, const SimConst& c, int nsteps, int block, int grid) {
  const Dev d = dev_of(x);
  uint32_t* queue = x.d.flags + 8;
  hipError_t e = launch_fill_u32(queue, 1, 0u, x.stream);
  if (e != hipSuccess) return e;
  const uint32_t nr = (uint32_t)x.nscenes;
  switch (block) {
    case 1024:
      k_persistent<1024><<<grid, 1024, 0, x.stream>>>(d, c, x.d.runs, nr, queue, x.p.iterations,
                                                      nsteps);
      break;
    case 512:
      k_persistent<512><<<grid, 512, 0, x.stream>>>(d, c, x.d.runs, nr, queue, x.p.iterations,
                                                    nsteps);
      break;
    default:
      k_persistent<256><<<grid, 256, 0, x.stream>>>(d, c, x.d.runs, nr, queue, x.p.iterations,
                                                    nsteps);
      break;
  }
  return hipGetLastError();
}

int persistent_max_blocks_per_cu(int block) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (block) {
    case 1024:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_persistent<1024>, 1024, 0);
      break;
    case 512:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_persistent<512>, 512, 0);
      break;
    default:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_persistent<256>, 256, 0);
      break;
  }
  return (e == hipSuccess && nb > 0) ? nb : 1;
}

hipError_t launch_unpermute(Ctx& x, const float4* ids, const float4* src, int first, int count,
                            float* dst) {
  k_unpermute<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, ids, src, first, count,
                                                        dst);
  return hipGetLastError();
}

hipError_t launch_unpermute_u32(Ctx& x, const float4* ids, const uint32_t* src, uint32_t* dst) {
  k_unpermute_u32<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, ids, src, dst);
  return hipGetLastError();
}

hipError_t launch_nbr_export(Ctx& x, uint32_t* counts, uint32_t* ids, int stride) {
  k_nbr_export<<<blocks_for(x.n), kBlock, 0, x.stream>>>(dev_of(x), counts, ids, stride);
  return hipGetLastError();
}

hipError_t launch_nbr_stats(Ctx& x, unsigned long long* sum, uint32_t* mx) {
  unsigned nb = blocks_for(x.n);
  if (nb > 1024) nb = 1024;
  k_nbr_stats<<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, x.d.ncnt, sum, mx);
  return hipGetLastError();
}

// Load-time density: hash the committed positions (no move), sort, then sum
// poly6 over the 27 cells.  Leaves the state re-sorted (posB/velB hold it).
hipError_t launch_sort_committed(Ctx& x, const SimConst& c) {
  hipError_t e = launch_fill_u32(x.d.cellCount, x.ncell, 0u, x.stream);
  if (e != hipSuccess) return e;
  const unsigned nb = blocks_for(x.n);
  Dev d = dev_of(x);
  if (multi(x))
    k_predict<true, false><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0]);
  else
    k_predict<false, false><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0]);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(x)) != hipSuccess) return e;
  return launch_scatter_rank(x);
}

hipError_t launch_density_field(Ctx& x, const SimConst& c, int scene, const float* pts,
                                const DensityLattice* lattice, uint32_t m, float* out) {
  if (m == 0) return hipSuccess;
  const Dev d = dev_of(x);
  const SceneGrid& g = x.grids[(size_t)scene];
  const unsigned nb = (unsigned)((m + kBlock - 1) / kBlock);
  if (lattice)
    k_density_field<true><<<nb, kBlock, 0, x.stream>>>(d, c, g, nullptr, *lattice, m, out);
  else
    k_density_field<false><<<nb, kBlock, 0, x.stream>>>(d, c, g, pts, DensityLattice{}, m, out);
  return hipGetLastError();
}

hipError_t launch_densities(Ctx& x, const SimConst& c, float* out) {
  hipError_t e = launch_sort_committed(x, c);
  if (e != hipSuccess) return e;
  const unsigned nb = blocks_for(x.n);
  Dev d = dev_of(x);
  if (multi(x))
    k_densities<true><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0], out);
  else
    k_densities<false><<<nb, kBlock, 0, x.stream>>>(d, c, x.grids[0], out);
  return hipGetLastError();
}

}  // namespace pbf
