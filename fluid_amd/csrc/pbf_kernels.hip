// pbf_kernels.hip -- gfx950 kernels of the PBF step and their launchers.
//
// One step (the reference's Particles::timeStep, src/particles.cpp:250-297):
//   k_predict        gravity, x* = x, clamp_response, cell hash, histogram
//   k_scan_*         wave64 exclusive scan of the cell histogram
//   k_scatter_ids    counting-sort scatter of particle ids by cell
//   k_rank_reorder   deterministic in-cell rank by id, gather into sorted order
//   k_nbr_build      27-cell search -> tiled ELL neighbour lists (fixed per step)
//   k_lambda         density / lambda pass            (x S)
//   k_deltap         delta-p + clamp pass              (x S; last one also v)
//   k_xsph           XSPH viscosity, vorticity, density
//   k_vort_commit    vorticity confinement, x <- x*
// Particles stay in the order (cell, id), so the lists of a particle are in
// ascending sorted index, i.e. the same order a CPU visits the 27 cells in.
#include <hip/hip_runtime.h>

#include "pbf_internal.h"

namespace pbf {

namespace {

constexpr int kBlock = 256;
// Neighbour loads in flight per group (build-time tunables).
#ifndef PBF_UNROLL
#define PBF_UNROLL 4
#endif
#ifndef PBF_UNROLL_XSPH
#define PBF_UNROLL_XSPH 8
#endif
constexpr int kUnroll = PBF_UNROLL;           // lambda, delta-p, vorticity
constexpr int kUnrollXsph = PBF_UNROLL_XSPH;  // XSPH gathers two float4 per neighbour

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

// Block-local exclusive prefix of a 1024-element tile; tile totals to partials.
__global__ __launch_bounds__(kBlock) void k_scan_block(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, uint32_t n,
                                                       uint32_t* __restrict__ partials) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  if (base + kScanItems <= n) {
    uint4 q = *reinterpret_cast<const uint4*>(in + base);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++) v[k] = (base + k < n) ? in[base + k] : 0u;
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

// ---------------------------------------------------------- grid build --
template <bool MULTI>
__device__ __forceinline__ SceneGrid grid_of(uint32_t i, const SceneGrid& g0,
                                             const SceneGrid* __restrict__ grids,
                                             const uint32_t* __restrict__ slotScene) {
  if (MULTI) return grids[slotScene[i]];
  return g0;
}

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
template <bool MULTI, bool MOVE>
__global__ __launch_bounds__(kBlock) void k_predict(
    uint32_t n, const float4* __restrict__ pos, float4* __restrict__ vel, float4* __restrict__ pred,
    uint32_t* __restrict__ cell, uint32_t* __restrict__ arrival, uint32_t* __restrict__ cellCount,
    SimConst c, SceneGrid g0, const SceneGrid* __restrict__ grids,
    const uint32_t* __restrict__ slotScene) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const SceneGrid g = grid_of<MULTI>(i, g0, grids, slotScene);
  const float4 x = pos[i];
  float3 p = make_float3(x.x, x.y, x.z);
  if (MOVE) {
    float4 v = vel[i];
    v.y = v.y - c.gdt;
    p = clamp_response(g, c, p, make_float3(v.x, v.y, v.z));
    vel[i] = v;
  }
  pred[i] = make_float4(p.x, p.y, p.z, 0.0f);
  const uint32_t cl = cell_lin(g, p);
  cell[i] = cl;
  arrival[i] = atomicAdd(&cellCount[cl], 1u);
}

__global__ __launch_bounds__(kBlock) void k_scatter_ids(uint32_t n, const uint32_t* __restrict__ cell,
                                                        const uint32_t* __restrict__ arrival,
                                                        const uint32_t* __restrict__ cellStart,
                                                        const float4* __restrict__ pos,
                                                        uint32_t* __restrict__ tmpId) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t slot = cellStart[cell[i]] + arrival[i];
  if (slot < n) tmpId[slot] = id_of(pos[i]);
}

// Rank of a particle inside its cell = number of cell members with a smaller
// id; the gather target is cellStart + rank, which makes the order (cell, id)
// independent of atomic arrival order.
__global__ __launch_bounds__(kBlock) void k_rank_reorder(
    uint32_t n, const uint32_t* __restrict__ cell, const uint32_t* __restrict__ cellStart,
    const uint32_t* __restrict__ tmpId, const float4* __restrict__ posA,
    const float4* __restrict__ velA, const float4* __restrict__ predP, float4* __restrict__ posB,
    float4* __restrict__ velB, float4* __restrict__ predQ, uint32_t* __restrict__ cellSorted) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = cell[i];
  const uint32_t s = cellStart[c];
  const uint32_t e = min(cellStart[c + 1], n);
  const float4 x = posA[i];
  const uint32_t id = id_of(x);
  uint32_t rank = 0;
  for (uint32_t t = s; t < e; t++) rank += (tmpId[t] < id) ? 1u : 0u;
  const uint32_t dst = s + rank;
  if (dst >= n) return;
  posB[dst] = x;
  velB[dst] = velA[i];
  predQ[dst] = predP[i];
  cellSorted[dst] = c;
}

// Neighbour sets of src/particles.cpp:258-265 (|x*_i - x*_j|^2 <= H2, j != i),
// found in the 27 cells around i.  For each (dz, dy) the three cells dx =
// -1..1 are contiguous in sorted order, so the search walks 9 index ranges in
// ascending order.  Found indices collect in a 4-entry register chunk that
// is stored whole (one 16-byte store) when full.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_nbr_build(
    uint32_t n, const float4* __restrict__ pred, const uint32_t* __restrict__ cellStart,
    uint32_t* __restrict__ nbr, uint32_t* __restrict__ ncnt, uint32_t* __restrict__ flags, int K4,
    float h2, SceneGrid g0, const SceneGrid* __restrict__ grids,
    const uint32_t* __restrict__ slotScene) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const SceneGrid g = grid_of<MULTI>(i, g0, grids, slotScene);
  const float4 xi = pred[i];
  int cc[3];
  cell_lin(g, make_float3(xi.x, xi.y, xi.z), cc);
  const int x0 = cc[0] > 0 ? cc[0] - 1 : 0;
  const int x1 = cc[0] + 1 < g.dims[0] ? cc[0] + 1 : g.dims[0] - 1;
  uint32_t cnt = 0;
  uint32_t buf[kChunk];
#pragma unroll
  for (int u = 0; u < kChunk; u++) buf[u] = 0u;
  auto flush = [&](int chunk) {
    if (chunk < K4)
      *reinterpret_cast<uint4*>(nbr + nbr_chunk(i, chunk, K4)) =
          make_uint4(buf[0], buf[1], buf[2], buf[3]);
  };
  auto test = [&](uint32_t j, const float4& xj) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float d2 = rx * rx + ry * ry + rz * rz;
    if (d2 <= h2 && j != i) {
      const uint32_t slot = cnt & 3u;
#pragma unroll
      for (int u = 0; u < kChunk; u++) buf[u] = (slot == (uint32_t)u) ? j : buf[u];
      cnt++;
      if ((cnt & 3u) == 0u) flush((int)(cnt >> 2) - 1);
    }
  };
  for (int dz = -1; dz <= 1; dz++) {
    const int z = cc[2] + dz;
    if (z < 0 || z >= g.dims[2]) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int y = cc[1] + dy;
      if (y < 0 || y >= g.dims[1]) continue;
      const uint32_t row =
          g.cell_base + ((uint32_t)z * (uint32_t)g.dims[1] + (uint32_t)y) * (uint32_t)g.dims[0];
      const uint32_t s = cellStart[row + (uint32_t)x0];
      const uint32_t e = min(cellStart[row + (uint32_t)x1 + 1u], n);
      uint32_t j = s;
      for (; j + 4 <= e; j += 4) {
        const float4 a = pred[j], b = pred[j + 1], c4 = pred[j + 2], d = pred[j + 3];
        test(j, a);
        test(j + 1, b);
        test(j + 2, c4);
        test(j + 3, d);
      }
      for (; j < e; j++) test(j, pred[j]);
    }
  }
  if (cnt & 3u) flush((int)(cnt >> 2));
  const uint32_t cap = (uint32_t)K4 * kChunk;
  if (cnt > cap) {
    atomicMax(&flags[0], cnt);
    cnt = cap;
  }
  ncnt[i] = cnt;
}

// ------------------------------------------------------------ solver ----
// newtonStepCalculateLambda (src/particles.cpp:185-204).  out = (x*, lambda).
__global__ __launch_bounds__(kBlock) void k_lambda(uint32_t n, const float4* __restrict__ in,
                                                   const uint32_t* __restrict__ nbr,
                                                   const uint32_t* __restrict__ ncnt, int K4,
                                                   SimConst c, float4* __restrict__ out) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = in[i];
  const int cnt = (int)ncnt[i];
  float dens = 0.0f, denom = 0.0f, Gx = 0.0f, Gy = 0.0f, Gz = 0.0f;
  for_each_neighbor<kUnroll>(
      i, cnt, K4, nbr,[&](uint32_t j) { return in[j]; },
      [&](uint32_t, const float4& xj) {
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float w = poly6_sel(c, r2);
        float3 gr = spiky_sel(c, rx, ry, rz, r2);
        dens = dens + w;
        gr.x = gr.x * c.inv_rho0;
        gr.y = gr.y * c.inv_rho0;
        gr.z = gr.z * c.inv_rho0;
        Gx = Gx + gr.x;
        Gy = Gy + gr.y;
        Gz = Gz + gr.z;
        denom = denom + (gr.x * gr.x + gr.y * gr.y + gr.z * gr.z);
      });
  const float ci = dens / c.rho0 - 1.0f;
  denom = denom + (Gx * Gx + Gy * Gy + Gz * Gz);
  const float lambda = -ci / (denom + c.eps);
  out[i] = make_float4(xi.x, xi.y, xi.z, lambda);
}

// newtonStepUpdatePosition (src/particles.cpp:206-213) with s_corr (:195)
// and clamp (:51-84).  in = (x*, lambda), out = (x*', 0).  The last
// iteration also forms v = (x* - x) / dt (updateVelocity, :215-217).
template <bool MULTI, bool LAST>
__global__ __launch_bounds__(kBlock) void k_deltap(
    uint32_t n, const float4* __restrict__ in, const uint32_t* __restrict__ nbr,
    const uint32_t* __restrict__ ncnt, int K4, SimConst c, SceneGrid g0,
    const SceneGrid* __restrict__ grids, const uint32_t* __restrict__ slotScene,
    float4* __restrict__ out, const float4* __restrict__ pos, float4* __restrict__ vel) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = in[i];
  const int cnt = (int)ncnt[i];
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  for_each_neighbor<kUnroll>(
      i, cnt, K4, nbr,[&](uint32_t j) { return in[j]; },
      [&](uint32_t, const float4& xj) {
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float w = poly6_sel(c, r2);
        const float s = c.negk * ipow(w * c.S, c.n_tensile);
        const float3 gr = spiky_sel(c, rx, ry, rz, r2);
        const float sc = (xi.w + xj.w) + s;
        dx = dx + sc * gr.x;
        dy = dy + sc * gr.y;
        dz = dz + sc * gr.z;
      });
  dx = dx * c.inv_rho0;
  dy = dy * c.inv_rho0;
  dz = dz * c.inv_rho0;
  const SceneGrid g = grid_of<MULTI>(i, g0, grids, slotScene);
  const float3 p = clamp_dp(g, c, make_float3(xi.x, xi.y, xi.z), make_float3(dx, dy, dz));
  out[i] = make_float4(p.x, p.y, p.z, 0.0f);
  if (LAST) {
    const float4 x = pos[i];
    vel[i] = make_float4(c.inv_dt * (p.x - x.x), c.inv_dt * (p.y - x.y), c.inv_dt * (p.z - x.z),
                         0.0f);
  }
}

// calculateVorticityApplyXSPHViscosity (src/particles.cpp:219-234), Jacobi:
// every neighbour velocity is the pre-XSPH v.  out = (x*, |omega|),
// vx = (v + C*visc, density), vort = (omega, 0).
__global__ __launch_bounds__(kBlock) void k_xsph(uint32_t n, const float4* __restrict__ in,
                                                 const float4* __restrict__ vel,
                                                 const uint32_t* __restrict__ nbr,
                                                 const uint32_t* __restrict__ ncnt, int K4,
                                                 SimConst c, float4* __restrict__ out,
                                                 float4* __restrict__ vx,
                                                 float4* __restrict__ vort) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = in[i];
  const float4 vi = vel[i];
  const int cnt = (int)ncnt[i];
  float ox = 0.0f, oy = 0.0f, oz = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, dens = 0.0f;
  struct XV {
    float4 x, v;
  };
  for_each_neighbor<kUnrollXsph>(
      i, cnt, K4, nbr,[&](uint32_t j) { return XV{in[j], vel[j]}; },
      [&](uint32_t, const XV& d) {
        const float u0 = d.v.x - vi.x, u1 = d.v.y - vi.y, u2 = d.v.z - vi.z;
        const float rx = xi.x - d.x.x, ry = xi.y - d.x.y, rz = xi.z - d.x.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float3 gr = spiky_sel(c, rx, ry, rz, r2);
        ox = ox + (u1 * gr.z - u2 * gr.y);
        oy = oy + (u2 * gr.x - u0 * gr.z);
        oz = oz + (u0 * gr.y - u1 * gr.x);
        const float w = poly6_sel(c, r2);
        sx = sx + u0 * w;
        sy = sy + u1 * w;
        sz = sz + u2 * w;
        dens = dens + w;
      });
  const float wn = sqrtf(ox * ox + oy * oy + oz * oz);
  out[i] = make_float4(xi.x, xi.y, xi.z, wn);
  vx[i] = make_float4(vi.x + c.xc * sx, vi.y + c.xc * sy, vi.z + c.xc * sz, dens);
  vort[i] = make_float4(ox, oy, oz, 0.0f);
}

// applyVorticity + updatePosition (src/particles.cpp:236-248).
// in = (x*, |omega|).  Commits posA = (x*, id), velA = (v, density).
__global__ __launch_bounds__(kBlock) void k_vort_commit(
    uint32_t n, const float4* __restrict__ in, const float4* __restrict__ vort,
    const float4* __restrict__ vx, const float4* __restrict__ posB,
    const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ ncnt, int K4, SimConst c,
    float4* __restrict__ posA, float4* __restrict__ velA) {
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = in[i];
  const int cnt = (int)ncnt[i];
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  for_each_neighbor<kUnroll>(
      i, cnt, K4, nbr,[&](uint32_t j) { return in[j]; },
      [&](uint32_t, const float4& xj) {
        const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float3 gr = spiky_sel(c, rx, ry, rz, r2);
        ex = ex + xj.w * gr.x;
        ey = ey + xj.w * gr.y;
        ez = ez + xj.w * gr.z;
      });
  float4 v = vx[i];
  const float e2 = ex * ex + ey * ey + ez * ez;
  if (sqrtf(e2) > c.eps_d) {
    const float rn = 1.0f / sqrtf(e2);
    const float u0 = rn * ex, u1 = rn * ey, u2 = rn * ez;
    const float4 w = vort[i];
    const float c0 = u1 * w.z - u2 * w.y;
    const float c1 = u2 * w.x - u0 * w.z;
    const float c2 = u0 * w.y - u1 * w.x;
    v.x = v.x + c.dtv * c0;
    v.y = v.y + c.dtv * c1;
    v.z = v.z + c.dtv * c2;
  }
  posA[i] = make_float4(xi.x, xi.y, xi.z, posB[i].w);
  velA[i] = v;
}

// Load-time density (src/particles.cpp:440-444 -> :158-163): every particle
// within the 27 cells, self included, r = x_j - x_i.  Written by id.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_densities(
    uint32_t n, const float4* __restrict__ pred, const float4* __restrict__ posB,
    const uint32_t* __restrict__ cellStart, SimConst c, SceneGrid g0,
    const SceneGrid* __restrict__ grids, const uint32_t* __restrict__ slotScene,
    float* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const SceneGrid g = grid_of<MULTI>(i, g0, grids, slotScene);
  const float4 xi = pred[i];
  int cc[3];
  cell_lin(g, make_float3(xi.x, xi.y, xi.z), cc);
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
      const uint32_t s = cellStart[row + (uint32_t)x0];
      const uint32_t e = min(cellStart[row + (uint32_t)x1 + 1u], n);
      for (uint32_t j = s; j < e; j++) {
        const float4 xj = pred[j];
        const float rx = xj.x - xi.x, ry = xj.y - xi.y, rz = xj.z - xi.z;
        dens = dens + poly6(c, rx * rx + ry * ry + rz * rz);
      }
    }
  }
  out[id_of(posB[i])] = dens;
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

__global__ __launch_bounds__(kBlock) void k_nbr_export(uint32_t n, const float4* __restrict__ ids,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ ncnt, int K4,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ out, int stride) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = id_of(ids[i]);
  const int cnt = (int)ncnt[i];
  counts[id] = (uint32_t)cnt;
  const int m = cnt < stride ? cnt : stride;
  for (int k = 0; k < m; k++) out[(size_t)id * stride + k] = id_of(ids[nbr[nbr_slot(i, k, K4)]]);
}

// Sum and maximum of the neighbour counts (out[0..1] = sum lo/hi, out[2] = max).
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

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ------------------------------------------------------------ launchers --
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + kBlock - 1) / kBlock;
  unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
  k_fill_u32<<<grid, kBlock, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials,
                           hipStream_t s) {
  const uint32_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb > 0) k_scan_block<<<nb, kBlock, 0, s>>>(in, out, n, partials);
  k_scan_partials<<<1, 1024, 0, s>>>(partials, nb, out, n);
  if (nb > 1) k_scan_add<<<blocks_for(n), kBlock, 0, s>>>(out, n, partials);
  return hipGetLastError();
}

static bool multi(const Ctx& x) { return x.nscenes > 1; }

hipError_t launch_predict(Ctx& x, const SimConst& c) {
  DevState& d = x.d;
  hipError_t e = launch_fill_u32(d.cellCount, x.ncell, 0u, x.stream);
  if (e != hipSuccess) return e;
  const unsigned nb = blocks_for(x.n);
  if (multi(x))
    k_predict<true, true><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.posA, d.velA, d.predP,
                                                       d.cell, d.arrival, d.cellCount, c,
                                                       x.grids[0], d.grids, d.slotScene);
  else
    k_predict<false, true><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.posA, d.velA, d.predP,
                                                        d.cell, d.arrival, d.cellCount, c,
                                                        x.grids[0], d.grids, d.slotScene);
  return hipGetLastError();
}

hipError_t launch_scan(Ctx& x) {
  return launch_scan_u32(x.d.cellCount, x.d.cellStart, x.ncell, x.d.blockSums, x.stream);
}

hipError_t launch_scatter_rank(Ctx& x) {
  DevState& d = x.d;
  const unsigned nb = blocks_for(x.n);
  k_scatter_ids<<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.cell, d.arrival, d.cellStart,
                                             d.posA, d.tmpId);
  k_rank_reorder<<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.cell, d.cellStart, d.tmpId,
                                              d.posA, d.velA, d.predP, d.posB, d.velB, d.predQ,
                                              d.cellSorted);
  return hipGetLastError();
}

hipError_t launch_nbr_build(Ctx& x, const SimConst& c) {
  DevState& d = x.d;
  const unsigned nb = blocks_for(x.n);
  if (multi(x))
    k_nbr_build<true><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.predQ, d.cellStart, d.nbr,
                                                   d.ncnt, d.flags, k4_of(x), c.h2,
                                                   x.grids[0], d.grids, d.slotScene);
  else
    k_nbr_build<false><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.predQ, d.cellStart, d.nbr,
                                                    d.ncnt, d.flags, k4_of(x), c.h2,
                                                    x.grids[0], d.grids, d.slotScene);
  return hipGetLastError();
}

hipError_t launch_lambda(Ctx& x, const SimConst& c, const float4* in, float4* out) {
  k_lambda<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, in, x.d.nbr, x.d.ncnt,
                                                     k4_of(x), c, out);
  return hipGetLastError();
}

hipError_t launch_deltap(Ctx& x, const SimConst& c, const float4* in, float4* out, bool last) {
  DevState& d = x.d;
  const unsigned nb = blocks_for(x.n);
  const uint32_t n = (uint32_t)x.n;
  const int K = k4_of(x);
  if (multi(x)) {
    if (last)
      k_deltap<true, true><<<nb, kBlock, 0, x.stream>>>(n, in, d.nbr, d.ncnt, K, c, x.grids[0],
                                                        d.grids, d.slotScene, out, d.posB, d.velB);
    else
      k_deltap<true, false><<<nb, kBlock, 0, x.stream>>>(n, in, d.nbr, d.ncnt, K, c, x.grids[0],
                                                         d.grids, d.slotScene, out, d.posB,
                                                         d.velB);
  } else {
    if (last)
      k_deltap<false, true><<<nb, kBlock, 0, x.stream>>>(n, in, d.nbr, d.ncnt, K, c, x.grids[0],
                                                         d.grids, d.slotScene, out, d.posB,
                                                         d.velB);
    else
      k_deltap<false, false><<<nb, kBlock, 0, x.stream>>>(n, in, d.nbr, d.ncnt, K, c, x.grids[0],
                                                          d.grids, d.slotScene, out, d.posB,
                                                          d.velB);
  }
  return hipGetLastError();
}

hipError_t launch_xsph(Ctx& x, const SimConst& c, const float4* in, float4* out) {
  DevState& d = x.d;
  k_xsph<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, in, d.velB, d.nbr, d.ncnt,
                                                   k4_of(x), c, out, d.vxsph, d.vort);
  return hipGetLastError();
}

hipError_t launch_vort_commit(Ctx& x, const SimConst& c, const float4* in) {
  DevState& d = x.d;
  k_vort_commit<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, in, d.vort, d.vxsph,
                                                          d.posB, d.nbr, d.ncnt,
                                                          k4_of(x), c, d.posA, d.velA);
  return hipGetLastError();
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
  k_nbr_export<<<blocks_for(x.n), kBlock, 0, x.stream>>>((uint32_t)x.n, x.d.posB, x.d.nbr,
                                                         x.d.ncnt, k4_of(x), counts, ids,
                                                         stride);
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
hipError_t launch_densities(Ctx& x, const SimConst& c, float* out) {
  DevState& d = x.d;
  hipError_t e = launch_fill_u32(d.cellCount, x.ncell, 0u, x.stream);
  if (e != hipSuccess) return e;
  const unsigned nb = blocks_for(x.n);
  if (multi(x))
    k_predict<true, false><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.posA, d.velA, d.predP,
                                                        d.cell, d.arrival, d.cellCount, c,
                                                        x.grids[0], d.grids, d.slotScene);
  else
    k_predict<false, false><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.posA, d.velA, d.predP,
                                                         d.cell, d.arrival, d.cellCount, c,
                                                         x.grids[0], d.grids, d.slotScene);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_scan(x)) != hipSuccess) return e;
  if ((e = launch_scatter_rank(x)) != hipSuccess) return e;
  if (multi(x))
    k_densities<true><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.predQ, d.posB, d.cellStart, c,
                                                   x.grids[0], d.grids, d.slotScene, out);
  else
    k_densities<false><<<nb, kBlock, 0, x.stream>>>((uint32_t)x.n, d.predQ, d.posB, d.cellStart,
                                                    c, x.grids[0], d.grids, d.slotScene, out);
  return hipGetLastError();
}

}  // namespace pbf
