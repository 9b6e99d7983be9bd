// pbf_adjoint.hip -- reverse-mode derivatives of the PBF step (differentiable
// extension; the reference has no counterpart).
//
// Forward step (src/particles.cpp:250-297), per particle i, neighbour sets
// N_i fixed from the predicted positions:
//   v1 = v - g dt y,  x*_0 = x + v1 dt                        (predict)
//   x*_{m+1} = x*_m + dp(x*_m),  m = 0..S-1                   (solver)
//     dp_i = (1/rho0) sum_j (l_i + l_j + s_ij) g_ij,  g_ij = gradW_spiky(x*_i - x*_j)
//     l_i = -C_i / (D_i + eps),  C_i = rho_i/rho0 - 1,  rho_i = sum_j W_ij
//     D_i = sum_j |g_ij/rho0|^2 + |G_i|^2,  G_i = sum_j g_ij/rho0
//     s_ij = -k (W_ij S)^n
//   v2 = (x*_S - x)/dt                                         (velocity)
//   v3_i = v2_i + c sum_j (v2_j - v2_i) W_ij                   (XSPH)
//   w_i = sum_j (v2_j - v2_i) x g_ij                           (vorticity)
//   eta_i = sum_j |w_j| g_ij,  N_i = eta_i/|eta_i|
//   v4_i = v3_i + dt eps_v (N_i x w_i),  x_new = x*_S,  v_new = v4
// The adjoint uses the analytic derivatives of these expressions: the spiky
// gradient g(r) = phi(|r|) r has Hessian H = phi I + (phi'/|r|) r r^T, the
// poly6 gradient is -6 c1 (h^2 - r^2)^2 r / h^9.  Neighbour relations are
// symmetric and g_ji = -g_ij, H_ji = H_ij, so every pair term is gathered by
// particle i from its own list (no atomics, deterministic).  Wall collision
// is differentiated as the identity; contacts are counted and reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "pbf_internal.h"

namespace pbf {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ float3 f3(const float4& a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) {
  return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ float3 operator-(float3 a, float3 b) {
  return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float3 operator*(float s, float3 a) {
  return make_float3(s * a.x, s * a.y, s * a.z);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 u, float3 v) {
  return make_float3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}

// Kernel values of one pair, r = x_i - x_j.
struct Pair {
  float3 r;
  float rl;     // |r|
  float W;      // poly6
  float3 gP;    // poly6 gradient
  float3 g;     // spiky gradient
  float phi;    // g = phi r
  float dphi;   // phi' / |r|  (H = phi I + dphi r r^T)
};

__device__ __forceinline__ Pair pair_of(const SimConst& c, float3 xi, float3 xj) {
  Pair p;
  p.r = xi - xj;
  const float r2 = dot3(p.r, p.r);
  p.rl = sqrtf(r2);
  if (r2 < c.h2) {
    const float t = c.h2 - r2;
    p.W = c.c1 * t * t * t / c.h9;
    p.gP = (-6.0f * c.c1 * t * t / c.h9) * p.r;
  } else {
    p.W = 0.0f;
    p.gP = make_float3(0.f, 0.f, 0.f);
  }
  if (p.rl < c.h && p.rl >= c.eps_d) {
    const float q = c.h - p.rl;
    p.phi = c.m3c2 * q * q / (c.h6 * p.rl);
    // phi'(rho) = -m3c2/h^6 (h - rho)(h + rho)/rho^2
    p.dphi = (-c.m3c2 / c.h6) * q * (c.h + p.rl) / (r2 * p.rl);
    p.g = p.phi * p.r;
  } else {
    p.phi = 0.0f;
    p.dphi = 0.0f;
    p.g = make_float3(0.f, 0.f, 0.f);
  }
  return p;
}

__device__ __forceinline__ float3 hess(const Pair& p, float3 v) {
  return p.phi * v + (p.dphi * dot3(p.r, v)) * p.r;
}

struct AdjDev {
  const float4* ids;    // posB: sorted x with ids in w
  const uint32_t* nbr;
  const uint32_t* ncnt;
  int K4;
  uint32_t n;
};

__device__ __forceinline__ uint32_t idx() { return blockIdx.x * kB + threadIdx.x; }

// The adjoint's neighbour passes can give each particle L consecutive lanes
// (L = 1, 2, 4): lane q of particle i visits the chunks ch = q, q + L, ...
// of its list (and the partial last chunk if it falls to q) and the L
// partial sums are added with a fixed butterfly, so the result is the same
// on every run.  Small scenes use L = 4 to put four times as many waves on
// the chip; a lane's sums then hold a quarter of the terms.
template <int L, class Body>
__device__ __forceinline__ void nbrs_split(const AdjDev& a, uint32_t i, int q, Body body) {
  const int cnt = (int)a.ncnt[i];
  const int nfull = cnt >> 2;
  for (int ch = q; ch < nfull; ch += L) {
    const uint4 e = *reinterpret_cast<const uint4*>(a.nbr + nbr_chunk(i, ch, a.K4));
    body(e.x);
    body(e.y);
    body(e.z);
    body(e.w);
  }
  if (q == nfull % L)
    for (int k = nfull * 4; k < cnt; k++) body(a.nbr[nbr_slot(i, k, a.K4)]);
}

// Sum over the L lanes of a particle (every lane gets the same value:
// floating-point addition is commutative).
template <int L>
__device__ __forceinline__ float lane_sum(float v) {
  for (int off = L / 2; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
  return v;
}
template <int L>
__device__ __forceinline__ float3 lane_sum(float3 v) {
  return make_float3(lane_sum<L>(v.x), lane_sum<L>(v.y), lane_sum<L>(v.z));
}

__global__ __launch_bounds__(kB) void k_load_state(uint32_t n, const float4* __restrict__ x,
                                                   const float4* __restrict__ v,
                                                   float4* __restrict__ posA,
                                                   float4* __restrict__ velA) {
  const uint32_t i = idx();
  if (i >= n) return;
  const float4 p = x[i];
  posA[i] = make_float4(p.x, p.y, p.z, __uint_as_float(i));
  const float4 q = v[i];
  velA[i] = make_float4(q.x, q.y, q.z, 0.0f);
}

// Sorted -> load order (dst[id] = src[i]) and back (dst[i] = src[id]).
__global__ __launch_bounds__(kB) void k_to_id(uint32_t n, const float4* __restrict__ ids,
                                              const float4* __restrict__ src,
                                              float4* __restrict__ dst) {
  const uint32_t i = idx();
  if (i < n) dst[__float_as_uint(ids[i].w)] = src[i];
}
// Two load-order arrays gathered into sorted order (ids: the sorted state's
// load ids) by one launch.
__global__ __launch_bounds__(kB) void k_from_id2(uint32_t n, const float4* __restrict__ ids,
                                                 const float4* __restrict__ src0,
                                                 float4* __restrict__ dst0,
                                                 const float4* __restrict__ src1,
                                                 float4* __restrict__ dst1) {
  const uint32_t i = idx();
  if (i >= n) return;
  const uint32_t j = __float_as_uint(ids[i].w);
  dst0[i] = src0[j];
  dst1[i] = src1[j];
}
// Committed state (any order, load ids in posA.w) against a load-order
// reference (rx, rv): counts the particles whose x or v differ in any bit.
__global__ __launch_bounds__(kB) void k_count_diff(uint32_t n, const float4* __restrict__ posA,
                                                   const float4* __restrict__ velA,
                                                   const float4* __restrict__ rx,
                                                   const float4* __restrict__ rv,
                                                   uint32_t* __restrict__ count) {
  const uint32_t i = idx();
  if (i >= n) return;
  const float4 p = posA[i], v = velA[i];
  const uint32_t id = __float_as_uint(p.w);
  const float4 a = rx[id], b = rv[id];
  const bool same = __float_as_uint(p.x) == __float_as_uint(a.x) &&
                    __float_as_uint(p.y) == __float_as_uint(a.y) &&
                    __float_as_uint(p.z) == __float_as_uint(a.z) &&
                    __float_as_uint(v.x) == __float_as_uint(b.x) &&
                    __float_as_uint(v.y) == __float_as_uint(b.y) &&
                    __float_as_uint(v.z) == __float_as_uint(b.z);
  if (!same) atomicAdd(count, 1u);
}
__global__ __launch_bounds__(kB) void k_copy4(uint32_t n, const float4* __restrict__ src,
                                              float4* __restrict__ dst) {
  const uint32_t i = idx();
  if (i < n) dst[i] = src[i];
}

// Vorticity confinement, part 1: eta_bar and the direct part of w_bar.
// xw = (x*_S, |w|), w = (w, .), vb4 = dL/dv_new.
template <int L>
__global__ __launch_bounds__(kB) void k_adj_vort_a(AdjDev a, SimConst c,
                                                   const float4* __restrict__ xw,
                                                   const float4* __restrict__ w,
                                                   const float4* __restrict__ vb4,
                                                   float4* __restrict__ etabar,
                                                   float4* __restrict__ wbar) {
  const uint32_t i = idx() / L;
  const int lq = (int)(idx() % L);
  if (i >= a.n) return;
  const float3 xi = f3(xw[i]);
  float3 eta = make_float3(0.f, 0.f, 0.f);
  nbrs_split<L>(a, i, lq, [&](uint32_t j) {
    const float4 q = xw[j];
    const Pair p = pair_of(c, xi, f3(q));
    eta = eta + q.w * p.g;
  });
  eta = lane_sum<L>(eta);
  if (lq != 0) return;
  const float e2 = dot3(eta, eta);
  float3 eb = make_float3(0.f, 0.f, 0.f), wb = make_float3(0.f, 0.f, 0.f);
  if (sqrtf(e2) > c.eps_d) {
    const float rn = 1.0f / sqrtf(e2);
    const float3 N = rn * eta;
    const float3 av = c.dtv * f3(vb4[i]);
    const float3 wi = f3(w[i]);
    const float3 Nb = cross3(wi, av);       // d(a.(N x w))/dN
    eb = rn * (Nb - dot3(N, Nb) * N);       // through N = eta/|eta|
    wb = cross3(av, N);                     // d(a.(N x w))/dw
  }
  etabar[i] = make_float4(eb.x, eb.y, eb.z, 0.f);
  wbar[i] = make_float4(wb.x, wb.y, wb.z, 0.f);
}

// Vorticity confinement, part 2: w_bar through |w_j| in eta_j, and x*_bar
// through g in eta: H_ij (|w_j| eta_bar_i - |w_i| eta_bar_j).  xs_acc starts
// from dL/dx_new (x_new = x*_S).
template <int L>
__global__ __launch_bounds__(kB) void k_adj_vort_b(AdjDev a, SimConst c,
                                                   const float4* __restrict__ xw,
                                                   const float4* __restrict__ w,
                                                   const float4* __restrict__ etabar,
                                                   float4* __restrict__ wbar,
                                                   const float4* __restrict__ xb_new,
                                                   float4* __restrict__ xs_acc) {
  const uint32_t i = idx() / L;
  const int lq = (int)(idx() % L);
  if (i >= a.n) return;
  const float4 xwi = xw[i];
  const float3 xi = f3(xwi);
  const float wn = xwi.w;
  const float3 ebi = f3(etabar[i]);
  float s = 0.0f;
  float3 xs = make_float3(0.f, 0.f, 0.f);
  nbrs_split<L>(a, i, lq, [&](uint32_t j) {
    const float4 q = xw[j];
    const Pair p = pair_of(c, xi, f3(q));
    const float3 ebj = f3(etabar[j]);
    s = s - dot3(ebj, p.g);
    xs = xs + hess(p, q.w * ebi - wn * ebj);
  });
  s = lane_sum<L>(s);
  xs = lane_sum<L>(xs);
  if (lq != 0) return;
  float3 wb = f3(wbar[i]);
  if (wn > 0.0f) wb = wb + (s / wn) * f3(w[i]);
  wbar[i] = make_float4(wb.x, wb.y, wb.z, 0.f);
  const float3 x0 = f3(xb_new[i]) + xs;
  xs_acc[i] = make_float4(x0.x, x0.y, x0.z, 0.f);
}

// XSPH and vorticity: v2_bar and x*_bar; then v2 = (x*_S - x)/dt.
// ybar = x*_S bar (all terms), xstart = -v2_bar/dt (the x part of v2).
template <int L>
__global__ __launch_bounds__(kB) void k_adj_xsph(AdjDev a, SimConst c,
                                                 const float4* __restrict__ xs,
                                                 const float4* __restrict__ v2,
                                                 const float4* __restrict__ vb3,
                                                 const float4* __restrict__ wbar,
                                                 const float4* __restrict__ xs_acc,
                                                 float4* __restrict__ ybar,
                                                 float4* __restrict__ xstart,
                                                 uint32_t* __restrict__ contacts, SceneGrid g) {
  const uint32_t i = idx() / L;
  const int lq = (int)(idx() % L);
  if (i >= a.n) return;
  const float3 xi = f3(xs[i]);
  const float3 vi = f3(v2[i]);
  const float3 bi = f3(vb3[i]);
  const float3 wi = f3(wbar[i]);
  float3 v2b = make_float3(0.f, 0.f, 0.f);
  float3 xb = make_float3(0.f, 0.f, 0.f);
  nbrs_split<L>(a, i, lq, [&](uint32_t j) {
    const Pair p = pair_of(c, xi, f3(xs[j]));
    const float3 u = f3(v2[j]) - vi;
    const float3 bj = f3(vb3[j]);
    const float3 ws = wi + f3(wbar[j]);
    v2b = v2b + (c.xc * p.W) * (bj - bi) - cross3(p.g, ws);
    xb = xb + (c.xc * dot3(u, bi - bj)) * p.gP + hess(p, cross3(ws, u));
  });
  v2b = bi + lane_sum<L>(v2b);
  xb = lane_sum<L>(xb);
  if (lq != 0) return;
  const float3 y = f3(xs_acc[i]) + xb + c.inv_dt * v2b;
  ybar[i] = make_float4(y.x, y.y, y.z, 0.f);
  const float3 xst = -c.inv_dt * v2b;
  xstart[i] = make_float4(xst.x, xst.y, xst.z, 0.f);
  if (xi.x <= g.lo_e[0] || xi.y <= g.lo_e[1] || xi.z <= g.lo_e[2] || xi.x >= g.hi_e[0] ||
      xi.y >= g.hi_e[1] || xi.z >= g.hi_e[2])
    atomicAdd(contacts, 1u);
}

// Solver iteration m, part 1: recompute rho, G, D, lambda at x*_m and
// lambda_bar = sum_j g_ij . (e_i - e_j), e = ybar / rho0; then
// C_bar = -lambda_bar/(D+eps), D_bar = -lambda_bar lambda/(D+eps),
// rho_bar = C_bar/rho0.  aux = (lambda, rho_bar, D_bar, .), Gb = (G, .).
template <int L>
__global__ __launch_bounds__(kB) void k_adj_lambda(AdjDev a, SimConst c,
                                                   const float4* __restrict__ xs,
                                                   const float4* __restrict__ ybar,
                                                   float4* __restrict__ aux,
                                                   float4* __restrict__ Gb,
                                                   uint32_t* __restrict__ contacts, SceneGrid g) {
  const uint32_t i = idx() / L;
  const int lq = (int)(idx() % L);
  if (i >= a.n) return;
  const float3 xi = f3(xs[i]);
  const float3 ei = c.inv_rho0 * f3(ybar[i]);
  float rho = 0.f, D = 0.f, lb = 0.f;
  float3 G = make_float3(0.f, 0.f, 0.f);
  nbrs_split<L>(a, i, lq, [&](uint32_t j) {
    const Pair p = pair_of(c, xi, f3(xs[j]));
    rho = rho + p.W;
    const float3 gs = c.inv_rho0 * p.g;
    G = G + gs;
    D = D + dot3(gs, gs);
    lb = lb + dot3(p.g, ei - c.inv_rho0 * f3(ybar[j]));
  });
  rho = lane_sum<L>(rho);
  G = lane_sum<L>(G);
  D = lane_sum<L>(D);
  lb = lane_sum<L>(lb);
  if (lq != 0) return;
  D = D + dot3(G, G);
  const float C = rho / c.rho0 - 1.0f;
  const float lam = -C / (D + c.eps);
  const float Cb = -lb / (D + c.eps);
  const float Db = -lb * lam / (D + c.eps);
  aux[i] = make_float4(lam, Cb * c.inv_rho0, Db, 0.f);
  Gb[i] = make_float4(G.x, G.y, G.z, 0.f);
  if (xi.x <= g.lo_e[0] || xi.y <= g.lo_e[1] || xi.z <= g.lo_e[2] || xi.x >= g.hi_e[0] ||
      xi.y >= g.hi_e[1] || xi.z >= g.hi_e[2])
    atomicAdd(contacts, 1u);
}

// Solver iteration m, part 2: x*_m bar = x*_{m+1} bar + (d dp / d x*_m)^T e.
// Per pair, with e_i = ybar_i/rho0, de = e_i - e_j:
//   H (l_i + l_j + s) de                        (g in dp)
//   + s'(W) (de . g) gP                          (s in dp)
//   + (rho_bar_i + rho_bar_j) gP                 (W in rho)
//   + H [2 (D_bar_i + D_bar_j) g/rho0^2 + 2 (D_bar_i G_i - D_bar_j G_j)/rho0]  (g in D)
// and dL/dk += sum_j (e_i . g_ij) (-(W S)^n).
template <int L>
__global__ __launch_bounds__(kB) void k_adj_pos(AdjDev a, SimConst c,
                                                const float4* __restrict__ xs,
                                                const float4* __restrict__ ybar,
                                                const float4* __restrict__ aux,
                                                const float4* __restrict__ Gb,
                                                float4* __restrict__ ynext,
                                                double* __restrict__ kpart) {
  const uint32_t i = idx() / L;
  const int lq = (int)(idx() % L);
  if (i >= a.n) return;
  const float3 xi = f3(xs[i]);
  const float3 ei = c.inv_rho0 * f3(ybar[i]);
  const float4 ai = aux[i];
  const float3 Gi = f3(Gb[i]);
  const float r02 = c.inv_rho0 * c.inv_rho0;
  float3 acc = make_float3(0.f, 0.f, 0.f);
  float kp = 0.f;
  nbrs_split<L>(a, i, lq, [&](uint32_t j) {
    const Pair p = pair_of(c, xi, f3(xs[j]));
    const float3 ej = c.inv_rho0 * f3(ybar[j]);
    const float4 aj = aux[j];
    const float3 de = ei - ej;
    const float ws = p.W * c.S;
    const float pw = ipow(ws, c.n_tensile);
    const float s = c.negk * pw;
    const float sd = c.n_tensile > 0 ? c.negk * (float)c.n_tensile * c.S * ipow(ws, c.n_tensile - 1)
                                     : 0.0f;
    const float3 hv = (ai.x + aj.x + s) * de + (2.0f * (ai.z + aj.z) * r02) * p.g +
                      (2.0f * c.inv_rho0) * (ai.z * Gi - aj.z * f3(Gb[j]));
    acc = acc + hess(p, hv) + (sd * dot3(de, p.g) + ai.y + aj.y) * p.gP;
    kp = kp + dot3(ei, p.g) * (-pw);
  });
  acc = lane_sum<L>(acc);
  kp = lane_sum<L>(kp);
  if (lq != 0) return;
  const float3 y = f3(ybar[i]) + acc;
  ynext[i] = make_float4(y.x, y.y, y.z, 0.f);
  kpart[i] = (double)kp;
}

// Predict: x*_0 = x + (v - g dt y) dt -> x_bar = x*_0 bar + xstart, v_bar = dt x*_0 bar.
__global__ __launch_bounds__(kB) void k_adj_predict(AdjDev a, SimConst c,
                                                    const float4* __restrict__ ybar,
                                                    const float4* __restrict__ xstart,
                                                    float4* __restrict__ xbar_id,
                                                    float4* __restrict__ vbar_id) {
  const uint32_t i = idx();
  if (i >= a.n) return;
  const uint32_t id = __float_as_uint(a.ids[i].w);
  const float3 y = f3(ybar[i]);
  const float3 xb = y + f3(xstart[i]);
  const float3 vb = c.dt * y;
  xbar_id[id] = make_float4(xb.x, xb.y, xb.z, 0.f);
  vbar_id[id] = make_float4(vb.x, vb.y, vb.z, 0.f);
}

// Deterministic sum of n doubles in two launches with a fixed order:
// k_sum_double_part gives workgroup b the sum of slice b of kSumParts equal
// slices, then k_sum_double adds the kSumParts partial sums.
constexpr int kSumParts = 64;
__global__ __launch_bounds__(kB) void k_sum_double_part(uint32_t n, const double* __restrict__ in,
                                                        double* __restrict__ part) {
  __shared__ double sh[kB];
  const uint32_t per = (n + kSumParts - 1) / kSumParts;
  const uint32_t b0 = blockIdx.x * per, b1 = min(n, b0 + per);
  double s = 0.0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += kB) s += in[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Deterministic sum of n doubles (one workgroup, fixed order).
__global__ __launch_bounds__(kB) void k_sum_double(uint32_t n, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  __shared__ double part[kB];
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += kB) s += in[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace
}  // namespace pbf

// ------------------------------------------------------------- host side --
// What the backward pass of one step reads of its forward pass, all in the
// step's sorted order: the sorted ids and state (posB, velB after the
// solve), x* after the solve (predQ) and after XSPH with |w| (predP), w
// (vort), the lists, and x*_m for m = 0..S (stash).
struct AdjTape {
  float4 *posB = nullptr, *predP = nullptr, *predQ = nullptr, *vort = nullptr, *velB = nullptr;
  float4* stash = nullptr;
  uint32_t *nbr = nullptr, *ncnt = nullptr;
};

struct pbf_adj {
  pbf_ctx* ctx = nullptr;  // plain context that runs the forward passes
  int nsteps = 0, every = 1, nckpt = 0;
  int64_t n = 0, npad = 0;
  float k = 0.0f;
  bool have_forward = false;
  float4 *ckx = nullptr, *ckv = nullptr;      // checkpoints, load order
  float4 *fx = nullptr, *fv = nullptr;        // final state of the forward pass, load order
  std::vector<AdjTape> tape;                  // one slot per step of a segment
  float4 *xbar = nullptr, *vbar = nullptr;    // running adjoints, load order
  float4 *yb = nullptr, *yb2 = nullptr, *xst = nullptr, *xsacc = nullptr, *vb3 = nullptr;
  float4 *etab = nullptr, *wb = nullptr, *aux = nullptr, *Gb = nullptr, *xbn = nullptr;
  double *kpart = nullptr, *ksteps = nullptr;
  double* kslice = nullptr;  // partial sums of the dL/dk reduction
  int lanes = 1;             // lanes per particle in the adjoint neighbour passes
  uint32_t* contacts = nullptr;
  double backward_ms = 0.0;
  int64_t last_contacts = 0;
  int64_t last_replay_diff = 0;  // particles whose replay differed from the forward pass
  std::string err;
};

namespace {

thread_local std::string g_adj_err = "no error";

int afail(pbf_adj* a, int code, const std::string& m) {
  if (a) a->err = m;
  g_adj_err = m;
  // Consume the runtime's last error of a failed call (see fail() in
  // pbf_capi.hip).
  if (code == PBF_ERR_HIP || code == PBF_ERR_NOMEM) (void)hipGetLastError();
  return code;
}

#define AHIP(a, call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return afail((a), PBF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

pbf::Ctx* X(pbf_adj* a) { return reinterpret_cast<pbf::Ctx*>(a->ctx); }

template <class T>
void afree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void free_all(pbf_adj* a) {
  afree(a->ckx); afree(a->ckv); afree(a->fx); afree(a->fv);
  for (AdjTape& t : a->tape) {
    afree(t.posB); afree(t.predP); afree(t.predQ); afree(t.vort); afree(t.velB);
    afree(t.stash); afree(t.nbr); afree(t.ncnt);
  }
  afree(a->xbar); afree(a->vbar); afree(a->yb); afree(a->yb2); afree(a->xst);
  afree(a->xsacc); afree(a->vb3); afree(a->etab); afree(a->wb); afree(a->aux);
  afree(a->Gb); afree(a->xbn); afree(a->kpart); afree(a->ksteps); afree(a->contacts);
  afree(a->kslice);
}

// Loads a load-order state (x, v) into the context as its committed state.
int load_state(pbf_adj* a, const float4* x, const float4* v) {
  pbf::Ctx* c = X(a);
  pbf::k_load_state<<<pbf::nblk(a->n), pbf::kB, 0, c->stream>>>((uint32_t)a->n, x, v, c->d.posA,
                                                                c->d.velA);
  AHIP(a, hipGetLastError());
  return PBF_OK;
}

// Saves the context's committed state in load order.
int save_state(pbf_adj* a, float4* x, float4* v) {
  pbf::Ctx* c = X(a);
  pbf::k_to_id<<<pbf::nblk(a->n), pbf::kB, 0, c->stream>>>((uint32_t)a->n, c->d.posA, c->d.posA, x);
  pbf::k_to_id<<<pbf::nblk(a->n), pbf::kB, 0, c->stream>>>((uint32_t)a->n, c->d.posA, c->d.velA, v);
  AHIP(a, hipGetLastError());
  return PBF_OK;
}

// One forward step from the context's committed state.  With a tape slot,
// the step's sorted state, lists and x*_m (m = 0..S) are written straight
// into the slot (the context's buffers are swapped for the slot's during
// the step) for the backward pass to read.
int forward_step(pbf_adj* a, const pbf::SimConst& k, AdjTape* slot) {
  pbf::Ctx* c = X(a);
  pbf::DevState& d = c->d;
  const uint32_t n = (uint32_t)a->n;
  const size_t np = (size_t)a->npad;
  const pbf::DevState own = d;
  if (slot) {
    d.posB = slot->posB;
    d.predP = slot->predP;
    d.predQ = slot->predQ;
    d.vort = slot->vort;
    d.velB = slot->velB;
    d.nbr = slot->nbr;
    d.ncnt = slot->ncnt;
  }
  auto run = [&]() -> int {
    AHIP(a, pbf::launch_predict(*c, k));
    AHIP(a, pbf::launch_scan(*c));
    AHIP(a, pbf::launch_scatter_rank(*c));
    // The list build and the first λ pass in one launch, as in the plain
    // step; neither changes predQ, so the stash of x* can follow.
    AHIP(a, pbf::launch_nbr_lambda(*c, k, d.predP));
    if (slot) pbf::k_copy4<<<pbf::nblk(n), pbf::kB, 0, c->stream>>>(n, d.predQ, slot->stash);
    for (int it = 0; it < c->p.iterations; it++) {
      const bool last = it == c->p.iterations - 1;
      if (it > 0) AHIP(a, pbf::launch_lambda(*c, k, d.predQ, d.predP, true));
      AHIP(a, pbf::launch_deltap(*c, k, d.predP, d.predQ, last));
      // The last iteration writes x* and v as XSPH records; the adjoint reads
      // them from predQ and velB.
      if (last) AHIP(a, pbf::launch_unpack_record(*c, d.predQ, d.velB));
      if (slot) {
        float4* st = slot->stash + (size_t)(it + 1) * np;
        if (last)
          pbf::k_copy4<<<pbf::nblk(n), pbf::kB, 0, c->stream>>>(n, d.predQ, st);
        else
          AHIP(a, pbf::launch_unpack_xs3(*c, st));  // x* of a pass before the last
      }
    }
    AHIP(a, pbf::launch_xsph(*c, k, d.predP));
    AHIP(a, pbf::launch_vort_commit(*c, k, d.predP));
    AHIP(a, hipGetLastError());
    return PBF_OK;
  };
  const int rc = run();
  d = own;
  return rc;
}

// Backward of step t, whose forward pass is recorded in tape slot T, with L
// lanes per particle in the neighbour passes.
// xbar/vbar (load order) hold dL/d(state after the step) on entry and
// dL/d(state before the step) on exit; ksteps[t] receives this step's dL/dk.
template <int L>
int backward_step_l(pbf_adj* a, const pbf::SimConst& k, int t, const AdjTape& T) {
  pbf::Ctx* c = X(a);
  const uint32_t n = (uint32_t)a->n;
  const size_t np = (size_t)a->npad;
  const AdjTape d = T;
  pbf::AdjDev ad{d.posB, d.nbr, d.ncnt, pbf::k4_of(*c), n};
  const unsigned nb = pbf::nblk(n), nbl = pbf::nblk((int64_t)n * L);
  hipStream_t s = c->stream;
  const pbf::SceneGrid g = c->grids[0];
  // dL/dv_new and dL/dx_new in sorted order.
  pbf::k_from_id2<<<nb, pbf::kB, 0, s>>>(n, d.posB, a->vbar, a->vb3, a->xbar, a->xbn);
  pbf::k_adj_vort_a<L><<<nbl, pbf::kB, 0, s>>>(ad, k, d.predP, d.vort, a->vb3, a->etab, a->wb);
  pbf::k_adj_vort_b<L><<<nbl, pbf::kB, 0, s>>>(ad, k, d.predP, d.vort, a->etab, a->wb, a->xbn,
                                              a->xsacc);
  pbf::k_adj_xsph<L><<<nbl, pbf::kB, 0, s>>>(ad, k, d.predQ, d.velB, a->vb3, a->wb, a->xsacc,
                                            a->yb, a->xst, a->contacts, g);
  AHIP(a, hipGetLastError());
  const int S = c->p.iterations;
  for (int it = S - 1; it >= 0; it--) {
    const float4* xs = d.stash + (size_t)it * np;
    pbf::k_adj_lambda<L><<<nbl, pbf::kB, 0, s>>>(ad, k, xs, a->yb, a->aux, a->Gb, a->contacts, g);
    pbf::k_adj_pos<L><<<nbl, pbf::kB, 0, s>>>(ad, k, xs, a->yb, a->aux, a->Gb, a->yb2,
                                             a->kpart + (size_t)it * np);
    AHIP(a, hipGetLastError());
    std::swap(a->yb, a->yb2);
  }
  // dL/dk of this step: one fixed-order sum over every iteration's partials.
  pbf::k_sum_double_part<<<pbf::kSumParts, pbf::kB, 0, s>>>((uint32_t)(np * (size_t)S), a->kpart,
                                                            a->kslice);
  pbf::k_sum_double<<<1, pbf::kB, 0, s>>>((uint32_t)pbf::kSumParts, a->kslice, a->ksteps + t);
  pbf::k_adj_predict<<<nb, pbf::kB, 0, s>>>(ad, k, a->yb, a->xst, a->xbar, a->vbar);
  AHIP(a, hipGetLastError());
  return PBF_OK;
}

int backward_step(pbf_adj* a, const pbf::SimConst& k, int t, const AdjTape& T) {
  switch (a->lanes) {
    case 4: return backward_step_l<4>(a, k, t, T);
    case 2: return backward_step_l<2>(a, k, t, T);
    default: return backward_step_l<1>(a, k, t, T);
  }
}

}  // namespace

extern "C" {

static int pbf_adj_create_impl(const pbf_params* p, int64_t n, int32_t nsteps, int32_t every,
                               pbf_adj** out) {
  if (!out) return afail(nullptr, PBF_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n < 1 || nsteps < 1 || every < 1)
    return afail(nullptr, PBF_ERR_INVALID, "n, nsteps and checkpoint_every must be >= 1");
  pbf_adj* a = new pbf_adj();
  int rc = pbf_create(p, &a->ctx);
  if (rc != PBF_OK) {
    std::string m = pbf_last_error(nullptr);
    delete a;
    return afail(nullptr, rc, m);
  }
  a->n = n;
  a->npad = (n + 63) / 64 * 64;
  // Four lanes per particle while that still leaves under ~4M threads
  // (small scenes are latency-bound); PBF_ADJ_LANES = 1, 2 or 4 overrides.
  a->lanes = n <= (1 << 20) ? 4 : 1;
  if (const char* e = getenv("PBF_ADJ_LANES")) {
    const int l = atoi(e);
    if (l == 1 || l == 2 || l == 4) a->lanes = l;
  }
  a->nsteps = nsteps;
  a->every = every;
  a->nckpt = (nsteps + every - 1) / every;
  const int S = p->iterations;
  const size_t np = (size_t)a->npad;
  hipError_t e = hipSuccess;
#define AA(ptr, cnt) \
  if (e == hipSuccess) e = hipMalloc((void**)&ptr, (cnt) * sizeof(*ptr))
  AA(a->ckx, np * a->nckpt); AA(a->ckv, np * a->nckpt); AA(a->fx, np); AA(a->fv, np);
  // The tape: one slot per step of a checkpoint segment.
  const size_t K = (size_t)pbf::k4_of(*X(a)) * pbf::kChunk;
  a->tape.resize((size_t)every);
  for (AdjTape& t : a->tape) {
    AA(t.posB, np); AA(t.predP, np + 64); AA(t.predQ, np + 64); AA(t.vort, np); AA(t.velB, np);
    AA(t.stash, np * (S + 1)); AA(t.nbr, np * K); AA(t.ncnt, np);
  }
  AA(a->xbar, np); AA(a->vbar, np); AA(a->yb, np); AA(a->yb2, np); AA(a->xst, np);
  AA(a->xsacc, np); AA(a->vb3, np); AA(a->etab, np); AA(a->wb, np); AA(a->aux, np);
  AA(a->Gb, np); AA(a->xbn, np);
  AA(a->kpart, np * (size_t)S); AA(a->ksteps, 2 * (size_t)nsteps); AA(a->contacts, 4);
  AA(a->kslice, (size_t)pbf::kSumParts);
  // The partials of the padding slots n..npad-1 are never written: zero
  // them once.
  if (e == hipSuccess) e = hipMemset(a->kpart, 0, np * (size_t)S * sizeof(double));
#undef AA
  if (e != hipSuccess) {
    free_all(a);
    pbf_destroy(a->ctx);
    delete a;
    return afail(nullptr, PBF_ERR_NOMEM, std::string("adjoint buffers: ") + hipGetErrorString(e));
  }
  *out = a;
  return PBF_OK;
}

PBF_API int pbf_adj_create(const pbf_params* p, int64_t n, int32_t nsteps, int32_t every,
                           pbf_adj** out) {
  try {
    return pbf_adj_create_impl(p, n, nsteps, every, out);
  } catch (const std::bad_alloc&) {
    return afail(nullptr, PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return afail(nullptr, PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API int pbf_adj_destroy(pbf_adj* a) {
  if (!a) return PBF_OK;
  pbf::Ctx* c = X(a);
  if (c && c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(a);
  pbf_destroy(a->ctx);
  delete a;
  return PBF_OK;
}

PBF_API const char* pbf_adj_last_error(const pbf_adj* a) {
  if (a && !a->err.empty()) return a->err.c_str();
  return g_adj_err.c_str();
}

static int pbf_adj_forward_impl(pbf_adj* a, const float* x0, const float* v0, float k_tensile,
                                float* xT, float* vT) {
  if (!a || !x0 || !v0) return afail(a, PBF_ERR_INVALID, "NULL argument");
  pbf::Ctx* c = X(a);
  c->p.k_tensile = k_tensile;
  a->k = k_tensile;
  int rc = pbf_load(a->ctx, x0, v0, a->n);
  if (rc != PBF_OK) return afail(a, rc, pbf_last_error(a->ctx));
  const pbf::SimConst k = pbf::consts_for(*c, c->p.dt);
  for (int t = 0; t < a->nsteps; t++) {
    if (t % a->every == 0) {
      const size_t o = (size_t)(t / a->every) * (size_t)a->npad;
      if ((rc = save_state(a, a->ckx + o, a->ckv + o)) != PBF_OK) return rc;
    }
    if ((rc = forward_step(a, k, nullptr)) != PBF_OK) return rc;
  }
  // The final state, against which the backward pass checks its replay of
  // the last segment.
  if ((rc = save_state(a, a->fx, a->fv)) != PBF_OK) return rc;
  rc = pbf_sync(a->ctx);
  if (rc != PBF_OK) return afail(a, rc, pbf_last_error(a->ctx));
  if (xT || vT) {
    rc = pbf_read(a->ctx, xT ? xT : nullptr, vT ? vT : nullptr, nullptr);
    if (rc != PBF_OK) return afail(a, rc, pbf_last_error(a->ctx));
  }
  a->have_forward = true;
  return PBF_OK;
}

PBF_API int pbf_adj_forward(pbf_adj* a, const float* x0, const float* v0, float k_tensile,
                            float* xT, float* vT) {
  try {
    return pbf_adj_forward_impl(a, x0, v0, k_tensile, xT, vT);
  } catch (const std::bad_alloc&) {
    return afail(a, PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return afail(a, PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_adj_backward_impl(pbf_adj* a, const float* xbar_T, const float* vbar_T, float* xbar0,
                                 float* vbar0, double* kbar) {
  if (!a || !xbar_T) return afail(a, PBF_ERR_INVALID, "NULL argument");
  if (!a->have_forward) return afail(a, PBF_ERR_STATE, "run pbf_adj_forward first");
  pbf::Ctx* c = X(a);
  hipStream_t s = c->stream;
  const size_t n = (size_t)a->n, np = (size_t)a->npad;
  std::vector<float4> hx(n), hv(n);
  for (size_t i = 0; i < n; i++) {
    hx[i] = make_float4(xbar_T[3 * i], xbar_T[3 * i + 1], xbar_T[3 * i + 2], 0.f);
    hv[i] = vbar_T ? make_float4(vbar_T[3 * i], vbar_T[3 * i + 1], vbar_T[3 * i + 2], 0.f)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  hipEvent_t e0, e1;
  AHIP(a, hipEventCreate(&e0));
  AHIP(a, hipEventCreate(&e1));
  AHIP(a, hipEventRecord(e0, s));
  AHIP(a, hipMemcpyAsync(a->xbar, hx.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
  AHIP(a, hipMemcpyAsync(a->vbar, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, s));
  AHIP(a, pbf::launch_fill_u32(a->contacts, 2, 0u, s));
  const pbf::SimConst k = pbf::consts_for(*c, c->p.dt);
  int rc = PBF_OK;
  for (int seg = a->nckpt - 1; seg >= 0; seg--) {
    const int t0 = seg * a->every;
    const int len = std::min(a->every, a->nsteps - t0);
    // Replay the segment from its checkpoint, recording every step.
    if ((rc = load_state(a, a->ckx + (size_t)seg * np, a->ckv + (size_t)seg * np)) != PBF_OK)
      return rc;
    for (int t = 0; t < len; t++)
      if ((rc = forward_step(a, k, &a->tape[(size_t)t])) != PBF_OK) return rc;
    // The replay must end where the forward pass was at the next checkpoint
    // (or at its end): every particle's x and v, bit for bit.
    {
      const bool last_seg = seg + 1 == a->nckpt;
      const float4* rx = last_seg ? a->fx : a->ckx + (size_t)(seg + 1) * np;
      const float4* rv = last_seg ? a->fv : a->ckv + (size_t)(seg + 1) * np;
      pbf::k_count_diff<<<pbf::nblk(a->n), pbf::kB, 0, s>>>((uint32_t)a->n, c->d.posA, c->d.velA,
                                                           rx, rv, a->contacts + 1);
      AHIP(a, hipGetLastError());
    }
    // Reverse through the segment.
    for (int t = len - 1; t >= 0; t--)
      if ((rc = backward_step(a, k, t0 + t, a->tape[(size_t)t])) != PBF_OK) return rc;
  }
  AHIP(a, hipEventRecord(e1, s));
  AHIP(a, hipStreamSynchronize(s));
  float ms = 0.f;
  AHIP(a, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  a->backward_ms = ms;
  // dL/dk: the per-step totals, folded in step order.
  double kb = 0.0;
  {
    std::vector<double> ks((size_t)a->nsteps);
    AHIP(a, hipMemcpy(ks.data(), a->ksteps, ks.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 0; t < a->nsteps; t++) kb += ks[(size_t)t];
  }
  uint32_t cont[2] = {0, 0};
  AHIP(a, hipMemcpy(cont, a->contacts, sizeof cont, hipMemcpyDeviceToHost));
  a->last_contacts = cont[0];
  a->last_replay_diff = cont[1];
  if (xbar0 || vbar0) {
    AHIP(a, hipMemcpy(hx.data(), a->xbar, n * sizeof(float4), hipMemcpyDeviceToHost));
    AHIP(a, hipMemcpy(hv.data(), a->vbar, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
      if (xbar0) { xbar0[3 * i] = hx[i].x; xbar0[3 * i + 1] = hx[i].y; xbar0[3 * i + 2] = hx[i].z; }
      if (vbar0) { vbar0[3 * i] = hv[i].x; vbar0[3 * i + 1] = hv[i].y; vbar0[3 * i + 2] = hv[i].z; }
    }
  }
  if (kbar) *kbar = kb;
  return PBF_OK;
}

PBF_API int pbf_adj_backward(pbf_adj* a, const float* xbar_T, const float* vbar_T, float* xbar0,
                             float* vbar0, double* kbar) {
  try {
    return pbf_adj_backward_impl(a, xbar_T, vbar_T, xbar0, vbar0, kbar);
  } catch (const std::bad_alloc&) {
    return afail(a, PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return afail(a, PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API int pbf_adj_stats(const pbf_adj* a, double* backward_ms, int64_t* wall_contacts) {
  if (!a) return afail(nullptr, PBF_ERR_INVALID, "NULL");
  if (backward_ms) *backward_ms = a->backward_ms;
  if (wall_contacts) *wall_contacts = a->last_contacts;
  return PBF_OK;
}

PBF_API int pbf_adj_replay_check(const pbf_adj* a, int64_t* mismatched) {
  if (!a || !mismatched) return afail(nullptr, PBF_ERR_INVALID, "NULL");
  *mismatched = a->last_replay_diff;
  return PBF_OK;
}

}  // extern "C"
