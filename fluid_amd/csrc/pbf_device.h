// pbf_device.h -- device-side arithmetic of the PBF step for gfx950.
//
// Every expression keeps the operation order of the reference's FP64 code
// (src/particles.cpp:51-248 on CGL::Vector3D, CGL/include/CGL/vector3D.h):
// the build compiles with -ffp-contract=off and correctly rounded division
// and square root, so that the FP32 results are reproducible bit for bit by
// a CPU that evaluates the same expressions in IEEE single precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

// Step constants, derived on the host once per (params, dt).
struct SimConst {
  float h, h2, h6, h9;   // H, H*H, intpow<6>(H), intpow<9>(H)
  float c1;              // poly6 literal 1.56668147106
  float m3c2;            // -3 * 4.774648292756860
  float S;               // tensile_instability_scale (particles.cpp:151)
  float inv_rho0, rho0;  // 1/rho0, rho0
  float eps;             // EPSILON
  float negk;            // -K
  float xc;              // XSPH C
  float dtv;             // dt * VORTICITY_EPSILON
  float inv_dt, dt;      // 1/dt, dt
  float gdt;             // g * dt
  float eps_d;           // EPS_D
  int n_tensile;         // N
  float r9;              // 1/h^9 rounded to nearest (for div_seq)
  float div_hi;          // div_seq is checked for numerators below div_hi
  int fast_div;          // poly6 divides by h^9 with div_seq
  int fast_rcp;          // spiky takes its reciprocal with rcp_seq
  int fast_sqrt;         // spiky takes its square root with sqrt_rn
};

// Grid and collision box of one scene.
struct SceneGrid {
  float lo[3], hi[3];      // box
  float lo_e[3], hi_e[3];  // lo + EPS_D, hi - EPS_D (the final per-axis clamp)
  float inv_cell;          // 1 / (h * cell_scale)
  int dims[3];             // nx, ny, nz
  uint32_t cell_base;      // first global cell of this scene
  uint32_t stop;           // analytic stop-plane walls, bit 2*axis+side
};

// poly6_kernel (src/particles.cpp:134-141).
__device__ __forceinline__ float poly6(const SimConst& c, float r2) {
  if (r2 >= c.h2) return 0.0f;
  float t = c.h2 - r2;
  return (c.c1 * ((t * t) * t)) / c.h9;
}

// grad_spiky_kernel (src/particles.cpp:143-149); r2 = (rx*rx+ry*ry)+rz*rz.
__device__ __forceinline__ float3 spiky(const SimConst& c, float rx, float ry, float rz, float r2) {
  float rl = sqrtf(r2);
  if (rl >= c.h || rl < c.eps_d) return make_float3(0.0f, 0.0f, 0.0f);
  float q = c.h - rl;
  float coef = c.m3c2 * (q * q);
  float rc = 1.0f / (c.h6 * rl);
  return make_float3(rc * (coef * rx), rc * (coef * ry), rc * (coef * rz));
}

// intpow<n>(b) = ((b*b)*b)... (src/bsdf.h:49-57).
__device__ __forceinline__ float ipow(float b, int n) {
  if (n <= 0) return 1.0f;
  float r = b;
  for (int k = 1; k < n; k++) r = r * b;
  return r;
}

// intpow with the exponent known at compile time (NT >= 0), the same
// products in the same order; NT < 0 takes n at run time.
template <int NT>
__device__ __forceinline__ float ipow_n(float b, int n) {
  if (NT < 0) return ipow(b, n);
  if (NT == 0) return 1.0f;
  float r = b;
#pragma unroll
  for (int k = 1; k < (NT > 0 ? NT : 1); k++) r = r * b;
  return r;
}

// std::min / std::max as the reference's per-axis clamp uses them.
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float comp(const float3& v, int a) {
  return a == 0 ? v.x : (a == 1 ? v.y : v.z);
}

// Triangle-mesh collider (the reference's scene BVH, src/particles.cpp:51-132
// with src/bvh.cpp:142-192 and src/static_scene/triangle.cpp:21-80).
// Triangle k: tri[6k] = (p1, original index bits), tri[6k+1] = e1 = p2 - p1,
// tri[6k+2] = e2 = p3 - p1, tri[6k+3..5] = vertex normals n1, n2, n3.
// BVH node, in depth-first preorder: node[2i] = (lo, escape), node[2i+1] =
// (hi, b) with integer bits; escape is the first node after i's subtree;
// b >= 2^31 marks a leaf of the (b >> 27) & 7 triangles from b & (2^27 - 1),
// otherwise (b = 0) the left child is node i + 1.  Node boxes are padded
// outwards, so the box test never drops a triangle that the exact test
// would accept.
struct MeshView {
  const float4* tri = nullptr;
  const float4* node = nullptr;
  uint32_t ntri = 0;  // 0: the analytic box walls
  uint32_t nnode = 0;
  float skin = 0.0f;  // clearance a hit leaves along the face normal
  // One byte per cell of the scene grid: nonzero where a triangle's box,
  // grown by 3 margins, overlaps the cell or one of its 26 neighbours
  // (NULL: always walk the BVH).
  const uint8_t* flag = nullptr;
  float margin = 0.0f;
};

__device__ __forceinline__ float3 f3(const float4& a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float dot3(float3 u, float3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
__device__ __forceinline__ float3 cross3(float3 u, float3 v) {
  return make_float3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}

// Whether the hit point o + t d lies within hp of the box of the triangle
// (p, p + e1, p + e2).  Every exact hit does; the test rejects a rounding
// accept far from the triangle (a near-zero determinant), so that a walk
// that skips far triangles (the BVH's padded boxes, the flagged cells)
// finds what a scan of every triangle finds.
__device__ __forceinline__ bool hit_in_box(float3 o, float3 d, float t, const float4& p,
                                           float3 e1, float3 e2, float hp) {
  const float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
  const float pa[3] = {p.x, p.y, p.z}, a1[3] = {e1.x, e1.y, e1.z}, a2[3] = {e2.x, e2.y, e2.z};
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float q = oa[a] + t * da[a];
    const float v1 = pa[a] + a1[a], v2 = pa[a] + a2[a];
    float lo = pa[a] < v1 ? pa[a] : v1, hi = pa[a] > v1 ? pa[a] : v1;
    lo = v2 < lo ? v2 : lo;
    hi = v2 > hi ? v2 : hi;
    in = in && (q >= lo - hp && q <= hi + hp);
  }
  return in;
}

// Nearest triangle hit of the segment o + t d, 0 <= t <= tmax: the smallest
// t, ties to the lowest original triangle index, so the result does not
// depend on the traversal order.  Each triangle test is Triangle::intersect
// in float with the reference's operation order.  Returns the triangle's
// original index or -1; *t, *n (w n1 + u n2 + v n3) and the face normal
// *nf = e1 x e2 on a hit.
__device__ inline int mesh_nearest(const MeshView& m, float3 o, float3 d, float tmax, float* t_out,
                                   float3* n_out, float3* nf_out) {
  float bt = tmax;
  int best = -1;
  // The slab tests' reciprocals of the direction, once per query.
  const float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z};
  float inva[3];
#pragma unroll
  for (int a = 0; a < 3; a++) inva[a] = da[a] == 0.0f ? 0.0f : 1.0f / da[a];
  // Stackless walk: a missed box or a finished leaf continues at the escape
  // node, a hit inner box at its left child (the next node).  Every move
  // goes to a larger index, so the walk ends within nnode nodes.
  uint32_t ni = 0;
  while (ni < m.nnode) {
    const float4 lo = m.node[2 * ni], hi = m.node[2 * ni + 1];
    const uint32_t esc = __float_as_uint(lo.w), b = __float_as_uint(hi.w);
    // Slab test against [0, bt]; an axis the segment runs parallel to only
    // needs the origin inside the slab.
    float t0 = 0.0f, t1 = bt;
    bool miss = false;
    const float la[3] = {lo.x, lo.y, lo.z}, ha[3] = {hi.x, hi.y, hi.z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (da[a] == 0.0f) {
        miss = miss || oa[a] < la[a] || oa[a] > ha[a];
      } else {
        const float inv = inva[a];
        float ta = (la[a] - oa[a]) * inv, tb = (ha[a] - oa[a]) * inv;
        if (ta > tb) {
          const float tt = ta;
          ta = tb;
          tb = tt;
        }
        t0 = fmaxf(t0, ta);
        t1 = fminf(t1, tb);
      }
    }
    // Rounding of the slab bounds is covered by the padded boxes; accept
    // a box whose entry is within a relative 1e-6 of the current bound.
    if (miss || t0 > t1 * 1.000001f + 1e-30f) {
      ni = esc;
      continue;
    }
    if (b & 0x80000000u) {
      const uint32_t a = b & 0x7ffffffu, end = a + ((b >> 27) & 7u);
      for (uint32_t k = a; k < end; k++) {
        const float4 P = m.tri[6 * (size_t)k];
        const float3 e1 = f3(m.tri[6 * (size_t)k + 1]), e2 = f3(m.tri[6 * (size_t)k + 2]);
        const float3 sv = make_float3(o.x - P.x, o.y - P.y, o.z - P.z);
        const float3 s1 = cross3(d, e2), s2 = cross3(sv, e1);
        const float det = dot3(s1, e1);
        if (det == 0.0f) continue;
        const float num = dot3(s2, e2);
        // Rejections the division below would make anyway, without it:
        // t < 0 (num and det of opposite signs; a zero or NaN num is not
        // taken), and t > bt where num/det exceeds bt by more than the
        // roundings of the two products below and of the quotient (1e-6 >
        // 16 ulp; both products normal, so each is within half an ulp), so
        // t = RN(num/det) > bt for certain.
        if ((num < 0.0f && det > 0.0f) || (num > 0.0f && det < 0.0f)) continue;
        const float btm = bt * 1.000001f, lim = fabsf(det) * btm;
        if (btm >= 0x1p-126f && lim >= 0x1p-126f && fabsf(num) > lim) continue;
        const float t = num / det;
        const int idx = (int)__float_as_uint(P.w);
        if (t < 0.0f || t > bt || (t == bt && best >= 0 && idx > best)) continue;
        const float u = dot3(s1, sv) / det;
        const float v = dot3(s2, d) / det;
        const float w = (1.0f - u) - v;
        if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f || w < 0.0f || w > 1.0f) continue;
        if (!hit_in_box(o, d, t, P, e1, e2, m.margin * 0.5f)) continue;
        const float3 n1 = f3(m.tri[6 * (size_t)k + 3]), n2 = f3(m.tri[6 * (size_t)k + 4]),
                     n3 = f3(m.tri[6 * (size_t)k + 5]);
        bt = t;
        best = idx;
        *n_out = make_float3((w * n1.x + u * n2.x) + v * n3.x, (w * n1.y + u * n2.y) + v * n3.y,
                             (w * n1.z + u * n2.z) + v * n3.z);
        *nf_out = cross3(e1, e2);
      }
      ni = esc;
    } else {
      ni = ni + 1;
    }
  }
  *t_out = bt;
  return best;
}

// Whether the segment o + t d, 0 <= t <= l, can hit the mesh.  A hit point
// lies within margin/2 of its triangle's box (hit_in_box); the cells near a
// triangle cover every triangle box grown by 3 margins, so a segment whose
// box, grown by one margin, touches none of them holds no such point (the
// float cell indexing errs far less than a margin).  The box spans at most
// two cells per axis, one of them the start point's, so one byte, the start
// cell's (flags dilated by one cell), decides.  Longer segments, and
// non-finite ones, walk the BVH.
__device__ __forceinline__ bool mesh_may_hit(const MeshView& m, const SceneGrid& g, float3 o,
                                             float3 d, float l) {
  if (!m.flag) return true;
  const float s[3] = {o.x, o.y, o.z};
  const float e[3] = {o.x + d.x * l, o.y + d.y * l, o.z + d.z * l};
  int cs[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float lo = fminf(s[a], e[a]) - m.margin, hi = fmaxf(s[a], e[a]) + m.margin;
    if (!(fabsf(lo) < 3.0e38f) || !(fabsf(hi) < 3.0e38f) || !(fabsf(s[a]) < 3.0e38f) ||
        !(fabsf(e[a]) < 3.0e38f))
      return true;
    const float f0 = floorf((lo - g.lo[a]) * g.inv_cell), f1 = floorf((hi - g.lo[a]) * g.inv_cell);
    const float fs = floorf((s[a] - g.lo[a]) * g.inv_cell);
    const int c0 = f0 < 0.0f ? 0 : (f0 < (float)g.dims[a] ? (int)f0 : g.dims[a] - 1);
    const int c1 = f1 < 0.0f ? 0 : (f1 < (float)g.dims[a] ? (int)f1 : g.dims[a] - 1);
    cs[a] = fs < 0.0f ? 0 : (fs < (float)g.dims[a] ? (int)fs : g.dims[a] - 1);
    if (c1 - c0 > 1) return true;
  }
  return m.flag[((size_t)cs[2] * (size_t)g.dims[1] + (size_t)cs[1]) * (size_t)g.dims[0] +
                (size_t)cs[0]] != 0;
}

// Box collision along dir for a displacement D of length l (= total).
// Analytic stop planes first (strict 0 < t < l, as src/particles.cpp:59-74),
// then the nearest mesh wall (0 <= t <= l, as the BVH triangle test); on a hit
// move to t - EPS_D, and in prediction slide once along a mesh wall for half
// the remaining length (:114-124).  Ends with the per-axis clamp (:129-131).
// Analytic stop planes of the box (strict 0 < t < l, src/particles.cpp:59-74).
__device__ __forceinline__ bool stop_planes(const SceneGrid& g, const float* pa, const float* da,
                                            float* l) {
  bool intersect = false;
#pragma unroll
  for (int a = 2; a >= 0; a--) {
    if (da[a] == 0.0f) continue;
    int hw = da[a] > 0.0f;
    if (!((g.stop >> (2 * a + hw)) & 1u)) continue;
    float plane = hw ? g.hi[a] : g.lo[a];
    float pt = (plane - pa[a]) / da[a];
    if (pt > 0.0f && pt < *l) {
      *l = pt;
      intersect = true;
    }
  }
  return intersect;
}

// After a hit on a face with normal nf, step off it by `skin` along the
// unit normal, on the side the segment came from.  EPS_D (1e-11) is far
// below float resolution: the reference's back-off along the ray alone can
// leave a float position on or behind the face, from where the next move
// passes through it.
__device__ __forceinline__ void push_off(float* pa, float3 d, float3 nf, float skin) {
  if (!(skin > 0.0f)) return;
  const float dn = dot3(d, nf);
  const float k = skin / sqrtf(dot3(nf, nf));
  const float s = dn > 0.0f ? -k : k;
  pa[0] = pa[0] + s * nf.x;
  pa[1] = pa[1] + s * nf.y;
  pa[2] = pa[2] + s * nf.z;
}

// Mesh collision: the stop planes, then the nearest triangle within the
// (shortened) segment; on a hit move to t - EPS_D (and off the face by the
// skin) and, in prediction, slide once along (D - (D.n) n).unit() for half
// the remaining length, against the mesh only (src/particles.cpp:111-124).
// Ends with the per-axis clamp.
__device__ inline float3 collide_mesh(const SceneGrid& g, const SimConst& c, const MeshView& m,
                                      float3 p, float3 D, float3 dir, float l, float total,
                                      bool slide) {
  float pa[3] = {p.x, p.y, p.z};
  const float da[3] = {dir.x, dir.y, dir.z};
  const bool intersect = stop_planes(g, pa, da, &l);
  float th;
  float3 n = make_float3(0.0f, 0.0f, 0.0f), nf = n;
  const int hit = mesh_may_hit(m, g, p, dir, l) ? mesh_nearest(m, p, dir, l, &th, &n, &nf) : -1;
  if (hit >= 0) l = th;
  if (hit >= 0 || intersect) {
    const float s = l - c.eps_d;
    pa[0] = pa[0] + s * da[0];
    pa[1] = pa[1] + s * da[1];
    pa[2] = pa[2] + s * da[2];
    if (hit >= 0) push_off(pa, dir, nf, m.skin);
    if (slide && hit >= 0 && !intersect && dot3(dir, n) > -1.0f) {
      const float dn = dot3(D, n);
      const float T[3] = {D.x - dn * n.x, D.y - dn * n.y, D.z - dn * n.z};
      const float tl2 = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
      if (tl2 > 0.0f) {
        const float rn = 1.0f / sqrtf(tl2);
        const float3 d2 = make_float3(rn * T[0], rn * T[1], rn * T[2]);
        float l2 = (total - l) * 0.5f, t2;
        float3 n2, nf2;
        const float3 p2 = make_float3(pa[0], pa[1], pa[2]);
        const int hit2 =
            mesh_may_hit(m, g, p2, d2, l2) ? mesh_nearest(m, p2, d2, l2, &t2, &n2, &nf2) : -1;
        if (hit2 >= 0) l2 = t2;
        const float s2 = l2 - c.eps_d;
        pa[0] = pa[0] + s2 * d2.x;
        pa[1] = pa[1] + s2 * d2.y;
        pa[2] = pa[2] + s2 * d2.z;
        if (hit2 >= 0) push_off(pa, d2, nf2, m.skin);
      }
    }
  } else {
    pa[0] = pa[0] + D.x;
    pa[1] = pa[1] + D.y;
    pa[2] = pa[2] + D.z;
  }
  return make_float3(smax(g.lo_e[0], smin(g.hi_e[0], pa[0])),
                     smax(g.lo_e[1], smin(g.hi_e[1], pa[1])),
                     smax(g.lo_e[2], smin(g.hi_e[2], pa[2])));
}

__device__ __forceinline__ float3 collide(const SceneGrid& g, const SimConst& c, float3 p,
                                          float3 D, float3 dir, float l, float total,
                                          bool slide) {
  float pa[3] = {p.x, p.y, p.z};
  const float da[3] = {dir.x, dir.y, dir.z};
  const bool intersect = stop_planes(g, pa, da, &l);
  int hit = -1;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (da[a] == 0.0f) continue;
    int hw = da[a] > 0.0f;
    if ((g.stop >> (2 * a + hw)) & 1u) continue;
    float plane = hw ? g.hi[a] : g.lo[a];
    float t = (plane - pa[a]) / da[a];
    if (t >= 0.0f && t <= l) {
      l = t;
      hit = a;
    }
  }
  if (hit >= 0 || intersect) {
    float s = l - c.eps_d;
    pa[0] = pa[0] + s * da[0];
    pa[1] = pa[1] + s * da[1];
    pa[2] = pa[2] + s * da[2];
    if (slide && hit >= 0 && !intersect && -fabsf(da[hit]) > -1.0f) {
      float T[3] = {D.x, D.y, D.z};
      T[hit] = 0.0f;
      float tl2 = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
      if (tl2 > 0.0f) {
        float rn = 1.0f / sqrtf(tl2);
        float d2[3] = {rn * T[0], rn * T[1], rn * T[2]};
        float l2 = (total - l) * 0.5f;
#pragma unroll
        for (int a = 0; a < 3; a++) {
          if (d2[a] == 0.0f) continue;
          int hw = d2[a] > 0.0f;
          if ((g.stop >> (2 * a + hw)) & 1u) continue;
          float plane = hw ? g.hi[a] : g.lo[a];
          float t = (plane - pa[a]) / d2[a];
          if (t >= 0.0f && t <= l2) l2 = t;
        }
        float s2 = l2 - c.eps_d;
        pa[0] = pa[0] + s2 * d2[0];
        pa[1] = pa[1] + s2 * d2[1];
        pa[2] = pa[2] + s2 * d2[2];
      }
    }
  } else {
    pa[0] = pa[0] + D.x;
    pa[1] = pa[1] + D.y;
    pa[2] = pa[2] + D.z;
  }
  return make_float3(smax(g.lo_e[0], smin(g.hi_e[0], pa[0])),
                     smax(g.lo_e[1], smin(g.hi_e[1], pa[1])),
                     smax(g.lo_e[2], smin(g.hi_e[2], pa[2])));
}

// clamp_response (src/particles.cpp:87-132): prediction with slide.
__device__ __forceinline__ float3 clamp_response(const SceneGrid& g, const SimConst& c,
                                                 const MeshView& m, float3 p, float3 v) {
  float3 D = make_float3(v.x * c.dt, v.y * c.dt, v.z * c.dt);
  float total = sqrtf(D.x * D.x + D.y * D.y + D.z * D.z);
  if (total <= c.eps_d) return p;
  float rl = 1.0f / total;
  float3 dir = make_float3(rl * D.x, rl * D.y, rl * D.z);
  if (m.ntri) return collide_mesh(g, c, m, p, D, dir, total, total, true);
  return collide(g, c, p, D, dir, total, total, true);
}

// clamp (src/particles.cpp:51-84): position correction, no slide.
__device__ __forceinline__ float3 clamp_dp(const SceneGrid& g, const SimConst& c,
                                           const MeshView& m, float3 p, float3 dp) {
  float l = sqrtf(dp.x * dp.x + dp.y * dp.y + dp.z * dp.z);
  if (l <= c.eps_d) return p;
  float rl = 1.0f / l;
  float3 dir = make_float3(rl * dp.x, rl * dp.y, rl * dp.z);
  if (m.ntri) return collide_mesh(g, c, m, p, dp, dir, l, l, false);
  return collide(g, c, p, dp, dir, l, l, false);
}

// floor((x - lo) * inv_cell) clamped to [0, n-1]; NaN maps to 0.
__device__ __forceinline__ int axis_cell(float x, float lo, float inv, int n) {
  float f = floorf((x - lo) * inv);
  if (!(f >= 0.0f)) return 0;
  if (!(f < (float)n)) return n - 1;
  return (int)f;
}

// Neighbour lists: chunked ELL.  A wave's 64 particles own one contiguous
// tile of K4 = K/4 chunks; chunk c of a particle holds its neighbours
// 4c..4c+3 as 16 contiguous bytes, and the 64 lanes' chunks c are adjacent
// (1 KiB).  A builder writes a whole chunk with one 16-byte store; a reader
// fetches four indices with one 16-byte load, and the 64 lanes of the wave
// then read one fully used 1 KiB block.
constexpr int kChunk = 4;

__device__ __forceinline__ size_t nbr_chunk(uint32_t i, int c, int K4) {
  return (((size_t)(i >> 6) * (size_t)K4 + (size_t)c) * 64u + (i & 63u)) * kChunk;
}

__device__ __forceinline__ size_t nbr_slot(uint32_t i, int k, int K4) {
  return nbr_chunk(i, k >> 2, K4) + (size_t)(k & 3);
}

// One 4-entry list chunk.  Lists are read once per pass, so the load is
// non-temporal: it does not displace gathered neighbour state in the L1
// (λ -4%, Δp -4%, build + first λ -3%).
#ifndef PBF_NT_INDEX
#define PBF_NT_INDEX 1
#endif
__device__ __forceinline__ uint4 load_chunk(const uint32_t* __restrict__ p) {
#if PBF_NT_INDEX
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(q.x, q.y, q.z, q.w);
#else
  return *reinterpret_cast<const uint4*>(p);
#endif
}


// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remapping the hardware block index so that every XCD runs contiguous
// ranges of logical blocks keeps the cell-sorted neighbourhoods a wave
// gathers from inside one L2.  Up to kXcdRun blocks per XCD, each XCD takes
// one contiguous eighth of the blocks.  Above that, the XCDs take runs of
// kXcdRun consecutive blocks in turn: within each group of 8 kXcdRun
// blocks, hardware block r (the (r/8)-th of XCD r mod 8) maps to
// (r mod 8) kXcdRun + r/8, and blocks past the last full group keep their
// index.  The runs spread a compressed region's blocks over the XCDs
// (measured on the 4M two-sphere run: 4.09 against 4.19 ms per step, its
// steps after the floor impact 4.42 against 4.62; the 262,144-particle dam
// break, 1,024 blocks, keeps one range per XCD: 1.33 ms per step against
// 1.44 with no remap).  Bijective on [0, nb) either way.
#ifndef PBF_XCD_SWIZZLE
#define PBF_XCD_SWIZZLE 1
#endif
constexpr uint32_t kXcdRun = 256;
__host__ __device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  if (!PBF_XCD_SWIZZLE) return b;
  const uint32_t q = nb >> 3;
  if (q > kXcdRun) {
    constexpr uint32_t S = 8u * kXcdRun;
    if (b >= (nb / S) * S) return b;
    const uint32_t r = b % S;
    return (b - r) + (r & 7u) * kXcdRun + (r >> 3);
  }
  const uint32_t rem = nb & 7u;
  const uint32_t x = b & 7u, k = b >> 3;
  return x * q + (x < rem ? x : rem) + k;
}

// ---- correctly rounded square root, reciprocal and division -------------
// The neighbour loops need IEEE-rounded sqrt and division (the CPU path does
// them in IEEE single precision).  These sequences give the same bits with
// fewer instructions than the general expansions.  Each is used only when a
// flag of SimConst says that an exhaustive check on the device (every
// significand of every binade of the stated range) matched the IEEE result
// and the host has shown that every argument the kernels pass lies in that
// range; otherwise the IEEE expansion runs.  The flags are kernel arguments,
// so the choice is a uniform branch.
//
// sqrt_rn: from the hardware reciprocal square root y ~ 1/sqrt(x), the root
//   s = x*y corrected once by its residual: s + (x - s*s) * y/2 (the Newton
//   step for s^2 = x, with FMA).  Checked on [2^-96, 2^128).  At 0 it gives
//   NaN (0 * inf) and below 2^-96 NaN or a value below 2^-47 (also
//   checked); spiky_sel treats both as under EPS_D, which the host requires
//   to be at least 2^-47, as the IEEE root is.
__device__ __forceinline__ float sqrt_rn(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float s = x * y;
  const float hy = 0.5f * y;
  const float r = fmaf(-s, s, x);
  return fmaf(r, hy, s);
}

// rcp_seq: hardware reciprocal refined by one Newton step with FMA.
//   Checked on [2^-100, 2^100]; the spiky kernel passes h^6 |r| with
//   EPS_D <= |r| < h, or 1.
__device__ __forceinline__ float rcp_seq(float z) {
  const float y = __builtin_amdgcn_rcpf(z);
  const float e = fmaf(-z, y, 1.0f);
  return fmaf(e, y, y);
}

// div_seq: a / b as a * RN(1/b) with two FMA residual corrections.
//   Checked for b = h^9 on 0 and on the binades from that of amin up to
//   div_hi, where the poly6 numerator c1 (h^2 - r^2)^3 is 0 or lies in
//   [amin, amax] (amin from the smallest positive h^2 - r^2); the host
//   requires amin to be normal and amax below div_hi.
__device__ __forceinline__ float div_seq(float a, float b, float rb) {
  float q = a * rb;
  float e = fmaf(-q, b, a);
  q = fmaf(e, rb, q);
  e = fmaf(-q, b, a);
  return fmaf(e, rb, q);
}

// poly6 / spiky for the neighbour loops: the same expressions as poly6() and
// spiky(), with the reference's cut-offs applied by selection (no divergent
// branches).  Outside the cut-off poly6 divides 0 (giving +0) and spiky
// scales by a zero coefficient, giving a +-0 gradient: the callers only add
// these terms to sums that start at +0, where adding +-0 changes no bit
// (a round-to-nearest sum is -0 only when both terms are -0).
__device__ __forceinline__ float poly6_sel(const SimConst& c, float r2) {
  const float t = c.h2 - r2;
  const float a = (r2 >= c.h2) ? 0.0f : c.c1 * ((t * t) * t);
  if (c.fast_div) return div_seq(a, c.h9, c.r9);
  return a / c.h9;
}

// poly6 with the reference's branch (most candidates of a query point lie
// beyond h, and a wave skips the division when all its lanes do) and the
// verified short division for the rest: the bits of poly6().
__device__ __forceinline__ float poly6_br(const SimConst& c, float r2) {
  if (r2 >= c.h2) return 0.0f;
  const float t = c.h2 - r2;
  const float a = c.c1 * ((t * t) * t);
  if (c.fast_div) return div_seq(a, c.h9, c.r9);
  return a / c.h9;
}

__device__ __forceinline__ float3 spiky_sel(const SimConst& c, float rx, float ry, float rz,
                                            float r2) {
  const float rl = c.fast_sqrt ? sqrt_rn(r2) : sqrtf(r2);
  // (rl >= h) || (rl < EPS_D), written to hold for a NaN root as well.
  const bool zero = !(rl < c.h) || !(rl >= c.eps_d);
  const float q = c.h - rl;
  const float coef = zero ? 0.0f : c.m3c2 * (q * q);
  const float z = zero ? 1.0f : c.h6 * rl;
  const float rc = c.fast_rcp ? rcp_seq(z) : 1.0f / z;
  return make_float3(rc * (coef * rx), rc * (coef * ry), rc * (coef * rz));
}

// Element j of a particle array, addressed with a 32-bit byte offset from
// the array base (arrays hold fewer than 2^28 float4), which lets the
// compiler use the scalar-base + 32-bit-offset form of the global loads.
template <class T>
__device__ __forceinline__ const T& at(const T* __restrict__ p, uint32_t j) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + j * (uint32_t)sizeof(T));
}

// Visits the neighbours of particle i in list order.  Groups of U
// independent index loads, then U independent gathers, are issued before the
// U bodies run, so the memory latency of a group overlaps; body(j, data) is
// still called for k = 0, 1, 2, ... in order, keeping the summation order.
template <int U, class Load, class Body>
__device__ __forceinline__ void for_each_neighbor(uint32_t i, int cnt, int K4,
                                                  const uint32_t* __restrict__ nbr, Load load,
                                                  Body body) {
  static_assert(U == 4 || U == 8, "groups of 4 or 8 indices");
  int k = 0;
  for (; k + U <= cnt; k += U) {
    uint32_t j[U];
    const uint4 a = *reinterpret_cast<const uint4*>(nbr + nbr_slot(i, k, K4));
    j[0] = a.x; j[1] = a.y; j[2] = a.z; j[3] = a.w;
    if (U == 8) {
      const uint4 b = *reinterpret_cast<const uint4*>(nbr + nbr_slot(i, k + 4, K4));
      j[U - 4] = b.x; j[U - 3] = b.y; j[U - 2] = b.z; j[U - 1] = b.w;
    }
    decltype(load(0u)) d[U];
#pragma unroll
    for (int u = 0; u < U; u++) d[u] = load(j[u]);
#pragma unroll
    for (int u = 0; u < U; u++) body(j[u], d[u]);
  }
  for (; k < cnt; k++) {
    const uint32_t j = nbr[nbr_slot(i, k, K4)];
    body(j, load(j));
  }
}

// The same visit, software-pipelined over 4-entry chunks: while the bodies
// of chunk c run, the gathers of chunk c+1 and the index load of chunk c+2
// are in flight, so a wave waits for one memory round trip per chunk at most
// instead of two.  The partial last chunk comes from the same pipeline;
// its gathers and bodies are predicated.
#ifndef PBF_PIPE
#define PBF_PIPE 1
#endif
template <class Load, class Body>
__device__ __forceinline__ void for_each_neighbor_pipe(uint32_t i, int cnt, int K4,
                                                       const uint32_t* __restrict__ nbr,
                                                       Load load, Body body) {
  using D = decltype(load(0u));
  auto idx = [&](int ch) { return load_chunk(nbr + nbr_chunk(i, ch, K4)); };
  auto gather = [&](const uint4& q, D* d) {
    d[0] = load(q.x);
    d[1] = load(q.y);
    d[2] = load(q.z);
    d[3] = load(q.w);
  };
  const int nfull = cnt >> 2, nch = (cnt + 3) >> 2;
  uint4 a = make_uint4(i, i, i, i), an = a;
  D d[4];
  if (nch > 0) a = idx(0);
  if (nfull > 0) gather(a, d);
  if (nch > 1) an = idx(1);
  for (int ch = 0; ch < nfull; ch++) {
    D dn[4];
    uint4 ann = an;
    if (ch + 1 < nfull) gather(an, dn);
    if (ch + 2 < nch) ann = idx(ch + 2);
    body(a.x, d[0]);
    body(a.y, d[1]);
    body(a.z, d[2]);
    body(a.w, d[3]);
    a = an;
    an = ann;
#pragma unroll
    for (int u = 0; u < 4; u++) d[u] = dn[u];
  }
  const int rem = cnt & 3;
  if (rem > 0) {
    D t0 = load(a.x), t1, t2;
    if (rem > 1) t1 = load(a.y);
    if (rem > 2) t2 = load(a.z);
    body(a.x, t0);
    if (rem > 1) body(a.y, t1);
    if (rem > 2) body(a.z, t2);
  }
}

// Four lanes per particle (one DPP quad).  Lane k = lane % 4 of particle
// i's quad walks chunks k, k + 4, k + 8, ... of its list (4 entries each),
// gathering and evaluating their NS terms with term(j, data, t); each round
// the four lanes hold chunks 4r .. 4r + 3, and every lane of the quad adds
// their sixteen terms to sums[] in list order (chunk 4r's four entries,
// then chunk 4r + 1's, ...), taking its partners' terms with DPP quad_perm
// broadcasts of lanes 0, 1, 2, 3.  Entries past the list end (the padding
// slots of the last chunk, which hold a particle index, and the chunks past
// the last, which read particle i) contribute +0: a sum that starts at +0
// is never -0, and s + (+0) = s for every other s, so the sums are those of
// the lane-per-particle loops bit for bit.  A quad's lanes share i and cnt,
// so they run the loop in step.  A particle's list is walked as four
// chunk streams at once: a list of m entries takes ceil(m/16) rounds of
// dependent loads instead of ceil(m/4), which is what sets the walk of the
// longest lists (the index load of a lane's next chunk is in flight while
// a round's terms are evaluated).
__device__ __forceinline__ float quad_lane(float t, int k) {
  int r;
  switch (k) {  // the DPP control must be a constant
    case 0: r = __builtin_amdgcn_update_dpp(0, __float_as_int(t), 0x00, 0xf, 0xf, false); break;
    case 1: r = __builtin_amdgcn_update_dpp(0, __float_as_int(t), 0x55, 0xf, 0xf, false); break;
    case 2: r = __builtin_amdgcn_update_dpp(0, __float_as_int(t), 0xAA, 0xf, 0xf, false); break;
    default: r = __builtin_amdgcn_update_dpp(0, __float_as_int(t), 0xFF, 0xf, 0xf, false); break;
  }
  return __int_as_float(r);
}

template <int NS, class Load, class Term>
__device__ __forceinline__ void quad_sums(uint32_t i, int cnt, int K4,
                                          const uint32_t* __restrict__ nbr, Load load, Term term,
                                          float* sums) {
  using D = decltype(load(0u));
  const int k = (int)(threadIdx.x & 3u);
#pragma unroll
  for (int q = 0; q < NS; q++) sums[q] = 0.0f;
  auto idx = [&](int ch) { return load_chunk(nbr + nbr_chunk(i, ch, K4)); };
  const int nch = (cnt + 3) >> 2;
  const int rounds = (nch + 3) >> 2;
  uint4 e = make_uint4(i, i, i, i);
  if (k < nch) e = idx(k);
  for (int r = 0; r < rounds; r++) {
    const int ch = 4 * r + k;
    const D d0 = load(e.x), d1 = load(e.y), d2 = load(e.z), d3 = load(e.w);
    uint4 en = make_uint4(i, i, i, i);
    if (ch + 4 < nch) en = idx(ch + 4);
    float t[4][NS];
    term(e.x, d0, t[0]);
    term(e.y, d1, t[1]);
    term(e.z, d2, t[2]);
    term(e.w, d3, t[3]);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (4 * ch + u >= cnt) {
#pragma unroll
        for (int q = 0; q < NS; q++) t[u][q] = 0.0f;
      }
    }
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int q = 0; q < NS; q++) sums[q] = sums[q] + quad_lane(t[u][q], kk);
      }
    }
    e = en;
  }
}

// Chooses the visit of the lists: pipelined or grouped by U.
template <int U, class Load, class Body>
__device__ __forceinline__ void visit_neighbors(uint32_t i, int cnt, int K4,
                                                const uint32_t* __restrict__ nbr, Load load,
                                                Body body) {
  if (PBF_PIPE && U == 4)
    for_each_neighbor_pipe(i, cnt, K4, nbr, load, body);
  else
    for_each_neighbor<U>(i, cnt, K4, nbr, load, body);
}

}  // namespace pbf
