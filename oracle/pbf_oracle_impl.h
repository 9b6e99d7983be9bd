/*
 * pbf_oracle_impl.h -- body of the CPU restatement, instantiated by
 * pbf_oracle.c once with REAL=float (suffix _f) and once with REAL=double
 * (suffix _d).  TEST INFRASTRUCTURE ONLY (see pbf_oracle.h).
 *
 * Expression order follows the reference exactly; the comment on each
 * function names the reference lines it restates.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define DER CAT(orc_der, SFX)

typedef struct {
  REAL h, h2, h6, h9, c1, m3c2, S, inv_rho0, rho0, eps, negk, xc, dtv, inv_dt,
      dt, gdt, cell, inv_cell, eps_d;
  REAL lo[3], hi[3], lo_e[3], hi_e[3];
  int dims[3];
  int n_tensile;
  int iters;
  unsigned stop;
  const orc_params *P; /* for the mesh collider */
  REAL mesh_hp;        /* half the mesh padding: a hit point must lie this close to the triangle's box */
} DER;

/* poly6_kernel, src/particles.cpp:134-141:
 *   r2 >= H2 ? 0 : 1.56668147106 * intpow<3>(H2 - r2) / intpow<9>(H) */
static inline REAL FN(poly6_)(const DER *d, REAL r2) {
  if (r2 >= d->h2) return (REAL)0;
  REAL t = d->h2 - r2;
  return (d->c1 * ((t * t) * t)) / d->h9;
}

/* grad_spiky_kernel, src/particles.cpp:143-149:
 *   -3 * 4.774648292756860 * intpow<2>(H - r) * r_vec / (intpow<6>(H) * r)
 * The vector division multiplies by the reciprocal (vector3D.h:79-82). */
static inline void FN(spiky_)(const DER *d, REAL rx, REAL ry, REAL rz, REAL r2, REAL *g) {
  REAL rl = SQRT(r2);
  if (rl >= d->h || rl < d->eps_d) {
    g[0] = (REAL)0; g[1] = (REAL)0; g[2] = (REAL)0;
    return;
  }
  REAL q = d->h - rl;
  REAL coef = d->m3c2 * (q * q);
  REAL rc = (REAL)1 / (d->h6 * rl);
  g[0] = rc * (coef * rx);
  g[1] = rc * (coef * ry);
  g[2] = rc * (coef * rz);
}

/* intpow<n> (src/bsdf.h:49-57): ((b*b)*b)*b... ; intpow<0> = 1. */
static inline REAL FN(ipow_)(REAL b, int n) {
  if (n <= 0) return (REAL)1;
  REAL r = b;
  for (int k = 1; k < n; k++) r = r * b;
  return r;
}

static void FN(der_init_)(const orc_params *p, REAL dt, DER *d) {
  d->P = p;
  d->h = (REAL)p->h;
  d->h2 = d->h * d->h;
  d->h6 = FN(ipow_)(d->h, 6);
  d->h9 = FN(ipow_)(d->h, 9);
  d->c1 = (REAL)p->poly6_coef;
  d->m3c2 = (REAL)(-3) * (REAL)p->spiky_coef;
  d->eps_d = (REAL)p->eps_d;
  /* Particle::tensile_instability_scale, src/particles.cpp:151:
   *   1 / poly6_kernel(Vector3D(0, 0, dq * H)) */
  {
    REAL z = (REAL)p->dq * d->h;
    REAL zero = (REAL)0;
    REAL r2 = (zero * zero + zero * zero) + z * z;
    d->S = (REAL)1 / FN(poly6_)(d, r2);
  }
  d->rho0 = (REAL)p->rest_density;
  d->inv_rho0 = (REAL)1 / d->rho0;
  d->eps = (REAL)p->epsilon;
  d->negk = -(REAL)p->k_tensile;
  d->xc = (REAL)p->xsph_c;
  d->dt = dt;
  d->dtv = dt * (REAL)p->vort_eps;
  d->inv_dt = (REAL)1 / dt;
  d->gdt = (REAL)p->gravity * dt;
  d->cell = d->h * (REAL)p->cell_scale;
  d->inv_cell = (REAL)1 / d->cell;
  for (int a = 0; a < 3; a++) {
    d->lo[a] = (REAL)p->box_lo[a];
    d->hi[a] = (REAL)p->box_hi[a];
    d->lo_e[a] = d->lo[a] + d->eps_d;
    d->hi_e[a] = d->hi[a] - d->eps_d;
    REAL ext = FLOOR((d->hi[a] - d->lo[a]) * d->inv_cell);
    int nd = (ext >= (REAL)0 && ext < (REAL)1e6) ? (int)ext + 1 : 1;
    d->dims[a] = nd;
  }
  d->n_tensile = p->n_tensile;
  d->iters = p->iterations;
  d->stop = p->wall_stop_mask;
  /* Mesh padding: 1e-5 of the largest coordinate of the triangles' vertices
   * and the box (as pbf_set_mesh computes it for the BVH boxes). */
  d->mesh_hp = (REAL)0;
  if (p->mesh_nt > 0) {
    REAL scale = (REAL)0;
    for (int64_t k = 0; k < 3 * p->mesh_nt; k++)
      for (int a = 0; a < 3; a++) {
        REAL q = (REAL)p->mesh_pos[3 * (size_t)p->mesh_tri[k] + a];
        q = q < (REAL)0 ? -q : q;
        scale = q > scale ? q : scale;
      }
    for (int a = 0; a < 3; a++) {
      REAL l = d->lo[a] < (REAL)0 ? -d->lo[a] : d->lo[a];
      REAL u = d->hi[a] < (REAL)0 ? -d->hi[a] : d->hi[a];
      scale = l > scale ? l : scale;
      scale = u > scale ? u : scale;
    }
    d->mesh_hp = ((REAL)1e-5f * (scale > (REAL)1e-6f ? scale : (REAL)1e-6f)) * (REAL)0.5;
  }
}

void FN(orc_derive_)(const orc_params *p, REAL *out, int *dims) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  out[ORC_H] = d.h; out[ORC_H2] = d.h2; out[ORC_H6] = d.h6; out[ORC_H9] = d.h9;
  out[ORC_C1] = d.c1; out[ORC_M3C2] = d.m3c2; out[ORC_TENSILE_S] = d.S;
  out[ORC_INV_RHO0] = d.inv_rho0; out[ORC_RHO0] = d.rho0; out[ORC_EPS] = d.eps;
  out[ORC_NEGK] = d.negk; out[ORC_XSPH_C] = d.xc; out[ORC_DT_VORT] = d.dtv;
  out[ORC_INV_DT] = d.inv_dt; out[ORC_DT] = d.dt; out[ORC_GDT] = d.gdt;
  out[ORC_CELL] = d.cell; out[ORC_INV_CELL] = d.inv_cell; out[ORC_EPS_D] = d.eps_d;
  for (int a = 0; a < 3; a++) {
    out[ORC_LO_E0 + a] = d.lo_e[a];
    out[ORC_HI_E0 + a] = d.hi_e[a];
    dims[a] = d.dims[a];
  }
}

REAL FN(orc_poly6_)(const orc_params *p, REAL r2) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  return FN(poly6_)(&d, r2);
}

void FN(orc_spiky_)(const orc_params *p, const REAL *r, REAL *g) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  REAL r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  FN(spiky_)(&d, r[0], r[1], r[2], r2, g);
}

/* std::min / std::max semantics (a NaN second argument yields the first). */
static inline REAL FN(smin_)(REAL a, REAL b) { return (b < a) ? b : a; }
static inline REAL FN(smax_)(REAL a, REAL b) { return (a < b) ? b : a; }

static inline REAL FN(dot3_)(const REAL *u, const REAL *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

static inline void FN(cross3_)(const REAL *u, const REAL *v, REAL *w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

/* Nearest hit of the segment o + t dir, 0 <= t <= tmax, with the mesh:
 * Triangle::intersect (src/static_scene/triangle.cpp:52-80) on every
 * triangle in ascending index, keeping the smallest t and, on equal t, the
 * lowest index (the result of any traversal order of a BVH, cf.
 * src/bvh.cpp:165-192).  n = w n1 + u n2 + v n3.  Returns the index or -1. */
static int FN(mesh_nearest_)(const DER *d, const REAL *o, const REAL *dir, REAL tmax, REAL *t_out,
                            REAL *n_out, REAL *nf_out) {
  const orc_params *P = d->P;
  REAL bt = tmax;
  int best = -1;
  for (int64_t k = 0; k < P->mesh_nt; k++) {
    const int32_t *tv = P->mesh_tri + 3 * k;
    REAL p1[3], e1[3], e2[3], sv[3], s1[3], s2[3];
    for (int a = 0; a < 3; a++) {
      p1[a] = (REAL)P->mesh_pos[3 * (size_t)tv[0] + a];
      e1[a] = (REAL)P->mesh_pos[3 * (size_t)tv[1] + a] - p1[a];
      e2[a] = (REAL)P->mesh_pos[3 * (size_t)tv[2] + a] - p1[a];
      sv[a] = o[a] - p1[a];
    }
    FN(cross3_)(dir, e2, s1);
    FN(cross3_)(sv, e1, s2);
    REAL det = FN(dot3_)(s1, e1);
    if (det == (REAL)0) continue;
    REAL t = FN(dot3_)(s2, e2) / det;
    if (t < (REAL)0 || t > bt || (t == bt && best >= 0)) continue;
    REAL u = FN(dot3_)(s1, sv) / det;
    REAL v = FN(dot3_)(s2, dir) / det;
    REAL w = ((REAL)1 - u) - v;
    if (u < (REAL)0 || u > (REAL)1 || v < (REAL)0 || v > (REAL)1 || w < (REAL)0 || w > (REAL)1)
      continue;
    /* The hit point must lie within mesh_hp of the triangle's box: true of
     * every exact hit; it rejects a rounding accept far from the triangle
     * (det near 0), so that a BVH that skips far triangles finds the same. */
    int inbox = 1;
    for (int a = 0; a < 3; a++) {
      REAL q = o[a] + t * dir[a];
      REAL v1 = p1[a] + e1[a], v2 = p1[a] + e2[a];
      REAL lo = p1[a] < v1 ? p1[a] : v1, hi = p1[a] > v1 ? p1[a] : v1;
      lo = v2 < lo ? v2 : lo;
      hi = v2 > hi ? v2 : hi;
      if (!(q >= lo - d->mesh_hp && q <= hi + d->mesh_hp)) inbox = 0;
    }
    if (!inbox) continue;
    bt = t;
    best = (int)k;
    FN(cross3_)(e1, e2, nf_out);
    for (int a = 0; a < 3; a++)
      n_out[a] = (w * (REAL)P->mesh_nrm[3 * (size_t)tv[0] + a] +
                  u * (REAL)P->mesh_nrm[3 * (size_t)tv[1] + a]) +
                 v * (REAL)P->mesh_nrm[3 * (size_t)tv[2] + a];
  }
  *t_out = bt;
  return best;
}

/* Analytic stop planes (strict 0 < t < l), src/particles.cpp:59-74. */
static int FN(stop_planes_)(const DER *d, const REAL *p, const REAL *dir, REAL *l) {
  int intersect = 0;
  for (int a = 2; a >= 0; a--) {
    if (dir[a] == (REAL)0) continue;
    int hw = dir[a] > (REAL)0;
    if (!((d->stop >> (2 * a + hw)) & 1u)) continue;
    REAL plane = hw ? d->hi[a] : d->lo[a];
    REAL pt = (plane - p[a]) / dir[a];
    if (pt > (REAL)0 && pt < *l) {
      *l = pt;
      intersect = 1;
    }
  }
  return intersect;
}

/* After a hit on a face with normal nf, step off it by the skin along the
 * unit normal, on the side the segment came from (EPS_D alone is below float
 * resolution; see the HIP collider). */
static void FN(push_off_)(REAL *p, const REAL *dir, const REAL *nf, REAL skin) {
  if (!(skin > (REAL)0)) return;
  REAL dn = FN(dot3_)(dir, nf);
  REAL k = skin / SQRT(FN(dot3_)(nf, nf));
  REAL s = dn > (REAL)0 ? -k : k;
  p[0] = p[0] + s * nf[0];
  p[1] = p[1] + s * nf[1];
  p[2] = p[2] + s * nf[2];
}

/* clamp / clamp_response against the mesh (src/particles.cpp:51-132): stop
 * planes, nearest triangle, move to t - EPS_D; in prediction slide once
 * along (D - (D.n) n).unit() for half the remaining length against the
 * mesh only; then the per-axis clamp.  A zero tangent does not slide. */
static void FN(collide_mesh_)(const DER *d, REAL *p, const REAL *D, const REAL *dir, REAL l,
                              REAL total, int slide) {
  const REAL skin = (REAL)d->P->mesh_skin;
  int intersect = FN(stop_planes_)(d, p, dir, &l);
  REAL th, n[3] = {0, 0, 0}, nf[3] = {0, 0, 0};
  int hit = FN(mesh_nearest_)(d, p, dir, l, &th, n, nf);
  if (hit >= 0) l = th;
  if (hit >= 0 || intersect) {
    REAL s = l - d->eps_d;
    p[0] = p[0] + s * dir[0];
    p[1] = p[1] + s * dir[1];
    p[2] = p[2] + s * dir[2];
    if (hit >= 0) FN(push_off_)(p, dir, nf, skin);
    if (slide && hit >= 0 && !intersect && FN(dot3_)(dir, n) > (REAL)-1) {
      REAL dn = FN(dot3_)(D, n);
      REAL T[3] = {D[0] - dn * n[0], D[1] - dn * n[1], D[2] - dn * n[2]};
      REAL tl2 = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
      if (tl2 > (REAL)0) {
        REAL rn = (REAL)1 / SQRT(tl2);
        REAL d2[3] = {rn * T[0], rn * T[1], rn * T[2]};
        REAL l2 = (total - l) * (REAL)0.5, t2, n2[3], nf2[3];
        int hit2 = FN(mesh_nearest_)(d, p, d2, l2, &t2, n2, nf2);
        if (hit2 >= 0) l2 = t2;
        REAL s2 = l2 - d->eps_d;
        p[0] = p[0] + s2 * d2[0];
        p[1] = p[1] + s2 * d2[1];
        p[2] = p[2] + s2 * d2[2];
        if (hit2 >= 0) FN(push_off_)(p, d2, nf2, skin);
      }
    }
  } else {
    p[0] = p[0] + D[0];
    p[1] = p[1] + D[1];
    p[2] = p[2] + D[2];
  }
  for (int a = 0; a < 3; a++) p[a] = FN(smax_)(d->lo_e[a], FN(smin_)(d->hi_e[a], p[a]));
}

/* Ray clip against the collision box.  Restates the shared body of clamp
 * (src/particles.cpp:51-84) and clamp_response (:87-132) for an axis-aligned
 * box [lo, hi] whose walls are either analytic stop planes (the reference's
 * z = 1 and y = 1.49 planes, :59-74, strict 0 < t < l, no slide) or mesh walls
 * with inward normals (the reference's BVH walls, any/nearest hit with
 * 0 <= t <= l as in marching_triangle.cpp:31-34; in prediction they slide
 * once along the wall for half the remaining length, :116-124).  Ends with the
 * per-axis clamp to [lo + EPS_D, hi - EPS_D] (:81-83, :129-131). */
static void FN(collide_)(const DER *d, REAL *p, const REAL *D, const REAL *dir, REAL l,
                         REAL total, int slide) {
  if (d->P && d->P->mesh_nt > 0) {
    FN(collide_mesh_)(d, p, D, dir, l, total, slide);
    return;
  }
  int intersect = 0;
  for (int a = 2; a >= 0; a--) {
    if (dir[a] == (REAL)0) continue;
    int hw = dir[a] > (REAL)0;
    if (!((d->stop >> (2 * a + hw)) & 1u)) continue;
    REAL plane = hw ? d->hi[a] : d->lo[a];
    REAL pt = (plane - p[a]) / dir[a];
    if (pt > (REAL)0 && pt < l) {
      l = pt;
      intersect = 1;
    }
  }
  int hit = -1;
  for (int a = 0; a < 3; a++) {
    if (dir[a] == (REAL)0) continue;
    int hw = dir[a] > (REAL)0;
    if ((d->stop >> (2 * a + hw)) & 1u) continue;
    REAL plane = hw ? d->hi[a] : d->lo[a];
    REAL t = (plane - p[a]) / dir[a];
    if (t >= (REAL)0 && t <= l) {
      l = t;
      hit = a;
    }
  }
  if (hit >= 0 || intersect) {
    REAL s = l - d->eps_d;
    p[0] = p[0] + s * dir[0];
    p[1] = p[1] + s * dir[1];
    p[2] = p[2] + s * dir[2];
    /* dot(d, n) > -1 with n the inward axis normal is -|d_a| > -1. */
    if (slide && hit >= 0 && !intersect && -FABS(dir[hit]) > (REAL)-1) {
      /* (delta_p - dot(delta_p, n) n).unit(): the axis component drops out
       * exactly.  A zero tangent (the reference would divide by zero) does
       * not slide. */
      REAL T[3] = {D[0], D[1], D[2]};
      T[hit] = (REAL)0;
      REAL tl2 = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
      if (tl2 > (REAL)0) {
        REAL rn = (REAL)1 / SQRT(tl2);
        REAL d2[3] = {rn * T[0], rn * T[1], rn * T[2]};
        REAL l2 = (total - l) * (REAL)0.5;
        for (int a = 0; a < 3; a++) {
          if (d2[a] == (REAL)0) continue;
          int hw = d2[a] > (REAL)0;
          if ((d->stop >> (2 * a + hw)) & 1u) continue;
          REAL plane = hw ? d->hi[a] : d->lo[a];
          REAL t = (plane - p[a]) / d2[a];
          if (t >= (REAL)0 && t <= l2) l2 = t;
        }
        REAL s2 = l2 - d->eps_d;
        p[0] = p[0] + s2 * d2[0];
        p[1] = p[1] + s2 * d2[1];
        p[2] = p[2] + s2 * d2[2];
      }
    }
  } else {
    p[0] = p[0] + D[0];
    p[1] = p[1] + D[1];
    p[2] = p[2] + D[2];
  }
  for (int a = 0; a < 3; a++) p[a] = FN(smax_)(d->lo_e[a], FN(smin_)(d->hi_e[a], p[a]));
}

/* clamp_response, src/particles.cpp:87-132. */
static void FN(clamp_response_)(const DER *d, REAL *p, const REAL *v, REAL dt) {
  REAL D[3] = {v[0] * dt, v[1] * dt, v[2] * dt};
  REAL total = SQRT(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
  if (total <= d->eps_d) return;
  REAL rl = (REAL)1 / total;
  REAL dir[3] = {rl * D[0], rl * D[1], rl * D[2]};
  FN(collide_)(d, p, D, dir, total, total, 1);
}

/* clamp, src/particles.cpp:51-84. */
static void FN(clamp_)(const DER *d, REAL *p, const REAL *dp) {
  REAL l = SQRT(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
  if (l <= d->eps_d) return;
  REAL rl = (REAL)1 / l;
  REAL dir[3] = {rl * dp[0], rl * dp[1], rl * dp[2]};
  FN(collide_)(d, p, dp, dir, l, l, 0);
}

void FN(orc_clamp_response_)(const orc_params *p, REAL *pos, const REAL *vel, REAL dt) {
  DER d;
  FN(der_init_)(p, dt, &d);
  FN(clamp_response_)(&d, pos, vel, dt);
}

void FN(orc_clamp_)(const orc_params *p, REAL *pos, const REAL *dp) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  FN(clamp_)(&d, pos, dp);
}

/* Grid cell of one coordinate: floor((x - lo) * inv_cell), clamped. */
static inline int FN(axis_cell_)(REAL x, REAL lo, REAL inv, int n) {
  REAL f = FLOOR((x - lo) * inv);
  if (!(f >= (REAL)0)) return 0;
  if (!(f < (REAL)n)) return n - 1;
  return (int)f;
}

static inline void FN(cell3_)(const DER *d, const REAL *x, int *c) {
  c[0] = FN(axis_cell_)(x[0], d->lo[0], d->inv_cell, d->dims[0]);
  c[1] = FN(axis_cell_)(x[1], d->lo[1], d->inv_cell, d->dims[1]);
  c[2] = FN(axis_cell_)(x[2], d->lo[2], d->inv_cell, d->dims[2]);
}

static inline uint32_t FN(cell_lin_)(const DER *d, const int *c) {
  return ((uint32_t)c[2] * (uint32_t)d->dims[1] + (uint32_t)c[1]) * (uint32_t)d->dims[0] +
         (uint32_t)c[0];
}

void FN(orc_cells_)(const orc_params *p, size_t n, const REAL *xs, uint32_t *cell) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  for (size_t i = 0; i < n; i++) {
    int c[3];
    FN(cell3_)(&d, xs + 3 * i, c);
    cell[i] = FN(cell_lin_)(&d, c);
  }
}

static inline REAL FN(dist2_)(const REAL *a, const REAL *b) {
  REAL rx = a[0] - b[0], ry = a[1] - b[1], rz = a[2] - b[2];
  return rx * rx + ry * ry + rz * rz;
}

/* Counting sort of particle indices by cell, stable in index. */
typedef struct {
  uint32_t *start; /* C+1 */
  uint32_t *slot;  /* n */
  size_t ncell;
} CAT(orc_grid, SFX);

static int FN(grid_build_)(const DER *d, size_t n, const REAL *xs, CAT(orc_grid, SFX) *g) {
  size_t C = (size_t)d->dims[0] * d->dims[1] * d->dims[2];
  g->ncell = C;
  g->start = (uint32_t *)calloc(C + 1, sizeof(uint32_t));
  g->slot = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *cell = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *fill = (uint32_t *)calloc(C + 1, sizeof(uint32_t));
  if (!g->start || !g->slot || !cell || !fill) {
    free(g->start); free(g->slot); free(cell); free(fill);
    return -1;
  }
  for (size_t i = 0; i < n; i++) {
    int c[3];
    FN(cell3_)(d, xs + 3 * i, c);
    cell[i] = FN(cell_lin_)(d, c);
    g->start[cell[i] + 1]++;
  }
  for (size_t c = 0; c < C; c++) g->start[c + 1] += g->start[c];
  for (size_t i = 0; i < n; i++) g->slot[g->start[cell[i]] + fill[cell[i]]++] = (uint32_t)i;
  free(cell);
  free(fill);
  return 0;
}

static void FN(grid_free_)(CAT(orc_grid, SFX) *g) {
  free(g->start);
  free(g->slot);
}

/* Visit the 27 cells around x in ascending linear cell order (dz, dy, dx
 * from -1 to 1), particles within a cell in ascending index.  Calls
 * body(j) for every candidate. */
#define GRID_FOREACH(d, g, x, JVAR, BODY)                                        \
  do {                                                                           \
    int c0_[3];                                                                  \
    FN(cell3_)(d, x, c0_);                                                       \
    for (int dz_ = -1; dz_ <= 1; dz_++) {                                        \
      int z_ = c0_[2] + dz_;                                                     \
      if (z_ < 0 || z_ >= (d)->dims[2]) continue;                                \
      for (int dy_ = -1; dy_ <= 1; dy_++) {                                      \
        int y_ = c0_[1] + dy_;                                                   \
        if (y_ < 0 || y_ >= (d)->dims[1]) continue;                              \
        for (int dx_ = -1; dx_ <= 1; dx_++) {                                    \
          int x_ = c0_[0] + dx_;                                                 \
          if (x_ < 0 || x_ >= (d)->dims[0]) continue;                            \
          int cc_[3] = {x_, y_, z_};                                             \
          uint32_t lin_ = FN(cell_lin_)(d, cc_);                                 \
          for (uint32_t t_ = (g)->start[lin_]; t_ < (g)->start[lin_ + 1]; t_++) { \
            uint32_t JVAR = (g)->slot[t_];                                       \
            BODY                                                                 \
          }                                                                      \
        }                                                                        \
      }                                                                          \
    }                                                                            \
  } while (0)

/* Neighbour sets, src/particles.cpp:254-265: N_i = { j != i :
 * |x*_i - x*_j|^2 <= H2 } (inclusive).  BRUTE keeps the reference's
 * ascending-j order; GRID lists the same set in (cell, index) order. */
long FN(orc_neighbors_)(const orc_params *p, int mode, size_t n, const REAL *xs,
                        uint32_t *off, uint32_t *list, size_t cap) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  uint32_t *cnt = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
  if (!cnt) return -2;
  CAT(orc_grid, SFX) g;
  g.start = NULL;
  g.slot = NULL;
  if (mode == ORC_NBR_GRID && FN(grid_build_)(&d, n, xs, &g) != 0) {
    free(cnt);
    return -2;
  }
  REAL h2 = d.h2;
  if (mode == ORC_NBR_BRUTE) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; i++) {
      uint32_t c = 0;
      for (size_t j = 0; j < n; j++)
        if (j != (size_t)i && FN(dist2_)(xs + 3 * i, xs + 3 * j) <= h2) c++;
      cnt[i] = c;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
      uint32_t c = 0;
      GRID_FOREACH(&d, &g, xs + 3 * i, j, {
        if (j != (uint32_t)i && FN(dist2_)(xs + 3 * i, xs + 3 * (size_t)j) <= h2) c++;
      });
      cnt[i] = c;
    }
  }
  off[0] = 0;
  for (size_t i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i];
  long total = (long)off[n];
  if ((size_t)total > cap) {
    free(cnt);
    if (mode == ORC_NBR_GRID) FN(grid_free_)(&g);
    return -1;
  }
  if (mode == ORC_NBR_BRUTE) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; i++) {
      uint32_t k = off[i];
      for (size_t j = 0; j < n; j++)
        if (j != (size_t)i && FN(dist2_)(xs + 3 * i, xs + 3 * j) <= h2) list[k++] = (uint32_t)j;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
      uint32_t k = off[i];
      GRID_FOREACH(&d, &g, xs + 3 * i, j, {
        if (j != (uint32_t)i && FN(dist2_)(xs + 3 * i, xs + 3 * (size_t)j) <= h2) list[k++] = j;
      });
    }
    FN(grid_free_)(&g);
  }
  free(cnt);
  return total;
}

/* applyForceVelocity, src/particles.cpp:175-183: v.y -= g*dt; x* = x;
 * clamp_response(x*, v, dt). */
static void FN(predict_)(const DER *d, size_t n, const REAL *x, REAL *v, REAL *xs) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    REAL *vi = v + 3 * i;
    vi[1] = vi[1] - d->gdt;
    REAL *p = xs + 3 * i;
    p[0] = x[3 * i];
    p[1] = x[3 * i + 1];
    p[2] = x[3 * i + 2];
    FN(clamp_response_)(d, p, vi, d->dt);
  }
}

void FN(orc_predict_)(const orc_params *p, size_t n, const REAL *x, REAL *v, REAL *xs, REAL dt) {
  DER d;
  FN(der_init_)(p, dt, &d);
  FN(predict_)(&d, n, x, v, xs);
}

/* newtonStepCalculateLambda, src/particles.cpp:185-204. */
static void FN(lambda_)(const DER *d, size_t n, const REAL *xs, const uint32_t *off,
                        const uint32_t *list, REAL *rho, REAL *lambda) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    const REAL *xi = xs + 3 * i;
    REAL dens = 0, denom = 0, G[3] = {0, 0, 0};
    for (uint32_t k = off[i]; k < off[i + 1]; k++) {
      const REAL *xj = xs + 3 * (size_t)list[k];
      REAL rx = xi[0] - xj[0], ry = xi[1] - xj[1], rz = xi[2] - xj[2];
      REAL r2 = rx * rx + ry * ry + rz * rz;
      REAL w = FN(poly6_)(d, r2);
      REAL g[3];
      FN(spiky_)(d, rx, ry, rz, r2, g);
      dens = dens + w;
      g[0] = g[0] * d->inv_rho0;
      g[1] = g[1] * d->inv_rho0;
      g[2] = g[2] * d->inv_rho0;
      G[0] = G[0] + g[0];
      G[1] = G[1] + g[1];
      G[2] = G[2] + g[2];
      denom = denom + (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    }
    REAL c = dens / d->rho0 - (REAL)1;
    denom = denom + (G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
    lambda[i] = -c / (denom + d->eps);
    if (rho) rho[i] = dens;
  }
}

void FN(orc_lambda_)(const orc_params *p, size_t n, const REAL *xs, const uint32_t *off,
                     const uint32_t *list, REAL *rho, REAL *lambda) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  FN(lambda_)(&d, n, xs, off, list, rho, lambda);
}

/* newtonStepUpdatePosition, src/particles.cpp:206-213 (Jacobi: reads only
 * the lambda-pass state), with s_corr from :195. */
static void FN(deltap_)(const DER *d, size_t n, const REAL *xs, const REAL *lambda,
                        const uint32_t *off, const uint32_t *list, REAL *xs_out) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    const REAL *xi = xs + 3 * i;
    REAL li = lambda[i];
    REAL dp[3] = {0, 0, 0};
    for (uint32_t k = off[i]; k < off[i + 1]; k++) {
      uint32_t j = list[k];
      const REAL *xj = xs + 3 * (size_t)j;
      REAL rx = xi[0] - xj[0], ry = xi[1] - xj[1], rz = xi[2] - xj[2];
      REAL r2 = rx * rx + ry * ry + rz * rz;
      REAL w = FN(poly6_)(d, r2);
      REAL s = d->negk * FN(ipow_)(w * d->S, d->n_tensile);
      REAL g[3];
      FN(spiky_)(d, rx, ry, rz, r2, g);
      REAL sc = (li + lambda[j]) + s;
      dp[0] = dp[0] + sc * g[0];
      dp[1] = dp[1] + sc * g[1];
      dp[2] = dp[2] + sc * g[2];
    }
    dp[0] = dp[0] * d->inv_rho0;
    dp[1] = dp[1] * d->inv_rho0;
    dp[2] = dp[2] * d->inv_rho0;
    REAL *po = xs_out + 3 * i;
    po[0] = xi[0];
    po[1] = xi[1];
    po[2] = xi[2];
    FN(clamp_)(d, po, dp);
  }
}

void FN(orc_deltap_)(const orc_params *p, size_t n, const REAL *xs, const REAL *lambda,
                     const uint32_t *off, const uint32_t *list, REAL *xs_out) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  FN(deltap_)(&d, n, xs, lambda, off, list, xs_out);
}

/* One particle of calculateVorticityApplyXSPHViscosity,
 * src/particles.cpp:219-234, reading neighbour velocities from vel. */
static inline void FN(xsph_one_)(const DER *d, size_t i, const REAL *xs, const REAL *vel,
                                 const uint32_t *off, const uint32_t *list, const REAL *vi,
                                 REAL *v_out, REAL *omega, REAL *rho) {
  const REAL *xi = xs + 3 * i;
  REAL om[3] = {0, 0, 0}, visc[3] = {0, 0, 0}, dens = 0;
  for (uint32_t k = off[i]; k < off[i + 1]; k++) {
    uint32_t j = list[k];
    const REAL *vj = vel + 3 * (size_t)j;
    REAL u0 = vj[0] - vi[0], u1 = vj[1] - vi[1], u2 = vj[2] - vi[2];
    const REAL *xj = xs + 3 * (size_t)j;
    REAL rx = xi[0] - xj[0], ry = xi[1] - xj[1], rz = xi[2] - xj[2];
    REAL r2 = rx * rx + ry * ry + rz * rz;
    REAL g[3];
    FN(spiky_)(d, rx, ry, rz, r2, g);
    om[0] = om[0] + (u1 * g[2] - u2 * g[1]);
    om[1] = om[1] + (u2 * g[0] - u0 * g[2]);
    om[2] = om[2] + (u0 * g[1] - u1 * g[0]);
    REAL w = FN(poly6_)(d, r2);
    visc[0] = visc[0] + u0 * w;
    visc[1] = visc[1] + u1 * w;
    visc[2] = visc[2] + u2 * w;
    dens = dens + w;
  }
  v_out[0] = vi[0] + d->xc * visc[0];
  v_out[1] = vi[1] + d->xc * visc[1];
  v_out[2] = vi[2] + d->xc * visc[2];
  omega[3 * i] = om[0];
  omega[3 * i + 1] = om[1];
  omega[3 * i + 2] = om[2];
  if (rho) rho[i] = dens;
}

/* updateVelocity + calculateVorticityApplyXSPHViscosity, src/particles.cpp:
 * 215-234 in the loop :285-288.  JACOBI: every v_i = (x*_i - x_i)/dt first,
 * XSPH into v_out.  SEQUENTIAL: the reference's in-place sweep in index
 * order, where neighbour j < i is already final and j > i still holds its
 * predicted velocity v_pred. */
static void FN(xsph_)(const DER *d, int mode, size_t n, const REAL *x, const REAL *xs,
                      const REAL *v_pred, const uint32_t *off, const uint32_t *list,
                      REAL *v_out, REAL *omega, REAL *rho, REAL *scratch) {
  if (mode == ORC_XSPH_JACOBI) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
      for (int a = 0; a < 3; a++) scratch[3 * i + a] = d->inv_dt * (xs[3 * i + a] - x[3 * i + a]);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
      FN(xsph_one_)(d, (size_t)i, xs, scratch, off, list, scratch + 3 * i, v_out + 3 * i, omega,
                    rho);
  } else {
    for (size_t i = 0; i < 3 * n; i++) scratch[i] = v_pred[i];
    for (size_t i = 0; i < n; i++) {
      REAL vi[3];
      for (int a = 0; a < 3; a++) vi[a] = d->inv_dt * (xs[3 * i + a] - x[3 * i + a]);
      for (int a = 0; a < 3; a++) scratch[3 * i + a] = vi[a];
      FN(xsph_one_)(d, i, xs, scratch, off, list, vi, scratch + 3 * i, omega, rho);
    }
    for (size_t i = 0; i < 3 * n; i++) v_out[i] = scratch[i];
  }
}

void FN(orc_xsph_)(const orc_params *p, int mode, size_t n, const REAL *x, const REAL *xs,
                   const REAL *v_pred, const uint32_t *off, const uint32_t *list, REAL *v_out,
                   REAL *omega, REAL *rho, REAL dt) {
  DER d;
  FN(der_init_)(p, dt, &d);
  REAL *scratch = (REAL *)malloc((3 * n + 3) * sizeof(REAL));
  if (!scratch) return;
  FN(xsph_)(&d, mode, n, x, xs, v_pred, off, list, v_out, omega, rho, scratch);
  free(scratch);
}

/* applyVorticity + updatePosition, src/particles.cpp:236-248 in :290-294:
 * eta = sum |w_j| gradW_ij; if |eta| > EPS_D, v += dt*eps_v*(eta.unit() x
 * w_i); x <- x*. */
static void FN(vorticity_)(const DER *d, size_t n, const REAL *xs, const REAL *omega,
                           const uint32_t *off, const uint32_t *list, REAL *v, REAL *x) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    const REAL *xi = xs + 3 * i;
    REAL eta[3] = {0, 0, 0};
    for (uint32_t k = off[i]; k < off[i + 1]; k++) {
      uint32_t j = list[k];
      const REAL *xj = xs + 3 * (size_t)j;
      REAL rx = xi[0] - xj[0], ry = xi[1] - xj[1], rz = xi[2] - xj[2];
      REAL r2 = rx * rx + ry * ry + rz * rz;
      REAL g[3];
      FN(spiky_)(d, rx, ry, rz, r2, g);
      const REAL *wj = omega + 3 * (size_t)j;
      REAL wn = SQRT(wj[0] * wj[0] + wj[1] * wj[1] + wj[2] * wj[2]);
      eta[0] = eta[0] + wn * g[0];
      eta[1] = eta[1] + wn * g[1];
      eta[2] = eta[2] + wn * g[2];
    }
    REAL e2 = eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2];
    if (SQRT(e2) > d->eps_d) {
      REAL rn = (REAL)1 / SQRT(e2);
      REAL u0 = rn * eta[0], u1 = rn * eta[1], u2 = rn * eta[2];
      const REAL *wi = omega + 3 * i;
      REAL c0 = u1 * wi[2] - u2 * wi[1];
      REAL c1 = u2 * wi[0] - u0 * wi[2];
      REAL c2 = u0 * wi[1] - u1 * wi[0];
      v[3 * i] = v[3 * i] + d->dtv * c0;
      v[3 * i + 1] = v[3 * i + 1] + d->dtv * c1;
      v[3 * i + 2] = v[3 * i + 2] + d->dtv * c2;
    }
    x[3 * i] = xi[0];
    x[3 * i + 1] = xi[1];
    x[3 * i + 2] = xi[2];
  }
}

void FN(orc_vorticity_)(const orc_params *p, size_t n, const REAL *xs, const REAL *omega,
                        const uint32_t *off, const uint32_t *list, REAL *v, REAL *x, REAL dt) {
  DER d;
  FN(der_init_)(p, dt, &d);
  FN(vorticity_)(&d, n, xs, omega, off, list, v, x);
}

/* Particles::timeStep(dt), src/particles.cpp:250-297. */
int FN(orc_step_)(const orc_params *p, int nbr_mode, int xsph_mode, size_t n, REAL *x, REAL *v,
                  REAL *rho, REAL dt, int nsteps) {
  DER d;
  FN(der_init_)(p, dt, &d);
  size_t n3 = 3 * n + 3;
  REAL *xs = (REAL *)malloc(n3 * sizeof(REAL));
  REAL *xs2 = (REAL *)malloc(n3 * sizeof(REAL));
  REAL *lam = (REAL *)malloc((n + 1) * sizeof(REAL));
  REAL *vpred = (REAL *)malloc(n3 * sizeof(REAL));
  REAL *vout = (REAL *)malloc(n3 * sizeof(REAL));
  REAL *om = (REAL *)malloc(n3 * sizeof(REAL));
  REAL *scratch = (REAL *)malloc(n3 * sizeof(REAL));
  uint32_t *off = (uint32_t *)malloc((n + 1) * sizeof(uint32_t));
  size_t cap = 48 * (n + 1);
  uint32_t *list = (uint32_t *)malloc(cap * sizeof(uint32_t));
  int rc = 0;
  if (!xs || !xs2 || !lam || !vpred || !vout || !om || !scratch || !off || !list) {
    rc = -1;
    goto done;
  }
  for (int s = 0; s < nsteps; s++) {
    FN(predict_)(&d, n, x, v, xs);
    for (;;) {
      long tot = FN(orc_neighbors_)(p, nbr_mode, n, xs, off, list, cap);
      if (tot >= 0) break;
      if (tot == -2) { rc = -1; goto done; }
      free(list);
      cap = 2 * cap;
      list = (uint32_t *)malloc(cap * sizeof(uint32_t));
      if (!list) { rc = -1; goto done; }
    }
    for (int it = 0; it < d.iters; it++) {
      FN(lambda_)(&d, n, xs, off, list, NULL, lam);
      FN(deltap_)(&d, n, xs, lam, off, list, xs2);
      REAL *t = xs;
      xs = xs2;
      xs2 = t;
    }
    memcpy(vpred, v, 3 * n * sizeof(REAL));
    FN(xsph_)(&d, xsph_mode, n, x, xs, vpred, off, list, vout, om, rho, scratch);
    FN(vorticity_)(&d, n, xs, om, off, list, vout, x);
    memcpy(v, vout, 3 * n * sizeof(REAL));
  }
done:
  free(xs); free(xs2); free(lam); free(vpred); free(vout); free(om); free(scratch);
  free(off); free(list);
  return rc;
}

/* Particles::estimateDensities, src/particles.cpp:440-444 -> :158-163:
 * rho_i = sum over ALL particles (self included) of W(x_j - x_i).  BRUTE
 * keeps ascending j; GRID visits the 27 cells. */
void FN(orc_estimate_densities_)(const orc_params *p, int mode, size_t n, const REAL *x,
                                 REAL *rho) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  if (mode == ORC_NBR_BRUTE) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; i++) {
      REAL s = 0;
      for (size_t j = 0; j < n; j++) s = s + FN(poly6_)(&d, FN(dist2_)(x + 3 * j, x + 3 * i));
      rho[i] = s;
    }
    return;
  }
  CAT(orc_grid, SFX) g;
  if (FN(grid_build_)(&d, n, x, &g) != 0) return;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)n; i++) {
    REAL s = 0;
    GRID_FOREACH(&d, &g, x + 3 * i, j,
                 { s = s + FN(poly6_)(&d, FN(dist2_)(x + 3 * (size_t)j, x + 3 * i)); });
    rho[i] = s;
  }
  FN(grid_free_)(&g);
}

/* estimateDensityAt, src/particles.cpp:446-453: density of the particles at
 * m query points, sum_j poly6(x_j - p).  BRUTE sums over every particle in
 * ascending j (the reference's order); GRID sums over the 27 cells around
 * p in (cell, index) order.  Queries outside the box take the clamped
 * boundary cell, whose neighbourhood still holds every particle within h. */
void FN(orc_density_at_)(const orc_params *p, int mode, size_t n, const REAL *x, size_t m,
                         const REAL *pts, REAL *out) {
  DER d;
  FN(der_init_)(p, (REAL)p->dt, &d);
  if (mode == ORC_NBR_BRUTE) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long q = 0; q < (long)m; q++) {
      REAL s = 0;
      for (size_t j = 0; j < n; j++) s = s + FN(poly6_)(&d, FN(dist2_)(x + 3 * j, pts + 3 * q));
      out[q] = s;
    }
    return;
  }
  CAT(orc_grid, SFX) g;
  if (FN(grid_build_)(&d, n, x, &g) != 0) return;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < (long)m; q++) {
    REAL s = 0;
    GRID_FOREACH(&d, &g, pts + 3 * q, j,
                 { s = s + FN(poly6_)(&d, FN(dist2_)(x + 3 * (size_t)j, pts + 3 * q)); });
    out[q] = s;
  }
  FN(grid_free_)(&g);
}

#undef DER
#undef FN
#undef CAT
#undef CAT2
