/*
 * pbf_oracle.h -- CPU restatement of the reference Position Based Fluids step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product library never links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's
 * Particles::timeStep (src/particles.cpp:250-301) with the reference's
 * compile-time constants (src/particles.cpp:7-48) turned into run-time
 * parameters.  Every arithmetic expression keeps the reference's operation
 * order (CGL/include/CGL/vector3D.h:64-150: division of a vector by a scalar is
 * a multiply by the reciprocal, norm2 is (x*x+y*y)+z*z, intpow is repeated
 * multiplication, src/bsdf.h:49-57).  Each entry point exists twice:
 * suffix _f (IEEE single, the parity target of the GPU path) and _d (IEEE
 * double, used for the finite-difference gradient checks).
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp (see Makefile).
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical parameters.  Values are given in double and are
 * converted to the working precision (float or double) once, by
 * orc_derive_*. */
typedef struct orc_params {
  double h;             /* smoothing radius, reference H (particles.cpp:26) */
  double rest_density;  /* rho0, reference Particle::rest_density */
  int    iterations;    /* solver iterations, reference NEWTON_NUM_STEPS (:34) */
  double epsilon;       /* constraint relaxation, reference EPSILON (:36) */
  double k_tensile;     /* artificial pressure coefficient, reference K (:38) */
  int    n_tensile;     /* artificial pressure exponent, reference N (:40) */
  double dq;            /* tensile reference distance as a fraction of h (:151) */
  double xsph_c;        /* XSPH viscosity, reference C (:44) */
  double vort_eps;      /* vorticity confinement, reference VORTICITY_EPSILON (:42) */
  double gravity;       /* acceleration along -y, reference 10 (:180) */
  double dt;            /* time step, reference DEFAULT_DELTA_T (:24) */
  double box_lo[3];     /* collision box */
  double box_hi[3];
  unsigned wall_stop_mask; /* bit w set: wall w is an analytic stop plane
                              (w = 2*axis + (0 lo, 1 hi)); other walls are
                              mesh walls that slide once in prediction */
  double cell_scale;    /* grid cell edge = h * cell_scale (>= 1) */
  double poly6_coef;    /* 315/(64 pi) as the reference literal (:140) */
  double spiky_coef;    /* 15/pi as the reference literal (:148) */
  double eps_d;         /* EPS_D (CGL/include/CGL/misc.h:11) */
  /* Triangle-mesh collider in place of the box's mesh walls (mesh_nt 0:
     the box walls).  Vertex positions and normals (nv*3) and vertex
     indices (nt*3); converted to the working precision where used. */
  const double *mesh_pos;
  const double *mesh_nrm;
  const int32_t *mesh_tri;
  int64_t mesh_nt;
  double mesh_skin;      /* clearance a mesh hit leaves along the face normal */
} orc_params;

/* Derived constants in working precision, in the order they are computed. */
enum {
  ORC_H = 0, ORC_H2, ORC_H6, ORC_H9, ORC_C1, ORC_M3C2, ORC_TENSILE_S,
  ORC_INV_RHO0, ORC_RHO0, ORC_EPS, ORC_NEGK, ORC_XSPH_C, ORC_DT_VORT,
  ORC_INV_DT, ORC_DT, ORC_GDT, ORC_CELL, ORC_INV_CELL, ORC_EPS_D,
  ORC_LO_E0, ORC_LO_E1, ORC_LO_E2, ORC_HI_E0, ORC_HI_E1, ORC_HI_E2,
  ORC_NDERIVED
};

/* Neighbour-list modes. */
enum { ORC_NBR_GRID = 0,   /* uniform grid, list in (cell, index) order */
       ORC_NBR_BRUTE = 1 };/* all pairs, ascending index: the reference */
/* XSPH modes. */
enum { ORC_XSPH_JACOBI = 0,      /* all velocities first, then XSPH */
       ORC_XSPH_SEQUENTIAL = 1 };/* the reference's in-place sweep */

void orc_default_params(orc_params *p);

/* Derived constants, grid dimensions. */
void orc_derive_f(const orc_params *p, float *out /*ORC_NDERIVED*/, int *dims /*3*/);
void orc_derive_d(const orc_params *p, double *out, int *dims);

/* Kernel functions (K1, K2). */
float  orc_poly6_f(const orc_params *p, float r2);
double orc_poly6_d(const orc_params *p, double r2);
void   orc_spiky_f(const orc_params *p, const float *r, float *g);
void   orc_spiky_d(const orc_params *p, const double *r, double *g);

/* Collision of one particle (C2 with slide, C1 without). p and v in place. */
void orc_clamp_response_f(const orc_params *p, float *pos, const float *vel, float dt);
void orc_clamp_response_d(const orc_params *p, double *pos, const double *vel, double dt);
void orc_clamp_f(const orc_params *p, float *pos, const float *dp);
void orc_clamp_d(const orc_params *p, double *pos, const double *dp);

/* Cell index of each position (row-major (z*ny+y)*nx+x). */
void orc_cells_f(const orc_params *p, size_t n, const float *xs, uint32_t *cell);

/* Neighbour lists in CSR form.  offsets has n+1 entries; list must hold
 * cap entries.  Returns the number of entries, or -1 if cap is too small. */
long orc_neighbors_f(const orc_params *p, int mode, size_t n, const float *xs,
                     uint32_t *offsets, uint32_t *list, size_t cap);
long orc_neighbors_d(const orc_params *p, int mode, size_t n, const double *xs,
                     uint32_t *offsets, uint32_t *list, size_t cap);

/* Single passes over a given neighbour list (x arrays are n*3). */
void orc_predict_f(const orc_params *p, size_t n, const float *x, float *v, float *xs, float dt);
void orc_predict_d(const orc_params *p, size_t n, const double *x, double *v, double *xs, double dt);
void orc_lambda_f(const orc_params *p, size_t n, const float *xs, const uint32_t *off,
                  const uint32_t *list, float *rho, float *lambda);
void orc_lambda_d(const orc_params *p, size_t n, const double *xs, const uint32_t *off,
                  const uint32_t *list, double *rho, double *lambda);
void orc_deltap_f(const orc_params *p, size_t n, const float *xs, const float *lambda,
                  const uint32_t *off, const uint32_t *list, float *xs_out);
void orc_deltap_d(const orc_params *p, size_t n, const double *xs, const double *lambda,
                  const uint32_t *off, const uint32_t *list, double *xs_out);
/* P6: v (in) is the velocity before updateVelocity (only read in the
 * sequential mode for j > i); writes v_out (after XSPH), omega, rho. */
void orc_xsph_f(const orc_params *p, int mode, size_t n, const float *x, const float *xs,
                const float *v_pred, const uint32_t *off, const uint32_t *list,
                float *v_out, float *omega, float *rho, float dt);
void orc_xsph_d(const orc_params *p, int mode, size_t n, const double *x, const double *xs,
                const double *v_pred, const uint32_t *off, const uint32_t *list,
                double *v_out, double *omega, double *rho, double dt);
/* P7: vorticity confinement and commit x <- xs. */
void orc_vorticity_f(const orc_params *p, size_t n, const float *xs, const float *omega,
                     const uint32_t *off, const uint32_t *list, float *v, float *x, float dt);
void orc_vorticity_d(const orc_params *p, size_t n, const double *xs, const double *omega,
                     const uint32_t *off, const uint32_t *list, double *v, double *x, double dt);

/* Whole step (Particles::timeStep).  x, v: n*3 in/out; rho: n out (may be
 * NULL).  Returns 0, or -1 on allocation failure. */
int orc_step_f(const orc_params *p, int nbr_mode, int xsph_mode, size_t n,
               float *x, float *v, float *rho, float dt, int nsteps);
int orc_step_d(const orc_params *p, int nbr_mode, int xsph_mode, size_t n,
               double *x, double *v, double *rho, double dt, int nsteps);

/* Load-time density (Particles::estimateDensities, includes self). */
void orc_estimate_densities_f(const orc_params *p, int mode, size_t n, const float *x, float *rho);
void orc_estimate_densities_d(const orc_params *p, int mode, size_t n, const double *x, double *rho);

/* Density field at m query points pts[m*3] (Particles::estimateDensityAt,
 * src/particles.cpp:446-453; self not special: a query at a particle's
 * position includes its W(0)). */
void orc_density_at_f(const orc_params *p, int mode, size_t n, const float *x, size_t m,
                      const float *pts, float *out);
void orc_density_at_d(const orc_params *p, int mode, size_t n, const double *x, size_t m,
                      const double *pts, double *out);

/* Threads used by the OpenMP loops (0 = runtime default). */
void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
