/*
 * pbf_oracle.c -- instantiates the CPU restatement (pbf_oracle_impl.h) in
 * single and double precision.  TEST INFRASTRUCTURE ONLY (see pbf_oracle.h).
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants of the reference, src/particles.cpp:7-48 and :151, the Cornell
 * box of dae/sky/CBempty.dae as the clamp bounds of src/particles.cpp:81-83,
 * and EPS_D of CGL/include/CGL/misc.h:11. */
void orc_default_params(orc_params *p) {
  memset(p, 0, sizeof(*p));
  p->h = 0.3;
  p->rest_density = 1000.0;
  p->iterations = 12;
  p->epsilon = 2.0;
  p->k_tensile = 0.0001;
  p->n_tensile = 4;
  p->dq = 0.1;
  p->xsph_c = 0.001;
  p->vort_eps = 0.001;
  p->gravity = 10.0;
  p->dt = 0.016;
  p->box_lo[0] = -1.0; p->box_lo[1] = 0.0;  p->box_lo[2] = -1.0;
  p->box_hi[0] = 1.0;  p->box_hi[1] = 1.49; p->box_hi[2] = 1.0;
  p->wall_stop_mask = (1u << 3) | (1u << 5); /* y = 1.49 and z = 1 planes */
  p->cell_scale = 1.0 + 1.0 / 1024.0;
  p->poly6_coef = 1.56668147106;
  p->spiky_coef = 4.774648292756860;
  p->eps_d = 0.00000000001;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define REAL float
#define SFX f
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#include "pbf_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FLOOR
#undef FABS

#define REAL double
#define SFX d
#define SQRT sqrt
#define FLOOR floor
#define FABS fabs
#include "pbf_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FLOOR
#undef FABS
