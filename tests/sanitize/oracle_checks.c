/*
 * Runs the CPU restatement (oracle/) through every mode on small scenes;
 * built with AddressSanitizer and UBSan by tests/test_sanitize.py.  Checks
 * only that everything runs to the end with finite results (the values are
 * pinned by tests/test_oracle.py); a sanitizer finding aborts.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pbf_oracle.h"

static int finite_all(const double *a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!isfinite(a[i])) return 0;
  return 1;
}

static int finite_allf(const float *a, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!isfinite(a[i])) return 0;
  return 1;
}

int main(void) {
  orc_params p;
  orc_default_params(&p);
  /* A 9x9x9 block near the floor of the reference's Cornell box, falling. */
  const size_t side = 9, n = side * side * side;
  double *x = malloc(3 * n * sizeof(double)), *v = malloc(3 * n * sizeof(double));
  double *rho = malloc(n * sizeof(double));
  float *xf = malloc(3 * n * sizeof(float)), *vf = malloc(3 * n * sizeof(float));
  float *rhof = malloc(n * sizeof(float));
  if (!x || !v || !rho || !xf || !vf || !rhof) return 2;
  int fail = 0;
  for (int mesh = 0; mesh < 2; mesh++) {
    /* The mesh case: the box floor and one wall as four triangles. */
    static const double mpos[] = {-1, 0.02, -1, 1, 0.02, -1, 1, 0.02, 1, -1, 0.02, 1,
                                  -0.98, 0, -1, -0.98, 1.4, -1, -0.98, 1.4, 1, -0.98, 0, 1};
    static const double mnrm[] = {0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0,
                                  1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
    static const int32_t mtri[] = {0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7};
    p.mesh_pos = mesh ? mpos : NULL;
    p.mesh_nrm = mesh ? mnrm : NULL;
    p.mesh_tri = mesh ? mtri : NULL;
    p.mesh_nt = mesh ? 4 : 0;
    p.mesh_skin = mesh ? 1e-6 : 0.0;
    for (int nbr = 0; nbr < 2; nbr++)
      for (int xsph = 0; xsph < 2; xsph++) {
        for (size_t i = 0; i < n; i++) {
          const size_t a = i % side, b = (i / side) % side, c = i / (side * side);
          x[3 * i] = -0.6 + 0.15 * (double)a + 1e-4 * (double)(i % 7);
          x[3 * i + 1] = 0.05 + 0.15 * (double)b;
          x[3 * i + 2] = -0.6 + 0.15 * (double)c - 1e-4 * (double)(i % 5);
          v[3 * i] = -2.0;
          v[3 * i + 1] = -1.0;
          v[3 * i + 2] = 0.5;
          for (int k = 0; k < 3; k++) {
            xf[3 * i + k] = (float)x[3 * i + k];
            vf[3 * i + k] = (float)v[3 * i + k];
          }
        }
        if (orc_step_d(&p, nbr, xsph, n, x, v, rho, p.dt, 12) != 0) fail = 1;
        if (orc_step_f(&p, nbr, xsph, n, xf, vf, rhof, (float)p.dt, 12) != 0) fail = 1;
        if (!finite_all(x, 3 * n) || !finite_all(v, 3 * n) || !finite_allf(xf, 3 * n) ||
            !finite_allf(vf, 3 * n)) {
          fprintf(stderr, "oracle_checks: non-finite state (mesh %d nbr %d xsph %d)\n", mesh, nbr,
                  xsph);
          fail = 1;
        }
      }
  }
  /* Neighbour lists with a capacity too small (the caller retries), then
   * enough; load-time densities and query points inside and outside. */
  p.mesh_nt = 0;
  uint32_t *off = malloc((n + 1) * sizeof(uint32_t)), *lst = malloc(64 * n * sizeof(uint32_t));
  if (!off || !lst) return 2;
  for (int mode = 0; mode < 2; mode++) {
    if (orc_neighbors_d(&p, mode, n, x, off, lst, 5) != -1) fail = 1;
    if (orc_neighbors_d(&p, mode, n, x, off, lst, 64 * n) < 0) fail = 1;
    orc_estimate_densities_d(&p, mode, n, x, rho);
    const double pts[] = {0, 0.5, 0, 3, 3, 3, -5, -5, -5, 0.9, 0.01, 0.9};
    double out[4];
    orc_density_at_d(&p, mode, n, x, 4, pts, out);
    if (!finite_all(rho, n) || !finite_all(out, 4)) fail = 1;
  }
  free(x); free(v); free(rho); free(xf); free(vf); free(rhof); free(off); free(lst);
  if (fail) return 1;
  printf("oracle_checks: all modes ran\n");
  return 0;
}
