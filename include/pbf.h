/*
 * pbf.h -- C ABI of the MI355X Position Based Fluids step.
 *
 * This is the drop-in boundary underneath the reference's `Particles`
 * load/step interface (src/particles.h:105-140).  Plain C types only: host
 * pointers to float arrays, sizes, and an opaque context.  All entry points
 * return PBF_OK (0) or a negative PBF_ERR_* code and never exit; the message
 * of the last failure is available from pbf_last_error().
 *
 * Arrays exchanged with the host are in load order ("particle id" order):
 * particle i of pbf_load is entry i of every read-back.  Positions and
 * velocities are n*3 floats (x, y, z interleaved).
 *
 * One context is not re-entrant.  Every entry point is synchronous at
 * return (work runs on the context's own HIP stream).
 */
#ifndef PBF_H
#define PBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_API __attribute__((visibility("default")))

enum {
  PBF_OK = 0,
  PBF_ERR_INVALID = -1,   /* bad argument (NULL, n = 0, non-finite, dt <= 0 ...) */
  PBF_ERR_HIP = -2,       /* HIP runtime failure or no device */
  PBF_ERR_OVERFLOW = -3,  /* a particle had more neighbours than max_neighbors */
  PBF_ERR_STATE = -4,     /* call out of order (e.g. step before load) */
  PBF_ERR_OUT_OF_BOX = -5,/* a loaded particle lies outside the collision box */
  PBF_ERR_NOMEM = -6      /* device or host allocation failed */
};

/* Physical and numerical parameters.  The reference fixes these as
 * compile-time macros (src/particles.cpp:7-48); here they are run-time
 * values.  pbf_params_default() fills the reference's own values. */
typedef struct pbf_params {
  float h;               /* smoothing radius H (particles.cpp:26) */
  float rest_density;    /* rho0 (Particles(double rest_density), particles.h:114) */
  int32_t iterations;    /* solver iterations NEWTON_NUM_STEPS (:34) */
  float epsilon;         /* constraint relaxation EPSILON (:36) */
  float k_tensile;       /* artificial pressure K (:38) */
  int32_t n_tensile;     /* artificial pressure exponent N (:40) */
  float dq;              /* tensile reference distance / h (:151, 0.1) */
  float xsph_c;          /* XSPH viscosity C (:44) */
  float vort_eps;        /* vorticity confinement VORTICITY_EPSILON (:42) */
  float gravity;         /* acceleration along -y (:180, 10) */
  float dt;              /* default time step DEFAULT_DELTA_T (:24) */
  float box_lo[3];       /* collision box, lower corner */
  float box_hi[3];       /* collision box, upper corner */
  uint32_t wall_stop_mask; /* bit (2*axis + side): that wall is an analytic
                              stop plane (no slide, strict crossing); the
                              other walls behave as the reference's mesh walls */
  float cell_scale;      /* grid cell edge = h * cell_scale, >= 1 */
  int32_t max_neighbors; /* neighbour-list capacity per particle, rounded up
                            to a multiple of 4; more neighbours fail the step
                            with PBF_ERR_OVERFLOW */
  float poly6_coef;      /* 315/(64 pi) as the reference literal (:140) */
  float spiky_coef;      /* 15/pi as the reference literal (:148) */
  float eps_d;           /* EPS_D (CGL/include/CGL/misc.h:11) */
} pbf_params;

typedef struct pbf_ctx pbf_ctx;

/* Fills the reference's constants: H=0.3, 12 iterations, EPSILON=2, K=1e-4,
 * N=4, dq=0.1, C=1e-3, VORTICITY_EPSILON=1e-3, g=10, dt=0.016, the Cornell
 * box [-1,1]x[0,1.49]x[-1,1] with stop planes y=1.49 and z=1
 * (src/particles.cpp:7-48, :59-83). */
PBF_API void pbf_params_default(pbf_params *p);

/* Creates a context on the current HIP device.  Replaces the reference's
 * constructor Particles(double rest_density) (src/particles.h:114). */
PBF_API int pbf_create(const pbf_params *p, pbf_ctx **out);

/* Multi-scene context: nscenes independent scenes with counts[s] particles
 * and boxes box_lo/box_hi[3*s..3*s+2]; all scenes share the physics of p.
 * Loaded state is the concatenation of the scenes in order.  No counterpart
 * in the reference (batched extension). */
PBF_API int pbf_create_batch(const pbf_params *p, int32_t nscenes, const int64_t *counts,
                             const float *box_lo, const float *box_hi, pbf_ctx **out);

PBF_API int pbf_destroy(pbf_ctx *ctx);

/* Uploads n particles (positions xyz, velocities v, each n*3 floats; v may
 * be NULL for zero).  Replaces the reference's repeated addParticle
 * (src/particles.h:118-120, called from src/application.cpp:334-340).
 * Resets simulate_time to 0. */
PBF_API int pbf_load(pbf_ctx *ctx, const float *xyz, const float *v, int64_t n);

/* Advances nsteps steps of length dt.  Replaces Particles::timeStep(double)
 * (src/particles.cpp:250-297). */
PBF_API int pbf_step_dt(pbf_ctx *ctx, float dt, int32_t nsteps);

/* Advances nsteps steps of the default dt.  Replaces Particles::timeStep()
 * (src/particles.cpp:299-301). */
PBF_API int pbf_step(pbf_ctx *ctx, int32_t nsteps);

/* Triangle-mesh collider, in place of the box's mesh walls: the reference's
 * scene BVH that clamp/clamp_response ray-cast against (src/particles.cpp:
 * 51-132, Particles::bvh set by PathTracer::build_accel, src/pathtracer.cpp:
 * 266; traversal src/bvh.cpp:142-192; triangle test src/static_scene/
 * triangle.cpp:21-80).  pos, nrm: nv*3 vertex positions and normals (the
 * mesh's vertex normals, interpolated w n1 + u n2 + v n3 at the hit as the
 * reference does); tri: nt*3 vertex indices.  With a mesh, a displacement
 * stops at the nearest triangle hit with 0 <= t <= l (ties: lowest triangle
 * index) at t - EPS_D, then steps `skin` off the face along its normal,
 * and prediction slides once along it; the stop planes of wall_stop_mask
 * and the final per-axis clamp to the box stay.  skin >= 0: the reference
 * uses EPS_D alone (1e-11, meaningful in FP64); in FP32 a skin well above
 * the rounding of the coordinates (e.g. 2^-17 of the scene size) keeps
 * particles from ending on or behind a face and passing through it on the
 * next move; skin 0 keeps the reference's formula.  nt = 0 removes the
 * mesh (back to the box walls).  Single-scene contexts only. */
PBF_API int pbf_set_mesh(pbf_ctx *ctx, const float *pos, const float *nrm, int64_t nv,
                         const int32_t *tri, int64_t nt, float skin);

/* Execution schedules (no effect on the arithmetic: every schedule runs the
 * same per-particle code in the same order and gives the same bits).
 * No counterpart in the reference (batched extension). */
enum {
  PBF_SCHED_FLAT = 0,       /* one launch per pass over all scenes (default) */
  PBF_SCHED_GRAPH = 1,      /* the flat launches of a step captured once in a
                               hipGraph and replayed for every step of a call
                               but its first and last (plain launches) */
  PBF_SCHED_PERSISTENT = 2  /* one launch for all steps: each workgroup runs
                               whole scenes, largest first, all passes */
};
/* Selects the schedule; block is the workgroup size of the persistent
 * schedule (256, 512 or 1024; ignored otherwise). */
PBF_API int pbf_set_schedule(pbf_ctx *ctx, int32_t schedule, int32_t block);

/* Enqueues nsteps steps on the context stream and returns without waiting.
 * Neighbour-list overflow is reported by the next pbf_sync. */
PBF_API int pbf_step_async(pbf_ctx *ctx, float dt, int32_t nsteps);
/* Waits for the context stream; reports overflow since the last sync. */
PBF_API int pbf_sync(pbf_ctx *ctx);

/* Copies the committed state back to the host in load order; any pointer
 * may be NULL.  rho is the SPH density of the last step's XSPH pass (the
 * value the reference keeps in Particle::density, src/particles.cpp:222-231).
 * Replaces the reads of ps[i]->getPosition() / velocity / density by the
 * CPU consumers (src/particles.cpp:305, :449, src/particles.h:36-52).
 * Arrays of 4 MB or more return through two 64 MB pinned host slabs,
 * allocated on first use and kept by the context until pbf_destroy. */
PBF_API int pbf_read(pbf_ctx *ctx, float *xyz, float *v, float *rho);

/* Load-time density including self, the reference's
 * Particles::estimateDensities (src/particles.cpp:440-444 -> :158-163),
 * evaluated on the grid.  Writes n floats in load order. */
PBF_API int pbf_estimate_densities(pbf_ctx *ctx, float *rho);

/* Checkpoint / resume.  The reference keeps state only in its XML input
 * (src/application.cpp:302-344, ~100 bytes of text per particle); a binary
 * snapshot restores a run exactly: loading a snapshot into a context made
 * with the same parameters (and, batched, the same scenes) and stepping on
 * gives the bits of the uninterrupted run.  File layout (little-endian):
 * a 64-byte header {char magic[8] = "PBFSNAP2"; uint32 header_bytes = 64;
 * uint32 params_bytes = sizeof(pbf_params); int64 n; int32 nscenes;
 * int32 0; double simulate_time; uint64 checksum of the payload;
 * 16 zero bytes}, then the payload: pbf_params, int64 count[nscenes],
 * float box_lo[3], box_hi[3] per scene, float x[n*3], v[n*3], rho[n]
 * (load order).  The checksum runs the FNV-1a-64 step (basis
 * 14695981039346656037, prime 1099511628211) on 64-bit little-endian
 * words in four lanes: word l of every 32-byte block of the payload goes to
 * lane l, the bytes after the last whole block go byte by byte into lane 0,
 * and the checksum is byte-serial FNV-1a-64 of the four lane values (32
 * bytes, lane 0 first).  Files with magic "PBFSNAP1" (byte-serial FNV-1a
 * of the payload from the basis 1469598103934665603) still load.  Loading checks magic, sizes, checksum,
 * parameters and scenes (PBF_ERR_INVALID on any mismatch) and restores
 * simulate_time. */
PBF_API int pbf_save_snapshot(pbf_ctx *ctx, const char *path);
PBF_API int pbf_load_snapshot(pbf_ctx *ctx, const char *path);

/* Density field of the committed particles, the reference's
 * Particles::estimateDensityAt (src/particles.cpp:446-453) as called by the
 * marching-cubes consumers generateGridCell (:309-323, 8 corners per cube)
 * and getVertexNormal (:404-418, 6 points per vertex).  out[q] = sum over
 * the particles of scene `scene` (0 unless batched) of poly6(x_j - p_q),
 * evaluated on the cell grid: particles of the 27 cells around p_q in
 * (cell, id) order.  pts: m*3 floats.  Like pbf_estimate_densities it
 * re-sorts the committed state (the positions do not change).  The first
 * call allocates device buffers for m results and m points (grown when a
 * later call needs more), so repeated small queries allocate nothing;
 * results of 4 MB or more return through the pinned slabs of pbf_read.
 * The context keeps these until pbf_destroy. */
PBF_API int pbf_density_at(pbf_ctx *ctx, int32_t scene, const float *pts, int64_t m, float *out);

/* The same field on the lattice of marching-cubes corners: point (ix, iy,
 * iz) is (min(hi[0], lo[0] + ix*step), min(hi[1], lo[1] + iy*step),
 * min(hi[2], lo[2] + iz*step)) in float, for i < counts[axis]; out is
 * [counts[2]][counts[1]][counts[0]] (x fastest).  The reference's
 * getSurfacePrims (:350-390) samples corners xmin + ix*fStepSize and
 * min(xmax, x1 + fStepSize) for ix <= (int)((xmax-xmin)/fStepSize), i.e.
 * counts = (int)((max-min)/step) + 2 per axis. */
PBF_API int pbf_density_lattice(pbf_ctx *ctx, int32_t scene, const float lo[3], const float hi[3],
                                float step, const int32_t counts[3], float *out);

/* Steps caller-owned device arrays without allocating: xyz and v are
 * DEVICE pointers to n*3 floats each (load order), advanced in place by
 * nsteps steps of length dt with the physics, box and grid of ctx (a
 * single-scene context from pbf_create; its own loaded state, if any, and
 * its simulate_time are left as they are).  The results are the bits
 * pbf_load + pbf_step_dt + pbf_read give for the same input.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops).  Every
 * buffer of the step lives in one fixed 64 MiB device array of the library
 * (each device holds its own copy while the library is loaded); n beyond what it
 * holds at these parameters (about 55,000 particles at the default
 * max_neighbors of 256, 115,000 at 96) fails with PBF_ERR_INVALID and a message giving the
 * limit.  No hipMalloc or hipFree is called.  All calls on one device share
 * the workspace: calls that overlap on one device (two host threads) race.
 * Synchronous at return; bad input is reported like pbf_load
 * (PBF_ERR_INVALID for non-finite values, PBF_ERR_OUT_OF_BOX) before any
 * step runs, and list overflow like pbf_step_dt. */
PBF_API int pbf_step_device_noalloc(pbf_ctx *ctx, float *xyz, float *v, int64_t n, float dt,
                                    int32_t nsteps);

/* Particle::simulate_time of the reference (src/particles.h:109). */
PBF_API double pbf_simulate_time(const pbf_ctx *ctx);
PBF_API int64_t pbf_count(const pbf_ctx *ctx);

/* Message of the last failure on ctx (or of the last failed pbf_create*
 * when ctx is NULL).  Never NULL. */
PBF_API const char *pbf_last_error(const pbf_ctx *ctx);

/* ---- differentiable extension (no counterpart in the reference) --------
 * Reverse-mode derivatives of nsteps steps of one scene (collision treated
 * as the identity; pbf_adj_stats counts wall contacts).  The forward pass
 * keeps (x, v) every checkpoint_every steps; the backward pass replays each
 * segment once from its checkpoint, recording every step's sorted state,
 * neighbour lists and solver iterates in a device tape of checkpoint_every
 * slots (about (5 + 4 (iterations + 1)) * 16 + 4 (max_neighbors + 1) bytes
 * per particle each), then runs hand-written adjoint kernels of the lambda,
 * delta-p, XSPH and vorticity passes over the tape in reverse. */
typedef struct pbf_adj pbf_adj;
PBF_API int pbf_adj_create(const pbf_params *p, int64_t n, int32_t nsteps,
                           int32_t checkpoint_every, pbf_adj **out);
PBF_API int pbf_adj_destroy(pbf_adj *adj);
/* Runs the steps from (x0, v0) with tensile coefficient k_tensile; writes the
 * final state (xT, vT may be NULL).  n*3 floats each, load order. */
PBF_API int pbf_adj_forward(pbf_adj *adj, const float *x0, const float *v0, float k_tensile,
                            float *xT, float *vT);
/* Given dL/dx_T (and dL/dv_T, may be NULL), writes dL/dx_0, dL/dv_0 (n*3
 * each, may be NULL) and dL/dk. */
PBF_API int pbf_adj_backward(pbf_adj *adj, const float *xbar_T, const float *vbar_T,
                             float *xbar_0, float *vbar_0, double *kbar);
/* Wall time of the last backward pass (HIP events, ms) and the number of
 * particle-iterations found on a wall during it. */
PBF_API int pbf_adj_stats(const pbf_adj *adj, double *backward_ms, int64_t *wall_contacts);
/* The backward pass replays each checkpoint segment and checks that the
 * replay ends where the forward pass was at the next checkpoint (at its end
 * for the last segment).  Writes the number of particle-checks of the last
 * backward pass whose x or v differed in any bit (0: the recorded tape is
 * the forward pass's trajectory). */
PBF_API int pbf_adj_replay_check(const pbf_adj *adj, int64_t *mismatched);
PBF_API const char *pbf_adj_last_error(const pbf_adj *adj);

/* ---- test hooks (inspection of the last step; no counterpart) ---------- */

/* Neighbour lists built by the last step, in load order: counts[n]; ids
 * receives, for particle i, its neighbours' load-order ids in the order the
 * passes visit them, at ids[i*stride .. i*stride+counts[i]).  Neighbour sets
 * are those of src/particles.cpp:258-265. */
PBF_API int pbf_neighbors(pbf_ctx *ctx, uint32_t *counts, uint32_t *ids, int32_t stride);

/* Sum and maximum of the neighbour counts of the last step (measurement). */
PBF_API int pbf_neighbor_stats(pbf_ctx *ctx, uint64_t *total, uint32_t *max_count);

/* Grid cell index of each particle at the last neighbour build, load order. */
PBF_API int pbf_cells(pbf_ctx *ctx, uint32_t *cells);

/* Grid dimensions of scene s (nx, ny, nz). */
PBF_API int pbf_grid_dims(const pbf_ctx *ctx, int32_t scene, int32_t *dims);

/* Runs one step but stops after `stage` (see PBF_STAGE_*), leaving the
 * intermediate buffers readable through pbf_debug_read.  The context must be
 * reloaded before the next pbf_step. */
enum {
  PBF_STAGE_PREDICT = 1,   /* x* and v after applyForceVelocity */
  PBF_STAGE_NEIGHBORS = 2, /* + neighbour lists */
  PBF_STAGE_LAMBDA0 = 3,   /* + lambda of iteration 0 */
  PBF_STAGE_DELTAP0 = 4,   /* + x* after iteration 0 */
  PBF_STAGE_SOLVE = 5,     /* + all iterations, v = (x* - x)/dt */
  PBF_STAGE_XSPH = 6       /* + XSPH velocity, vorticity, density */
};
PBF_API int pbf_step_debug(pbf_ctx *ctx, float dt, int32_t stage);

enum {
  PBF_BUF_PRED = 1,   /* x* (xyz) and lambda (w) */
  PBF_BUF_VEL = 2,    /* current velocity (xyz) */
  PBF_BUF_VXSPH = 3,  /* velocity after XSPH (xyz), density (w) */
  PBF_BUF_OMEGA = 4,  /* vorticity (xyz) */
  PBF_BUF_POS = 5     /* committed position (xyz) */
};
/* Reads buffer `which` as n*4 floats in load order. */
PBF_API int pbf_debug_read(pbf_ctx *ctx, int32_t which, float *out4);

/* Passes of one step, for pbf_step_profiled. */
enum {
  PBF_PASS_PREDICT = 0,   /* gravity, prediction + collision, cell hash, histogram */
  PBF_PASS_SCAN = 1,      /* exclusive scan of the cell histogram */
  PBF_PASS_SORT = 2,      /* counting-sort scatter + in-cell rank + reorder */
  PBF_PASS_NEIGHBORS = 3, /* neighbour-list build */
  PBF_PASS_LAMBDA = 4,    /* density / lambda (one launch per iteration) */
  PBF_PASS_DELTAP = 5,    /* delta-p + collision (one launch per iteration) */
  PBF_PASS_XSPH = 6,      /* XSPH viscosity + vorticity */
  PBF_PASS_VORTICITY = 7, /* vorticity confinement + commit */
  PBF_NPASS = 8
};
/* Runs nsteps steps like pbf_step_dt, timing every pass with HIP events on
 * the context stream; pass_ms[p] receives the summed milliseconds of pass p
 * and pass_launches[p] (may be NULL) its number of launches.  Measurement
 * hook: the state advances exactly as with pbf_step_dt. */
PBF_API int pbf_step_profiled(pbf_ctx *ctx, float dt, int32_t nsteps, float *pass_ms,
                              int32_t *pass_launches, int32_t npass);

/* Derived constants in float, same order as the oracle's ORC_* indices
 * (h, h2, h^6, h^9, ...), and the number written (25).  With cap >= 28 it
 * also writes, as 0/1, whether the short division by h^9, reciprocal and
 * square root sequences passed the exhaustive device check at create time
 * and are in use (28).  Setting PBF_IEEE_MATH in the environment before
 * pbf_create keeps the IEEE expansions; both give the same bits. */
PBF_API int pbf_debug_derived(const pbf_ctx *ctx, float *out, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
