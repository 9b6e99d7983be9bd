// pbf_internal.h -- context layout and kernel launchers shared by the C ABI
// (pbf_capi.hip) and the kernels (pbf_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_device.h"

namespace pbf {

// One scene of a multi-scene context as the persistent schedule sees it.
struct SceneRun {
  uint32_t scene;   // index into the grid table
  uint32_t offset;  // first particle
  uint32_t count;   // particles
  uint32_t ncells;  // cells of its grid
};

// Device state of one context.  All particle arrays are float4 and are kept
// in cell-sorted order between steps; posA.w carries the particle's load
// index (its id) as integer bits.
struct DevState {
  float4* posA = nullptr;   // committed x (xyz), id (w)
  float4* posB = nullptr;   // sorted copy of x for the current step
  float4* velA = nullptr;   // committed v (xyz), density (w)
  float4* velB = nullptr;   // sorted v for the current step
  float4* predP = nullptr;  // x* ping (xyz), lambda or |omega| (w)
  float4* predQ = nullptr;  // x* pong
  float* xs3 = nullptr;     // x* packed as 3 floats (12-byte records), for the gathers
  float4* vort = nullptr;   // omega (xyz)
  float4* vxsph = nullptr;  // v after XSPH (xyz), density (w)
  float4* xv = nullptr;     // (x*, v) pairs after the solve, 2 per particle
  uint32_t* cell = nullptr;       // cell of particle, unsorted
  uint32_t* arrival = nullptr;    // arrival rank within its cell
  uint32_t* cellSorted = nullptr; // cell of particle, sorted
  uint32_t* tmpId = nullptr;      // ids scattered by cell, unordered within a cell
  uint32_t* cellCount = nullptr;  // C
  uint32_t* cellStart = nullptr;  // C + 1, exclusive prefix
  uint32_t* blockSums = nullptr;  // scan partials
  uint32_t* nbr = nullptr;        // tiled ELL neighbour lists, npad * K
  uint32_t* ncnt = nullptr;       // neighbour counts
  uint32_t* flags = nullptr;      // [0] max neighbour count seen beyond K
  SceneGrid* grids = nullptr;     // per-scene grids (multi-scene)
  uint32_t* slotScene = nullptr;  // scene of each sorted slot (multi-scene)
  SceneRun* runs = nullptr;       // scenes, largest first (persistent schedule)
  float* stage = nullptr;         // read-back staging, n * 4 floats
  float4* meshTri = nullptr;      // collider triangles, 6 float4 each (pbf_device.h MeshView)
  float4* meshNode = nullptr;     // collider BVH, 2 float4 per node
  uint8_t* meshFlag = nullptr;    // scene-grid cells near a BVH leaf box (MeshView::flag)
  uint32_t* dense = nullptr;      // particles with long lists, queued by the build (n)
  uint32_t* dense_b = nullptr;    // particles in crowded cells, queued by the sort (n)
};

// Lanes per particle of each list pass, by state size: one wave (64 lanes)
// for states of at most wave_max particles, one DPP quad (4 lanes; single
// scenes; for the vorticity pass that carries the next step's prediction
// only up to kVortPredictQuadMax) up to quad_max, else one lane.
// Same bits in every case.  The bounds sit between the sizes where the
// kernel time per launch crosses over (rocprofv3 kernel traces of lattice
// blocks at 0.5 h spacing, 4 iterations; microseconds per launch, wave /
// quad / lane; the quad build's figure includes its separate first lambda
// launch):
//   build + first lambda   8,192: 13.6 / 22.6 / 42.6  16,384: 20.3 / 23.6 / 43.6
//                         24,576: 27.2 / 25.4 / 44.5  65,536: -    / 32.7 / 46.9
//                        131,072: -    / 55.6 / 54.4
//   lambda                 8,192:  5.1 / 6.7 / 12.5   16,384:  7.3 / 6.8 / 12.5
//                         65,536:  -   / 10.3 / 13.0 131,072:  -   / 16.1 / 15.0
//   delta-p                8,192:  6.7 / 6.9 / 13.0   16,384:  9.9 / 6.9 / 12.9
//                         65,536:  -   / 10.6 / 13.8 131,072:  -   / 17.3 / 15.4
//   XSPH                   8,192:  5.5 / 7.8 / 13.9   16,384:  8.3 / 7.7 / 13.9
//                         65,536:  -   / 12.6 / 15.2 131,072:  -   / 21.2 / 19.1
//   vorticity              8,192: 10.3 / -   / 11.6   16,384: 13.8 / -    / 11.9
//   vorticity commit      16,384:  -   / 5.7 / 10.9   32,768:  -   / 6.6 / 11.3
//                         65,536:  -   / 8.9 / 11.9   98,304:  -   / 13.6 / 12.6
// PBF_WIDE_MAX (every pass's wave_max), PBF_QUAD_MAX (every quad_max) and
// PBF_RULE_<PASS>=wave_max,quad_max override them at context creation; 0
// turns a variant off.
enum ListPass : int { kPassBuild, kPassLambda, kPassDeltap, kPassXsph, kPassVort, kListPasses };
struct PassRule {
  int64_t wave_max;
  int64_t quad_max;
};
constexpr PassRule kPassRules[kListPasses] = {
    {20480, 114688},  // build + first lambda
    {12288, 114688},  // lambda
    {8192, 114688},   // delta-p
    {12288, 114688},  // XSPH
    {12288, 81920},   // vorticity (+ commit, + prediction)
};
// The vorticity pass that carries the next step's prediction takes its quad
// only up to this many particles: its histogram slots cost a quad one
// atomic per particle, where a lane per particle takes one per run of equal
// cells (microseconds per launch, quad / lane: 16,384 9.4 / 12.1, 32,768
// 11.8 / 12.2, 65,536 17.5 / 12.5).
constexpr int64_t kVortPredictQuadMax = 24576;

// Dense split: in single-scene states of at most kDenseMaxN particles whose
// list build runs one lane per particle, the build queues the particles
// with more than dense_list neighbours (count in flags[kDenseSlot]), and the
// per-lane lambda, delta-p and XSPH launches leave them to quad workgroups
// appended to the same launch.  A quad walks a list as four chunk streams
// at once (quad_sums), so a compressed region's longest lists take a
// quarter of the dependent rounds of the lone waves that otherwise set
// these launches' length; the other particles keep one lane each.  Same
// bits (the quad sums are the per-lane sums).  PBF_DENSE_LIST overrides the
// threshold (0: off).  Measured on the 262,144-particle dam break (200
// steps): 1.32 ms per step without, 0.97 at 128 (0.99 at 64, 1.07 at 256),
// 0.93 with the build split below as well; uniform blocks of 131,072 and
// 262,144 particles (no list above 128, no cell above 32) within 0.6%; the
// 4,096,000-particle run 0.6% slower, hence the bound on n.
constexpr uint32_t kDenseList = 128;
constexpr int64_t kDenseMaxN = int64_t{1} << 20;
constexpr int kDenseSlot = 12;
// The list build is split the same way, by what is known before it: the
// sort queues the particles whose own cell holds more than dense_cell
// particles (count in flags[kDenseBuildSlot]), and quad workgroups appended
// to the per-lane build + first lambda launch build their lists with a quad
// each (k_nbr_quad's walk) and run their first lambda on the same quad.
// PBF_DENSE_CELL overrides the threshold (0: off).
constexpr uint32_t kDenseCell = 32;
constexpr int kDenseBuildSlot = 13;

struct Ctx {
  pbf_params p{};
  int nscenes = 1;
  std::vector<int64_t> counts;    // particles per scene
  std::vector<int64_t> offsets;   // first particle of each scene
  std::vector<SceneGrid> grids;   // host copy
  int64_t n = 0;                  // total particles
  int64_t npad = 0;               // n rounded up to 64
  uint32_t ncell = 0;             // total cells
  uint32_t nblocks_scan = 0;
  int block = 256;                // threads per block of the particle kernels
  hipStream_t stream = nullptr;
  int device = 0;
  DevState d;
  size_t alloc_n = 0, alloc_cells = 0;
  double sim_time = 0.0;
  bool loaded = false;
  bool debug_dirty = false;
  bool has_nbrs = false;
  bool field_ready = false;          // committed state sorted by cell, cell table current
  uint32_t mesh_ntri = 0;            // triangles of the mesh collider (0: box walls)
  uint32_t mesh_nnode = 0;           // BVH nodes
  float mesh_skin = 0.0f;            // clearance a mesh hit leaves along the face normal
  float mesh_margin = 0.0f;          // MeshView::margin
  const float4* dbg_pred = nullptr;  // x* buffer current at the last debug stop
  const float4* dbg_vel = nullptr;   // velocity buffer current at the last debug stop
  int schedule = PBF_SCHED_FLAT;     // PBF_SCHED_*
  int pblock = 256;                  // workgroup size of the persistent schedule
  int ncu = 0;                       // compute units of the device
  int fast_div = 0;                  // short division by h^9 verified exact
  int fast_rcp = 0;                  // short reciprocal verified exact
  int fast_sqrt = 0;                 // short square root verified exact
  bool small_sort = true;            // single-workgroup scan + sort for small states
  PassRule rules[kListPasses] = {kPassRules[0], kPassRules[1], kPassRules[2], kPassRules[3],
                                 kPassRules[4]};
  uint32_t dense_list = kDenseList;  // dense split threshold (0: off)
  uint32_t dense_cell = kDenseCell;  // build split threshold (0: off)
  // Whether the prediction stores the advanced velocity in velA and the sort
  // moves it into velB with the particle.  A full step reads neither (its
  // velocity comes from x* - x and the commit rewrites velA), so run_step
  // clears this for full steps; the debug stops and the adjoint keep it.
  bool sorted_v = true;
  // Density-field output on the device, kept between calls, and two pinned
  // host slabs through which large results return (allocated on first use).
  float* field_dev = nullptr;
  size_t field_cap = 0;
  float* field_pts = nullptr;        // query points on the device, kept too
  size_t field_pts_cap = 0;          // points
  float* stage_host[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  uint32_t* flag_host = nullptr;     // pinned word the per-call overflow check reads into
  hipGraph_t graph = nullptr;        // captured step (PBF_SCHED_GRAPH)
  hipGraphExec_t graph_exec = nullptr;
  float graph_dt = 0.0f;
  hipGraph_t graph1 = nullptr;       // captured whole step, for calls of one step
  hipGraphExec_t graph1_exec = nullptr;
  float graph1_dt = 0.0f;
  bool step_graph = true;            // flat schedule: replay graph1 when dt repeats
  float last_dt1 = -1.0f;            // dt of the previous one-step call
  std::string err;
};

// Chunks of 4 neighbour slots per particle: the list capacity is
// max_neighbors rounded up to a multiple of 4.
inline int k4_of(const Ctx& x) { return (x.p.max_neighbors + 3) / 4; }

// Host-side derivation of the step constants, in float, in the order the
// reference evaluates them.
void derive_constants(const pbf_params& p, float dt, SimConst* c);
void derive_grid(const pbf_params& p, const float* lo, const float* hi, SceneGrid* g);

// Kernel launchers (pbf_kernels.hip).  Each returns hipSuccess or the launch
// error.
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
hipError_t launch_predict(Ctx& x, const SimConst& c);
hipError_t launch_scan(Ctx& x);
// Counting-sort scatter and (cell, id) rank; pred is where the unsorted
// predictions are (NULL: predP).
hipError_t launch_scatter_rank(Ctx& x, const float4* pred = nullptr);
// Scan, scatter and rank in one single-workgroup launch, for states of at
// most kSortSmallMax particles over at most kSortSmallCells cells (measured:
// 8 us at 720 particles, 18 us at 2,106, but 36 us at 4,096, where the five
// launches of the general path take less).
constexpr int64_t kSortSmallMax = 2560;
constexpr uint32_t kSortSmallCells = 1u << 15;
bool sort_in_one_workgroup(const Ctx& x);
hipError_t launch_sort_small(Ctx& x, const float4* pred = nullptr);
hipError_t launch_nbr_build(Ctx& x, const SimConst& c);
// packed: the x* records (xs3) hold in's xyz, as after a delta-p pass.
hipError_t launch_lambda(Ctx& x, const SimConst& c, const float4* in, float4* out, bool packed);
// Neighbour build and the first λ pass (in = predQ) in one launch.
hipError_t launch_nbr_lambda(Ctx& x, const SimConst& c, float4* out);
hipError_t launch_deltap(Ctx& x, const SimConst& c, const float4* in, float4* out, bool last);
// XSPH reads x* and v from the records the last delta-p wrote.
hipError_t launch_xsph(Ctx& x, const SimConst& c, float4* out);
// x* and v of the last iteration's records into float4 arrays.
hipError_t launch_unpack_record(Ctx& x, float4* pos, float4* vel);
// x* of the 12-byte records (after a delta-p pass before the last) into pos.
hipError_t launch_unpack_xs3(Ctx& x, float4* pos);
hipError_t launch_vort_commit(Ctx& x, const SimConst& c, const float4* in);
// Vorticity and commit fused with the next step's prediction (zeroes the
// histogram first); the next step then starts at the scan.
hipError_t launch_vort_commit_predict(Ctx& x, const SimConst& c, const float4* in);
hipError_t launch_unpermute(Ctx& x, const float4* ids, const float4* src, int first, int count,
                            float* dst);
hipError_t launch_unpermute_u32(Ctx& x, const float4* ids, const uint32_t* src, uint32_t* dst);
hipError_t launch_nbr_export(Ctx& x, uint32_t* counts, uint32_t* ids, int stride);
hipError_t launch_densities(Ctx& x, const SimConst& c, float* out);
// Sorts the committed state by cell (x into predQ/posB, v into velB) and
// builds the cell table, as the load-time density pass does.
hipError_t launch_sort_committed(Ctx& x, const SimConst& c);
// Lattice of density query points: (min(hi, lo + i*step), ...), i < n.
struct DensityLattice {
  float lo[3], hi[3], step;
  uint32_t n[3];
};
// Density field of scene `scene` at m points (pts, or the lattice when
// given); needs launch_sort_committed first.
hipError_t launch_density_field(Ctx& x, const SimConst& c, int scene, const float* pts,
                                const DensityLattice* lattice, uint32_t m, float* out);
hipError_t launch_persistent(Ctx& x, const SimConst& c, int nsteps, int block, int grid);
hipError_t launch_verify_math(float b, float rb, int div_e_lo, int div_e_hi, uint32_t* bad,
                              hipStream_t s);
// Step constants of a context, with its verified short-division flags.
SimConst consts_for(const Ctx& x, float dt);
int persistent_max_blocks_per_cu(int block);
hipError_t launch_nbr_stats(Ctx& x, unsigned long long* sum, uint32_t* mx);
// Exclusive scan of in[0..n) into out[0..n], out[n] = total; with clear_in
// the counts are zero afterwards.
hipError_t launch_scan_u32(uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials,
                           hipStream_t s, bool clear_in = false);

}  // namespace pbf
