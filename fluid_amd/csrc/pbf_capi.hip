// pbf_capi.hip -- the C ABI of include/pbf.h: context lifetime, validation,
// device allocation, step orchestration and read-back.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <type_traits>
#include <vector>

#include "pbf_host.h"
#include "pbf_internal.h"

using pbf::Ctx;
using pbf::SceneGrid;
using pbf::SimConst;
namespace host = pbf::host;
static_assert(sizeof(host::F4) == sizeof(float4) && alignof(float4) % alignof(host::F4) == 0,
              "host records have the device's float4 layout");

namespace {

thread_local std::string g_err = "no error";

int fail(Ctx* x, int code, const std::string& msg) {
  if (x) x->err = msg;
  g_err = msg;
  // A failed runtime call (an allocation that found no memory, say) stays
  // the thread's last HIP error; consume it here, or the next launch check
  // (hipGetLastError after a launch) would report it again.
  if (code == PBF_ERR_HIP || code == PBF_ERR_NOMEM) (void)hipGetLastError();
  return code;
}

#define HIPCHK(x, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((x), PBF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool is_finite_f(float v) { return isfinite(v) != 0; }

int validate_params(const pbf_params* p, std::string* why) {
  char buf[256];
  if (!p) {
    *why = "params is NULL";
    return PBF_ERR_INVALID;
  }
  const float vals[] = {p->h,        p->rest_density, p->epsilon,   p->k_tensile, p->dq,
                        p->xsph_c,   p->vort_eps,     p->gravity,   p->dt,        p->cell_scale,
                        p->poly6_coef, p->spiky_coef, p->eps_d};
  for (float v : vals)
    if (!is_finite_f(v)) {
      *why = "non-finite parameter";
      return PBF_ERR_INVALID;
    }
  if (!(p->h > 0.0f) || !(p->rest_density > 0.0f) || !(p->dt > 0.0f) || !(p->cell_scale >= 1.0f) ||
      !(p->epsilon >= 0.0f) || !(p->eps_d >= 0.0f)) {
    *why = "h, rest_density, dt must be > 0, cell_scale >= 1, epsilon and eps_d >= 0";
    return PBF_ERR_INVALID;
  }
  if (p->iterations < 1 || p->iterations > 1000 || p->n_tensile < 0 || p->n_tensile > 16 ||
      p->max_neighbors < 1 || p->max_neighbors > 4096) {
    snprintf(buf, sizeof buf,
             "iterations %d (1..1000), n_tensile %d (0..16), max_neighbors %d (1..4096)",
             p->iterations, p->n_tensile, p->max_neighbors);
    *why = buf;
    return PBF_ERR_INVALID;
  }
  return PBF_OK;
}

int validate_box(const float* lo, const float* hi, std::string* why) {
  for (int a = 0; a < 3; a++)
    if (!is_finite_f(lo[a]) || !is_finite_f(hi[a]) || !(hi[a] > lo[a])) {
      *why = "box must satisfy lo < hi on every axis";
      return PBF_ERR_INVALID;
    }
  return PBF_OK;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void free_device(Ctx* x) {
  pbf::DevState& d = x->d;
  dfree(d.posA); dfree(d.posB); dfree(d.velA); dfree(d.velB);
  dfree(d.predP); dfree(d.predQ); dfree(d.xs3); dfree(d.vort); dfree(d.vxsph); dfree(d.xv);
  dfree(d.cell); dfree(d.arrival); dfree(d.cellSorted); dfree(d.tmpId);
  dfree(d.cellCount); dfree(d.cellStart); dfree(d.blockSums);
  dfree(d.nbr); dfree(d.ncnt); dfree(d.flags);
  dfree(d.grids); dfree(d.slotScene); dfree(d.runs); dfree(d.stage); dfree(d.dense);
  dfree(d.dense_b);
  x->alloc_n = 0;
  x->alloc_cells = 0;
}

// Drops the captured step graph (buffers, dt or schedule changed).
void drop_graph(Ctx* x) {
  if (x->graph_exec) (void)hipGraphExecDestroy(x->graph_exec);
  if (x->graph) (void)hipGraphDestroy(x->graph);
  x->graph_exec = nullptr;
  x->graph = nullptr;
  if (x->graph1_exec) (void)hipGraphExecDestroy(x->graph1_exec);
  if (x->graph1) (void)hipGraphDestroy(x->graph1);
  x->graph1_exec = nullptr;
  x->graph1 = nullptr;
}

template <class T>
hipError_t dalloc(T*& p, size_t count) {
  return hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
}

int ensure_device(Ctx* x) {
  if (x->alloc_n == (size_t)x->n && x->alloc_cells == (size_t)x->ncell && x->d.posA) return PBF_OK;
  free_device(x);
  pbf::DevState& d = x->d;
  const size_t n = (size_t)x->n, np = (size_t)x->npad, C = (size_t)x->ncell;
  const size_t K = (size_t)pbf::k4_of(*x) * pbf::kChunk;
  const size_t nb = (C + 1023) / 1024 + 1;
  hipError_t e = hipSuccess;
#define DA(ptr, cnt) \
  if (e == hipSuccess) e = dalloc(ptr, cnt)
  DA(d.posA, np); DA(d.posB, np); DA(d.velA, np); DA(d.velB, np);
  DA(d.predP, np + 64); DA(d.predQ, np + 64); DA(d.vort, np); DA(d.vxsph, np); DA(d.xv, 2 * np);
  DA(d.xs3, 3 * (np + 64));
  DA(d.cell, np); DA(d.arrival, np); DA(d.cellSorted, np); DA(d.tmpId, np);
  DA(d.cellCount, C + 4); DA(d.cellStart, C + 4); DA(d.blockSums, nb);
  DA(d.nbr, np * K); DA(d.ncnt, np); DA(d.flags, 64);
  DA(d.grids, (size_t)x->nscenes); DA(d.slotScene, np); DA(d.runs, (size_t)x->nscenes);
  DA(d.stage, 4 * np);
  DA(d.dense, np);
  DA(d.dense_b, np);
#undef DA
  if (e != hipSuccess) {
    free_device(x);
    return fail(x, PBF_ERR_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
  }
  x->alloc_n = n;
  x->alloc_cells = C;
  return PBF_OK;
}

// Verdicts of the exhaustive device check of the short square root,
// reciprocal and division sequences, per device (and per h^9 for the
// division).  A failed or skipped check leaves the IEEE expansions on.
struct MathVerdict {
  bool checked = false;
  int rcp = 0, sqrt = 0;
  std::map<uint64_t, int> div;  // keyed by the bits of h^9 and the lowest binade
};
std::mutex g_math_mu;
std::map<int, MathVerdict> g_math;

int verify_math(Ctx* x) {
  if (getenv("PBF_IEEE_MATH")) return PBF_OK;  // keep the IEEE expansions
  const SimConst k = pbf::consts_for(*x, x->p.dt);
  // The poly6 numerator c1 (h^2 - r^2)^3 is 0 or lies in [amin, amax]:
  // h^2 - r^2 >= h^2 - pred(h^2) for every float r^2 < h^2, and rounding is
  // monotonic.
  const float t = k.h2, amax = k.c1 * ((t * t) * t);
  const float tmin = k.h2 - nextafterf(k.h2, 0.0f), amin = k.c1 * ((tmin * tmin) * tmin);
  const bool div_range = amin >= 0x1p-126f && amax < k.div_hi;
  const int e_hi = ilogbf(k.div_hi), e_lo = div_range ? ilogbf(amin) : e_hi;
  uint32_t hb9;
  memcpy(&hb9, &k.h9, sizeof hb9);
  const uint64_t key = ((uint64_t)hb9 << 32) | (uint32_t)(e_lo + 1024);
  std::lock_guard<std::mutex> lock(g_math_mu);
  MathVerdict& v = g_math[x->device];
  auto it = v.div.find(key);
  if (!v.checked || it == v.div.end()) {
    uint32_t* bad = nullptr;
    HIPCHK(x, hipMalloc((void**)&bad, 4 * sizeof(uint32_t)));
    uint32_t hb[4] = {1, 1, 1, 1};
    hipError_t e = pbf::launch_fill_u32(bad, 4, 0u, x->stream);
    if (e == hipSuccess) e = pbf::launch_verify_math(k.h9, k.r9, e_lo, e_hi, bad, x->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hb, bad, sizeof hb, hipMemcpyDeviceToHost, x->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
    (void)hipFree(bad);
    HIPCHK(x, e);
    v.checked = true;
    v.rcp = hb[1] == 0;
    v.sqrt = hb[2] == 0;
    it = v.div.emplace(key, e_lo < e_hi && hb[0] == 0).first;
  }
  // The arguments the kernels pass must lie in the checked ranges: spiky's
  // reciprocal argument is h^6 |r| with EPS_D <= |r| < h (or 1), and the
  // short square root of an argument below 2^-96 (NaN or under 2^-47) only
  // meets the EPS_D cut-off.
  const float zlo = k.h6 * k.eps_d, zhi = k.h6 * k.h;
  x->fast_div = it->second && div_range;
  x->fast_rcp = v.rcp && zlo >= 0x1p-100f && zhi < 0x1p100f;
  x->fast_sqrt = v.sqrt && k.eps_d >= 0x1p-47f;
  return PBF_OK;
}

int finish_create(Ctx* x, pbf_ctx** out) {
  uint64_t total = 0;
  for (auto& g : x->grids) {
    if (total > (1ull << 30)) break;
    g.cell_base = (uint32_t)total;
    total += (uint64_t)g.dims[0] * (uint64_t)g.dims[1] * (uint64_t)g.dims[2];
  }
  if (total > (1ull << 30)) {
    std::string m = "grid too large (more than 2^30 cells): enlarge cell_scale or shrink the box";
    delete x;
    return fail(nullptr, PBF_ERR_INVALID, m);
  }
  x->ncell = (uint32_t)total;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&x->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    std::string m = std::string("HIP unavailable: ") + hipGetErrorString(e);
    delete x;
    return fail(nullptr, PBF_ERR_HIP, m);
  }
  x->device = dev;
  // PBF_NO_SMALL_SORT keeps the multi-launch scan and sort for small states
  // too (same bits; for measurement).
  x->small_sort = getenv("PBF_NO_SMALL_SORT") == nullptr;
  // PBF_STEP_GRAPH=0: one-step calls of the flat schedule keep plain launches.
  if (const char* w = getenv("PBF_STEP_GRAPH")) x->step_graph = w[0] != '0';
  if (const char* w = getenv("PBF_WIDE_MAX"))
    for (pbf::PassRule& r : x->rules) r.wave_max = atoll(w);
  if (const char* w = getenv("PBF_QUAD_MAX"))
    for (pbf::PassRule& r : x->rules) r.quad_max = atoll(w);
  if (const char* w = getenv("PBF_DENSE_LIST")) x->dense_list = (uint32_t)atoll(w);
  if (const char* w = getenv("PBF_DENSE_CELL")) x->dense_cell = (uint32_t)atoll(w);
  // PBF_RULE_<PASS>=wave_max,quad_max sets one pass's bounds.
  static const char* const kRuleEnv[pbf::kListPasses] = {
      "PBF_RULE_BUILD", "PBF_RULE_LAMBDA", "PBF_RULE_DELTAP", "PBF_RULE_XSPH", "PBF_RULE_VORT"};
  for (int k = 0; k < pbf::kListPasses; k++) {
    long long wm = 0, qm = 0;
    if (const char* w = getenv(kRuleEnv[k]))
      if (sscanf(w, "%lld,%lld", &wm, &qm) == 2) x->rules[k] = {wm, qm};
  }
  if (verify_math(x) != PBF_OK) {
    std::string m = x->err;
    (void)hipStreamDestroy(x->stream);
    delete x;
    return fail(nullptr, PBF_ERR_HIP, m);
  }
  *out = reinterpret_cast<pbf_ctx*>(x);
  return PBF_OK;
}

Ctx* C(pbf_ctx* c) { return reinterpret_cast<Ctx*>(c); }
const Ctx* C(const pbf_ctx* c) { return reinterpret_cast<const Ctx*>(c); }

// Per-pass HIP events of a profiled step (recorded on the context stream).
struct PassEvents {
  std::vector<hipEvent_t> ev;  // 2 per launch
  std::vector<int> pass;       // PBF_PASS_* of each launch
  int used = 0;
};

int mark(Ctx* x, PassEvents* pe, int pass, bool begin) {
  if (!pe) return PBF_OK;
  if (begin) {
    if ((size_t)(pe->used + 1) * 2 > pe->ev.size()) {
      hipEvent_t a, b;
      HIPCHK(x, hipEventCreate(&a));
      HIPCHK(x, hipEventCreate(&b));
      pe->ev.push_back(a);
      pe->ev.push_back(b);
      pe->pass.push_back(pass);
    }
    pe->pass[(size_t)pe->used] = pass;
    HIPCHK(x, hipEventRecord(pe->ev[(size_t)pe->used * 2], x->stream));
  } else {
    HIPCHK(x, hipEventRecord(pe->ev[(size_t)pe->used * 2 + 1], x->stream));
    pe->used++;
  }
  return PBF_OK;
}

#define PASS(pass, call)                               \
  do {                                                 \
    int r_ = mark(x, pe, (pass), true);                \
    if (r_ != PBF_OK) return r_;                       \
    HIPCHK(x, call);                                   \
    if ((r_ = mark(x, pe, (pass), false)) != PBF_OK) return r_; \
  } while (0)

// One step, optionally stopping after a debug stage (0 = full step).
// predicted: the previous step of the same call already ran this step's
// prediction; predict_next: fuse the next step's prediction into this
// step's last pass (both steps of one call, same dt).
int run_step(Ctx* x, const SimConst& c, int stage, PassEvents* pe = nullptr,
             bool predicted = false, bool predict_next = false) {
  pbf::DevState& d = x->d;
  // A full step neither stores the advanced velocity nor sorts it; the debug
  // stops keep both (they read them back).  Restored on every return.
  struct SortedV {
    Ctx* x;
    ~SortedV() { x->sorted_v = true; }
  } restore{x};
  x->sorted_v = stage != 0;
  if (!predicted) PASS(PBF_PASS_PREDICT, pbf::launch_predict(*x, c));
  // A prediction fused into the previous step's last pass sits in vort.
  if (x->small_sort && pbf::sort_in_one_workgroup(*x)) {
    PASS(PBF_PASS_SORT, pbf::launch_sort_small(*x, predicted ? d.vort : nullptr));
  } else {
    PASS(PBF_PASS_SCAN, pbf::launch_scan(*x));
    PASS(PBF_PASS_SORT, pbf::launch_scatter_rank(*x, predicted ? d.vort : nullptr));
  }
  x->dbg_pred = d.predQ;
  x->dbg_vel = d.velB;
  if (stage == PBF_STAGE_PREDICT) return PBF_OK;
  float4* cur = d.predQ;
  float4* other = d.predP;
#ifndef PBF_FUSE_NBR_LAMBDA
#define PBF_FUSE_NBR_LAMBDA 1
#endif
  const bool fuse = PBF_FUSE_NBR_LAMBDA && stage != PBF_STAGE_NEIGHBORS;
  if (fuse)
    PASS(PBF_PASS_NEIGHBORS, pbf::launch_nbr_lambda(*x, c, other));
  else
    PASS(PBF_PASS_NEIGHBORS, pbf::launch_nbr_build(*x, c));
  x->has_nbrs = true;
  if (stage == PBF_STAGE_NEIGHBORS) return PBF_OK;
  for (int it = 0; it < x->p.iterations; it++) {
    if (!(fuse && it == 0))
      PASS(PBF_PASS_LAMBDA, pbf::launch_lambda(*x, c, cur, other, it > 0));
    if (stage == PBF_STAGE_LAMBDA0 && it == 0) {
      x->dbg_pred = other;
      return PBF_OK;
    }
    const bool last = it == x->p.iterations - 1;
    PASS(PBF_PASS_DELTAP, pbf::launch_deltap(*x, c, other, cur, last));
    // The last iteration leaves x* and v only in the XSPH records; a debug
    // stop unpacks them where the earlier stages keep them.
    if ((stage == PBF_STAGE_DELTAP0 && it == 0) || (stage == PBF_STAGE_SOLVE && last)) {
      if (last)
        HIPCHK(x, pbf::launch_unpack_record(*x, cur, d.velB));
      else
        HIPCHK(x, pbf::launch_unpack_xs3(*x, cur));
      x->dbg_pred = cur;
      return PBF_OK;
    }
  }
  PASS(PBF_PASS_XSPH, pbf::launch_xsph(*x, c, other));
  x->dbg_pred = other;
  if (stage == PBF_STAGE_XSPH) return PBF_OK;
  if (predict_next)
    PASS(PBF_PASS_VORTICITY, pbf::launch_vort_commit_predict(*x, c, other));
  else
    PASS(PBF_PASS_VORTICITY, pbf::launch_vort_commit(*x, c, other));
  x->dbg_vel = d.velA;
  return PBF_OK;
}

int check_ready(Ctx* x) {
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (!x->loaded) return fail(x, PBF_ERR_STATE, "no particles loaded");
  HIPCHK(x, hipSetDevice(x->device));
  return PBF_OK;
}

using host::run_parts;

// Pinned slabs of kStageSlab floats through which large results return.
constexpr size_t kStageSlab = (size_t)16 << 20;

hipError_t stage_alloc(Ctx* x) {
  hipError_t e = hipSuccess;
  for (int s = 0; s < 2 && e == hipSuccess; s++) {
    e = hipHostMalloc((void**)&x->stage_host[s], kStageSlab * sizeof(float), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x->stage_ev[s], hipEventDisableTiming);
  }
  return e;
}

void stage_free(Ctx* x) {
  for (int s = 0; s < 2; s++) {
    if (x->stage_host[s]) (void)hipHostFree(x->stage_host[s]);
    if (x->stage_ev[s]) (void)hipEventDestroy(x->stage_ev[s]);
    x->stage_host[s] = nullptr;
    x->stage_ev[s] = nullptr;
  }
}

// Copies m floats from device memory to out: slab k+1 travels to its pinned
// buffer while host threads copy slab k into out (a plain copy to pageable
// memory is staged and copied by one thread, and takes every first-touch
// page fault of a new array on that thread).
hipError_t stage_copy_back(Ctx* x, const float* dsrc, size_t m, float* out) {
  if (!x->stage_host[0]) {
    const hipError_t e = stage_alloc(x);
    if (e != hipSuccess) {
      stage_free(x);
      return e;
    }
  }
  const size_t nslab = (m + kStageSlab - 1) / kStageSlab;
  const unsigned hw = std::thread::hardware_concurrency();
  // At least 1M floats (4 MB) per host thread, at most 8 threads.
  const size_t by_size = std::max<size_t>(1, std::min(m, kStageSlab) >> 20);
  const unsigned cap = std::max(1u, std::min(8u, hw ? hw : 1u));
  const unsigned nthr = (unsigned)std::min<size_t>(by_size, cap);
  for (size_t s = 0; s <= nslab; s++) {
    if (s < nslab) {
      const size_t off = s * kStageSlab, cnt = std::min(kStageSlab, m - off);
      hipError_t e = hipMemcpyAsync(x->stage_host[s & 1], dsrc + off, cnt * sizeof(float),
                                    hipMemcpyDeviceToHost, x->stream);
      if (e == hipSuccess) e = hipEventRecord(x->stage_ev[s & 1], x->stream);
      if (e != hipSuccess) return e;
    }
    if (s == 0) continue;
    const size_t prev = s - 1, off = prev * kStageSlab, cnt = std::min(kStageSlab, m - off);
    hipError_t e = hipEventSynchronize(x->stage_ev[prev & 1]);
    if (e != hipSuccess) return e;
    const float* src = x->stage_host[prev & 1];
    const size_t per = (cnt + nthr - 1) / nthr;
    run_parts((int64_t)nthr, [=](int64_t t) {
      const size_t b = (size_t)t * per;
      if (b < cnt) memcpy(out + off + b, src + b, std::min(per, cnt - b) * sizeof(float));
    });
  }
  return hipSuccess;
}

// Device-to-host copy of `bytes` (a whole number of floats) on the context
// stream, synchronous at return; copies of 4 MB or more go through the
// pinned slabs.
int copy_out(Ctx* x, const float* dsrc, void* host, size_t bytes) {
  if (bytes >= ((size_t)4 << 20)) {
    HIPCHK(x, stage_copy_back(x, dsrc, bytes / sizeof(float), static_cast<float*>(host)));
    return PBF_OK;
  }
  HIPCHK(x, hipMemcpyAsync(host, dsrc, bytes, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  return PBF_OK;
}

}  // namespace

namespace pbf {

// Derived constants in float, each in the reference's evaluation order:
// intpow by repeated multiplication (src/bsdf.h:49-57), the tensile scale
// 1/poly6((0,0,dq*h)) (src/particles.cpp:151), 1/rho0 for the vector
// divisions (vector3D.h:100-102), g*dt (:180), dt*eps_v (:242).
void derive_constants(const pbf_params& p, float dt, SimConst* c) {
  c->h = p.h;
  c->h2 = p.h * p.h;
  float h6 = p.h, h9 = p.h;
  for (int k = 1; k < 6; k++) h6 = h6 * p.h;
  for (int k = 1; k < 9; k++) h9 = h9 * p.h;
  c->h6 = h6;
  c->h9 = h9;
  c->c1 = p.poly6_coef;
  c->m3c2 = -3.0f * p.spiky_coef;
  c->eps_d = p.eps_d;
  const float z = p.dq * p.h, zero = 0.0f;
  const float r2 = (zero * zero + zero * zero) + z * z;
  float w = 0.0f;
  if (r2 < c->h2) {
    const float t = c->h2 - r2;
    w = (c->c1 * ((t * t) * t)) / c->h9;
  }
  c->S = 1.0f / w;
  c->rho0 = p.rest_density;
  c->inv_rho0 = 1.0f / p.rest_density;
  c->eps = p.epsilon;
  c->negk = -p.k_tensile;
  c->xc = p.xsph_c;
  c->dt = dt;
  c->dtv = dt * p.vort_eps;
  c->inv_dt = 1.0f / dt;
  c->gdt = p.gravity * dt;
  c->n_tensile = p.n_tensile;
  // Short division by h^9: numerators below div_hi keep a * (1/h^9) below
  // 2^100.  The flags stay off until a device has checked the sequences
  // (consts_for).
  c->r9 = 1.0f / h9;
  int e_hi = 100;
  if (c->r9 > 0.0f && isfinite(c->r9)) e_hi = std::min(100, 100 - ilogbf(c->r9) - 1);
  c->div_hi = ldexpf(1.0f, std::max(e_hi, -126));
  c->fast_div = 0;
  c->fast_rcp = 0;
  c->fast_sqrt = 0;
}

SimConst consts_for(const Ctx& x, float dt) {
  SimConst k;
  derive_constants(x.p, dt, &k);
  k.fast_div = x.fast_div;
  k.fast_rcp = x.fast_rcp;
  k.fast_sqrt = x.fast_sqrt;
  return k;
}

void derive_grid(const pbf_params& p, const float* lo, const float* hi, SceneGrid* g) {
  const float cell = p.h * p.cell_scale;
  const float inv = 1.0f / cell;
  for (int a = 0; a < 3; a++) {
    g->lo[a] = lo[a];
    g->hi[a] = hi[a];
    g->lo_e[a] = lo[a] + p.eps_d;
    g->hi_e[a] = hi[a] - p.eps_d;
    const float ext = floorf((hi[a] - lo[a]) * inv);
    g->dims[a] = (ext >= 0.0f && ext < 1e6f) ? (int)ext + 1 : 1;
  }
  g->inv_cell = inv;
  g->stop = p.wall_stop_mask;
  g->cell_base = 0;
}

}  // namespace pbf

extern "C" {

PBF_API void pbf_params_default(pbf_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->h = 0.3f;
  p->rest_density = 1000.0f;
  p->iterations = 12;
  p->epsilon = 2.0f;
  p->k_tensile = 0.0001f;
  p->n_tensile = 4;
  p->dq = 0.1f;
  p->xsph_c = 0.001f;
  p->vort_eps = 0.001f;
  p->gravity = 10.0f;
  p->dt = 0.016f;
  p->box_lo[0] = -1.0f; p->box_lo[1] = 0.0f;  p->box_lo[2] = -1.0f;
  p->box_hi[0] = 1.0f;  p->box_hi[1] = 1.49f; p->box_hi[2] = 1.0f;
  p->wall_stop_mask = (1u << 3) | (1u << 5);
  p->cell_scale = 1.0f + 1.0f / 1024.0f;
  p->max_neighbors = 256;
  p->poly6_coef = 1.56668147106f;
  p->spiky_coef = 4.774648292756860f;
  p->eps_d = 0.00000000001f;
}

static int pbf_create_impl(const pbf_params* p, pbf_ctx** out) {
  std::string why;
  if (!out) return fail(nullptr, PBF_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int rc = validate_params(p, &why);
  if (rc == PBF_OK) rc = validate_box(p->box_lo, p->box_hi, &why);
  if (rc != PBF_OK) return fail(nullptr, rc, why);
  Ctx* x = new Ctx();
  x->p = *p;
  x->nscenes = 1;
  x->counts.assign(1, 0);
  x->offsets.assign(1, 0);
  x->grids.resize(1);
  pbf::derive_grid(*p, p->box_lo, p->box_hi, &x->grids[0]);
  return finish_create(x, out);
}

PBF_API int pbf_create(const pbf_params* p, pbf_ctx** out) {
  try {
    return pbf_create_impl(p, out);
  } catch (const std::bad_alloc&) {
    return fail(nullptr, PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(nullptr, PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_create_batch_impl(const pbf_params* p, int32_t nscenes, const int64_t* counts,
                                 const float* box_lo, const float* box_hi, pbf_ctx** out) {
  std::string why;
  if (!out) return fail(nullptr, PBF_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int rc = validate_params(p, &why);
  if (rc != PBF_OK) return fail(nullptr, rc, why);
  if (nscenes < 1 || !counts || !box_lo || !box_hi)
    return fail(nullptr, PBF_ERR_INVALID, "nscenes >= 1 and counts/box_lo/box_hi are required");
  Ctx* x = new Ctx();
  x->p = *p;
  x->nscenes = nscenes;
  x->counts.assign(counts, counts + nscenes);
  x->offsets.resize(nscenes);
  x->grids.resize(nscenes);
  int64_t tot = 0;
  for (int s = 0; s < nscenes; s++) {
    if (counts[s] < 1 || validate_box(box_lo + 3 * s, box_hi + 3 * s, &why) != PBF_OK) {
      delete x;
      return fail(nullptr, PBF_ERR_INVALID,
                  "scene " + std::to_string(s) + ": count must be >= 1 and lo < hi");
    }
    x->offsets[s] = tot;
    tot += counts[s];
    pbf::derive_grid(*p, box_lo + 3 * s, box_hi + 3 * s, &x->grids[s]);
  }
  if (tot >= (int64_t)1 << 31) {
    delete x;
    return fail(nullptr, PBF_ERR_INVALID, "more than 2^31 particles");
  }
  return finish_create(x, out);
}

PBF_API int pbf_create_batch(const pbf_params* p, int32_t nscenes, const int64_t* counts,
                             const float* box_lo, const float* box_hi, pbf_ctx** out) {
  try {
    return pbf_create_batch_impl(p, nscenes, counts, box_lo, box_hi, out);
  } catch (const std::bad_alloc&) {
    return fail(nullptr, PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(nullptr, PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API int pbf_destroy(pbf_ctx* c) {
  Ctx* x = C(c);
  if (!x) return PBF_OK;
  (void)hipSetDevice(x->device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  drop_graph(x);
  free_device(x);
  dfree(x->d.meshTri);
  dfree(x->d.meshNode);
  dfree(x->d.meshFlag);
  stage_free(x);
  if (x->flag_host) (void)hipHostFree(x->flag_host);
  dfree(x->field_dev);
  dfree(x->field_pts);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return PBF_OK;
}

namespace {
int load_impl(Ctx* x, const float* xyz, const float* v, const float* rho, int64_t n);
}

static int pbf_load_impl(pbf_ctx* c, const float* xyz, const float* v, int64_t n) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  return load_impl(x, xyz, v, nullptr, n);
}

PBF_API int pbf_load(pbf_ctx* c, const float* xyz, const float* v, int64_t n) {
  try {
    return pbf_load_impl(c, xyz, v, n);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

namespace {
// pbf_load, with the last densities (rho, may be NULL) restored as well.
int load_impl(Ctx* x, const float* xyz, const float* v, const float* rho, int64_t n) {
  if (!xyz || n < 1) return fail(x, PBF_ERR_INVALID, "xyz must be non-NULL and n >= 1");
  if (n >= (int64_t)1 << 31) return fail(x, PBF_ERR_INVALID, "more than 2^31 particles");
  if (x->nscenes > 1) {
    int64_t tot = 0;
    for (auto k : x->counts) tot += k;
    if (tot != n)
      return fail(x, PBF_ERR_INVALID,
                  "n = " + std::to_string(n) + " but the scenes hold " + std::to_string(tot));
  }
  // Validate every particle against its scene box and pack the float4
  // records (host threads; the first failure a sequential pass would meet).
  std::vector<float> boxes;
  for (const SceneGrid& g : x->grids) {
    boxes.insert(boxes.end(), g.lo, g.lo + 3);
    boxes.insert(boxes.end(), g.hi, g.hi + 3);
  }
  host::Packed pk;
  host::validate_pack(xyz, v, rho, n, x->offsets, boxes, 8, &pk);
  if (pk.code == PBF_ERR_INVALID)
    return fail(x, PBF_ERR_INVALID, "particle " + std::to_string(pk.first) + " is not finite");
  if (pk.code == PBF_ERR_OUT_OF_BOX)
    return fail(x, PBF_ERR_OUT_OF_BOX,
                "particle " + std::to_string(pk.first) + " lies outside the collision box");
  // The input is accepted.  From here on the device buffers are replaced or
  // overwritten, so the context counts as unloaded until every upload has
  // succeeded: after a failed allocation or copy, pbf_step and pbf_read
  // return PBF_ERR_STATE instead of running on freed or partial buffers.
  HIPCHK(x, hipSetDevice(x->device));
  x->loaded = false;
  x->has_nbrs = false;
  x->field_ready = false;
  drop_graph(x);
  if (x->nscenes == 1) x->counts[0] = n;
  x->n = n;
  x->npad = (n + 63) / 64 * 64;
  int rc = ensure_device(x);
  if (rc != PBF_OK) return rc;
  pbf::DevState& d = x->d;
  HIPCHK(x, hipMemcpyAsync(d.posA, pk.pos.data(), (size_t)n * sizeof(float4),
                           hipMemcpyHostToDevice, x->stream));
  HIPCHK(x, hipMemcpyAsync(d.velA, pk.vel.data(), (size_t)n * sizeof(float4),
                           hipMemcpyHostToDevice, x->stream));
  HIPCHK(x, hipMemcpyAsync(d.grids, x->grids.data(), x->grids.size() * sizeof(SceneGrid),
                           hipMemcpyHostToDevice, x->stream));
  HIPCHK(x, hipMemcpyAsync(d.slotScene, pk.slot_scene.data(), (size_t)n * sizeof(uint32_t),
                           hipMemcpyHostToDevice, x->stream));
  // Scene runs for the persistent schedule, largest first (longest
  // processing time first balances the workgroups).
  std::vector<pbf::SceneRun> runs((size_t)x->nscenes);
  for (int s = 0; s < x->nscenes; s++) {
    const SceneGrid& g = x->grids[(size_t)s];
    runs[(size_t)s] = {(uint32_t)s, (uint32_t)x->offsets[(size_t)s], (uint32_t)x->counts[(size_t)s],
                       (uint32_t)g.dims[0] * (uint32_t)g.dims[1] * (uint32_t)g.dims[2]};
  }
  std::stable_sort(runs.begin(), runs.end(), [](const pbf::SceneRun& a, const pbf::SceneRun& b) {
    return a.count > b.count;
  });
  HIPCHK(x, hipMemcpyAsync(d.runs, runs.data(), runs.size() * sizeof(pbf::SceneRun),
                           hipMemcpyHostToDevice, x->stream));
  HIPCHK(x, pbf::launch_fill_u32(d.flags, 64, 0u, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  drop_graph(x);
  x->sim_time = 0.0;
  x->loaded = true;
  x->debug_dirty = false;
  x->has_nbrs = false;
  x->field_ready = false;
  return PBF_OK;
}
}  // namespace

namespace {

// Waits for the stream, then reports (and clears) a neighbour-list overflow.
int sync_check(Ctx* x) {
  // The flag word returns through a pinned word of the context: a small
  // copy into pageable memory is staged by the runtime on every call.
  if (!x->flag_host) HIPCHK(x, hipHostMalloc((void**)&x->flag_host, 64, hipHostMallocDefault));
  HIPCHK(x, hipMemcpyAsync(x->flag_host, x->d.flags, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  const uint32_t over = x->flag_host[0];
  if (over == 0) return PBF_OK;
  HIPCHK(x, pbf::launch_fill_u32(x->d.flags, 1, 0u, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  return fail(x, PBF_ERR_OVERFLOW,
              "a particle had " + std::to_string(over) + " neighbours; the list capacity is " +
                  std::to_string(pbf::k4_of(*x) * pbf::kChunk) + " (max_neighbors " +
                  std::to_string(x->p.max_neighbors) + ")");
}

// Enqueues nsteps steps with the context's schedule.
int enqueue_steps(Ctx* x, float dt, int nsteps) {
  if (nsteps == 0) return PBF_OK;
  x->field_ready = false;
  const SimConst k = pbf::consts_for(*x, dt);
  if (x->schedule == PBF_SCHED_PERSISTENT) {
    const int per = pbf::persistent_max_blocks_per_cu(x->pblock);
    const int grid = std::max(1, std::min(x->nscenes, std::max(1, x->ncu) * per));
    HIPCHK(x, pbf::launch_persistent(*x, k, nsteps, x->pblock, grid));
    x->has_nbrs = true;
    x->dbg_vel = x->d.velA;
  } else if (x->schedule == PBF_SCHED_GRAPH && nsteps >= 3) {
    // The first step (it predicts) and the last (it commits without
    // predicting) are plain launches; every step between replays one
    // captured step that starts from the prediction fused into the previous
    // step and ends with the next one, as the flat schedule runs them.
    int rc = run_step(x, k, 0, nullptr, false, true);
    if (rc != PBF_OK) return rc;
    if (!x->graph_exec || x->graph_dt != dt) {
      drop_graph(x);
      HIPCHK(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
      rc = run_step(x, k, 0, nullptr, true, true);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(x->stream, &g);
      if (rc != PBF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(x, PBF_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
      x->graph = g;
      HIPCHK(x, hipGraphInstantiate(&x->graph_exec, g, nullptr, nullptr, 0));
      x->graph_dt = dt;
    }
    for (int s = 1; s + 1 < nsteps; s++) HIPCHK(x, hipGraphLaunch(x->graph_exec, x->stream));
    rc = run_step(x, k, 0, nullptr, true, false);
    if (rc != PBF_OK) return rc;
  } else if (nsteps == 1 &&
             (x->schedule == PBF_SCHED_GRAPH ||
              (x->schedule == PBF_SCHED_FLAT && x->step_graph && dt == x->last_dt1))) {
    // A call of one step (the reference's timeStep per frame) replays one
    // captured whole step, from the prediction to the plain commit: one
    // launch from the host instead of one per pass.  The flat schedule does
    // so once a second one-step call in a row has the same dt, so that a
    // caller who changes dt every call is not made to capture every call.
    x->last_dt1 = dt;
    if (!x->graph1_exec || x->graph1_dt != dt) {
      if (x->graph1_exec) (void)hipGraphExecDestroy(x->graph1_exec);
      if (x->graph1) (void)hipGraphDestroy(x->graph1);
      x->graph1_exec = nullptr;
      x->graph1 = nullptr;
      HIPCHK(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
      const int rc = run_step(x, k, 0, nullptr, false, false);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(x->stream, &g);
      if (rc != PBF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(x, PBF_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
      x->graph1 = g;
      HIPCHK(x, hipGraphInstantiate(&x->graph1_exec, g, nullptr, nullptr, 0));
      x->graph1_dt = dt;
    }
    HIPCHK(x, hipGraphLaunch(x->graph1_exec, x->stream));
  } else {
    // Flat schedule, and the graph schedule for calls of two steps.
    if (nsteps == 1) x->last_dt1 = dt;
    for (int s = 0; s < nsteps; s++) {
      const int rc = run_step(x, k, 0, nullptr, s > 0, s + 1 < nsteps);
      if (rc != PBF_OK) return rc;
    }
  }
  for (int s = 0; s < nsteps; s++) x->sim_time += (double)dt;
  return PBF_OK;
}

int check_step_args(Ctx* x, float dt, int32_t nsteps) {
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  if (!is_finite_f(dt) || !(dt > 0.0f) || nsteps < 0)
    return fail(x, PBF_ERR_INVALID, "dt must be finite and > 0, nsteps >= 0");
  return PBF_OK;
}

}  // namespace

static int pbf_step_async_impl(pbf_ctx* c, float dt, int32_t nsteps) {
  Ctx* x = C(c);
  int rc = check_step_args(x, dt, nsteps);
  if (rc != PBF_OK) return rc;
  return enqueue_steps(x, dt, nsteps);
}

PBF_API int pbf_step_async(pbf_ctx* c, float dt, int32_t nsteps) {
  try {
    return pbf_step_async_impl(c, dt, nsteps);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API int pbf_sync(pbf_ctx* c) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  return sync_check(x);
}

static int pbf_step_dt_impl(pbf_ctx* c, float dt, int32_t nsteps) {
  Ctx* x = C(c);
  int rc = check_step_args(x, dt, nsteps);
  if (rc != PBF_OK) return rc;
  if ((rc = enqueue_steps(x, dt, nsteps)) != PBF_OK) return rc;
  return sync_check(x);
}

PBF_API int pbf_step_dt(pbf_ctx* c, float dt, int32_t nsteps) {
  try {
    return pbf_step_dt_impl(c, dt, nsteps);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_set_schedule_impl(pbf_ctx* c, int32_t schedule, int32_t block) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (schedule < PBF_SCHED_FLAT || schedule > PBF_SCHED_PERSISTENT)
    return fail(x, PBF_ERR_INVALID, "unknown schedule");
  if (schedule == PBF_SCHED_PERSISTENT && block != 256 && block != 512 && block != 1024)
    return fail(x, PBF_ERR_INVALID, "persistent block must be 256, 512 or 1024");
  drop_graph(x);
  x->schedule = schedule;
  if (schedule == PBF_SCHED_PERSISTENT) x->pblock = block;
  return PBF_OK;
}

PBF_API int pbf_set_schedule(pbf_ctx* c, int32_t schedule, int32_t block) {
  try {
    return pbf_set_schedule_impl(c, schedule, block);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_step_profiled_impl(pbf_ctx* c, float dt, int32_t nsteps, float* pass_ms,
                                  int32_t* pass_launches, int32_t npass) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  if (!is_finite_f(dt) || !(dt > 0.0f) || nsteps < 0 || !pass_ms || npass < PBF_NPASS)
    return fail(x, PBF_ERR_INVALID, "dt > 0, nsteps >= 0, pass_ms[PBF_NPASS] required");
  const SimConst k = pbf::consts_for(*x, dt);
  for (int i = 0; i < npass; i++) {
    pass_ms[i] = 0.0f;
    if (pass_launches) pass_launches[i] = 0;
  }
  x->field_ready = false;
  HIPCHK(x, pbf::launch_fill_u32(x->d.flags, 1, 0u, x->stream));
  // Events of every launch of every step, read after one synchronisation at
  // the end, so the steps run back to back as in pbf_step.
  PassEvents pe;
  for (int s = 0; s < nsteps && rc == PBF_OK; s++) {
    rc = run_step(x, k, 0, &pe, s > 0, s + 1 < nsteps);
    if (rc == PBF_OK) x->sim_time += (double)dt;
  }
  if (rc == PBF_OK) {
    hipError_t e = hipStreamSynchronize(x->stream);
    if (e != hipSuccess) rc = fail(x, PBF_ERR_HIP, hipGetErrorString(e));
  }
  for (int l = 0; l < pe.used && rc == PBF_OK; l++) {
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, pe.ev[(size_t)l * 2], pe.ev[(size_t)l * 2 + 1]);
    if (e != hipSuccess) rc = fail(x, PBF_ERR_HIP, hipGetErrorString(e));
    pass_ms[pe.pass[(size_t)l]] += ms;
    if (pass_launches) pass_launches[pe.pass[(size_t)l]] += 1;
  }
  for (hipEvent_t e : pe.ev) (void)hipEventDestroy(e);
  if (rc != PBF_OK) return rc;
  return sync_check(x);
}

PBF_API int pbf_step_profiled(pbf_ctx* c, float dt, int32_t nsteps, float* pass_ms,
                              int32_t* pass_launches, int32_t npass) {
  try {
    return pbf_step_profiled_impl(c, dt, nsteps, pass_ms, pass_launches, npass);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_step_impl(pbf_ctx* c, int32_t nsteps) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  return pbf_step_dt(c, x->p.dt, nsteps);
}

PBF_API int pbf_step(pbf_ctx* c, int32_t nsteps) {
  try {
    return pbf_step_impl(c, nsteps);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_step_debug_impl(pbf_ctx* c, float dt, int32_t stage) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  if (!is_finite_f(dt) || !(dt > 0.0f) || stage < PBF_STAGE_PREDICT || stage > PBF_STAGE_XSPH)
    return fail(x, PBF_ERR_INVALID, "dt must be > 0 and stage a PBF_STAGE_* value");
  const SimConst k = pbf::consts_for(*x, dt);
  x->field_ready = false;
  HIPCHK(x, pbf::launch_fill_u32(x->d.flags, 1, 0u, x->stream));
  drop_graph(x);
  x->debug_dirty = true;
  rc = run_step(x, k, stage);
  if (rc != PBF_OK) return rc;
  HIPCHK(x, hipStreamSynchronize(x->stream));
  return PBF_OK;
}

PBF_API int pbf_step_debug(pbf_ctx* c, float dt, int32_t stage) {
  try {
    return pbf_step_debug_impl(c, dt, stage);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_read_impl(pbf_ctx* c, float* xyz, float* v, float* rho) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  const size_t n = (size_t)x->n;
  pbf::DevState& d = x->d;
  if (xyz) {
    HIPCHK(x, pbf::launch_unpermute(*x, d.posA, d.posA, 0, 3, d.stage));
    if ((rc = copy_out(x, d.stage, xyz, n * 3 * sizeof(float))) != PBF_OK) return rc;
  }
  if (v) {
    HIPCHK(x, pbf::launch_unpermute(*x, d.posA, d.velA, 0, 3, d.stage));
    if ((rc = copy_out(x, d.stage, v, n * 3 * sizeof(float))) != PBF_OK) return rc;
  }
  if (rho) {
    HIPCHK(x, pbf::launch_unpermute(*x, d.posA, d.velA, 3, 1, d.stage));
    if ((rc = copy_out(x, d.stage, rho, n * sizeof(float))) != PBF_OK) return rc;
  }
  return PBF_OK;
}

PBF_API int pbf_read(pbf_ctx* c, float* xyz, float* v, float* rho) {
  try {
    return pbf_read_impl(c, xyz, v, rho);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

}  // extern "C"

extern "C" {

static int pbf_set_mesh_impl(pbf_ctx* c, const float* pos, const float* nrm, int64_t nv,
                             const int32_t* tri, int64_t nt, float skin) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (!is_finite_f(skin) || skin < 0.0f) return fail(x, PBF_ERR_INVALID, "skin must be >= 0");
  if (nt < 0 || nt >= ((int64_t)1 << 27)) return fail(x, PBF_ERR_INVALID, "0 <= nt < 2^27 required");
  if (nt > 0 && x->nscenes != 1)
    return fail(x, PBF_ERR_INVALID, "a mesh collider needs a single-scene context");
  if (nt > 0 && (!pos || !nrm || !tri || nv < 3 || nv >= ((int64_t)1 << 31)))
    return fail(x, PBF_ERR_INVALID, "pos, nrm, tri required and 3 <= nv < 2^31");
  for (int64_t i = 0; nt > 0 && i < 3 * nv; i++)
    if (!is_finite_f(pos[i]) || !is_finite_f(nrm[i]))
      return fail(x, PBF_ERR_INVALID, "non-finite mesh vertex or normal");
  for (int64_t i = 0; i < 3 * nt; i++)
    if (tri[i] < 0 || tri[i] >= nv)
      return fail(x, PBF_ERR_INVALID, "triangle " + std::to_string(i / 3) + " has a bad vertex index");
  HIPCHK(x, hipSetDevice(x->device));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  drop_graph(x);
  dfree(x->d.meshTri);
  dfree(x->d.meshNode);
  dfree(x->d.meshFlag);
  x->mesh_ntri = 0;
  if (nt == 0) return PBF_OK;
  // BVH, triangle records and cell flags on the host (pbf_host.h).
  // PBF_MESH_CELLS=0 leaves the cells out (every segment walks the BVH).
  const pbf::SceneGrid& g = x->grids[0];
  host::MeshGrid mg;
  for (int a = 0; a < 3; a++) {
    mg.lo[a] = g.lo[a];
    mg.hi[a] = g.hi[a];
    mg.dims[a] = g.dims[a];
  }
  mg.inv_cell = g.inv_cell;
  const char* cells_env = getenv("PBF_MESH_CELLS");
  const bool use_cells = !(cells_env && cells_env[0] == '0');
  host::MeshHost mh;
  const std::string why = host::build_mesh(pos, nrm, tri, nt, mg, use_cells, &mh);
  if (!why.empty()) return fail(x, PBF_ERR_INVALID, why);
  const size_t ncell = mh.flags.size();
  hipError_t e = dalloc(x->d.meshTri, mh.tris.size());
  if (e == hipSuccess) e = dalloc(x->d.meshNode, mh.nodes.size());
  if (e == hipSuccess && use_cells) e = dalloc(x->d.meshFlag, ncell);
  if (e == hipSuccess && use_cells)
    e = hipMemcpy(x->d.meshFlag, mh.flags.data(), ncell, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(x->d.meshTri, mh.tris.data(), mh.tris.size() * sizeof(float4),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(x->d.meshNode, mh.nodes.data(), mh.nodes.size() * sizeof(float4),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dfree(x->d.meshTri);
    dfree(x->d.meshNode);
    dfree(x->d.meshFlag);
    return fail(x, PBF_ERR_NOMEM, std::string("mesh upload: ") + hipGetErrorString(e));
  }
  x->mesh_ntri = (uint32_t)nt;
  x->mesh_nnode = mh.nnode;
  x->mesh_skin = skin;
  x->mesh_margin = mh.pad;
  return PBF_OK;
}

PBF_API int pbf_set_mesh(pbf_ctx* c, const float* pos, const float* nrm, int64_t nv,
                         const int32_t* tri, int64_t nt, float skin) {
  try {
    return pbf_set_mesh_impl(c, pos, nrm, nv, tri, nt, skin);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

}  // extern "C"

namespace {
// Snapshot format and checksum: pbf_host.h.
using host::SnapBody;
using host::SnapHeader;

void scene_layout(const Ctx* x, SnapBody* b) {
  b->p = x->p;
  b->counts = x->counts;
  b->boxes.clear();
  for (const SceneGrid& g : x->grids) {
    b->boxes.insert(b->boxes.end(), g.lo, g.lo + 3);
    b->boxes.insert(b->boxes.end(), g.hi, g.hi + 3);
  }
}
}  // namespace

extern "C" {

static int pbf_save_snapshot_impl(pbf_ctx* c, const char* path) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!path) return fail(x, PBF_ERR_INVALID, "path is NULL");
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  SnapBody b;
  scene_layout(x, &b);
  const size_t n = (size_t)x->n;
  b.x.resize(3 * n);
  b.v.resize(3 * n);
  b.rho.resize(n);
  if ((rc = pbf_read(c, b.x.data(), b.v.data(), b.rho.data())) != PBF_OK) return rc;
  const std::string why = host::write_snapshot(path, b, x->n, x->sim_time);
  if (!why.empty()) return fail(x, PBF_ERR_INVALID, why);
  return PBF_OK;
}

PBF_API int pbf_save_snapshot(pbf_ctx* c, const char* path) {
  try {
    return pbf_save_snapshot_impl(c, path);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_load_snapshot_impl(pbf_ctx* c, const char* path) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (!path) return fail(x, PBF_ERR_INVALID, "path is NULL");
  SnapBody b, mine;
  scene_layout(x, &mine);
  SnapHeader hd{};
  const std::string why = host::read_snapshot(path, mine, &hd, &b);
  if (!why.empty())
    return fail(x, PBF_ERR_INVALID, why.rfind("cannot open", 0) == 0 ? why : why + ": " + path);
  const int rc = load_impl(x, b.x.data(), b.v.data(), b.rho.data(), hd.n);
  if (rc != PBF_OK) return rc;
  x->sim_time = hd.sim_time;
  return PBF_OK;
}

PBF_API int pbf_load_snapshot(pbf_ctx* c, const char* path) {
  try {
    return pbf_load_snapshot_impl(c, path);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_estimate_densities_impl(pbf_ctx* c, float* rho) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!rho) return fail(x, PBF_ERR_INVALID, "rho is NULL");
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  const SimConst k = pbf::consts_for(*x, x->p.dt);
  HIPCHK(x, pbf::launch_densities(*x, k, x->d.stage));
  if ((rc = copy_out(x, x->d.stage, rho, (size_t)x->n * sizeof(float))) != PBF_OK) return rc;
  // The density pass re-sorted the committed state into posB/velB.
  drop_graph(x);
  std::swap(x->d.posA, x->d.posB);
  std::swap(x->d.velA, x->d.velB);
  x->has_nbrs = false;
  x->field_ready = true;
  return PBF_OK;
}

PBF_API int pbf_estimate_densities(pbf_ctx* c, float* rho) {
  try {
    return pbf_estimate_densities_impl(c, rho);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

// Density field of one scene at m query points (pts) or on a lattice.  The
// committed state is re-sorted by cell first, as pbf_estimate_densities does.
static int density_field(Ctx* x, int32_t scene, const float* pts, const pbf::DensityLattice* L,
                         int64_t m, float* out) {
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (x->debug_dirty) return fail(x, PBF_ERR_STATE, "state left by pbf_step_debug: reload first");
  if (scene < 0 || scene >= x->nscenes) return fail(x, PBF_ERR_INVALID, "scene out of range");
  if (m < 0 || m > (int64_t)INT32_MAX) return fail(x, PBF_ERR_INVALID, "0 <= m < 2^31 required");
  if (!out || (!L && m > 0 && !pts)) return fail(x, PBF_ERR_INVALID, "NULL argument");
  if (!L)
    for (int64_t i = 0; i < 3 * m; i++)
      if (!is_finite_f(pts[i])) return fail(x, PBF_ERR_INVALID, "non-finite query point");
  if (m == 0) return PBF_OK;
  const SimConst k = pbf::consts_for(*x, x->p.dt);
  hipError_t e = hipSuccess;
  if ((size_t)m > x->field_cap) {
    dfree(x->field_dev);
    x->field_cap = 0;
    e = dalloc(x->field_dev, (size_t)m);
    if (e == hipSuccess) x->field_cap = (size_t)m;
  }
  // Query points go to a device buffer the context keeps, so that repeated
  // small queries (a few points per call) allocate nothing.
  if (e == hipSuccess && !L && (size_t)m > x->field_pts_cap) {
    dfree(x->field_pts);
    x->field_pts_cap = 0;
    e = dalloc(x->field_pts, 3 * (size_t)m);
    if (e == hipSuccess) x->field_pts_cap = (size_t)m;
  }
  float* dp = L ? nullptr : x->field_pts;
  if (e == hipSuccess && !L)
    e = hipMemcpyAsync(dp, pts, (size_t)m * 3 * sizeof(float), hipMemcpyHostToDevice, x->stream);
  // Sorted once per committed state: later queries reuse the cell table.
  const bool sort = !x->field_ready;
  if (e == hipSuccess && sort) e = pbf::launch_sort_committed(*x, k);
  if (e == hipSuccess)
    e = pbf::launch_density_field(*x, k, scene, dp, L, (uint32_t)m, x->field_dev);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(x->stream);
    x->field_ready = false;
    return fail(x, PBF_ERR_HIP, std::string("density field: ") + hipGetErrorString(e));
  }
  // Small results (a few query points) take a plain copy, large ones the
  // pinned slabs; either way the stream has finished at return.
  rc = copy_out(x, x->field_dev, out, (size_t)m * sizeof(float));
  if (rc != PBF_OK) {
    x->field_ready = false;
    return rc;
  }
  if (sort) {
    // The sort left the committed state, re-sorted, in posB/velB.
    drop_graph(x);
    std::swap(x->d.posA, x->d.posB);
    std::swap(x->d.velA, x->d.velB);
    x->has_nbrs = false;
    x->field_ready = true;
  }
  return PBF_OK;
}

static int pbf_density_at_impl(pbf_ctx* c, int32_t scene, const float* pts, int64_t m, float* out) {
  return density_field(C(c), scene, pts, nullptr, m, out);
}

PBF_API int pbf_density_at(pbf_ctx* c, int32_t scene, const float* pts, int64_t m, float* out) {
  try {
    return pbf_density_at_impl(c, scene, pts, m, out);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

static int pbf_density_lattice_impl(pbf_ctx* c, int32_t scene, const float* lo, const float* hi,
                                    float step, const int32_t* counts, float* out) {
  Ctx* x = C(c);
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (!lo || !hi || !counts) return fail(x, PBF_ERR_INVALID, "NULL argument");
  pbf::DensityLattice L{};
  int64_t m = 1;
  if (!is_finite_f(step) || !(step > 0.0f)) return fail(x, PBF_ERR_INVALID, "step > 0 required");
  L.step = step;
  for (int a = 0; a < 3; a++) {
    if (!is_finite_f(lo[a]) || !is_finite_f(hi[a]) || counts[a] < 1)
      return fail(x, PBF_ERR_INVALID, "finite lo/hi and counts >= 1 required");
    L.lo[a] = lo[a];
    L.hi[a] = hi[a];
    L.n[a] = (uint32_t)counts[a];
    m *= counts[a];
    if (m > (int64_t)INT32_MAX) return fail(x, PBF_ERR_INVALID, "lattice of 2^31 points or more");
  }
  return density_field(x, scene, nullptr, &L, m, out);
}

PBF_API int pbf_density_lattice(pbf_ctx* c, int32_t scene, const float* lo, const float* hi,
                                float step, const int32_t* counts, float* out) {
  try {
    return pbf_density_lattice_impl(c, scene, lo, hi, step, counts, out);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

// ---- stepping caller-owned device arrays with no allocation -------------
// Every buffer the step needs beyond the caller's positions and velocities
// (sorted copies, predictions, cell tables, neighbour lists, flags) lives in
// one fixed device array.  Each device that runs the library holds its own
// copy, at its own address, until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for the device.
namespace {

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;  // device -> address of its copy

// Address of this device's copy of the workspace (looked up once per device).
int scratch_address(Ctx* x, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch.find(x->device);
  if (it == g_scratch.end()) {
    void* ptr = nullptr;
    HIPCHK(x, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(x, PBF_ERR_HIP, "no device address for the step workspace");
    it = g_scratch.emplace(x->device, static_cast<unsigned char*>(ptr)).first;
  }
  *out = it->second;
  return PBF_OK;
}

// Carves the device state of n particles out of [base, base + cap): the
// same arrays ensure_device allocates, each aligned to 256 bytes.  With
// base NULL it only measures; returns the bytes used.
size_t carve_state(unsigned char* base, size_t n, size_t ncell, size_t K, pbf::DevState* d) {
  const size_t np = (n + 63) / 64 * 64;
  size_t off = 0;
  auto take = [&](auto*& p, size_t count) {
    off = (off + 255) & ~(size_t)255;
    p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base ? base + off : nullptr);
    off += std::max<size_t>(count, 1) * sizeof(*p);
  };
  take(d->posA, np); take(d->posB, np); take(d->velA, np); take(d->velB, np);
  take(d->predP, np + 64); take(d->predQ, np + 64); take(d->vort, np); take(d->vxsph, np);
  take(d->xv, 2 * np); take(d->xs3, 3 * (np + 64));
  take(d->cell, np); take(d->arrival, np); take(d->cellSorted, np); take(d->tmpId, np);
  take(d->cellCount, ncell + 4); take(d->cellStart, ncell + 4);
  take(d->blockSums, (ncell + 1023) / 1024 + 1);
  take(d->nbr, np * K); take(d->ncnt, np); take(d->flags, 64);
  take(d->grids, 1); take(d->slotScene, np); take(d->runs, 1);
  d->stage = nullptr;
  d->dense = nullptr;  // no dense split here
  d->dense_b = nullptr;
  return off;
}

// Caller arrays (load order, 3 floats per particle) into the committed
// state: posA = (x, id), velA = (v, 0).  flags[2] / flags[3] receive the
// lowest index of a non-finite particle / a particle outside the box.
__global__ __launch_bounds__(256) void k_noalloc_pack(uint32_t n, const float* __restrict__ xyz,
                                                      const float* __restrict__ v,
                                                      float4* __restrict__ posA,
                                                      float4* __restrict__ velA, SceneGrid g,
                                                      uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float p[3] = {xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]};
  const float q[3] = {v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]};
  bool finite = true, inside = true;
  for (int a = 0; a < 3; a++) {
    finite = finite && isfinite(p[a]) && isfinite(q[a]);
    inside = inside && p[a] >= g.lo[a] && p[a] <= g.hi[a];
  }
  if (!finite) atomicMin(flags + 2, i);
  else if (!inside) atomicMin(flags + 3, i);
  posA[i] = make_float4(p[0], p[1], p[2], __uint_as_float(i));
  velA[i] = make_float4(q[0], q[1], q[2], 0.0f);
}

// Committed state (cell order, ids in posA.w) back into the caller arrays.
__global__ __launch_bounds__(256) void k_noalloc_unpack(uint32_t n, const float4* __restrict__ posA,
                                                        const float4* __restrict__ velA,
                                                        float* __restrict__ xyz,
                                                        float* __restrict__ v) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 p = posA[i], q = velA[i];
  const size_t o = 3 * (size_t)__float_as_uint(p.w);
  xyz[o] = p.x; xyz[o + 1] = p.y; xyz[o + 2] = p.z;
  v[o] = q.x; v[o + 1] = q.y; v[o + 2] = q.z;
}

// Restores the fields of a context that a workspace run borrows.
struct Borrow {
  Ctx* x;
  pbf::DevState d;
  int64_t n, npad;
  const float4 *dbg_pred, *dbg_vel;
  bool has_nbrs;
  explicit Borrow(Ctx* c)
      : x(c), d(c->d), n(c->n), npad(c->npad), dbg_pred(c->dbg_pred), dbg_vel(c->dbg_vel),
        has_nbrs(c->has_nbrs) {}
  ~Borrow() {
    x->d = d;
    x->n = n;
    x->npad = npad;
    x->dbg_pred = dbg_pred;
    x->dbg_vel = dbg_vel;
    x->has_nbrs = has_nbrs;
  }
};

int step_noalloc(Ctx* x, float* xyz, float* v, int64_t n, float dt, int32_t nsteps) {
  if (!x) return fail(nullptr, PBF_ERR_INVALID, "context is NULL");
  if (x->nscenes != 1) return fail(x, PBF_ERR_INVALID, "single-scene contexts only");
  if (!xyz || !v || n < 1) return fail(x, PBF_ERR_INVALID, "xyz and v must be non-NULL and n >= 1");
  if (n >= (int64_t)1 << 31) return fail(x, PBF_ERR_INVALID, "more than 2^31 particles");
  if (!is_finite_f(dt) || !(dt > 0.0f) || nsteps < 0)
    return fail(x, PBF_ERR_INVALID, "dt must be finite and > 0, nsteps >= 0");
  const size_t K = (size_t)pbf::k4_of(*x) * pbf::kChunk;
  pbf::DevState probe;
  const size_t need = carve_state(nullptr, (size_t)n, x->ncell, K, &probe);
  if (need > kNoAllocBytes) {
    // The largest n that fits at these parameters.
    int64_t lo = 0, hi = n;
    while (lo + 1 < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      (carve_state(nullptr, (size_t)mid, x->ncell, K, &probe) <= kNoAllocBytes ? lo : hi) = mid;
    }
    return fail(x, PBF_ERR_INVALID,
                "n = " + std::to_string(n) + " needs " + std::to_string(need) +
                    " bytes of the fixed step workspace, which holds " +
                    std::to_string(kNoAllocBytes) + " (at most " + std::to_string(lo) +
                    " particles with this grid and max_neighbors)");
  }
  HIPCHK(x, hipSetDevice(x->device));
  unsigned char* base = nullptr;
  int rc = scratch_address(x, &base);
  if (rc != PBF_OK) return rc;
  Borrow keep(x);
  pbf::DevState d = x->d;  // keeps the mesh collider's buffers
  carve_state(base, (size_t)n, x->ncell, K, &d);
  x->d = d;
  x->n = n;
  x->npad = (n + 63) / 64 * 64;
  const unsigned nb = (unsigned)((n + 255) / 256);
  // flags[0]: neighbour overflow; [2], [3]: first bad particle (none: ~0).
  HIPCHK(x, pbf::launch_fill_u32(d.flags, 2, 0u, x->stream));
  HIPCHK(x, pbf::launch_fill_u32(d.flags + 2, 2, 0xffffffffu, x->stream));
  k_noalloc_pack<<<nb, 256, 0, x->stream>>>((uint32_t)n, xyz, v, d.posA, d.velA, x->grids[0],
                                             d.flags);
  HIPCHK(x, hipGetLastError());
  uint32_t bad[4] = {0, 0, 0, 0};
  HIPCHK(x, hipMemcpyAsync(bad, d.flags, sizeof bad, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  if (bad[2] != 0xffffffffu)
    return fail(x, PBF_ERR_INVALID, "particle " + std::to_string(bad[2]) + " is not finite");
  if (bad[3] != 0xffffffffu)
    return fail(x, PBF_ERR_OUT_OF_BOX,
                "particle " + std::to_string(bad[3]) + " lies outside the collision box");
  const SimConst k = pbf::consts_for(*x, dt);
  for (int s = 0; s < nsteps; s++) {
    rc = run_step(x, k, 0, nullptr, s > 0, s + 1 < nsteps);
    if (rc != PBF_OK) return rc;
  }
  k_noalloc_unpack<<<nb, 256, 0, x->stream>>>((uint32_t)n, d.posA, d.velA, xyz, v);
  HIPCHK(x, hipGetLastError());
  HIPCHK(x, hipMemcpyAsync(bad, d.flags, sizeof bad, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  if (bad[0] != 0)
    return fail(x, PBF_ERR_OVERFLOW,
                "a particle had " + std::to_string(bad[0]) + " neighbours; the list capacity is " +
                    std::to_string(K) + " (max_neighbors " + std::to_string(x->p.max_neighbors) +
                    ")");
  return PBF_OK;
}

}  // namespace

PBF_API int pbf_step_device_noalloc(pbf_ctx* c, float* xyz, float* v, int64_t n, float dt,
                                    int32_t nsteps) {
  try {
    return step_noalloc(C(c), xyz, v, n, dt, nsteps);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API double pbf_simulate_time(const pbf_ctx* c) {
  const Ctx* x = C(c);
  return x ? x->sim_time : 0.0;
}

PBF_API int64_t pbf_count(const pbf_ctx* c) {
  const Ctx* x = C(c);
  return x ? x->n : 0;
}

PBF_API const char* pbf_last_error(const pbf_ctx* c) {
  const Ctx* x = C(c);
  if (x && !x->err.empty()) return x->err.c_str();
  return g_err.c_str();
}

static int pbf_neighbors_impl(pbf_ctx* c, uint32_t* counts, uint32_t* ids, int32_t stride) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!x->has_nbrs) return fail(x, PBF_ERR_STATE, "no neighbour lists: run a step first");
  if (!counts || !ids || stride < 1) return fail(x, PBF_ERR_INVALID, "counts/ids/stride");
  const size_t n = (size_t)x->n;
  uint32_t *dc = nullptr, *di = nullptr;
  HIPCHK(x, hipMalloc((void**)&dc, n * sizeof(uint32_t)));
  hipError_t e = hipMalloc((void**)&di, n * (size_t)stride * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipFree(dc);
    return fail(x, PBF_ERR_NOMEM, "neighbour export buffer");
  }
  e = pbf::launch_nbr_export(*x, dc, di, stride);
  if (e == hipSuccess) e = hipMemcpyAsync(counts, dc, n * sizeof(uint32_t), hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ids, di, n * (size_t)stride * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  (void)hipFree(dc);
  (void)hipFree(di);
  if (e != hipSuccess) return fail(x, PBF_ERR_HIP, std::string("neighbour export: ") + hipGetErrorString(e));
  return PBF_OK;
}

PBF_API int pbf_neighbors(pbf_ctx* c, uint32_t* counts, uint32_t* ids, int32_t stride) {
  try {
    return pbf_neighbors_impl(c, counts, ids, stride);
  } catch (const std::bad_alloc&) {
    return fail(C(c), PBF_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(C(c), PBF_ERR_INVALID, std::string("internal error: ") + e.what());
  }
}

PBF_API int pbf_neighbor_stats(pbf_ctx* c, uint64_t* total, uint32_t* max_count) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!x->has_nbrs) return fail(x, PBF_ERR_STATE, "no neighbour lists: run a step first");
  if (!total || !max_count) return fail(x, PBF_ERR_INVALID, "total/max_count");
  // flags[2..3] hold the 64-bit sum, flags[4] the maximum.
  uint32_t* f = x->d.flags;
  HIPCHK(x, pbf::launch_fill_u32(f + 2, 3, 0u, x->stream));
  HIPCHK(x, pbf::launch_nbr_stats(*x, reinterpret_cast<unsigned long long*>(f + 2), f + 4));
  uint32_t h[3];
  if ((rc = copy_out(x, reinterpret_cast<const float*>(f + 2), h, sizeof h)) != PBF_OK) return rc;
  *total = (uint64_t)h[0] | ((uint64_t)h[1] << 32);
  *max_count = h[2];
  return PBF_OK;
}

PBF_API int pbf_cells(pbf_ctx* c, uint32_t* cells) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!x->has_nbrs) return fail(x, PBF_ERR_STATE, "no grid: run a step first");
  if (!cells) return fail(x, PBF_ERR_INVALID, "cells is NULL");
  uint32_t* dst = reinterpret_cast<uint32_t*>(x->d.stage);
  HIPCHK(x, pbf::launch_unpermute_u32(*x, x->d.posB, x->d.cellSorted, dst));
  return copy_out(x, x->d.stage, cells, (size_t)x->n * sizeof(uint32_t));
}

PBF_API int pbf_grid_dims(const pbf_ctx* c, int32_t scene, int32_t* dims) {
  const Ctx* x = C(c);
  if (!x || !dims || scene < 0 || scene >= x->nscenes)
    return fail(const_cast<Ctx*>(x), PBF_ERR_INVALID, "bad scene or dims");
  for (int a = 0; a < 3; a++) dims[a] = x->grids[(size_t)scene].dims[a];
  return PBF_OK;
}

PBF_API int pbf_debug_read(pbf_ctx* c, int32_t which, float* out4) {
  Ctx* x = C(c);
  int rc = check_ready(x);
  if (rc != PBF_OK) return rc;
  if (!out4) return fail(x, PBF_ERR_INVALID, "out4 is NULL");
  pbf::DevState& d = x->d;
  const float4* src = nullptr;
  const float4* ids = x->debug_dirty ? d.posB : d.posA;
  switch (which) {
    case PBF_BUF_PRED: src = x->dbg_pred; break;
    case PBF_BUF_VEL: src = x->debug_dirty ? x->dbg_vel : d.velA; break;
    case PBF_BUF_VXSPH: src = d.vxsph; break;
    case PBF_BUF_OMEGA: src = d.vort; break;
    case PBF_BUF_POS: src = d.posA; ids = d.posA; break;
    default: return fail(x, PBF_ERR_INVALID, "unknown buffer");
  }
  if (!src) return fail(x, PBF_ERR_STATE, "buffer not produced yet");
  HIPCHK(x, pbf::launch_unpermute(*x, ids, src, 0, 4, d.stage));
  return copy_out(x, d.stage, out4, (size_t)x->n * 4 * sizeof(float));
}

PBF_API int pbf_debug_derived(const pbf_ctx* c, float* out, int32_t cap) {
  const Ctx* x = C(c);
  if (!x || !out || cap < 25) return fail(const_cast<Ctx*>(x), PBF_ERR_INVALID, "cap >= 25");
  const SimConst k = pbf::consts_for(*x, x->p.dt);
  const SceneGrid& g = x->grids[0];
  const float vals[25] = {k.h,   k.h2,   k.h6,     k.h9,   k.c1,  k.m3c2, k.S,
                          k.inv_rho0, k.rho0, k.eps, k.negk, k.xc, k.dtv,  k.inv_dt,
                          k.dt,  k.gdt,  x->p.h * x->p.cell_scale, g.inv_cell, k.eps_d,
                          g.lo_e[0], g.lo_e[1], g.lo_e[2], g.hi_e[0], g.hi_e[1], g.hi_e[2]};
  memcpy(out, vals, sizeof vals);
  if (cap < 28) return 25;
  out[25] = (float)k.fast_div;
  out[26] = (float)k.fast_rcp;
  out[27] = (float)k.fast_sqrt;
  return 28;
}

}  // extern "C"
