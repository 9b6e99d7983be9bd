// pbf_host.h -- the host-side logic of the C ABI that needs no device: the
// validation and packing of a load, the snapshot file format and its
// checksum, and the triangle-mesh BVH build.  Plain C++17 with no HIP
// headers, so that it also builds on its own (the CPU tests compile it with
// AddressSanitizer and UBSan and drive it with damaged inputs).
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pbf.h"

namespace pbf {
namespace host {

// Four floats, laid out as HIP's float4 (the device arrays are float4).
struct F4 {
  float x, y, z, w;
};

inline bool finite_f(float v) { return std::isfinite(v); }

inline float bits_f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Runs f(0) .. f(n-1): f(0) on the calling thread and the rest on threads
// of their own; a part whose thread cannot be started runs on the calling
// thread, and every started thread is joined before returning.  Both vectors
// are sized before the first thread starts, so nothing below can throw
// while a started thread is still joinable.
template <class F>
void run_parts(int64_t n, F f) {
  std::vector<std::thread> th;
  std::vector<int64_t> inline_parts;
  th.reserve((size_t)std::max<int64_t>(n, 1));
  inline_parts.reserve((size_t)std::max<int64_t>(n, 1));
  for (int64_t t = 1; t < n; t++) {
    try {
      th.emplace_back(f, t);
    } catch (...) {
      inline_parts.push_back(t);
    }
  }
  f(0);
  for (int64_t t : inline_parts) f(t);
  for (auto& t : th) t.join();
}

// ---- load: validation and packing -----------------------------------------
// A load of n particles (xyz, v, rho: n*3, n*3, n floats; v and rho may be
// NULL) over the scenes that start at offsets[s], with boxes lo = boxes[6s..],
// hi = boxes[6s+3..].  Packs pos = (x, id bits), vel = (v, rho) and the
// scene of every particle, on up to 8 host threads over contiguous ranges.
// The error reported is the one a sequential pass would meet first: the
// lowest failing particle, and in it the first failing axis (finite, then
// inside the box).
struct Packed {
  std::vector<uint32_t> slot_scene;
  std::vector<F4> pos, vel;
  int code = PBF_OK;  // PBF_OK, PBF_ERR_INVALID (non-finite) or PBF_ERR_OUT_OF_BOX
  int64_t first = 0;  // the failing particle
};

inline void validate_pack(const float* xyz, const float* v, const float* rho, int64_t n,
                          const std::vector<int64_t>& offsets, const std::vector<float>& boxes,
                          int max_threads, Packed* out) {
  const int nscenes = (int)offsets.size();
  out->slot_scene.assign((size_t)n, 0u);
  out->pos.assign((size_t)n, F4{0, 0, 0, 0});
  out->vel.assign((size_t)n, F4{0, 0, 0, 0});
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned cap = (unsigned)std::max(1, max_threads);
  const int64_t nthr =
      n < ((int64_t)1 << 16) ? 1 : (int64_t)std::max(1u, std::min(cap, hw ? hw : 1u));
  const int64_t per = (n + nthr - 1) / nthr;
  std::vector<int64_t> bad_at((size_t)nthr, n);
  std::vector<int> bad_code((size_t)nthr, PBF_OK);
  auto work = [&](int64_t t) {
    const int64_t b = t * per, e = std::min(n, b + per);
    int s = 0;
    while (s + 1 < nscenes && offsets[(size_t)s + 1] <= b) s++;
    for (int64_t i = b; i < e; i++) {
      while (s + 1 < nscenes && offsets[(size_t)s + 1] <= i) s++;
      const float* lo = &boxes[6 * (size_t)s];
      const float* hi = lo + 3;
      out->slot_scene[(size_t)i] = (uint32_t)s;
      for (int a = 0; a < 3; a++) {
        const float p = xyz[3 * i + a];
        const float q = v ? v[3 * i + a] : 0.0f;
        int code = PBF_OK;
        if (!finite_f(p) || !finite_f(q)) code = PBF_ERR_INVALID;
        else if (p < lo[a] || p > hi[a]) code = PBF_ERR_OUT_OF_BOX;
        if (code != PBF_OK) {
          bad_at[(size_t)t] = i;
          bad_code[(size_t)t] = code;
          return;
        }
      }
      out->pos[(size_t)i] = F4{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], bits_f((uint32_t)i)};
      const float r = rho ? rho[i] : 0.0f;
      out->vel[(size_t)i] = v ? F4{v[3 * i], v[3 * i + 1], v[3 * i + 2], r} : F4{0, 0, 0, r};
    }
  };
  run_parts(nthr, work);
  out->code = PBF_OK;
  out->first = n;
  for (int64_t t = 0; t < nthr; t++)
    if (bad_code[(size_t)t] != PBF_OK && bad_at[(size_t)t] < out->first) {
      out->first = bad_at[(size_t)t];
      out->code = bad_code[(size_t)t];
    }
}

// ---- snapshots --------------------------------------------------------------
// File layout (little-endian), see pbf_save_snapshot in pbf.h.  Version 2
// checksums the payload with Fnv4 (written); version 1 files, checksummed
// with byte-serial FNV-1a-64, still load.
constexpr char kSnapMagic[8] = {'P', 'B', 'F', 'S', 'N', 'A', 'P', '2'};
constexpr char kSnapMagic1[8] = {'P', 'B', 'F', 'S', 'N', 'A', 'P', '1'};
struct SnapHeader {
  char magic[8];
  uint32_t header_bytes;  // sizeof(SnapHeader) = 64
  uint32_t params_bytes;  // sizeof(pbf_params)
  int64_t n;
  int32_t nscenes;
  int32_t reserved0;
  double sim_time;
  uint64_t checksum;      // Fnv4 (version 2) of every byte after the header
  uint64_t reserved1[2];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");

// FNV-1a-64 prime and offset basis; version 1 files used the basis
// 1469598103934665603 instead.
constexpr uint64_t kFnvPrime = 1099511628211ull, kFnvBasis = 14695981039346656037ull;
constexpr uint64_t kFnvBasisV1 = 1469598103934665603ull;

// Byte-serial FNV-1a-64 from basis h0.
struct Fnv {
  uint64_t h;
  explicit Fnv(uint64_t h0 = kFnvBasis) : h(h0) {}
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * kFnvPrime;
  }
  uint64_t value() const { return h; }
};

// Fnv4: the FNV-1a step on 64-bit little-endian words in four lanes, so the
// four multiply chains run side by side.  Over the payload as one byte
// stream: 32-byte block k gives word l (bytes 32k + 8l ..) to lane l; the
// bytes after the last whole block go byte by byte into lane 0; the value is
// byte-serial FNV-1a-64 of the four lane values (32 bytes, lane 0 first).
// Every lane step is a bijection, so any change confined to one lane's words
// changes the value.
struct Fnv4 {
  uint64_t h[4] = {kFnvBasis, kFnvBasis, kFnvBasis, kFnvBasis};
  unsigned char buf[32];
  size_t nbuf = 0;
  void block(const unsigned char* b) {
    uint64_t w[4];
    memcpy(w, b, 32);
    for (int l = 0; l < 4; l++) h[l] = (h[l] ^ w[l]) * kFnvPrime;
  }
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    while (n > 0 && nbuf > 0) {
      buf[nbuf++] = *b++;
      n--;
      if (nbuf == 32) {
        block(buf);
        nbuf = 0;
      }
    }
    for (; n >= 32; n -= 32, b += 32) block(b);
    while (n > 0) {
      buf[nbuf++] = *b++;
      n--;
    }
  }
  uint64_t value() const {
    uint64_t t = h[0];
    for (size_t i = 0; i < nbuf; i++) t = (t ^ buf[i]) * kFnvPrime;
    const uint64_t lanes[4] = {t, h[1], h[2], h[3]};
    Fnv f;
    f.add(lanes, sizeof lanes);
    return f.value();
  }
};

// The payload after the header: params, per-scene counts and boxes, x, v, rho.
struct SnapBody {
  pbf_params p{};
  std::vector<int64_t> counts;
  std::vector<float> boxes;  // lo[3], hi[3] per scene
  std::vector<float> x, v, rho;
  template <class F>
  void each(F f) {
    f(&p, sizeof p);
    f(counts.data(), counts.size() * sizeof(int64_t));
    f(boxes.data(), boxes.size() * sizeof(float));
    f(x.data(), x.size() * sizeof(float));
    f(v.data(), v.size() * sizeof(float));
    f(rho.data(), rho.size() * sizeof(float));
  }
};

// Bytes of a file with n particles in nscenes scenes.
inline uint64_t snapshot_bytes(int64_t n, int32_t nscenes) {
  return sizeof(SnapHeader) + sizeof(pbf_params) + (uint64_t)nscenes * (8 + 24) +
         (uint64_t)n * 28;
}

// Writes b (n particles) with simulate_time t; returns "" or the reason.
inline std::string write_snapshot(const char* path, SnapBody& b, int64_t n, double t) {
  SnapHeader hd{};
  memcpy(hd.magic, kSnapMagic, 8);
  hd.header_bytes = sizeof(SnapHeader);
  hd.params_bytes = sizeof(pbf_params);
  hd.n = n;
  hd.nscenes = (int32_t)b.counts.size();
  hd.sim_time = t;
  Fnv4 f;
  b.each([&](const void* p, size_t k) { f.add(p, k); });
  hd.checksum = f.value();
  FILE* fp = fopen(path, "wb");
  if (!fp) return std::string("cannot open ") + path + " for writing";
  bool ok = fwrite(&hd, sizeof hd, 1, fp) == 1;
  b.each([&](const void* p, size_t k) { ok = ok && (k == 0 || fwrite(p, 1, k, fp) == k); });
  ok = (fclose(fp) == 0) && ok;
  if (!ok) return std::string("write to ") + path + " failed";
  return "";
}

// Reads a snapshot into b and hd, checking it against the layout of the
// context it is loaded into (mine: its params, per-scene counts and boxes;
// with several scenes the counts must match).  Returns "" or the reason.
// The file size is checked against the header before any payload buffer is
// sized, so a damaged header cannot ask for more memory than the file holds.
inline std::string read_snapshot(const char* path, const SnapBody& mine, SnapHeader* hd,
                                 SnapBody* b) {
  FILE* fp = fopen(path, "rb");
  if (!fp) return std::string("cannot open ") + path;
  const int nscenes = (int)mine.counts.size();
  *hd = SnapHeader{};
  const bool head_ok = fread(hd, sizeof *hd, 1, fp) == 1;
  std::string why;
  const bool v1 = head_ok && memcmp(hd->magic, kSnapMagic1, 8) == 0;
  long long fsize = -1;
  if (head_ok && fseek(fp, 0, SEEK_END) == 0) {
    fsize = ftell(fp);
    if (fseek(fp, (long)sizeof(SnapHeader), SEEK_SET) != 0) fsize = -1;
  }
  if (!head_ok || (memcmp(hd->magic, kSnapMagic, 8) != 0 && !v1) ||
      hd->header_bytes != sizeof(SnapHeader))
    why = "not a snapshot file";
  else if (hd->params_bytes != sizeof(pbf_params))
    why = "snapshot written with a different pbf_params layout";
  else if (hd->nscenes != nscenes || hd->n < 1 || hd->n >= ((int64_t)1 << 31))
    why = "snapshot holds " + std::to_string(hd->nscenes) + " scene(s), the context " +
          std::to_string(nscenes);
  else if (fsize < 0 || (uint64_t)fsize != snapshot_bytes(hd->n, hd->nscenes))
    why = "snapshot file truncated or too long";
  if (why.empty()) {
    const size_t n = (size_t)hd->n;
    b->counts.resize((size_t)hd->nscenes);
    b->boxes.resize(6 * (size_t)hd->nscenes);
    b->x.resize(3 * n);
    b->v.resize(3 * n);
    b->rho.resize(n);
    bool ok = true;
    b->each([&](void* p, size_t k) { ok = ok && (k == 0 || fread(p, 1, k, fp) == k); });
    uint64_t sum;
    if (v1) {
      Fnv f(kFnvBasisV1);
      b->each([&](const void* p, size_t k) { f.add(p, k); });
      sum = f.value();
    } else {
      Fnv4 f;
      b->each([&](const void* p, size_t k) { f.add(p, k); });
      sum = f.value();
    }
    if (!ok || fgetc(fp) != EOF) why = "snapshot file truncated or too long";
    else if (sum != hd->checksum) why = "snapshot checksum mismatch";
  }
  fclose(fp);
  if (why.empty()) {
    if (memcmp(&b->p, &mine.p, sizeof(pbf_params)) != 0)
      why = "snapshot written with different parameters";
    else if (nscenes > 1 && b->counts != mine.counts)
      why = "snapshot scene particle counts differ from the context's";
    else if (memcmp(b->boxes.data(), mine.boxes.data(), b->boxes.size() * sizeof(float)) != 0)
      why = "snapshot scene boxes differ from the context's";
  }
  return why;
}

// ---- mesh collider: BVH build ------------------------------------------------
// At each node, triangles whose box diagonal exceeds half the node box's go
// to a subtree of their own (the first child, index order kept), so a few
// large faces (the walls of a box around a finely meshed obstacle) do not
// widen every box on the way to the small ones: a walk near such an
// obstacle then visits ~32 nodes instead of ~125.  Otherwise (none or all
// of them large), a median split on the longest axis of the centroid
// bounds (ties by index, so the tree is deterministic).  Leaves hold at
// most 4 triangles.  Nodes are
// stored in depth-first preorder, so an inner node's left child follows it,
// and each node records its escape: the first node after its subtree.  The
// device walks the tree without a stack (a missed box or a finished leaf
// continues at the escape, a hit inner box at the next node).  Node boxes
// are padded outwards by 1e-5 of the scene scale so the float slab test is
// conservative.
struct MeshBuild {
  std::vector<uint32_t> order;   // triangle indices, leaf order
  std::vector<float> cen, lo, hi;  // per triangle: centroid, box
  std::vector<F4> nodes;         // 2 per node
  float pad = 0.0f;

  // Builds the subtree of order[b, e) and returns its node's index; the
  // subtree occupies [index, nodes.size()/2).  Iterative (an explicit stack
  // of pending ranges), so the depth of a degenerate mesh cannot exhaust
  // the host stack.
  void build(uint32_t b0, uint32_t e0) {
    struct Item {
      uint32_t b, e, node;
      bool done;
    };
    std::vector<Item> stack{{b0, e0, 0u, false}};
    while (!stack.empty()) {
      Item it = stack.back();
      stack.pop_back();
      if (it.done) {
        // Subtree complete: its escape is the next free node.
        const uint32_t ua = (uint32_t)(nodes.size() / 2);
        nodes[2 * it.node].w = bits_f(ua);
        continue;
      }
      const uint32_t ni = (uint32_t)(nodes.size() / 2);
      nodes.resize(nodes.size() + 2);
      float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
      float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t i = it.b; i < it.e; i++) {
        const uint32_t k = order[i];
        for (int a = 0; a < 3; a++) {
          bl[a] = std::min(bl[a], lo[3 * k + a]);
          bh[a] = std::max(bh[a], hi[3 * k + a]);
          cl[a] = std::min(cl[a], cen[3 * k + a]);
          ch[a] = std::max(ch[a], cen[3 * k + a]);
        }
      }
      uint32_t ub = 0;  // inner node: 0; leaf: 2^31 | count << 27 | first triangle
      nodes[2 * ni] = F4{bl[0] - pad, bl[1] - pad, bl[2] - pad, 0.0f};
      stack.push_back({it.b, it.e, ni, true});
      if (it.e - it.b <= 4) {
        ub = 0x80000000u | ((it.e - it.b) << 27) | it.b;
      } else {
        float nd2 = 0.0f;
        for (int a = 0; a < 3; a++) nd2 += (bh[a] - bl[a]) * (bh[a] - bl[a]);
        const auto large = [&](uint32_t k) {
          float d2 = 0.0f;
          for (int a = 0; a < 3; a++) d2 += (hi[3 * k + a] - lo[3 * k + a]) * (hi[3 * k + a] - lo[3 * k + a]);
          return d2 > 0.25f * nd2;
        };
        uint32_t mid = (uint32_t)(std::stable_partition(order.begin() + it.b,
                                                        order.begin() + it.e, large) -
                                  order.begin());
        if (mid == it.b || mid == it.e) {
          int ax = 0;
          for (int a = 1; a < 3; a++)
            if (ch[a] - cl[a] > ch[ax] - cl[ax]) ax = a;
          mid = it.b + (it.e - it.b) / 2;
          std::nth_element(order.begin() + it.b, order.begin() + mid, order.begin() + it.e,
                           [&](uint32_t x, uint32_t y) {
                             const float cx = cen[3 * x + ax], cy = cen[3 * y + ax];
                             return cx < cy || (cx == cy && x < y);
                           });
        }
        // Left child (node ni + 1) first: pushed last.
        stack.push_back({mid, it.e, 0u, false});
        stack.push_back({it.b, mid, 0u, false});
      }
      nodes[2 * ni + 1] = F4{bh[0] + pad, bh[1] + pad, bh[2] + pad, bits_f(ub)};
    }
  }
};

// The scene grid a mesh's cell flags refer to.
struct MeshGrid {
  float lo[3], hi[3];  // scene box
  float inv_cell;
  int dims[3];
};

// What the device reads of a mesh: triangle records in leaf order (p1 with
// the original index bits, e1 = p2 - p1, e2 = p3 - p1, the three vertex
// normals), BVH nodes, one flag byte per grid cell near a triangle (empty
// without cells), the box padding and the node count.
struct MeshHost {
  std::vector<F4> tris, nodes;
  std::vector<uint8_t> flags;
  float pad = 0.0f;
  uint32_t nnode = 0;
};

// Builds the device form of the mesh (nv vertices pos/nrm, nt triangles of
// vertex indices tri, checked by the caller).  Returns "" or the reason.
inline std::string build_mesh(const float* pos, const float* nrm, const int32_t* tri, int64_t nt,
                              const MeshGrid& g, bool use_cells, MeshHost* out) {
  MeshBuild mb;
  const size_t n = (size_t)nt;
  mb.order.resize(n);
  mb.cen.resize(3 * n);
  mb.lo.resize(3 * n);
  mb.hi.resize(3 * n);
  float scale = 0.0f;
  for (size_t k = 0; k < n; k++) {
    mb.order[k] = (uint32_t)k;
    for (int a = 0; a < 3; a++) {
      const float q0 = pos[3 * (size_t)tri[3 * k] + a], q1 = pos[3 * (size_t)tri[3 * k + 1] + a],
                  q2 = pos[3 * (size_t)tri[3 * k + 2] + a];
      mb.lo[3 * k + a] = std::min(q0, std::min(q1, q2));
      mb.hi[3 * k + a] = std::max(q0, std::max(q1, q2));
      mb.cen[3 * k + a] = (q0 + q1 + q2) / 3.0f;
      scale = std::max(scale, std::max(fabsf(mb.lo[3 * k + a]), fabsf(mb.hi[3 * k + a])));
    }
  }
  for (int a = 0; a < 3; a++) scale = std::max(scale, std::max(fabsf(g.lo[a]), fabsf(g.hi[a])));
  mb.pad = 1e-5f * std::max(scale, 1e-6f);
  mb.build(0, (uint32_t)n);
  // The device walk moves to a larger index at every node, so it ends after
  // at most nnode nodes; check the links that guarantee it.
  const uint32_t nnode = (uint32_t)(mb.nodes.size() / 2);
  for (uint32_t i = 0; i < nnode; i++) {
    uint32_t esc, b;
    memcpy(&esc, &mb.nodes[2 * i].w, 4);
    memcpy(&b, &mb.nodes[2 * i + 1].w, 4);
    const bool leaf = (b & 0x80000000u) != 0;
    if (!(esc > i && esc <= nnode) || (!leaf && i + 1 >= nnode) ||
        (leaf && (((b >> 27) & 7u) == 0 || (b & 0x7ffffffu) + ((b >> 27) & 7u) > n)))
      return "mesh BVH: inconsistent node links";
  }
  out->tris.assign(6 * n, F4{0, 0, 0, 0});
  for (size_t i = 0; i < n; i++) {
    const uint32_t k = mb.order[i];
    const float* p1 = pos + 3 * (size_t)tri[3 * k];
    const float* p2 = pos + 3 * (size_t)tri[3 * k + 1];
    const float* p3 = pos + 3 * (size_t)tri[3 * k + 2];
    out->tris[6 * i] = F4{p1[0], p1[1], p1[2], bits_f(k)};
    out->tris[6 * i + 1] = F4{p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2], 0.0f};
    out->tris[6 * i + 2] = F4{p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2], 0.0f};
    for (int j = 0; j < 3; j++) {
      const float* q = nrm + 3 * (size_t)tri[3 * k + j];
      out->tris[6 * i + 3 + j] = F4{q[0], q[1], q[2], 0.0f};
    }
  }
  // Cells of the scene grid near a triangle: a segment whose box (grown by
  // pad) touches none of them skips the BVH walk.  Triangle boxes are grown
  // by 3 pad here (a hit point lies within pad/2 of its triangle's box), far
  // more than the float cell indexing on the device can err; if marking
  // would visit more than 64 cells per grid cell, every cell is flagged (the
  // walk always runs).  The flags are then dilated by one cell.
  const size_t ncell = use_cells ? (size_t)g.dims[0] * (size_t)g.dims[1] * (size_t)g.dims[2] : 0;
  out->flags.assign(ncell, 0);
  if (use_cells) {
    std::vector<uint8_t>& flag = out->flags;
    size_t work = 0;
    bool all = false;
    auto cell = [&](double v, int a) {
      const double f = std::floor((v - (double)g.lo[a]) * (double)g.inv_cell);
      if (!(f >= 0.0)) return 0;
      if (!(f < (double)g.dims[a])) return g.dims[a] - 1;
      return (int)f;
    };
    for (size_t k = 0; k < n && !all; k++) {
      int c0[3], c1[3];
      size_t cells = 1;
      for (int a = 0; a < 3; a++) {
        c0[a] = cell((double)mb.lo[3 * k + a] - 3.0 * mb.pad, a);
        c1[a] = cell((double)mb.hi[3 * k + a] + 3.0 * mb.pad, a);
        cells *= (size_t)(c1[a] - c0[a] + 1);
      }
      work += cells;
      if (work > 64 * ncell) {
        all = true;
        break;
      }
      for (int z = c0[2]; z <= c1[2]; z++)
        for (int y = c0[1]; y <= c1[1]; y++)
          for (int xx = c0[0]; xx <= c1[0]; xx++)
            flag[((size_t)z * g.dims[1] + y) * g.dims[0] + xx] = 1;
    }
    if (all) std::fill(flag.begin(), flag.end(), (uint8_t)1);
    // Dilate by one cell on every axis (three passes of a 3-wide OR): the
    // device then reads only the start cell of a displacement.
    const size_t nx = (size_t)g.dims[0], ny = (size_t)g.dims[1], nz = (size_t)g.dims[2];
    const size_t stride[3] = {1, nx, nx * ny}, dim[3] = {nx, ny, nz};
    std::vector<uint8_t> tmp(ncell);
    for (int a = 0; a < 3 && !all; a++) {
      for (size_t c = 0; c < ncell; c++) {
        const size_t k = (c / stride[a]) % dim[a];
        uint8_t v = flag[c];
        if (k > 0) v |= flag[c - stride[a]];
        if (k + 1 < dim[a]) v |= flag[c + stride[a]];
        tmp[c] = v;
      }
      flag.swap(tmp);
    }
  }
  out->nodes = std::move(mb.nodes);
  out->pad = mb.pad;
  out->nnode = nnode;
  return "";
}

}  // namespace host
}  // namespace pbf
