// Drives the device-free host code of the C ABI (fluid_amd/csrc/pbf_host.h)
// with valid and damaged inputs; built with AddressSanitizer and UBSan by
// tests/test_sanitize.py.  Exits 0 when every check holds; a failed check
// prints its line and exits 1; a sanitizer finding aborts.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "pbf_host.h"

using namespace pbf::host;

static int g_fail = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      fprintf(stderr, "host_checks:%d: check failed: %s\n", __LINE__, #cond); \
      g_fail = 1;                                                    \
    }                                                                \
  } while (0)

static uint32_t id_bits(const F4& q) {
  uint32_t u;
  memcpy(&u, &q.w, 4);
  return u;
}

static void check_run_parts() {
  for (int64_t n : {1, 2, 7, 33}) {
    std::vector<std::atomic<int>> hits((size_t)n);
    for (auto& h : hits) h = 0;
    run_parts(n, [&](int64_t t) { hits[(size_t)t]++; });
    for (auto& h : hits) CHECK(h == 1);
  }
}

// Three scenes of unequal size, valid input, then planted bad particles.
static void check_validate_pack() {
  const int64_t n = 200000;
  const std::vector<int64_t> offsets = {0, 70000, 150000};
  const std::vector<float> boxes = {0, 0, 0, 1, 1, 1,  -2, -2, -2, 2, 2, 2,  5, 5, 5, 6, 7, 8};
  std::mt19937 rng(5);
  std::vector<float> x(3 * n), v(3 * n), rho(n);
  for (int64_t i = 0; i < n; i++) {
    const int s = i < 70000 ? 0 : (i < 150000 ? 1 : 2);
    for (int a = 0; a < 3; a++) {
      const float lo = boxes[6 * s + a], hi = boxes[6 * s + 3 + a];
      x[3 * i + a] = lo + (hi - lo) * std::uniform_real_distribution<float>(0, 1)(rng);
      v[3 * i + a] = std::uniform_real_distribution<float>(-1, 1)(rng);
    }
    rho[i] = (float)i;
  }
  // Exactly on the faces is inside.
  x[0] = 0.0f;
  x[3 * 69999 + 2] = 1.0f;
  for (int threads : {1, 8}) {
    Packed pk;
    validate_pack(x.data(), v.data(), rho.data(), n, offsets, boxes, threads, &pk);
    CHECK(pk.code == PBF_OK && pk.first == n);
    CHECK(pk.slot_scene[0] == 0 && pk.slot_scene[69999] == 0 && pk.slot_scene[70000] == 1 &&
          pk.slot_scene[n - 1] == 2);
    for (int64_t i : {(int64_t)0, (int64_t)123456, n - 1}) {
      CHECK(id_bits(pk.pos[(size_t)i]) == (uint32_t)i);
      CHECK(pk.pos[(size_t)i].x == x[3 * i] && pk.vel[(size_t)i].z == v[3 * i + 2]);
      CHECK(pk.vel[(size_t)i].w == rho[i]);
    }
    Packed nov;
    validate_pack(x.data(), nullptr, nullptr, n, offsets, boxes, threads, &nov);
    CHECK(nov.code == PBF_OK && nov.vel[5].x == 0.0f && nov.vel[5].w == 0.0f);
  }
  // The lowest bad particle wins, whatever thread found it; within one
  // particle the first failing axis decides (finite before box).
  std::vector<float> y = x;
  y[3 * 190000 + 1] = NAN;        // scene 2, late
  y[3 * 80000 + 0] = 9.0f;        // scene 1: outside [-2, 2]
  y[3 * 150000 + 2] = 4.0f;       // scene 2: outside [5, 8]
  for (int threads : {1, 3, 8}) {
    Packed pk;
    validate_pack(y.data(), v.data(), nullptr, n, offsets, boxes, threads, &pk);
    CHECK(pk.code == PBF_ERR_OUT_OF_BOX && pk.first == 80000);
  }
  std::vector<float> w = v;
  w[3 * 80000 + 2] = INFINITY;    // the same particle: axis 0 is out of the box first
  Packed pk;
  validate_pack(y.data(), w.data(), nullptr, n, offsets, boxes, 8, &pk);
  CHECK(pk.code == PBF_ERR_OUT_OF_BOX && pk.first == 80000);
  w[3 * 80000 + 0] = NAN;         // now axis 0 is not finite
  validate_pack(y.data(), w.data(), nullptr, n, offsets, boxes, 8, &pk);
  CHECK(pk.code == PBF_ERR_INVALID && pk.first == 80000);
}

static std::vector<unsigned char> slurp(const std::string& path) {
  std::vector<unsigned char> b;
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return b;
  int c;
  while ((c = fgetc(f)) != EOF) b.push_back((unsigned char)c);
  fclose(f);
  return b;
}

static void spit(const std::string& path, const std::vector<unsigned char>& b) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return;
  if (!b.empty()) fwrite(b.data(), 1, b.size(), f);
  fclose(f);
}

static void check_snapshots(const std::string& dir) {
  SnapBody b;
  memset(&b.p, 0, sizeof b.p);
  b.p.h = 0.1f;
  b.p.iterations = 4;
  b.counts = {400, 250, 350};
  b.boxes = {0, 0, 0, 1, 1, 1, -1, -1, -1, 1, 1, 1, 0, 0, 0, 3, 3, 3};
  const int64_t n = 1000;
  b.x.resize(3 * n);
  b.v.resize(3 * n);
  b.rho.resize(n);
  for (int64_t i = 0; i < 3 * n; i++) {
    b.x[(size_t)i] = 0.001f * (float)i;
    b.v[(size_t)i] = -0.5f + 0.0003f * (float)i;
  }
  for (int64_t i = 0; i < n; i++) b.rho[(size_t)i] = 1000.0f + (float)i;
  const std::string path = dir + "/ok.pbfsnap";
  CHECK(write_snapshot(path.c_str(), b, n, 1.25) == "");
  SnapBody mine;
  mine.p = b.p;
  mine.counts = b.counts;
  mine.boxes = b.boxes;
  SnapHeader hd;
  SnapBody got;
  CHECK(read_snapshot(path.c_str(), mine, &hd, &got) == "");
  CHECK(hd.n == n && hd.nscenes == 3 && hd.sim_time == 1.25);
  CHECK(got.x == b.x && got.v == b.v && got.rho == b.rho && got.counts == b.counts);
  CHECK(slurp(path).size() == snapshot_bytes(n, 3));

  const std::vector<unsigned char> good = slurp(path);
  auto expect = [&](const std::vector<unsigned char>& bytes, const char* why) {
    const std::string p2 = dir + "/bad.pbfsnap";
    spit(p2, bytes);
    SnapHeader h2;
    SnapBody g2;
    const std::string r = read_snapshot(p2.c_str(), mine, &h2, &g2);
    if (r.find(why) == std::string::npos) {
      fprintf(stderr, "host_checks: expected \"%s\", got \"%s\"\n", why, r.c_str());
      g_fail = 1;
    }
  };
  std::vector<unsigned char> t = good;
  t.pop_back();
  expect(t, "truncated or too long");
  t = good;
  t.push_back(0);
  expect(t, "truncated or too long");
  expect(std::vector<unsigned char>(good.begin(), good.begin() + 40), "not a snapshot file");
  expect(std::vector<unsigned char>(), "not a snapshot file");
  t = good;
  t[0] = 'X';
  expect(t, "not a snapshot file");
  t = good;
  t[good.size() - 5] ^= 0x10;  // a bit of rho
  expect(t, "checksum mismatch");
  // A header that claims 2^30 particles: refused by size, before any buffer
  // of that size is allocated.
  t = good;
  const int64_t huge = (int64_t)1 << 30;
  memcpy(t.data() + 16, &huge, 8);
  expect(t, "truncated or too long");
  t = good;
  const int32_t two = 2;
  memcpy(t.data() + 24, &two, 4);
  expect(t, "scene(s)");
  // Parameters and scenes of another context.
  SnapBody other = mine;
  other.p.h = 0.2f;
  SnapHeader h3;
  SnapBody g3;
  CHECK(read_snapshot(path.c_str(), other, &h3, &g3).find("different parameters") !=
        std::string::npos);
  other = mine;
  other.counts = {401, 249, 350};
  CHECK(read_snapshot(path.c_str(), other, &h3, &g3).find("particle counts") != std::string::npos);
  other = mine;
  other.boxes[4] = 2.0f;
  CHECK(read_snapshot(path.c_str(), other, &h3, &g3).find("boxes differ") != std::string::npos);
  CHECK(read_snapshot((dir + "/missing.pbfsnap").c_str(), mine, &h3, &g3).find("cannot open") == 0);
  // Version 1: byte-serial FNV-1a from the old basis, still accepted.
  t = good;
  memcpy(t.data(), kSnapMagic1, 8);
  Fnv f(kFnvBasisV1);
  f.add(t.data() + sizeof(SnapHeader), t.size() - sizeof(SnapHeader));
  const uint64_t sum = f.value();
  memcpy(t.data() + 40, &sum, 8);
  spit(dir + "/v1.pbfsnap", t);
  SnapBody g4;
  CHECK(read_snapshot((dir + "/v1.pbfsnap").c_str(), mine, &h3, &g4) == "" && g4.rho == b.rho);
  // Fnv4 does not depend on how the stream is cut into add() calls.
  std::vector<unsigned char> stream(good.begin(), good.end());
  Fnv4 whole;
  whole.add(stream.data(), stream.size());
  std::mt19937 rng(3);
  for (int trial = 0; trial < 20; trial++) {
    Fnv4 parts;
    size_t at = 0;
    while (at < stream.size()) {
      const size_t k = std::min(stream.size() - at, (size_t)(rng() % 97));
      parts.add(stream.data() + at, k);
      at += k;
    }
    CHECK(parts.value() == whole.value());
  }
}

// Random triangle soups and degenerate meshes: every triangle in exactly one
// leaf, leaf boxes contain their triangles, inner boxes their children,
// every cell holding a vertex flagged.
static void check_mesh() {
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  for (int trial = 0; trial < 6; trial++) {
    const int nt = trial == 0 ? 1 : (trial == 1 ? 5 : (trial == 2 ? 3000 : 4000));
    const bool same = trial >= 4;  // every triangle at the same place
    const int nv = 3 * nt;
    std::vector<float> pos(3 * nv), nrm(3 * nv);
    std::vector<int32_t> tri(3 * nt);
    for (int i = 0; i < nv; i++)
      for (int a = 0; a < 3; a++) {
        pos[3 * i + a] = same ? 0.5f + 0.01f * (float)(i % 3) * (float)(a + 1) : 3.0f * u(rng);
        nrm[3 * i + a] = a == 1 ? 1.0f : 0.0f;
      }
    for (int k = 0; k < 3 * nt; k++) tri[k] = trial == 5 ? k % 3 : k;
    MeshGrid g{{0, 0, 0}, {3, 3, 3}, 10.0f, {30, 30, 30}};
    MeshHost mh;
    CHECK(build_mesh(pos.data(), nrm.data(), tri.data(), nt, g, true, &mh) == "");
    CHECK(mh.tris.size() == 6 * (size_t)nt && mh.nnode * 2 == mh.nodes.size());
    std::vector<int> seen(nt, 0);
    for (uint32_t i = 0; i < mh.nnode; i++) {
      uint32_t b;
      memcpy(&b, &mh.nodes[2 * i + 1].w, 4);
      const F4 lo = mh.nodes[2 * i], hi = mh.nodes[2 * i + 1];
      if (b & 0x80000000u) {
        const uint32_t first = b & 0x7ffffffu, cnt = (b >> 27) & 7u;
        for (uint32_t t = first; t < first + cnt; t++) {
          uint32_t k;
          memcpy(&k, &mh.tris[6 * t].w, 4);
          CHECK(k < (uint32_t)nt);
          if (k < (uint32_t)nt) seen[k]++;
          for (int j = 0; j < 3; j++) {
            const float* q = &pos[3 * (size_t)tri[3 * k + j]];
            CHECK(q[0] >= lo.x && q[0] <= hi.x && q[1] >= lo.y && q[1] <= hi.y && q[2] >= lo.z &&
                  q[2] <= hi.z);
          }
        }
      } else {
        const F4 clo = mh.nodes[2 * (i + 1)], chi = mh.nodes[2 * (i + 1) + 1];
        CHECK(clo.x >= lo.x && clo.y >= lo.y && clo.z >= lo.z && chi.x <= hi.x && chi.y <= hi.y &&
              chi.z <= hi.z);
      }
    }
    for (int k = 0; k < nt; k++) CHECK(seen[k] == 1);
    for (int i = 0; i < nv; i++) {
      int c[3];
      for (int a = 0; a < 3; a++) c[a] = std::min(29, std::max(0, (int)(pos[3 * i + a] * 10.0f)));
      CHECK(mh.flags[((size_t)c[2] * 30 + c[1]) * 30 + c[0]] == 1);
    }
    MeshHost nocells;
    CHECK(build_mesh(pos.data(), nrm.data(), tri.data(), nt, g, false, &nocells) == "");
    CHECK(nocells.flags.empty() && nocells.nodes.size() == mh.nodes.size());
  }
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  check_run_parts();
  check_validate_pack();
  check_snapshots(dir);
  check_mesh();
  if (g_fail) return 1;
  printf("host_checks: all checks passed\n");
  return 0;
}
