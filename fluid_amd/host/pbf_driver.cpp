// pbf_driver.cpp -- headless simulate-and-save driver over the C++ facade.
//
// Does what the reference's `pathtracer -p particles.xml -d T` path does for
// the simulation (src/main.cpp:110-117, src/application.cpp:302-344,
// src/pathtracer.cpp:444-480): load the particle XML, construct Particles
// with the file's rest density, add every particle, estimateDensities, then
// call timeStep() until simulate_time has advanced by T seconds, printing the
// average density of every step.  The rendering of each frame stays with the
// reference's CPU path tracer; this driver writes the final particles back in
// the same XML format instead (-o), for the CPU consumers.
//
// Checkpoint/resume: -s writes a binary snapshot of the final state, and
// -r resumes from one (the particles of -p then only give the rest
// density, which must be the snapshot's).
//
//   pbf_driver -p particles.xml [-d seconds] [-t dt] [-o out.xml] [-s out.snap]
//              [-r in.snap] [-q]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pbf_particles.hpp"

namespace {

struct Vec3 {
  double x, y, z;
  Vec3(double a = 0, double b = 0, double c = 0) : x(a), y(b), z(c) {}
};

bool text_of(const std::string& s, size_t from, size_t to, const char* tag, std::string* out,
             size_t* end) {
  const std::string open = std::string("<") + tag + ">", close = std::string("</") + tag + ">";
  size_t a = s.find(open, from);
  if (a == std::string::npos || a >= to) return false;
  a += open.size();
  size_t b = s.find(close, a);
  if (b == std::string::npos || b > to) return false;
  *out = s.substr(a, b - a);
  if (end) *end = b + close.size();
  return true;
}

bool vec_of(const std::string& t, Vec3* v) {
  std::stringstream ss(t);  // like Application::stov (src/application.cpp:293-300)
  return static_cast<bool>(ss >> v->x >> v->y >> v->z);
}

// The reference's particle XML: <particles><density/><ps><particle><pos/><v/>.
int load_xml(const char* path, double* rho0, std::vector<Vec3>* pos, std::vector<Vec3>* vel) {
  std::ifstream f(path);
  if (!f) {
    std::fprintf(stderr, "[ERROR] particle file %s not found\n", path);
    return 2;
  }
  std::stringstream buf;
  buf << f.rdbuf();
  const std::string s = buf.str();
  if (s.find("<particles>") == std::string::npos) {
    std::fprintf(stderr, "[ERROR] %s: not a particles file\n", path);
    return 2;
  }
  std::string t;
  if (!text_of(s, 0, s.size(), "density", &t, nullptr)) {
    std::fprintf(stderr, "[ERROR] %s: no <density>\n", path);
    return 2;
  }
  *rho0 = (double)std::strtof(t.c_str(), nullptr);  // stof, src/application.cpp:331
  size_t at = 0;
  for (;;) {
    size_t a = s.find("<particle>", at);
    if (a == std::string::npos) break;
    size_t b = s.find("</particle>", a);
    if (b == std::string::npos) {
      std::fprintf(stderr, "[ERROR] %s: unterminated <particle>\n", path);
      return 2;
    }
    std::string ps, vs;
    Vec3 p, v;
    if (!text_of(s, a, b, "pos", &ps, nullptr) || !text_of(s, a, b, "v", &vs, nullptr) ||
        !vec_of(ps, &p) || !vec_of(vs, &v)) {
      std::fprintf(stderr, "[ERROR] %s: malformed particle at byte %zu\n", path, a);
      return 2;
    }
    pos->push_back(p);
    vel->push_back(v);
    at = b;
  }
  return 0;
}

int save_xml(const char* path, double rho0, pbf::ParticlesT<Vec3>& ps) {
  std::FILE* f = std::fopen(path, "w");
  if (!f) return 1;
  std::fprintf(f, "<particles>\n  <density>%.9g</density>\n  <ps>\n", rho0);
  for (size_t i = 0; i < ps.size(); i++) {
    const Vec3 p = ps.getPosition(i), v = ps.getVelocity(i);
    std::fprintf(f,
                 "    <particle>\n      <pos>%.9g %.9g %.9g</pos>\n      <v>%.9g %.9g %.9g</v>\n"
                 "    </particle>\n",
                 p.x, p.y, p.z, v.x, v.y, v.z);
  }
  std::fprintf(f, "  </ps>\n</particles>\n");
  return std::fclose(f) == 0 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  const char* in = nullptr;
  const char* out = nullptr;
  const char* snap_out = nullptr;
  const char* snap_in = nullptr;
  double seconds = 1.0, dt = 0.0;
  bool quiet = false;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "-p") && i + 1 < argc) in = argv[++i];
    else if (!std::strcmp(argv[i], "-d") && i + 1 < argc) seconds = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "-t") && i + 1 < argc) dt = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
    else if (!std::strcmp(argv[i], "-s") && i + 1 < argc) snap_out = argv[++i];
    else if (!std::strcmp(argv[i], "-r") && i + 1 < argc) snap_in = argv[++i];
    else if (!std::strcmp(argv[i], "-q")) quiet = true;
    else {
      std::fprintf(stderr,
                   "usage: %s -p particles.xml [-d seconds] [-t dt] [-o out.xml] [-s out.snap] "
                   "[-r in.snap] [-q]\n",
                   argv[0]);
      return 2;
    }
  }
  if (!in) {
    std::fprintf(stderr, "[Warning] use -p <particle_file_path>\n");
    return 2;
  }
  double rho0 = 0;
  std::vector<Vec3> pos, vel;
  int rc = load_xml(in, &rho0, &pos, &vel);
  if (rc) return rc;
  try {
    pbf::ParticlesT<Vec3> particles(rho0, nullptr, quiet);
    if (snap_in) {
      particles.loadSnapshot(snap_in);
      std::printf("[Fluid Simulation] resumed %zu particles at t=%.6f from %s\n",
                  particles.size(), particles.simulate_time, snap_in);
    } else {
      for (size_t i = 0; i < pos.size(); i++) particles.addParticle(pos[i], vel[i]);
      particles.estimateDensities();
      std::printf("[Fluid Simulation] %zu particles, rest density %g\n", pos.size(), rho0);
    }
    const double start = particles.simulate_time;
    int steps = 0;
    while (particles.simulate_time < start + seconds) {
      if (dt > 0) particles.timeStep(dt);
      else particles.timeStep();
      steps++;
      if (!quiet) {
        double ds = 0.0;
        for (size_t i = 0; i < particles.size(); i++) ds += particles.getDensity(i);
        std::printf("step %d t=%.4f avg rho: %.6g\n", steps, particles.simulate_time,
                    ds / (double)particles.size());
      }
    }
    double cx = 0, cy = 0, cz = 0;
    for (size_t i = 0; i < particles.size(); i++) {
      const Vec3 p = particles.getPosition(i);
      cx += p.x; cy += p.y; cz += p.z;
    }
    const double np = (double)particles.size();
    std::printf("done: %d steps, simulate_time %.6f, centroid %.9g %.9g %.9g\n", steps,
                particles.simulate_time, cx / np, cy / np, cz / np);
    if (snap_out) particles.saveSnapshot(snap_out);
    if (out && save_xml(out, rho0, particles) != 0) {
      std::fprintf(stderr, "[ERROR] cannot write %s\n", out);
      return 1;
    }
  } catch (const pbf::Error& e) {
    std::fprintf(stderr, "[ERROR] %s (code %d)\n", e.what(), e.code());
    return 1;
  }
  return 0;
}
