// pbf_particles.hpp -- C++ facade with the reference's Particles interface
// over the C ABI of pbf.h.
//
// Mirrors struct Particles of the reference (src/particles.h:105-140) for the
// members its callers use: the constructor Particles(double rest_density)
// (src/application.cpp:307, :332), addParticle (:338), estimateDensities
// (:341), timeStep() / timeStep(double) (src/pathtracer.cpp:455, :468),
// simulate_time (:448-475), the bvh member (assigned at
// src/pathtracer.cpp:266) and paramsString (src/particles.cpp:420-438).
// Positions, velocities and densities are read back lazily into host
// shadows for the CPU consumers (marching cubes, the path tracer, the GL
// viewer), which in the reference read ps[i]->getPosition()
// (src/particles.cpp:305, :449).
//
// Vec3 is the caller's vector type: it needs members x, y, z convertible to
// double and a constructor Vec3(double, double, double) -- CGL::Vector3D
// (CGL/include/CGL/vector3D.h) qualifies.
#pragma once

#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pbf.h"

namespace pbf {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

template <class Vec3>
class ParticlesT {
 public:
  // The reference's parameter set (src/particles.cpp:7-48) unless `params` is
  // given; rest density as in Particles(double rest_density).  Prints the
  // parameter banner like the reference's constructor unless quiet.
  explicit ParticlesT(double rest_density = 1000.0, const pbf_params* params = nullptr,
                      bool quiet = false)
      : rest_density(rest_density) {
    if (params) {
      p_ = *params;
    } else {
      pbf_params_default(&p_);
    }
    p_.rest_density = (float)rest_density;
    check(pbf_create(&p_, &ctx_), nullptr);
    if (!quiet) {
      std::fputs(paramsString().c_str(), stdout);
      std::fflush(stdout);
    }
  }
  ~ParticlesT() { pbf_destroy(ctx_); }
  ParticlesT(const ParticlesT&) = delete;
  ParticlesT& operator=(const ParticlesT&) = delete;

  // -- the reference's interface ------------------------------------------
  // Appends to the live state: after timeStep the host copy is refreshed
  // first, so the next upload carries the simulated positions, not the ones
  // of the last load (src/particles.h:118-120 appends to the live vector).
  void addParticle(Vec3 pos, Vec3 v) {
    sync();
    x_.push_back((float)pos.x); x_.push_back((float)pos.y); x_.push_back((float)pos.z);
    v_.push_back((float)v.x);   v_.push_back((float)v.y);   v_.push_back((float)v.z);
    rho_.push_back(0.0f);
    dirty_ = true;
  }

  void timeStep(double delta_t) {
    upload();
    check(pbf_step_dt(ctx_, (float)delta_t, 1), ctx_);
    simulate_time += delta_t;
    stale_ = true;
    surfaceUpToTimestep = false;
  }
  void timeStep() { timeStep((double)p_.dt); }

  // Load-time SPH density of every particle, self included
  // (src/particles.cpp:440-444).
  void estimateDensities() {
    upload();
    sync();
    check(pbf_estimate_densities(ctx_, rho_.data()), ctx_);
  }

  // Checkpoint / resume through binary snapshots (pbf_save_snapshot).  A
  // context built with the same parameters resumes bit for bit.
  void saveSnapshot(const char* path) {
    upload();
    check(pbf_save_snapshot(ctx_, path), ctx_);
  }
  void loadSnapshot(const char* path) {
    check(pbf_load_snapshot(ctx_, path), ctx_);
    const size_t n = (size_t)pbf_count(ctx_);
    x_.assign(3 * n, 0.0f);
    v_.assign(3 * n, 0.0f);
    rho_.assign(n, 0.0f);
    simulate_time = pbf_simulate_time(ctx_);
    dirty_ = false;
    stale_ = true;
    surfaceUpToTimestep = false;
  }

  // SPH density at an arbitrary point (src/particles.cpp:446-453), on the
  // GPU over the cell grid of the committed state.  One call per point works
  // (the grid is built once per state), but the marching-cubes consumers
  // should ask for all their points at once (estimateDensitiesAt,
  // densityLattice).
  double estimateDensityAt(Vec3 pos) {
    const float q[3] = {(float)pos.x, (float)pos.y, (float)pos.z};
    float out = 0.0f;
    upload();
    check(pbf_density_at(ctx_, 0, q, 1, &out), ctx_);
    return out;
  }

  // The density at many points: generateGridCell's 8 corners
  // (src/particles.cpp:309-323) and getVertexNormal's 6 samples (:404-418).
  std::vector<double> estimateDensitiesAt(const std::vector<Vec3>& pts) {
    std::vector<float> q(3 * pts.size()), out(pts.size());
    for (size_t i = 0; i < pts.size(); i++) {
      q[3 * i] = (float)pts[i].x;
      q[3 * i + 1] = (float)pts[i].y;
      q[3 * i + 2] = (float)pts[i].z;
    }
    upload();
    check(pbf_density_at(ctx_, 0, q.data(), (int64_t)pts.size(), out.data()), ctx_);
    return std::vector<double>(out.begin(), out.end());
  }

  // The density on getSurfacePrims' corner lattice (src/particles.cpp:
  // 350-390): counts[a] = (int)((max-min)/step) + 2 points per axis at
  // min(max, min + i*step); x fastest.
  std::vector<float> densityLattice(Vec3 lo, Vec3 hi, double step, int32_t counts[3]) {
    const float l[3] = {(float)lo.x, (float)lo.y, (float)lo.z};
    const float u[3] = {(float)hi.x, (float)hi.y, (float)hi.z};
    std::vector<float> out((size_t)counts[0] * (size_t)counts[1] * (size_t)counts[2]);
    upload();
    check(pbf_density_lattice(ctx_, 0, l, u, (float)step, counts, out.data()), ctx_);
    return out;
  }

  std::string paramsString() const {
    std::stringstream ss;
    ss << "Fluid simulation parameters: " << std::endl
       << "\tTime step: " << p_.dt << std::endl
       << "\tSPH Density estimate radius H: " << p_.h << std::endl
       << "\tNewton steps: " << p_.iterations << std::endl
       << "\tConstraint relaxation epsilon: " << p_.epsilon << std::endl
       << "\tTensile artificial pressure coefficient K: " << p_.k_tensile << std::endl
       << "\tTensile artificial pressure exponent N: " << p_.n_tensile << std::endl
       << "\tVorticity confinement coefficient epsilon: " << p_.vort_eps << std::endl
       << "\tViscosity coefficient C: " << p_.xsph_c << std::endl
       << "\tSolver: MI355X HIP (libpbf_hip)" << std::endl;
    return ss.str();
  }

  // -- state for CPU consumers ------------------------------------------------
  size_t size() const { return x_.size() / 3; }
  Vec3 getPosition(size_t i) {
    sync();
    return Vec3(x_[3 * i], x_[3 * i + 1], x_[3 * i + 2]);
  }
  Vec3 getVelocity(size_t i) {
    sync();
    return Vec3(v_[3 * i], v_[3 * i + 1], v_[3 * i + 2]);
  }
  double getDensity(size_t i) {
    sync();
    return rho_[i];
  }
  const std::vector<float>& positions() {
    sync();
    return x_;
  }
  const std::vector<float>& velocities() {
    sync();
    return v_;
  }
  const pbf_params& params() const { return p_; }
  pbf_ctx* context() { return ctx_; }

  // -- members the reference's callers touch ----------------------------
  double simulate_time = 0.0;
  const double rest_density;
  void* bvh = nullptr;  // collision is the analytic box of the parameters
  bool surfaceUpToTimestep = false;

 private:
  void check(int rc, const pbf_ctx* c) {
    if (rc < 0) throw Error(rc, pbf_last_error(c));
  }
  void upload() {
    if (!dirty_) return;
    if (size() == 0) throw Error(PBF_ERR_STATE, "no particles added");
    check(pbf_load(ctx_, x_.data(), v_.data(), (int64_t)size()), ctx_);
    dirty_ = false;
    stale_ = false;
  }
  void sync() {
    if (!stale_) return;
    check(pbf_read(ctx_, x_.data(), v_.data(), rho_.data()), ctx_);
    stale_ = false;
  }

  pbf_params p_{};
  pbf_ctx* ctx_ = nullptr;
  std::vector<float> x_, v_, rho_;
  bool dirty_ = false;
  bool stale_ = false;
};

}  // namespace pbf
