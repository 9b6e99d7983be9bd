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
  void addParticle(Vec3 pos, Vec3 v) {
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

  // SPH density at an arbitrary point over the read-back positions
  // (src/particles.cpp:446-453); a CPU consumer, like marching cubes.
  double estimateDensityAt(Vec3 pos) {
    sync();
    const double h = p_.h, h2 = h * h;
    double h9 = h;
    for (int k = 1; k < 9; k++) h9 *= h;
    double density = 0.0;
    for (size_t i = 0; i < size(); i++) {
      const double dx = x_[3 * i] - pos.x, dy = x_[3 * i + 1] - pos.y, dz = x_[3 * i + 2] - pos.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < h2) {
        const double t = h2 - r2;
        density += (double)p_.poly6_coef * (t * t * t) / h9;
      }
    }
    return density;
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
