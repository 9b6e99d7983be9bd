// The reference's callers of Particles, compiled against the drop-in header.
//
// The replacement src/particles.h derives
// CGL::Particles from pbf::ParticlesT<CGL::Vector3D>.  This file is that
// header plus two functions that make the calls the reference makes, with
// the reference's own CGL::Vector3D, CGL::Color and tinyxml2 headers on the
// include path:
//   * Application::load_particles (src/application.cpp:302-344): XML parse,
//     `new Particles(d)`, addParticle(pos, v) per particle,
//     estimateDensities();
//   * PathTracer::fluid_simulate_time / fluid_simulate_time_step
//     (src/pathtracer.cpp:444-480): simulate_time read, timeStep() until it
//     has advanced by delta_t; the bvh member assigned by build_accel (:266).
// bvh.h and marching.h pull in the GL headers (glew.h -> GL/glu.h, absent
// here), so the two types the replacement header names from them are only
// declared; nothing in the facade dereferences them.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "CGL/color.h"
#include "CGL/tinyxml2.h"
#include "CGL/vector3D.h"
#include "pbf_particles.hpp"

namespace CGL {
class BVHAccel;  // src/bvh.h
namespace StaticScene {
class Primitive;  // src/static_scene/primitive.h
}

// The replacement src/particles.h.
struct Particles : pbf::ParticlesT<Vector3D> {
  using Base = pbf::ParticlesT<Vector3D>;
  BVHAccel* bvh = nullptr;
  std::vector<StaticScene::Primitive*> surface;
  explicit Particles(double rest_density = 1000.0) : Base(rest_density) {}
  void redraw(const Color& c) {
    (void)c;
    for (size_t i = 0; i < size(); i++) (void)getPosition(i);
  }
};
}  // namespace CGL

// Every member of the facade, instantiated with the reference's vector type.
template class pbf::ParticlesT<CGL::Vector3D>;

using CGL::Particles;
using CGL::Vector3D;

namespace {

Vector3D stov(const char* s) {
  double x = 0, y = 0, z = 0;
  std::stringstream ss(s);
  ss >> x >> y >> z;
  return Vector3D(x, y, z);
}

}  // namespace

// The calls of Application::load_particles (src/application.cpp:302-344).
Particles* load_particles(const char* filename) {
  using namespace tinyxml2;
  XMLDocument doc;
  doc.LoadFile(filename);
  if (doc.Error()) return new Particles();
  XMLElement* root = doc.FirstChildElement("particles");
  if (!root) return nullptr;
  XMLElement* density = root->FirstChildElement("density");
  XMLElement* ps = root->FirstChildElement("ps");
  float d = std::stof(density->GetText());
  Particles* particles = new Particles(d);
  XMLElement* p = ps->FirstChildElement("particle");
  while (p) {
    Vector3D pos = stov(p->FirstChildElement("pos")->GetText());
    Vector3D v = stov(p->FirstChildElement("v")->GetText());
    particles->addParticle(pos, v);
    p = p->NextSiblingElement("particle");
  }
  particles->estimateDensities();
  return particles;
}

// The calls of PathTracer::build_accel and fluid_simulate_time
// (src/pathtracer.cpp:266, :444-480).
void fluid_simulate_time(Particles* fluid_particles, CGL::BVHAccel* bvh, double delta_t) {
  fluid_particles->bvh = bvh;
  double start_t = fluid_particles->simulate_time;
  while (fluid_particles->simulate_time < start_t + delta_t) {
    fluid_particles->timeStep();
  }
  fluid_particles->redraw(CGL::Color(1, 1, 1, 1));
  std::cout << fluid_particles->paramsString();
  Vector3D q = fluid_particles->getPosition(0);
  (void)fluid_particles->estimateDensityAt(q);
}
