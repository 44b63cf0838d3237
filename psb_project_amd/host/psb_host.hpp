// psb_host.hpp -- C++ host API over the C-ABI (include/psb_sph.h) that keeps
// the reference simulator's class interface for the SPH / PCISPH path:
//
//   SimulationData_SPHSystem / _PCI_SPHSystem  materials, with the
//        Fluid_Material XML load / save (SimulationData_SPHSystem.cpp:115-160)
//   SPHParticle / PCI_SPHParticle             particles (SphParticle.h:8-67)
//   BoxCollision, ForceExt_Constante          collider and constant force
//   EmitterBox                                lattice emitter (EmitterBox.cpp:29-47)
//   System -> SPHSystem -> PCI_SPHSystem       init / update / collide /
//        emitParticles / addForce / addObjectCollision / setDt (System.h:16-89)
//   ParticleExporter_XML / ParticleLoader_XML  particle files
//        (ParticleExporter_XML.cpp:131-190, ParticleLoader_XML.cpp:121-180)
//   ParticleExporter_Txt / ParticleLoader_Txt  "SPH n" text files
//        (ParticleExporter_Txt.cpp:26-34)
//   SPHSystem::saveCheckpoint / loadCheckpoint  bit-exact resume (C-ABI)
//   SurfaceSPH                                 iso-surface mesh (SurfaceSPH.cpp:29-190)
//
// Differences that are deliberate: no GL (System() needs none), errors throw
// std::runtime_error with psb_last_error() instead of being ignored, the host
// particle mirror is refreshed on demand (getParticles / syncParticles)
// instead of after every step, and the solver mode is selectable
// (PSB_MODE_REFERENCE, the default, reproduces the reference's semantics).
// Header only; link libpsb_sph.so.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "psb_sph.h"

namespace psbhost {

inline void check(psb_status s) {
  if (s != PSB_OK) throw std::runtime_error(psb_last_error());
}

struct Vector3 {
  float x = 0.f, y = 0.f, z = 0.f;
  Vector3() = default;
  Vector3(float a, float b, float c) : x(a), y(b), z(c) {}
  Vector3 operator+(const Vector3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  Vector3 operator-(const Vector3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  Vector3 operator*(float s) const { return {x * s, y * s, z * s}; }
};

// ---------------------------------------------------------------------------
// Materials
// ---------------------------------------------------------------------------
// Attributes of the first element named `tag` in a small XML file (the
// Fluid_Material format: one element, its settings as attributes).
inline bool read_xml_attributes(const char* filename, const std::string& tag,
                                std::map<std::string, std::string>& out) {
  std::ifstream f(filename);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  const size_t b = s.find("<" + tag);
  if (b == std::string::npos) return false;
  const size_t e = s.find('>', b);
  if (e == std::string::npos) return false;
  size_t i = b + 1 + tag.size();
  while (i < e) {
    while (i < e && std::isspace((unsigned char)s[i])) ++i;
    const size_t eq = s.find('=', i);
    if (eq == std::string::npos || eq > e) break;
    std::string key = s.substr(i, eq - i);
    while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
    const size_t q0 = s.find_first_of("\"'", eq);
    if (q0 == std::string::npos || q0 > e) break;
    const size_t q1 = s.find(s[q0], q0 + 1);
    if (q1 == std::string::npos) break;
    out[key] = s.substr(q0 + 1, q1 - q0 - 1);
    i = q1 + 1;
  }
  return true;
}

// SimulationData_SPHSystem (SimulationData_SPHSystem.h:10-62): water by default
// (SimulationData_SPHSystem.cpp:5-16).
class SimulationData_SPHSystem {
 public:
  SimulationData_SPHSystem() = default;
  // derived quantities as SimulationData_SPHSystem.cpp:19-31
  SimulationData_SPHSystem(float particleRadius, float mass, Vector3 color, float restDensity,
                           float viscosity, float surfaceTension, float gasStiffness,
                           float kernelParticles)
      : particleMass(mass), color(color), restDensity(restDensity), viscosity(viscosity),
        surfaceTension(surfaceTension), gasStiffness(gasStiffness),
        kernelParticles(kernelParticles) {
    (void)particleRadius;
    threshold = std::sqrt(restDensity / kernelParticles);
    supportRadius = std::pow((3 * mass * kernelParticles) / (4 * (float)M_PI * restDensity), 0.333f);
    this->particleRadius = std::pow(3 * mass / (4 * (float)M_PI * restDensity), 0.333f);
  }
  virtual ~SimulationData_SPHSystem() = default;

  bool loadConfiguration(const char* filename) {
    std::map<std::string, std::string> a;
    if (!read_xml_attributes(filename, "Fluid_Material", a)) return false;
    auto get = [&](const char* k, float& v) {
      auto it = a.find(k);
      if (it != a.end()) v = std::stof(it->second);
    };
    get("visualizationRadius", particleRadius);
    get("particle_mass", particleMass);
    get("restDensity", restDensity);
    get("viscosity", viscosity);
    get("surfaceTension", surfaceTension);
    get("threshold", threshold);
    get("gasStiffness", gasStiffness);
    get("kernelParticles", kernelParticles);
    get("supportRadius", supportRadius);
    return true;
  }
  bool saveConfiguration(const char* filename) const {
    std::ofstream f(filename);
    if (!f) return false;
    f.precision(9);
    f << "<?xml version=\"1.0\" ?>\n<Fluid_Material name=\"Fluid_Material\""
      << " visualizationRadius=\"" << particleRadius << "\" particle_mass=\"" << particleMass
      << "\" restDensity=\"" << restDensity << "\" viscosity=\"" << viscosity
      << "\" surfaceTension=\"" << surfaceTension << "\" threshold=\"" << threshold
      << "\" gasStiffness=\"" << gasStiffness << "\" kernelParticles=\"" << kernelParticles
      << "\" supportRadius=\"" << supportRadius << "\" />\n";
    return (bool)f;
  }

  float getParticleRadius() const { return particleRadius; }
  float getParticleMass() const { return particleMass; }
  float getRestDensity() const { return restDensity; }
  float getViscosity() const { return viscosity; }
  float getSurfaceTension() const { return surfaceTension; }
  float getThreshold() const { return threshold; }
  float getGasStiffness() const { return gasStiffness; }
  float getKernelParticles() const { return kernelParticles; }
  float getSupportRadius() const { return supportRadius; }
  void setParticleRadius(float v) { particleRadius = v; }
  void setParticleMass(float v) { particleMass = v; }
  void setRestDensity(float v) { restDensity = v; }
  void setViscosity(float v) { viscosity = v; }
  void setSurfaceTension(float v) { surfaceTension = v; }
  void setThreshold(float v) { threshold = v; }
  void setGasStiffness(float v) { gasStiffness = v; }
  void setKernelParticles(float v) { kernelParticles = v; }
  void setSupportRadius(float v) { supportRadius = v; }

 protected:
  float particleRadius = 0.0457f / 2, particleMass = 0.02f;
  Vector3 color{0, 0, 1};
  float restDensity = 998.29f, viscosity = 3.5f, surfaceTension = 0.0728f, threshold = 7.065f;
  float gasStiffness = 3.f, kernelParticles = 20.f, supportRadius = 0.0457f;
};

// SimulationData_PCI_SPHSystem (SimulationData_PCI_SPHSystem.h:7-23): same data
class SimulationData_PCI_SPHSystem : public SimulationData_SPHSystem {
 public:
  using SimulationData_SPHSystem::SimulationData_SPHSystem;
};

// ---------------------------------------------------------------------------
// Particles, collider, force, emitter
// ---------------------------------------------------------------------------
class SPHParticle {
 public:
  SPHParticle(Vector3 pos, Vector3 vel, const SimulationData_SPHSystem* data)
      : pos(pos), vel(vel), data(data) {}
  virtual ~SPHParticle() = default;
  Vector3 getPos() const { return pos; }
  Vector3 getVel() const { return vel; }
  Vector3 getVelInterAv() const { return velInterAv; }
  Vector3 getVelInterAp() const { return velInterAp; }
  float getDensity() const { return density; }
  float getPressure() const { return pressure; }
  float getMass() const { return data->getParticleMass(); }
  float getInteractionRadius() const { return data->getSupportRadius(); }
  float getRestDensity() const { return data->getRestDensity(); }
  const SimulationData_SPHSystem* getData() const { return data; }
  void setPos(Vector3 v) { pos = v; }
  void setVel(Vector3 v) { vel = v; }
  void setVelInterAv(Vector3 v) { velInterAv = v; }
  void setVelInterAp(Vector3 v) { velInterAp = v; }
  void setDensity(float v) { density = v; }
  void setPressure(float v) { pressure = v; }

 private:
  Vector3 pos, vel, velInterAv, velInterAp;
  float density = 0.f, pressure = 0.f;
  const SimulationData_SPHSystem* data;
};

class PCI_SPHParticle : public SPHParticle {
 public:
  using SPHParticle::SPHParticle;
};

// BoxCollision::create(origin, length, width, depth, restitution, container)
// (BoxCollision.cpp:22-30): the box centred at origin.
class BoxCollision {
 public:
  BoxCollision(Vector3 origin, float length, float width, float depth, float elast,
               bool container = true)
      : min(origin - Vector3(length, width, depth) * 0.5f),
        max(origin + Vector3(length, width, depth) * 0.5f), elast(elast), container(container) {}
  static BoxCollision* fromBounds(Vector3 lo, Vector3 hi, float elast) {
    return new BoxCollision((lo + hi) * 0.5f, hi.x - lo.x, hi.y - lo.y, hi.z - lo.z, elast);
  }
  Vector3 getMin() const { return min; }
  Vector3 getMax() const { return max; }
  float getElast() const { return elast; }
  bool getContainer() const { return container; }

 private:
  Vector3 min, max;
  float elast;
  bool container;
};

// SphereCollision(origin, radius, elast, friction, is_container)
// (SphereCollision.cpp:20-22): a static sphere collider; friction is kept
// and unused, as in the reference.
class SphereCollision {
 public:
  SphereCollision(Vector3 origin, float radius, float elast, float friction = 0.f,
                  bool is_container = true)
      : origin(origin), radius(radius), elast(elast), friction(friction), container(is_container) {}
  psb_collider desc() const {
    psb_collider k{};
    k.type = PSB_COLLIDER_SPHERE;
    k.container = container ? 1 : 0;
    k.elast = elast;
    k.radius = radius;
    k.origin[0] = origin.x; k.origin[1] = origin.y; k.origin[2] = origin.z;
    return k;
  }
  Vector3 getCenter() const { return origin; }
  float getRadius() const { return radius; }

 private:
  Vector3 origin;
  float radius, elast, friction;
  bool container;
};

// PlaneCollision::create(origin, length, width, elast, friction, container,
// orientation) (PlaneCollision.cpp:29-36 -> Quadrilateral::create,
// Quadrilateral.cpp:37-77): an axis-aligned length x width rectangle
// centred at origin, corners in the reference's order, normal along the
// orientation's axis.
class PlaneCollision {
 public:
  void create(Vector3 origin, float length, float width, float elast, float friction,
              bool is_container, Vector3 orientation) {
    (void)friction; (void)is_container;
    this->elast = elast;
    const float L = length / 2, W = width / 2;
    Vector3 v[4];
    if (orientation.x == -1 || orientation.x == 1) {
      v[0] = Vector3(origin.x, origin.y - L, origin.z - W);
      v[1] = Vector3(origin.x, origin.y + L, origin.z - W);
      v[2] = Vector3(origin.x, origin.y + L, origin.z + W);
      v[3] = Vector3(origin.x, origin.y - L, origin.z + W);
      normal = Vector3(orientation.x, 0, 0);
    } else if (orientation.y == -1 || orientation.y == 1) {
      v[0] = Vector3(origin.x - L, origin.y, origin.z - W);
      v[1] = Vector3(origin.x + L, origin.y, origin.z - W);
      v[2] = Vector3(origin.x + L, origin.y, origin.z + W);
      v[3] = Vector3(origin.x - L, origin.y, origin.z + W);
      normal = Vector3(0, orientation.y, 0);
    } else if (orientation.z == -1 || orientation.z == 1) {
      v[0] = Vector3(origin.x - L, origin.y - W, origin.z);
      v[1] = Vector3(origin.x + L, origin.y - W, origin.z);
      v[2] = Vector3(origin.x + L, origin.y + W, origin.z);
      v[3] = Vector3(origin.x - L, origin.y + W, origin.z);
      normal = Vector3(0, 0, orientation.z);
    } else {
      throw std::runtime_error("PlaneCollision::create: orientation must be an axis direction");
    }
    for (int i = 0; i < 4; ++i) vertex[i] = v[i];
  }
  psb_collider desc() const {
    psb_collider k{};
    k.type = PSB_COLLIDER_PLANE;
    k.elast = elast;
    for (int i = 0; i < 4; ++i) {
      k.corners[i][0] = vertex[i].x; k.corners[i][1] = vertex[i].y; k.corners[i][2] = vertex[i].z;
    }
    k.normal[0] = normal.x; k.normal[1] = normal.y; k.normal[2] = normal.z;
    return k;
  }
  Vector3 getVertex(int i) const { return vertex[i]; }
  Vector3 getNormale() const { return normal; }

 private:
  Vector3 vertex[4], normal;
  float elast = 0.f;
};

// ForceExt_Constante(direction, amplitude) (ForceExt_Constante.cu:15-24)
class ForceExt_Constante {
 public:
  ForceExt_Constante(Vector3 direction, float amplitude) : direction(direction), amplitude(amplitude) {}
  Vector3 getDirection() const { return direction; }
  float getAmplitude() const { return amplitude; }

 private:
  Vector3 direction;
  float amplitude;
};

// EmitterBox::emitParticles (EmitterBox.cpp:29-47): a lattice of spacing
// 2 particleRadius from center - size/2 while <= center + size/2, z outer,
// x inner, the same float stepping as the reference.
class EmitterBox {
 public:
  EmitterBox(Vector3 center, float sizeX, float sizeY, float sizeZ,
             const SimulationData_SPHSystem* data, Vector3 worldPosition = {})
      : center(center), sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ), data(data),
        worldPosition(worldPosition) {}
  std::vector<SPHParticle*> emitParticles() const {
    std::vector<SPHParticle*> out;
    const float step = data->getParticleRadius() * 2;
    for (float z = center.z - sizeZ / 2; z <= center.z + sizeZ / 2; z += step)
      for (float y = center.y - sizeY / 2; y <= center.y + sizeY / 2; y += step)
        for (float x = center.x - sizeX / 2; x <= center.x + sizeX / 2; x += step)
          out.push_back(new SPHParticle(Vector3(x, y, z) + worldPosition, Vector3(), data));
    return out;
  }

 private:
  Vector3 center;
  float sizeX, sizeY, sizeZ;
  const SimulationData_SPHSystem* data;
  Vector3 worldPosition;
};

// The reference packs every particle's own material into per-particle
// arrays (SphSystem.cpp:157-167, PciSphSystem.cpp:103-113, Emitter.cpp:141-175)
// and mixes them per pair (PciSphSystem_Kernel.cu:51,79,101,138,144,
// SphKernel.cu:17-71).  collectMaterials groups the particles by material:
// particles sharing one SimulationData object, or objects with equal
// values, form one material.  One material runs the system's fast path;
// several (up to PSB_MAX_MATERIALS, reference mode) go to
// psb_system_set_materials with one index per particle.
struct MaterialSet {
  std::vector<const SimulationData_SPHSystem*> mats;   // distinct, first-use order
  std::vector<uint8_t> index;                          // per particle
};
inline bool sameMaterial(const SimulationData_SPHSystem* a, const SimulationData_SPHSystem* b) {
  return a == b ||
         (a->getParticleMass() == b->getParticleMass() &&
          a->getSupportRadius() == b->getSupportRadius() &&
          a->getRestDensity() == b->getRestDensity() &&
          a->getGasStiffness() == b->getGasStiffness() &&
          a->getViscosity() == b->getViscosity() &&
          a->getSurfaceTension() == b->getSurfaceTension() &&
          a->getThreshold() == b->getThreshold());
}
inline MaterialSet collectMaterials(const std::vector<SPHParticle*>& ps, const char* who) {
  MaterialSet m;
  m.index.resize(ps.size());
  for (size_t k = 0; k < ps.size(); ++k) {
    const SimulationData_SPHSystem* d = ps[k]->getData();
    size_t j = 0;
    while (j < m.mats.size() && !sameMaterial(m.mats[j], d)) ++j;
    if (j == m.mats.size()) {
      if (m.mats.size() == PSB_MAX_MATERIALS)
        throw std::invalid_argument(std::string(who) + ": more than " +
                                    std::to_string(PSB_MAX_MATERIALS) + " materials");
      m.mats.push_back(d);
    }
    m.index[k] = (uint8_t)j;
  }
  return m;
}
inline void setMaterial(psb_sph_params& p, const SimulationData_SPHSystem* d) {
  p.mass = d->getParticleMass();
  p.support_radius = d->getSupportRadius();
  p.rest_density = d->getRestDensity();
  p.gas_stiffness = d->getGasStiffness();
  p.viscosity = d->getViscosity();
  p.surface_tension = d->getSurfaceTension();
  p.threshold = d->getThreshold();
}

// ---------------------------------------------------------------------------
// Systems
// ---------------------------------------------------------------------------
class System {
 public:
  virtual ~System() = default;
  virtual void init(std::vector<SPHParticle*> particles) = 0;
  virtual void update() = 0;
  virtual void emitParticles() {}
  virtual void collide() {}   // fused into update()
  std::vector<SPHParticle*> getParticles() { syncParticles(); return particles; }
  void addObjectCollision(BoxCollision* b) { box.reset(b); }
  // colliders after the box, in the order added
  void addObjectCollision(SphereCollision* c) { colliders.push_back(c->desc()); delete c; }
  void addObjectCollision(PlaneCollision* c) { colliders.push_back(c->desc()); delete c; }
  void addForce(ForceExt_Constante* f) { forces.emplace_back(f); }
  int getNumParticles() const { return (int)particles.size(); }
  double getDt() const { return dt; }
  void setDt(double v) { dt = v; }
  virtual void syncParticles() {}

 protected:
  std::vector<SPHParticle*> particles;
  std::unique_ptr<BoxCollision> box;
  std::vector<psb_collider> colliders;
  std::vector<std::unique_ptr<ForceExt_Constante>> forces;
  double dt = 0.01;   // SphSystem.cpp:9
};

// SPHSystem (SphSystem.h): WCSPH on the GPU through psb_system.
class SPHSystem : public System {
 public:
  explicit SPHSystem(int device = 0) {
    check(psb_init(device, &ctx));
    check(psb_default_params(&prm));
  }
  ~SPHSystem() override {
    if (sys) psb_system_destroy(sys);
    if (ctx) psb_finalize(ctx);
    for (auto* p : particles) delete p;
  }
  SPHSystem(const SPHSystem&) = delete;
  SPHSystem& operator=(const SPHSystem&) = delete;

  // Solver settings beyond the reference's fixed ones (mode, tolerance ...):
  // edit before init().
  psb_sph_params& params() { return prm; }

  void init(std::vector<SPHParticle*> ps) override {
    particles = std::move(ps);
    if (particles.empty()) throw std::runtime_error("SPHSystem::init without particles");
    const MaterialSet ms = collectMaterials(particles, "SPHSystem::init");
    if (ms.mats.size() > 1 && prm.mode != PSB_MODE_REFERENCE)
      throw std::invalid_argument("SPHSystem::init: particles of several materials need "
                                  "PSB_MODE_REFERENCE");
    // the system carries the material with the largest support radius (it
    // sizes the neighbour search); the box clamp takes particle 0's radius
    // (SphSystem.cpp:326)
    const SimulationData_SPHSystem* d = ms.mats[0];
    for (const SimulationData_SPHSystem* e : ms.mats)
      if (e->getSupportRadius() > d->getSupportRadius()) d = e;
    prm.solver = solver();
    setMaterial(prm, d);
    prm.particle_radius = particles[0]->getData()->getParticleRadius();
    prm.dt = (float)dt;
    Vector3 g;
    for (auto& f : forces) g = g + f->getDirection() * f->getAmplitude();
    prm.gravity[0] = g.x; prm.gravity[1] = g.y; prm.gravity[2] = g.z;
    if (box) {
      prm.box_enabled = 1;
      const Vector3 lo = box->getMin(), hi = box->getMax();
      prm.box_min[0] = lo.x; prm.box_min[1] = lo.y; prm.box_min[2] = lo.z;
      prm.box_max[0] = hi.x; prm.box_max[1] = hi.y; prm.box_max[2] = hi.z;
      prm.elast = box->getElast();
      // grid over the box, cell = lattice spacing when set, else h
      const float cell = prm.lattice_spacing > 0.f ? prm.lattice_spacing : prm.support_radius;
      check(psb_create_grid_desc_box(lo, hi, cell));
    }
    const uint32_t n = (uint32_t)particles.size();
    if (sys) { psb_system_destroy(sys); sys = nullptr; }
    // capacity: the fluid plus the rigid spheres' boundary particles
    const uint32_t nsph = (uint32_t)sphere_centers.size();
    const uint32_t extra = nsph ? nsph * psb_sphere_particle_count(&prm, &sphere_prm) : 0u;
    check(psb_system_create(ctx, &prm, n + extra, &sys));
    for (const psb_collider& k : colliders) check(psb_system_add_collider(sys, &k));
    if (lanes >= 0) check(psb_system_set_corrector_lanes(sys, lanes));
    if (compact >= 0) check(psb_system_set_compact_list(sys, compact));
    if (halo >= 0) check(psb_system_set_halo(sys, halo));
    if (mlp) check(psb_system_set_mlp(sys, mlp));
    if (nsph) {
      std::vector<float> c(3 * (size_t)nsph), u(3 * (size_t)nsph, 0.f);
      for (uint32_t k = 0; k < nsph; ++k) {
        c[3 * k] = sphere_centers[k].x; c[3 * k + 1] = sphere_centers[k].y; c[3 * k + 2] = sphere_centers[k].z;
        if (k < sphere_velocities.size()) {
          u[3 * k] = sphere_velocities[k].x; u[3 * k + 1] = sphere_velocities[k].y;
          u[3 * k + 2] = sphere_velocities[k].z;
        }
      }
      check(psb_system_set_spheres(sys, nsph, c.data(), u.data(), &sphere_prm));
    }
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
      const Vector3 p = particles[i]->getPos(), u = particles[i]->getVel();
      x[3 * i] = p.x; x[3 * i + 1] = p.y; x[3 * i + 2] = p.z;
      v[3 * i] = u.x; v[3 * i + 1] = u.y; v[3 * i + 2] = u.z;
    }
    check(psb_system_set_particles(sys, n, x.data(), x.data(), v.data(), v.data(), v.data(),
                                   v.data()));
    if (ms.mats.size() > 1) {
      std::vector<psb_sph_params> table(ms.mats.size(), prm);
      for (size_t k = 0; k < ms.mats.size(); ++k) setMaterial(table[k], ms.mats[k]);
      check(psb_system_set_materials(sys, (uint32_t)table.size(), table.data(), ms.index.data(),
                                     n));
    }
    dirty = false;
  }

  void update() override {
    if (!sys) throw std::runtime_error("SPHSystem::update before init");
    check(psb_system_update(sys, 1));
    dirty = true;
  }

  // host mirror of positions, velocities, density and pressure, by particle id
  void syncParticles() override {
    if (!dirty || !sys) return;
    const uint32_t n = psb_system_count(sys);
    std::vector<float> x(4 * (size_t)n), v(4 * (size_t)n), vh(4 * (size_t)n), rho(n), p(n);
    std::vector<uint32_t> ids(n);
    check(psb_system_get_particles_f32(sys, x.data(), v.data(), vh.data(), rho.data(), p.data(),
                                       ids.data()));
    for (uint32_t k = 0; k < n; ++k) {
      if (ids[k] >= particles.size()) continue;
      SPHParticle* q = particles[ids[k]];
      q->setPos(Vector3(x[4 * k], x[4 * k + 1], x[4 * k + 2]));
      q->setVel(Vector3(v[4 * k], v[4 * k + 1], v[4 * k + 2]));
      q->setVelInterAp(Vector3(vh[4 * k], vh[4 * k + 1], vh[4 * k + 2]));
      q->setDensity(rho[k]);
      q->setPressure(p[k]);
    }
    dirty = false;
  }

  psb_step_stats stats() const {
    psb_step_stats s{};
    check(psb_system_get_stats(sys, &s));
    return s;
  }
  psb_system handle() const { return sys; }
  psb_context context() const { return ctx; }
  Vector3 gravity() const {
    Vector3 g;
    for (auto& f : forces) g = g + f->getDirection() * f->getAmplitude();
    return g;
  }

  // Binary checkpoint of the whole system state (psb_system_save_checkpoint);
  // loadCheckpoint needs init() with the same setup first and continues the
  // run bit for bit.
  void saveCheckpoint(const char* filename) const {
    if (!sys) throw std::runtime_error("saveCheckpoint before init");
    check(psb_system_save_checkpoint(sys, filename));
  }
  void loadCheckpoint(const char* filename) {
    if (!sys) throw std::runtime_error("loadCheckpoint before init");
    check(psb_system_load_checkpoint(sys, filename));
    dirty = true;
  }

  // Rigid DEM spheres coupled to the fluid through boundary particles
  // (psb_system_set_spheres; no counterpart in the reference's SPH classes).
  // Set before init(): the system's capacity then holds their boundary
  // particles too, which follow the fluid and are not in getParticles().
  void setSpheres(const std::vector<Vector3>& centers, const psb_sphere_params& sp,
                  const std::vector<Vector3>& velocities = {}) {
    sphere_centers = centers;
    sphere_velocities = velocities;
    sphere_prm = sp;
  }
  struct SphereState {
    std::vector<Vector3> centers, velocities, fluid_forces;
  };
  // centres, velocities and the last step's fluid force of every sphere
  SphereState getSpheres() const {
    SphereState o;
    if (!sys) return o;
    const uint32_t m = psb_system_num_spheres(sys);
    std::vector<float> c(3 * (size_t)m), v(3 * (size_t)m), f(3 * (size_t)m);
    if (m) check(psb_system_get_spheres(sys, c.data(), v.data(), f.data()));
    for (uint32_t k = 0; k < m; ++k) {
      o.centers.emplace_back(c[3 * k], c[3 * k + 1], c[3 * k + 2]);
      o.velocities.emplace_back(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
      o.fluid_forces.emplace_back(f[3 * k], f[3 * k + 1], f[3 * k + 2]);
    }
    return o;
  }

  // Learned warm start (params().warm_start = PSB_WARM_MLP): the network
  // that predicts each step's starting pressure; the caller keeps it alive.
  void setWarmStartModel(psb_mlp m) {
    mlp = m;
    if (sys) check(psb_system_set_mlp(sys, m));
  }

  // Speed knobs of the corrector passes (results do not depend on them):
  // threads per particle (0 = by size, 1, 2), the 16-bit neighbour list,
  // neighbour rows staged in LDS.  Applied at init() when set before it.
  void setCorrectorLanes(int l) { lanes = l; if (sys) check(psb_system_set_corrector_lanes(sys, l)); }
  void setCompactList(bool on) { compact = on; if (sys) check(psb_system_set_compact_list(sys, on)); }
  void setHalo(bool on) { halo = on; if (sys) check(psb_system_set_halo(sys, on)); }

  // State in a given storage order (ids: particle index of each slot), e.g.
  // from ParticleLoader_XML::restore; marks the host mirror stale.
  void setState(const std::vector<float>& pos4, const std::vector<float>& vel4,
                const std::vector<float>& vh4, const std::vector<float>& pressure,
                const std::vector<uint32_t>& ids) {
    if (!sys) throw std::runtime_error("setState before init");
    const uint32_t n = (uint32_t)ids.size();
    if (n != particles.size()) throw std::runtime_error("setState: particle count differs from init()");
    check(psb_system_set_state(sys, n, pos4.data(), vel4.data(), vh4.data(), pressure.data(),
                               ids.data()));
    dirty = true;
  }

 protected:
  virtual int32_t solver() const { return PSB_SOLVER_WCSPH; }
  psb_status psb_create_grid_desc_box(Vector3 lo, Vector3 hi, float cell) {
    prm.grid.min[0] = lo.x; prm.grid.min[1] = lo.y; prm.grid.min[2] = lo.z;
    prm.grid.cell_size = cell;
    prm.grid.dims[0] = (int32_t)std::ceil((hi.x - lo.x) / cell - 1e-6f);
    prm.grid.dims[1] = (int32_t)std::ceil((hi.y - lo.y) / cell - 1e-6f);
    prm.grid.dims[2] = (int32_t)std::ceil((hi.z - lo.z) / cell - 1e-6f);
    return PSB_OK;
  }

  psb_context ctx = nullptr;
  psb_system sys = nullptr;
  psb_sph_params prm{};
  bool dirty = false;
  std::vector<Vector3> sphere_centers, sphere_velocities;
  psb_sphere_params sphere_prm{};
  psb_mlp mlp = nullptr;
  int lanes = -1, compact = -1, halo = -1;   // -1: library default
};

// PCI_SPHSystem (PciSphSystem.h): PCISPH on the GPU through psb_system.
class PCI_SPHSystem : public SPHSystem {
 public:
  using SPHSystem::SPHSystem;

 protected:
  int32_t solver() const override { return PSB_SOLVER_PCISPH; }
};

// SceneBatch: many independent PCISPH scenes of one material advanced by one
// launch per phase (psb_batch_*; no counterpart in the reference, which runs
// one System per process).  Each scene stops at its own tolerance and equals
// the same scene run alone, bit for bit.
class SceneBatch {
 public:
  SceneBatch(psb_context ctx, const psb_sph_params& material,
             const std::vector<psb_scene_desc>& scenes) {
    check(psb_batch_create(ctx, &material, (uint32_t)scenes.size(), scenes.data(), &b));
  }
  ~SceneBatch() { if (b) psb_batch_destroy(b); }
  SceneBatch(const SceneBatch&) = delete;
  SceneBatch& operator=(const SceneBatch&) = delete;

  uint32_t numScenes() const { return psb_batch_num_scenes(b); }
  // FP32 float4 arrays of one scene (vel4 / vh4 may be null: zero)
  void setParticles(uint32_t scene, const float* pos4, const float* vel4 = nullptr,
                    const float* vh4 = nullptr) {
    check(psb_batch_set_particles(b, scene, pos4, vel4, vh4));
  }
  void update(int32_t steps = 1) { check(psb_batch_update(b, steps)); }
  psb_step_stats stats(uint32_t scene) const {
    psb_step_stats s{};
    check(psb_batch_get_stats(b, scene, &s));
    return s;
  }
  // one scene's state in its sorted order, with the particle ids
  void getParticles(uint32_t scene, uint32_t count, std::vector<float>& pos4,
                    std::vector<float>& vel4, std::vector<uint32_t>& ids) const {
    pos4.assign(4 * (size_t)count, 0.f);
    vel4.assign(4 * (size_t)count, 0.f);
    ids.assign(count, 0u);
    check(psb_batch_get_particles(b, scene, pos4.data(), vel4.data(), nullptr, nullptr, nullptr,
                                  ids.data()));
  }
  void setWarmStartModel(psb_mlp m) { check(psb_batch_set_mlp(b, m)); }
  // launch shapes per phase (PSB_BATCH_PHASES), as psb_batch_autotune returns them
  std::vector<int32_t> autotune(int32_t steps = 2) {
    std::vector<int32_t> bs(PSB_BATCH_PHASES, 0);
    check(psb_batch_autotune(b, steps, bs.data()));
    return bs;
  }
  void setBlockSizes(const std::vector<int32_t>& bs) { check(psb_batch_set_block_sizes(b, bs.data())); }
  void setGraphs(bool on) { check(psb_batch_set_graphs(b, on)); }
  void setLatency(bool on) { check(psb_batch_set_latency(b, on)); }
  struct Latency {
    double p50_ms = 0, p99_ms = 0, max_ms = 0;
    uint64_t samples = 0;
  };
  Latency latency() const {
    Latency l;
    check(psb_batch_get_latency(b, &l.p50_ms, &l.p99_ms, &l.max_ms, &l.samples));
    return l;
  }
  psb_batch handle() const { return b; }

 private:
  psb_batch b = nullptr;
};

// ---------------------------------------------------------------------------
// Surface
// ---------------------------------------------------------------------------
// SurfaceSPH (SurfaceSPH.h, SurfaceSPH.cpp:29-190): extract(S, isoLevel, step)
// fills the indexed mesh -- unique vertices, their normals, and the triangle
// list (three indices per triangle) -- from the system's fluid particles.
class SurfaceSPH {
 public:
  SurfaceSPH() { check(psb_surface_default_params(&prm)); }
  ~SurfaceSPH() {
    if (surf) psb_surface_destroy(surf);
  }
  SurfaceSPH(const SurfaceSPH&) = delete;
  SurfaceSPH& operator=(const SurfaceSPH&) = delete;

  // the node spacing, kernel scale and margin (SurfaceSPH.cpp:31-32 by default)
  psb_surface_params& params() { return prm; }

  bool extract(SPHSystem* S, double isoLevel, unsigned step = 0) {
    (void)step;
    if (!S || !S->handle()) return false;
    if (!surf) check(psb_surface_create(S->context(), &surf));
    prm.iso_level = (float)isoLevel;
    uint32_t nv = 0, nt = 0;
    check(psb_surface_extract(surf, S->handle(), &prm, &nv, &nt));
    std::vector<float> v(3 * (size_t)nv), n(3 * (size_t)nv);
    std::vector<uint32_t> t(3 * (size_t)nt);
    check(psb_surface_get_mesh(surf, v.data(), n.data(), t.data()));
    vertexs_cpu.resize(nv);
    normales_cpu.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) {
      vertexs_cpu[i] = Vector3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
      normales_cpu[i] = Vector3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    }
    indexs_cpu.assign(t.begin(), t.end());
    return true;
  }

  const std::vector<Vector3>& getVertexs() const { return vertexs_cpu; }
  const std::vector<Vector3>& getNormales() const { return normales_cpu; }
  const std::vector<uint32_t>& getIndexs() const { return indexs_cpu; }

 private:
  psb_surface surf = nullptr;
  psb_surface_params prm{};
  std::vector<Vector3> vertexs_cpu, normales_cpu;
  std::vector<uint32_t> indexs_cpu;
};

// ---------------------------------------------------------------------------
// Particle files
// ---------------------------------------------------------------------------
// Storage-order snapshot of a system: float4 arrays, scalars, ids.
struct SystemSnapshot {
  uint32_t n = 0;
  std::vector<float> pos4, vel4, vh4, vav4, density, pressure;
  std::vector<uint32_t> ids;
};

inline SystemSnapshot snapshot(const SPHSystem& S) {
  SystemSnapshot o;
  psb_system sys = S.handle();
  if (!sys) throw std::runtime_error("snapshot before init");
  o.n = psb_system_count(sys);
  const size_t n = o.n;
  o.pos4.resize(4 * n); o.vel4.resize(4 * n); o.vh4.resize(4 * n); o.vav4.resize(4 * n);
  o.density.resize(n); o.pressure.resize(n); o.ids.resize(n);
  check(psb_system_get_particles_f32(sys, o.pos4.data(), o.vel4.data(), o.vh4.data(),
                                     o.density.data(), o.pressure.data(), o.ids.data()));
  psb_system_arrays a{};
  check(psb_system_get_arrays(sys, &a));
  if (n) check(psb_copy_array_from_device(S.context(), o.vav4.data(), a.vel_inter_av4, 0, 16 * n));
  return o;
}

// ParticleExporter_XML (ParticleExporter_XML.cpp:15-37, 131-190):
//   <Particles name="SPH System" dt count> <Gravity X Y Z Amp/>
//   <Particle index mass radius kernelParticles interactionRadius restDensity
//    gasStiffness threshold surfaceTension viscosity>
//     <position/> <velocity/> <velocityInterAv/> <velocityInterAp/> <color/>
// Particles are written in the system's storage order with index = particle
// id.  Differences: numbers carry 9 significant digits (an FP32 round trip;
// TinyXML writes %f) and each Particle also carries its pressure, so that
// ParticleLoader_XML::restore continues the run bit for bit.  Files it
// writes load with the reference's loader.
class ParticleExporter_XML {
 public:
  bool _export(const char* filename, SPHSystem* S) const {
    const SystemSnapshot o = snapshot(*S);
    FILE* f = std::fopen(filename, "w");
    if (!f) return false;
    const std::vector<SPHParticle*>& ps = S->getParticles();
    if (ps.size() < o.n) { std::fclose(f); throw std::runtime_error("export: system holds more particles than init()"); }
    const Vector3 g = S->gravity();
    const float amp = std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
    const Vector3 dir = amp > 0.f ? g * (1.0f / amp) : Vector3();
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n<Particles name=\"SPH System\" dt=\"%.9g\" count=\"%u\">\n",
                 S->getDt(), o.n);
    std::fprintf(f, "    <Gravity X=\"%.9g\" Y=\"%.9g\" Z=\"%.9g\" Amp=\"%.9g\" />\n", dir.x, dir.y,
                 dir.z, amp);
    auto vec = [&](const char* tag, const float* v) {
      std::fprintf(f, "        <%s X=\"%.9g\" Y=\"%.9g\" Z=\"%.9g\" />\n", tag, v[0], v[1], v[2]);
    };
    for (uint32_t k = 0; k < o.n; ++k) {
      if (o.ids[k] >= ps.size()) { std::fclose(f); throw std::runtime_error("export: particle id beyond init()"); }
      const SimulationData_SPHSystem* d = ps[o.ids[k]]->getData();
      std::fprintf(f,
                   "    <Particle index=\"%u\" mass=\"%.9g\" radius=\"%.9g\" kernelParticles=\"%.9g\" "
                   "interactionRadius=\"%.9g\" restDensity=\"%.9g\" gasStiffness=\"%.9g\" "
                   "threshold=\"%.9g\" surfaceTension=\"%.9g\" viscosity=\"%.9g\" "
                   "pressure=\"%.9g\" density=\"%.9g\">\n",
                   o.ids[k], d->getParticleMass(), d->getParticleRadius(), d->getKernelParticles(),
                   d->getSupportRadius(), d->getRestDensity(), d->getGasStiffness(),
                   d->getThreshold(), d->getSurfaceTension(), d->getViscosity(), o.pressure[k],
                   o.density[k]);
      vec("position", &o.pos4[4 * k]);
      vec("velocity", &o.vel4[4 * k]);
      vec("velocityInterAv", &o.vav4[4 * k]);
      vec("velocityInterAp", &o.vh4[4 * k]);
      const float col[3] = {0.f, 0.f, 1.f};
      vec("color", col);
      std::fprintf(f, "    </Particle>\n");
    }
    std::fprintf(f, "</Particles>\n");
    return std::fclose(f) == 0;
  }
};

// Streaming reader of the Particles / Particle / child-element structure.
namespace detail {
struct XmlElem {
  std::string name;
  std::map<std::string, std::string> attr;
  bool closing = false, self_closing = false;
};
// Next element tag at or after pos (skips text, comments, declarations).
inline bool next_tag(const std::string& s, size_t& pos, XmlElem& e) {
  while (true) {
    const size_t lt = s.find('<', pos);
    if (lt == std::string::npos) return false;
    if (s.compare(lt, 4, "<!--") == 0) {
      const size_t end = s.find("-->", lt);
      if (end == std::string::npos) return false;
      pos = end + 3;
      continue;
    }
    if (s.compare(lt, 2, "<?") == 0 || s.compare(lt, 2, "<!") == 0) {
      const size_t end = s.find('>', lt);
      if (end == std::string::npos) return false;
      pos = end + 1;
      continue;
    }
    const size_t gt = s.find('>', lt);
    if (gt == std::string::npos) return false;
    e = XmlElem{};
    size_t i = lt + 1;
    if (s[i] == '/') { e.closing = true; ++i; }
    size_t j = i;
    while (j < gt && !std::isspace((unsigned char)s[j]) && s[j] != '/') ++j;
    e.name = s.substr(i, j - i);
    e.self_closing = s[gt - 1] == '/';
    i = j;
    while (i < gt) {
      while (i < gt && (std::isspace((unsigned char)s[i]) || s[i] == '/')) ++i;
      const size_t eq = s.find('=', i);
      if (eq == std::string::npos || eq > gt) break;
      std::string key = s.substr(i, eq - i);
      while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
      const size_t q0 = s.find_first_of("\"'", eq);
      if (q0 == std::string::npos || q0 > gt) break;
      const size_t q1 = s.find(s[q0], q0 + 1);
      if (q1 == std::string::npos) break;
      e.attr[key] = s.substr(q0 + 1, q1 - q0 - 1);
      i = q1 + 1;
    }
    pos = gt + 1;
    return true;
  }
}
inline double attr_d(const XmlElem& e, const char* k, double dflt = 0.0) {
  auto it = e.attr.find(k);
  return it == e.attr.end() ? dflt : std::strtod(it->second.c_str(), nullptr);
}
inline bool read_file(const char* filename, std::string& out) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}
}  // namespace detail

// ParticleLoader_XML (ParticleLoader_XML.cpp:18-37, 121-180).
//   load:    SPHParticles in file order, each with a material made of its
//            per-particle attributes (the loader owns the materials);
//   restore: the state of a file written by ParticleExporter_XML into a
//            system init()ed with the same particles: storage order = file
//            order, ids = index, with pressure -- the run continues bit for
//            bit.  Spheres are not part of the file (use the checkpoint).
class ParticleLoader_XML {
 public:
  struct Record {
    uint32_t index = 0;
    Vector3 pos, vel, vav, vap, color;
    float pressure = 0.f, density = 0.f;
    bool has_pressure = false;
    float mass = 0, radius = 0, kernelParticles = 0, interactionRadius = 0, restDensity = 0,
          gasStiffness = 0, threshold = 0, surfaceTension = 0, viscosity = 0;
  };

  // every Particle of an "SPH System" file; false when the file or its
  // root is not one
  bool read(const char* filename, std::vector<Record>& out, double* dt = nullptr) const {
    std::string s;
    if (!detail::read_file(filename, s)) return false;
    size_t pos = 0;
    detail::XmlElem e;
    bool root = false;
    while (detail::next_tag(s, pos, e)) {
      if (e.closing) continue;
      if (e.name == "Particles") {
        auto it = e.attr.find("name");
        if (it == e.attr.end() || it->second != "SPH System") return false;
        if (dt) *dt = detail::attr_d(e, "dt", 0.0);
        root = true;
        continue;
      }
      if (!root || e.name != "Particle") continue;
      Record r;
      r.index = (uint32_t)detail::attr_d(e, "index", (double)out.size());
      r.mass = (float)detail::attr_d(e, "mass");
      r.radius = (float)detail::attr_d(e, "radius");
      r.kernelParticles = (float)detail::attr_d(e, "kernelParticles");
      r.interactionRadius = (float)detail::attr_d(e, "interactionRadius");
      r.restDensity = (float)detail::attr_d(e, "restDensity");
      r.gasStiffness = (float)detail::attr_d(e, "gasStiffness");
      r.threshold = (float)detail::attr_d(e, "threshold");
      r.surfaceTension = (float)detail::attr_d(e, "surfaceTension");
      r.viscosity = (float)detail::attr_d(e, "viscosity");
      r.has_pressure = e.attr.count("pressure") != 0;
      r.pressure = (float)detail::attr_d(e, "pressure");
      r.density = (float)detail::attr_d(e, "density");
      if (!e.self_closing) {
        detail::XmlElem c;
        while (detail::next_tag(s, pos, c)) {
          if (c.closing && c.name == "Particle") break;
          const Vector3 v((float)detail::attr_d(c, "X"), (float)detail::attr_d(c, "Y"),
                          (float)detail::attr_d(c, "Z"));
          if (c.name == "position") r.pos = v;
          else if (c.name == "velocity") r.vel = v;
          else if (c.name == "velocityInterAv") r.vav = v;
          else if (c.name == "velocityInterAp") r.vap = v;
          else if (c.name == "color") r.color = v;
        }
      }
      out.push_back(r);
    }
    return root;
  }

  std::vector<SPHParticle*> load(const char* filename) {
    std::vector<Record> rs;
    std::vector<SPHParticle*> out;
    if (!read(filename, rs)) return out;
    for (const Record& r : rs) {
      auto m = std::make_unique<SimulationData_SPHSystem>();
      m->setParticleMass(r.mass);
      m->setParticleRadius(r.radius);
      m->setKernelParticles(r.kernelParticles);
      m->setSupportRadius(r.interactionRadius);
      m->setRestDensity(r.restDensity);
      m->setGasStiffness(r.gasStiffness);
      m->setThreshold(r.threshold);
      m->setSurfaceTension(r.surfaceTension);
      m->setViscosity(r.viscosity);
      SPHParticle* p = new SPHParticle(r.pos, r.vel, m.get());
      p->setVelInterAv(r.vav);
      p->setVelInterAp(r.vap);
      p->setPressure(r.pressure);
      p->setDensity(r.density);
      materials.push_back(std::move(m));
      out.push_back(p);
    }
    return out;
  }

  bool restore(SPHSystem* S, const char* filename) const {
    std::vector<Record> rs;
    if (!read(filename, rs)) return false;
    const size_t n = rs.size();
    std::vector<float> x(4 * n, 0.f), v(4 * n, 0.f), vh(4 * n, 0.f), p(n, 0.f);
    std::vector<uint32_t> ids(n);
    for (size_t k = 0; k < n; ++k) {
      const Record& r = rs[k];
      x[4 * k] = r.pos.x; x[4 * k + 1] = r.pos.y; x[4 * k + 2] = r.pos.z;
      v[4 * k] = r.vel.x; v[4 * k + 1] = r.vel.y; v[4 * k + 2] = r.vel.z;
      vh[4 * k] = r.vap.x; vh[4 * k + 1] = r.vap.y; vh[4 * k + 2] = r.vap.z;
      p[k] = r.pressure;
      ids[k] = r.index;
    }
    S->setState(x, v, vh, p, ids);
    return true;
  }

 private:
  std::vector<std::unique_ptr<SimulationData_SPHSystem>> materials;
};

// ParticleExporter_Txt, SPH branch (ParticleExporter_Txt.cpp:26-34):
//   "SPH <count>" then "-p x y z -m mass -d density" per particle, by id
//   (the reference prints %lf; 9 significant digits here).
class ParticleExporter_Txt {
 public:
  bool _export(const char* filename, SPHSystem* S) const {
    const SystemSnapshot o = snapshot(*S);
    const std::vector<SPHParticle*>& ps = S->getParticles();
    if (ps.size() != o.n) throw std::runtime_error("export: particle count differs from init()");
    std::vector<uint32_t> slot(o.n);
    for (uint32_t k = 0; k < o.n; ++k) slot[o.ids[k]] = k;
    FILE* f = std::fopen(filename, "w");
    if (!f) return false;
    std::fprintf(f, "SPH %u\n", o.n);
    for (uint32_t i = 0; i < o.n; ++i) {
      const uint32_t k = slot[i];
      std::fprintf(f, "-p %.9g %.9g %.9g -m %.9g -d %.9g\n", o.pos4[4 * k], o.pos4[4 * k + 1],
                   o.pos4[4 * k + 2], ps[i]->getMass(), o.density[k]);
    }
    return std::fclose(f) == 0;
  }
};

// Reader of the text format above (the reference reads text files of its
// simple particle system only, ParticleLoader_Txt.cpp:17-43): positions,
// masses and densities by particle id, velocities zero.
class ParticleLoader_Txt {
 public:
  struct Record { Vector3 pos; float mass = 0.f, density = 0.f; };
  bool read(const char* filename, std::vector<Record>& out) const {
    FILE* f = std::fopen(filename, "r");
    if (!f) return false;
    unsigned count = 0;
    if (std::fscanf(f, "SPH %u", &count) != 1) { std::fclose(f); return false; }
    out.resize(count);
    for (unsigned i = 0; i < count; ++i) {
      Record& r = out[i];
      if (std::fscanf(f, " -p %f %f %f -m %f -d %f", &r.pos.x, &r.pos.y, &r.pos.z, &r.mass,
                      &r.density) != 5) {
        std::fclose(f);
        out.resize(i);
        return false;
      }
    }
    std::fclose(f);
    return true;
  }
};

}  // namespace psbhost
