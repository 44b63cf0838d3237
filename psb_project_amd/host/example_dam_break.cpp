// example_dam_break.cpp -- the PCISPH dam break driven through the C++ host
// API (psb_host.hpp), as a PSB_Project application would drive
// PCI_SPHSystem: material from a Fluid_Material file or code, an EmitterBox
// lattice, a BoxCollision, gravity as ForceExt_Constante, init, update.
//
//   example_dam_break --material FILE --print-material      (no GPU needed)
//   example_dam_break --save-material FILE                  (no GPU needed)
//   example_dam_break --print-xml FILE | --print-txt FILE   (no GPU needed)
//       particle files read by ParticleLoader_XML::load / ParticleLoader_Txt,
//       one JSON line per particle
//   example_dam_break --check-materials A.xml B.xml         (no GPU needed)
//       two EmitterBox blocks, one per material, grouped as SPHSystem::init
//       groups them: prints "materials K particles N"
//   example_dam_break --two-fluid A.xml B.xml --steps K --out FILE
//       a block of each material side by side in a reference-mode
//       PCI_SPHSystem (per-particle materials), K steps; FILE holds the
//       record below followed by the material count, the materials
//       (psb_sph_params each) and one material index byte per particle
//   example_dam_break --steps K --side N --out FILE [--config]
//       runs K steps of an N^3 block (spacing 5 mm) and writes the binary
//       record: params (psb_sph_params), n, initial float4 positions, final
//       float4 positions / velocities (by particle id), iterations of the
//       last step.
//       --batch runs the scene as scene 0 of a two-scene SceneBatch instead
//       (with --config) and writes scene 0's result;
//       --spheres adds two rigid DEM spheres (SPHSystem::setSpheres) and
//       prints their parameters and final state;
//       --colliders adds a SphereCollision obstacle at the block's front and
//       a PlaneCollision plate under it (printed as JSON lines "collider")
//       --surface FILE writes the SurfaceSPH mesh after the K steps (binary:
//       nv, nt, vertices, normals, triangles)
//   state files, written after the K steps / read before them:
//       --export-xml FILE / --export-txt FILE   ParticleExporter_XML / _Txt
//       --checkpoint FILE                       binary checkpoint
//       --resume-xml FILE                       ParticleLoader_XML::restore
//       --resume FILE                           binary checkpoint
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psb_host.hpp"

using namespace psbhost;

static void print_material(const SimulationData_SPHSystem& m) {
  std::printf("{\"particle_radius\": %.9g, \"particle_mass\": %.9g, \"rest_density\": %.9g, "
              "\"viscosity\": %.9g, \"surface_tension\": %.9g, \"threshold\": %.9g, "
              "\"gas_stiffness\": %.9g, \"kernel_particles\": %.9g, \"support_radius\": %.9g}\n",
              m.getParticleRadius(), m.getParticleMass(), m.getRestDensity(), m.getViscosity(),
              m.getSurfaceTension(), m.getThreshold(), m.getGasStiffness(),
              m.getKernelParticles(), m.getSupportRadius());
}

int main(int argc, char** argv) {
  std::string material, save, out, export_xml, export_txt, checkpoint, resume_xml, resume;
  std::string print_xml, print_txt, surface_out, check_a, check_b, two_a, two_b;
  bool print = false, config = false, colliders = false, spheres = false, batch = false;
  int steps = 10, side = 12;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
    if (a == "--material") material = next();
    else if (a == "--save-material") save = next();
    else if (a == "--print-material") print = true;
    else if (a == "--steps") steps = std::stoi(next());
    else if (a == "--side") side = std::stoi(next());
    else if (a == "--out") out = next();
    else if (a == "--config") config = true;
    else if (a == "--colliders") colliders = true;
    else if (a == "--spheres") spheres = true;
    else if (a == "--batch") batch = true;
    else if (a == "--export-xml") export_xml = next();
    else if (a == "--export-txt") export_txt = next();
    else if (a == "--checkpoint") checkpoint = next();
    else if (a == "--resume-xml") resume_xml = next();
    else if (a == "--resume") resume = next();
    else if (a == "--print-xml") print_xml = next();
    else if (a == "--print-txt") print_txt = next();
    else if (a == "--surface") surface_out = next();
    else if (a == "--check-materials") { check_a = next(); check_b = next(); }
    else if (a == "--two-fluid") { two_a = next(); two_b = next(); }
    else { std::fprintf(stderr, "unknown argument %s\n", a.c_str()); return 2; }
  }
  if (!check_a.empty()) {
    SimulationData_PCI_SPHSystem ma, mb;
    if (!ma.loadConfiguration(check_a.c_str()) || !mb.loadConfiguration(check_b.c_str())) {
      std::fprintf(stderr, "cannot read the materials\n");
      return 1;
    }
    std::vector<SPHParticle*> ps = EmitterBox({0.1f, 0.1f, 0.1f}, 0.05f, 0.05f, 0.05f, &ma).emitParticles();
    std::vector<SPHParticle*> pb = EmitterBox({0.3f, 0.1f, 0.1f}, 0.05f, 0.05f, 0.05f, &mb).emitParticles();
    ps.insert(ps.end(), pb.begin(), pb.end());
    const MaterialSet m = collectMaterials(ps, "SPHSystem::init");
    std::printf("materials %zu particles %zu\n", m.mats.size(), ps.size());
    for (auto* q : ps) delete q;
    return 0;
  }
  if (!two_a.empty()) {
    SimulationData_PCI_SPHSystem ma, mb;
    if (!ma.loadConfiguration(two_a.c_str()) || !mb.loadConfiguration(two_b.c_str())) {
      std::fprintf(stderr, "cannot read the materials\n");
      return 1;
    }
    PCI_SPHSystem sys;   // reference mode: the reference's three corrector iterations
    sys.setDt(0.005);
    sys.addForce(new ForceExt_Constante(Vector3(0, -1, 0), 9.81f));
    sys.addObjectCollision(BoxCollision::fromBounds(Vector3(0, 0, 0), Vector3(1.2f, 0.8f, 0.8f), 0.f));
    std::vector<SPHParticle*> ps = EmitterBox({0.25f, 0.2f, 0.4f}, 0.4f, 0.3f, 0.4f, &ma).emitParticles();
    std::vector<SPHParticle*> pb = EmitterBox({0.65f, 0.2f, 0.4f}, 0.4f, 0.3f, 0.4f, &mb).emitParticles();
    ps.insert(ps.end(), pb.begin(), pb.end());
    const MaterialSet m = collectMaterials(ps, "two-fluid");
    const uint32_t n = (uint32_t)ps.size();
    std::vector<float> x0(4 * (size_t)n, 0.f);
    for (uint32_t i = 0; i < n; ++i) {
      const Vector3 q = ps[i]->getPos();
      x0[4 * i] = q.x; x0[4 * i + 1] = q.y; x0[4 * i + 2] = q.z;
    }
    sys.init(ps);
    for (int k = 0; k < steps; ++k) sys.update();
    const psb_step_stats st = sys.stats();
    std::printf("two-fluid materials %u particles %u steps %d iterations %d\n",
                psb_system_num_materials(sys.handle()), n, steps, st.iterations);
    if (out.empty()) return 0;
    std::vector<SPHParticle*> got = sys.getParticles();
    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) return 1;
    const psb_sph_params prm = sys.params();
    std::fwrite(&prm, sizeof(prm), 1, f);
    std::fwrite(&n, 4, 1, f);
    std::fwrite(x0.data(), 4, x0.size(), f);
    std::vector<float> x(4 * (size_t)n, 0.f), v(4 * (size_t)n, 0.f);
    for (uint32_t i = 0; i < n; ++i) {
      const Vector3 q = got[i]->getPos(), u = got[i]->getVel();
      x[4 * i] = q.x; x[4 * i + 1] = q.y; x[4 * i + 2] = q.z;
      v[4 * i] = u.x; v[4 * i + 1] = u.y; v[4 * i + 2] = u.z;
    }
    std::fwrite(x.data(), 4, x.size(), f);
    std::fwrite(v.data(), 4, v.size(), f);
    const int32_t it = st.iterations;
    std::fwrite(&it, 4, 1, f);
    const uint32_t count = (uint32_t)m.mats.size();
    std::fwrite(&count, 4, 1, f);
    for (const SimulationData_SPHSystem* d : m.mats) {
      psb_sph_params q = prm;
      setMaterial(q, d);
      std::fwrite(&q, sizeof(q), 1, f);
    }
    std::fwrite(m.index.data(), 1, m.index.size(), f);
    std::fclose(f);
    return 0;
  }
  SimulationData_PCI_SPHSystem water;
  if (!material.empty() && !water.loadConfiguration(material.c_str())) {
    std::fprintf(stderr, "cannot read %s\n", material.c_str());
    return 1;
  }
  if (print) { print_material(water); return 0; }
  if (!print_xml.empty()) {
    ParticleLoader_XML loader;
    std::vector<ParticleLoader_XML::Record> rs;
    double dt = 0;
    if (!loader.read(print_xml.c_str(), rs, &dt)) { std::fprintf(stderr, "not an SPH System file\n"); return 1; }
    std::vector<SPHParticle*> ps = loader.load(print_xml.c_str());
    std::printf("{\"count\": %zu, \"dt\": %.9g}\n", ps.size(), dt);
    for (size_t i = 0; i < ps.size(); ++i) {
      const SPHParticle* q = ps[i];
      const Vector3 x = q->getPos(), v = q->getVel(), a = q->getVelInterAv(), b = q->getVelInterAp();
      std::printf("{\"index\": %u, \"pos\": [%.9g, %.9g, %.9g], \"vel\": [%.9g, %.9g, %.9g], "
                  "\"vav\": [%.9g, %.9g, %.9g], \"vap\": [%.9g, %.9g, %.9g], \"mass\": %.9g, "
                  "\"support_radius\": %.9g, \"rest_density\": %.9g, \"viscosity\": %.9g}\n",
                  rs[i].index, x.x, x.y, x.z, v.x, v.y, v.z, a.x, a.y, a.z, b.x, b.y, b.z,
                  q->getMass(), q->getInteractionRadius(), q->getRestDensity(),
                  q->getData()->getViscosity());
      delete ps[i];
    }
    return 0;
  }
  if (!print_txt.empty()) {
    std::vector<ParticleLoader_Txt::Record> rs;
    if (!ParticleLoader_Txt().read(print_txt.c_str(), rs)) { std::fprintf(stderr, "not an SPH text file\n"); return 1; }
    std::printf("{\"count\": %zu}\n", rs.size());
    for (const auto& r : rs)
      std::printf("{\"pos\": [%.9g, %.9g, %.9g], \"mass\": %.9g, \"density\": %.9g}\n", r.pos.x,
                  r.pos.y, r.pos.z, r.mass, r.density);
    return 0;
  }
  if (!save.empty()) return water.saveConfiguration(save.c_str()) ? 0 : 1;

  try {
    // the 5 mm lattice of the PCISPH dam break: h = 2 s, mass giving rho0 on a
    // perfect lattice, particle radius s / 2
    const float s = 0.005f, h = 2 * s;
    water.setSupportRadius(h);
    water.setParticleRadius(0.5f * s);
    water.setRestDensity(1000.f);
    water.setViscosity(0.01f);
    water.setSurfaceTension(0.f);
    double acc = 0;
    for (int a = -2; a <= 2; ++a)
      for (int b = -2; b <= 2; ++b)
        for (int c = -2; c <= 2; ++c) {
          const double d2 = (a * a + b * b + c * c) * (double)s * s;
          if (d2 <= (double)h * h) acc += std::pow((double)h * h - d2, 3);
        }
    water.setParticleMass((float)(1000.0 / (315.0 / (64.0 * M_PI * std::pow((double)h, 9)) * acc)));

    PCI_SPHSystem sys;
    psb_sph_params& p = sys.params();
    if (config) {
      p.mode = PSB_MODE_CONFIG;
      p.eos = PSB_EOS_LINEAR;
      p.tolerance = 0.01f;
      p.min_iterations = 3;
      p.max_iterations = 50;
      p.delta_local_cap = 1;
      p.retry_levels = 2;
      p.max_neighbors = 96;
    }
    p.lattice_spacing = s;
    sys.setDt(5e-4);
    sys.addForce(new ForceExt_Constante(Vector3(0, -1, 0), 9.81f));
    const float L = side * s;
    sys.addObjectCollision(BoxCollision::fromBounds(Vector3(0, 0, 0), Vector3(4 * L, 2 * L, 2 * L), 0.5f));
    if (colliders) {
      SphereCollision* ob = new SphereCollision(Vector3(1.2f * L, 0.2f * L, 0.5f * L), 0.25f * L,
                                                0.f, 0.f, false);
      PlaneCollision* plate = new PlaneCollision();
      plate->create(Vector3(0.4f * L, 0.08f * L, 0.5f * L), 0.8f * L, L, 0.f, 0.f, false,
                    Vector3(0, 1, 0));
      for (const psb_collider& k : {ob->desc(), plate->desc()}) {
        std::printf("{\"collider\": %d, \"container\": %d, \"elast\": %.9g, \"radius\": %.9g, "
                    "\"origin\": [%.9g, %.9g, %.9g], \"corners\": [", k.type, k.container, k.elast,
                    k.radius, k.origin[0], k.origin[1], k.origin[2]);
        for (int v = 0; v < 4; ++v)
          std::printf("%s[%.9g, %.9g, %.9g]", v ? ", " : "", k.corners[v][0], k.corners[v][1],
                      k.corners[v][2]);
        std::printf("], \"normal\": [%.9g, %.9g, %.9g]}\n", k.normal[0], k.normal[1], k.normal[2]);
      }
      sys.addObjectCollision(ob);
      sys.addObjectCollision(plate);
    }
    if (spheres) {
      // two rigid DEM spheres in the flow's path
      psb_sphere_params sp{};
      sp.radius = 0.2f * L; sp.density = 500.f; sp.spring = 1e4f; sp.damping = 0.3f;
      const std::vector<Vector3> centers = {Vector3(1.6f * L, 0.25f * L, 0.5f * L),
                                            Vector3(2.2f * L, 0.25f * L, 1.2f * L)};
      sys.setSpheres(centers, sp);
      for (const Vector3& c : centers)
        std::printf("{\"sphere\": [%.9g, %.9g, %.9g], \"radius\": %.9g, \"density\": %.9g, "
                    "\"spring\": %.9g, \"damping\": %.9g}\n", c.x, c.y, c.z, sp.radius, sp.density,
                    sp.spring, sp.damping);
    }
    EmitterBox emitter(Vector3(0.5f * L, 0.5f * L, 0.5f * L), L - s, L - s, L - s, &water);
    std::vector<SPHParticle*> ps = emitter.emitParticles();
    const uint32_t n = (uint32_t)ps.size();
    std::vector<float> x0(4 * (size_t)n, 0.f);
    for (uint32_t i = 0; i < n; ++i) {
      const Vector3 q = ps[i]->getPos();
      x0[4 * i] = q.x; x0[4 * i + 1] = q.y; x0[4 * i + 2] = q.z;
    }
    sys.init(ps);
    if (!resume_xml.empty() && !ParticleLoader_XML().restore(&sys, resume_xml.c_str())) {
      std::fprintf(stderr, "cannot restore %s\n", resume_xml.c_str());
      return 1;
    }
    if (!resume.empty()) sys.loadCheckpoint(resume.c_str());
    if (batch) {
      // the same scene as scene 0 of a SceneBatch, beside a second scene (the
      // block's first half in the same box); scene 0's result goes to --out
      const psb_sph_params m = sys.params();
      psb_scene_desc d{};
      d.count = n;
      for (int a = 0; a < 3; ++a) { d.box_min[a] = m.box_min[a]; d.box_max[a] = m.box_max[a]; }
      d.grid = m.grid;
      psb_scene_desc d1 = d;
      d1.count = n / 2;
      SceneBatch sb(sys.context(), m, {d, d1});
      sb.setParticles(0, x0.data());
      sb.setParticles(1, x0.data());
      sb.update(steps);
      std::vector<float> bx, bv;
      std::vector<uint32_t> ids;
      sb.getParticles(0, n, bx, bv, ids);
      const psb_step_stats st = sb.stats(0);
      std::printf("batch scenes %u particles %u steps %d iterations %d\n", sb.numScenes(), n, steps,
                  st.iterations);
      if (!out.empty()) {
        FILE* f = std::fopen(out.c_str(), "wb");
        if (!f) return 1;
        std::fwrite(&m, sizeof(m), 1, f);
        std::fwrite(&n, 4, 1, f);
        std::fwrite(x0.data(), 4, x0.size(), f);
        std::vector<float> x(4 * (size_t)n, 0.f), v(4 * (size_t)n, 0.f);
        for (uint32_t k = 0; k < n; ++k)
          for (int a = 0; a < 3; ++a) {
            x[4 * (size_t)ids[k] + a] = bx[4 * (size_t)k + a];
            v[4 * (size_t)ids[k] + a] = bv[4 * (size_t)k + a];
          }
        std::fwrite(x.data(), 4, x.size(), f);
        std::fwrite(v.data(), 4, v.size(), f);
        const int32_t it = st.iterations;
        std::fwrite(&it, 4, 1, f);
        std::fclose(f);
      }
      return 0;
    }
    for (int k = 0; k < steps; ++k) sys.update();
    if (!export_xml.empty() && !ParticleExporter_XML()._export(export_xml.c_str(), &sys)) return 1;
    if (!export_txt.empty() && !ParticleExporter_Txt()._export(export_txt.c_str(), &sys)) return 1;
    if (!checkpoint.empty()) sys.saveCheckpoint(checkpoint.c_str());
    if (!surface_out.empty()) {
      SurfaceSPH surf;
      surf.params().cell_size = s;
      if (!surf.extract(&sys, 0.5, steps)) return 1;
      FILE* f = std::fopen(surface_out.c_str(), "wb");
      if (!f) return 1;
      const uint32_t nv = (uint32_t)surf.getVertexs().size(), nt = (uint32_t)surf.getIndexs().size() / 3;
      std::fwrite(&nv, 4, 1, f);
      std::fwrite(&nt, 4, 1, f);
      for (const Vector3& q : surf.getVertexs()) { const float a3[3] = {q.x, q.y, q.z}; std::fwrite(a3, 4, 3, f); }
      for (const Vector3& q : surf.getNormales()) { const float a3[3] = {q.x, q.y, q.z}; std::fwrite(a3, 4, 3, f); }
      std::fwrite(surf.getIndexs().data(), 4, surf.getIndexs().size(), f);
      std::fclose(f);
    }
    if (spheres) {
      const PCI_SPHSystem::SphereState ss = sys.getSpheres();
      for (size_t k = 0; k < ss.centers.size(); ++k)
        std::printf("{\"sphere_final\": [%.9g, %.9g, %.9g], \"velocity\": [%.9g, %.9g, %.9g]}\n",
                    ss.centers[k].x, ss.centers[k].y, ss.centers[k].z, ss.velocities[k].x,
                    ss.velocities[k].y, ss.velocities[k].z);
    }
    std::vector<SPHParticle*> res = sys.getParticles();
    const psb_step_stats st = sys.stats();
    std::printf("particles %u steps %d iterations %d max_error %.6g\n", n, steps, st.iterations,
                st.max_error);
    if (!out.empty()) {
      FILE* f = std::fopen(out.c_str(), "wb");
      if (!f) return 1;
      const psb_sph_params prm = sys.params();
      std::fwrite(&prm, sizeof(prm), 1, f);
      std::fwrite(&n, 4, 1, f);
      std::fwrite(x0.data(), 4, x0.size(), f);
      std::vector<float> x(4 * (size_t)n, 0.f), v(4 * (size_t)n, 0.f);
      for (uint32_t i = 0; i < n; ++i) {
        const Vector3 q = res[i]->getPos(), u = res[i]->getVel();
        x[4 * i] = q.x; x[4 * i + 1] = q.y; x[4 * i + 2] = q.z;
        v[4 * i] = u.x; v[4 * i + 1] = u.y; v[4 * i + 2] = u.z;
      }
      std::fwrite(x.data(), 4, x.size(), f);
      std::fwrite(v.data(), 4, v.size(), f);
      const int32_t it = st.iterations;
      std::fwrite(&it, 4, 1, f);
      std::fclose(f);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
