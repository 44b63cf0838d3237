// bodies.h -- rigid DEM spheres (internal).
#pragma once
#include "psb_internal.h"

namespace psb {

struct SphereArgs {
  const float4* c_in;      // centres at the start of the step
  const float4* v_in;      // velocities at the start of the step
  float4* c_out;           // centres at the end of the step
  float4* v_out;
  float4* f_fluid;         // fluid force of the step, per sphere
  float4* fb;              // fluid force per boundary particle (by boundary id)
  const float4* offsets;   // boundary-particle offsets from the centre [per]
  uint32_t count, per;
  float radius, mass, spring, damping, wall_damping;
  float box_min[3], box_max[3];   // walls for the sphere contacts
};

psb_status launch_bodies(uint32_t n, const Consts& c, const uint32_t* cnt, const uint32_t* list,
                         const float4* X, const uint32_t* ids, const SphereArgs& a, float4* bx,
                         float4* bv, float4* bvh, uint32_t* bid, float* bp, hipStream_t s);

// lattice points (i, j, k) * s strictly inside radius r, in (k, j, i) order
inline std::vector<float> sphere_offsets(double r, double s) {
  std::vector<float> o;
  const int m = (int)std::ceil(r / s);
  for (int k = -m; k <= m; ++k)
    for (int j = -m; j <= m; ++j)
      for (int i = -m; i <= m; ++i) {
        const double d2 = ((double)i * i + (double)j * j + (double)k * k) * s * s;
        if (d2 < r * r) {
          o.push_back((float)(i * s)); o.push_back((float)(j * s)); o.push_back((float)(k * s));
          o.push_back(0.f);
        }
      }
  return o;
}

}  // namespace psb
