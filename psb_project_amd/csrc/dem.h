// dem.h -- DEM contact law shared by the standalone interaction kernel and
// the rigid-sphere coupling.
#pragma once
#include "psb_internal.h"

namespace psb {

// Force on A from B (collideParticle, CudaParticle_Kernel.cu:10-38), with
// d = |x_B - x_A|, n = (x_B - x_A) * (1/d) (double3's '/' multiplies by the
// reciprocal), R = r_A + r_B.
//   REF:  -k (R - d) n           for 0 < d < 2R (the dropped damping, shear
//                                and attraction terms are not restated)
//   else: (-k (R - d) + c v_n) n for 0 < d < R, v_n = (v_B - v_A) . n
template <bool REF>
__device__ __forceinline__ void dem_pair(const float4& xa, const float4& xb, const float4& vb,
                                         const float4& va, float ra, float rb, float k, float c,
                                         float& fx, float& fy, float& fz) {
  fx = 0.f; fy = 0.f; fz = 0.f;
  const float dx = xb.x - xa.x, dy = xb.y - xa.y, dz = xb.z - xa.z;
  const float dist = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
  if (!(dist > 0.f)) return;
  const float R = ra + rb;
  if (REF ? !(dist < R * 2.0f) : !(dist < R)) return;
  const float inv = 1.0f / dist;
  const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
  float s = -k * (R - dist);
  if (!REF) {
    const float vn = dot3(vb.x - va.x, vb.y - va.y, vb.z - va.z, nx, ny, nz);
    s = s + c * vn;
  }
  fx = s * nx; fy = s * ny; fz = s * nz;
}

}  // namespace psb
