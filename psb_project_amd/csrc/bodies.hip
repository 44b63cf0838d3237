// bodies.hip -- rigid DEM spheres coupled to the PCISPH fluid.
//
// Each sphere is a rigid set of boundary particles (lattice points inside
// it) that live in the fluid's particle arrays with ids >= n_fluid.  During
// the corrector they are static neighbours: they add to the fluid density
// and push the fluid with the fluid particle's own pressure (boundary
// pressure 0).  After the corrector, per step:
//   k_body_react      force of the fluid on every boundary particle: the
//                     opposite of the pressure force it exerted, at the
//                     final predicted positions and pressures;
//   k_body_step       one wave per sphere: fluid force summed in
//                     boundary-id order, gravity, DEM contacts with every
//                     other sphere (index order) and the six box walls,
//                     symplectic Euler;
//   k_body_particles  boundary particles follow their sphere (no rotation).
#include "bodies.h"
#include "dem.h"
#include "psb_internal.h"

namespace psb {

__global__ __launch_bounds__(256) void k_body_react(uint32_t n, Consts c,
                                                    const uint32_t* __restrict__ cnt,
                                                    const uint32_t* __restrict__ list,
                                                    const float4* __restrict__ X,
                                                    const uint32_t* __restrict__ ids,
                                                    float4* __restrict__ fb) {
  const uint32_t p = particle_index();
  if (p >= n || !is_boundary(c, ids, p)) return;
  const float4 xb = X[p];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  // boundary neighbours carry w = 0 and add nothing
  for_neighbors<kUnroll>(p, n, cnt[p], list, X, [&](uint32_t, const float4& xq) {
    const float dx = xq.x - xb.x, dy = xq.y - xb.y, dz = xq.z - xb.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (d2 > 0.f && d2 < c.h2) {
      const float dd = sqrtf(d2);
      const float invd = 1.0f / dd;
      const float hd = c.h - dd;
      const float gw = c.c_spiky * (hd * hd);
      const float cc = (gw * invd) * xq.w;
      sx = __builtin_fmaf(dx, cc, sx);
      sy = __builtin_fmaf(dy, cc, sy);
      sz = __builtin_fmaf(dz, cc, sz);
    }
  });
  // the fluid particle q received m kp p_q gradW(x_q - x_b); b gets the opposite
  const float s = -(c.m * c.kp);
  fb[ids[p] - c.n_fluid] = make_float4(sx * s, sy * s, sz * s, 0.f);
}

__device__ __forceinline__ float lane_value(float v, uint32_t u) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)u));
}

// One wave per sphere: the lanes load the boundary forces and evaluate the
// sphere pairs 64 at a time, and every lane then adds them up lane by lane,
// in boundary-id and sphere-index order (each sum in the same serial order as a single loop).
__global__ __launch_bounds__(256) void k_body_step(SphereArgs a, Consts c) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (j >= a.count) return;   // j is uniform over the wave
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (uint32_t l0 = 0; l0 < a.per; l0 += 64u) {
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    if (l0 + lane < a.per) f = a.fb[(size_t)j * a.per + l0 + lane];
    const uint32_t m = min(64u, a.per - l0);
    for (uint32_t u = 0; u < m; ++u) {
      fx = fx + lane_value(f.x, u);
      fy = fy + lane_value(f.y, u);
      fz = fz + lane_value(f.z, u);
    }
  }
  if (lane == 0) a.f_fluid[j] = make_float4(fx, fy, fz, 0.f);
  fx = fx + a.mass * c.gx;
  fy = fy + a.mass * c.gy;
  fz = fz + a.mass * c.gz;
  const float4 cj = a.c_in[j], vj = a.v_in[j];
  for (uint32_t i0 = 0; i0 < a.count; i0 += 64u) {
    const uint32_t i = i0 + lane;
    float gx = 0.f, gy = 0.f, gz = 0.f;
    if (i < a.count && i != j)
      dem_pair<false>(cj, a.c_in[i], a.v_in[i], vj, a.radius, a.radius, a.spring, a.damping, gx,
                      gy, gz);
    const uint32_t m = min(64u, a.count - i0);
    for (uint32_t u = 0; u < m; ++u) {
      if (i0 + u == j) continue;
      fx = fx + lane_value(gx, u);
      fy = fy + lane_value(gy, u);
      fz = fz + lane_value(gz, u);
    }
  }
  // walls: signed distance d of the centre (positive inside), overlap
  // R - d along the inward normal, dashpot on the normal velocity
  const float pc[3] = {cj.x, cj.y, cj.z}, pv[3] = {vj.x, vj.y, vj.z};
  float fw[3] = {0.f, 0.f, 0.f};
  for (int ax = 0; ax < 3; ++ax) {
    for (int side = 0; side < 2; ++side) {
      const float d = side == 0 ? pc[ax] - a.box_min[ax] : a.box_max[ax] - pc[ax];
      if (d < a.radius) {
        const float nin = side == 0 ? 1.0f : -1.0f;
        const float sw = a.spring * (a.radius - d) - a.wall_damping * (pv[ax] * nin);
        fw[ax] = fw[ax] + sw * nin;
      }
    }
  }
  fx = fx + fw[0]; fy = fy + fw[1]; fz = fz + fw[2];
  const float inv_m = 1.0f / a.mass;
  const float vx = vj.x + (fx * inv_m) * c.dt, vy = vj.y + (fy * inv_m) * c.dt,
              vzz = vj.z + (fz * inv_m) * c.dt;
  if (lane != 0) return;
  a.v_out[j] = make_float4(vx, vy, vzz, 0.f);
  a.c_out[j] = make_float4(cj.x + vx * c.dt, cj.y + vy * c.dt, cj.z + vzz * c.dt, 0.f);
}

__global__ __launch_bounds__(256) void k_body_particles(uint32_t n, Consts c, SphereArgs a,
                                                        const uint32_t* __restrict__ ids,
                                                        float4* __restrict__ bx,
                                                        float4* __restrict__ bv,
                                                        float4* __restrict__ bvh,
                                                        uint32_t* __restrict__ bid,
                                                        float* __restrict__ bp) {
  const uint32_t p = particle_index();
  if (p >= n || !is_boundary(c, ids, p)) return;
  const uint32_t k = ids[p] - c.n_fluid;
  const uint32_t j = k / a.per, l = k - j * a.per;
  const float4 o = a.offsets[l], cn = a.c_out[j], v = a.v_out[j];
  bx[p] = make_float4(cn.x + o.x, cn.y + o.y, cn.z + o.z, 0.f);
  bv[p] = v;
  bvh[p] = v;
  bid[p] = ids[p];
  bp[p] = 0.f;
}

psb_status launch_bodies(uint32_t n, const Consts& c, const uint32_t* cnt, const uint32_t* list,
                         const float4* X, const uint32_t* ids, const SphereArgs& a, float4* bx,
                         float4* bv, float4* bvh, uint32_t* bid, float* bp, hipStream_t s) {
  if (!a.count) return PSB_OK;
  hipLaunchKernelGGL(k_body_react, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, c, cnt, list, X,
                     ids, a.fb);
  hipLaunchKernelGGL(k_body_step, dim3(blocks_for((uint64_t)a.count * 64u)), dim3(kBlock), 0, s, a, c);
  hipLaunchKernelGGL(k_body_particles, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, c, a, ids, bx,
                     bv, bvh, bid, bp);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

}  // namespace psb
