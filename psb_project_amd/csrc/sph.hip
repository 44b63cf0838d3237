// sph.hip -- WCSPH kernels for gfx950 and the pointwise step kernels.
//
// Reference: densityEvaluation / internalForces / integrateSPH_LeapFrog /
// interpolateSPH_velocities (SphKernel.cu:9-135), evaluate_force_constante
// (ForceExt_Constante.cu:15-24), collisionBox (BoxCollision_Kernel.cu:7-34).
//
// The fused system step runs two neighbour passes per WCSPH step:
//   k_neighbors<DENS> (grid.hip) with the density epilogue inside: list,
//       rho, p = EOS(rho), x.w = p / rho^2, v.w = 1 / rho
//   k_wcsph_forces: pressure + viscosity (+ surface) force densities, gravity
//       g*rho, leapfrog, box, velocity interpolation, all in registers; the
//       new state goes to the B buffers.
// Arithmetic follows the canonical FP32 order documented in oracle.cpp (the
// same expressions, fused multiply-adds only where __builtin_fmaf is written).
#include "psb_internal.h"
#include "pcisph.h"
#include "sph_kernels.h"

namespace psb {

// Density from an existing list (standalone evaluate_densities_forces);
// copies x, v into the scratch vec4 arrays with their w lanes filled.
__global__ __launch_bounds__(256) void k_density_list(uint32_t n, Consts c,
                                                      const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ list,
                                                      const float4* __restrict__ pos,
                                                      const float4* __restrict__ vel,
                                                      float* __restrict__ rho,
                                                      float* __restrict__ pres,
                                                      float4* __restrict__ xs,
                                                      float4* __restrict__ vs) {
  const uint32_t p = particle_index();
  if (p >= n) return;
  const float4 xi = pos[p];
  float S = 0.f;
  const uint32_t m = cnt[p];
  for (uint32_t k = 0; k < m; ++k) {
    const float4 xj = pos[list[(uint64_t)k * n + p]];
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    const float w = c.h2 - d2;
    S = __builtin_fmaf(w * w, w, S);
  }
  const float r = S * c.mc;
  const float pr = eos(c, r);
  rho[p] = r;
  pres[p] = pr;
  xs[p] = make_float4(xi.x, xi.y, xi.z, pr / (r * r));
  const float4 vi = vel[p];
  vs[p] = make_float4(vi.x, vi.y, vi.z, 1.0f / r);
}

struct WcsphOut {
  // fused step
  float4* bx; float4* bv; float4* bvh; uint32_t* bid; float* bp;
  const float4* avh; const uint32_t* aid; const float* pres;
  // standalone (may be null)
  float4* normals; float4* fP; float4* fV; float4* fS; float4* F;
  uint32_t* flags;
};

// internalForces (SphKernel.cu:32-91) [+ gravity, leapfrog, box, interpolation]
template <bool SURF, bool STEP>
__global__ __launch_bounds__(256) void k_wcsph_forces(uint32_t n, Consts c,
                                                      const uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ list,
                                                      const float4* __restrict__ xs,
                                                      const float4* __restrict__ vs,
                                                      const float* __restrict__ rho, WcsphOut o) {
  const uint32_t p = particle_index();
  if (p >= n) return;
  PairSums s = internal_sums<true, SURF>(c, p, n, cnt, list, xs, vs);
  const float ri = rho[p];
  const float cpf = -(c.m * ri);
  float fP[3] = {s.fp[0] * cpf, s.fp[1] * cpf, s.fp[2] * cpf};
  float fV[3] = {s.fv[0] * c.cv, s.fv[1] * c.cv, s.fv[2] * c.cv};
  float fS[3] = {0.f, 0.f, 0.f}, nrm[3] = {0.f, 0.f, 0.f};
  if (SURF) surface_finish(c, s, c.sigma, fS, nrm);
  float F0 = (fV[0] + fS[0]) + fP[0];
  float F1 = (fV[1] + fS[1]) + fP[1];
  float F2 = (fV[2] + fS[2]) + fP[2];
  if (!STEP) {
    if (o.normals) o.normals[p] = make_float4(nrm[0], nrm[1], nrm[2], 0.f);
    if (o.fP) o.fP[p] = make_float4(fP[0], fP[1], fP[2], 0.f);
    if (o.fV) o.fV[p] = make_float4(fV[0], fV[1], fV[2], 0.f);
    if (o.fS) o.fS[p] = make_float4(fS[0], fS[1], fS[2], 0.f);
    if (o.F) o.F[p] = make_float4(F0, F1, F2, 0.f);
    return;
  }
  if (!(ri > 0.f)) atomicOr(o.flags, PSB_FLAG_ZERO_DENSITY);
  // evaluate_force_constante with the density as weight (SphSystem.cpp:312)
  F0 = F0 + c.gx * ri; F1 = F1 + c.gy * ri; F2 = F2 + c.gz * ri;
  // integrateSPH_LeapFrog (SphKernel.cu:118-121)
  const float inv = 1.0f / ri;
  float4 vh = o.avh[p];
  const float4 va = vh;
  vh.x = vh.x + (F0 * inv) * c.dt;
  vh.y = vh.y + (F1 * inv) * c.dt;
  vh.z = vh.z + (F2 * inv) * c.dt;
  const float4 x0 = xs[p];
  float4 x = x0;
  x.x = x.x + vh.x * c.dt; x.y = x.y + vh.y * c.dt; x.z = x.z + vh.z * c.dt;
  finish_particle(c, p, x, vh, va, x0, o.bx, o.bv, o.bvh);
  o.bid[p] = o.aid[p];
  o.bp[p] = o.pres[p];
}

// ---- pointwise kernels ------------------------------------------------------
__global__ __launch_bounds__(256) void k_force_constante(uint32_t n, float4* __restrict__ acc,
                                                         const float* __restrict__ wgt, float gx,
                                                         float gy, float gz) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = acc[i];
  const float m = wgt[i];
  a.x = a.x + gx * m; a.y = a.y + gy * m; a.z = a.z + gz * m;
  acc[i] = a;
}

__global__ __launch_bounds__(256) void k_integrate(uint32_t n, float4* __restrict__ vav,
                                                   float4* __restrict__ vap, float4* __restrict__ xo,
                                                   float4* __restrict__ xn,
                                                   const float4* __restrict__ F,
                                                   const float* __restrict__ rho, float dt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 vh = vap[i];
  vav[i] = vh;
  const float4 f = F[i];
  const float inv = 1.0f / rho[i];
  vh.x = vh.x + (f.x * inv) * dt; vh.y = vh.y + (f.y * inv) * dt; vh.z = vh.z + (f.z * inv) * dt;
  vap[i] = vh;
  float4 x = xn[i];
  xo[i] = x;
  x.x = x.x + vh.x * dt; x.y = x.y + vh.y * dt; x.z = x.z + vh.z * dt;
  xn[i] = x;
}

__global__ __launch_bounds__(256) void k_interpolate(uint32_t n, const float4* __restrict__ vav,
                                                     const float4* __restrict__ vap,
                                                     float4* __restrict__ vo,
                                                     float4* __restrict__ vn) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vo[i] = vn[i];
  const float4 a = vav[i], b = vap[i];
  vn[i] = make_float4((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, vn[i].w);
}

__global__ __launch_bounds__(256) void k_box(uint32_t n, Consts c, float4* __restrict__ x,
                                             float4* __restrict__ v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 xi = x[i], vi = v[i];
  box_collide(c, xi, vi);
  x[i] = xi;
  v[i] = vi;
}

__global__ __launch_bounds__(256) void k_collider(uint32_t n, ColliderDev k, float dt,
                                                  const float4* __restrict__ x0,
                                                  float4* __restrict__ x, float4* __restrict__ v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 xi = x[i], vi = v[i];
  collide_one(k, dt, x0[i], xi, vi);
  x[i] = xi;
  v[i] = vi;
}

// ---- launchers used by system.hip --------------------------------------------
psb_status launch_density_list(uint32_t n, const Consts& c, psb_neighbors nb, const float4* pos,
                               const float4* vel, float* rho, float* pres, float4* xs, float4* vs,
                               hipStream_t s) {
  if (!n) return PSB_OK;
  hipLaunchKernelGGL(k_density_list, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, c, nb.count,
                     nb.list, pos, vel, rho, pres, xs, vs);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

EpiArgs make_epilogue(const Consts& c, int pcisph, int warm, int xw_prev, float* pres,
                      const float* pprev, float4* xs, float4* vs, const uint32_t* ids) {
  EpiArgs e{};
  e.on = 1;
  e.pcisph = pcisph; e.warm = warm; e.xw_prev = xw_prev;
  e.c = c;
  e.pres = pres; e.pprev = pprev; e.xs = xs; e.vs = vs; e.ids = ids;
  return e;
}

psb_status launch_wcsph_step_forces(uint32_t n, const Consts& c, bool surf, const uint32_t* cnt,
                                    const uint32_t* list, const float4* xs, const float4* vs,
                                    const float* rho, const float* pres, const float4* avh,
                                    const uint32_t* aid, float4* bx, float4* bv, float4* bvh,
                                    uint32_t* bid, float* bp, uint32_t* flags, hipStream_t s) {
  if (!n) return PSB_OK;
  WcsphOut o{};
  o.bx = bx; o.bv = bv; o.bvh = bvh; o.bid = bid; o.bp = bp; o.avh = avh; o.aid = aid;
  o.pres = pres; o.flags = flags;
  if (surf)
    hipLaunchKernelGGL((k_wcsph_forces<true, true>), dim3(blocks_for(n)), dim3(kBlock), 0, s, n, c,
                       cnt, list, xs, vs, rho, o);
  else
    hipLaunchKernelGGL((k_wcsph_forces<false, true>), dim3(blocks_for(n)), dim3(kBlock), 0, s, n,
                       c, cnt, list, xs, vs, rho, o);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

}  // namespace psb

using namespace psb;

extern "C" {

psb_status psb_evaluate_densities_forces(psb_context ctx, const psb_sph_params* prm,
                                         psb_neighbors nb, const float* pos4, const float* vel4,
                                         uint32_t n, float* density, float* pressure,
                                         float* normals4, float* f_pressure4,
                                         float* f_viscosity4, float* f_surface4,
                                         float* forces_accum4) {
  PSB_ARG(ctx && prm && (n == 0 || (pos4 && vel4 && density && pressure)), "null argument");
  PSB_ARG(nb.n == n && nb.count && nb.list, "neighbour list stride must equal n");
  Consts c{};
  PSB_TRY(make_consts(*prm, &c, nullptr));
  if (!n) return PSB_OK;
  PSB_HIP(hipSetDevice(ctx->device));
  PSB_TRY(ensure_scratch(ctx, 32 * (size_t)n));
  float4* xs = reinterpret_cast<float4*>(ctx->scratch);
  float4* vs = xs + n;
  hipLaunchKernelGGL(k_density_list, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n, c,
                     nb.count, nb.list, reinterpret_cast<const float4*>(pos4),
                     reinterpret_cast<const float4*>(vel4), density, pressure, xs, vs);
  WcsphOut o{};
  o.normals = reinterpret_cast<float4*>(normals4);
  o.fP = reinterpret_cast<float4*>(f_pressure4);
  o.fV = reinterpret_cast<float4*>(f_viscosity4);
  o.fS = reinterpret_cast<float4*>(f_surface4);
  o.F = reinterpret_cast<float4*>(forces_accum4);
  if (prm->surface_tension != 0.f)
    hipLaunchKernelGGL((k_wcsph_forces<true, false>), dim3(blocks_for(n)), dim3(kBlock), 0,
                       ctx->stream, n, c, nb.count, nb.list, xs, vs, density, o);
  else
    hipLaunchKernelGGL((k_wcsph_forces<false, false>), dim3(blocks_for(n)), dim3(kBlock), 0,
                       ctx->stream, n, c, nb.count, nb.list, xs, vs, density, o);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_force_constante(psb_context ctx, uint32_t n, float* accum4, const float* weight,
                               const float direction[3], float amplitude) {
  PSB_ARG(ctx && direction && (n == 0 || (accum4 && weight)), "null argument");
  if (!n) return PSB_OK;
  // direction.x * amplitude in float (ForceExt_Constante.cu:19-21)
  const float gx = direction[0] * amplitude, gy = direction[1] * amplitude,
              gz = direction[2] * amplitude;
  hipLaunchKernelGGL(k_force_constante, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n,
                     reinterpret_cast<float4*>(accum4), weight, gx, gy, gz);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_integrate_sph(psb_context ctx, float* vav4, float* vap4, float* old_pos4,
                             float* new_pos4, const float* forces4, const float* density, float dt,
                             uint32_t n) {
  PSB_ARG(ctx && (n == 0 || (vav4 && vap4 && old_pos4 && new_pos4 && forces4 && density)),
          "null argument");
  if (!n) return PSB_OK;
  hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n,
                     reinterpret_cast<float4*>(vav4), reinterpret_cast<float4*>(vap4),
                     reinterpret_cast<float4*>(old_pos4), reinterpret_cast<float4*>(new_pos4),
                     reinterpret_cast<const float4*>(forces4), density, dt);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_interpolate_velocities(psb_context ctx, const float* vav4, const float* vap4,
                                      float* old_vel4, float* new_vel4, uint32_t n) {
  PSB_ARG(ctx && (n == 0 || (vav4 && vap4 && old_vel4 && new_vel4)), "null argument");
  if (!n) return PSB_OK;
  hipLaunchKernelGGL(k_interpolate, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n,
                     reinterpret_cast<const float4*>(vav4), reinterpret_cast<const float4*>(vap4),
                     reinterpret_cast<float4*>(old_vel4), reinterpret_cast<float4*>(new_vel4));
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_collision_box(psb_context ctx, float* new_pos4, const float* old_vel4,
                             float* new_vel4, float radius, float dt, uint32_t n,
                             const float box_min[3], const float box_max[3], int32_t container,
                             float elast) {
  (void)old_vel4; (void)dt; (void)container;
  PSB_ARG(ctx && box_min && box_max && (n == 0 || (new_pos4 && new_vel4)), "null argument");
  if (!n) return PSB_OK;
  Consts c{};
  for (int a = 0; a < 3; ++a) {
    c.box_lo[a] = box_min[a] + radius;
    c.box_hi[a] = box_max[a] - radius;
  }
  c.one_plus_e = 1.0f + elast;
  hipLaunchKernelGGL(k_box, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n, c,
                     reinterpret_cast<float4*>(new_pos4), reinterpret_cast<float4*>(new_vel4));
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

static psb_status run_collider(psb_context ctx, const psb_collider& in, float dt, uint32_t n,
                               const float* old_pos4, float* new_pos4, float* new_vel4) {
  ColliderDev k;
  PSB_TRY(make_collider(in, &k));
  PSB_ARG(dt > 0.f, "dt must be > 0");
  if (!n) return PSB_OK;
  hipLaunchKernelGGL(k_collider, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n, k, dt,
                     reinterpret_cast<const float4*>(old_pos4),
                     reinterpret_cast<float4*>(new_pos4), reinterpret_cast<float4*>(new_vel4));
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_collision_sphere(psb_context ctx, const float* old_pos4, float* new_pos4,
                                const float* old_vel4, float* new_vel4, float radius_particle,
                                float dt, uint32_t n, const float origin[3], float radius,
                                float elast, int32_t container) {
  (void)old_vel4; (void)radius_particle;
  PSB_ARG(ctx && origin && (n == 0 || (old_pos4 && new_pos4 && new_vel4)), "null argument");
  psb_collider k{};
  k.type = PSB_COLLIDER_SPHERE;
  k.container = container;
  k.elast = elast;
  k.radius = radius;
  for (int a = 0; a < 3; ++a) k.origin[a] = origin[a];
  return run_collider(ctx, k, dt, n, old_pos4, new_pos4, new_vel4);
}

psb_status psb_collision_plane(psb_context ctx, const float* old_pos4, float* new_pos4,
                               const float* old_vel4, float* new_vel4, const float a[3],
                               const float b[3], const float c[3], const float d[3],
                               const float normal[3], float elast, float dt, uint32_t n) {
  (void)old_vel4;
  PSB_ARG(ctx && a && b && c && d && normal && (n == 0 || (old_pos4 && new_pos4 && new_vel4)),
          "null argument");
  psb_collider k{};
  k.type = PSB_COLLIDER_PLANE;
  k.elast = elast;
  for (int i = 0; i < 3; ++i) {
    k.corners[0][i] = a[i]; k.corners[1][i] = b[i]; k.corners[2][i] = c[i]; k.corners[3][i] = d[i];
    k.normal[i] = normal[i];
  }
  return run_collider(ctx, k, dt, n, old_pos4, new_pos4, new_vel4);
}

}  // extern "C"
