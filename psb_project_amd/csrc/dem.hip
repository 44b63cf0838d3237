// dem.hip -- DEM contact forces (interactionSystem, CudaSystem.cu:6-14;
// collideParticles / collideParticle, CudaParticle_Kernel.cu:10-55).
//
// One thread per particle A walks A's neighbour list (ELL, list[k n + A]) and
// accumulates the contact force of every neighbour B into a register, then
// adds it to forces[A] once: the same sum order as the reference's
// forces[A] = forces[A] + collideParticle(...) per neighbour, without its
// read-modify-write of global memory per pair.
#include "dem.h"
#include "psb_internal.h"

namespace psb {

template <bool REF>
__global__ __launch_bounds__(256) void k_interaction(uint32_t n, const float4* __restrict__ pos,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ forces,
                                                     const float* __restrict__ radius,
                                                     const float* __restrict__ spring,
                                                     const float* __restrict__ damping,
                                                     const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ list) {
  const uint32_t a = particle_index();
  if (a >= n) return;
  const float4 xa = pos[a];
  const float4 va = REF ? make_float4(0.f, 0.f, 0.f, 0.f) : vel[a];
  const float ra = radius[a], k = spring[a], c = REF ? 0.f : damping[a];
  float4 f = forces[a];
  const uint32_t m = cnt[a];
  for (uint32_t j = 0; j < m; ++j) {
    const uint32_t b = list[(uint64_t)j * n + a];
    const float4 xb = pos[b];
    float fx, fy, fz;
    dem_pair<REF>(xa, xb, REF ? va : vel[b], va, ra, radius[b], k, c, fx, fy, fz);
    f.x = f.x + fx;
    f.y = f.y + fy;
    f.z = f.z + fz;
  }
  forces[a] = f;
}

}  // namespace psb

using namespace psb;

extern "C" {

psb_status psb_interaction_system(psb_context ctx, int32_t mode, const float* pos4,
                                  const float* vel4, float* forces4, const float* radius,
                                  const float* spring, const float* damping, const float* shear,
                                  const float* attraction, psb_neighbors nb, uint32_t n) {
  (void)shear;
  (void)attraction;
  PSB_ARG(ctx, "null context");
  PSB_ARG(mode == PSB_MODE_REFERENCE || mode == PSB_MODE_CONFIG, "mode");
  if (!n) return PSB_OK;
  PSB_ARG(pos4 && forces4 && radius && spring && nb.count && nb.list, "null array");
  PSB_ARG(mode == PSB_MODE_REFERENCE || (vel4 && damping), "config mode needs vel4 and damping");
  PSB_ARG(nb.n == n, "neighbour list stride must equal n");
  PSB_HIP(hipSetDevice(ctx->device));
  const dim3 g(blocks_for(n)), t(kBlock);
  const float4* p = reinterpret_cast<const float4*>(pos4);
  const float4* v = reinterpret_cast<const float4*>(vel4);
  float4* f = reinterpret_cast<float4*>(forces4);
  if (mode == PSB_MODE_REFERENCE)
    hipLaunchKernelGGL(k_interaction<true>, g, t, 0, ctx->stream, n, p, v, f, radius, spring,
                       damping, nb.count, nb.list);
  else
    hipLaunchKernelGGL(k_interaction<false>, g, t, 0, ctx->stream, n, p, v, f, radius, spring,
                       damping, nb.count, nb.list);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

}  // extern "C"
