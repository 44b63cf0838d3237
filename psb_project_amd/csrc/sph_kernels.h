// sph_kernels.h -- per-particle neighbour sums shared by the WCSPH and PCISPH
// kernels (device code).
#pragma once
#include "psb_internal.h"

namespace psb {

struct PairSums {
  float fp[3], fv[3], fs, nn[3];
};

// Pair loop of internalForces (SphKernel.cu:47-78) and of
// compute_Viscosity_SurfaceTension_Force_PCI (PciSphSystem_Kernel.cu:134-158)
// with uniform material: xs[q].w = p_q/rho_q^2 (used when PRESS),
// vs[q].w = 1/rho_q.  Sums exclude d = 0 (the particle itself).  LI: list
// accessor (List32, List16) over the particle's m entries.
template <bool PRESS, bool SURF, class LI>
__device__ __forceinline__ PairSums internal_sums_l(const Consts& c, uint32_t p, uint32_t n,
                                                    uint32_t m, const LI& list,
                                                    const float4* __restrict__ xs,
                                                    const float4* __restrict__ vs) {
  PairSums s;
  s.fp[0] = s.fp[1] = s.fp[2] = 0.f;
  s.fv[0] = s.fv[1] = s.fv[2] = 0.f;
  s.fs = 0.f;
  s.nn[0] = s.nn[1] = s.nn[2] = 0.f;
  const float4 xi = xs[p], vi = vs[p];
  const float pvi = xi.w;
  for_pairs2<kUnroll>(p, n, m, list, xs, vs, [&](const float4& xj, const float4& vj) {
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (d2 > 0.f) {
      float dd, invd = 0.f;
      if (PRESS) dist_and_inverse(d2, dd, invd);
      else dd = sqrt_rn(d2);
      const float hd = c.h - dd;
      const float rinvj = vj.w;
      if (PRESS) {
        const float gw = c.c_spiky * (hd * hd);
        const float cc = gw * (pvi + xj.w);
        s.fp[0] = __builtin_fmaf(dx * invd, cc, s.fp[0]);
        s.fp[1] = __builtin_fmaf(dy * invd, cc, s.fp[1]);
        s.fp[2] = __builtin_fmaf(dz * invd, cc, s.fp[2]);
      }
      const float lv = hd * rinvj;
      s.fv[0] = __builtin_fmaf(vj.x - vi.x, lv, s.fv[0]);
      s.fv[1] = __builtin_fmaf(vj.y - vi.y, lv, s.fv[1]);
      s.fv[2] = __builtin_fmaf(vj.z - vi.z, lv, s.fv[2]);
      if (SURF) {
        const float a = c.h2 - d2;
        const float b = c.h2x3 - 7.0f * d2;
        s.fs = __builtin_fmaf(a * b, rinvj, s.fs);
        const float an = (a * a) * rinvj;
        s.nn[0] = __builtin_fmaf(dx, an, s.nn[0]);
        s.nn[1] = __builtin_fmaf(dy, an, s.nn[1]);
        s.nn[2] = __builtin_fmaf(dz, an, s.nn[2]);
      }
    }
  });
  return s;
}

template <bool PRESS, bool SURF>
__device__ __forceinline__ PairSums internal_sums(const Consts& c, uint32_t p, uint32_t n,
                                                  const uint32_t* __restrict__ cnt,
                                                  const uint32_t* __restrict__ list,
                                                  const float4* __restrict__ xs,
                                                  const float4* __restrict__ vs) {
  return internal_sums_l<PRESS, SURF>(c, p, n, cnt[p], List32{list, n}, xs, vs);
}

// SphKernel.cu:80-88: f_S = (n (fs * -sigma)) / |n| when |n| >= l.
__device__ __forceinline__ void surface_finish(const Consts& c, const PairSums& s, float sigma,
                                               float fS[3], float nrm[3]) {
  const float fsum = s.fs * c.cs;
  const float n0 = s.nn[0] * c.cn, n1 = s.nn[1] * c.cn, n2 = s.nn[2] * c.cn;
  const float ln = sqrtf(dot3(n0, n1, n2, n0, n1, n2));
  if (ln >= c.l) {
    const float tt = fsum * (-sigma);
    const float r = 1.0f / ln;
    fS[0] = (n0 * tt) * r; fS[1] = (n1 * tt) * r; fS[2] = (n2 * tt) * r;
    nrm[0] = n0 / ln; nrm[1] = n1 / ln; nrm[2] = n2 / ln;
  } else {
    fS[0] = fS[1] = fS[2] = 0.f;
    nrm[0] = nrm[1] = nrm[2] = 0.f;
  }
}

// collide (box, then the added colliders in order, on position and half-step
// velocity, SphSystem.cpp:323-328) then interpolateSPH_velocities
// (SphKernel.cu:128-135); writes the B state.  x0: start-of-step position.
__device__ __forceinline__ void finish_particle(const Consts& c, uint32_t p, float4 x, float4 vh,
                                                float4 va, float4 x0, float4* __restrict__ bx,
                                                float4* __restrict__ bv,
                                                float4* __restrict__ bvh) {
  if (c.box_enabled) box_collide(c, x, vh);
  for (int k = 0; k < c.ncoll; ++k) collide_one(c.coll[k], c.dt, x0, x, vh);
  bx[p] = make_float4(x.x, x.y, x.z, 0.f);
  bvh[p] = make_float4(vh.x, vh.y, vh.z, 0.f);
  bv[p] = make_float4((va.x + vh.x) * 0.5f, (va.y + vh.y) * 0.5f, (va.z + vh.z) * 0.5f, 0.f);
}

}  // namespace psb
