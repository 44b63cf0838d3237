// mixed.hip -- reference-mode WCSPH and PCISPH steps with per-particle
// materials (a table of at most kMaxMaterials, one index per particle).
//
// One thread per particle over the 32-bit list in the WCSPH step, four in
// the PCISPH passes (lane l sums the entries k = l mod 4, the partial sums
// added as (s0 + s1) + (s2 + s3)); the
// material table (3.3 KB) is copied into LDS by every workgroup.  Canonical FP32 order (the oracle
// runs the same expressions; fma only where written):
//   list:      j kept when dot3(d) <= own[a].h2 at the start positions
//   density:   S = fma((w * w) * w, P.mc, S), w = P.h2 - d2 (no clamp)
//   WCSPH p:   k (rho - rho0); x.w = p / rho^2, v.w = 1 / rho
//   pair, d2 > 0: dd = sqrt(d2), invd = 1 / dd, hd = P.h - dd
//     pressure  c = (P.csm * (hd * hd)) * (pv_i + pv_j); fp = fma(dx * invd, c, fp)
//     viscosity lv = ((P.cvm * hd) * rinv_j) * P.mu;   fv = fma(v_j - v_i, lv, fv)
//     surface   a = P.h2 - d2, b = P.h2x3 - 7 d2, cj = P.csf * rinv_j
//               fs = fma(a * b, cj, fs); nn = fma(dx, (a * a) * cj, nn)
//   surface finish: |nn| >= l: fS = (nn * (fs * -sigma)) * (1 / |nn|)
//   PCISPH W1/W2 (d2 > 0): g = P.cs * (hd * hd), u = (dx * invd) * g,
//     W1 += u, W2 += dot3(u, u); p += e / (beta (dot3(W1, W1) + W2)) when
//     rho > rho0 and W2 > 0.
// Reference: SphKernel.cu:9-91 (WCSPH), PciSphSystem_Kernel.cu:36-170
// (PCISPH), PciSphSystem.cu:8-68 (the three-iteration loop).
#include "mixed.h"
#include "sph_kernels.h"

namespace psb {

// Copies the table into LDS (every thread of the workgroup takes part).
__device__ __forceinline__ void load_table(const MatTable* __restrict__ g, MatTable& l) {
  const float* src = reinterpret_cast<const float*>(g);
  float* dst = reinterpret_cast<float*>(&l);
  for (uint32_t i = threadIdx.x; i < sizeof(MatTable) / 4; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

__device__ __forceinline__ const MatPair& pair_of(const MatTable& t, uint32_t a, uint32_t b) {
  return t.pair[a * kMaxMaterials + b];
}

__global__ __launch_bounds__(256) void k_mix_prepare(MixArgs a) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t g = particle_index();
  if (g == 0)
    for (int k = 0; k < kMaxIterSlots; ++k) a.st->err[k] = 0u;
  if (a.es.base)
    for (uint32_t k = g; k < (uint32_t)(kErrSlots * kMaxIterSlots); k += gridDim.x * blockDim.x)
      a.es.base[(size_t)(k / kMaxIterSlots) * a.es.stride + k % kMaxIterSlots] = 0u;
  // kMixLanes lanes per particle test four consecutive entries per round; a
  // ballot over the group gives each kept entry its place, so the list is
  // compacted in place in its own order (a round writes only to positions
  // its group has already read)
  const uint32_t p = g / kMixLanes, lane = g % kMixLanes;
  if (p >= a.n) return;   // a particle's four lanes leave together
  const uint32_t m = a.mat_by_id[a.aid[p]];
  if (lane == 0u) a.ms[p] = (uint8_t)m;
  const float h2 = t.own[m].h2;
  const float4 xi = a.ax[p];
  const uint32_t c = a.cnt[p];
  const uint32_t first = (threadIdx.x & 63u) & ~(kMixLanes - 1u);
  const uint32_t below = (1u << lane) - 1u;
  uint32_t kept = 0;
  for (uint32_t k0 = 0; k0 < c; k0 += kMixLanes) {
    const uint32_t k = k0 + lane;
    bool keep = false;
    uint32_t j = 0;
    if (k < c) {
      j = a.list[(uint64_t)k * a.n + p];
      const float4 xj = a.ax[j];
      const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      keep = dot3(dx, dy, dz, dx, dy, dz) <= h2;
    }
    const uint32_t grp = (uint32_t)(__ballot(keep) >> first) & ((1u << kMixLanes) - 1u);
    if (keep) a.list[(uint64_t)(kept + __popc(grp & below)) * a.n + p] = j;
    kept += __popc(grp);
  }
  if (lane == 0u) a.cnt[p] = kept;
}

// The PCISPH passes take kMixLanes threads per particle: lane l sums the
// list entries k = l (mod 4) in increasing k and the four partial sums are
// added as (s0 + s1) + (s2 + s3), the same bits on every lane of the group.
__device__ __forceinline__ float mix_total(float v) {
  const float x = v + __shfl_xor(v, 1, 64);
  return x + __shfl_xor(x, 2, 64);
}

// Density over positions X (start or predicted); L lanes per particle (1 or
// kMixLanes, this thread being `lane` of them).
template <uint32_t L>
__device__ __forceinline__ float mixed_density(const MatTable& t, const MixArgs& a, uint32_t p,
                                               uint32_t lane, uint32_t mi,
                                               const float4* __restrict__ X) {
  const float4 xi = X[p];
  float S = 0.f;
  const uint32_t c = a.cnt[p];
  for (uint32_t k = lane; k < c; k += L) {
    const uint32_t j = a.list[(uint64_t)k * a.n + p];
    const float4 xj = X[j];
    const MatPair& P = pair_of(t, mi, a.ms[j]);
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float w = P.h2 - dot3(dx, dy, dz, dx, dy, dz);
    S = __builtin_fmaf((w * w) * w, P.mc, S);
  }
  return L == 1 ? S : mix_total(S);
}

struct MixSums {
  float fp[3], fv[3], fs, nn[3];
};

// Pair sums of internalForces / compute_Viscosity_SurfaceTension_Force_PCI
// on the start positions (x.w = p / rho^2 when PRESS, v.w = 1 / rho).
template <bool PRESS, uint32_t L>
__device__ __forceinline__ MixSums mixed_sums(const MatTable& t, const MixArgs& a, uint32_t p,
                                              uint32_t lane, uint32_t mi) {
  MixSums s;
  s.fp[0] = s.fp[1] = s.fp[2] = 0.f;
  s.fv[0] = s.fv[1] = s.fv[2] = 0.f;
  s.fs = 0.f;
  s.nn[0] = s.nn[1] = s.nn[2] = 0.f;
  const float4 xi = a.ax[p], vi = a.av[p];
  const uint32_t c = a.cnt[p];
  for (uint32_t k = lane; k < c; k += L) {
    const uint32_t j = a.list[(uint64_t)k * a.n + p];
    const float4 xj = a.ax[j], vj = a.av[j];
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (!(d2 > 0.f)) continue;
    const MatPair& P = pair_of(t, mi, a.ms[j]);
    float dd, invd;
    dist_and_inverse(d2, dd, invd);
    const float hd = P.h - dd;
    const float rinvj = vj.w;
    if (PRESS) {
      const float cc = (P.csm * (hd * hd)) * (xi.w + xj.w);
      s.fp[0] = __builtin_fmaf(dx * invd, cc, s.fp[0]);
      s.fp[1] = __builtin_fmaf(dy * invd, cc, s.fp[1]);
      s.fp[2] = __builtin_fmaf(dz * invd, cc, s.fp[2]);
    }
    const float lv = ((P.cvm * hd) * rinvj) * P.mu;
    s.fv[0] = __builtin_fmaf(vj.x - vi.x, lv, s.fv[0]);
    s.fv[1] = __builtin_fmaf(vj.y - vi.y, lv, s.fv[1]);
    s.fv[2] = __builtin_fmaf(vj.z - vi.z, lv, s.fv[2]);
    const float aa = P.h2 - d2;
    const float bb = P.h2x3 - 7.0f * d2;
    const float cj = P.csf * rinvj;
    s.fs = __builtin_fmaf(aa * bb, cj, s.fs);
    const float an = (aa * aa) * cj;
    s.nn[0] = __builtin_fmaf(dx, an, s.nn[0]);
    s.nn[1] = __builtin_fmaf(dy, an, s.nn[1]);
    s.nn[2] = __builtin_fmaf(dz, an, s.nn[2]);
  }
  if (L > 1) {
    for (int i = 0; i < 3; ++i) {
      if (PRESS) s.fp[i] = mix_total(s.fp[i]);
      s.fv[i] = mix_total(s.fv[i]);
      s.nn[i] = mix_total(s.nn[i]);
    }
    s.fs = mix_total(s.fs);
  }
  return s;
}

__device__ __forceinline__ void mixed_surface(const MixSums& s, float l, float sigma,
                                              float fS[3]) {
  const float ln = sqrtf(dot3(s.nn[0], s.nn[1], s.nn[2], s.nn[0], s.nn[1], s.nn[2]));
  if (ln >= l) {
    const float tt = s.fs * (-sigma);
    const float r = 1.0f / ln;
    fS[0] = (s.nn[0] * tt) * r; fS[1] = (s.nn[1] * tt) * r; fS[2] = (s.nn[2] * tt) * r;
  } else {
    fS[0] = fS[1] = fS[2] = 0.f;
  }
}

// densityEvaluation (SphKernel.cu:9-27) on the start positions.
__global__ __launch_bounds__(256) void k_mix_wcsph_density(MixArgs a) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t p = particle_index();
  if (p >= a.n) return;
  const uint32_t mi = a.ms[p];
  const float r = mixed_density<1>(t, a, p, 0u, mi, a.ax);
  const float pr = t.own[mi].k * (r - t.own[mi].rho0);
  a.rho[p] = r;
  a.pres[p] = pr;
  a.ax[p].w = pr / (r * r);
  a.av[p].w = 1.0f / r;
}

// internalForces (SphKernel.cu:32-91), gravity g rho (SphSystem.cpp:312),
// integrateSPH_LeapFrog (:112-124), box and colliders, interpolation.
__global__ __launch_bounds__(256) void k_mix_wcsph_forces(MixArgs a) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t p = particle_index();
  if (p >= a.n) return;
  const uint32_t mi = a.ms[p];
  const MixSums s = mixed_sums<true, 1>(t, a, p, 0u, mi);
  const float ri = a.rho[p];
  const float cpf = -ri;
  float fS[3];
  mixed_surface(s, t.own[mi].l, t.own[mi].sigma, fS);
  float F0 = (s.fv[0] + fS[0]) + s.fp[0] * cpf;
  float F1 = (s.fv[1] + fS[1]) + s.fp[1] * cpf;
  float F2 = (s.fv[2] + fS[2]) + s.fp[2] * cpf;
  if (!(ri > 0.f)) atomicOr(a.flags, PSB_FLAG_ZERO_DENSITY);
  const Consts& c = a.c;
  F0 = F0 + c.gx * ri; F1 = F1 + c.gy * ri; F2 = F2 + c.gz * ri;
  const float inv = 1.0f / ri;
  float4 vh = a.avh[p];
  const float4 va = vh;
  vh.x = vh.x + (F0 * inv) * c.dt;
  vh.y = vh.y + (F1 * inv) * c.dt;
  vh.z = vh.z + (F2 * inv) * c.dt;
  const float4 x0 = a.ax[p];
  float4 x = x0;
  x.x = x.x + vh.x * c.dt; x.y = x.y + vh.y * c.dt; x.z = x.z + vh.z * c.dt;
  finish_particle(c, p, x, vh, va, x0, a.bx, a.bv, a.bvh);
  a.bid[p] = a.aid[p];
  a.bp[p] = a.pres[p];
}

// compute_Density_PCI on the start positions (PciSphSystem_Kernel.cu:70-85);
// the pressure restarts from zero (:131).
__global__ __launch_bounds__(256) void k_mix_pci_density0(MixArgs a) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t g = particle_index();
  const uint32_t p = g / kMixLanes, lane = g % kMixLanes;
  if (p >= a.n) return;   // a particle's four lanes leave together
  const float r = mixed_density<kMixLanes>(t, a, p, lane, a.ms[p], a.ax);
  if (lane != 0u) return;
  a.rho[p] = r;
  a.pres[p] = 0.f;
  a.av[p].w = 1.0f / r;
}

// compute_Viscosity_SurfaceTension_Force_PCI (PciSphSystem_Kernel.cu:117-170:
// sigma 0.0728, F_acc = (0, -9.81 rho, 0)) and the first prediction
// integrateSPH_LeapFrog_Forces with f_P = 0 (SphKernel.cu:95-109).
__global__ __launch_bounds__(256) void k_mix_pci_nonpressure(MixArgs a, float4* __restrict__ xp) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t g = particle_index();
  const uint32_t p = g / kMixLanes, lane = g % kMixLanes;
  if (p >= a.n) return;   // a particle's four lanes leave together
  const uint32_t mi = a.ms[p];
  const MixSums s = mixed_sums<false, kMixLanes>(t, a, p, lane, mi);
  if (lane != 0u) return;
  float fS[3];
  mixed_surface(s, t.own[mi].l, 0.0728f, fS);
  const float ri = a.rho[p];
  const float fa0 = 0.f + s.fv[0], fa1 = (-9.81f * ri) + s.fv[1], fa2 = 0.f + s.fv[2];
  a.fa[p] = make_float4(fa0, fa1, fa2, 0.f);
  a.fs[p] = make_float4(fS[0], fS[1], fS[2], 0.f);
  const float F0 = (fa0 + 0.f) + fS[0], F1 = (fa1 + 0.f) + fS[1], F2 = (fa2 + 0.f) + fS[2];
  const float inv = 1.0f / ri;
  const float dt = a.c.dt;
  const float4 x0 = a.ax[p], vh0 = a.avh[p];
  const float v0 = vh0.x + (F0 * inv) * dt, v1 = vh0.y + (F1 * inv) * dt,
              v2 = vh0.z + (F2 * inv) * dt;
  xp[p] = make_float4(x0.x + v0 * dt, x0.y + v1 * dt, x0.z + v2 * dt, 0.f);
}

// compute_Density_PCI + compute_Pressure_PCI (PciSphSystem_Kernel.cu:36-85)
// on predicted positions.
__global__ __launch_bounds__(256) void k_mix_pci_density(MixArgs a, float4* __restrict__ X,
                                                         int it) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t g = particle_index();
  const uint32_t p = g / kMixLanes, lane = g % kMixLanes;
  const bool valid = p < a.n;
  const uint32_t pp = valid ? p : 0u;
  const uint32_t mi = a.ms[pp];
  const float4 xi = X[pp];
  float S = 0.f, W1x = 0.f, W1y = 0.f, W1z = 0.f, W2 = 0.f;
  const uint32_t c = valid ? a.cnt[p] : 0u;
  for (uint32_t k = lane; k < c; k += kMixLanes) {
    const uint32_t j = a.list[(uint64_t)k * a.n + p];
    const float4 xj = X[j];
    const MatPair& P = pair_of(t, mi, a.ms[j]);
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    const float w = P.h2 - d2;
    S = __builtin_fmaf((w * w) * w, P.mc, S);
    if (d2 > 0.f) {
      float dd, invd;
      dist_and_inverse(d2, dd, invd);
      const float hd = P.h - dd;
      const float gw = P.cs * (hd * hd);
      const float ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
      W1x = W1x + ux; W1y = W1y + uy; W1z = W1z + uz;
      W2 = W2 + dot3(ux, uy, uz, ux, uy, uz);
    }
  }
  S = mix_total(S);
  W1x = mix_total(W1x); W1y = mix_total(W1y); W1z = mix_total(W1z);
  W2 = mix_total(W2);
  uint32_t err = 0u;
  if (valid && lane == 0u) {
    const MatOwn& o = t.own[mi];
    const float r = S;
    const float e = r - o.rho0;
    float pr = a.pres[p];
    if (r > o.rho0 && W2 > 0.f)
      pr = pr + e / (o.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2));
    a.pres[p] = pr;
    a.rho_pred[p] = r;
    X[p].w = pr / (r * r);
    err = __float_as_uint((e > 0.f ? e : 0.f) * o.inv_rho0);
  }
  // one atomic per wave, spread over kErrSlots words (4 lanes per particle
  // make 4x the waves; on one word their atomics serialised)
  err = wave_max_u(err);
  if ((threadIdx.x & 63u) == 0u && err)
    atomicMax(a.es.base ? &a.es.base[(size_t)(blockIdx.x % kErrSlots) * a.es.stride + it]
                        : &a.st->err[it],
              err);
}

__global__ __launch_bounds__(256) void k_mix_pci_force(MixArgs a, const float4* __restrict__ X,
                                                       float4* __restrict__ Xn, int fin,
                                                       int iters) {
  __shared__ MatTable t;
  load_table(a.tbl, t);
  const uint32_t g = particle_index();
  const uint32_t p = g / kMixLanes, lane = g % kMixLanes;
  if (fin && g == 0) {
    PciStatus* st = a.st;
    uint32_t e = st->err[iters - 1];
    if (a.es.base) {
      for (int k = 0; k < kErrSlots; ++k)
        e = max(e, a.es.base[(size_t)k * a.es.stride + iters - 1]);
      st->err[iters - 1] = e;
    }
    st->last_iters = iters;
    st->last_err = e;
    st->total_iters += (unsigned long long)iters;
    if (e > st->worst_err) st->worst_err = e;
  }
  // the four lanes of a particle share one wave (n * 4 threads, groups of
  // four aligned), so a whole group leaves here together before the shuffles
  if (p >= a.n) return;
  const uint32_t mi = a.ms[p];
  const float4 xi = X[p];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  const uint32_t c = a.cnt[p];
  for (uint32_t k = lane; k < c; k += kMixLanes) {
    const uint32_t j = a.list[(uint64_t)k * a.n + p];
    const float4 xj = X[j];
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (!(d2 > 0.f)) continue;
    const MatPair& P = pair_of(t, mi, a.ms[j]);
    float dd, invd;
    dist_and_inverse(d2, dd, invd);
    const float hd = P.h - dd;
    const float cc = (P.csm * (hd * hd)) * (xi.w + xj.w);
    sx = __builtin_fmaf(dx * invd, cc, sx);
    sy = __builtin_fmaf(dy * invd, cc, sy);
    sz = __builtin_fmaf(dz * invd, cc, sz);
  }
  sx = mix_total(sx);
  sy = mix_total(sy);
  sz = mix_total(sz);
  if (lane != 0u) return;
  const float rp = a.rho_pred[p];
  const float cpf = -rp;
  const float4 f = a.fa[p], fs = a.fs[p];
  const float F0 = (f.x + sx * cpf) + fs.x, F1 = (f.y + sy * cpf) + fs.y,
              F2 = (f.z + sz * cpf) + fs.z;
  const float inv = 1.0f / rp;
  const Consts& cst = a.c;
  const float4 x0 = a.ax[p], vh0 = a.avh[p];
  float4 vh = make_float4(vh0.x + (F0 * inv) * cst.dt, vh0.y + (F1 * inv) * cst.dt,
                          vh0.z + (F2 * inv) * cst.dt, 0.f);
  float4 x = make_float4(x0.x + vh.x * cst.dt, x0.y + vh.y * cst.dt, x0.z + vh.z * cst.dt, 0.f);
  if (!fin) {
    Xn[p] = x;
    return;
  }
  finish_particle(cst, p, x, vh, vh0, x0, a.bx, a.bv, a.bvh);
  a.bid[p] = a.aid[p];
  a.bp[p] = a.pres[p];
}

// ---- host ----------------------------------------------------------------
psb_status make_mat_table(const psb_sph_params& sys, const psb_sph_params* mats, uint32_t count,
                          MatTable* out) {
  PSB_ARG(mats && count >= 1 && count <= (uint32_t)kMaxMaterials, "1 to PSB_MAX_MATERIALS materials");
  *out = MatTable{};
  const double pi = 3.14159265358979323846;
  const double dt = sys.dt;
  for (uint32_t a = 0; a < count; ++a) {
    const psb_sph_params& q = mats[a];
    PSB_ARG(q.support_radius > 0.f && q.support_radius <= sys.support_radius,
            "material support_radius must be in (0, the system's support_radius]");
    PSB_ARG(q.mass > 0.f && q.rest_density > 0.f, "material mass and rest_density must be > 0");
    const double h = q.support_radius, m = q.mass, rho0 = q.rest_density;
    MatOwn& o = out->own[a];
    o.h2 = (float)(h * h);
    o.rho0 = (float)rho0;
    o.inv_rho0 = (float)(1.0 / rho0);
    o.k = q.gas_stiffness;
    o.l = q.threshold;
    o.sigma = q.surface_tension;
    o.beta = (float)(2.0 * std::pow(m * dt / rho0, 2));
  }
  for (uint32_t a = 0; a < count; ++a)
    for (uint32_t b = 0; b < count; ++b) {
      const psb_sph_params &qa = mats[a], &qb = mats[b];
      const double hk = sys.solver == PSB_SOLVER_PCISPH
                            ? (double)std::max(qa.support_radius, qb.support_radius)
                            : (double)qb.support_radius;
      const double mb = qb.mass;
      MatPair& P = out->pair[a * kMaxMaterials + b];
      P.h = (float)hk;
      P.h2 = (float)(hk * hk);
      P.h2x3 = (float)(3.0 * hk * hk);
      P.mc = (float)(mb * 315.0 / (64.0 * pi * std::pow(hk, 9)));
      P.cs = (float)(-45.0 / (pi * std::pow(hk, 6)));
      P.csm = (float)(mb * (-45.0 / (pi * std::pow(hk, 6))));
      P.cvm = (float)(mb * 45.0 / (pi * std::pow(hk, 6)));
      P.csf = (float)(mb * (-945.0 / (32.0 * pi * std::pow(hk, 9))));
      P.mu = (float)(((double)qa.viscosity + (double)qb.viscosity) / 2.0);
    }
  return PSB_OK;
}

psb_status launch_mix_prepare(const MixArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_mix_prepare, dim3(blocks_for((uint64_t)std::max(a.n, 1u) * kMixLanes)),
                     dim3(kBlock), 0, s, a);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_mix_wcsph(const MixArgs& a, hipStream_t s) {
  if (!a.n) return PSB_OK;
  hipLaunchKernelGGL(k_mix_wcsph_density, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a);
  PSB_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mix_wcsph_forces, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_mix_pci_prologue(const MixArgs& a, float4* xp, hipStream_t s) {
  if (!a.n) return PSB_OK;
  hipLaunchKernelGGL(k_mix_pci_density0, dim3(blocks_for((uint64_t)a.n * kMixLanes)), dim3(kBlock),
                     0, s, a);
  PSB_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mix_pci_nonpressure, dim3(blocks_for((uint64_t)a.n * kMixLanes)),
                     dim3(kBlock), 0, s, a, xp);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_mix_pci_density(const MixArgs& a, float4* X, int it, hipStream_t s) {
  PSB_ARG(it >= 0 && it < kMaxIterSlots, "iteration out of range");
  if (!a.n) return PSB_OK;
  hipLaunchKernelGGL(k_mix_pci_density, dim3(blocks_for((uint64_t)a.n * kMixLanes)), dim3(kBlock),
                     0, s, a, X, it);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_mix_pci_force(const MixArgs& a, const float4* X, float4* Xn, bool fin,
                                int iters, hipStream_t s) {
  PSB_ARG(iters >= 1 && iters <= kMaxIterSlots, "iteration count out of range");
  hipLaunchKernelGGL(k_mix_pci_force, dim3(blocks_for((uint64_t)std::max(a.n, 1u) * kMixLanes)),
                     dim3(kBlock), 0, s, a, X, Xn, fin ? 1 : 0, iters);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

}  // namespace psb
