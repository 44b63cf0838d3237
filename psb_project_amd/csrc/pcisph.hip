// pcisph.hip -- PCISPH predict-correct kernels for gfx950.
//
// Reference: pci_SPH_update (PciSphSystem.cu:8-68) and its kernels
// compute_Density_PCI / compute_Pressure_PCI / compute_Pressure_Force_PCI /
// compute_Viscosity_SurfaceTension_Force_PCI / restore_Pos_Vel_PCI
// (PciSphSystem_Kernel.cu:36-202) and integrateSPH_LeapFrog_Forces
// (SphKernel.cu:95-109).  The reference runs, per iteration, five kernels
// separated by host synchronisations; here one iteration is two kernels:
//
//   D(k): predicted density at X (the predicted positions), pressure update,
//         per-particle compression error max(0, rho - rho0)/rho0 reduced to one
//         atomicMax per workgroup into err[k];
//   F(k): pressure force at X, then either the next prediction (written to
//         the other X buffer: the "restore" of the reference is implicit, the
//         prediction always starts from the saved start state) or, on the
//         final iteration, the leapfrog update + box + velocity interpolation.
//
// F(k) decides on the device whether iteration k was the last one
// (iterations >= min and err[k] <= tolerance, or iterations == max) and
// records it in PciStatus.final_iter; every later D/F launch of the same step
// returns immediately, so the host can queue iterations ahead without
// waiting for each error.
#include <cstring>
#include <mutex>
#include <string>

#include "psb_internal.h"
#include "pcisph.h"
#include "sph_kernels.h"

namespace psb {

// one block per scene; block 0 also clears the batch's finished-scene count
__global__ void k_pci_reset(PciStatus* st0, uint32_t* done, ErrSlots es) {
  pci_reset_scene(st0, blockIdx.x, done, es);
}

__global__ void k_pci_retry_reset(PciStatus* st0, ErrSlots es, const uint32_t* scenes) {
  pci_reset_scene(st0, scenes ? scenes[blockIdx.x] : blockIdx.x, nullptr, es, true);
}

__global__ void k_pci_stats_reset(PciStatus* st0) {
  PciStatus* st = st0 + blockIdx.x;
  if (threadIdx.x == 0) {
    st->total_iters = 0;
    st->worst_err = 0u;
    st->last_iters = 0;
    st->last_err = 0u;
    st->final_iter = -1;
  }
}

// The config-mode pair sums of the non-pressure pass in one walk over
// particle p's list: the viscosity (and surface) sums of internal_sums, and
// the pressure scaling delta_i -- the template delta, capped by the
// per-particle value 1/(beta (W1.W1 + W2)) of compute_Pressure_PCI
// (PciSphSystem_Kernel.cu:44-63) evaluated once at the start positions
// (rigid-sphere boundary particles, static with v.w = 0, add to W1 only).
// FEAT: also the warm-start features 1..7 of p (velocity divergence,
// neighbour count, previous pressure, |v|, previous pressure gradient);
// xs[q].w then holds the previous-step pressure, vs[q].w 1/rho.  Each sum
// keeps its own sequential order over the list.
template <bool SURF, bool FEAT, class LI>
__device__ __forceinline__ PairSums np_sums(const Consts& c, uint32_t p, uint32_t n, uint32_t m,
                                            const LI& list, const float4* __restrict__ xs,
                                            const float4* __restrict__ vs,
                                            float* __restrict__ feat, float& dl) {
  PairSums s;
  s.fp[0] = s.fp[1] = s.fp[2] = 0.f;
  s.fv[0] = s.fv[1] = s.fv[2] = 0.f;
  s.fs = 0.f;
  s.nn[0] = s.nn[1] = s.nn[2] = 0.f;
  const float4 xi = xs[p], vi = vs[p];
  const bool local = FEAT || c.delta_local_cap;
  const bool bodies = c.n_fluid != kNoBodies;
  float W1x = 0.f, W1y = 0.f, W1z = 0.f, W2 = 0.f;
  float div = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  for_pairs2<kUnroll>(p, n, m, list, xs, vs, [&](const float4& xj, const float4& vj) {
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (d2 > 0.f) {
      float dd, invd = 0.f;
      if (local) dist_and_inverse(d2, dd, invd);
      else dd = sqrt_rn(d2);
      const float hd = c.h - dd;
      const float rinvj = vj.w;
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
      if (local && d2 < c.h2) {
        const float gw = c.c_spiky * (hd * hd);
        const float ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
        W1x = W1x + ux; W1y = W1y + uy; W1z = W1z + uz;
        const bool moves = (FEAT || bodies) ? vj.w > 0.f : true;
        if (moves) W2 = W2 + dot3(ux, uy, uz, ux, uy, uz);
        if (FEAT) {
          const float vol = c.m * vj.w;
          const float t = dot3(vj.x - vi.x, vj.y - vi.y, vj.z - vi.z, ux, uy, uz);
          div = __builtin_fmaf(t, vol, div);
          const float dp = (xj.w - xi.w) * vol;
          gx = __builtin_fmaf(ux, dp, gx);
          gy = __builtin_fmaf(uy, dp, gy);
          gz = __builtin_fmaf(uz, dp, gz);
        }
      }
    }
  });
  if (FEAT) {
    feat[1 * (size_t)n + p] = div * c.dt;
    feat[2 * (size_t)n + p] = (float)m * 0.03125f;
    feat[3 * (size_t)n + p] = xi.w * c.inv_ps;
    feat[4 * (size_t)n + p] = sqrtf(dot3(vi.x, vi.y, vi.z, vi.x, vi.y, vi.z)) * c.dt_over_h;
    feat[5 * (size_t)n + p] = gx * c.h_over_ps;
    feat[6 * (size_t)n + p] = gy * c.h_over_ps;
    feat[7 * (size_t)n + p] = gz * c.h_over_ps;
  }
  dl = c.delta;
  if (!c.delta_local_cap) return s;
  if (c.wall) {
    // the static wall particles add to sum gradW, not to sum |gradW|^2
    float S, wgx, wgy, wgz;
    wall_sums<false, true>(c, xi.x, xi.y, xi.z, S, wgx, wgy, wgz);
    W1x = W1x + wgx; W1y = W1y + wgy; W1z = W1z + wgz;
  }
  const float den = c.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2);
  if (den > 0.f) {
    const float loc = 1.0f / den;
    if (loc < c.delta) dl = loc;
  }
  return s;
}

// compute_Viscosity_SurfaceTension_Force_PCI (PciSphSystem_Kernel.cu:117-170)
// + first prediction (integrateSPH_LeapFrog_Forces with f_P = 0).
template <bool REF, bool SURF, bool WARM, bool OUT, bool FEAT, class M>
__global__ __launch_bounds__(M::kMaxThreads) void k_pci_nonpressure(
    M map, uint32_t n, Consts c, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ list, const float4* __restrict__ xs,
    const float4* __restrict__ vs, const float* __restrict__ rho, const float4* __restrict__ avh,
    const float* __restrict__ pres, PciBuffers b, PciOut o) {
  uint32_t p, sc;
  bool lead;
  if (!map.locate(n, p, sc, lead)) return;
  map.patch(c, sc);
  if (!REF && is_boundary(c, b.aid, p)) {
    // rigid-sphere boundary particle: static during the corrector
    const float4 x = xs[p];
    b.anp[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    b.dl[p] = 0.f;
    b.xp0[p] = make_float4(x.x, x.y, x.z, 0.f);
    if (FEAT)
      for (int k = 1; k < 8; ++k) b.feat[(size_t)k * n + p] = 0.f;
    return;
  }
  // the compact list unless this scene's offsets overflowed it (or, once
  // they have, the search writes the grouped 32-bit list instead)
  const PciStatus* stc = b.st ? b.st + map.scene_of_block() : nullptr;
  const bool c16 = b.list16 && stc && stc->wide == 0u && !(M::kWideList && c.list32g && stc->want32);
  const uint32_t m = cnt[p];
  PairSums s;
  float dl = c.delta;
  if (REF) {
    s = c16 ? internal_sums_l<false, SURF>(c, p, n, m, List16{b.list16, n}, xs, vs)
            : internal_sums_l<false, SURF>(c, p, n, m, List32{list, n}, xs, vs);
  } else {
    s = c16 ? np_sums<SURF, FEAT>(c, p, n, m, List16{b.list16, n}, xs, vs, b.feat, dl)
            : np_sums<SURF, FEAT>(c, p, n, m, List32{list, n}, xs, vs, b.feat, dl);
  }
  const float ri = rho[p];
  const float sigma = REF ? 0.0728f : c.sigma;
  float fV[3] = {s.fv[0] * c.cv, s.fv[1] * c.cv, s.fv[2] * c.cv};
  float fS[3] = {0.f, 0.f, 0.f}, nrm[3] = {0.f, 0.f, 0.f};
  if (SURF) surface_finish(c, s, sigma, fS, nrm);
  const float4 x = xs[p], vh0 = avh[p];
  float4 xp;
  if (REF) {
    // forcesAccum = (0, -9.81 rho, 0) (PciSphSystem_Kernel.cu:168);
    // predictor F = ((Facc + fV) + fP) + fS with fP = 0 (:127, SphKernel.cu:103)
    const float gy = -9.81f;
    const float fa0 = 0.f + fV[0], fa1 = (gy * ri) + fV[1], fa2 = 0.f + fV[2];
    b.fa[p] = make_float4(fa0, fa1, fa2, 0.f);
    b.fs[p] = make_float4(fS[0], fS[1], fS[2], 0.f);
    const float F0 = (fa0 + 0.f) + fS[0], F1 = (fa1 + 0.f) + fS[1], F2 = (fa2 + 0.f) + fS[2];
    const float inv = 1.0f / ri;
    const float v0 = vh0.x + (F0 * inv) * c.dt, v1 = vh0.y + (F1 * inv) * c.dt,
                v2 = vh0.z + (F2 * inv) * c.dt;
    xp = make_float4(x.x + v0 * c.dt, x.y + v1 * c.dt, x.z + v2 * c.dt, 0.f);
    if (OUT) {
      if (o.F) o.F[p] = make_float4(0.f, gy * ri, 0.f, 0.f);
    }
  } else {
    b.dl[p] = dl;
    const float inv = 1.0f / ri;
    const float a0 = (fV[0] + fS[0]) * inv + c.gx, a1 = (fV[1] + fS[1]) * inv + c.gy,
                a2 = (fV[2] + fS[2]) * inv + c.gz;
    b.anp[p] = make_float4(a0, a1, a2, 0.f);
    if (WARM) {
      xp = make_float4(x.x, x.y, x.z, pres[p]);
    } else {
      const float v0 = vh0.x + a0 * c.dt, v1 = vh0.y + a1 * c.dt, v2 = vh0.z + a2 * c.dt;
      xp = make_float4(x.x + v0 * c.dt, x.y + v1 * c.dt, x.z + v2 * c.dt, 0.f);
    }
    if (OUT) {
      if (o.F) o.F[p] = make_float4(a0, a1, a2, 0.f);
    }
  }
  b.xp0[p] = xp;
  if (OUT) {
    if (o.normals) o.normals[p] = make_float4(nrm[0], nrm[1], nrm[2], 0.f);
    if (o.fV) o.fV[p] = make_float4(fV[0], fV[1], fV[2], 0.f);
    if (o.fS) o.fS[p] = make_float4(fS[0], fS[1], fS[2], 0.f);
  }
}

// compute_Density_PCI + compute_Pressure_PCI (PciSphSystem_Kernel.cu:36-85)
// in reference mode; template-delta update p = max(0, p + delta (rho - rho0))
// in config mode.  X[p].w <- p/rho^2 (reference) or p (config).
// Copies a TileMap tile's halo (X of every staged slot) into LDS; returns the
// staged count, 0 when the tile keeps the global path.  Every thread of the
// workgroup calls it (it ends with a barrier).
template <class M>
__device__ __forceinline__ uint32_t stage_halo(const M& map, const float4* __restrict__ X,
                                               float4* hx) {
  const uint32_t t = map.tile();
  const HaloTile& T = map.tiles[t];
  const uint32_t tot = T.ok ? (uint32_t)T.total : 0u;
  const uint32_t* __restrict__ sr = map.src + (size_t)t * kHaloCap;
  for (uint32_t k = threadIdx.x; k < tot; k += blockDim.x) hx[k] = X[sr[k]];
  __syncthreads();
  return tot;
}

template <bool REF, class M>
__global__ __launch_bounds__(M::kMaxThreads) void k_pci_density(
    M map, uint32_t n, Consts c, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ list, float4* X, float* __restrict__ pres,
    float* __restrict__ rho_pred, const float* __restrict__ dl, PciStatus* __restrict__ st0, int k,
    float* __restrict__ feat, int dflags, const uint32_t* __restrict__ ids,
    const uint16_t* __restrict__ list16, ErrSlots es) {
  __shared__ float red[M::kMaxThreads / 64];
  __shared__ float4 hx[M::kHalo ? kHaloCap : 1];
  PciStatus* __restrict__ st = st0 + map.scene_of_block();   // a block serves one scene
  {
    const int fi = st->final_iter;
    if (fi >= 0 && fi < k) return;
  }
  uint32_t staged = 0;   // staged particle count (TileMap tiles whose halo fits)
  if constexpr (M::kHalo) staged = stage_halo(map, X, hx);
  uint32_t p, sc;
  bool lead;
  const bool valid = map.locate(n, p, sc, lead);
  map.patch(c, sc);
  float err = 0.f;
  constexpr int L = M::kLanes;
  const uint32_t lane = L == 1 ? 0u : (threadIdx.x & 1u);
  const bool writer = lane == 0;   // the pair's other lane only sums
  const bool bnd = !REF && valid && is_boundary(c, ids, p);
  if (bnd && writer) {
    // rigid-sphere boundary particle: no pressure, no error
    reinterpret_cast<float*>(X)[4 * (size_t)p + 3] = 0.f;
    pres[p] = 0.f;
    rho_pred[p] = c.rho0;
    if (dflags & kDensWriteFeature) feat[p] = 0.f;
  }
  if (valid && !bnd) {
    const float4 xi = X[p];
    float S2[2] = {0.f, 0.f}, W1x2[2] = {0.f, 0.f}, W1y2[2] = {0.f, 0.f}, W1z2[2] = {0.f, 0.f};
    float W22[2] = {0.f, 0.f};
    auto pair = [&](int a, const float4& xj) {
      const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const float d2 = dot3(dx, dy, dz, dx, dy, dz);
      float w = c.h2 - d2;
      if (!REF) w = fmaxf(w, 0.f);
      S2[a] = __builtin_fmaf(w * w, w, S2[a]);
      if (REF && d2 > 0.f) {
        float dd, invd;
        dist_and_inverse(d2, dd, invd);
        const float hd = c.h - dd;
        const float gw = c.c_spiky * (hd * hd);
        const float ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
        W1x2[a] = W1x2[a] + ux; W1y2[a] = W1y2[a] + uy; W1z2[a] = W1z2[a] + uz;
        W22[a] = W22[a] + dot3(ux, uy, uz, ux, uy, uz);
      }
    };
    if (staged)
      for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, ListSlot{list16, n, staged - 1}, hx, pair);
    else if (M::kWideList && list16 && c.list32g && st->want32)
      for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List32G{c.list32g, n}, X, pair);
    else if (list16 && st->wide == 0u)
      for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List16{list16, n}, X, pair);
    else
      for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List32{list, n}, X, pair);
    const float S = split_total<L>(S2, lane);
    float W1x = 0.f, W1y = 0.f, W1z = 0.f, W2 = 0.f;
    if (REF) {
      W1x = split_total<L>(W1x2, lane); W1y = split_total<L>(W1y2, lane);
      W1z = split_total<L>(W1z2, lane); W2 = split_total<L>(W22, lane);
    }
    float r = S * c.mc;
    if (!REF && c.wall) {
      float Sw, gx, gy, gz;
      wall_sums<true, false>(c, xi.x, xi.y, xi.z, Sw, gx, gy, gz);
      r = r + Sw * c.mc;
    }
    float pr = pres[p];
    const float e = r - c.rho0;
    if (REF) {
      if (r > c.rho0 && W2 > 0.f) pr = pr + e / (c.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2));
      if (writer) reinterpret_cast<float*>(X)[4 * (size_t)p + 3] = pr / (r * r);
    } else {
      if ((dflags & kDensWriteFeature) && writer) feat[p] = (e * c.inv_rho0) * 100.0f;
      if (!(dflags & kDensSkipUpdate)) {
        pr = pr + dl[p] * e;
        if (pr < 0.f) pr = 0.f;
      }
      if (writer) reinterpret_cast<float*>(X)[4 * (size_t)p + 3] = pr;
    }
    if (writer) {
      rho_pred[p] = r;
      pres[p] = pr;
    }
    err = (e > 0.f ? e : 0.f) * c.inv_rho0;
  }
  err = wave_max<64>(err);
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = err;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
    if (m > 0.f) {
      uint32_t* dst = es.base ? es.base + (size_t)(blockIdx.x % kErrSlots) * es.stride +
                                    (size_t)map.scene_of_block() * kMaxIterSlots + k
                              : &st->err[k];
      atomicMax(dst, __float_as_uint(m));
    }
  }
}

// compute_Pressure_Force_PCI (PciSphSystem_Kernel.cu:88-114) +
// integrateSPH_LeapFrog_Forces from the start state (SphKernel.cu:95-109,
// restore_Pos_Vel_PCI :190-202).  warm_pass: the pass before the first
// density evaluation that turns a warm-start pressure into a force.
template <bool REF, bool STEP, class M>
__global__ __launch_bounds__(M::kMaxThreads) void k_pci_force(
    M map, uint32_t n, Consts c, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ list, const float4* __restrict__ X, float4* __restrict__ Xn,
    const float* __restrict__ rho_pred, const float* __restrict__ pres, PciBuffers b,
    PciStatus* __restrict__ st0, int k, int warm_pass, PciOut o,
    const uint16_t* __restrict__ list16) {
  uint32_t p, sc;
  bool lead;
  const bool valid = map.locate(n, p, sc, lead);
  map.patch(c, sc);
  PciStatus* __restrict__ st = st0 + map.scene_of_block();
  bool fin = false;
  float err = 0.f;
  if (!warm_pass) {
    const int fi = st->final_iter;
    if (fi >= 0 && fi < k) return;
    uint32_t nbmax = 0u;
    if (o.es.base) {
      // one vector load per wave: lanes 0..15 read iteration k's error
      // slots, lanes 16..31 the neighbour-count slots; wave maxima
      const uint32_t ln = threadIdx.x & 63u;
      const size_t e0 = (size_t)map.scene_of_block() * kMaxIterSlots;
      uint32_t v = 0u, vc = 0u;
      if (ln < (uint32_t)kErrSlots) v = o.es.base[(size_t)ln * o.es.stride + e0 + k];
      else if (ln < 2u * kErrSlots)
        vc = o.es.base[(size_t)(ln - kErrSlots) * o.es.stride + e0 + kCountEntry];
      err = __uint_as_float(wave_max_u(v));
      nbmax = wave_max_u(vc);
    } else {
      err = __uint_as_float(st->err[k]);
    }
    const int it = k + 1;
    // smallest error of iterations 0 .. k-1 (+inf at k = 0)
    const float minp = __uint_as_float(st->min_err[(k + 1) & 1]);
    const bool diverged = !REF && c.diverge_stop && it >= c.min_it &&
                          err > kDivergeFloor * c.tolerance && err > kDivergeFactor * minp;
    fin = (it >= c.min_it && (REF || err <= c.tolerance)) || it >= c.max_it || diverged;
    if (lead && !REF) st->min_err[k & 1] = __float_as_uint(fminf(minp, err));
    if (fin && lead) {
      st->final_iter = k;
      st->last_iters = it;
      st->last_err = __float_as_uint(err);
      atomicAdd(&st->total_iters, (unsigned long long)it);
      // an attempt that will be redone as substeps leaves the worst error
      // and the flag to the attempts that are kept
      const bool retried = !REF && err > c.tolerance && c.retry_ok;
      if (!retried) atomicMax(&st->worst_err, __float_as_uint(err));
      // the neighbour search's largest count of this step -> flags[1]
      if (nbmax) atomicMax(o.flags + 1, nbmax);
      if (!REF && err > c.tolerance && !retried) atomicOr(o.flags, PSB_FLAG_NOT_CONVERGED);
      // tell the host, which stops queueing speculative iterations once
      // every scene has finished
      bool all = true;
      if (o.done) all = atomicAdd(o.done, 1u) + 1u == o.num_scenes;
      if (all && o.host_final) {
        const int32_t hv = (!REF && err > c.tolerance) ? (k | kFinalFailed) : k;
        __hip_atomic_store(o.host_final, hv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    if (lead && o.host_progress)
      __hip_atomic_store(o.host_progress, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __shared__ float4 hx[M::kHalo ? kHaloCap : 1];
  uint32_t staged = 0;
  if constexpr (M::kHalo) staged = stage_halo(map, X, hx);
  if (!valid) return;
  const float4 xi = X[p];
  constexpr int L = M::kLanes;
  const uint32_t lane = L == 1 ? 0u : (threadIdx.x & 1u);
  if (!REF && is_boundary(c, b.aid, p)) {
    // rigid-sphere boundary particle: static during the corrector; its
    // end-of-step state is written by the sphere update
    if (!fin && lane == 0) Xn[p] = make_float4(xi.x, xi.y, xi.z, 0.f);
    return;
  }
  float sx2[2] = {0.f, 0.f}, sy2[2] = {0.f, 0.f}, sz2[2] = {0.f, 0.f};
  auto pair = [&](int a, const float4& xj) {
    // the pressure term wj is read with the position (one 16-byte gather)
    float wj = xj.w;
    asm volatile("" : "+v"(wj));
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (REF) {
      if (d2 > 0.f) {
        float dd, invd;
        dist_and_inverse(d2, dd, invd);
        const float hd = c.h - dd;
        const float gw = c.c_spiky * (hd * hd);
        const float cc = gw * (xi.w + wj);
        sx2[a] = __builtin_fmaf(dx * invd, cc, sx2[a]);
        sy2[a] = __builtin_fmaf(dy * invd, cc, sy2[a]);
        sz2[a] = __builtin_fmaf(dz * invd, cc, sz2[a]);
      }
    } else {
      // (h - d)^2 / d (p_i + p_j) with r = 1/d from rsqrt_nr and d = d2 r;
      // no branch: past h the clamp makes (h - d) zero, and the self entry
      // (dx = 0) adds zero.  c_spiky kp multiplies the sum afterwards.
      const float d2c = fmaxf(d2, kMinPairD2);
      const float r = rsqrt_nr(d2c);
      const float hd = fmaxf(__builtin_fmaf(-d2c, r, c.h), 0.f);
      const float cc = ((hd * hd) * r) * (xi.w + wj);
      sx2[a] = __builtin_fmaf(dx, cc, sx2[a]);
      sy2[a] = __builtin_fmaf(dy, cc, sy2[a]);
      sz2[a] = __builtin_fmaf(dz, cc, sz2[a]);
    }
  };
  if (staged)
    for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, ListSlot{list16, n, staged - 1}, hx, pair);
  else if (M::kWideList && list16 && c.list32g && st->want32)
    for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List32G{c.list32g, n}, X, pair);
  else if (list16 && st->wide == 0u)
    for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List16{list16, n}, X, pair);
  else
    for_neighbors_split<L, kUnroll>(p, n, cnt[p], lane, List32{list, n}, X, pair);
  float sx = split_total<L>(sx2, lane), sy = split_total<L>(sy2, lane),
        sz = split_total<L>(sz2, lane);
  if (lane != 0) return;   // the pair's second lane only summed
  const float4 x0 = b.ax[p], vh0 = b.avh[p];
  float4 vh;
  float fP0 = 0.f, fP1 = 0.f, fP2 = 0.f;
  if (REF) {
    const float rp = rho_pred[p];
    const float cpf = -(c.m * rp);
    fP0 = sx * cpf; fP1 = sy * cpf; fP2 = sz * cpf;
    const float4 a = b.fa[p], s = b.fs[p];
    const float F0 = (a.x + fP0) + s.x, F1 = (a.y + fP1) + s.y, F2 = (a.z + fP2) + s.z;
    const float inv = 1.0f / rp;
    vh = make_float4(vh0.x + (F0 * inv) * c.dt, vh0.y + (F1 * inv) * c.dt,
                     vh0.z + (F2 * inv) * c.dt, 0.f);
  } else {
    const float4 a = b.anp[p];
    fP0 = sx * c.kps; fP1 = sy * c.kps; fP2 = sz * c.kps;
    if (c.wall) {
      // wall particles push with the particle's own pressure only (p_b = 0:
      // F = -m psi_b p_i / rho0^2 gradW, psi_b = m on the fluid lattice)
      float S, gx, gy, gz;
      wall_sums<false, true>(c, xi.x, xi.y, xi.z, S, gx, gy, gz);
      const float kpw = c.kp * xi.w;
      fP0 = __builtin_fmaf(gx, kpw, fP0);
      fP1 = __builtin_fmaf(gy, kpw, fP1);
      fP2 = __builtin_fmaf(gz, kpw, fP2);
    }
    float a0 = a.x + fP0, a1 = a.y + fP1, a2 = a.z + fP2;
    vh = make_float4(vh0.x + a0 * c.dt, vh0.y + a1 * c.dt, vh0.z + a2 * c.dt, 0.f);
  }
  float4 x = make_float4(x0.x + vh.x * c.dt, x0.y + vh.y * c.dt, x0.z + vh.z * c.dt, 0.f);
  if (!fin) {
    Xn[p] = x;
    return;
  }
  if (STEP) {
    finish_particle(c, p, x, vh, vh0, x0, o.bx, o.bv, o.bvh);
    o.bid[p] = b.aid[p];
    o.bp[p] = pres[p];
  } else {
    // standalone pci_SPH_update: new state without collision / interpolation
    o.pos[p] = make_float4(x.x, x.y, x.z, 0.f);
    o.old_pos[p] = make_float4(x0.x, x0.y, x0.z, 0.f);
    o.vav[p] = vh0;
    o.vap[p] = vh;
    if (o.fP) o.fP[p] = make_float4(fP0, fP1, fP2, 0.f);
  }
}

// ---- launchers ----------------------------------------------------------------
psb_status launch_pci_retry_reset(PciStatus* st, hipStream_t s, ErrSlots es,
                                  const uint32_t* scenes, uint32_t count) {
  if (!count) return PSB_OK;
  hipLaunchKernelGGL(k_pci_retry_reset, dim3(count), dim3(256), 0, s, st, es, scenes);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_pci_reset(PciStatus* st, bool stats, hipStream_t s, uint32_t scenes,
                            uint32_t* done, ErrSlots es) {
  if (stats) hipLaunchKernelGGL(k_pci_stats_reset, dim3(scenes), dim3(64), 0, s, st);
  hipLaunchKernelGGL(k_pci_reset, dim3(scenes), dim3(256), 0, s, st, done, es);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

template <bool R_, bool S_, bool W_, bool O_, bool F_, class M>
static void np_launch(const M& map, const Shape& sh, hipStream_t s, uint32_t n, const Consts& c,
                      const uint32_t* cnt, const uint32_t* list, const float4* xs,
                      const float4* vs, const float* rho, const float4* avh, const float* pres,
                      const PciBuffers& b, const PciOut& o) {
  hipLaunchKernelGGL((k_pci_nonpressure<R_, S_, W_, O_, F_, M>), sh.grid, sh.block, 0, s, map, n, c,
                     cnt, list, xs, vs, rho, avh, pres, b, o);
}

template <bool R_, bool S_, bool W_, bool O_, class M>
static void np_launch_f(bool feat, const M& map, const Shape& sh, hipStream_t s, uint32_t n,
                        const Consts& c, const uint32_t* cnt, const uint32_t* list,
                        const float4* xs, const float4* vs, const float* rho, const float4* avh,
                        const float* pres, const PciBuffers& b, const PciOut& o) {
  if (feat) np_launch<R_, S_, W_, O_, true>(map, sh, s, n, c, cnt, list, xs, vs, rho, avh, pres, b, o);
  else np_launch<R_, S_, W_, O_, false>(map, sh, s, n, c, cnt, list, xs, vs, rho, avh, pres, b, o);
}

template <class M>
static psb_status nonpressure_impl(const M& map, const Shape& sh, uint32_t n, const Consts& c,
                                   bool surf, bool warm, bool out, const uint32_t* cnt,
                                   const uint32_t* list, const float4* xs, const float4* vs,
                                   const float* rho, const float4* avh, const float* pres,
                                   const PciBuffers& b, const PciOut& o, hipStream_t s) {
  if (!n) return PSB_OK;
  const bool ref = c.mode == PSB_MODE_REFERENCE;
  const bool feat = !ref && b.feat != nullptr;
#define PSB_NP(R_, S_, W_, O_) \
  np_launch_f<R_, S_, W_, O_>(feat, map, sh, s, n, c, cnt, list, xs, vs, rho, avh, pres, b, o)
  if (ref) {
    if (out) PSB_NP(true, true, false, true); else PSB_NP(true, true, false, false);
  } else if (surf) {
    if (warm) { if (out) PSB_NP(false, true, true, true); else PSB_NP(false, true, true, false); }
    else { if (out) PSB_NP(false, true, false, true); else PSB_NP(false, true, false, false); }
  } else {
    if (warm) { if (out) PSB_NP(false, false, true, true); else PSB_NP(false, false, true, false); }
    else { if (out) PSB_NP(false, false, false, true); else PSB_NP(false, false, false, false); }
  }
#undef PSB_NP
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

template <class M>
static psb_status density_impl(const M& map, const Shape& sh, uint32_t n, const Consts& c,
                               const uint32_t* cnt, const uint32_t* list, float4* X, float* pres,
                               float* rho_pred, const float* dl, PciStatus* st, int k,
                               float* feat, int dflags, const uint32_t* ids, hipStream_t s,
                               const uint16_t* l16, ErrSlots es) {
  if (!n) return PSB_OK;
  if (c.mode == PSB_MODE_REFERENCE)
    hipLaunchKernelGGL((k_pci_density<true, M>), sh.grid, sh.block, 0, s, map, n, c, cnt, list, X,
                       pres, rho_pred, dl, st, k, feat, 0, ids, l16, es);
  else
    hipLaunchKernelGGL((k_pci_density<false, M>), sh.grid, sh.block, 0, s, map, n, c, cnt, list, X,
                       pres, rho_pred, dl, st, k, feat, dflags, ids, l16, es);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

template <class M>
static psb_status force_impl(const M& map, const Shape& sh, uint32_t n, const Consts& c, bool step,
                             const uint32_t* cnt, const uint32_t* list, const float4* X,
                             float4* Xn, const float* rho_pred, const float* pres,
                             const PciBuffers& b, PciStatus* st, int k, int warm_pass,
                             const PciOut& o, hipStream_t s, const uint16_t* l16) {
  if (!n) return PSB_OK;
  const bool ref = c.mode == PSB_MODE_REFERENCE;
#define PSB_F(R_, S_)                                                                           \
  hipLaunchKernelGGL((k_pci_force<R_, S_, M>), sh.grid, sh.block, 0, s, map, n, c, cnt, list, X, \
                     Xn, rho_pred, pres, b, st, k, warm_pass, o, l16)
  if (ref) { if (step) PSB_F(true, true); else PSB_F(true, false); }
  else { if (step) PSB_F(false, true); else PSB_F(false, false); }
#undef PSB_F
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

static Shape single_shape(uint32_t n, int lanes = 1) {
  return Shape{dim3(blocks_for((uint64_t)n * lanes)), dim3(kBlock)};
}

static Shape tile_shape(uint32_t n) {
  return Shape{dim3((n + kHaloTile - 1) / kHaloTile), dim3(kHaloTile)};
}

psb_status launch_pci_nonpressure(uint32_t n, const Consts& c, bool surf, bool warm, bool out,
                                  const uint32_t* cnt, const uint32_t* list, const float4* xs,
                                  const float4* vs, const float* rho, const float4* avh,
                                  const float* pres, const PciBuffers& b, const PciOut& o,
                                  hipStream_t s) {
  return nonpressure_impl(SingleMap{}, single_shape(n), n, c, surf, warm, out, cnt, list, xs, vs,
                          rho, avh, pres, b, o, s);
}

psb_status launch_pci_density(uint32_t n, const Consts& c, const uint32_t* cnt,
                              const uint32_t* list, float4* X, float* pres, float* rho_pred,
                              const float* dl, PciStatus* st, int k, float* feat, int dflags,
                              hipStream_t s, const uint32_t* ids, int lanes,
                              const uint16_t* list16, ErrSlots es, const HaloTile* tiles,
                              const uint32_t* halo_src) {
  if (tiles && lanes == 1 && list16)
    return density_impl(TileMap{tiles, halo_src}, tile_shape(n), n, c, cnt, list, X, pres,
                        rho_pred, dl, st, k, feat, dflags, ids, s, list16, es);
  if (lanes == 2)
    return density_impl(SingleMapT<2>{}, single_shape(n, 2), n, c, cnt, list, X, pres, rho_pred,
                        dl, st, k, feat, dflags, ids, s, list16, es);
  return density_impl(SingleMap{}, single_shape(n), n, c, cnt, list, X, pres, rho_pred, dl, st, k,
                      feat, dflags, ids, s, list16, es);
}

psb_status launch_pci_force(uint32_t n, const Consts& c, bool step, const uint32_t* cnt,
                            const uint32_t* list, const float4* X, float4* Xn,
                            const float* rho_pred, const float* pres, const PciBuffers& b,
                            PciStatus* st, int k, int warm_pass, const PciOut& o, hipStream_t s,
                            int lanes, const uint16_t* list16, const HaloTile* tiles,
                            const uint32_t* halo_src) {
  if (tiles && lanes == 1 && list16)
    return force_impl(TileMap{tiles, halo_src}, tile_shape(n), n, c, step, cnt, list, X, Xn,
                      rho_pred, pres, b, st, k, warm_pass, o, s, list16);
  if (lanes == 2)
    return force_impl(SingleMapT<2>{}, single_shape(n, 2), n, c, step, cnt, list, X, Xn, rho_pred,
                      pres, b, st, k, warm_pass, o, s, list16);
  return force_impl(SingleMap{}, single_shape(n), n, c, step, cnt, list, X, Xn, rho_pred, pres, b,
                    st, k, warm_pass, o, s, list16);
}

psb_status launch_pci_nonpressure_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                        const Consts& c, bool surf, bool warm,
                                        const uint32_t* cnt, const uint32_t* list,
                                        const float4* xs, const float4* vs, const float* rho,
                                        const float4* avh, const float* pres,
                                        const PciBuffers& b, const PciOut& o, hipStream_t s) {
  return nonpressure_impl(map, sh, n, c, surf, warm, false, cnt, list, xs, vs, rho, avh, pres, b,
                          o, s);
}

psb_status launch_pci_density_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                    const Consts& c, const uint32_t* cnt, const uint32_t* list,
                                    float4* X, float* pres, float* rho_pred, const float* dl,
                                    PciStatus* st, int k, float* feat, int dflags,
                                    hipStream_t s, int lanes, const uint16_t* list16,
                                    ErrSlots es) {
  if (lanes == 2)
    return density_impl(BatchMapT<2>{map.sc, map.blk_scene, map.bsel}, sh, n, c, cnt, list, X,
                        pres, rho_pred, dl, st, k, feat, dflags, nullptr, s, list16, es);
  return density_impl(map, sh, n, c, cnt, list, X, pres, rho_pred, dl, st, k, feat, dflags,
                      nullptr, s, list16, es);
}

psb_status launch_pci_force_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                  const Consts& c, const uint32_t* cnt, const uint32_t* list,
                                  const float4* X, float4* Xn, const float* rho_pred,
                                  const float* pres, const PciBuffers& b, PciStatus* st, int k,
                                  int warm_pass, const PciOut& o, hipStream_t s, int lanes,
                                  const uint16_t* list16) {
  if (lanes == 2)
    return force_impl(BatchMapT<2>{map.sc, map.blk_scene, map.bsel}, sh, n, c, true, cnt, list, X,
                      Xn, rho_pred, pres, b, st, k, warm_pass, o, s, list16);
  return force_impl(map, sh, n, c, true, cnt, list, X, Xn, rho_pred, pres, b, st, k, warm_pass, o,
                    s, list16);
}

// Append every `stride`-th particle (p = offset, offset + stride, ...) of the
// step: features and target = converged pressure / pressure_scale.
__global__ __launch_bounds__(256) void k_record(const float* __restrict__ feat,
                                                const float* __restrict__ pres, uint32_t n,
                                                uint32_t stride, uint32_t offset, float inv_ps,
                                                float* __restrict__ dfeat,
                                                float* __restrict__ dtarget, uint64_t dcap,
                                                uint64_t base, uint32_t count) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const uint64_t p = (uint64_t)offset + (uint64_t)j * stride;
  const uint64_t d = base + j;
  if (p >= n || d >= dcap) return;
#pragma unroll
  for (int k = 0; k < 8; ++k) dfeat[(uint64_t)k * dcap + d] = feat[(uint64_t)k * n + p];
  dtarget[d] = pres[p] * inv_ps;
}

psb_status launch_record(const float* feat, const float* pres, uint32_t n, uint32_t stride,
                         uint32_t offset, float inv_ps, float* dfeat, float* dtarget,
                         uint64_t dcap, uint64_t base, uint32_t count, hipStream_t s) {
  if (!count) return PSB_OK;
  hipLaunchKernelGGL(k_record, dim3(blocks_for(count)), dim3(kBlock), 0, s, feat, pres, n, stride,
                     offset, inv_ps, dfeat, dtarget, dcap, base, count);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

}  // namespace psb

using namespace psb;

namespace psb {

// ---------------------------------------------------------------------------
// Static scratch of psb_pci_sph_update_noalloc: one array per GPU, held until
// the process ends, sized for PSB_NOALLOC_MAX_PARTICLES.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a host-side
// stand-in); it passes the device address that hipGetSymbolAddress returned
// for the current device at psb_init.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = pci_update_scratch_bytes(PSB_NOALLOC_MAX_PARTICLES);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

psb_status static_scratch_setup(int device) {
  PSB_ARG(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return PSB_OK;
  void* ptr = nullptr;
  PSB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));   // device `device` is current
  g_scratch_ptr[device] = static_cast<char*>(ptr);
  return PSB_OK;
}

static char* static_scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

// The solve of psb_pci_sph_update on a caller-chosen scratch area of at
// least pci_update_scratch_bytes(n) bytes (null: the context's scratch,
// grown on demand).
static psb_status pci_sph_update_on(psb_context ctx, const psb_sph_params* prm, psb_neighbors nb,
                                    float* old_pos4, float* pos4, const float* vel4,
                                    float* vel_inter_av4, float* vel_inter_ap4, float* density,
                                    float* pressure, float* normals4, float* f_viscosity4,
                                    float* f_surface4, float* f_pressure4, float* forces_accum4,
                                    uint32_t n, int32_t* iterations, float* max_error,
                                    char* fixed_scratch) {
  PSB_ARG(ctx && prm, "null argument");
  PSB_ARG(n == 0 || (old_pos4 && pos4 && vel4 && vel_inter_av4 && vel_inter_ap4 && density &&
                     pressure),
          "null array");
  PSB_ARG(nb.n == n && nb.count && nb.list, "neighbour list stride must equal n");
  PSB_ARG(prm->warm_start != PSB_WARM_MLP, "the MLP warm start needs a psb_system");
  Consts c{};
  PSB_TRY(make_consts(*prm, &c, nullptr));
  if (!n) return PSB_OK;
  PSB_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t v4 = 16 * (size_t)n;
  char* base = fixed_scratch;
  if (!base) {
    PSB_TRY(ensure_scratch(ctx, pci_update_scratch_bytes(n)));
    base = (char*)ctx->scratch;
  }
  float4* xs = (float4*)base;
  float4* vs = (float4*)(base + v4);
  PciBuffers b{};
  b.fa = (float4*)(base + 2 * v4);
  b.fs = (float4*)(base + 3 * v4);
  b.anp = (float4*)(base + 4 * v4);
  b.xp0 = (float4*)(base + 5 * v4);
  float4* xp1 = (float4*)(base + 6 * v4);
  float* tmp_p = (float*)(base + 7 * v4);
  b.dl = tmp_p + n;
  PciStatus* st = (PciStatus*)(((uintptr_t)(base + 7 * v4 + 8 * (size_t)n) + 255) & ~(uintptr_t)255);
  b.ax = reinterpret_cast<const float4*>(pos4);
  b.avh = reinterpret_cast<const float4*>(vel_inter_ap4);
  b.aid = nullptr;
  const bool ref = prm->mode == PSB_MODE_REFERENCE;
  const bool warm = !ref && prm->warm_start == PSB_WARM_PREVIOUS;
  PSB_TRY(launch_pci_reset(st, true, s));
  PSB_HIP(hipMemsetAsync(ctx->d_flags, 0, 4, s));
  // start density (PciSphSystem.cu:17), scratch copies with v.w = 1/rho
  PSB_TRY(launch_density_list(n, c, nb, reinterpret_cast<const float4*>(pos4),
                              reinterpret_cast<const float4*>(vel4), density, tmp_p, xs, vs, s));
  if (!warm) PSB_HIP(hipMemsetAsync(pressure, 0, 4 * (size_t)n, s));
  PciOut o{};
  o.normals = reinterpret_cast<float4*>(normals4);
  o.fV = reinterpret_cast<float4*>(f_viscosity4);
  o.fS = reinterpret_cast<float4*>(f_surface4);
  o.F = reinterpret_cast<float4*>(forces_accum4);
  o.fP = reinterpret_cast<float4*>(f_pressure4);
  o.pos = reinterpret_cast<float4*>(pos4);
  o.old_pos = reinterpret_cast<float4*>(old_pos4);
  o.vav = reinterpret_cast<float4*>(vel_inter_av4);
  o.vap = reinterpret_cast<float4*>(vel_inter_ap4);
  o.flags = ctx->d_flags;
  const bool surf = ref || prm->surface_tension != 0.f;
  PSB_TRY(launch_pci_nonpressure(n, c, surf, warm, true, nb.count, nb.list, xs, vs, density,
                                 b.avh, pressure, b, o, s));
  float4* cur = b.xp0;
  float4* nxt = xp1;
  if (warm) {
    PSB_TRY(launch_pci_force(n, c, false, nb.count, nb.list, cur, nxt, density, pressure, b, st, -1,
                             1, o, s));
    std::swap(cur, nxt);
  }
  int iters = 0;
  float err = 0.f;
  for (int k = 0; k < c.max_it; ++k) {
    PSB_TRY(launch_pci_density(n, c, nb.count, nb.list, cur, pressure, density, b.dl, st, k,
                               nullptr, 0, s));
    PSB_TRY(launch_pci_force(n, c, false, nb.count, nb.list, cur, nxt, density, pressure, b, st, k,
                             0, o, s));
    std::swap(cur, nxt);
    if (k + 1 >= c.min_it) {
      PSB_HIP(hipMemcpyAsync(ctx->h_pinned, st, 16, hipMemcpyDeviceToHost, s));
      PSB_HIP(hipStreamSynchronize(s));
      const PciStatus* hs = reinterpret_cast<const PciStatus*>(ctx->h_pinned);
      if (hs->final_iter >= 0) {
        iters = hs->last_iters;
        std::memcpy(&err, &hs->last_err, 4);
        break;
      }
    }
  }
  if (iterations) *iterations = iters;
  if (max_error) *max_error = err;
  return PSB_OK;
}

}  // namespace psb

extern "C" {

psb_status psb_pci_sph_update(psb_context ctx, const psb_sph_params* prm, psb_neighbors nb,
                              float* old_pos4, float* pos4, const float* vel4,
                              float* vel_inter_av4, float* vel_inter_ap4, float* density,
                              float* pressure, float* normals4, float* f_viscosity4,
                              float* f_surface4, float* f_pressure4, float* forces_accum4,
                              uint32_t n, int32_t* iterations, float* max_error) {
  return pci_sph_update_on(ctx, prm, nb, old_pos4, pos4, vel4, vel_inter_av4, vel_inter_ap4,
                           density, pressure, normals4, f_viscosity4, f_surface4, f_pressure4,
                           forces_accum4, n, iterations, max_error, nullptr);
}

psb_status psb_pci_sph_update_noalloc(psb_context ctx, const psb_sph_params* prm,
                                      psb_neighbors nb, float* old_pos4, float* pos4,
                                      const float* vel4, float* vel_inter_av4,
                                      float* vel_inter_ap4, float* density, float* pressure,
                                      float* normals4, float* f_viscosity4, float* f_surface4,
                                      float* f_pressure4, float* forces_accum4, uint32_t n,
                                      int32_t* iterations, float* max_error) {
  PSB_ARG(ctx, "null argument");
  if (n > PSB_NOALLOC_MAX_PARTICLES) {
    set_error("psb_pci_sph_update_noalloc: n = " + std::to_string(n) + " exceeds the static "
              "scratch's maximum of " + std::to_string(PSB_NOALLOC_MAX_PARTICLES) +
              " particles (use psb_pci_sph_update)");
    return PSB_ERR_CAPACITY;
  }
  char* base = static_scratch_ptr(ctx->device);
  if (!base) { set_error("static scratch not set up on this device (psb_init)"); return PSB_ERR_STATE; }
  return pci_sph_update_on(ctx, prm, nb, old_pos4, pos4, vel4, vel_inter_av4, vel_inter_ap4,
                           density, pressure, normals4, f_viscosity4, f_surface4, f_pressure4,
                           forces_accum4, n, iterations, max_error, base);
}

}  // extern "C"
