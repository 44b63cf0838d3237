// system.hip -- fused SPH / PCISPH system step (SPHSystem::update,
// PCI_SPHSystem::update) on one device stream.
//
// State lives in two sorted buffers:
//   A = start-of-step state in this step's sorted order (written by the
//       grid build's reorder); after the step it holds m_dPos[0], m_dVel[0]
//       and m_dVelInterAv of the reference (old position, old velocity and
//       start half-step velocity);
//   B = end-of-step state in the same order (m_dPos[1], m_dVel[1],
//       m_dVelInterAp).  The next grid build hashes B and reorders B -> A.
// The reference's 4 device-to-device "restore" copies per PCISPH step
// (PciSphSystem.cpp:200-203) and its 4-5 device-to-host copies per step
// (SphSystem.cpp:333,341-344) disappear: predictions always start from A,
// and the host reads state only when asked.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bodies.h"
#include "mlp.h"
#include "pcisph.h"
#include "mixed.h"
#include "psb_internal.h"

namespace psb {

enum Phase {
  PH_GRID, PH_NEIGHBORS, PH_EPILOGUE, PH_FORCES, PH_NONPRESSURE, PH_PCI_DENSITY, PH_PCI_FORCE,
  PH_MLP, PH_RECORD, PH_MIX_PROLOGUE, PH_MIX_DENSITY, PH_MIX_FORCE, PH_COUNT
};
// grid_build = hash/count + 3-kernel scan + scatter + stable rank/reorder
static const char* kPhaseNames[PH_COUNT] = {
    "grid_build", "neighbors_density", "density_epilogue", "wcsph_forces", "pci_nonpressure",
    "pci_density", "pci_force", "mlp_infer", "record", "mix_pci_prologue", "mix_pci_density",
    "mix_pci_force"};

psb_status grid_create(psb_context ctx, const psb_grid_desc& d, uint32_t capacity, psb_grid* out);

}  // namespace psb

using namespace psb;

struct psb_dataset_s {
  psb_context ctx = nullptr;
  float* feat = nullptr;    // [8][capacity]
  float* target = nullptr;  // [capacity]
  uint64_t capacity = 0, count = 0;
};

struct psb_system_s {
  psb_context ctx = nullptr;
  psb_sph_params prm{};
  Consts c{};
  GridDev gd{};
  uint32_t cap = 0, n = 0;
  psb_grid grid = nullptr;
  // A / B state
  float4 *ax = nullptr, *av = nullptr, *avh = nullptr, *bx = nullptr, *bv = nullptr, *bvh = nullptr;
  uint32_t *aid = nullptr, *bid = nullptr;
  float *ap = nullptr, *bp = nullptr;
  float *rho = nullptr, *pres = nullptr, *rho_pred = nullptr;
  float4 *fa = nullptr, *fs = nullptr, *anp = nullptr, *xp0 = nullptr, *xp1 = nullptr;
  float* dl = nullptr;            // config-mode PCISPH per-particle delta
  float* feat = nullptr;          // warm-start features [8][cap] (lazily allocated)
  psb_mlp mlp = nullptr;          // network of PSB_WARM_MLP
  psb_dataset rec = nullptr;      // training-set recorder
  uint32_t rec_stride = 1;
  int64_t step_counter = 0;
  uint32_t* ncount = nullptr;
  uint32_t* nlist = nullptr;
  uint16_t* nlist16 = nullptr;     // compact twin of nlist for the corrector passes (PCISPH)
  uint32_t* nlist32g = nullptr;    // grouped 32-bit twin, used where compact offsets overflow
  bool grouped32 = false;          // the last update call ended on the grouped 32-bit list
  bool compact = true;
  // neighbour rows staged in LDS by the corrector passes (single-lane, compact list)
  bool halo = false;
  HaloTile* tiles = nullptr;
  uint32_t* halo_src = nullptr;
  uint32_t* flags = nullptr;      // [0] flags, [1] max neighbour count
  PciStatus* st = nullptr;
  ErrSlots es{};                  // spread per-iteration error maxima (PCISPH)
  int32_t* h_final = nullptr;     // host-mapped final_iter per step slot
  int32_t* d_final = nullptr;
  bool has_state = false;
  bool stepped = false;
  psb_step_stats stats{};
  // timing
  int32_t timing_every = 0;   // record events on every k-th step (0: off)
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  struct Rec { int phase; size_t e0; int step; int iter; };
  std::vector<Rec> recs;
  std::vector<int> step_final;    // host-known final iteration per step of this update
  double t_total[PH_COUNT] = {};
  int64_t t_count[PH_COUNT] = {};
  // rigid DEM spheres (boundary particles after the fluid)
  uint32_t n_fluid = 0;
  uint32_t n_spheres = 0, per_sphere = 0;
  psb_sphere_params sphere{};
  std::vector<float> sphere_c, sphere_v, sphere_off;   // host: centres, velocities, offsets
  float4* d_sc[2] = {nullptr, nullptr};
  float4* d_sv[2] = {nullptr, nullptr};
  float4* d_sf = nullptr;
  float4* d_soff = nullptr;
  float4* d_fb = nullptr;
  int sphere_cur = 0;
  int32_t corrector_lanes = 0;   // threads per particle in the corrector passes (0: by size)
  int32_t cus = 0;               // compute units of the device
  std::vector<ColliderDev> colliders;   // after the box, in order
  ColliderDev* d_coll = nullptr;
  // per-particle materials (reference mode): table, material of each id,
  // material of each sorted particle; nmat = 0: the system's one material
  MatTable* d_tbl = nullptr;
  uint8_t* mat_by_id = nullptr;
  uint8_t* ms = nullptr;
  uint32_t nmat = 0;
};

namespace psb {

static psb_status ev_get(psb_system sys, size_t* idx) {
  if (sys->ev_used + 2 > sys->ev.size()) {
    size_t add = std::max<size_t>(256, sys->ev.size());
    for (size_t i = 0; i < add; ++i) {
      hipEvent_t e;
      PSB_HIP(hipEventCreate(&e));
      sys->ev.push_back(e);
    }
  }
  *idx = sys->ev_used;
  sys->ev_used += 2;
  return PSB_OK;
}

// Wrap a launch sequence in timing events when timing is on.
struct Timed {
  psb_system sys;
  int phase, step, iter;
  size_t e0 = 0;
  bool on;
  Timed(psb_system s, int ph, int st = 0, int it = -1)
      : sys(s), phase(ph), step(st), iter(it),
        on(s->timing_every > 0 && s->step_counter % s->timing_every == 0) {
    if (on && ev_get(sys, &e0) == PSB_OK) hipEventRecord(sys->ev[e0], sys->ctx->stream);
    else on = false;
  }
  ~Timed() {
    if (!on) return;
    hipEventRecord(sys->ev[e0 + 1], sys->ctx->stream);
    sys->recs.push_back({phase, e0, step, iter});
  }
};

static psb_status collect_timing(psb_system sys) {
  if (sys->recs.empty()) { sys->ev_used = 0; return PSB_OK; }
  PSB_HIP(hipStreamSynchronize(sys->ctx->stream));
  for (const auto& r : sys->recs) {
    if (r.iter >= 0 && r.step < (int)sys->step_final.size() && r.iter > sys->step_final[r.step])
      continue;  // an iteration launched ahead and skipped on the device
    float ms = 0.f;
    PSB_HIP(hipEventElapsedTime(&ms, sys->ev[r.e0], sys->ev[r.e0 + 1]));
    sys->t_total[r.phase] += ms;
    sys->t_count[r.phase] += 1;
  }
  sys->recs.clear();
  sys->ev_used = 0;
  return PSB_OK;
}

static psb_status sys_alloc(psb_system s) {
  const size_t N = s->cap ? s->cap : 1;
  auto A = [&](void** p, size_t b) -> psb_status { PSB_HIP(hipMalloc(p, b)); return PSB_OK; };
  PSB_TRY(A((void**)&s->ax, 16 * N)); PSB_TRY(A((void**)&s->av, 16 * N)); PSB_TRY(A((void**)&s->avh, 16 * N));
  PSB_TRY(A((void**)&s->bx, 16 * N)); PSB_TRY(A((void**)&s->bv, 16 * N)); PSB_TRY(A((void**)&s->bvh, 16 * N));
  PSB_TRY(A((void**)&s->aid, 4 * N)); PSB_TRY(A((void**)&s->bid, 4 * N));
  PSB_TRY(A((void**)&s->ap, 4 * N)); PSB_TRY(A((void**)&s->bp, 4 * N));
  PSB_TRY(A((void**)&s->rho, 4 * N)); PSB_TRY(A((void**)&s->pres, 4 * N)); PSB_TRY(A((void**)&s->rho_pred, 4 * N));
  if (s->prm.solver == PSB_SOLVER_PCISPH) {
    if (s->prm.mode == PSB_MODE_REFERENCE) {
      PSB_TRY(A((void**)&s->fa, 16 * N)); PSB_TRY(A((void**)&s->fs, 16 * N));
    } else {
      PSB_TRY(A((void**)&s->anp, 16 * N));
      PSB_TRY(A((void**)&s->dl, 4 * N));
    }
    PSB_TRY(A((void**)&s->xp0, 16 * N)); PSB_TRY(A((void**)&s->xp1, 16 * N));
    PSB_TRY(A((void**)&s->nlist16, l16_bytes(N, (uint32_t)s->prm.max_neighbors)));
    // offsets can overflow the compact list only beyond 32,768 particles
    if (N > (size_t)kCompactReach + 1) {
      PSB_TRY(A((void**)&s->nlist32g, l32g_bytes(N, (uint32_t)s->prm.max_neighbors)));
      s->c.list32g = s->nlist32g;
    }
    s->es.stride = err_slot_stride(1);
    PSB_TRY(A((void**)&s->es.base, 4 * (size_t)kErrSlots * s->es.stride));
  }
  PSB_TRY(A((void**)&s->ncount, 4 * N));
  PSB_TRY(A((void**)&s->nlist, 4 * N * (size_t)s->prm.max_neighbors));
  PSB_TRY(A((void**)&s->flags, 64));
  PSB_TRY(A((void**)&s->st, sizeof(PciStatus)));
  // [slot]: final iteration of an attempt; [64 + slot]: its force-pass progress
  PSB_HIP(hipHostMalloc((void**)&s->h_final, 128 * sizeof(int32_t),
                        hipHostMallocMapped | hipHostMallocCoherent));
  PSB_HIP(hipHostGetDevicePointer((void**)&s->d_final, s->h_final, 0));
  for (int i = 0; i < 128; ++i) s->h_final[i] = -1;
  PSB_HIP(hipMemsetAsync(s->flags, 0, 64, s->ctx->stream));
  PSB_HIP(hipMemsetAsync(s->st, 0, sizeof(PciStatus), s->ctx->stream));
  return PSB_OK;
}

static psb_status grid_phase(psb_system s, int step, const StepReset* reset = nullptr) {
  ReorderArrays ra;
  ra.sx = s->bx; ra.sv = s->bv; ra.svh = s->bvh; ra.sid = s->bid; ra.sp = s->bp;
  ra.dx = s->ax; ra.dv = s->av; ra.dvh = s->avh; ra.did = s->aid; ra.dp = s->ap;
  Timed t(s, PH_GRID, step);
  return grid_build(s->grid, s->bx, s->n, &ra, s->ctx->stream, reset);
}

static psb_status wcsph_step(psb_system s, int step) {
  hipStream_t st = s->ctx->stream;
  PSB_TRY(grid_phase(s, step));
  psb_neighbors nb{s->ncount, s->nlist, (uint32_t)s->prm.max_neighbors, s->n};
  {
    // neighbour search + density + EOS epilogue in one pass
    Timed t(s, PH_NEIGHBORS, step);
    const EpiArgs ep = make_epilogue(s->c, 0, 0, 0, s->pres, nullptr, s->ax, s->av, nullptr);
    PSB_TRY(launch_neighbors(s->grid, s->ax, s->n, s->c.h2, s->gd.R, nb, s->rho, s->c.mc,
                             s->flags, s->flags + 1, st, Compact16{}, ErrSlots{}, nullptr, &ep));
  }
  {
    Timed t(s, PH_FORCES, step);
    PSB_TRY(launch_wcsph_step_forces(s->n, s->c, s->prm.surface_tension != 0.f, s->ncount,
                                     s->nlist, s->ax, s->av, s->rho, s->pres, s->avh, s->aid, s->bx,
                                     s->bv, s->bvh, s->bid, s->bp, s->flags, st));
  }
  return PSB_OK;
}

// ---- per-particle materials (reference mode) -------------------------------
static MixArgs mix_args(psb_system s) {
  MixArgs a{};
  a.n = s->n; a.c = s->c; a.tbl = s->d_tbl; a.cnt = s->ncount; a.list = s->nlist;
  a.mat_by_id = s->mat_by_id; a.ms = s->ms;
  a.ax = s->ax; a.av = s->av; a.avh = s->avh; a.aid = s->aid;
  a.bx = s->bx; a.bv = s->bv; a.bvh = s->bvh; a.bid = s->bid; a.bp = s->bp;
  a.rho = s->rho; a.pres = s->pres; a.rho_pred = s->rho_pred; a.fa = s->fa; a.fs = s->fs;
  a.flags = s->flags; a.st = s->st; a.es = s->es;
  return a;
}

// Lists over the system's radius (the largest material's), then each
// particle's cut to its own material's radius (UniformGridKernel.cu:82).
static psb_status mixed_lists(psb_system s, int step) {
  PSB_TRY(grid_phase(s, step));
  psb_neighbors nb{s->ncount, s->nlist, (uint32_t)s->prm.max_neighbors, s->n};
  Timed t(s, PH_NEIGHBORS, step);
  PSB_TRY(launch_neighbors(s->grid, s->ax, s->n, s->c.h2, s->gd.R, nb, s->rho, s->c.mc,
                           s->flags, s->flags + 1, s->ctx->stream));
  return launch_mix_prepare(mix_args(s), s->ctx->stream);
}

// SPHSystem::update with per-particle materials (SphSystem.cpp:273-297).
static psb_status mixed_wcsph_step(psb_system s, int step) {
  PSB_TRY(mixed_lists(s, step));
  Timed t(s, PH_FORCES, step);
  return launch_mix_wcsph(mix_args(s), s->ctx->stream);
}

// PCI_SPHSystem::update with per-particle materials: pci_SPH_update's three
// predict-correct iterations (PciSphSystem.cu:8-68), collision, interpolation.
static psb_status mixed_pcisph_step(psb_system s, int step) {
  PSB_TRY(mixed_lists(s, step));
  const MixArgs a = mix_args(s);
  hipStream_t st = s->ctx->stream;
  {
    Timed t(s, PH_MIX_PROLOGUE, step);
    PSB_TRY(launch_mix_pci_prologue(a, s->xp0, st));
  }
  constexpr int kIters = 3;
  float4* cur = s->xp0;
  float4* nxt = s->xp1;
  for (int it = 0; it < kIters; ++it) {
    {
      Timed t(s, PH_MIX_DENSITY, step);
      PSB_TRY(launch_mix_pci_density(a, cur, it, st));
    }
    Timed t(s, PH_MIX_FORCE, step);
    PSB_TRY(launch_mix_pci_force(a, cur, nxt, it == kIters - 1, kIters, st));
    std::swap(cur, nxt);
  }
  return PSB_OK;
}

// Threads per particle in the density (D) and force (F) passes, as set, or
// by default two while one lane per particle would leave at most 4
// wavefronts per SIMD (n <= 1024 x CUs) or once the system reads the grouped
// 32-bit list (its last update call ended with it), else one.  Measured on
// MI355X, event-timed: 262,144 particles, D 16.9 -> 16.4 us and F 19.3 ->
// 18.2 us with two.  2.1 M (DEM scene) on the ungrouped 32-bit list two
// lanes won in D early (184 -> 172 us) but lost later in the run (46.6 ->
// 61.9 us after 400 steps) and in F (213 -> 220 us); on the grouped list
// they win throughout: steps 305-405, D 195.5 -> 160.7 us, F 200.4 -> 179.3
// us, 4.89 -> 4.40 ms per step (early steps 1.834 against 1.848 ms).  LDS
// staging runs with one lane.
struct Lanes {
  int d, f;
};
static Lanes corrector_lanes(psb_system s) {
  if (s->corrector_lanes) {
    const int l = s->corrector_lanes == 2 ? 2 : 1;
    return Lanes{l, l};
  }
  if (s->halo) return Lanes{1, 1};
  const bool small = s->cus > 0 && (uint64_t)s->n <= 1024ull * (uint64_t)s->cus;
  const int l = small || s->grouped32 ? 2 : 1;
  return Lanes{l, l};
}

// Constants of a retry substep at level L (dt / 2^L): everything make_consts
// derives from dt (beta, the template delta, the warm-start feature scales)
// recomputed, delta relaxed by 2^-L, the system's patched fields (bodies,
// colliders) kept.  The deepest level may iterate up to kRetryIterFactor x
// max_iterations: its relaxed corrector converges more slowly.
static psb_status level_consts(psb_system s, int level, Consts* out) {
  if (level == 0) { *out = s->c; return PSB_OK; }
  psb_sph_params p = s->prm;
  const float scale = std::ldexp(1.0f, -level);
  p.dt = s->prm.dt * scale;
  if (p.delta > 0.f) p.delta = s->prm.delta / (scale * scale);
  // the corrector under-relaxed with the step: a step that failed over-
  // corrects as readily at any dt (the pressure's displacement scales with
  // delta dt^2), so each level halves delta on top of params.relaxation
  p.relaxation = (s->prm.relaxation > 0.f ? s->prm.relaxation : 1.0f) * scale;
  PSB_TRY(make_consts(p, out, nullptr));
  out->n_fluid = s->c.n_fluid;
  out->coll = s->c.coll;
  out->ncoll = s->c.ncoll;
  out->list32g = s->c.list32g;
  out->retry_ok = level < s->prm.retry_levels ? 1 : 0;
  out->diverge_stop = 1;
  if (level == s->prm.retry_levels) out->max_it = retry_max_iterations(s->prm);
  return PSB_OK;
}

// Until the force pass of iteration `it` of the attempt in `slot` has made
// its convergence decision (its first workgroup stores `it` into host-mapped
// memory, after the final iteration when there is one) or the attempt has
// finished: host polling instead of an event recorded after every pass, whose
// completion signal held the next density pass back by ~6 us.  A queue that
// stays slow past 200 ms is waited for with a stream synchronisation, after
// which one of the two words is set.
static psb_status wait_progress(psb_system s, int slot, int it) {
  const int32_t* prog = s->h_final + 64 + slot;
  const int32_t* fin = s->h_final + slot;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spin = 1;; ++spin) {
    if (__atomic_load_n(prog, __ATOMIC_ACQUIRE) >= it || __atomic_load_n(fin, __ATOMIC_ACQUIRE) >= 0)
      return PSB_OK;
    if ((spin & 4095u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
      PSB_HIP(hipStreamSynchronize(s->ctx->stream));
      return PSB_OK;
    }
    __builtin_ia32_pause();
  }
}

// One attempt at a PCISPH step with constants c.  from_a: restart from the
// start state A of the current step (a retry: the grid, the lists and the
// densities of A are still valid, the pressure starts at 0); otherwise the
// step begins with the grid build of B.  plain: a retry substep, which runs
// the plain corrector (p = 0 start, no warm start, features or recorder).
// Returns the final iteration and whether it ended above the tolerance.
static psb_status pcisph_attempt(psb_system s, int step, const Consts& c, bool from_a,
                                 bool plain, int* final_out, bool* failed_out) {
  hipStream_t st = s->ctx->stream;
  const Lanes lanes = corrector_lanes(s);
  const bool ref = s->prm.mode == PSB_MODE_REFERENCE;
  const bool warm = !ref && !plain && s->prm.warm_start == PSB_WARM_PREVIOUS;
  const bool mlp = !ref && !plain && s->prm.warm_start == PSB_WARM_MLP;
  const bool feat = !ref && !plain && (mlp || s->rec != nullptr);
  const int att = (int)s->step_final.size();   // attempt index of this update call
  const int slot = att & 63;
  s->step_final.push_back(-1);
  const StepReset reset{s->st, nullptr, s->es, 1u};
  if (!from_a) {
    // the status reset rides on the grid build's first kernel
    PSB_TRY(grid_phase(s, att, &reset));
  } else {
    PSB_TRY(launch_pci_retry_reset(s->st, st, s->es));
    PSB_HIP(hipMemsetAsync(s->pres, 0, 4 * (size_t)s->n, st));
  }
  psb_neighbors nb{s->ncount, s->nlist, (uint32_t)s->prm.max_neighbors, s->n};
  Compact16 c16;
  if (s->compact && s->nlist16) {
    c16.list16 = s->nlist16;
    c16.wide = &s->st->wide;
    c16.list32g = s->nlist32g;
    c16.want32 = &s->st->want32;
  }
  const uint16_t* l16 = c16.list16;
  // halo staging: single lane, compact list
  const HaloTile* tiles =
      (s->halo && l16 && lanes.d == 1 && lanes.f == 1 && s->tiles && s->halo_src) ? s->tiles
                                                                                     : nullptr;
  if (!from_a) {
    if (tiles) PSB_TRY(launch_halo(s->grid, s->ax, s->n, s->gd.R, s->tiles, s->halo_src, st));
    // neighbour search + density + epilogue in one pass
    Timed t(s, PH_NEIGHBORS, att);
    const EpiArgs ep = make_epilogue(c, 1, warm ? 1 : 0, feat ? 1 : 0, s->pres, s->ap, s->ax,
                                     s->av, s->aid);
    PSB_TRY(launch_neighbors(s->grid, s->ax, s->n, c.h2, s->gd.R, nb, s->rho, c.mc,
                             s->flags, s->flags + 1, st, c16, s->es, tiles, &ep));
  }
  PciBuffers b{};
  b.fa = s->fa; b.fs = s->fs; b.anp = s->anp; b.xp0 = s->xp0; b.dl = s->dl;
  b.feat = feat ? s->feat : nullptr;
  b.ax = s->ax; b.avh = s->avh; b.aid = s->aid;
  // staged tiles hold LDS slots in the compact list, which only the
  // corrector passes decode: the non-pressure pass then reads the 32-bit list
  b.list16 = tiles ? nullptr : l16;
  b.st = s->st;
  PciOut o{};
  o.bx = s->bx; o.bv = s->bv; o.bvh = s->bvh; o.bid = s->bid; o.bp = s->bp; o.flags = s->flags;
  o.host_final = ref ? nullptr : s->d_final + slot;
  o.host_progress = ref ? nullptr : s->d_final + 64 + slot;
  o.es = s->es;
  const bool surf = ref || s->prm.surface_tension != 0.f;
  {
    Timed t(s, PH_NONPRESSURE, att);
    PSB_TRY(launch_pci_nonpressure(s->n, c, surf, warm, false, s->ncount, s->nlist, s->ax, s->av,
                                   s->rho, s->avh, s->pres, b, o, st));
  }
  float4* cur = s->xp0;
  float4* nxt = s->xp1;
  if (warm) {
    Timed t(s, PH_PCI_FORCE, att);
    PSB_TRY(launch_pci_force(s->n, c, true, s->ncount, s->nlist, cur, nxt, s->rho_pred, s->pres,
                             b, s->st, -1, 1, o, st, lanes.f, l16, tiles, s->halo_src));
    std::swap(cur, nxt);
  }
  const int min_it = c.min_it, max_it = c.max_it;
  int final_known = -1;
  if (!ref) s->h_final[slot] = s->h_final[64 + slot] = -1;
  for (int k = 0; k < max_it; ++k) {
    {
      Timed t(s, PH_PCI_DENSITY, att, k);
      const int dflags = (k == 0 && feat) ? (kDensWriteFeature | (mlp ? kDensSkipUpdate : 0)) : 0;
      PSB_TRY(launch_pci_density(s->n, c, s->ncount, s->nlist, cur, s->pres, s->rho_pred, s->dl,
                                 s->st, k, s->feat, dflags, st, s->aid, lanes.d, l16, s->es, tiles,
                                 s->halo_src));
    }
    if (k == 0 && mlp) {
      // the network sets the starting pressure (and the w lane of the
      // predicted positions the force pass gathers)
      Timed t(s, PH_MLP, att);
      PSB_TRY(mlp_infer(s->mlp, s->feat, s->n, s->pres, cur, 1, c.pscale, st,
                        s->n_spheres ? s->aid : nullptr, c.n_fluid));
    }
    {
      Timed t(s, PH_PCI_FORCE, att, k);
      PSB_TRY(launch_pci_force(s->n, c, true, s->ncount, s->nlist, cur, nxt, s->rho_pred,
                               s->pres, b, s->st, k, 0, o, st, lanes.f, l16, tiles, s->halo_src));
    }
    std::swap(cur, nxt);
    if (ref) continue;
    if (k >= min_it) {
      // iteration k is queued; wait for the force pass of k-1 to start and
      // stop queueing once it finished the step
      PSB_TRY(wait_progress(s, slot, k - 1));
      int32_t f = __atomic_load_n(&s->h_final[slot], __ATOMIC_ACQUIRE);
      if (f >= 0) { final_known = f; break; }
    }
  }
  if (ref) final_known = 2;
  if (final_known < 0) {
    // finished by the last queued iteration (k == max-1 or the min-th one)
    PSB_HIP(hipStreamSynchronize(st));
    final_known = __atomic_load_n(&s->h_final[slot], __ATOMIC_ACQUIRE);
    if (final_known < 0) final_known = max_it - 1;
  }
  *failed_out = (final_known & kFinalFailed) != 0;
  final_known &= kFinalFailed - 1;
  s->step_final[att] = final_known;
  *final_out = final_known;
  (void)step;
  return PSB_OK;
}

// After an accepted attempt: the rigid spheres (config 3) and the
// training-set recorder.
static psb_status pcisph_finish(psb_system s, const Consts& c, int final_known, bool plain) {
  hipStream_t st = s->ctx->stream;
  const bool ref = s->prm.mode == PSB_MODE_REFERENCE;
  const bool warm = !ref && !plain && s->prm.warm_start == PSB_WARM_PREVIOUS;
  const bool mlp = !ref && !plain && s->prm.warm_start == PSB_WARM_MLP;
  const bool feat = !ref && !plain && (mlp || s->rec != nullptr);
  const int att = (int)s->step_final.size() - 1;
  if (s->n_spheres) {
    // the pressure force of the final iteration used the positions its
    // density pass read
    float4* xfin = ((final_known + (warm ? 1 : 0)) & 1) ? s->xp1 : s->xp0;
    SphereArgs a{};
    a.c_in = s->d_sc[s->sphere_cur]; a.v_in = s->d_sv[s->sphere_cur];
    a.c_out = s->d_sc[s->sphere_cur ^ 1]; a.v_out = s->d_sv[s->sphere_cur ^ 1];
    a.f_fluid = s->d_sf; a.fb = s->d_fb; a.offsets = s->d_soff;
    a.count = s->n_spheres; a.per = s->per_sphere;
    a.radius = s->sphere.radius;
    a.mass = (float)(s->sphere.density * 4.0 / 3.0 * M_PI * std::pow((double)s->sphere.radius, 3));
    a.spring = s->sphere.spring;
    a.damping = s->sphere.damping;
    a.wall_damping = s->sphere.damping;
    for (int ax = 0; ax < 3; ++ax) { a.box_min[ax] = s->prm.box_min[ax]; a.box_max[ax] = s->prm.box_max[ax]; }
    PSB_TRY(launch_bodies(s->n, c, s->ncount, s->nlist, xfin, s->aid, a, s->bx, s->bv, s->bvh,
                          s->bid, s->bp, st));
    s->sphere_cur ^= 1;
  }
  if (s->rec && feat) {
    // features and converged pressure of every rec_stride-th particle
    psb_dataset d = s->rec;
    const uint32_t stride = s->rec_stride;
    const uint32_t offset = (uint32_t)(s->step_counter % stride);
    uint64_t cnt = offset < s->n ? (uint64_t)(s->n - offset + stride - 1) / stride : 0;
    if (d->count + cnt > d->capacity) cnt = d->capacity - d->count;
    if (cnt) {
      Timed t(s, PH_RECORD, att);
      PSB_TRY(launch_record(s->feat, s->pres, s->n, stride, offset, c.inv_ps, d->feat,
                            d->target, d->capacity, d->count, (uint32_t)cnt, st));
      d->count += cnt;
    }
  }
  return PSB_OK;
}

// Advance one step of dt / 2^level.  A step whose corrector ends above the
// tolerance is redone from its start state A as two steps of half the size
// (params.retry_levels deep); the failed attempt's end state in B is
// overwritten by the first half, which reuses A's grid and lists.  Substeps
// (level > 0) run the plain corrector.
static psb_status pcisph_advance(psb_system s, int step, int level, bool from_a) {
  Consts c{};
  PSB_TRY(level_consts(s, level, &c));
  int fin = 0;
  bool failed = false;
  PSB_TRY(pcisph_attempt(s, step, c, from_a, level > 0, &fin, &failed));
  if (failed && c.retry_ok) {
    if (level == 0) s->stats.retried_steps += 1;
    s->stats.substeps += 2;
    PSB_TRY(pcisph_advance(s, step, level + 1, true));
    PSB_TRY(pcisph_advance(s, step, level + 1, false));
    return PSB_OK;
  }
  return pcisph_finish(s, c, fin, level > 0);
}

static psb_status pcisph_step(psb_system s, int step) { return pcisph_advance(s, step, 0, false); }

static psb_status ensure_features(psb_system s) {
  if (s->feat) return PSB_OK;
  PSB_HIP(hipMalloc(&s->feat, 8 * 4 * (size_t)(s->cap ? s->cap : 1)));
  return PSB_OK;
}

}  // namespace psb

extern "C" {

psb_status psb_system_create(psb_context ctx, const psb_sph_params* prm, uint32_t capacity,
                             psb_system* out) {
  PSB_ARG(ctx && prm && out, "null argument");
  *out = nullptr;
  PSB_ARG(capacity <= kMaxParticles, "capacity above 2^27 particles");
  psb_system s = new psb_system_s();
  s->ctx = ctx;
  s->prm = *prm;
  s->cap = capacity;
  psb_status stt = make_consts(*prm, &s->c, &s->gd);
  s->c.retry_ok = (prm->mode == PSB_MODE_CONFIG && prm->solver == PSB_SOLVER_PCISPH &&
                   prm->retry_levels > 0) ? 1 : 0;
  if (stt == PSB_OK) {
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) stt = hip_fail(e, "hipSetDevice", __FILE__, __LINE__);
  }
  if (stt == PSB_OK) {
    hipError_t e = hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    if (e != hipSuccess) stt = hip_fail(e, "hipDeviceGetAttribute", __FILE__, __LINE__);
  }
  if (stt == PSB_OK) stt = grid_create(ctx, prm->grid, capacity, &s->grid);
  if (stt == PSB_OK) stt = sys_alloc(s);
  if (stt == PSB_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) stt = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
  }
  if (stt != PSB_OK) { psb_system_destroy(s); return stt; }
  *out = s;
  return PSB_OK;
}

psb_status psb_system_destroy(psb_system s) {
  if (!s) return PSB_OK;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  void* ptrs[] = {s->ax, s->av, s->avh, s->bx, s->bv, s->bvh, s->aid, s->bid, s->ap, s->bp,
                  s->rho, s->pres, s->rho_pred, s->fa, s->fs, s->anp, s->xp0, s->xp1, s->dl,
                  s->ncount, s->nlist, s->nlist16, s->nlist32g, s->flags, s->st, s->feat,
                  s->es.base,
                  s->tiles, s->halo_src, s->d_tbl, s->mat_by_id, s->ms};
  for (void* p : ptrs) if (p) hipFree(p);
  if (s->h_final) hipHostFree(s->h_final);
  for (void* p : {(void*)s->d_sc[0], (void*)s->d_sc[1], (void*)s->d_sv[0], (void*)s->d_sv[1],
                  (void*)s->d_sf, (void*)s->d_soff, (void*)s->d_fb, (void*)s->d_coll})
    if (p) hipFree(p);
  for (auto e : s->ev) hipEventDestroy(e);
  if (s->grid) psb_finalize_grid(s->grid);
  delete s;
  return PSB_OK;
}

psb_status psb_system_get_params(psb_system s, psb_sph_params* out) {
  PSB_ARG(s && out, "null argument");
  *out = s->prm;
  return PSB_OK;
}

uint32_t psb_system_count(psb_system s) { return s ? s->n : 0; }
uint32_t psb_system_num_spheres(psb_system s) { return s ? s->n_spheres : 0; }

// Boundary particles of the rigid spheres after the n_fluid fluid particles
// (ids n_fluid ..), at rest relative to their sphere; sphere state uploaded.
static psb_status append_spheres(psb_system s, uint32_t n_fluid) {
  s->n_fluid = n_fluid;
  s->c.n_fluid = kNoBodies;
  if (!s->n_spheres) return PSB_OK;
  const uint32_t nb = s->n_spheres * s->per_sphere;
  if ((uint64_t)n_fluid + nb > s->cap) {
    set_error("capacity must hold the fluid and the sphere boundary particles");
    return PSB_ERR_CAPACITY;
  }
  std::vector<float> x(4 * (size_t)nb), v(4 * (size_t)nb, 0.f), zeros(nb, 0.f);
  std::vector<uint32_t> ids(nb);
  std::vector<float> c4(4 * (size_t)s->n_spheres, 0.f), v4(4 * (size_t)s->n_spheres, 0.f);
  for (uint32_t j = 0; j < s->n_spheres; ++j)
    for (int a = 0; a < 3; ++a) {
      c4[4 * j + a] = s->sphere_c[3 * j + a];
      v4[4 * j + a] = s->sphere_v[3 * j + a];
    }
  for (uint32_t j = 0; j < s->n_spheres; ++j)
    for (uint32_t l = 0; l < s->per_sphere; ++l) {
      const size_t k = (size_t)j * s->per_sphere + l;
      for (int a = 0; a < 3; ++a) {
        x[4 * k + a] = c4[4 * j + a] + s->sphere_off[4 * l + a];
        v[4 * k + a] = v4[4 * j + a];
      }
      x[4 * k + 3] = 0.f;
      ids[k] = n_fluid + (uint32_t)k;
    }
  hipStream_t st = s->ctx->stream;
  const size_t b4 = 16 * (size_t)nb, b1 = 4 * (size_t)nb;
  for (float4* dst : {s->bx, s->ax})
    PSB_HIP(hipMemcpyAsync(dst + n_fluid, x.data(), b4, hipMemcpyHostToDevice, st));
  for (float4* dst : {s->bv, s->av, s->bvh, s->avh})
    PSB_HIP(hipMemcpyAsync(dst + n_fluid, v.data(), b4, hipMemcpyHostToDevice, st));
  for (uint32_t* dst : {s->bid, s->aid})
    PSB_HIP(hipMemcpyAsync(dst + n_fluid, ids.data(), b1, hipMemcpyHostToDevice, st));
  for (float* dst : {s->bp, s->pres, s->rho, s->rho_pred})
    PSB_HIP(hipMemcpyAsync(dst + n_fluid, zeros.data(), b1, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->d_sc[0], c4.data(), 16 * (size_t)s->n_spheres, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->d_sv[0], v4.data(), 16 * (size_t)s->n_spheres, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemsetAsync(s->d_sf, 0, 16 * (size_t)s->n_spheres, st));
  PSB_HIP(hipStreamSynchronize(st));
  s->sphere_cur = 0;
  s->n = n_fluid + nb;
  s->c.n_fluid = n_fluid;
  s->nmat = 0;
  return PSB_OK;
}

static psb_status upload_state(psb_system s, uint32_t n, const std::vector<float>& x,
                               const std::vector<float>& v, const std::vector<float>& vh,
                               const std::vector<float>& xo, const std::vector<float>& vo,
                               const std::vector<float>& vav) {
  hipStream_t st = s->ctx->stream;
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  std::vector<float> zeros(n, 0.f);
  PSB_HIP(hipMemcpyAsync(s->bx, x.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->bv, v.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->bvh, vh.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->ax, xo.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->av, vo.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->avh, vav.data(), 16 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->bid, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->aid, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->bp, zeros.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->pres, zeros.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->rho, zeros.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->rho_pred, zeros.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipStreamSynchronize(st));
  s->n = n;
  PSB_TRY(append_spheres(s, n));
  s->has_state = true;
  s->nmat = 0;   // per-particle materials refer to the previous particles
  s->stepped = false;
  std::memset(&s->stats, 0, sizeof(s->stats));
  return PSB_OK;
}

psb_status psb_system_set_particles(psb_system s, uint32_t n, const double* old_pos,
                                    const double* new_pos, const double* old_vel,
                                    const double* new_vel, const double* vel_inter_av,
                                    const double* vel_inter_ap) {
  PSB_ARG(s, "null system");
  if (n > s->cap) { set_error("n exceeds system capacity"); return PSB_ERR_CAPACITY; }
  PSB_HIP(hipSetDevice(s->ctx->device));
  auto pack = [&](const double* src) {
    std::vector<float> v(4 * (size_t)n, 0.f);
    if (src)
      for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) v[4 * (size_t)i + a] = (float)src[3 * (size_t)i + a];
    return v;
  };
  return upload_state(s, n, pack(new_pos), pack(new_vel), pack(vel_inter_ap), pack(old_pos),
                      pack(old_vel), pack(vel_inter_av));
}

psb_status psb_system_set_particles_device(psb_system s, uint32_t n, const float* pos4,
                                           const float* vel4, const float* vh4) {
  PSB_ARG(s && (pos4 || !n), "null argument");
  if (n > s->cap) { set_error("n exceeds system capacity"); return PSB_ERR_CAPACITY; }
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const size_t b4 = 16 * (size_t)n;
  PSB_HIP(hipMemcpyAsync(s->bx, pos4, b4, hipMemcpyDeviceToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->ax, pos4, b4, hipMemcpyDeviceToDevice, st));
  if (vel4) {
    PSB_HIP(hipMemcpyAsync(s->bv, vel4, b4, hipMemcpyDeviceToDevice, st));
    PSB_HIP(hipMemcpyAsync(s->av, vel4, b4, hipMemcpyDeviceToDevice, st));
  } else {
    PSB_HIP(hipMemsetAsync(s->bv, 0, b4, st));
    PSB_HIP(hipMemsetAsync(s->av, 0, b4, st));
  }
  if (vh4) {
    PSB_HIP(hipMemcpyAsync(s->bvh, vh4, b4, hipMemcpyDeviceToDevice, st));
    PSB_HIP(hipMemcpyAsync(s->avh, vh4, b4, hipMemcpyDeviceToDevice, st));
  } else {
    PSB_HIP(hipMemsetAsync(s->bvh, 0, b4, st));
    PSB_HIP(hipMemsetAsync(s->avh, 0, b4, st));
  }
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  PSB_HIP(hipMemcpyAsync(s->bid, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(s->aid, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemsetAsync(s->bp, 0, 4 * (size_t)n, st));
  PSB_HIP(hipMemsetAsync(s->pres, 0, 4 * (size_t)n, st));
  PSB_HIP(hipMemsetAsync(s->rho, 0, 4 * (size_t)n, st));
  PSB_HIP(hipMemsetAsync(s->rho_pred, 0, 4 * (size_t)n, st));
  PSB_HIP(hipStreamSynchronize(st));
  s->n = n;
  PSB_TRY(append_spheres(s, n));
  s->has_state = true;
  s->nmat = 0;   // per-particle materials refer to the previous particles
  s->stepped = false;
  std::memset(&s->stats, 0, sizeof(s->stats));
  return PSB_OK;
}

psb_status psb_system_update(psb_system s, int32_t steps) {
  PSB_ARG(s && steps >= 0, "bad argument");
  if (!s->has_state) { set_error("psb_system_update before psb_system_set_particles"); return PSB_ERR_STATE; }
  if (s->prm.solver == PSB_SOLVER_PCISPH && s->prm.mode == PSB_MODE_CONFIG &&
      s->prm.warm_start == PSB_WARM_MLP) {
    if (!s->mlp) { set_error("PSB_WARM_MLP without psb_system_set_mlp"); return PSB_ERR_STATE; }
    PSB_ARG(s->c.min_it >= 2, "PSB_WARM_MLP needs min_iterations >= 2 (the first density pass "
                              "only measures)");
  }
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  PSB_HIP(hipMemsetAsync(s->flags, 0, 8, st));
  PSB_HIP(hipMemsetAsync(s->grid->flags, 0, 4, st));
  PSB_TRY(launch_pci_reset(s->st, true, st, 1, nullptr, s->es));
  s->step_final.clear();
  s->stats.retried_steps = 0;
  s->stats.substeps = 0;
  for (int k = 0; k < steps && s->n > 0; ++k) {
    if (s->nmat) {
      if (s->prm.solver == PSB_SOLVER_PCISPH) PSB_TRY(mixed_pcisph_step(s, k));
      else PSB_TRY(mixed_wcsph_step(s, k));
    } else if (s->prm.solver == PSB_SOLVER_PCISPH) {
      PSB_TRY(pcisph_step(s, k));
    } else {
      PSB_TRY(wcsph_step(s, k));
    }
    s->step_counter += 1;
    if (s->recs.size() > 4096) PSB_TRY(collect_timing(s));
  }
  // statistics
  uint32_t fl[2] = {0, 0}, gf = 0;
  PciStatus hs;
  PSB_HIP(hipMemcpyAsync(fl, s->flags, 8, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(&gf, s->grid->flags, 4, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(&hs, s->st, 32, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  PSB_TRY(collect_timing(s));
  s->stats.flags = fl[0] | gf;
  s->stats.max_neighbors_seen = (int32_t)fl[1];
  s->stats.steps = steps;
  if (s->prm.solver == PSB_SOLVER_PCISPH && steps > 0) {
    s->grouped32 = hs.want32 != 0u && s->nlist32g && s->compact;
    s->stats.iterations = hs.last_iters;
    std::memcpy(&s->stats.max_error, &hs.last_err, 4);
    s->stats.total_iterations = (int64_t)hs.total_iters;
    std::memcpy(&s->stats.worst_error, &hs.worst_err, 4);
  } else {
    s->stats.iterations = 0;
    s->stats.max_error = 0.f;
    s->stats.total_iterations = 0;
    s->stats.worst_error = 0.f;
  }
  if (steps > 0) s->stepped = true;
  return PSB_OK;
}

psb_status psb_system_set_materials(psb_system s, uint32_t count, const psb_sph_params* materials,
                                    const uint8_t* index, uint32_t n) {
  PSB_ARG(s, "null system");
  if (count == 0) { s->nmat = 0; return PSB_OK; }
  PSB_ARG(s->prm.mode == PSB_MODE_REFERENCE, "per-particle materials need PSB_MODE_REFERENCE");
  PSB_ARG(s->prm.solver == PSB_SOLVER_PCISPH || s->prm.eos == PSB_EOS_LINEAR,
          "per-particle materials with WCSPH need PSB_EOS_LINEAR");
  if (!s->has_state) {
    set_error("psb_system_set_materials before psb_system_set_particles");
    return PSB_ERR_STATE;
  }
  PSB_ARG(index && n == s->n, "one material index per particle of the last set_particles");
  PSB_ARG(s->n_spheres == 0, "per-particle materials and rigid spheres do not combine");
  for (uint32_t i = 0; i < n; ++i) PSB_ARG(index[i] < count, "material index out of range");
  MatTable tbl;
  PSB_TRY(make_mat_table(s->prm, materials, count, &tbl));
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const size_t N = s->cap ? s->cap : 1;
  if (!s->d_tbl) PSB_HIP(hipMalloc((void**)&s->d_tbl, sizeof(MatTable)));
  if (!s->mat_by_id) PSB_HIP(hipMalloc((void**)&s->mat_by_id, N));
  if (!s->ms) PSB_HIP(hipMalloc((void**)&s->ms, N));
  PSB_HIP(hipMemcpyAsync(s->d_tbl, &tbl, sizeof(MatTable), hipMemcpyHostToDevice, st));
  if (n) PSB_HIP(hipMemcpyAsync(s->mat_by_id, index, n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipStreamSynchronize(st));
  s->nmat = count;
  return PSB_OK;
}

uint32_t psb_system_num_materials(psb_system s) { return s ? s->nmat : 0; }

psb_status psb_system_get_stats(psb_system s, psb_step_stats* out) {
  PSB_ARG(s && out, "null argument");
  *out = s->stats;
  return PSB_OK;
}

psb_status psb_system_get_particles_f32(psb_system s, float* pos4, float* vel4, float* vh4,
                                        float* density, float* pressure, uint32_t* ids) {
  PSB_ARG(s, "null system");
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const size_t n = s->n;
  const float* dens = (s->prm.solver == PSB_SOLVER_PCISPH) ? s->rho_pred : s->rho;
  if (pos4) PSB_HIP(hipMemcpyAsync(pos4, s->bx, 16 * n, hipMemcpyDeviceToHost, st));
  if (vel4) PSB_HIP(hipMemcpyAsync(vel4, s->bv, 16 * n, hipMemcpyDeviceToHost, st));
  if (vh4) PSB_HIP(hipMemcpyAsync(vh4, s->bvh, 16 * n, hipMemcpyDeviceToHost, st));
  if (density) PSB_HIP(hipMemcpyAsync(density, dens, 4 * n, hipMemcpyDeviceToHost, st));
  if (pressure) PSB_HIP(hipMemcpyAsync(pressure, s->pres, 4 * n, hipMemcpyDeviceToHost, st));
  if (ids) PSB_HIP(hipMemcpyAsync(ids, s->bid, 4 * n, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  return PSB_OK;
}

psb_status psb_system_get_particles(psb_system s, double* old_pos, double* new_pos,
                                    double* old_vel, double* new_vel, double* vav, double* vap,
                                    double* density, double* pressure) {
  PSB_ARG(s, "null system");
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const size_t n = s->n;
  std::vector<uint32_t> ids(n);
  PSB_HIP(hipMemcpyAsync(ids.data(), s->bid, 4 * n, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  auto fetch4 = [&](const float4* src, double* dst) -> psb_status {
    if (!dst) return PSB_OK;
    std::vector<float> t(4 * n);
    PSB_HIP(hipMemcpyAsync(t.data(), src, 16 * n, hipMemcpyDeviceToHost, st));
    PSB_HIP(hipStreamSynchronize(st));
    for (size_t k = 0; k < n; ++k)
      for (int a = 0; a < 3; ++a) dst[3 * (size_t)ids[k] + a] = t[4 * k + a];
    return PSB_OK;
  };
  auto fetch1 = [&](const float* src, double* dst) -> psb_status {
    if (!dst) return PSB_OK;
    std::vector<float> t(n);
    PSB_HIP(hipMemcpyAsync(t.data(), src, 4 * n, hipMemcpyDeviceToHost, st));
    PSB_HIP(hipStreamSynchronize(st));
    for (size_t k = 0; k < n; ++k) dst[ids[k]] = t[k];
    return PSB_OK;
  };
  PSB_TRY(fetch4(s->bx, new_pos));
  PSB_TRY(fetch4(s->ax, old_pos));
  PSB_TRY(fetch4(s->bv, new_vel));
  PSB_TRY(fetch4(s->av, old_vel));
  PSB_TRY(fetch4(s->bvh, vap));
  PSB_TRY(fetch4(s->avh, vav));
  PSB_TRY(fetch1(s->prm.solver == PSB_SOLVER_PCISPH ? s->rho_pred : s->rho, density));
  PSB_TRY(fetch1(s->pres, pressure));
  return PSB_OK;
}

psb_status psb_system_get_arrays(psb_system s, psb_system_arrays* out) {
  PSB_ARG(s && out, "null argument");
  out->pos4 = (const float*)s->bx;
  out->old_pos4 = (const float*)s->ax;
  out->vel4 = (const float*)s->bv;
  out->old_vel4 = (const float*)s->av;
  out->vel_inter_ap4 = (const float*)s->bvh;
  out->vel_inter_av4 = (const float*)s->avh;
  out->density = s->prm.solver == PSB_SOLVER_PCISPH ? s->rho_pred : s->rho;
  out->pressure = s->pres;
  out->ids = s->bid;
  out->hash = s->grid->hash;
  out->perm = s->grid->perm;
  out->cell_start = s->grid->start;
  out->neighbors = psb_neighbors{s->ncount, s->nlist, (uint32_t)s->prm.max_neighbors, s->n};
  out->ncells = s->grid->ncells;
  return PSB_OK;
}

psb_status psb_system_set_timing(psb_system s, int32_t every) {
  PSB_ARG(s && every >= 0, "bad argument");
  s->timing_every = every;
  return PSB_OK;
}

int32_t psb_system_get_timing(psb_system s, int32_t cap, const char** names, double* total_ms,
                              int64_t* launches) {
  if (!s) return 0;
  for (int i = 0; i < PH_COUNT && i < cap; ++i) {
    if (names) names[i] = kPhaseNames[i];
    if (total_ms) total_ms[i] = s->t_total[i];
    if (launches) launches[i] = s->t_count[i];
  }
  return PH_COUNT;
}

psb_status psb_system_reset_timing(psb_system s) {
  PSB_ARG(s, "null system");
  for (int i = 0; i < PH_COUNT; ++i) { s->t_total[i] = 0; s->t_count[i] = 0; }
  return PSB_OK;
}

psb_status psb_system_time_corrector(psb_system s, int32_t reps, double* us_density,
                                     double* us_force) {
  PSB_ARG(s && reps > 0 && us_density && us_force, "bad argument");
  PSB_ARG(s->prm.solver == PSB_SOLVER_PCISPH && s->prm.mode == PSB_MODE_CONFIG,
          "psb_system_time_corrector needs a PCISPH config-mode system");
  if (!s->stepped || s->n == 0) { set_error("psb_system_time_corrector before a step"); return PSB_ERR_STATE; }
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const Lanes lanes = corrector_lanes(s);
  // iteration 0 of a corrector that never finalises: the force pass writes
  // the next prediction only
  Consts c = s->c;
  c.min_it = 2;
  c.max_it = kMaxIterSlots - 1;
  c.retry_ok = 0;
  const uint16_t* l16 = (s->compact && s->nlist16) ? s->nlist16 : nullptr;
  const HaloTile* tiles =
      (s->halo && l16 && lanes.d == 1 && lanes.f == 1 && s->tiles && s->halo_src) ? s->tiles
                                                                                     : nullptr;
  PciBuffers b{};
  b.fa = s->fa; b.fs = s->fs; b.anp = s->anp; b.xp0 = s->xp0; b.dl = s->dl;
  b.ax = s->ax; b.avh = s->avh; b.aid = s->aid;
  b.list16 = tiles ? nullptr : l16;
  b.st = s->st;
  PciOut o{};
  o.flags = s->flags;
  o.es = s->es;
  PSB_TRY(launch_pci_retry_reset(s->st, st, s->es));
  hipEvent_t e[3];
  for (auto& x : e) PSB_HIP(hipEventCreate(&x));
  psb_status stt = PSB_OK;
  auto run = [&]() -> psb_status {
    PSB_HIP(hipEventRecord(e[0], st));
    for (int r = 0; r < reps; ++r)
      PSB_TRY(launch_pci_density(s->n, c, s->ncount, s->nlist, s->xp0, s->pres, s->rho_pred,
                                 s->dl, s->st, 0, nullptr, 0, st, s->aid, lanes.d, l16, s->es,
                                 tiles, s->halo_src));
    PSB_HIP(hipEventRecord(e[1], st));
    for (int r = 0; r < reps; ++r)
      PSB_TRY(launch_pci_force(s->n, c, true, s->ncount, s->nlist, s->xp0, s->xp1, s->rho_pred,
                               s->pres, b, s->st, 0, 0, o, st, lanes.f, l16, tiles, s->halo_src));
    PSB_HIP(hipEventRecord(e[2], st));
    PSB_HIP(hipEventSynchronize(e[2]));
    float md = 0.f, mf = 0.f;
    PSB_HIP(hipEventElapsedTime(&md, e[0], e[1]));
    PSB_HIP(hipEventElapsedTime(&mf, e[1], e[2]));
    *us_density = 1e3 * (double)md / reps;
    *us_force = 1e3 * (double)mf / reps;
    return PSB_OK;
  };
  stt = run();
  for (auto& x : e) hipEventDestroy(x);
  return stt;
}

float psb_system_pressure_scale(psb_system s) { return s ? s->c.pscale : 0.f; }

uint32_t psb_sphere_particle_count(const psb_sph_params* prm, const psb_sphere_params* sp) {
  if (!prm || !sp || !(sp->radius > 0.f)) return 0;
  const double s = prm->lattice_spacing > 0.f ? (double)prm->lattice_spacing
                                              : 0.5 * (double)prm->support_radius;
  return (uint32_t)(sphere_offsets(sp->radius, s).size() / 4);
}

psb_status psb_system_set_spheres(psb_system s, uint32_t count, const float* centers3,
                                  const float* velocities3, const psb_sphere_params* sp) {
  PSB_ARG(s, "null system");
  PSB_ARG(s->prm.mode == PSB_MODE_CONFIG && s->prm.solver == PSB_SOLVER_PCISPH,
          "spheres need a PSB_MODE_CONFIG PCISPH system");
  PSB_ARG(count == 0 || (centers3 && sp), "null argument");
  PSB_HIP(hipSetDevice(s->ctx->device));
  for (void* p : {(void*)s->d_sc[0], (void*)s->d_sc[1], (void*)s->d_sv[0], (void*)s->d_sv[1],
                  (void*)s->d_sf, (void*)s->d_soff, (void*)s->d_fb})
    if (p) hipFree(p);
  s->d_sc[0] = s->d_sc[1] = s->d_sv[0] = s->d_sv[1] = s->d_sf = s->d_soff = s->d_fb = nullptr;
  s->n_spheres = 0;
  s->per_sphere = 0;
  s->has_state = false;   // set the particles again: the boundary particles follow the fluid
  if (!count) return PSB_OK;
  PSB_ARG(sp->radius > 0.f && sp->density > 0.f && sp->spring >= 0.f && sp->damping >= 0.f,
          "sphere radius and density must be > 0, spring and damping >= 0");
  const double lat = s->prm.lattice_spacing > 0.f ? (double)s->prm.lattice_spacing
                                                  : 0.5 * (double)s->prm.support_radius;
  s->sphere_off = sphere_offsets(sp->radius, lat);
  PSB_ARG(!s->sphere_off.empty(), "sphere radius below the lattice spacing");
  s->sphere = *sp;
  s->per_sphere = (uint32_t)(s->sphere_off.size() / 4);
  s->sphere_c.assign(centers3, centers3 + 3 * (size_t)count);
  s->sphere_v.assign(3 * (size_t)count, 0.f);
  if (velocities3) s->sphere_v.assign(velocities3, velocities3 + 3 * (size_t)count);
  const size_t b4 = 16 * (size_t)count;
  PSB_HIP(hipMalloc(&s->d_sc[0], b4)); PSB_HIP(hipMalloc(&s->d_sc[1], b4));
  PSB_HIP(hipMalloc(&s->d_sv[0], b4)); PSB_HIP(hipMalloc(&s->d_sv[1], b4));
  PSB_HIP(hipMalloc(&s->d_sf, b4));
  PSB_HIP(hipMalloc(&s->d_soff, 16 * (size_t)s->per_sphere));
  PSB_HIP(hipMalloc(&s->d_fb, 16 * (size_t)s->per_sphere * count));
  PSB_HIP(hipMemcpy(s->d_soff, s->sphere_off.data(), 16 * (size_t)s->per_sphere,
                    hipMemcpyHostToDevice));
  s->n_spheres = count;
  return PSB_OK;
}

psb_status psb_system_get_spheres(psb_system s, float* centers3, float* velocities3,
                                  float* fluid_forces3) {
  PSB_ARG(s, "null system");
  if (!s->n_spheres) return PSB_OK;
  PSB_HIP(hipSetDevice(s->ctx->device));
  const uint32_t S = s->n_spheres;
  std::vector<float> c(4 * (size_t)S), v(4 * (size_t)S), f(4 * (size_t)S);
  hipStream_t st = s->ctx->stream;
  PSB_HIP(hipMemcpyAsync(c.data(), s->d_sc[s->sphere_cur], 16 * (size_t)S, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(v.data(), s->d_sv[s->sphere_cur], 16 * (size_t)S, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(f.data(), s->d_sf, 16 * (size_t)S, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  for (uint32_t j = 0; j < S; ++j)
    for (int a = 0; a < 3; ++a) {
      if (centers3) centers3[3 * j + a] = c[4 * j + a];
      if (velocities3) velocities3[3 * j + a] = v[4 * j + a];
      if (fluid_forces3) fluid_forces3[3 * j + a] = f[4 * j + a];
    }
  return PSB_OK;
}

static psb_status upload_colliders(psb_system s) {
  PSB_HIP(hipSetDevice(s->ctx->device));
  PSB_HIP(hipStreamSynchronize(s->ctx->stream));   // no step still reads the old array
  if (s->d_coll) { PSB_HIP(hipFree(s->d_coll)); s->d_coll = nullptr; }
  if (!s->colliders.empty()) {
    const size_t nb = sizeof(ColliderDev) * s->colliders.size();
    PSB_HIP(hipMalloc(&s->d_coll, nb));
    PSB_HIP(hipMemcpy(s->d_coll, s->colliders.data(), nb, hipMemcpyHostToDevice));
  }
  s->c.coll = s->d_coll;
  s->c.ncoll = (int)s->colliders.size();
  return PSB_OK;
}

psb_status psb_system_add_collider(psb_system s, const psb_collider* k) {
  PSB_ARG(s && k, "null argument");
  PSB_ARG(s->colliders.size() < PSB_MAX_COLLIDERS, "at most PSB_MAX_COLLIDERS colliders");
  ColliderDev d;
  PSB_TRY(make_collider(*k, &d));
  s->colliders.push_back(d);
  return upload_colliders(s);
}

psb_status psb_system_clear_colliders(psb_system s) {
  PSB_ARG(s, "null system");
  s->colliders.clear();
  return upload_colliders(s);
}

psb_status psb_surface_extract(psb_surface surf, psb_system s, const psb_surface_params* prm,
                               uint32_t* num_vertices, uint32_t* num_triangles) {
  PSB_ARG(surf && s && prm, "null argument");
  if (!s->has_state) { set_error("psb_surface_extract before psb_system_set_particles"); return PSB_ERR_STATE; }
  const uint32_t limit = s->n_spheres ? s->n_fluid : 0xFFFFFFFFu;
  return psb_surface_extract_points(surf, (const float*)s->bx, s->bid, s->n, limit,
                                    s->prm.support_radius, prm, num_vertices, num_triangles);
}

psb_status psb_system_set_halo(psb_system s, int32_t on) {
  PSB_ARG(s, "null system");
  if (on && !s->tiles && s->nlist16) {
    // tile descriptors and staged-index tables, allocated on first use
    PSB_HIP(hipSetDevice(s->ctx->device));
    const size_t ntiles = ((s->cap ? s->cap : 1) + kHaloTile - 1) / kHaloTile;
    PSB_HIP(hipMalloc((void**)&s->tiles, sizeof(HaloTile) * ntiles));
    PSB_HIP(hipMalloc((void**)&s->halo_src, 4 * (size_t)kHaloCap * ntiles));
  }
  s->halo = on != 0;
  return PSB_OK;
}

psb_status psb_system_set_compact_list(psb_system s, int32_t on) {
  PSB_ARG(s, "null system");
  s->compact = on != 0;
  return PSB_OK;
}

psb_status psb_system_set_corrector_lanes(psb_system s, int32_t lanes) {
  PSB_ARG(s && lanes >= 0 && lanes <= 2, "lanes must be 0 (default), 1 or 2");
  s->corrector_lanes = lanes;
  return PSB_OK;
}

psb_status psb_system_set_mlp(psb_system s, psb_mlp m) {
  PSB_ARG(s, "null system");
  PSB_ARG(s->prm.mode == PSB_MODE_CONFIG && s->prm.solver == PSB_SOLVER_PCISPH,
          "the MLP warm start needs a PSB_MODE_CONFIG PCISPH system");
  if (m) PSB_TRY(ensure_features(s));
  s->mlp = m;
  return PSB_OK;
}

psb_status psb_system_get_features(psb_system s, const float** features) {
  PSB_ARG(s && features, "null argument");
  *features = s->feat;
  return PSB_OK;
}

psb_status psb_dataset_create(psb_context ctx, uint64_t capacity, psb_dataset* out) {
  PSB_ARG(ctx && out && capacity > 0, "bad argument");
  PSB_HIP(hipSetDevice(ctx->device));
  psb_dataset d = new psb_dataset_s();
  d->ctx = ctx;
  d->capacity = capacity;
  hipError_t e = hipMalloc(&d->feat, 8 * 4 * capacity);
  if (e == hipSuccess) e = hipMalloc(&d->target, 4 * capacity);
  if (e != hipSuccess) { psb_dataset_destroy(d); return hip_fail(e, "dataset alloc", __FILE__, __LINE__); }
  *out = d;
  return PSB_OK;
}

psb_status psb_dataset_destroy(psb_dataset d) {
  if (!d) return PSB_OK;
  hipSetDevice(d->ctx->device);
  hipStreamSynchronize(d->ctx->stream);
  hipFree(d->feat);
  hipFree(d->target);
  delete d;
  return PSB_OK;
}

psb_status psb_dataset_reset(psb_dataset d) {
  PSB_ARG(d, "null dataset");
  d->count = 0;
  return PSB_OK;
}

psb_status psb_dataset_info(psb_dataset d, const float** features, const float** targets,
                            uint64_t* count, uint64_t* capacity) {
  PSB_ARG(d, "null dataset");
  if (features) *features = d->feat;
  if (targets) *targets = d->target;
  if (count) *count = d->count;
  if (capacity) *capacity = d->capacity;
  return PSB_OK;
}

psb_status psb_system_set_state(psb_system s, uint32_t n, const float* pos4, const float* vel4,
                                const float* vel_half4, const float* pressure,
                                const uint32_t* ids) {
  PSB_ARG(s && (n == 0 || (pos4 && ids)), "null argument");
  PSB_ARG(n <= s->cap, "more particles than the capacity");
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    PSB_ARG(ids[i] < n && !seen[ids[i]], "ids must be a permutation of 0..n-1");
    seen[ids[i]] = 1;
  }
  PSB_HIP(hipSetDevice(s->ctx->device));
  std::vector<float> x(pos4, pos4 + 4 * (size_t)n);
  std::vector<float> v(4 * (size_t)n, 0.f), vh(4 * (size_t)n, 0.f);
  if (vel4) v.assign(vel4, vel4 + 4 * (size_t)n);
  if (vel_half4) vh.assign(vel_half4, vel_half4 + 4 * (size_t)n);
  PSB_TRY(upload_state(s, n, x, v, vh, x, v, vh));
  hipStream_t st = s->ctx->stream;
  if (n) {
    PSB_HIP(hipMemcpyAsync(s->bid, ids, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    PSB_HIP(hipMemcpyAsync(s->aid, ids, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    if (pressure) {
      PSB_HIP(hipMemcpyAsync(s->bp, pressure, 4 * (size_t)n, hipMemcpyHostToDevice, st));
      PSB_HIP(hipMemcpyAsync(s->pres, pressure, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    }
    PSB_HIP(hipStreamSynchronize(st));
  }
  return PSB_OK;
}

// ---- checkpoints ---------------------------------------------------------------
// File layout (little-endian): CkptHeader, psb_sph_params, psb_sphere_params,
// psb_step_stats; the state in storage order: bx bv bvh ax av avh (16 n B
// each), bid aid (4 n B each), bp ap rho rho_pred pres (4 n B each); the
// spheres' centres, velocities and fluid forces (16 B each); an FNV-1a
// 64-bit hash of every byte before it.
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version, params_bytes, sphere_bytes, stats_bytes;
  uint32_t n, n_fluid, n_spheres, per_sphere;
  uint32_t stepped, solver;
  int64_t step_counter;
};
constexpr char kCkptMagic[8] = {'P', 'S', 'B', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 2;   // 2: psb_sph_params.retry_levels, retry stats

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t nb) {
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < nb; ++i) { h ^= b[i]; h *= 1099511628211ull; }
  }
};

struct CkptArrays {
  std::vector<uint8_t> buf;
  // (device pointer, bytes) in file order
  static std::vector<std::pair<void*, size_t>> layout(psb_system s, uint32_t n) {
    const size_t v4 = 16 * (size_t)n, v1 = 4 * (size_t)n;
    return {{s->bx, v4},  {s->bv, v4},  {s->bvh, v4}, {s->ax, v4},       {s->av, v4},
            {s->avh, v4}, {s->bid, v1}, {s->aid, v1}, {s->bp, v1},       {s->ap, v1},
            {s->rho, v1}, {s->rho_pred, v1}, {s->pres, v1}};
  }
};

bool read_header(FILE* f, CkptHeader& h, psb_sph_params& prm, psb_sphere_params& sp,
                 psb_step_stats& stats, Fnv& fnv) {
  if (fread(&h, sizeof h, 1, f) != 1) return false;
  if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != kCkptVersion ||
      h.params_bytes != sizeof(psb_sph_params) || h.sphere_bytes != sizeof(psb_sphere_params) ||
      h.stats_bytes != sizeof(psb_step_stats))
    return false;
  if (fread(&prm, sizeof prm, 1, f) != 1 || fread(&sp, sizeof sp, 1, f) != 1 ||
      fread(&stats, sizeof stats, 1, f) != 1)
    return false;
  fnv.add(&h, sizeof h); fnv.add(&prm, sizeof prm); fnv.add(&sp, sizeof sp);
  fnv.add(&stats, sizeof stats);
  return true;
}
}  // namespace

psb_status psb_system_save_checkpoint(psb_system s, const char* path) {
  PSB_ARG(s && path, "null argument");
  if (!s->has_state) { set_error("psb_system_save_checkpoint before psb_system_set_particles"); return PSB_ERR_STATE; }
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const uint32_t n = s->n, S = s->n_spheres;
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.version = kCkptVersion;
  h.params_bytes = sizeof(psb_sph_params);
  h.sphere_bytes = sizeof(psb_sphere_params);
  h.stats_bytes = sizeof(psb_step_stats);
  h.n = n; h.n_fluid = S ? s->n_fluid : n; h.n_spheres = S; h.per_sphere = s->per_sphere;
  h.stepped = s->stepped ? 1u : 0u; h.solver = (uint32_t)s->prm.solver;
  h.step_counter = s->step_counter;
  auto lay = CkptArrays::layout(s, n);
  if (S) {
    lay.push_back({s->d_sc[s->sphere_cur], 16 * (size_t)S});
    lay.push_back({s->d_sv[s->sphere_cur], 16 * (size_t)S});
    lay.push_back({s->d_sf, 16 * (size_t)S});
  }
  size_t total = 0;
  for (auto& a : lay) total += a.first ? a.second : 0;
  std::vector<uint8_t> buf(total, 0);
  size_t off = 0;
  for (auto& a : lay) {
    // arrays a solver does not allocate (WCSPH: rho_pred) are stored as zeros
    if (a.first && a.second) PSB_HIP(hipMemcpyAsync(buf.data() + off, a.first, a.second, hipMemcpyDeviceToHost, st));
    off += a.first ? a.second : 0;
  }
  PSB_HIP(hipStreamSynchronize(st));
  FILE* f = std::fopen(path, "wb");
  if (!f) { set_error(std::string("cannot write ") + path); return PSB_ERR_IO; }
  Fnv fnv;
  auto put = [&](const void* p, size_t nb) { fnv.add(p, nb); return std::fwrite(p, 1, nb, f) == nb; };
  bool ok = put(&h, sizeof h) && put(&s->prm, sizeof s->prm) && put(&s->sphere, sizeof s->sphere) &&
            put(&s->stats, sizeof s->stats);
  // a presence mask records which arrays follow
  uint32_t mask = 0;
  for (size_t i = 0; i < lay.size(); ++i) if (lay[i].first) mask |= 1u << i;
  ok = ok && put(&mask, sizeof mask) && (total == 0 || put(buf.data(), total));
  const uint64_t hash = fnv.h;
  ok = ok && std::fwrite(&hash, sizeof hash, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { set_error(std::string("write failed: ") + path); return PSB_ERR_IO; }
  return PSB_OK;
}

psb_status psb_checkpoint_info(const char* path, psb_sph_params* params, uint32_t* count,
                               uint32_t* fluid_count, uint32_t* num_spheres) {
  PSB_ARG(path, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("cannot read ") + path); return PSB_ERR_IO; }
  CkptHeader h;
  psb_sph_params prm;
  psb_sphere_params sp;
  psb_step_stats stats;
  Fnv fnv;
  const bool ok = read_header(f, h, prm, sp, stats, fnv);
  std::fclose(f);
  if (!ok) { set_error(std::string("not a checkpoint of this library version: ") + path); return PSB_ERR_IO; }
  if (params) *params = prm;
  if (count) *count = h.n;
  if (fluid_count) *fluid_count = h.n_fluid;
  if (num_spheres) *num_spheres = h.n_spheres;
  return PSB_OK;
}

psb_status psb_system_load_checkpoint(psb_system s, const char* path) {
  PSB_ARG(s && path, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("cannot read ") + path); return PSB_ERR_IO; }
  CkptHeader h;
  psb_sph_params prm;
  psb_sphere_params sp;
  psb_step_stats stats;
  Fnv fnv;
  uint32_t mask = 0;
  if (!read_header(f, h, prm, sp, stats, fnv) || std::fread(&mask, sizeof mask, 1, f) != 1) {
    std::fclose(f);
    set_error(std::string("not a checkpoint of this library version: ") + path);
    return PSB_ERR_IO;
  }
  fnv.add(&mask, sizeof mask);
  if (std::memcmp(&prm, &s->prm, sizeof prm) != 0) {
    std::fclose(f);
    set_error("checkpoint parameters differ from the system's (psb_checkpoint_info returns them)");
    return PSB_ERR_INVALID_ARGUMENT;
  }
  if (h.n > s->cap) { std::fclose(f); set_error("checkpoint holds more particles than the capacity"); return PSB_ERR_CAPACITY; }
  auto lay = CkptArrays::layout(s, h.n);
  const uint32_t S = h.n_spheres;
  if (S) {
    lay.push_back({nullptr, 16 * (size_t)S});
    lay.push_back({nullptr, 16 * (size_t)S});
    lay.push_back({nullptr, 16 * (size_t)S});
  }
  size_t total = 0;
  for (size_t i = 0; i < lay.size(); ++i) if (mask >> i & 1u) total += lay[i].second;
  std::vector<uint8_t> buf(total);
  uint64_t hash = 0;
  const bool rd = (total == 0 || std::fread(buf.data(), 1, total, f) == total) &&
                  std::fread(&hash, sizeof hash, 1, f) == 1;
  std::fclose(f);
  if (!rd) { set_error(std::string("truncated checkpoint: ") + path); return PSB_ERR_IO; }
  fnv.add(buf.data(), total);
  if (fnv.h != hash) { set_error(std::string("checkpoint hash mismatch: ") + path); return PSB_ERR_IO; }
  PSB_HIP(hipSetDevice(s->ctx->device));
  // spheres: same parameters and count, state from the file
  if (S) {
    std::vector<float> c3(3 * (size_t)S, 0.f);
    PSB_TRY(psb_system_set_spheres(s, S, c3.data(), nullptr, &sp));
    if (s->per_sphere != h.per_sphere) { set_error("checkpoint sphere lattice differs"); return PSB_ERR_INVALID_ARGUMENT; }
  } else if (s->n_spheres) {
    PSB_TRY(psb_system_set_spheres(s, 0, nullptr, nullptr, nullptr));
  }
  hipStream_t st = s->ctx->stream;
  size_t off = 0;
  for (size_t i = 0; i < lay.size(); ++i) {
    if (!(mask >> i & 1u)) continue;
    void* dst = lay[i].first;
    if (i == lay.size() - 3 && S) dst = s->d_sc[0];
    if (i == lay.size() - 2 && S) dst = s->d_sv[0];
    if (i == lay.size() - 1 && S) dst = s->d_sf;
    if (dst && lay[i].second) PSB_HIP(hipMemcpyAsync(dst, buf.data() + off, lay[i].second, hipMemcpyHostToDevice, st));
    off += lay[i].second;
  }
  PSB_HIP(hipStreamSynchronize(st));
  if (S) {
    // host copies of the current centres / velocities (a later set_particles
    // re-appends the spheres there)
    const float* c4 = (const float*)(buf.data() + total - 48 * (size_t)S);
    const float* v4 = c4 + 4 * (size_t)S;
    for (uint32_t j = 0; j < S; ++j)
      for (int a = 0; a < 3; ++a) {
        s->sphere_c[3 * j + a] = c4[4 * j + a];
        s->sphere_v[3 * j + a] = v4[4 * j + a];
      }
    s->sphere_cur = 0;
  }
  s->n = h.n;
  s->n_fluid = h.n_fluid;
  s->c.n_fluid = S ? h.n_fluid : kNoBodies;
  s->step_counter = h.step_counter;
  s->stepped = h.stepped != 0;
  s->stats = stats;
  s->has_state = true;
  s->nmat = 0;   // per-particle materials refer to the previous particles
  return PSB_OK;
}

psb_status psb_system_set_recorder(psb_system s, psb_dataset d, uint32_t stride) {
  PSB_ARG(s, "null system");
  if (d) {
    PSB_ARG(stride > 0, "stride must be > 0");
    PSB_ARG(s->prm.mode == PSB_MODE_CONFIG && s->prm.solver == PSB_SOLVER_PCISPH,
            "recording needs a PSB_MODE_CONFIG PCISPH system");
    PSB_TRY(ensure_features(s));
  }
  s->rec = d;
  s->rec_stride = stride ? stride : 1;
  return PSB_OK;
}

}  // extern "C"
