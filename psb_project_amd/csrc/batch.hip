// batch.hip -- many independent PCISPH scenes advanced together.
//
// Layout: the scenes' particles are concatenated (scene-major) in every
// state array, and their grids are concatenated in one global cell space
// (scene s owns cells [cell_off_s, cell_off_s + nx ny nz)).  One counting
// sort over the global cells therefore keeps every scene's particles in its
// own range, sorted by its own cells.  Every per-particle phase is one launch
// over all scenes with workgroups aligned to scene boundaries (BatchMap): a
// workgroup serves exactly one scene, so the corrector's error reduction,
// early exit and convergence decision stay per scene, exactly as in the
// single-scene step, and results equal the scenes run one by one.
//
// Host control per step: the prologue (status reset, grid build, neighbour
// search + density, epilogue, non-pressure forces and the first
// min_iterations corrector iterations) is queued at once, from a HIP graph
// when graphs are on.  Later iterations are queued one ahead: the force pass
// of the last scene to converge stores the iteration into host-mapped
// memory, the host stops queueing, and speculative launches of finished
// scenes return at their first instruction.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mlp.h"
#include "pcisph.h"
#include "psb_internal.h"

namespace psb {

psb_status grid_create(psb_context ctx, const psb_grid_desc& d, uint32_t capacity, psb_grid* out);

enum BatchPhase { BP_HASH, BP_NEIGHBORS, BP_EPILOGUE, BP_NONPRESSURE, BP_DENSITY, BP_FORCE,
                  BP_COUNT };
static_assert(BP_COUNT == PSB_BATCH_PHASES, "phase count");

}  // namespace psb

using namespace psb;

struct psb_batch_s {
  psb_context ctx = nullptr;
  psb_sph_params prm{};
  Consts c{};
  GridDev gd{};                       // material grid (stencil width)
  uint32_t S = 0, n = 0;
  std::vector<BatchScene> hsc;
  BatchScene* dsc = nullptr;
  uint16_t* dblk[kBatchBlockSizes] = {};
  uint32_t nblk[kBatchBlockSizes] = {};
  int bsel[BP_COUNT] = {2, 2, 2, 2, 2, 2};   // 256 threads until tuned
  int lanes[BP_COUNT] = {1, 1, 1, 1, 1, 1};  // threads per particle (D and F: 1 or 2)
  psb_grid grid = nullptr;
  float4 *ax = nullptr, *av = nullptr, *avh = nullptr, *bx = nullptr, *bv = nullptr, *bvh = nullptr;
  uint32_t *aid = nullptr, *bid = nullptr;
  float *ap = nullptr, *bp = nullptr, *rho = nullptr, *pres = nullptr, *rho_pred = nullptr;
  float4 *anp = nullptr, *xp0 = nullptr, *xp1 = nullptr;
  float* dl = nullptr;
  float* feat = nullptr;
  uint32_t* ncount = nullptr;
  uint32_t* nlist = nullptr;
  uint16_t* nlist16 = nullptr;    // compact twin of nlist for the corrector passes
  uint32_t* flags = nullptr;          // [0] flags, [1] max neighbours
  PciStatus* st = nullptr;
  ErrSlots es{};                  // spread per-iteration error maxima            // [S]
  uint32_t* done = nullptr;
  int32_t* h_final = nullptr;         // host-mapped: [0] iteration at which every scene finished,
                                      // [1] the last force pass that made its decision
  int32_t* d_final = nullptr;
  PciStatus* h_stat = nullptr;        // pinned [2][S] status copies (latency, stats)
  psb_mlp mlp = nullptr;
  std::vector<char> has_state;
  // per-scene statistics of the last update
  std::vector<psb_step_stats> stats;
  uint32_t last_flags = 0, last_maxnb = 0;
  // graphs
  bool use_graphs = false;
  hipGraphExec_t g_pro = nullptr;
  std::vector<hipGraphExec_t> g_iter;
  // events: step start and per-iteration ends, double-buffered by step parity
  hipEvent_t ev_start[2] = {};
  hipEvent_t ev_copy[2] = {};
  std::vector<hipEvent_t> ev_iter[2];
  // latency: recorded on every lat_every-th step (0: off)
  int lat_every = 0;
  bool lat_on = false;
  int pending = -1;                   // parity of a step whose latencies are not yet read
  std::vector<double> lat_ms;
  // phase timing (autotune)
  bool tuning = false;
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  struct TRec { int phase; size_t e; };
  std::vector<TRec> trecs;
  double t_phase[BP_COUNT] = {};
  int64_t step_counter = 0;
  // retries (params.retry_levels): per level, the scenes being redone as
  // substeps and their block tables (scene records with blk0 for the subset)
  struct RetrySet {
    std::vector<uint32_t> scenes;
    BatchScene* dsc = nullptr;
    uint32_t* dlist = nullptr;
    uint16_t* dblk[kBatchBlockSizes] = {};
    uint32_t nblk[kBatchBlockSizes] = {};
  };
  RetrySet rs[kMaxRetryLevels + 1];
  std::vector<int32_t> retried, substeps;   // per scene, over the current update call
  std::vector<char> lat_retried[2];         // scenes redone in the step of each parity
  hipEvent_t ev_retry_end[2] = {};
  std::vector<hipEvent_t> ev_riter;         // iteration ends of retry attempts
  PciStatus* h_rstat = nullptr;             // pinned [S] status copy for the retry decision
};

namespace psb {

// block size 64 << bsel threads; particles per block (64 << bsel) / lanes,
// whose block table has index bsel - (lanes - 1)
static int btable(psb_batch b, int phase) { return b->bsel[phase] - (b->lanes[phase] - 1); }

static BatchMap bmap(psb_batch b, int phase) {
  const int t = btable(b, phase);
  return BatchMap{b->dsc, b->dblk[t], t};
}

static Shape bshape(psb_batch b, int phase) {
  return Shape{dim3(b->nblk[btable(b, phase)]), dim3(64u << b->bsel[phase])};
}

struct BTimed {
  psb_batch b;
  int phase;
  size_t e = 0;
  bool on;
  BTimed(psb_batch bb, int ph) : b(bb), phase(ph), on(bb->tuning) {
    if (!on) return;
    if (b->tev_used + 2 > b->tev.size()) {
      for (int i = 0; i < 256; ++i) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) { on = false; return; }
        b->tev.push_back(ev);
      }
    }
    e = b->tev_used;
    b->tev_used += 2;
    hipEventRecord(b->tev[e], b->ctx->stream);
  }
  ~BTimed() {
    if (!on) return;
    hipEventRecord(b->tev[e + 1], b->ctx->stream);
    b->trecs.push_back({phase, e});
  }
};

static psb_status collect_tuning(psb_batch b) {
  if (b->trecs.empty()) return PSB_OK;
  PSB_HIP(hipStreamSynchronize(b->ctx->stream));
  for (const auto& r : b->trecs) {
    float ms = 0.f;
    PSB_HIP(hipEventElapsedTime(&ms, b->tev[r.e], b->tev[r.e + 1]));
    b->t_phase[r.phase] += ms;
  }
  b->trecs.clear();
  b->tev_used = 0;
  return PSB_OK;
}

// predicted-position buffer read by iteration k's density pass
static float4* pred_buf(psb_batch b, int k) {
  const int warm = b->prm.warm_start == PSB_WARM_PREVIOUS ? 1 : 0;
  return ((k + warm) & 1) ? b->xp1 : b->xp0;
}

static void buffers(psb_batch b, PciBuffers& pb, PciOut& o) {
  pb = PciBuffers{};
  pb.anp = b->anp; pb.xp0 = b->xp0; pb.dl = b->dl;
  pb.feat = b->mlp ? b->feat : nullptr;
  pb.ax = b->ax; pb.avh = b->avh; pb.aid = b->aid;
  pb.list16 = b->nlist16; pb.st = b->st;
  o = PciOut{};
  o.bx = b->bx; o.bv = b->bv; o.bvh = b->bvh; o.bid = b->bid; o.bp = b->bp; o.flags = b->flags;
  o.host_final = b->d_final;
  o.host_progress = b->d_final + 1;
  o.done = b->done;
  o.num_scenes = b->S;
  o.es = b->es;
}

static psb_status enqueue_iter(psb_batch b, int k) {
  hipStream_t st = b->ctx->stream;
  const bool mlp = b->prm.warm_start == PSB_WARM_MLP;
  psb_neighbors nb{b->ncount, b->nlist, (uint32_t)b->prm.max_neighbors, b->n};
  PciBuffers pb; PciOut o;
  buffers(b, pb, o);
  float4* cur = pred_buf(b, k);
  float4* nxt = pred_buf(b, k + 1);
  {
    BTimed t(b, BP_DENSITY);
    const int dflags = (k == 0 && mlp) ? (kDensWriteFeature | kDensSkipUpdate) : 0;
    PSB_TRY(launch_pci_density_batch(bmap(b, BP_DENSITY), bshape(b, BP_DENSITY), b->n, b->c,
                                     nb.count, nb.list, cur, b->pres, b->rho_pred, b->dl, b->st,
                                     k, b->feat, dflags, st, b->lanes[BP_DENSITY], b->nlist16,
                                     b->es));
  }
  if (k == 0 && mlp) PSB_TRY(mlp_infer(b->mlp, b->feat, b->n, b->pres, cur, 1, b->c.pscale, st));
  {
    BTimed t(b, BP_FORCE);
    PSB_TRY(launch_pci_force_batch(bmap(b, BP_FORCE), bshape(b, BP_FORCE), b->n, b->c, nb.count,
                                   nb.list, cur, nxt, b->rho_pred, b->pres, pb, b->st, k, 0, o,
                                   st, b->lanes[BP_FORCE], b->nlist16));
  }
  return PSB_OK;
}

static psb_status enqueue_prologue(psb_batch b) {
  hipStream_t st = b->ctx->stream;
  const bool warm = b->prm.warm_start == PSB_WARM_PREVIOUS;
  const bool mlp = b->prm.warm_start == PSB_WARM_MLP;
  {
    // the per-step status reset rides on the hashing pass
    BTimed t(b, BP_HASH);
    ReorderArrays ra;
    ra.sx = b->bx; ra.sv = b->bv; ra.svh = b->bvh; ra.sid = b->bid; ra.sp = b->bp;
    ra.dx = b->ax; ra.dv = b->av; ra.dvh = b->avh; ra.did = b->aid; ra.dp = b->ap;
    const int s = b->bsel[BP_HASH];
    const StepReset reset{b->st, b->done, b->es, b->S};
    PSB_TRY(grid_build_batch(b->grid, bmap(b, BP_HASH), b->nblk[s], 64u << s, b->bx, b->n, &ra,
                             st, &reset));
  }
  psb_neighbors nb{b->ncount, b->nlist, (uint32_t)b->prm.max_neighbors, b->n};
  {
    BTimed t(b, BP_NEIGHBORS);
    const int s = b->bsel[BP_NEIGHBORS];
    Compact16 c16;
    c16.list16 = b->nlist16;
    c16.wide = &b->st[0].wide;
    c16.wide_stride = sizeof(PciStatus) / sizeof(uint32_t);
    // with the density epilogue inside (the epilogue phase launches nothing)
    const EpiArgs ep = make_epilogue(b->c, 1, warm ? 1 : 0, mlp ? 1 : 0, b->pres, b->ap, b->ax,
                                     b->av, nullptr);
    PSB_TRY(launch_neighbors_batch(b->grid, bmap(b, BP_NEIGHBORS), b->nblk[s], 64u << s, b->ax,
                                   b->n, b->c.h2, b->gd.R, nb, b->rho, b->c.mc, b->flags,
                                   b->flags + 1, st, c16, b->es, &ep));
  }
  PciBuffers pb; PciOut o;
  buffers(b, pb, o);
  {
    BTimed t(b, BP_NONPRESSURE);
    PSB_TRY(launch_pci_nonpressure_batch(bmap(b, BP_NONPRESSURE), bshape(b, BP_NONPRESSURE), b->n,
                                         b->c, b->prm.surface_tension != 0.f, warm, nb.count,
                                         nb.list, b->ax, b->av, b->rho, b->avh, b->pres, pb, o,
                                         st));
  }
  if (warm) {
    BTimed t(b, BP_FORCE);
    PSB_TRY(launch_pci_force_batch(bmap(b, BP_FORCE), bshape(b, BP_FORCE), b->n, b->c, nb.count,
                                   nb.list, b->xp0, b->xp1, b->rho_pred, b->pres, pb, b->st, -1, 1,
                                   o, st, b->lanes[BP_FORCE], b->nlist16));
  }
  for (int k = 0; k < b->c.min_it; ++k) PSB_TRY(enqueue_iter(b, k));
  return PSB_OK;
}

static void drop_graphs(psb_batch b) {
  if (b->g_pro) hipGraphExecDestroy(b->g_pro);
  b->g_pro = nullptr;
  for (auto g : b->g_iter) if (g) hipGraphExecDestroy(g);
  b->g_iter.clear();
}

template <class F>
static psb_status capture(psb_batch b, hipGraphExec_t* out, F&& enqueue) {
  hipStream_t st = b->ctx->stream;
  PSB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  psb_status r = enqueue();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (r != PSB_OK) { if (g) hipGraphDestroy(g); return r; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  return PSB_OK;
}

static psb_status run_prologue(psb_batch b) {
  if (!b->use_graphs) return enqueue_prologue(b);
  if (!b->g_pro) PSB_TRY(capture(b, &b->g_pro, [&] { return enqueue_prologue(b); }));
  PSB_HIP(hipGraphLaunch(b->g_pro, b->ctx->stream));
  return PSB_OK;
}

static psb_status run_iter(psb_batch b, int k) {
  if (!b->use_graphs) return enqueue_iter(b, k);
  if ((int)b->g_iter.size() <= k) b->g_iter.resize(k + 1, nullptr);
  if (!b->g_iter[k]) PSB_TRY(capture(b, &b->g_iter[k], [&] { return enqueue_iter(b, k); }));
  PSB_HIP(hipGraphLaunch(b->g_iter[k], b->ctx->stream));
  return PSB_OK;
}

// latencies of the step whose events carry parity `par`
static psb_status read_latency(psb_batch b, int par) {
  PSB_HIP(hipEventSynchronize(b->ev_copy[par]));
  // ev_copy precedes the step's retries; a redone scene's latency ends at
  // ev_retry_end, which may still be ahead on the stream
  const std::vector<char>& redo = b->lat_retried[par];
  if (std::find(redo.begin(), redo.end(), (char)1) != redo.end())
    PSB_HIP(hipEventSynchronize(b->ev_retry_end[par]));
  const PciStatus* hs = b->h_stat + (size_t)par * b->S;
  for (uint32_t s = 0; s < b->S; ++s) {
    if (!b->lat_retried[par].empty() && b->lat_retried[par][s]) {
      // redone as substeps: the step ends with the retries
      float ms = 0.f;
      PSB_HIP(hipEventElapsedTime(&ms, b->ev_start[par], b->ev_retry_end[par]));
      b->lat_ms.push_back(ms);
      continue;
    }
    const int f = hs[s].final_iter;
    if (f < 0 || f >= (int)b->ev_iter[par].size()) continue;
    float ms = 0.f;
    PSB_HIP(hipEventElapsedTime(&ms, b->ev_start[par], b->ev_iter[par][f]));
    b->lat_ms.push_back(ms);
  }
  return PSB_OK;
}

// ---------------------------------------------------------------------------
// Retries: a scene whose corrector ends above the tolerance is redone from
// its start state as two substeps of dt/2 (params.retry_levels deep), exactly
// as the single-system step does, with the other scenes left as they are.
// The launches cover only the redone scenes (their own block tables); the
// grid build of a second half runs over every scene (one counting sort over
// the global cells), which re-sorts the finished scenes' end states into A
// without changing them.
// ---------------------------------------------------------------------------
// Level constants start from make_consts, which leaves no collider or body
// behind (a batch has neither), so the final force pass of a retry walks no
// collider list.
static psb_status batch_level_consts(psb_batch b, int level, Consts* out) {
  if (level == 0) { *out = b->c; return PSB_OK; }
  psb_sph_params p = b->prm;
  const float scale = std::ldexp(1.0f, -level);
  p.dt = b->prm.dt * scale;
  if (p.delta > 0.f) p.delta = b->prm.delta / (scale * scale);
  // the corrector under-relaxed with the step: a step that failed over-
  // corrects as readily at any dt (the pressure's displacement scales with
  // delta dt^2), so each level halves delta on top of params.relaxation
  p.relaxation = (b->prm.relaxation > 0.f ? b->prm.relaxation : 1.0f) * scale;
  PSB_TRY(make_consts(p, out, nullptr));
  out->retry_ok = level < b->prm.retry_levels ? 1 : 0;
  out->diverge_stop = 1;
  if (level == b->prm.retry_levels) out->max_it = retry_max_iterations(b->prm);
  return PSB_OK;
}

static psb_status retry_set(psb_batch b, int level, const std::vector<uint32_t>& scenes) {
  psb_batch_s::RetrySet& R = b->rs[level];
  auto A = [&](void** p, size_t bytes) -> psb_status { PSB_HIP(hipMalloc(p, bytes)); return PSB_OK; };
  if (!R.dsc) {
    PSB_TRY(A((void**)&R.dsc, sizeof(BatchScene) * b->S));
    PSB_TRY(A((void**)&R.dlist, 4 * (size_t)b->S));
    for (int t = 0; t < kBatchBlockSizes; ++t) PSB_TRY(A((void**)&R.dblk[t], 2 * (size_t)b->nblk[t]));
  }
  R.scenes = scenes;
  std::vector<BatchScene> hs(b->hsc);
  for (int t = 0; t < kBatchBlockSizes; ++t) {
    const uint32_t bs = 64u << t;
    std::vector<uint16_t> tab;
    for (uint32_t s : scenes) {
      hs[s].blk0[t] = (uint32_t)tab.size();
      tab.insert(tab.end(), (hs[s].n + bs - 1) / bs, (uint16_t)s);
    }
    R.nblk[t] = (uint32_t)tab.size();
    PSB_HIP(hipMemcpy(R.dblk[t], tab.data(), 2 * tab.size(), hipMemcpyHostToDevice));
  }
  PSB_HIP(hipMemcpy(R.dsc, hs.data(), sizeof(BatchScene) * b->S, hipMemcpyHostToDevice));
  PSB_HIP(hipMemcpy(R.dlist, scenes.data(), 4 * scenes.size(), hipMemcpyHostToDevice));
  return PSB_OK;
}

static BatchMap rmap(psb_batch b, const psb_batch_s::RetrySet& R, int phase) {
  const int t = btable(b, phase);
  return BatchMap{R.dsc, R.dblk[t], t};
}

static Shape rshape(psb_batch b, const psb_batch_s::RetrySet& R, int phase) {
  return Shape{dim3(R.nblk[btable(b, phase)]), dim3(64u << b->bsel[phase])};
}

// Scenes of `scenes` whose last attempt ended above the tolerance.
static psb_status failed_scenes(psb_batch b, const std::vector<uint32_t>& scenes,
                                std::vector<uint32_t>* out) {
  PSB_HIP(hipMemcpyAsync(b->h_rstat, b->st, sizeof(PciStatus) * b->S, hipMemcpyDeviceToHost,
                         b->ctx->stream));
  PSB_HIP(hipStreamSynchronize(b->ctx->stream));
  out->clear();
  for (uint32_t s : scenes) {
    float e;
    std::memcpy(&e, &b->h_rstat[s].last_err, 4);
    if (e > b->prm.tolerance) out->push_back(s);
  }
  return PSB_OK;
}

// One plain-corrector attempt of dt / 2^level for the scenes of rs[level]:
// from their start state A (the first half of a redo) or from B after a
// grid build (the second half).
static psb_status retry_attempt(psb_batch b, int level, bool from_a) {
  const psb_batch_s::RetrySet& R = b->rs[level];
  const uint32_t cnt = (uint32_t)R.scenes.size();
  Consts c{};
  PSB_TRY(batch_level_consts(b, level, &c));
  hipStream_t st = b->ctx->stream;
  psb_neighbors nb{b->ncount, b->nlist, (uint32_t)b->prm.max_neighbors, b->n};
  PciBuffers pb; PciOut o;
  buffers(b, pb, o);
  pb.feat = nullptr;
  o.num_scenes = cnt;
  if (!from_a) {
    ReorderArrays ra;
    ra.sx = b->bx; ra.sv = b->bv; ra.svh = b->bvh; ra.sid = b->bid; ra.sp = b->bp;
    ra.dx = b->ax; ra.dv = b->av; ra.dvh = b->avh; ra.did = b->aid; ra.dp = b->ap;
    const int sh = b->bsel[BP_HASH];
    const StepReset reset{b->st, b->done, b->es, b->S};
    PSB_TRY(grid_build_batch(b->grid, bmap(b, BP_HASH), b->nblk[sh], 64u << sh, b->bx, b->n, &ra,
                             st, &reset));
    const int sn = b->bsel[BP_NEIGHBORS];
    Compact16 c16;
    c16.list16 = b->nlist16;
    c16.wide = &b->st[0].wide;
    c16.wide_stride = sizeof(PciStatus) / sizeof(uint32_t);
    const EpiArgs ep = make_epilogue(c, 1, 0, 0, b->pres, b->ap, b->ax, b->av, nullptr);
    PSB_TRY(launch_neighbors_batch(b->grid, rmap(b, R, BP_NEIGHBORS), R.nblk[sn], 64u << sn,
                                   b->ax, b->n, c.h2, b->gd.R, nb, b->rho, c.mc, b->flags,
                                   b->flags + 1, st, c16, b->es, &ep));
  } else {
    PSB_TRY(launch_pci_retry_reset(b->st, st, b->es, R.dlist, cnt));
    PSB_HIP(hipMemsetAsync(b->done, 0, 4, st));
    for (uint32_t s : R.scenes)
      PSB_HIP(hipMemsetAsync(b->pres + b->hsc[s].p0, 0, 4 * (size_t)b->hsc[s].n, st));
  }
  PSB_TRY(launch_pci_nonpressure_batch(rmap(b, R, BP_NONPRESSURE), rshape(b, R, BP_NONPRESSURE),
                                       b->n, c, b->prm.surface_tension != 0.f, false, nb.count,
                                       nb.list, b->ax, b->av, b->rho, b->avh, b->pres, pb, o, st));
  b->h_final[0] = b->h_final[1] = -1;
  for (int k = 0; k < c.max_it; ++k) {
    float4* cur = (k & 1) ? b->xp1 : b->xp0;
    float4* nxt = (k & 1) ? b->xp0 : b->xp1;
    PSB_TRY(launch_pci_density_batch(rmap(b, R, BP_DENSITY), rshape(b, R, BP_DENSITY), b->n, c,
                                     nb.count, nb.list, cur, b->pres, b->rho_pred, b->dl, b->st,
                                     k, nullptr, 0, st, b->lanes[BP_DENSITY], b->nlist16, b->es));
    PSB_TRY(launch_pci_force_batch(rmap(b, R, BP_FORCE), rshape(b, R, BP_FORCE), b->n, c,
                                   nb.count, nb.list, cur, nxt, b->rho_pred, b->pres, pb, b->st, k,
                                   0, o, st, b->lanes[BP_FORCE], b->nlist16));
    if (k + 1 >= c.min_it) PSB_HIP(hipEventRecord(b->ev_riter[k], st));
    if (k >= c.min_it) {
      PSB_HIP(hipEventSynchronize(b->ev_riter[k - 1]));
      if (__atomic_load_n(b->h_final, __ATOMIC_ACQUIRE) >= 0) break;
    }
  }
  PSB_HIP(hipStreamSynchronize(st));
  return PSB_OK;
}

static psb_status retry_redo(psb_batch b, const std::vector<uint32_t>& scenes, int level);

static psb_status retry_half(psb_batch b, const std::vector<uint32_t>& scenes, int level,
                             bool from_a) {
  PSB_TRY(retry_set(b, level, scenes));
  PSB_TRY(retry_attempt(b, level, from_a));
  if (level >= b->prm.retry_levels) return PSB_OK;
  std::vector<uint32_t> failed;
  PSB_TRY(failed_scenes(b, scenes, &failed));
  if (!failed.empty()) PSB_TRY(retry_redo(b, failed, level + 1));
  return PSB_OK;
}

// Redo `scenes` (failed at level - 1) as two substeps of dt / 2^level.
static psb_status retry_redo(psb_batch b, const std::vector<uint32_t>& scenes, int level) {
  for (uint32_t s : scenes) b->substeps[s] += 2;
  PSB_TRY(retry_half(b, scenes, level, true));
  PSB_TRY(retry_half(b, scenes, level, false));
  return PSB_OK;
}

// Until force pass `it` has made its decision (the first workgroup of every
// unfinished scene stores `it` into host-mapped memory) or every scene has
// finished: host polling instead of an event after every pass, whose
// completion signal held the next density pass back.  A queue that stays
// slow past 200 ms is waited for with a stream synchronisation.
static psb_status wait_progress(psb_batch b, int it) {
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spin = 1;; ++spin) {
    if (__atomic_load_n(b->h_final + 1, __ATOMIC_ACQUIRE) >= it ||
        __atomic_load_n(b->h_final, __ATOMIC_ACQUIRE) >= 0)
      return PSB_OK;
    if ((spin & 4095u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
      PSB_HIP(hipStreamSynchronize(b->ctx->stream));
      return PSB_OK;
    }
    __builtin_ia32_pause();
  }
}

static psb_status batch_step(psb_batch b, int64_t t) {
  hipStream_t st = b->ctx->stream;
  const int par = (int)(t & 1);
  const int min_it = b->c.min_it, max_it = b->c.max_it;
  // latency steps time every iteration with events; the others pace the
  // queue by the host-mapped progress word
  const bool lat = b->lat_on && b->lat_every > 0 && t % b->lat_every == 0;
  if (lat) PSB_HIP(hipEventRecord(b->ev_start[par], st));
  b->h_final[0] = b->h_final[1] = -1;
  PSB_TRY(run_prologue(b));
  if (lat) PSB_HIP(hipEventRecord(b->ev_iter[par][min_it - 1], st));
  if (b->lat_on && b->pending >= 0) {
    PSB_TRY(read_latency(b, b->pending));
    b->pending = -1;
  }
  int final_known = -1;
  for (int k = min_it; k < max_it; ++k) {
    PSB_TRY(run_iter(b, k));
    // iteration k is queued; wait for k-1 and stop queueing once every scene finished
    if (lat) {
      PSB_HIP(hipEventRecord(b->ev_iter[par][k], st));
      PSB_HIP(hipEventSynchronize(b->ev_iter[par][k - 1]));
    } else {
      PSB_TRY(wait_progress(b, k - 1));
    }
    const int32_t f = __atomic_load_n(b->h_final, __ATOMIC_ACQUIRE);
    if (f >= 0) { final_known = f; break; }
  }
  // finished by the last queued iteration: let its host-mapped store land
  // before the next step resets the flag
  if (final_known < 0) PSB_HIP(hipStreamSynchronize(st));
  if (lat) {
    // (before any retry: a second half's grid build resets every scene's
    // final iteration)
    PSB_HIP(hipMemcpyAsync(b->h_stat + (size_t)par * b->S, b->st, sizeof(PciStatus) * b->S,
                           hipMemcpyDeviceToHost, st));
    PSB_HIP(hipEventRecord(b->ev_copy[par], st));
    b->pending = par;
  }
  if (b->prm.retry_levels > 0) {
    std::vector<uint32_t> all(b->S), failed;
    for (uint32_t s = 0; s < b->S; ++s) all[s] = s;
    PSB_TRY(failed_scenes(b, all, &failed));
    if (lat) b->lat_retried[par].assign(b->S, 0);
    if (!failed.empty()) {
      for (uint32_t s : failed) {
        b->retried[s] += 1;
        if (lat) b->lat_retried[par][s] = 1;
      }
      PSB_TRY(retry_redo(b, failed, 1));
    }
    if (lat) PSB_HIP(hipEventRecord(b->ev_retry_end[par], st));
  }
  if (b->tuning) PSB_TRY(collect_tuning(b));
  return PSB_OK;
}

static psb_status batch_alloc(psb_batch b) {
  const size_t N = b->n;
  auto A = [&](void** p, size_t bytes) -> psb_status { PSB_HIP(hipMalloc(p, bytes)); return PSB_OK; };
  PSB_TRY(A((void**)&b->ax, 16 * N)); PSB_TRY(A((void**)&b->av, 16 * N)); PSB_TRY(A((void**)&b->avh, 16 * N));
  PSB_TRY(A((void**)&b->bx, 16 * N)); PSB_TRY(A((void**)&b->bv, 16 * N)); PSB_TRY(A((void**)&b->bvh, 16 * N));
  PSB_TRY(A((void**)&b->aid, 4 * N)); PSB_TRY(A((void**)&b->bid, 4 * N));
  PSB_TRY(A((void**)&b->ap, 4 * N)); PSB_TRY(A((void**)&b->bp, 4 * N));
  PSB_TRY(A((void**)&b->rho, 4 * N)); PSB_TRY(A((void**)&b->pres, 4 * N));
  PSB_TRY(A((void**)&b->rho_pred, 4 * N));
  PSB_TRY(A((void**)&b->anp, 16 * N)); PSB_TRY(A((void**)&b->dl, 4 * N));
  PSB_TRY(A((void**)&b->xp0, 16 * N)); PSB_TRY(A((void**)&b->xp1, 16 * N));
  PSB_TRY(A((void**)&b->ncount, 4 * N));
  PSB_TRY(A((void**)&b->nlist, 4 * N * (size_t)b->prm.max_neighbors));
  PSB_TRY(A((void**)&b->nlist16, l16_bytes(N, (uint32_t)b->prm.max_neighbors)));
  PSB_TRY(A((void**)&b->flags, 64));
  PSB_TRY(A((void**)&b->st, sizeof(PciStatus) * b->S));
  b->es.stride = err_slot_stride(b->S);
  PSB_TRY(A((void**)&b->es.base, 4 * (size_t)kErrSlots * b->es.stride));
  PSB_TRY(A((void**)&b->done, 64));
  PSB_TRY(A((void**)&b->dsc, sizeof(BatchScene) * b->S));
  PSB_HIP(hipHostMalloc((void**)&b->h_final, 64, hipHostMallocMapped | hipHostMallocCoherent));
  PSB_HIP(hipHostGetDevicePointer((void**)&b->d_final, b->h_final, 0));
  b->h_final[0] = b->h_final[1] = -1;
  PSB_HIP(hipHostMalloc((void**)&b->h_stat, 2 * sizeof(PciStatus) * b->S, hipHostMallocDefault));
  PSB_HIP(hipHostMalloc((void**)&b->h_rstat, sizeof(PciStatus) * b->S, hipHostMallocDefault));
  for (int p = 0; p < 2; ++p) PSB_HIP(hipEventCreate(&b->ev_retry_end[p]));
  for (int k = 0; k < kMaxIterSlots; ++k) {
    hipEvent_t e;
    PSB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->ev_riter.push_back(e);
  }
  hipStream_t st = b->ctx->stream;
  PSB_HIP(hipMemsetAsync(b->flags, 0, 64, st));
  PSB_HIP(hipMemsetAsync(b->st, 0, sizeof(PciStatus) * b->S, st));
  PSB_HIP(hipMemsetAsync(b->done, 0, 64, st));
  PSB_HIP(hipMemcpyAsync(b->dsc, b->hsc.data(), sizeof(BatchScene) * b->S, hipMemcpyHostToDevice, st));
  for (int i = 0; i < kBatchBlockSizes; ++i) {
    const uint32_t bs = 64u << i;
    std::vector<uint16_t> tab;
    for (uint32_t s = 0; s < b->S; ++s) {
      const uint32_t nbk = (b->hsc[s].n + bs - 1) / bs;
      tab.insert(tab.end(), nbk, (uint16_t)s);
    }
    b->nblk[i] = (uint32_t)tab.size();
    PSB_TRY(A((void**)&b->dblk[i], 2 * tab.size()));
    PSB_HIP(hipMemcpyAsync(b->dblk[i], tab.data(), 2 * tab.size(), hipMemcpyHostToDevice, st));
    PSB_HIP(hipStreamSynchronize(st));
  }
  for (int p = 0; p < 2; ++p) {
    PSB_HIP(hipEventCreate(&b->ev_start[p]));
    PSB_HIP(hipEventCreateWithFlags(&b->ev_copy[p], hipEventDisableTiming));
    for (int k = 0; k < kMaxIterSlots; ++k) {
      hipEvent_t e;
      PSB_HIP(hipEventCreate(&e));
      b->ev_iter[p].push_back(e);
    }
  }
  PSB_HIP(hipStreamSynchronize(st));
  return PSB_OK;
}

static psb_status fetch_stats(psb_batch b) {
  hipStream_t st = b->ctx->stream;
  uint32_t fl[2] = {0, 0}, gf = 0;
  PSB_HIP(hipMemcpyAsync(fl, b->flags, 8, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(&gf, b->grid->flags, 4, hipMemcpyDeviceToHost, st));
  std::vector<PciStatus> hs(b->S);
  PSB_HIP(hipMemcpyAsync(hs.data(), b->st, sizeof(PciStatus) * b->S, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  b->last_flags = fl[0] | gf;
  b->last_maxnb = fl[1];
  for (uint32_t s = 0; s < b->S; ++s) {
    psb_step_stats& o = b->stats[s];
    o.flags = b->last_flags;
    o.iterations = hs[s].last_iters;
    std::memcpy(&o.max_error, &hs[s].last_err, 4);
    o.total_iterations = (int64_t)hs[s].total_iters;
    std::memcpy(&o.worst_error, &hs[s].worst_err, 4);
    o.max_neighbors_seen = (int32_t)b->last_maxnb;
    o.retried_steps = b->retried[s];
    o.substeps = b->substeps[s];
  }
  return PSB_OK;
}

}  // namespace psb

extern "C" {

psb_status psb_batch_create(psb_context ctx, const psb_sph_params* material, uint32_t num_scenes,
                            const psb_scene_desc* scenes, psb_batch* out) {
  PSB_ARG(ctx && material && scenes && out, "null argument");
  *out = nullptr;
  PSB_ARG(num_scenes >= 1 && num_scenes <= 65535, "1 <= num_scenes <= 65535");
  PSB_ARG(material->mode == PSB_MODE_CONFIG && material->solver == PSB_SOLVER_PCISPH,
          "batches run PSB_MODE_CONFIG PCISPH");
  psb_batch b = new psb_batch_s();
  b->ctx = ctx;
  b->prm = *material;
  b->S = num_scenes;
  psb_status stt = make_consts(*material, &b->c, &b->gd);
  if (stt != PSB_OK) { delete b; return stt; }
  const double h = material->support_radius;
  const double s_lat = material->lattice_spacing > 0.f ? (double)material->lattice_spacing : 0.5 * h;
  uint64_t cells = 0, parts = 0;
  for (uint32_t i = 0; i < num_scenes && stt == PSB_OK; ++i) {
    const psb_scene_desc& d = scenes[i];
    if (d.count < 1) { set_error("every scene needs at least one particle"); stt = PSB_ERR_INVALID_ARGUMENT; break; }
    if (d.grid.cell_size != material->grid.cell_size || d.grid.dims[0] <= 0 || d.grid.dims[1] <= 0 ||
        d.grid.dims[2] <= 0) {
      set_error("scene grids need positive dims and the material's cell size");
      stt = PSB_ERR_INVALID_ARGUMENT;
      break;
    }
    BatchScene sc{};
    sc.grid = grid_dev(d.grid, (float)h);
    if (sc.grid.R != b->gd.R) { set_error("scene stencil width differs"); stt = PSB_ERR_INVALID_ARGUMENT; break; }
    for (int a = 0; a < 3; ++a) {
      sc.box_lo[a] = (float)(d.box_min[a] + material->particle_radius);
      sc.box_hi[a] = (float)(d.box_max[a] - material->particle_radius);
      sc.wall_o[a] = d.box_min[a];
      sc.wall_n[a] = (int)std::floor(((double)d.box_max[a] - d.box_min[a]) / s_lat + 1e-3);
    }
    sc.cell_off = (uint32_t)cells;
    sc.p0 = (uint32_t)parts;
    sc.n = d.count;
    cells += (uint64_t)d.grid.dims[0] * d.grid.dims[1] * d.grid.dims[2];
    parts += d.count;
    if (cells >= (1ull << 31) || parts > kMaxParticles) {
      set_error("batch exceeds 2^31 cells or 2^27 particles");
      stt = PSB_ERR_INVALID_ARGUMENT;
      break;
    }
    b->hsc.push_back(sc);
  }
  if (stt != PSB_OK) { delete b; return stt; }
  for (int i = 0; i < kBatchBlockSizes; ++i) {
    const uint32_t bs = 64u << i;
    uint32_t acc = 0;
    for (auto& sc : b->hsc) { sc.blk0[i] = acc; acc += (sc.n + bs - 1) / bs; }
  }
  b->n = (uint32_t)parts;
  b->has_state.assign(num_scenes, 0);
  b->stats.assign(num_scenes, psb_step_stats{});
  b->retried.assign(num_scenes, 0);
  b->substeps.assign(num_scenes, 0);
  b->c.retry_ok = material->retry_levels > 0 ? 1 : 0;
  {
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) stt = hip_fail(e, "hipSetDevice", __FILE__, __LINE__);
  }
  if (stt == PSB_OK) {
    psb_grid_desc gdesc{};
    gdesc.cell_size = 1.f;
    gdesc.dims[0] = (int32_t)cells; gdesc.dims[1] = 1; gdesc.dims[2] = 1;
    stt = grid_create(ctx, gdesc, b->n, &b->grid);
  }
  if (stt == PSB_OK) stt = batch_alloc(b);
  if (stt != PSB_OK) { psb_batch_destroy(b); return stt; }
  *out = b;
  return PSB_OK;
}

psb_status psb_batch_destroy(psb_batch b) {
  if (!b) return PSB_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  drop_graphs(b);
  void* ptrs[] = {b->ax, b->av, b->avh, b->bx, b->bv, b->bvh, b->aid, b->bid, b->ap, b->bp,
                  b->rho, b->pres, b->rho_pred, b->anp, b->dl, b->xp0, b->xp1, b->feat,
                  b->ncount, b->nlist, b->nlist16, b->flags, b->st, b->done, b->dsc, b->es.base};
  for (void* p : ptrs) if (p) hipFree(p);
  for (auto p : b->dblk) if (p) hipFree(p);
  if (b->h_final) hipHostFree(b->h_final);
  if (b->h_stat) hipHostFree(b->h_stat);
  if (b->h_rstat) hipHostFree(b->h_rstat);
  for (auto& R : b->rs) {
    if (R.dsc) hipFree(R.dsc);
    if (R.dlist) hipFree(R.dlist);
    for (auto p : R.dblk) if (p) hipFree(p);
  }
  for (int p = 0; p < 2; ++p) if (b->ev_retry_end[p]) hipEventDestroy(b->ev_retry_end[p]);
  for (auto e : b->ev_riter) hipEventDestroy(e);
  for (int p = 0; p < 2; ++p) {
    if (b->ev_start[p]) hipEventDestroy(b->ev_start[p]);
    if (b->ev_copy[p]) hipEventDestroy(b->ev_copy[p]);
    for (auto e : b->ev_iter[p]) hipEventDestroy(e);
  }
  for (auto e : b->tev) hipEventDestroy(e);
  if (b->grid) psb_finalize_grid(b->grid);
  delete b;
  return PSB_OK;
}

uint32_t psb_batch_num_scenes(psb_batch b) { return b ? b->S : 0; }

psb_status psb_batch_set_particles(psb_batch b, uint32_t scene, const float* pos4,
                                   const float* vel4, const float* vh4) {
  PSB_ARG(b && pos4, "null argument");
  PSB_ARG(scene < b->S, "scene index out of range");
  PSB_HIP(hipSetDevice(b->ctx->device));
  const BatchScene& sc = b->hsc[scene];
  const uint32_t n = sc.n, p0 = sc.p0;
  hipStream_t st = b->ctx->stream;
  std::vector<float> zeros4(4 * (size_t)n, 0.f);
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  const size_t b4 = 16 * (size_t)n;
  const float* v = vel4 ? vel4 : zeros4.data();
  const float* vh = vh4 ? vh4 : zeros4.data();
  PSB_HIP(hipMemcpyAsync(b->bx + p0, pos4, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->ax + p0, pos4, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->bv + p0, v, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->av + p0, v, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->bvh + p0, vh, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->avh + p0, vh, b4, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->bid + p0, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemcpyAsync(b->aid + p0, ids.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
  PSB_HIP(hipMemsetAsync(b->bp + p0, 0, 4 * (size_t)n, st));
  PSB_HIP(hipMemsetAsync(b->ap + p0, 0, 4 * (size_t)n, st));
  PSB_HIP(hipMemsetAsync(b->pres + p0, 0, 4 * (size_t)n, st));
  PSB_HIP(hipStreamSynchronize(st));
  b->has_state[scene] = 1;
  b->stats[scene] = psb_step_stats{};
  return PSB_OK;
}

psb_status psb_batch_set_mlp(psb_batch b, psb_mlp m) {
  PSB_ARG(b, "null batch");
  if (m && !b->feat) PSB_HIP(hipMalloc(&b->feat, 8 * 4 * (size_t)b->n));
  b->mlp = m;
  drop_graphs(b);   // graphs capture the network's arguments
  return PSB_OK;
}

psb_status psb_batch_set_graphs(psb_batch b, int32_t on) {
  PSB_ARG(b, "null batch");
  drop_graphs(b);
  b->use_graphs = on != 0;
  return PSB_OK;
}

psb_status psb_batch_set_latency(psb_batch b, int32_t every) {
  PSB_ARG(b && every >= 0, "bad argument");
  b->lat_every = every;
  b->lat_on = every != 0;
  b->pending = -1;
  b->lat_ms.clear();
  return PSB_OK;
}

psb_status psb_batch_get_latency(psb_batch b, double* p50_ms, double* p99_ms, double* max_ms,
                                 uint64_t* samples) {
  PSB_ARG(b, "null batch");
  if (b->lat_on && b->pending >= 0) {
    PSB_TRY(read_latency(b, b->pending));
    b->pending = -1;
  }
  std::vector<double> v = b->lat_ms;
  std::sort(v.begin(), v.end());
  auto q = [&](double f) {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, (size_t)std::ceil(f * v.size()) - (f > 0 ? 1 : 0));
    return v[i];
  };
  if (p50_ms) *p50_ms = q(0.5);
  if (p99_ms) *p99_ms = q(0.99);
  if (max_ms) *max_ms = v.empty() ? 0.0 : v.back();
  if (samples) *samples = v.size();
  return PSB_OK;
}

psb_status psb_batch_update(psb_batch b, int32_t steps) {
  PSB_ARG(b && steps >= 0, "bad argument");
  for (uint32_t s = 0; s < b->S; ++s)
    if (!b->has_state[s]) { set_error("psb_batch_update before every scene's particles were set"); return PSB_ERR_STATE; }
  if (b->prm.warm_start == PSB_WARM_MLP) {
    if (!b->mlp) { set_error("PSB_WARM_MLP without psb_batch_set_mlp"); return PSB_ERR_STATE; }
    PSB_ARG(b->c.min_it >= 2, "PSB_WARM_MLP needs min_iterations >= 2");
  }
  PSB_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  PSB_HIP(hipMemsetAsync(b->flags, 0, 8, st));
  PSB_HIP(hipMemsetAsync(b->grid->flags, 0, 4, st));
  PSB_TRY(launch_pci_reset(b->st, true, st, b->S, b->done, b->es));
  std::fill(b->retried.begin(), b->retried.end(), 0);
  std::fill(b->substeps.begin(), b->substeps.end(), 0);
  for (int k = 0; k < steps; ++k) PSB_TRY(batch_step(b, b->step_counter++));
  return fetch_stats(b);
}

psb_status psb_batch_get_stats(psb_batch b, uint32_t scene, psb_step_stats* out) {
  PSB_ARG(b && out && scene < b->S, "bad argument");
  *out = b->stats[scene];
  return PSB_OK;
}

psb_status psb_batch_get_particles(psb_batch b, uint32_t scene, float* pos4, float* vel4,
                                   float* vel_half4, float* density, float* pressure,
                                   uint32_t* ids) {
  PSB_ARG(b && scene < b->S, "bad argument");
  PSB_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  const BatchScene& sc = b->hsc[scene];
  const size_t n = sc.n, p0 = sc.p0;
  if (pos4) PSB_HIP(hipMemcpyAsync(pos4, b->bx + p0, 16 * n, hipMemcpyDeviceToHost, st));
  if (vel4) PSB_HIP(hipMemcpyAsync(vel4, b->bv + p0, 16 * n, hipMemcpyDeviceToHost, st));
  if (vel_half4) PSB_HIP(hipMemcpyAsync(vel_half4, b->bvh + p0, 16 * n, hipMemcpyDeviceToHost, st));
  if (density) PSB_HIP(hipMemcpyAsync(density, b->rho_pred + p0, 4 * n, hipMemcpyDeviceToHost, st));
  if (pressure) PSB_HIP(hipMemcpyAsync(pressure, b->bp + p0, 4 * n, hipMemcpyDeviceToHost, st));
  if (ids) PSB_HIP(hipMemcpyAsync(ids, b->bid + p0, 4 * n, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  return PSB_OK;
}

psb_status psb_batch_set_block_sizes(psb_batch b, const int32_t* block_sizes) {
  PSB_ARG(b && block_sizes, "null argument");
  int sel[BP_COUNT], ln[BP_COUNT];
  for (int p = 0; p < BP_COUNT; ++p) {
    const int32_t v = block_sizes[p];
    ln[p] = v < 0 ? 2 : 1;
    const int32_t t = v < 0 ? -v : v;
    sel[p] = -1;
    for (int k = 0; k < kBatchBlockSizes; ++k)
      if (t == (64 << k)) sel[p] = k;
    PSB_ARG(sel[p] >= 0, "block sizes must be 64, 128, 256 or 512");
    PSB_ARG(ln[p] == 1 || ((p == BP_DENSITY || p == BP_FORCE) && sel[p] >= 1),
            "two threads per particle only for pci_density / pci_force at 128..512 threads");
  }
  for (int p = 0; p < BP_COUNT; ++p) { b->bsel[p] = sel[p]; b->lanes[p] = ln[p]; }
  drop_graphs(b);
  return PSB_OK;
}

psb_status psb_batch_time_corrector(psb_batch b, int32_t reps, double* us_density,
                                    double* us_force) {
  PSB_ARG(b && reps > 0 && us_density && us_force, "bad argument");
  if (b->step_counter == 0) { set_error("psb_batch_time_corrector before a step"); return PSB_ERR_STATE; }
  PSB_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  // iteration 0 of a corrector that never finalises, every scene active: the
  // force pass writes the next prediction only
  Consts c = b->c;
  c.min_it = 2;
  c.max_it = kMaxIterSlots - 1;
  c.retry_ok = 0;
  c.diverge_stop = 0;
  PciBuffers pb; PciOut o;
  buffers(b, pb, o);
  pb.feat = nullptr;
  o.host_final = nullptr;
  o.host_progress = nullptr;
  o.done = nullptr;
  psb_neighbors nb{b->ncount, b->nlist, (uint32_t)b->prm.max_neighbors, b->n};
  PSB_TRY(launch_pci_retry_reset(b->st, st, b->es, nullptr, b->S));
  hipEvent_t e[3];
  for (auto& x : e) PSB_HIP(hipEventCreate(&x));
  auto run = [&]() -> psb_status {
    PSB_HIP(hipEventRecord(e[0], st));
    for (int r = 0; r < reps; ++r)
      PSB_TRY(launch_pci_density_batch(bmap(b, BP_DENSITY), bshape(b, BP_DENSITY), b->n, c,
                                       nb.count, nb.list, b->xp0, b->pres, b->rho_pred, b->dl,
                                       b->st, 0, nullptr, 0, st, b->lanes[BP_DENSITY], b->nlist16,
                                       b->es));
    PSB_HIP(hipEventRecord(e[1], st));
    for (int r = 0; r < reps; ++r)
      PSB_TRY(launch_pci_force_batch(bmap(b, BP_FORCE), bshape(b, BP_FORCE), b->n, c, nb.count,
                                     nb.list, b->xp0, b->xp1, b->rho_pred, b->pres, pb, b->st, 0, 0,
                                     o, st, b->lanes[BP_FORCE], b->nlist16));
    PSB_HIP(hipEventRecord(e[2], st));
    PSB_HIP(hipEventSynchronize(e[2]));
    float md = 0.f, mf = 0.f;
    PSB_HIP(hipEventElapsedTime(&md, e[0], e[1]));
    PSB_HIP(hipEventElapsedTime(&mf, e[1], e[2]));
    *us_density = 1e3 * (double)md / reps;
    *us_force = 1e3 * (double)mf / reps;
    return PSB_OK;
  };
  const psb_status stt = run();
  for (auto& x : e) hipEventDestroy(x);
  return stt;
}

psb_status psb_batch_autotune(psb_batch b, int32_t steps, int32_t* block_sizes) {
  PSB_ARG(b && steps >= 1, "bad argument");
  for (uint32_t s = 0; s < b->S; ++s)
    if (!b->has_state[s]) { set_error("psb_batch_autotune before every scene's particles were set"); return PSB_ERR_STATE; }
  PSB_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = b->ctx->stream;
  // snapshot of the end-of-step state every step starts from
  const size_t N = b->n;
  void* save = nullptr;
  const size_t bytes = 48 * N + 8 * N;
  PSB_HIP(hipMalloc(&save, bytes));
  char* sv = (char*)save;
  auto snap = [&](bool restore) -> psb_status {
    struct { void* p; size_t sz; } arr[] = {{b->bx, 16 * N}, {b->bv, 16 * N}, {b->bvh, 16 * N},
                                            {b->bid, 4 * N}, {b->bp, 4 * N}};
    size_t off = 0;
    for (auto& a : arr) {
      if (restore) PSB_HIP(hipMemcpyAsync(a.p, sv + off, a.sz, hipMemcpyDeviceToDevice, st));
      else PSB_HIP(hipMemcpyAsync(sv + off, a.p, a.sz, hipMemcpyDeviceToDevice, st));
      off += a.sz;
    }
    return PSB_OK;
  };
  psb_status r = snap(false);
  const bool graphs = b->use_graphs;
  const bool lat = b->lat_on;
  b->use_graphs = false;
  b->lat_on = false;
  drop_graphs(b);
  // candidates (threads per particle, block size index): every block size
  // with one thread per particle, and 128..512 threads with two for the
  // corrector passes (other phases run the one-thread variant again)
  const int kCand = kBatchBlockSizes + kBatchBlockSizes - 1;
  auto cand_lanes = [&](int cand, int p) {
    return (cand >= kBatchBlockSizes && (p == BP_DENSITY || p == BP_FORCE)) ? 2 : 1;
  };
  auto cand_bsel = [&](int cand) { return cand < kBatchBlockSizes ? cand : cand - kBatchBlockSizes + 1; };
  double best[BP_COUNT];
  int pick[BP_COUNT];
  for (int p = 0; p < BP_COUNT; ++p) { best[p] = 1e300; pick[p] = 2; }
  for (int cand = 0; cand < kCand && r == PSB_OK; ++cand) {
    for (int p = 0; p < BP_COUNT; ++p) { b->bsel[p] = cand_bsel(cand); b->lanes[p] = cand_lanes(cand, p); }
    r = snap(true);
    if (r != PSB_OK) break;
    std::memset(b->t_phase, 0, sizeof(b->t_phase));
    b->tuning = true;
    for (int k = 0; k < steps && r == PSB_OK; ++k) r = batch_step(b, k);
    b->tuning = false;
    if (r == PSB_OK) r = collect_tuning(b);
    for (int p = 0; p < BP_COUNT; ++p)
      if (b->t_phase[p] < best[p]) { best[p] = b->t_phase[p]; pick[p] = cand; }
  }
  b->tuning = false;
  b->trecs.clear();
  b->tev_used = 0;
  for (int p = 0; p < BP_COUNT; ++p) { b->bsel[p] = cand_bsel(pick[p]); b->lanes[p] = cand_lanes(pick[p], p); }
  if (r == PSB_OK) r = snap(true);
  if (r == PSB_OK) {
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) r = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
  }
  hipFree(save);
  if (r == PSB_OK) r = launch_pci_reset(b->st, true, st, b->S, b->done, b->es);
  b->use_graphs = graphs;
  b->lat_on = lat;
  if (block_sizes)   // threads per workgroup; negative: two threads per particle
    for (int p = 0; p < BP_COUNT; ++p)
      block_sizes[p] = (b->lanes[p] == 2 ? -1 : 1) * (64 << b->bsel[p]);
  return r;
}

}  // extern "C"
