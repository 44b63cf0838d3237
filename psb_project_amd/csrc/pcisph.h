// pcisph.h -- PCISPH device structures and launchers (internal).
#pragma once
#include "psb_internal.h"

namespace psb {

// Device-side state of one PCISPH solve.  The first 16 bytes are copied to
// the host to learn whether (and when) the corrector finished.
struct PciStatus {
  int32_t final_iter;              // -1 until a force pass finalised the step
  int32_t last_iters;              // corrector iterations of the last step
  uint32_t last_err;               // float bits: max compression error, last step
  uint32_t worst_err;              // float bits: largest final error since reset
  unsigned long long total_iters;  // iterations since reset
  uint32_t wide;                   // 1: some compact neighbour offset overflowed this step
  // 1 once a step's compact offsets overflowed (set by the next step's
  // reset, then kept): the neighbour search also writes the grouped 32-bit
  // list, which the corrector passes read while `wide` is set
  uint32_t want32;
  // float bits: smallest error of iterations 0..k, written by F(k) into
  // [k & 1] and read by F(k + 1) from there (each launch reads the other slot)
  uint32_t min_err[2];
  uint32_t err[kMaxIterSlots];     // float bits: max error of D(k)
};

// The corrector of a retry substep (config mode, p = 0 start) stops as
// failed once its error, after min_iterations, exceeds both kDivergeFactor
// times the smallest error of its earlier iterations and kDivergeFloor times
// the tolerance.  First attempts keep iterating to max_iterations: a warm
// start whose first error is small may bounce above twice it and converge.
constexpr float kDivergeFactor = 2.0f;
constexpr float kDivergeFloor = 4.0f;

// Per-step reset of one scene's status and error slots, by one workgroup.
__device__ __forceinline__ void pci_reset_scene(PciStatus* st0, uint32_t scene, uint32_t* done,
                                                ErrSlots es, bool keep_wide = false) {
  PciStatus* st = st0 + scene;
  for (int k = threadIdx.x; k < kMaxIterSlots; k += blockDim.x) st->err[k] = 0u;
  if (es.base)
    for (int k = threadIdx.x; k < kErrSlots * kMaxIterSlots; k += blockDim.x)
      es.base[(size_t)(k / kMaxIterSlots) * es.stride + (size_t)scene * kMaxIterSlots +
              (k % kMaxIterSlots)] = 0u;
  if (threadIdx.x == 0) {
    st->final_iter = -1;
    if (!keep_wide) {
      if (st->wide) st->want32 = 1u;
      st->wide = 0u;
    }
    st->min_err[0] = st->min_err[1] = 0x7F800000u;   // +inf
  }
  if (done && scene == 0 && threadIdx.x == 0) *done = 0u;
}

// The per-step reset as a job for the first kernel of the step (the grid
// build's hashing pass runs it on its first `scenes` workgroups).
struct StepReset {
  PciStatus* st;
  uint32_t* done;
  ErrSlots es;
  uint32_t scenes;
};

struct PciBuffers {
  float4* fa;    // reference mode: Facc + f_viscosity (force density)
  float4* fs;    // reference mode: f_surface
  float4* anp;   // config mode: non-pressure acceleration
  float4* xp0;   // first predicted positions
  float* dl;     // config mode: per-particle pressure scaling delta_i
  float* feat;   // warm-start features [8][n], or null when not computed
  const float4* ax;   // start-of-step positions
  const float4* avh;  // start-of-step half-step velocities
  const uint32_t* aid;
  // compact neighbour list for the non-pressure pass, used for the scenes
  // whose wide flag (st[scene].wide) is clear; null: the 32-bit list
  const uint16_t* list16;
  const PciStatus* st;
};

// Added to the iteration a force pass stores into PciOut.host_final when the
// step ends above the tolerance (config mode), so the host learns both when
// and whether the corrector finished.
constexpr int32_t kFinalFailed = 1 << 16;

struct PciOut {
  // fused step: end-of-step state
  float4* bx; float4* bv; float4* bvh; uint32_t* bid; float* bp;
  // standalone pci_SPH_update outputs
  float4* pos; float4* old_pos; float4* vav; float4* vap;
  float4* fP; float4* normals; float4* fV; float4* fS; float4* F;
  uint32_t* flags;
  // optional host-mapped slot that receives the final iteration index
  int32_t* host_final;
  // optional host-mapped word: the first workgroup of force pass k stores k
  // (after host_final, when k is final), so the host can pace its queue of
  // iterations without recording an event after every pass
  int32_t* host_progress;
  // batches: scenes finished this step; host_final is written by the last
  uint32_t* done;
  uint32_t num_scenes;
  ErrSlots es;
};

// Kernel launch shape (grid, workgroup).
struct Shape {
  dim3 grid, block;
};

// Scratch bytes of psb_pci_sph_update for n particles: seven float4 arrays
// (start-state copies, forcesAccum, surface force, non-pressure acceleration,
// two predictions), the start pressure and delta (2 x 4 B), the status
// aligned to 256 B.
constexpr size_t pci_update_scratch_bytes(size_t n) {
  return 7 * 16 * n + 8 * n + sizeof(PciStatus) + 256;
}
// per-device static scratch of psb_pci_sph_update_noalloc, looked up once per
// device with that device current (psb_init)
psb_status static_scratch_setup(int device);

// one status per scene (scenes = 1 for a single system); done: batch counter
psb_status launch_pci_reset(PciStatus* st, bool stats, hipStream_t s, uint32_t scenes = 1,
                            uint32_t* done = nullptr, ErrSlots es = ErrSlots{});
// reset of a step redone from its start state: the neighbour lists (and the
// compact list's overflow flag) of that state stay valid.  scenes: device
// list of `count` scene indices (null: scenes 0 .. count-1)
psb_status launch_pci_retry_reset(PciStatus* st, hipStream_t s, ErrSlots es,
                                  const uint32_t* scenes = nullptr, uint32_t count = 1);
psb_status launch_pci_nonpressure(uint32_t n, const Consts& c, bool surf, bool warm, bool out,
                                  const uint32_t* cnt, const uint32_t* list, const float4* xs,
                                  const float4* vs, const float* rho, const float4* avh,
                                  const float* pres, const PciBuffers& b, const PciOut& o,
                                  hipStream_t s);
// density-pass flags (config mode)
constexpr int kDensWriteFeature = 1;   // write feature 0 (first-prediction density error)
constexpr int kDensSkipUpdate = 2;     // leave the pressure to the MLP
psb_status launch_pci_density(uint32_t n, const Consts& c, const uint32_t* cnt,
                              const uint32_t* list, float4* X, float* pres, float* rho_pred,
                              const float* dl, PciStatus* st, int k, float* feat, int dflags,
                              hipStream_t s, const uint32_t* ids = nullptr, int lanes = 1,
                              const uint16_t* list16 = nullptr, ErrSlots es = ErrSlots{},
                              const HaloTile* tiles = nullptr, const uint32_t* halo_src = nullptr);
psb_status launch_record(const float* feat, const float* pres, uint32_t n, uint32_t stride,
                         uint32_t offset, float inv_ps, float* dfeat, float* dtarget,
                         uint64_t dcap, uint64_t base, uint32_t count, hipStream_t s);
psb_status launch_pci_force(uint32_t n, const Consts& c, bool step, const uint32_t* cnt,
                            const uint32_t* list, const float4* X, float4* Xn,
                            const float* rho_pred, const float* pres, const PciBuffers& b,
                            PciStatus* st, int k, int warm_pass, const PciOut& o, hipStream_t s,
                            int lanes = 1, const uint16_t* list16 = nullptr,
                            const HaloTile* tiles = nullptr, const uint32_t* halo_src = nullptr);

// batches (scene-aligned blocks; see BatchMap)
psb_status launch_pci_nonpressure_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                        const Consts& c, bool surf, bool warm,
                                        const uint32_t* cnt, const uint32_t* list,
                                        const float4* xs, const float4* vs, const float* rho,
                                        const float4* avh, const float* pres,
                                        const PciBuffers& b, const PciOut& o, hipStream_t s);
psb_status launch_pci_density_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                    const Consts& c, const uint32_t* cnt, const uint32_t* list,
                                    float4* X, float* pres, float* rho_pred, const float* dl,
                                    PciStatus* st, int k, float* feat, int dflags, hipStream_t s,
                                    int lanes = 1, const uint16_t* list16 = nullptr,
                                    ErrSlots es = ErrSlots{});
psb_status launch_pci_force_batch(const BatchMap& map, const Shape& sh, uint32_t n,
                                  const Consts& c, const uint32_t* cnt, const uint32_t* list,
                                  const float4* X, float4* Xn, const float* rho_pred,
                                  const float* pres, const PciBuffers& b, PciStatus* st, int k,
                                  int warm_pass, const PciOut& o, hipStream_t s, int lanes = 1,
                                  const uint16_t* list16 = nullptr);

// sph.hip
psb_status launch_density_list(uint32_t n, const Consts& c, psb_neighbors nb, const float4* pos,
                               const float4* vel, float* rho, float* pres, float4* xs, float4* vs,
                               hipStream_t s);
// the epilogue's arguments, for the fused path of launch_neighbors
EpiArgs make_epilogue(const Consts& c, int pcisph, int warm, int xw_prev, float* pres,
                      const float* pprev, float4* xs, float4* vs, const uint32_t* ids);
psb_status launch_wcsph_step_forces(uint32_t n, const Consts& c, bool surf, const uint32_t* cnt,
                                    const uint32_t* list, const float4* xs, const float4* vs,
                                    const float* rho, const float* pres, const float4* avh,
                                    const uint32_t* aid, float4* bx, float4* bv, float4* bvh,
                                    uint32_t* bid, float* bp, uint32_t* flags, hipStream_t s);

}  // namespace psb
