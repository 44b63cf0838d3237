// mixed.h -- per-particle materials of the reference-mode WCSPH and PCISPH
// steps (internal).
//
// The reference packs every particle's material into per-particle arrays
// (mass, interactionRadius, gasStiffness, restDensity, viscosity, threshold,
// surfaceTension; SphSystem.cpp:157-167, Emitter.cpp:141-175) and mixes them
// per pair: the WCSPH kernels take the neighbour's radius and mass
// (SphKernel.cu:17-21, 51-71), the PCISPH kernels max(h_i, h_j) and the
// neighbour's mass (PciSphSystem_Kernel.cu:51, 79, 101, 138), both the mean
// viscosity (mu_i + mu_j) / 2 (SphKernel.cu:62, PciSphSystem_Kernel.cu:144),
// and the neighbour search keeps j when |x_i - x_j| <= h_i
// (UniformGridKernel.cu:82).  Here a system holds a table of at most
// kMaxMaterials materials and one material index per particle; every pair
// constant is computed once on the host (in double, rounded to float) for
// each (material of i, material of j).
#pragma once
#include "psb_internal.h"
#include "pcisph.h"

namespace psb {

constexpr int kMaxMaterials = PSB_MAX_MATERIALS;
// threads per particle in the mixed PCISPH corrector passes (a power of two
// dividing 64: a particle's lanes share one wavefront)
constexpr uint32_t kMixLanes = 4;

// The particle's own material.
struct MatOwn {
  float h2;        // its search radius squared (the list keeps d^2 <= h2)
  float rho0, inv_rho0, k;
  float l, sigma;  // surface threshold and tension (WCSPH)
  float beta;      // 2 (m dt / rho0)^2 (PCISPH pressure update)
  float pad;
};

// Constants of a pair (material a of i, material b of j).  hk is h_b
// (WCSPH) or max(h_a, h_b) (PCISPH); m_b the neighbour's mass.
struct MatPair {
  float h, h2, h2x3;   // hk, hk^2, 3 hk^2
  float mc;            // m_b 315 / (64 pi hk^9)
  float cs;            // -45 / (pi hk^6)
  float csm;           // m_b (-45 / (pi hk^6))
  float cvm;           // m_b 45 / (pi hk^6)
  float csf;           // m_b (-945 / (32 pi hk^9))
  float mu;            // (mu_a + mu_b) / 2
  float pad[3];
};

struct MatTable {
  MatOwn own[kMaxMaterials];
  MatPair pair[kMaxMaterials * kMaxMaterials];   // [a * kMaxMaterials + b]
};

// Validates the materials against the system's parameters and fills the
// table.  Every material's support_radius must be positive and at most the
// system's (the neighbour search radius); solver PCISPH takes max(h_a, h_b).
psb_status make_mat_table(const psb_sph_params& sys, const psb_sph_params* mats, uint32_t count,
                          MatTable* out);

struct MixArgs {
  uint32_t n;
  Consts c;                           // box, colliders, dt, gravity of the system
  const MatTable* tbl;                // device copy
  uint32_t* cnt; uint32_t* list;      // 32-bit ELL list (filtered in place)
  const uint8_t* mat_by_id;           // material of each particle id
  uint8_t* ms;                        // material of each sorted particle
  float4* ax; float4* av; const float4* avh; const uint32_t* aid;
  float4* bx; float4* bv; float4* bvh; uint32_t* bid; float* bp;
  float* rho; float* pres; float* rho_pred;
  float4* fa; float4* fs;
  uint32_t* flags;
  PciStatus* st;
  ErrSlots es;                        // per-iteration error slots (PCISPH systems)
};

// After the neighbour search over the system radius: each particle's
// material, and its list cut to d^2 <= h2 of that material (order kept).
psb_status launch_mix_prepare(const MixArgs& a, hipStream_t s);
// densityEvaluation + internalForces with per-particle materials, gravity,
// leapfrog, box, colliders, interpolation -> B state.
psb_status launch_mix_wcsph(const MixArgs& a, hipStream_t s);
// compute_Density_PCI + compute_Viscosity_SurfaceTension_Force_PCI and the
// first prediction into xp.
psb_status launch_mix_pci_prologue(const MixArgs& a, float4* xp, hipStream_t s);
// compute_Density_PCI + compute_Pressure_PCI on predicted positions X
// (writes X.w = p / rho^2); iteration it's maximum error into st->err[it].
psb_status launch_mix_pci_density(const MixArgs& a, float4* X, int it, hipStream_t s);
// compute_Pressure_Force_PCI + integrateSPH_LeapFrog_Forces from the start
// state: the next prediction Xn, or (fin) the end-of-step B state and the
// step's statistics (iterations = iters, error of iteration iters - 1).
psb_status launch_mix_pci_force(const MixArgs& a, const float4* X, float4* Xn, bool fin,
                                int iters, hipStream_t s);

}  // namespace psb
