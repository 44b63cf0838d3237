/*
 * psb_sph.h -- C-ABI of the MI355X (gfx950) SPH / PCISPH timestep.
 *
 * Drop-in boundary for the CUDA launch-wrapper layer of PSB_Project
 * (the `extern "C"` functions declared in src/common.cuh, UniformGrid.cuh,
 * SphSystem.cuh, PciSphSystem.cuh, BoxCollision.cuh, ForceExt.cuh and
 * CudaSystem.cuh).  Every function here names the reference wrapper it
 * replaces.  Differences that are deliberate:
 *
 *   - every function returns a psb_status instead of void, and nothing is
 *     left unchecked (the reference never calls cudaGetLastError);
 *   - device arrays are FP32.  A "vec4" argument is a device array of
 *     n float4 records (x, y, z, w; 16-byte stride, w unused on input
 *     unless stated) replacing the reference's FP64 double3 AoS arrays
 *     (24-byte stride); a scalar array is n floats;
 *   - all work is ordered on the context's own HIP stream (default
 *     priority) rather than the legacy default stream; functions that hand
 *     results to the host synchronise that stream themselves;
 *   - the per-particle material arrays of the reference (mass, radius, k,
 *     restDensity, viscosity, threshold, surfaceTension) become one material
 *     per system, passed once in psb_sph_params, or (reference mode) a table
 *     of up to PSB_MAX_MATERIALS materials with one index per particle
 *     (psb_system_set_materials);
 *   - capacity, grid domain and neighbour-list width are arguments instead of
 *     the constants 65,536 / 1 m^3 / 200 / 64 of the reference.
 *
 * No torch or HIP types appear in these signatures; a stream is exposed
 * only as an opaque pointer (psb_context_stream).
 */
#ifndef PSB_SPH_H
#define PSB_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes and per-step flags                                           */
/* ------------------------------------------------------------------------ */
typedef int32_t psb_status;
#define PSB_OK                    0
#define PSB_ERR_INVALID_ARGUMENT  1
#define PSB_ERR_HIP               2  /* a HIP runtime call failed          */
#define PSB_ERR_CAPACITY          3  /* n larger than the allocated capacity */
#define PSB_ERR_NO_DEVICE         4
#define PSB_ERR_OUT_OF_MEMORY     5
#define PSB_ERR_STATE             6  /* call order (e.g. step before set)   */
#define PSB_ERR_IO                7  /* a file could not be read / written,
                                        or is not a valid checkpoint        */

/* Bits of psb_step_stats.flags, raised by the kernels on the device.       */
#define PSB_FLAG_NEIGHBOR_OVERFLOW  0x1u  /* > max_neighbors within h; list truncated */
#define PSB_FLAG_OUT_OF_GRID        0x2u  /* a particle outside the grid domain (clamped to an edge cell) */
#define PSB_FLAG_NONFINITE          0x4u  /* NaN/Inf position or density */
#define PSB_FLAG_NOT_CONVERGED      0x8u  /* PCISPH hit max_iterations above tolerance */
#define PSB_FLAG_ZERO_DENSITY       0x10u /* rho <= 0 for some particle */

/* Text of the last error raised on this thread (never NULL). */
const char* psb_last_error(void);

/* ------------------------------------------------------------------------ */
/* Parameters                                                                */
/* ------------------------------------------------------------------------ */
#define PSB_MODE_REFERENCE  0  /* reference semantics: linear EOS, 3 fixed PCISPH
                                  iterations, per-particle delta, sigma forced to
                                  0.0728 and g = -9.81 y in PCISPH, r = 0 box    */
#define PSB_MODE_CONFIG     1  /* Tait EOS, tolerance loop, template delta,
                                  box restitution, pressure clamped at 0        */

#define PSB_SOLVER_WCSPH    0  /* SPHSystem::update      (SphSystem.cpp:273-297)   */
#define PSB_SOLVER_PCISPH   1  /* PCI_SPHSystem::update  (PciSphSystem.cpp:163-204) */

#define PSB_EOS_LINEAR      0  /* p = k (rho - rho0)            SphKernel.cu:25 */
#define PSB_EOS_TAIT        1  /* p = B ((rho/rho0)^gamma - 1)                  */

#define PSB_WARM_NONE       0  /* PCISPH starts from p = 0 (PciSphSystem_Kernel.cu:131) */
#define PSB_WARM_PREVIOUS   1  /* PCISPH starts from the previous step's pressure       */
#define PSB_WARM_MLP        2  /* PCISPH starts from the per-particle MLP prediction    */

typedef struct psb_grid_desc {
  float   min[3];     /* UniformGrid::Min (UniformGrid.cuh:18)            */
  float   cell_size;  /* UniformGrid::sizeCell                            */
  int32_t dims[3];    /* nbCellsX/Y/Z = ceil(extent / cell) (UniformGrid.cu:27-29) */
  int32_t pad;
} psb_grid_desc;

/* Material (SimulationData_SPHSystem, SimulationData_SPHSystem.h:10-62),
 * collider, external force and solver settings of one system.             */
typedef struct psb_sph_params {
  int32_t mode;             /* PSB_MODE_*   */
  int32_t solver;           /* PSB_SOLVER_* */
  int32_t eos;              /* PSB_EOS_*    */
  int32_t warm_start;       /* PSB_WARM_*   */
  float   mass;             /* particle mass m                               */
  float   support_radius;   /* h, SPHParticle::interactionRadius             */
  float   rest_density;     /* rho0                                          */
  float   gas_stiffness;    /* k                                             */
  float   viscosity;        /* mu                                            */
  float   surface_tension;  /* sigma (0 disables the surface term)           */
  float   threshold;        /* l, |n| threshold of the surface term          */
  float   particle_radius;  /* radius used by the box collider               */
  float   tait_gamma;       /* gamma of the Tait EOS (7)                     */
  float   tait_b;           /* B of the Tait EOS; 0 -> k * rho0 / gamma       */
  float   gravity[3];       /* ForceExt_Constante direction * amplitude      */
  float   dt;               /* time step                                     */
  int32_t box_enabled;      /* BoxCollision present                          */
  float   box_min[3];       /* Box::getMin()                                 */
  float   box_max[3];       /* Box::getMax()                                 */
  float   elast;            /* restitution (ignored in PSB_MODE_REFERENCE)   */
  psb_grid_desc grid;
  int32_t max_neighbors;    /* ELL width of the neighbour list (reference: 200) */
  float   tolerance;        /* PCISPH max compression error / rho0 (0.01)     */
  int32_t min_iterations;   /* PCISPH minimum corrector iterations (3)        */
  int32_t max_iterations;   /* PCISPH maximum corrector iterations (50)       */
  float   delta;            /* PCISPH template scaling; 0 -> computed from a
                               full cubic-lattice neighbourhood of spacing
                               lattice_spacing                                */
  float   lattice_spacing;  /* s, used for the template delta                 */
  int32_t delta_local_cap;  /* PSB_MODE_CONFIG: 1 -> per particle, use
                               min(delta, 1/(beta (|sum gradW|^2 + sum |gradW|^2)))
                               over its start-of-step neighbours, so particles
                               with one-sided neighbourhoods (walls without
                               boundary particles, free surface) are not
                               over-corrected; 0 -> template delta only       */
  int32_t wall_model;       /* PSB_WALL_*: PCISPH config mode only           */
  float   relaxation;       /* PCISPH config mode: delta = relaxation * template
                               delta (<= 0 means 1).  Below 1 under-relaxes
                               the corrector: a deep column on a rigid floor
                               (boundary-particle walls, ~128 layers) makes the
                               full template over-correct and diverge; 0.6
                               holds there.  The warm-start feature scale
                               keeps the unrelaxed delta.                     */
  int32_t retry_levels;     /* PCISPH config mode, fused system step: a step
                               whose corrector ends above the tolerance after
                               max_iterations is redone from its start state
                               as two steps of dt/2 (the plain corrector,
                               p = 0 start, delta halved as well), each of
                               which may be redone as two of dt/4 (delta
                               quartered), down to dt / 2^retry_levels; the
                               step still advances dt.  0 disables (the
                               reference's fixed step).  At most 4.  A
                               substep's corrector whose error, after
                               min_iterations, grows past twice the smallest
                               error of its earlier iterations and past four
                               times the tolerance stops there as failed (it
                               diverges: iterating on would only build
                               pressure).                                     */
} psb_sph_params;

#define PSB_WALL_CLAMP      0  /* box walls act only through the collider (clamp) */
#define PSB_WALL_LATTICE    1  /* the box walls are also lattice boundary
                                  particles: implicit particles of mass m at
                                  box_min + (i + 1/2) lattice_spacing for every
                                  lattice index i outside the box on some axis
                                  (ceil(h / spacing) + 1 layers), so a fluid
                                  lattice at rest keeps rho0 up to the walls,
                                  edges and corners.  They add to the density
                                  and to sum gradW of delta_i, and push with
                                  the particle's own pressure only,
                                  -m psi_b p_i / rho0^2 gradW with psi_b = m
                                  (Akinci et al. 2012); they are
                                  computed in registers, never stored.  The
                                  collider stays as a safety net              */

/* Fill *p with the reference defaults of SPHParticle / SimulationData_SPHSystem
 * (SphParticle.cpp:7-22): water, h = 0.0457, rho0 = 998.29, k = 3,
 * mu = 3.5, sigma = 0.0728, l = 7.065; PSB_MODE_REFERENCE, WCSPH, dt 0.01,
 * grid [-0.5,0.5]^3 with cell = h, 200 neighbours, PCISPH 3 iterations.    */
psb_status psb_default_params(psb_sph_params* p);

/* Derived constants the kernels use (exposed for tests and the host API). */
typedef struct psb_derived {
  float h2, c_poly6, c_spiky, c_visc, c_surf_lap, c_surf_grad;
  float mc_poly6;      /* mass * c_poly6                  */
  float beta;          /* 2 (m dt / rho0)^2                */
  float delta;         /* template PCISPH delta            */
  float tait_b;
  int32_t stencil;     /* R = ceil(h / cell) cells         */
} psb_derived;
psb_status psb_derive(const psb_sph_params* p, psb_derived* out);

/* ------------------------------------------------------------------------ */
/* L0: device and memory helpers  (common.cuh:1-25, common.cu:16-107)        */
/* ------------------------------------------------------------------------ */
typedef struct psb_context_s* psb_context;

/* cudaInit (common.cu:36-55): select a device and create the stream.       */
psb_status psb_init(int32_t device, psb_context* out);
psb_status psb_finalize(psb_context ctx);
/* getCudaEnabledDeviceCount (common.cu:16) */
psb_status psb_device_count(int32_t* count);
void*      psb_context_stream(psb_context ctx);  /* hipStream_t as void*    */

/* allocateArray / freeArray (common.cu:58-66) */
psb_status psb_allocate_array(psb_context ctx, void** dev, size_t bytes);
psb_status psb_free_array(psb_context ctx, void* dev);
/* copyArrayToDevice / copyArrayFromDevice / copyArrayDeviceToDevice
 * (common.cu:68-93): byte offset applies to the device side.              */
psb_status psb_copy_array_to_device(psb_context ctx, void* dev, const void* host,
                                    size_t offset, size_t bytes);
psb_status psb_copy_array_from_device(psb_context ctx, void* host, const void* dev,
                                      size_t offset, size_t bytes);
psb_status psb_copy_array_device_to_device(psb_context ctx, void* dst, const void* src,
                                           size_t offset, size_t bytes);
psb_status psb_memset_array(psb_context ctx, void* dev, int32_t value, size_t bytes);
/* threadSync (common.cu:95) */
psb_status psb_thread_sync(psb_context ctx);
/* computeGridSize (common.cu:102-107) with an explicit block size.        */
psb_status psb_compute_grid_size(uint32_t n, uint32_t block,
                                 uint32_t* num_blocks, uint32_t* num_threads);

/* ------------------------------------------------------------------------ */
/* Uniform grid  (UniformGrid.cuh:16-53, UniformGrid.cu:17-99)               */
/* ------------------------------------------------------------------------ */
typedef struct psb_grid_s* psb_grid;

/* createGrid (UniformGrid.cu:17-49): Min = centre - extent/2,
 * dims = ceil(extent / cell).  `capacity` is the largest particle count
 * psb_store_particles will be called with.                                */
psb_status psb_create_grid(psb_context ctx, double xc, double yc, double zc,
                           double length, double width, double depth,
                           float cell_size, uint32_t capacity, psb_grid* out);
/* Same with an explicit origin and cell counts. */
psb_status psb_create_grid_desc(psb_context ctx, const psb_grid_desc* desc,
                                uint32_t capacity, psb_grid* out);
/* _finalizeGrid (UniformGrid.cu:53-61) */
psb_status psb_finalize_grid(psb_grid grid);
psb_status psb_grid_get_desc(psb_grid grid, psb_grid_desc* out);

/* reinitCells + storeParticles (UniformGrid.cu:64-87, UniformGridKernel.cu:30-55).
 * Replaces the atomic 64-slot buckets by a stable counting sort:
 *   hash[i]  = cx + nx (cy + ny cz),  c* = floor((x* - Min*) / cell) clamped;
 *   perm[k]  = index of the k-th particle in (hash, index) order;
 *   start[c] = number of particles with hash < c, for c in [0, ncells];
 *   the particles of cell c are perm[start[c] .. start[c+1]).
 * pos4 is a device vec4 array.  Results stay on the device (get_arrays). */
psb_status psb_store_particles(psb_grid grid, const float* pos4, uint32_t n);
/* Device pointers owned by the grid (valid until the next store/finalize):
 * hash[n], perm[n], cell_start[ncells + 1]; flags = device uint32.        */
psb_status psb_grid_get_arrays(psb_grid grid, const uint32_t** hash,
                               const uint32_t** perm, const uint32_t** cell_start,
                               uint64_t* ncells);
/* Gather src[perm[k]] -> dst[k] (the "reordered particle arrays").
 * elem_floats = 1 (scalar arrays) or 4 (vec4 arrays).                      */
psb_status psb_grid_reorder(psb_grid grid, const float* src, float* dst,
                            uint32_t n, int32_t elem_floats);

/* Neighbour list of the reference's partVoisine (SphKernel.cuh:6-10), in ELL
 * layout: the k-th neighbour of particle i is list[k * n + i], k < count[i],
 * in ascending (sorted) index order, self included; at most `capacity`
 * entries per particle.                                                    */
typedef struct psb_neighbors {
  uint32_t* count;     /* device [n]             */
  uint32_t* list;      /* device [capacity * n]  */
  uint32_t  capacity;
  uint32_t  n;
} psb_neighbors;

/* searchNeighbooring (UniformGrid.cu:91-99, UniformGridKernel.cu:58-95):
 * for each particle of sorted_pos4 (ordered by the last psb_store_particles,
 * e.g. via psb_grid_reorder), list every j with |x_i - x_j|^2 <= radius^2.
 * Stencil = ceil(radius / cell) cells per side.  Raises
 * PSB_FLAG_NEIGHBOR_OVERFLOW in *flags_out (host) if a list was truncated. */
psb_status psb_search_neighbors(psb_grid grid, const float* sorted_pos4, uint32_t n,
                                float radius, psb_neighbors nb, uint32_t* flags_out);

/* ------------------------------------------------------------------------ */
/* SPH kernels on caller arrays  (SphSystem.cuh:8-17)                         */
/* ------------------------------------------------------------------------ */

/* evaluate_densities_forces (SphSystem.cu:27-42 -> SphKernel.cu:9-91):
 * density (poly6) + EOS, then pressure, viscosity and surface-tension force
 * densities and their sum forces_accum4 = f_viscosity + f_surface + f_pressure.
 * normals4, f_*4 may be NULL when not wanted.                               */
psb_status psb_evaluate_densities_forces(psb_context ctx, const psb_sph_params* prm,
                                         psb_neighbors nb, const float* pos4,
                                         const float* vel4, uint32_t n,
                                         float* density, float* pressure,
                                         float* normals4, float* f_pressure4,
                                         float* f_viscosity4, float* f_surface4,
                                         float* forces_accum4);

/* evaluate_ForceExt_Constante (ForceExt.cu:24-34 -> ForceExt_Constante.cu:15-24):
 * accum4[i] += (direction * amplitude) * weight[i]  (the reference passes the
 * density array as weight for SPH, SphSystem.cpp:312).                     */
psb_status psb_force_constante(psb_context ctx, uint32_t n, float* accum4,
                               const float* weight, const float direction[3],
                               float amplitude);

/* integrateSPHSystem (SphSystem.cu:6-14 -> SphKernel.cu:112-124): leapfrog
 *   vel_av = vel_ap; vel_ap += (F / rho) dt; old_pos = new_pos;
 *   new_pos += vel_ap dt.                                                    */
psb_status psb_integrate_sph(psb_context ctx, float* vel_inter_av4, float* vel_inter_ap4,
                             float* old_pos4, float* new_pos4, const float* forces4,
                             const float* density, float dt, uint32_t n);

/* interpolateSPHVelocities (SphSystem.cu:17-24 -> SphKernel.cu:128-135):
 *   old_vel = new_vel; new_vel = (vel_av + vel_ap) / 2.                      */
psb_status psb_interpolate_velocities(psb_context ctx, const float* vel_inter_av4,
                                      const float* vel_inter_ap4, float* old_vel4,
                                      float* new_vel4, uint32_t n);

/* collisionSystem_Box_CUDA (BoxCollision.cu:10-18 -> BoxCollision_Kernel.cu:7-34):
 * clamp new_pos into [Min + r, Max - r] per axis and remove (1 + elast) times
 * the velocity component along the normalised sum of the hit face normals.
 * The reference ignores elast (r = 0, :27); pass elast = 0 for that.
 * old_vel4 is accepted and unused, as in the reference.                     */
psb_status psb_collision_box(psb_context ctx, float* new_pos4, const float* old_vel4,
                             float* new_vel4, float radius, float dt, uint32_t n,
                             const float box_min[3], const float box_max[3],
                             int32_t container, float elast);

/* Static colliders besides the box, on the same integrate -> collide seam.
 *   PSB_COLLIDER_SPHERE  SphereCollision (SphereCollision_Kernel.cu:5-80):
 *       container: particles beyond `radius` from `origin` go back to the
 *       surface; obstacle (container 0): particles inside go out to it.
 *   PSB_COLLIDER_PLANE   PlaneCollision (PlaneCollision_Kernel.cu:6-80): an
 *       axis-aligned quadrilateral `corners` A B C D with unit `normal`;
 *       particles on or behind it inside the corners' bounds go to where
 *       their step crossed it.
 * Both then set v <- v - (1 + r)(v.n) n, r = elast |x - x_hit| / (dt |v|).
 * Differences from the reference: FP32; the wrapper's swapped arguments
 * (SphereCollision.cu:15-16 passes dt as `container`) are not reproduced;
 * the obstacle branch, dead there (:28-45), pushes particles out along
 * their direction from the centre; the plane's bounds test uses 1 - |N|
 * (1 - N there misses every particle of a plane facing -axis); a
 * zero-length or plane-parallel step projects onto the plane instead of
 * dividing by zero.                                                       */
#define PSB_COLLIDER_SPHERE 1
#define PSB_COLLIDER_PLANE  2
#define PSB_MAX_COLLIDERS   16
typedef struct psb_collider {
  int32_t type;
  int32_t container;       /* sphere: 1 keeps particles inside            */
  float   elast;
  float   radius;          /* sphere                                       */
  float   origin[3];       /* sphere centre                                */
  float   corners[4][3];   /* plane: A B C D (Quadrilateral::create order) */
  float   normal[3];       /* plane: axis-aligned unit normal               */
} psb_collider;

/* collisionSystem_Sphere_CUDA (SphereCollision.cu:8-17) and collisionPlan_CUDA
 * (PlaneCollision.cu:8-22) on float4 arrays; old_vel4 and radius_particle
 * are accepted and unused, as in the reference.                             */
psb_status psb_collision_sphere(psb_context ctx, const float* old_pos4, float* new_pos4,
                                const float* old_vel4, float* new_vel4, float radius_particle,
                                float dt, uint32_t n, const float origin[3], float radius,
                                float elast, int32_t container);
psb_status psb_collision_plane(psb_context ctx, const float* old_pos4, float* new_pos4,
                               const float* old_vel4, float* new_vel4, const float a[3],
                               const float b[3], const float c[3], const float d[3],
                               const float normal[3], float elast, float dt, uint32_t n);

/* interactionSystem (CudaSystem.cu:6-14 -> collideParticles / collideParticle,
 * CudaParticle_Kernel.cu:10-55): DEM contact forces over each particle's
 * neighbour list, accumulated into forces4 (+=, in list order), with the
 * parameters of particle A (the list's owner).  d = |x_B - x_A|,
 * n = (x_B - x_A) / d, R = r_A + r_B:
 *   PSB_MODE_REFERENCE  F += -k (R - d) n for 0 < d < 2R: repulsive in
 *                       contact, attractive for R < d < 2R; the damping,
 *                       shear and attraction terms are dropped, as the
 *                       reference's non-mutating double3 += drops them
 *                       (double3.h:115-123), and R is rounded to float;
 *   PSB_MODE_CONFIG     spring-dashpot in contact only (0 < d < R):
 *                       F += -k (R - d) n + c ((v_B - v_A) . n) n, c = damping.
 * shear and attraction are accepted and unused in both modes.               */
psb_status psb_interaction_system(psb_context ctx, int32_t mode, const float* pos4,
                                  const float* vel4, float* forces4, const float* radius,
                                  const float* spring, const float* damping, const float* shear,
                                  const float* attraction, psb_neighbors nb, uint32_t n);

/* pci_SPH_update (PciSphSystem.cu:8-68): one PCISPH predict-correct solve.
 * On entry pos4/vel_inter_ap4 etc. hold the start-of-step state (the
 * reference's restore_* snapshot equals it); on return pos4 holds the final
 * predicted position, old_pos4 the start position, vel_inter_av4 the start
 * half-step velocity and vel_inter_ap4 the new one (collision and velocity
 * interpolation are separate calls, as in PciSphSystem.cpp:193-197).
 * PSB_MODE_REFERENCE runs exactly 3 iterations; PSB_MODE_CONFIG iterates
 * until max(0, rho - rho0)/rho0 <= tolerance (min/max iterations).
 * pressure is read as the initial pressure when warm_start != PSB_WARM_NONE.
 * iterations/max_error (host, may be NULL) report the solve.               */
psb_status psb_pci_sph_update(psb_context ctx, const psb_sph_params* prm, psb_neighbors nb,
                              float* old_pos4, float* pos4, const float* vel4,
                              float* vel_inter_av4, float* vel_inter_ap4,
                              float* density, float* pressure, float* normals4,
                              float* f_viscosity4, float* f_surface4, float* f_pressure4,
                              float* forces_accum4, uint32_t n,
                              int32_t* iterations, float* max_error);

/* Allocation-free psb_pci_sph_update.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as psb_pci_sph_update, but its scratch
 * (about 120 B per particle) is a static device array of each GPU, sized
 * for PSB_NOALLOC_MAX_PARTICLES particles (the reference's own capacity,
 * 65,536; 7.9 MB per GPU, held until the process ends), whose device
 * address psb_init looks up.  n above the maximum returns
 * PSB_ERR_CAPACITY.  The call allocates nothing; every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other
 * contexts or threads on the same device) race: serialise them.          */
#define PSB_NOALLOC_MAX_PARTICLES 65536u
psb_status psb_pci_sph_update_noalloc(psb_context ctx, const psb_sph_params* prm,
                                      psb_neighbors nb, float* old_pos4, float* pos4,
                                      const float* vel4, float* vel_inter_av4,
                                      float* vel_inter_ap4, float* density, float* pressure,
                                      float* normals4, float* f_viscosity4, float* f_surface4,
                                      float* f_pressure4, float* forces_accum4, uint32_t n,
                                      int32_t* iterations, float* max_error);

/* ------------------------------------------------------------------------ */
/* Fused system step (SPHSystem / PCI_SPHSystem)                              */
/* ------------------------------------------------------------------------ */
typedef struct psb_system_s* psb_system;

typedef struct psb_step_stats {
  uint32_t flags;          /* PSB_FLAG_* raised during the last update call */
  int32_t  iterations;     /* PCISPH corrector iterations of the last step  */
  float    max_error;      /* max(0, rho - rho0)/rho0 at the last iteration */
  int32_t  steps;          /* steps run by the last update call             */
  int64_t  total_iterations; /* corrector iterations over the last update call */
  float    worst_error;    /* largest final max_error over the last update call */
  int32_t  max_neighbors_seen; /* largest neighbour count of the last step   */
  int32_t  retried_steps;  /* steps of the last update call redone as substeps
                              (params.retry_levels)                          */
  int32_t  substeps;       /* substeps those retries ran                     */
} psb_step_stats;

/* SPHSystem::SPHSystem + init (SphSystem.cpp:6-113): allocate every device
 * buffer for `capacity` particles under the given parameters.             */
psb_status psb_system_create(psb_context ctx, const psb_sph_params* prm,
                             uint32_t capacity, psb_system* out);
psb_status psb_system_destroy(psb_system sys);
psb_status psb_system_get_params(psb_system sys, psb_sph_params* out);
/* SPHSystem::init(particles) / _initialize (SphSystem.cpp:50-201): host FP64
 * xyz triples (3 doubles per particle, the reference's m_h* layout).
 * Any pointer may be NULL (zeros).  Resets pressure and statistics.       */
psb_status psb_system_set_particles(psb_system sys, uint32_t n,
                                    const double* old_pos, const double* new_pos,
                                    const double* old_vel, const double* new_vel,
                                    const double* vel_inter_av, const double* vel_inter_ap);
/* Same from device vec4 arrays (no host round trip). */
psb_status psb_system_set_particles_device(psb_system sys, uint32_t n,
                                           const float* pos4, const float* vel4,
                                           const float* vel_inter_ap4);
uint32_t   psb_system_count(psb_system sys);
/* System::addObjectCollision (System.h:38) for sphere and plane colliders:
 * applied after the box each step, in the order added (the box, from
 * psb_sph_params, comes first).  Not part of checkpoints.                  */
psb_status psb_system_add_collider(psb_system sys, const psb_collider* collider);
psb_status psb_system_clear_colliders(psb_system sys);

/* Number of rigid spheres (psb_system_set_spheres or a loaded checkpoint).  */
uint32_t psb_system_num_spheres(psb_system sys);
/* Per-particle materials, the reference's per-particle arrays mass,
 * interactionRadius, gasStiffness, restDensity, viscosity, threshold and
 * surfaceTension (SphSystem.cpp:157-167, PciSphSystem.cpp:103-113, filled per
 * particle by Emitter::addParticle, Emitter.cpp:141-175).  Particle i of
 * the last psb_system_set_particles takes materials[index[i]] (n = that
 * particle count); the fields read from each material are mass,
 * support_radius, gas_stiffness, rest_density, viscosity, threshold and
 * surface_tension, everything else comes from the system's parameters.
 * Pairs mix as the reference's kernels do: WCSPH takes the neighbour's
 * radius and mass (SphKernel.cu:17-21, 51-71), PCISPH max(h_i, h_j) and the
 * neighbour's mass (PciSphSystem_Kernel.cu:51, 79, 101, 138), both the mean
 * viscosity; particle i's neighbours are those within its own radius
 * (UniformGridKernel.cu:82).  Reference mode only (WCSPH with
 * PSB_EOS_LINEAR); every support_radius must be at most the system's, which
 * sizes the neighbour search.  The steps then run separate one-thread-per-
 * particle kernels (PCISPH: the reference's three corrector iterations).
 * count = 0 returns to the system's one material; set_particles,
 * set_state, set_spheres and load_checkpoint clear the table too. */
#define PSB_MAX_MATERIALS 8
psb_status psb_system_set_materials(psb_system sys, uint32_t count,
                                    const psb_sph_params* materials,
                                    const uint8_t* index, uint32_t n);
/* Materials in use (0: the system's one material). */
uint32_t psb_system_num_materials(psb_system sys);
/* SPHSystem::update / PCI_SPHSystem::update, `steps` times. */
psb_status psb_system_update(psb_system sys, int32_t steps);
psb_status psb_system_get_stats(psb_system sys, psb_step_stats* out);
/* Host FP64 triples in the caller's original particle order (NULL skips):
 * m_hPos[0], m_hPos[1], m_hVel[0], m_hVel[1], m_hVelInterAv, m_hVelInterAp,
 * plus density and pressure (n doubles).                                   */
psb_status psb_system_get_particles(psb_system sys, double* old_pos, double* new_pos,
                                    double* old_vel, double* new_vel,
                                    double* vel_inter_av, double* vel_inter_ap,
                                    double* density, double* pressure);
/* Same, FP32, sorted order of the last step, plus the particle ids. */
psb_status psb_system_get_particles_f32(psb_system sys, float* pos4, float* vel4,
                                        float* vel_inter_ap4, float* density,
                                        float* pressure, uint32_t* ids);

/* State in a given storage order (the inverse of psb_system_get_particles_f32):
 * pos4 / vel4 / vel_inter_ap4 as float4 (w ignored), pressure (NULL: zero; the
 * PSB_WARM_PREVIOUS start), ids a permutation of 0..n-1 naming each particle.
 * Spheres set with psb_system_set_spheres are appended after the n fluid
 * particles, as by psb_system_set_particles.  No reference counterpart: it
 * restores a state written by ParticleExporter_XML-style exporters in the
 * storage order the next step sorts from, so that the run continues bit for
 * bit.                                                                    */
psb_status psb_system_set_state(psb_system sys, uint32_t n, const float* pos4,
                                const float* vel4, const float* vel_inter_ap4,
                                const float* pressure, const uint32_t* ids);

/* Checkpoint / resume.  The file holds the parameters, the sphere setup and
 * statistics, and every per-particle array that carries from one step to the
 * next, in storage order, with a 64-bit hash; a system loaded from it
 * continues exactly as the one that saved it (same results bit for bit).
 * Load needs a system created with the same psb_sph_params (byte for byte)
 * and the capacity for psb_checkpoint_info's count; MLP and recorder
 * settings are not part of the file.  Counterpart of the reference's
 * ParticleExporter_XML::_export_SPHParticles / ParticleLoader_XML::
 * loadSPHSystem (ParticleExporter_XML.cpp:131-190, ParticleLoader_XML.cpp:
 * 121-180), which keep positions and velocities only.                      */
psb_status psb_system_save_checkpoint(psb_system sys, const char* path);
psb_status psb_system_load_checkpoint(psb_system sys, const char* path);
psb_status psb_checkpoint_info(const char* path, psb_sph_params* params, uint32_t* count,
                               uint32_t* fluid_count, uint32_t* num_spheres);

/* ------------------------------------------------------------------------ */
/* Surface extraction: SurfaceSPH::extract (SurfaceSPH.cpp:29-190) --        */
/* evaluateIso_Kernel (UniformGridKernel.cu:98-138) on a node grid over the  */
/* particles' bounds plus a margin, marching cubes (MarchingCube_Kernel.cu:  */
/* 14-158), vertex normals.  Differences from the reference: FP32; the field */
/* sums every particle within r = h / scale (the reference scans floor(1/    */
/* scale) cells of size h, which misses part of that radius); one vertex per */
/* crossed grid edge shared by the triangles that use it (the reference      */
/* merges vertices by float equality on the host); normals are the direction */
/* of -grad of the extracted field (the reference sums an SPH colour         */
/* gradient six times over); the case table is generated from the contours  */
/* (same crossed edges and polygons as the reference's table in all 256      */
/* cases, fan triangulation).                                                */
/* ------------------------------------------------------------------------ */
typedef struct psb_surface_s* psb_surface;

typedef struct psb_surface_params {
  float cell_size;   /* node spacing (SurfaceSPH.cpp:31: 0.03)                  */
  float scale;       /* field radius r = support_radius / scale (0.65)          */
  float iso_level;   /* level of f(x) = sum_j (r^2 - |x - x_j|^2)^3 / r^6       */
  float margin;      /* node grid beyond the particles' bounds, in r (2)        */
} psb_surface_params;

psb_status psb_surface_default_params(psb_surface_params* params);
psb_status psb_surface_create(psb_context ctx, psb_surface* out);
psb_status psb_surface_destroy(psb_surface s);
/* Extract from the fluid particles of a system's current state. */
psb_status psb_surface_extract(psb_surface s, psb_system sys, const psb_surface_params* params,
                               uint32_t* num_vertices, uint32_t* num_triangles);
/* Extract from n device float4 positions; particles whose id (ids, device,
 * may be NULL) is >= id_limit are left out.                                 */
psb_status psb_surface_extract_points(psb_surface s, const float* pos4, const uint32_t* ids,
                                      uint32_t n, uint32_t id_limit, float support_radius,
                                      const psb_surface_params* params, uint32_t* num_vertices,
                                      uint32_t* num_triangles);
/* Host copies of the last mesh: vertices and unit normals (3 floats each),
 * triangles (3 vertex indices each, wound with the normal facing the low
 * side of the field).  Any pointer may be NULL.                             */
psb_status psb_surface_get_mesh(psb_surface s, float* vertices3, float* normals3,
                                uint32_t* triangles3);
/* The node field of the last extraction (x fastest) and its grid; node
 * (i, j, k) sits at origin + (i, j, k) * cell.  field may be NULL.         */
psb_status psb_surface_get_field(psb_surface s, float* field, uint32_t dims[3], float origin[3],
                                 float* cell);
/* The case table in use: per case (bit k: corner k <= iso, corners numbered
 * as MarchingCube_Kernel.cu:45-53) its triangle count and up to 15 edge
 * indices (16 entries per case, -1 padded).                                 */
psb_status psb_surface_case_table(uint8_t* triangle_counts, int8_t* triangle_edges);

/* Device pointers of the system's sorted state (valid until the next update). */
typedef struct psb_system_arrays {
  const float*    pos4;          /* m_dPos[1] after the last step   */
  const float*    old_pos4;      /* m_dPos[0]                       */
  const float*    vel4;          /* m_dVel[1]                       */
  const float*    old_vel4;      /* m_dVel[0]                       */
  const float*    vel_inter_ap4; /* m_dVelInterAp                   */
  const float*    vel_inter_av4; /* m_dVelInterAv                   */
  const float*    density;
  const float*    pressure;
  const uint32_t* ids;           /* caller index of each sorted slot */
  const uint32_t* hash;          /* grid hash in the previous order  */
  const uint32_t* perm;
  const uint32_t* cell_start;
  psb_neighbors   neighbors;
  uint64_t        ncells;
} psb_system_arrays;
psb_status psb_system_get_arrays(psb_system sys, psb_system_arrays* out);

/* Compact neighbour list for the PCISPH corrector passes (on by default):
 * the neighbour search also writes 16-bit offsets q - p, which the
 * per-iteration density and pressure-force passes read instead of the
 * 32-bit indices; a step where some offset exceeds 32767 falls back to the
 * 32-bit list by itself, and from the next step on the search writes a
 * grouped 32-bit list (four indices per 16-byte word) in its place.
 * Results do not depend on it.                                             */
psb_status psb_system_set_compact_list(psb_system sys, int32_t on);

/* Neighbour rows staged in LDS by the PCISPH corrector passes (off by
 * default; used with one thread per particle, psb_system_set_corrector_lanes
 * (sys, 1), and the compact list): the
 * sorted particles are cut into tiles of 512, each tile's neighbour rows are
 * copied into LDS once per pass and the neighbours read from there.  A tile
 * whose rows exceed 4,096 particles keeps the global path.  Results do not
 * depend on it.                                                            */
psb_status psb_system_set_halo(psb_system sys, int32_t on);

/* Threads per particle in the PCISPH corrector passes: 1, 2 (the neighbour
 * sums split into even and odd list positions, one lane each), or 0 (the
 * default: 2 while one lane per particle would leave at most 4 wavefronts
 * per SIMD, i.e. n <= 1024 x compute units, or once the previous update
 * call ended on the grouped 32-bit list, else 1; 1 with LDS staging on).
 * Results do not depend on it: the sums always run in the even + odd
 * order.                                                                   */
psb_status psb_system_set_corrector_lanes(psb_system sys, int32_t lanes);

/* Per-phase timing with HIP events on the system stream (off by default):
 * `every` = 0 disables it, k > 0 brackets every launch of every k-th step
 * (counted over the system's lifetime) with a pair of events, so k > 1
 * samples the kernels with 1/k of the event overhead.
 * Phase names: "hash_count", "scan", "scatter", "reorder", "neighbors_density",
 * "forces", "nonpressure", "pci_density", "pci_force", plus "mlp" etc.    */
psb_status psb_system_set_timing(psb_system sys, int32_t every);
/* Returns the number of phases; writes up to `cap` entries. */
int32_t    psb_system_get_timing(psb_system sys, int32_t cap, const char** names,
                                 double* total_ms, int64_t* launches);
psb_status psb_system_reset_timing(psb_system sys);
/* Kernel durations of the corrector passes on the system's current state
 * (no counterpart in the reference; a measurement helper for the
 * benchmark): the density pass and the force pass of iteration 0 are each
 * launched `reps` times back to back on the system's stream between two HIP
 * events, so the per-launch figure carries one event pair per `reps`
 * launches instead of one per launch.  Needs a PCISPH config-mode system
 * that has stepped; the replays rewrite only scratch (predicted positions,
 * corrector pressure), so the next update continues from the same state.
 * Outputs are microseconds per launch.                                     */
psb_status psb_system_time_corrector(psb_system sys, int32_t reps, double* us_density,
                                     double* us_force);

/* ------------------------------------------------------------------------ */
/* Learned PCISPH warm start.  No counterpart in the reference, whose        */
/* corrector always starts from p = 0 (PciSphSystem_Kernel.cu:131); the      */
/* unchanged corrector still enforces the tolerance, so a poor prediction    */
/* only costs iterations.                                                    */
/* ------------------------------------------------------------------------ */
typedef struct psb_mlp_s* psb_mlp;

/* Per-particle MLP 8 -> 64 -> 64 -> 1 (SiLU), 4,801 FP32 master parameters
 * (W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3), BF16 MFMA products
 * with FP32 accumulation.  He-uniform initialisation from `seed`.           */
psb_status psb_mlp_create(psb_context ctx, uint64_t seed, psb_mlp* out);
psb_status psb_mlp_destroy(psb_mlp m);
int32_t    psb_mlp_num_params(void);
psb_status psb_mlp_set_params(psb_mlp m, const float* host_params, int32_t count);
psb_status psb_mlp_get_params(psb_mlp m, float* host_params, int32_t count);
/* Input standardisation x' = (x - mean) / stddev and the output scale:
 * predicted pressure = pressure_scale * network output.                     */
psb_status psb_mlp_set_normalization(psb_mlp m, const float mean[8], const float stddev[8],
                                     float pressure_scale);
/* out[i] = network output (unscaled) for features[k * n + i] (device SoA). */
psb_status psb_mlp_forward(psb_mlp m, const float* features, uint32_t n, float* out);
/* One minibatch of MSE training on (features[k * num_samples + s],
 * targets[s]) for s = batch_idx[0 .. batch) (device arrays): gradients
 * accumulate on the device; apply_adam != 0 then takes one Adam step
 * (lr, beta 0.9 / 0.999, eps 1e-8) and clears them.  *loss_out (host, may be
 * NULL) receives the minibatch mean squared error before the update.       */
psb_status psb_mlp_train_step(psb_mlp m, const float* features, const float* targets,
                              uint64_t num_samples, const uint32_t* batch_idx, uint32_t batch,
                              float lr, int32_t apply_adam, float* loss_out);
psb_status psb_mlp_get_grad(psb_mlp m, float* host_grad, int32_t count);
psb_status psb_mlp_zero_grad(psb_mlp m);

/* Features of the warm start, per particle, computed by the PCISPH passes
 * (FP32, sorted order, SoA [8][n]):
 *   0: 100 (rho* - rho0)/rho0 at the first prediction (p = 0)
 *   1: dt sum_j (m/rho_j)(v_j - v_i).gradW_ij      (velocity divergence)
 *   2: neighbour count / 32
 *   3: previous-step pressure / pressure_scale
 *   4: |v_i| dt / h
 *   5-7: (h / pressure_scale) sum_j (m/rho_j)(p_j - p_i) gradW_ij  (previous-step pressure)
 * pressure_scale = delta_template * rho0 / 100 (the pressure correcting a
 * 1 % compression).                                                        */
float      psb_system_pressure_scale(psb_system sys);
/* Attach the network used when params.warm_start == PSB_WARM_MLP. */
psb_status psb_system_set_mlp(psb_system sys, psb_mlp m);
/* Device pointer to the feature buffer of the last step ([8][n]); NULL
 * until features were computed (MLP warm start or recording).             */
psb_status psb_system_get_features(psb_system sys, const float** features);

/* Training set: feature SoA [8][capacity] + target[capacity] on the device. */
typedef struct psb_dataset_s* psb_dataset;
psb_status psb_dataset_create(psb_context ctx, uint64_t capacity, psb_dataset* out);
psb_status psb_dataset_destroy(psb_dataset d);
psb_status psb_dataset_reset(psb_dataset d);
psb_status psb_dataset_info(psb_dataset d, const float** features, const float** targets,
                            uint64_t* count, uint64_t* capacity);
/* Record, after every PCISPH step, the features of every `stride`-th particle
 * (offset rotating with the step) and its converged pressure / pressure_scale
 * as the target.  Recording stops when the dataset is full.  NULL detaches. */
psb_status psb_system_set_recorder(psb_system sys, psb_dataset d, uint32_t stride);

/* ------------------------------------------------------------------------ */
/* Rigid DEM spheres coupled to a PCISPH system (PSB_MODE_CONFIG).           */
/* The reference couples particle systems only through interactionSystem   */
/* (above); here each sphere is also a rigid body of boundary particles:    */
/* the lattice points (i, j, k) * lattice_spacing strictly inside its       */
/* radius, mass m each.  They are listed as neighbours of the fluid, add to */
/* its density and to sum gradW of delta_i, and push it with the fluid      */
/* particle's own pressure (boundary pressure 0, Akinci et al. 2012).  The  */
/* sphere receives the opposite force, summed per sphere in boundary-id     */
/* order; with gravity and the DEM contacts of psb_interaction_system's     */
/* config law (sphere-sphere, and sphere-wall against the collider box) it  */
/* moves by symplectic Euler once per step, without rotation (pressure and  */
/* normal contacts exert no torque on a sphere).  The boundary particles    */
/* occupy ids n_fluid .. n_fluid + count * per_sphere - 1 after the fluid.  */
/* ------------------------------------------------------------------------ */
typedef struct psb_sphere_params {
  float radius;        /* m                                                  */
  float density;       /* kg/m^3 (mass = density 4/3 pi radius^3)           */
  float spring;        /* k_n, N/m                                           */
  float damping;       /* normal dashpot coefficient c, N s/m                */
} psb_sphere_params;

/* Set before psb_system_set_particles* (which appends the boundary
 * particles after the fluid); the capacity must hold fluid + boundary.
 * centers3 / velocities3: count x 3 floats (velocities may be NULL).      */
psb_status psb_system_set_spheres(psb_system sys, uint32_t count, const float* centers3,
                                  const float* velocities3, const psb_sphere_params* params);
/* Boundary particles per sphere for these params (lattice points inside). */
uint32_t   psb_sphere_particle_count(const psb_sph_params* prm, const psb_sphere_params* sp);
/* Sphere state after the last step: centres, velocities and the fluid
 * force of the last step (count x 3 floats each; any may be NULL).        */
psb_status psb_system_get_spheres(psb_system sys, float* centers3, float* velocities3,
                                  float* fluid_forces3);

/* ------------------------------------------------------------------------ */
/* Batched scenes.  No counterpart in the reference, which simulates one     */
/* System per process (one UniformGrid, 65,536 particles).  A batch advances */
/* many independent PCISPH scenes (PSB_MODE_CONFIG) that share one material  */
/* (h, rho0, m, dt, tolerance, iterations, warm start) and differ in         */
/* geometry: particle count, grid and collider box.  Every phase of the step */
/* is one launch over all scenes (workgroups aligned to scene boundaries);   */
/* each scene keeps its own corrector: it stops at its own tolerance and its */
/* blocks skip the remaining iterations of slower scenes.  Results equal     */
/* the same scenes run one by one as psb_system.                             */
/* ------------------------------------------------------------------------ */
typedef struct psb_batch_s* psb_batch;

typedef struct psb_scene_desc {
  uint32_t      count;        /* particles (fixed for the batch's life)      */
  uint32_t      pad;
  float         box_min[3];   /* collider box (BoxCollision, per scene)      */
  float         box_max[3];
  psb_grid_desc grid;         /* scene grid; cell_size = material grid's     */
} psb_scene_desc;

/* Scene-batch step phases, for psb_batch_autotune / psb_batch_get_timing.
 * The density epilogue runs inside the neighbors_density phase; its own
 * phase slot is kept for the layout and launches nothing.                  */
#define PSB_BATCH_PHASES 6  /* hash_count, neighbors_density, density_epilogue,
                               pci_nonpressure, pci_density, pci_force       */

psb_status psb_batch_create(psb_context ctx, const psb_sph_params* material, uint32_t num_scenes,
                            const psb_scene_desc* scenes, psb_batch* out);
psb_status psb_batch_destroy(psb_batch b);
uint32_t   psb_batch_num_scenes(psb_batch b);
/* Host FP32 float4 arrays of scene `scene` (count particles); vel4 / vh4 may
 * be NULL (zero).                                                          */
psb_status psb_batch_set_particles(psb_batch b, uint32_t scene, const float* pos4,
                                   const float* vel4, const float* vh4);
psb_status psb_batch_update(psb_batch b, int32_t steps);
/* Per-scene statistics of the last update (flags are batch-wide).          */
psb_status psb_batch_get_stats(psb_batch b, uint32_t scene, psb_step_stats* out);
/* State of one scene in its sorted order (host arrays, any may be NULL).   */
psb_status psb_batch_get_particles(psb_batch b, uint32_t scene, float* pos4, float* vel4,
                                   float* vel_half4, float* density, float* pressure,
                                   uint32_t* ids);
/* Learned warm start (material warm_start = PSB_WARM_MLP).                 */
psb_status psb_batch_set_mlp(psb_batch b, psb_mlp m);
/* Time every phase at 64, 128, 256 and 512 threads per workgroup, and the
 * corrector passes (pci_density, pci_force) also at 128, 256 and 512 threads
 * with two threads per particle, on `steps` steps from the current state each
 * (state restored afterwards) and keep the fastest per phase;
 * block_sizes[PSB_BATCH_PHASES] receives the choice (may be NULL): threads
 * per workgroup, negated for two threads per particle.  Results do not depend
 * on the choice.                                                            */
psb_status psb_batch_autotune(psb_batch b, int32_t steps, int32_t* block_sizes);
/* Kernel durations of the corrector passes over every scene of the batch on
 * its current state (no counterpart in the reference; a measurement helper
 * for the benchmark, as psb_system_time_corrector): the density and force
 * passes of iteration 0, every scene active, each launched `reps` times back
 * to back between two HIP events; microseconds per launch.  Needs a batch
 * that has stepped; the replays rewrite only scratch, so the next update
 * continues from the same state.                                           */
psb_status psb_batch_time_corrector(psb_batch b, int32_t reps, double* us_density,
                                    double* us_force);
/* Set the launch shapes directly, in the form psb_batch_autotune returns
 * (e.g. a stored tuning): 64, 128, 256 or 512 per phase; -128, -256 or -512
 * for two threads per particle in pci_density and pci_force.               */
psb_status psb_batch_set_block_sizes(psb_batch b, const int32_t* block_sizes);
/* Replay each step from HIP graphs: the step prologue (grid, neighbours,
 * non-pressure forces, first min_iterations corrector iterations) and one
 * graph per later iteration, captured on first use.                        */
psb_status psb_batch_set_graphs(psb_batch b, int32_t on);
/* Per-scene step latency (from the step's start to the end of the force
 * pass that finished that scene), collected on every `every`-th step while
 * enabled (0: off, 1: every step).  Sampled steps record an event after every
 * force pass, which delays the next density pass by a few microseconds; the
 * other steps pace their queue by polling host-mapped memory.               */
psb_status psb_batch_set_latency(psb_batch b, int32_t every);
psb_status psb_batch_get_latency(psb_batch b, double* p50_ms, double* p99_ms, double* max_ms,
                                 uint64_t* samples);

#ifdef __cplusplus
}
#endif
#endif /* PSB_SPH_H */
