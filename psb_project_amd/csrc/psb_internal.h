// psb_internal.h -- internal declarations of the gfx950 SPH/PCISPH library.
//
// Layout in HBM (all FP32, sorted particle order after each grid build):
//   vec4 arrays (16 B/particle): positions, velocities, half-step velocities,
//     predicted positions, per-particle force terms.  The w lane carries a
//     per-particle scalar that neighbour passes gather together with xyz
//     (p/rho^2, 1/rho or the PCISPH pressure), so one 16-byte load serves a
//     neighbour.
//   scalar arrays (4 B/particle): density, pressure, ids, grid hash/perm.
//   cell table: start[ncells + 1] (u32), count[ncells] (u32).
//   neighbour list: ELL, list[k * n + i] (u32), count[n].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/psb_sph.h"

namespace psb {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
psb_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define PSB_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) return ::psb::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define PSB_TRY(call)                                                        \
  do {                                                                       \
    psb_status s_ = (call);                                                  \
    if (s_ != PSB_OK) return s_;                                             \
  } while (0)

#define PSB_ARG(cond, msg)                                                   \
  do {                                                                       \
    if (!(cond)) { ::psb::set_error(msg); return PSB_ERR_INVALID_ARGUMENT; } \
  } while (0)

constexpr uint32_t kBlock = 256;          // threads per workgroup (4 wave64)
constexpr uint32_t kScanTile = 4096;      // cells per scan workgroup (256 x 16)
constexpr int kMaxIterSlots = 256;        // PCISPH iteration slots per step
constexpr int kMaxRetryLevels = 4;        // dt halvings of a failed PCISPH step
constexpr int kRetryIterFactor = 4;       // iteration cap of the deepest retry level
constexpr uint32_t kMaxParticles = 1u << 27;  // particles per system or batch (gather4)
// corrector iteration cap of the deepest retry level (params.retry_levels)
inline int retry_max_iterations(const psb_sph_params& p) {
  const int m = p.max_iterations * kRetryIterFactor;
  return m < kMaxIterSlots - 1 ? m : kMaxIterSlots - 1;
}

inline uint32_t blocks_for(uint64_t n, uint32_t b = kBlock) {
  return (uint32_t)((n + b - 1) / b);
}

// ---------------------------------------------------------------------------
// Constants the kernels read (derived once on the host, psb_derive)
// ---------------------------------------------------------------------------
struct GridDev {
  float minx, miny, minz, cell;
  int nx, ny, nz;
  int R;            // stencil half-width in cells
};

// Static collider after the box (psb_collider on the device)
struct ColliderDev {
  int type, container;
  float elast, radius;
  float o[3];            // sphere centre
  float a[3];            // plane corner A
  float mn[3], mx[3];    // plane corner bounds as the reference forms them
  float nrm[3];          // plane normal
};

struct Consts {
  float h, h2, h2x3, c_spiky, mc, beta, delta, tait_b, inv_rho0, rho0, k, kp;
  float kps;   // kp * c_spiky (config-mode pressure force, applied to the pair sum)
  float cv, cs, cn, m, dt, sigma, l, gx, gy, gz;
  float box_lo[3], box_hi[3], one_plus_e;
  float tolerance;
  int eos, mode, box_enabled, min_it, max_it;
  int delta_local_cap;
  // 1: a step ending above the tolerance will be redone as substeps (the
  // fused system step with params.retry_levels), so the force pass leaves
  // NOT_CONVERGED and the worst error to the attempt that is kept
  int retry_ok;
  // 1: a retry substep, whose corrector stops as failed when it diverges
  // (kDivergeFactor / kDivergeFloor)
  int diverge_stop;
  // warm-start features: pressure scale and feature normalisers
  float pscale, inv_ps, h_over_ps, dt_over_h, e_scale;
  // lattice walls (PSB_WALL_LATTICE): implicit boundary particles at
  // wall_o + (i + 1/2) s for every lattice index i outside [0, wall_n) on
  // some axis and inside [-wall_layers, wall_n + wall_layers) on all axes
  int wall;
  int wall_n[3], wall_layers;
  float wall_o[3];
  float wall_s, wall_inv_s, wall_r;   // spacing, 1/spacing, h/spacing
  // rigid-sphere boundary particles: ids >= n_fluid (kNoBodies: none)
  uint32_t n_fluid;
  // colliders applied after the box, in order (device array)
  const ColliderDev* coll;
  int ncoll;
  // grouped 32-bit neighbour list (l32g_at) of a single system whose compact
  // offsets overflow, or null
  const uint32_t* list32g;
};

psb_status make_consts(const psb_sph_params& p, Consts* c, GridDev* g);
// validated device form of a psb_collider
psb_status make_collider(const psb_collider& in, ColliderDev* out);

constexpr uint32_t kNoBodies = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Device helpers shared by all kernels
// ---------------------------------------------------------------------------
// Blocks are dispatched round-robin over the 8 XCDs of MI355X; give each XCD
// a contiguous range of particles so neighbour gathers stay in its L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t X = 8;
  const uint32_t q = nb / X, r = nb % X;
  const uint32_t x = b % X, k = b / X;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ uint32_t particle_index() {
  return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// Particle maps: which particle (and scene) a thread works on.
//
// SingleMap: one scene, particle = block * blockDim + thread (XCD remapped).
// BatchMap: many independent scenes in one launch.  Each scene's particles
// are contiguous (p0 .. p0 + n) and its blocks are contiguous too, padded to
// a whole number of workgroups, so every workgroup belongs to exactly one
// scene and block-level logic (error reduction, early exit, the convergence
// decision) stays per scene.  The scene supplies its grid, cell offset,
// collider box and lattice-wall origin; the material constants are shared.
// ---------------------------------------------------------------------------
constexpr int kBatchBlockSizes = 4;           // 64, 128, 256, 512 threads

struct BatchScene {
  GridDev grid;                  // scene grid (origin, cell, dims, stencil)
  float box_lo[3], box_hi[3];    // collider box (radius applied)
  float wall_o[3];               // lattice-wall origin (box_min)
  int wall_n[3];                 // lattice sites per axis inside the box
  uint32_t cell_off;             // first global cell of the scene
  uint32_t p0, n;                // particle range
  uint32_t blk0[kBatchBlockSizes];  // first block, per block size
  uint32_t pad;
};

// LANES: threads per particle (1, or 2 for the corrector passes, whose
// neighbour sums split into even and odd neighbours).
template <int LANES = 1>
struct SingleMapT {
  static constexpr int kMaxThreads = 256;
  static constexpr int kLanes = LANES;
  static constexpr bool kHalo = false;
  static constexpr bool kWideList = true;   // may read the grouped 32-bit list
  __device__ __forceinline__ bool locate(uint32_t n, uint32_t& p, uint32_t& s, bool& lead) const {
    p = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / LANES;
    s = 0;
    lead = blockIdx.x == 0 && threadIdx.x == 0;
    return p < n;
  }
  __device__ __forceinline__ void patch(Consts&, uint32_t) const {}
  __device__ __forceinline__ uint32_t scene_of_block() const { return 0; }
  __device__ __forceinline__ GridDev grid(const GridDev& g, uint32_t, uint32_t& off) const {
    off = 0;
    return g;
  }
};

using SingleMap = SingleMapT<1>;

// ---------------------------------------------------------------------------
// Neighbour rows staged in LDS (single scene).  The sorted particles are cut
// into tiles of kHaloTile; a tile's halo is every grid row (y', z') within
// the stencil of one of its particles, each row over the x cells
// [xlo, xhi] (the tile's x range widened by the stencil), i.e. a set of
// contiguous sorted-index ranges.  The corrector passes copy the halo into
// LDS once per launch and read every neighbour from there; the compact list
// then holds LDS slots instead of index offsets.  A tile whose halo exceeds
// the limits keeps the global path (ok = 0).
// ---------------------------------------------------------------------------
constexpr int kHaloTile = 512;      // particles per tile = threads per workgroup
constexpr int kHaloCap = 4096;      // staged particles per tile (64 KiB of float4)
constexpr int kHaloMaxZ = 13;       // z' layers
constexpr int kHaloMaxRows = 224;

struct HaloTile {
  int32_t ok, nrows, total, xlo, xhi, zlo, nz, pad;
  int32_t ylo[kHaloMaxZ], rowbase[kHaloMaxZ + 1];   // layer i: rows rowbase[i] .. rowbase[i+1)-1
  uint32_t gstart[kHaloMaxRows];     // sorted index of the row's first particle (cell xlo)
  uint32_t lbase[kHaloMaxRows + 1];  // LDS slot of that particle; lbase[nrows] = total
};

// slot of the staged copy of particle q (in row (y', z') of tile T) minus q
__device__ __forceinline__ int32_t halo_slot_base(const HaloTile& T, int y, int z) {
  const int i = z - T.zlo;
  const int r = T.rowbase[i] + (y - T.ylo[i]);
  return (int32_t)T.lbase[r] - (int32_t)T.gstart[r];
}

// TileMap: one scene, one workgroup per tile (blockDim = kHaloTile), tiles
// XCD remapped like SingleMap; the corrector passes stage the tile's halo.
struct TileMap {
  static constexpr int kMaxThreads = kHaloTile;
  static constexpr int kLanes = 1;
  static constexpr bool kHalo = true;
  static constexpr bool kWideList = false;
  const HaloTile* __restrict__ tiles;
  const uint32_t* __restrict__ src;  // [tile][kHaloCap]: sorted index staged in each slot
  __device__ __forceinline__ uint32_t tile() const { return xcd_block(blockIdx.x, gridDim.x); }
  __device__ __forceinline__ bool locate(uint32_t n, uint32_t& p, uint32_t& s, bool& lead) const {
    p = tile() * blockDim.x + threadIdx.x;
    s = 0;
    lead = blockIdx.x == 0 && threadIdx.x == 0;
    return p < n;
  }
  __device__ __forceinline__ void patch(Consts&, uint32_t) const {}
  __device__ __forceinline__ uint32_t scene_of_block() const { return 0; }
  __device__ __forceinline__ GridDev grid(const GridDev& g, uint32_t, uint32_t& off) const {
    off = 0;
    return g;
  }
};

template <int LANES = 1>
struct BatchMapT {
  static constexpr int kMaxThreads = 512;
  static constexpr int kLanes = LANES;
  static constexpr bool kHalo = false;
  static constexpr bool kWideList = false;
  const BatchScene* __restrict__ sc;
  const uint16_t* __restrict__ blk_scene;   // scene of each block, for this block size
  int bsel;                                 // particles per block = 64 << bsel
  __device__ __forceinline__ uint32_t scene_of_block() const {
    return blk_scene[xcd_block(blockIdx.x, gridDim.x)];
  }
  __device__ __forceinline__ bool locate(uint32_t, uint32_t& p, uint32_t& s, bool& lead) const {
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    s = blk_scene[b];
    const BatchScene& S = sc[s];
    const uint32_t lb = b - S.blk0[bsel];
    const uint32_t local = (lb * blockDim.x + threadIdx.x) / LANES;
    p = S.p0 + local;
    lead = lb == 0 && threadIdx.x == 0;
    return local < S.n;
  }
  __device__ __forceinline__ void patch(Consts& c, uint32_t s) const {
    const BatchScene& S = sc[s];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      c.box_lo[a] = S.box_lo[a];
      c.box_hi[a] = S.box_hi[a];
      c.wall_o[a] = S.wall_o[a];
      c.wall_n[a] = S.wall_n[a];
    }
  }
  __device__ __forceinline__ GridDev grid(const GridDev&, uint32_t s, uint32_t& off) const {
    off = sc[s].cell_off;
    return sc[s].grid;
  }
};

using BatchMap = BatchMapT<1>;

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return __builtin_fmaf(ax, bx, __builtin_fmaf(ay, by, az * bz));
}

// sorted particle p is a rigid-sphere boundary particle (ids: start-of-step ids)
__device__ __forceinline__ bool is_boundary(const Consts& c, const uint32_t* __restrict__ ids,
                                            uint32_t p) {
  return c.n_fluid != kNoBodies && ids[p] >= c.n_fluid;
}

// cell = floor((x - Min) / cell) clamped to the grid (UniformGridKernel.cu:48-50)
__device__ __forceinline__ int cell_coord(float x, float mn, float cell, int n, uint32_t& bad) {
  float t = floorf((x - mn) / cell);
  if (!(t >= 0.0f && t < (float)n)) {
    bad = 1u;
    t = (t >= (float)n) ? (float)(n - 1) : 0.0f;
  }
  return (int)t;
}

// BoxCollision_Kernel.cu:7-34 (restitution via one_plus_e)
__device__ __forceinline__ void box_collide(const Consts& c, float4& x, float4& v) {
  float px = x.x, py = x.y, pz = x.z, nx = 0.f, ny = 0.f, nz = 0.f;
  if (x.x >= c.box_hi[0]) { px = c.box_hi[0]; nx = -1.f; }
  if (x.y >= c.box_hi[1]) { py = c.box_hi[1]; ny = -1.f; }
  if (x.z >= c.box_hi[2]) { pz = c.box_hi[2]; nz = -1.f; }
  if (x.x <= c.box_lo[0]) { px = c.box_lo[0]; nx = 1.f; }
  if (x.y <= c.box_lo[1]) { py = c.box_lo[1]; ny = 1.f; }
  if (x.z <= c.box_lo[2]) { pz = c.box_lo[2]; nz = 1.f; }
  float ln2 = dot3(nx, ny, nz, nx, ny, nz);
  if (ln2 > 0.f) {
    float ln = sqrtf(ln2);
    nx = nx / ln; ny = ny / ln; nz = nz / ln;
    x.x = px; x.y = py; x.z = pz;
    float vn = c.one_plus_e * dot3(v.x, v.y, v.z, nx, ny, nz);
    v.x = v.x - vn * nx; v.y = v.y - vn * ny; v.z = v.z - vn * nz;
  }
}

// Velocity response and placement shared by the sphere and plane colliders
// (SphereCollision_Kernel.cu:62-76, PlaneCollision_Kernel.cu:62-76):
// r = elast |x - x_int| / (dt |v|) when |v| > 0 and elast > 0; with n the
// normalised contact normal v <- v - ((1 + r)(v.n)) n; x <- x_int.
__device__ __forceinline__ void collider_response(float elast, float dt, float4& x, float4& v,
                                                  float ix, float iy, float iz, float nx,
                                                  float ny, float nz) {
  const float vl = sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z));
  float r = 0.f;
  if (vl > 0.f && elast > 0.f) {
    const float ex = x.x - ix, ey = x.y - iy, ez = x.z - iz;
    r = (elast * sqrtf(dot3(ex, ey, ez, ex, ey, ez))) / (dt * vl);
  }
  const float nl = sqrtf(dot3(nx, ny, nz, nx, ny, nz));
  nx = nx / nl; ny = ny / nl; nz = nz / nl;
  const float vn = (1.0f + r) * dot3(v.x, v.y, v.z, nx, ny, nz);
  v.x = v.x - vn * nx; v.y = v.y - vn * ny; v.z = v.z - vn * nz;
  x.x = ix; x.y = iy; x.z = iz;
}

// detection_collisionSphere (SphereCollision_Kernel.cu:5-47) in FP32.
// Container: a particle at distance d >= R from the centre goes to the
// surface point along its direction, normal towards the centre.  Obstacle
// (d <= R): the surface point along its direction, normal outwards -- the
// reference's obstacle branch could only fire at d == R (:28-45) and would
// have used the opposite surface point.  d == R: stays, normal -OP/d.  A
// particle exactly at the centre of an obstacle leaves along +y.
__device__ __forceinline__ bool sphere_hit(const ColliderDev& k, const float4& x, float& ix,
                                           float& iy, float& iz, float& nx, float& ny,
                                           float& nz) {
  const float ox = x.x - k.o[0], oy = x.y - k.o[1], oz = x.z - k.o[2];
  const float d = sqrtf(dot3(ox, oy, oz, ox, oy, oz));
  if (k.container ? !(d >= k.radius) : !(d <= k.radius)) return false;
  if (d == k.radius) {
    ix = x.x; iy = x.y; iz = x.z;
    nx = -ox / d; ny = -oy / d; nz = -oz / d;
    return true;
  }
  float ux = 0.f, uy = 1.f, uz = 0.f;
  if (d > 0.f) { ux = ox / d; uy = oy / d; uz = oz / d; }
  ix = k.o[0] + ux * k.radius; iy = k.o[1] + uy * k.radius; iz = k.o[2] + uz * k.radius;
  if (k.container) { nx = k.o[0] - ix; ny = k.o[1] - iy; nz = k.o[2] - iz; }
  else { nx = ix - k.o[0]; ny = iy - k.o[1]; nz = iz - k.o[2]; }
  return true;
}

// detectionPlan (PlaneCollision_Kernel.cu:6-46) in FP32: inside the corner
// bounds (the normal's axis ignored through the factors 1 - |N|) and on or
// behind the plane, the particle goes to where its step x0 -> x crossed the
// plane.  The reference's factors 1 - N compare the normal's axis at twice
// its coordinate when N = -1, which no particle off the plane passes; a step
// of zero length or parallel to the plane (the reference divides by zero)
// projects x onto the plane instead.
__device__ __forceinline__ bool plane_hit(const ColliderDev& k, const float4& x0, const float4& x,
                                          float& ix, float& iy, float& iz, float& nx, float& ny,
                                          float& nz) {
  const float fx = 1.0f - fabsf(k.nrm[0]), fy = 1.0f - fabsf(k.nrm[1]), fz = 1.0f - fabsf(k.nrm[2]);
  const float px = x.x * fx, py = x.y * fy, pz = x.z * fz;
  if (!(px <= k.mx[0] * fx && px >= k.mn[0] * fx && py <= k.mx[1] * fy && py >= k.mn[1] * fy &&
        pz <= k.mx[2] * fz && pz >= k.mn[2] * fz))
    return false;
  const float d1 = dot3(x.x - k.a[0], x.y - k.a[1], x.z - k.a[2], k.nrm[0], k.nrm[1], k.nrm[2]);
  if (!(d1 <= 0.f)) return false;
  const float aon = dot3(k.a[0] - x0.x, k.a[1] - x0.y, k.a[2] - x0.z, k.nrm[0], k.nrm[1], k.nrm[2]);
  float dx = x.x - x0.x, dy = x.y - x0.y, dz = x.z - x0.z;
  const float dl = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
  float dn = 0.f;
  if (dl > 0.f) {
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    dn = dot3(dx, dy, dz, k.nrm[0], k.nrm[1], k.nrm[2]);
  }
  if (dn != 0.f) {
    const float t = aon / dn;
    ix = x0.x + t * dx; iy = x0.y + t * dy; iz = x0.z + t * dz;
  } else {
    ix = x.x - d1 * k.nrm[0]; iy = x.y - d1 * k.nrm[1]; iz = x.z - d1 * k.nrm[2];
  }
  nx = k.nrm[0]; ny = k.nrm[1]; nz = k.nrm[2];
  return true;
}

// one collider on (x, v) with start-of-step position x0
__device__ __forceinline__ void collide_one(const ColliderDev& k, float dt, const float4& x0,
                                            float4& x, float4& v) {
  float ix, iy, iz, nx, ny, nz;
  const bool hit = k.type == PSB_COLLIDER_SPHERE ? sphere_hit(k, x, ix, iy, iz, nx, ny, nz)
                                                 : plane_hit(k, x0, x, ix, iy, iz, nx, ny, nz);
  if (hit) collider_response(k.elast, dt, x, v, ix, iy, iz, nx, ny, nz);
}

// Index range [lo, hi] along one axis of the lattice sites within h of
// coordinate x (superset; the distance test decides), clipped to the wall
// layers [-L, n + L).
__device__ __forceinline__ void wall_range(const Consts& c, int a, float x, int& lo, int& hi) {
  lo = 0; hi = -1;
  const float u = (x - c.wall_o[a]) * c.wall_inv_s - 0.5f;
  float l = ceilf(u - c.wall_r), r = floorf(u + c.wall_r);
  if (!(l <= r)) return;  // also NaN
  l = fmaxf(l, (float)-c.wall_layers);
  r = fminf(r, (float)(c.wall_n[a] + c.wall_layers - 1));
  if (l <= r) { lo = (int)l; hi = (int)r; }
}

// Contributions of the lattice wall particles within h of (x, y, z): the
// poly6 sum S += (h^2 - d^2)^3 (DENS) and the spiky gradient sum
// g += gradW(x - x_b) (GRAD), sites visited x-major, then y, then z ascending.
// A particle whose candidate sites all lie inside the box skips the loop.
template <bool DENS, bool GRAD>
__device__ __forceinline__ void wall_sums(const Consts& c, float x, float y, float z, float& S,
                                          float& gx, float& gy, float& gz) {
  S = 0.f; gx = 0.f; gy = 0.f; gz = 0.f;
  int lx, hx, ly, hy, lz, hz;
  wall_range(c, 0, x, lx, hx);
  wall_range(c, 1, y, ly, hy);
  wall_range(c, 2, z, lz, hz);
  const int nx = c.wall_n[0], ny = c.wall_n[1], nz = c.wall_n[2];
  if (!(lx < 0 || hx >= nx || ly < 0 || hy >= ny || lz < 0 || hz >= nz)) return;
  for (int ix = lx; ix <= hx; ++ix) {
    const bool xin = ix >= 0 && ix < nx;
    const float dx = x - (c.wall_o[0] + ((float)ix + 0.5f) * c.wall_s);
    for (int iy = ly; iy <= hy; ++iy) {
      const bool yin = iy >= 0 && iy < ny;
      const float dy = y - (c.wall_o[1] + ((float)iy + 0.5f) * c.wall_s);
      // inside on x and y: only the z layers outside [0, nz) are walls
      const int za = lz, zb = (xin && yin) ? min(hz, -1) : hz;
      const int zc = (xin && yin) ? max(lz, nz) : hz + 1, zd = hz;
      for (int part = 0; part < 2; ++part) {
        const int z0 = part == 0 ? za : zc, z1 = part == 0 ? zb : zd;
        for (int iz = z0; iz <= z1; ++iz) {
          const float dz = z - (c.wall_o[2] + ((float)iz + 0.5f) * c.wall_s);
          const float d2 = dot3(dx, dy, dz, dx, dy, dz);
          if (d2 < c.h2) {
            if (DENS) {
              const float w = c.h2 - d2;
              S = __builtin_fmaf(w * w, w, S);
            }
            if (GRAD && d2 > 0.f) {
              const float dd = sqrtf(d2);
              const float invd = 1.0f / dd;
              const float hd = c.h - dd;
              const float gw = c.c_spiky * (hd * hd);
              gx = gx + (dx * invd) * gw;
              gy = gy + (dy * invd) * gw;
              gz = gz + (dz * invd) * gw;
            }
          }
        }
      }
    }
  }
}

// EOS (SphKernel.cu:25 linear; Tait); PSB_MODE_CONFIG clamps at zero.
__device__ __forceinline__ float eos(const Consts& c, float rho) {
  float pr;
  if (c.eos == PSB_EOS_TAIT) {
    float x = rho * c.inv_rho0;
    float x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x7 = x6 * x;
    pr = c.tait_b * (x7 - 1.0f);
  } else {
    pr = c.k * (rho - c.rho0);
  }
  if (c.mode == PSB_MODE_CONFIG && pr < 0.f) pr = 0.f;
  return pr;
}

// Per-particle density epilogue after the neighbour sums.  WCSPH: p =
// EOS(rho); x.w = p/rho^2; v.w = 1/rho.  PCISPH: p = 0 or the reordered
// previous pressure; v.w = 1/rho; with xw_prev, x.w = previous pressure
// (gathered by the warm-start feature sums); lattice walls add their density.
// Rigid-sphere boundary particles get p = 0 and v.w = 0: zero volume in
// every neighbour sum that weights by 1/rho_j (viscosity, surface normals,
// warm-start features).  The neighbour kernel runs it when `on`.
struct EpiArgs {
  int on, pcisph, warm, xw_prev;
  Consts c;
  float* pres;
  const float* pprev;
  float4* xs;
  float4* vs;
  const uint32_t* ids;
};
__device__ __forceinline__ void density_epilogue_one(const Consts& c, const EpiArgs& e, uint32_t p,
                                                     float r, float* __restrict__ rho) {
  if (!e.pcisph) {
    const float pr = eos(c, r);
    rho[p] = r;
    e.pres[p] = pr;
    e.xs[p].w = pr / (r * r);
    e.vs[p].w = 1.0f / r;
  } else if (is_boundary(c, e.ids, p)) {
    rho[p] = r;
    e.pres[p] = 0.f;
    if (e.xw_prev) e.xs[p].w = 0.f;
    e.vs[p].w = 0.f;
  } else {
    e.pres[p] = e.warm ? e.pprev[p] : 0.f;
    if (e.xw_prev) e.xs[p].w = e.pprev[p];
    float rr = r;
    if (c.wall) {
      const float4 x = e.xs[p];
      float S, gx, gy, gz;
      wall_sums<true, false>(c, x.x, x.y, x.z, S, gx, gy, gz);
      rr = r + S * c.mc;
    }
    rho[p] = rr;
    e.vs[p].w = 1.0f / rr;
  }
}

// Visit the ELL neighbour list of particle p in order.  The list loads and
// gathers of U neighbours are issued before any of them is consumed, so a
// wave keeps U dependent load chains in flight; f(q, xq) still sees the
// neighbours in list order, so every sum keeps its sequential association.
template <int U, class F>
__device__ __forceinline__ void for_neighbors(uint32_t p, uint32_t n, uint32_t m,
                                              const uint32_t* __restrict__ list,
                                              const float4* X, F&& f) {
  uint32_t k = 0;
  for (; k + U <= m; k += U) {
    uint32_t q[U];
    float4 xq[U];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = list[(uint64_t)(k + u) * n + p];
#pragma unroll
    for (int u = 0; u < U; ++u) xq[u] = X[q[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) f(q[u], xq[u]);
  }
  for (; k < m; ++k) {
    const uint32_t q = list[(uint64_t)k * n + p];
    f(q, X[q]);
  }
}

// Same with a second per-neighbour vec4 array gathered alongside.
template <int U, class F>
__device__ __forceinline__ void for_neighbors2(uint32_t p, uint32_t n, uint32_t m,
                                               const uint32_t* __restrict__ list,
                                               const float4* X, const float4* V, F&& f) {
  uint32_t k = 0;
  for (; k + U <= m; k += U) {
    uint32_t q[U];
    float4 xq[U], vq[U];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = list[(uint64_t)(k + u) * n + p];
#pragma unroll
    for (int u = 0; u < U; ++u) { xq[u] = X[q[u]]; vq[u] = V[q[u]]; }
#pragma unroll
    for (int u = 0; u < U; ++u) f(q[u], xq[u], vq[u]);
  }
  for (; k < m; ++k) {
    const uint32_t q = list[(uint64_t)k * n + p];
    f(q, X[q], V[q]);
  }
}

constexpr int kUnroll = 4;

// Correctly rounded sqrt(x) and 1/d on the fast path of the IEEE
// expansions: the hardware approximations (within 1 ulp) and the same
// correction steps, without the range scaling and special-value handling
// that x >= 2^-96 (so d = sqrt(x) >= 2^-48, a normal quotient 1/d) never
// needs.  sqrt: the candidates s - ulp, s, s + ulp are decided by the
// residuals fma(-(s -/+ ulp), s, x); 1/d: one Newton step on the
// reciprocal and two quotient corrections (Markstein's scheme).  The
// results equal sqrtf(x) and 1.0f / d bit for bit on that range.
constexpr float kFastSqrtMin = 0x1p-96f;
__device__ __forceinline__ float sqrt_rn_fast(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sdn = __int_as_float(__float_as_int(s) - 1);
  const float sup = __int_as_float(__float_as_int(s) + 1);
  const float rdn = __builtin_fmaf(-sdn, s, x);
  const float rup = __builtin_fmaf(-sup, s, x);
  const float t = rdn <= 0.f ? sdn : s;
  return rup > 0.f ? sup : t;
}
__device__ __forceinline__ float rcp_rn_fast(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  r = __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
  float q = r;
  q = __builtin_fmaf(__builtin_fmaf(-d, q, 1.0f), r, q);
  return __builtin_fmaf(__builtin_fmaf(-d, q, 1.0f), r, q);
}
// sqrtf(x), on the fast path for x >= kFastSqrtMin
__device__ __forceinline__ float sqrt_rn(float x) {
  return __builtin_expect(x >= kFastSqrtMin, 1) ? sqrt_rn_fast(x) : sqrtf(x);
}
// dd = sqrtf(d2), invd = 1.0f / dd for d2 > 0: the fast path always (no
// branch around it), redone by the full expansions in the rare lanes below
// kFastSqrtMin
__device__ __forceinline__ void dist_and_inverse(float d2, float& dd, float& invd) {
  dd = sqrt_rn_fast(d2);
  invd = rcp_rn_fast(dd);
  if (__builtin_expect(d2 < kFastSqrtMin, 0)) {
    dd = sqrtf(d2);
    invd = 1.0f / dd;
  }
}

// 1/sqrt(x) for normal x > 0 from fp32 adds, multiplies and fmas only, so a
// CPU restatement computes the same bits: the bit-pattern seed
// 0x5F375A86 - (bits(x) >> 1) (C. Lomont, "Fast inverse square root", 2003;
// relative error below 3.5 %) and three Newton steps y <- y (3/2 - (x/2) y^2)
// (error 3.5e-2 -> 1.8e-3 -> 4.7e-6 -> fp32 rounding).  No transcendental
// unit, no compare or select: 12 full-rate VALU instructions.
__host__ __device__ __forceinline__ float rsqrt_nr(float x) {
  int32_t i;
  __builtin_memcpy(&i, &x, 4);
  i = 0x5F375A86 - (i >> 1);
  float y;
  __builtin_memcpy(&y, &i, 4);
  const float hx = 0.5f * x;
  y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
  y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
  y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
  return y;
}
// Squared distances below this are taken at this value in the config-mode
// pressure force, so rsqrt_nr sees a normal input and the self entry
// (d2 = 0, dx = 0) adds exactly zero.
constexpr float kMinPairD2 = 0x1p-96f;

// Per-step maxima reported by many workgroups, spread over kErrSlots words
// kErrSlotStride or more apart (different memory channels): a workgroup
// takes slot blockIdx % kErrSlots, so the atomicMax do not serialise on one
// address (one address cost ~8.6 us per PCISPH density launch over 1024
// workgroups, and ~22 us per neighbour search over 4096 waves, on MI355X).
// Slot s, scene c, entry k: base[s * stride + c * kMaxIterSlots + k]; entry k
// < kMaxIterSlots - 1 is iteration k's compression error, entry
// kMaxIterSlots - 1 the step's largest neighbour count.  base == nullptr:
// single words instead (PciStatus::err, the count flag word).
constexpr int kErrSlots = 16;
constexpr uint32_t kErrSlotStride = 16384;   // words (64 KiB)
constexpr int kCountEntry = kMaxIterSlots - 1;
struct ErrSlots {
  uint32_t* base = nullptr;
  uint32_t stride = 0;
};
inline uint32_t err_slot_stride(uint32_t scenes) {
  const uint64_t need = (uint64_t)scenes * kMaxIterSlots;
  return (uint32_t)((need + kErrSlotStride - 1) / kErrSlotStride * kErrSlotStride);
}

// Neighbour list readers: the ELL list of 32-bit indices (entry k of
// particle p at k * n + p), and its compact twin of 16-bit entries that
// halves the list bytes the corrector passes stream every iteration.  The
// compact list keeps 8 consecutive entries of one particle in one 16-byte
// word, [k / 8][p][k % 8] (l16_at), so one vector load per thread brings 8
// indices.  Entry: offset q - p + 32768 (|q - p| <= 32767), or, in a tile
// staged in LDS, the slot of q's staged copy.
__host__ __device__ __forceinline__ uint64_t l16_at(uint32_t k, uint32_t p, uint32_t n) {
  return ((uint64_t)(k >> 3) * n + p) * 8u + (k & 7u);
}
constexpr int kL16Group = 8;
inline size_t l16_bytes(size_t n, uint32_t cap) {
  return 2u * (size_t)kL16Group * n * ((cap + kL16Group - 1) / kL16Group);
}
// Grouped 32-bit list: absolute indices, four consecutive entries of one
// particle in one 16-byte word, [k / 4][p][k % 4] (l32g_at), so a wave's
// load of one group is contiguous and a reader takes 8 entries in two
// loads; the corrector passes read it where the compact offsets overflow.
// Allocated in whole groups of 8 (the readers' unit).
__host__ __device__ __forceinline__ uint64_t l32g_at(uint32_t k, uint32_t p, uint32_t n) {
  return ((uint64_t)(k >> 2) * n + p) * 4u + (k & 3u);
}
inline size_t l32g_bytes(size_t n, uint32_t cap) {
  return 4u * 8u * n * ((cap + 7u) / 8u);
}
struct List32 {
  static constexpr bool kGrouped = false;
  static constexpr bool kGrouped32 = false;
  const uint32_t* __restrict__ l;
  uint32_t n;
  __device__ __forceinline__ uint32_t at(uint32_t k, uint32_t p) const {
    return l[(uint64_t)k * n + p];
  }
};
struct List32G {
  static constexpr bool kGrouped = false;
  static constexpr bool kGrouped32 = true;
  const uint32_t* __restrict__ l;
  uint32_t n;
  __device__ __forceinline__ uint32_t at(uint32_t k, uint32_t p) const {
    return l[l32g_at(k, p, n)];
  }
};
struct List16 {
  static constexpr bool kGrouped = true;
  static constexpr bool kGrouped32 = false;
  const uint16_t* __restrict__ l;
  uint32_t n;
  __device__ __forceinline__ uint32_t decode(uint32_t e, uint32_t p) const {
    return p + e - 32768u;
  }
  __device__ __forceinline__ uint32_t safe(uint32_t p) const { return p; }
  __device__ __forceinline__ uint32_t at(uint32_t k, uint32_t p) const {
    return decode(l[l16_at(k, p, n)], p);
  }
};
constexpr int kCompactReach = 32767;
// LDS slots of a staged tile (clamped to the staged count)
struct ListSlot {
  static constexpr bool kGrouped = true;
  static constexpr bool kGrouped32 = false;
  const uint16_t* __restrict__ l;
  uint32_t n;
  uint32_t last;
  __device__ __forceinline__ uint32_t decode(uint32_t e, uint32_t) const { return min(e, last); }
  __device__ __forceinline__ uint32_t safe(uint32_t) const { return 0u; }
  __device__ __forceinline__ uint32_t at(uint32_t k, uint32_t p) const {
    return decode(l[l16_at(k, p, n)], p);
  }
};

// Compact list output of the neighbour search: list16 (same ELL shape as
// the 32-bit list) and a flag per scene, wide[scene * wide_stride], set when
// some offset does not fit (the readers then take the 32-bit list).
struct Compact16 {
  uint16_t* list16 = nullptr;
  uint32_t* wide = nullptr;
  uint32_t wide_stride = 0;
  // grouped 32-bit list (l32g_at), written instead of list16 when *want32 is
  // set and no tile is staged (single system)
  uint32_t* list32g = nullptr;
  const uint32_t* want32 = nullptr;
};

// Neighbour sums of the corrector passes in their canonical order: two
// partial sums, over the even and over the odd list positions k, each in
// increasing k, added as even + odd.  LANES = 1: one thread keeps both
// (f(acc = k & 1, ...)); LANES = 2: lane l of the particle's pair keeps the
// sum over k = l (mod 2) (f(acc = 0, ...)) and split_total adds the pair.
// the 8 entries of one compact-list word, in entry order
__device__ __forceinline__ void unpack_l16(const uint4& c, uint32_t* e) {
  e[0] = c.x & 0xFFFFu; e[1] = c.x >> 16; e[2] = c.y & 0xFFFFu; e[3] = c.y >> 16;
  e[4] = c.z & 0xFFFFu; e[5] = c.z >> 16; e[6] = c.w & 0xFFFFu; e[7] = c.w >> 16;
}

// X[i] through a 32-bit byte offset (i < 2^27 particles: every system and
// batch is created below that), so the gather takes the array's base from
// scalar registers and the offset from one VGPR: one shift-add per
// neighbour instead of a 64-bit address computation.
__device__ __forceinline__ float4 gather4(const float4* X, uint32_t i) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(X) + (i << 4));
}

template <int LANES, int U, class LI, class F>
__device__ __forceinline__ void for_neighbors_split(uint32_t p, uint32_t n, uint32_t m,
                                                    uint32_t lane, const LI& list,
                                                    const float4* X, F&& f) {
  static_assert(U % 2 == 0, "U even keeps k & 1 static");
  if constexpr (LANES == 1 && LI::kGrouped) {
    // one 16-byte load per 8 entries, the next group's load issued before
    // the current group's gathers.  Full groups run without per-entry
    // tests (so every gather of a group is in flight at once); in the last,
    // partial group the entries past m select a safe index and are not
    // summed.
    const uint4* __restrict__ g4 = reinterpret_cast<const uint4*>(list.l);
    const uint32_t ng = (m + kL16Group - 1) / kL16Group, nfull = m / kL16Group;
    uint4 w = make_uint4(0u, 0u, 0u, 0u);
    if (ng) w = g4[p];
    for (uint32_t g = 0; g < nfull; ++g) {
      const uint4 cur = w;
      if (g + 1 < ng) w = g4[(uint64_t)(g + 1) * n + p];
      uint32_t e[8];
      unpack_l16(cur, e);
      float4 xq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xq[u] = gather4(X, list.decode(e[u], p));
#pragma unroll
      for (int u = 0; u < 8; ++u) f(u & 1, xq[u]);
    }
    if (nfull < ng) {
      uint32_t e[8];
      unpack_l16(w, e);
      const uint32_t r = m - nfull * kL16Group;   // 1..7 entries left
      float4 xq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        xq[u] = gather4(X, (uint32_t)u < r ? list.decode(e[u], p) : list.safe(p));
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if ((uint32_t)u < r) f(u & 1, xq[u]);
    }
  } else if constexpr (LANES == 2 && LI::kGrouped) {
    // two lanes per particle on the grouped list: both load the group's
    // word, lane l takes its entries k = l (mod 2) (4 gathers per group)
    const uint4* __restrict__ g4 = reinterpret_cast<const uint4*>(list.l);
    const uint32_t ng = (m + kL16Group - 1) / kL16Group, nfull = m / kL16Group;
    const uint32_t sh = lane << 4;   // this lane's half of each 32-bit pair of entries
    uint4 w = make_uint4(0u, 0u, 0u, 0u);
    if (ng) w = g4[p];
    // the prefetch index is clamped to the last group: an unconditional load
    // leaves no conditional value to carry (no register copies per group)
    for (uint32_t g = 0; g < nfull; ++g) {
      const uint4 cur = w;
      w = g4[(uint64_t)min(g + 1, ng - 1) * n + p];
      const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
      float4 xq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        xq[u] = gather4(X, list.decode(__builtin_amdgcn_ubfe(wd[u], sh, 16), p));
#pragma unroll
      for (int u = 0; u < 4; ++u) f(0, xq[u]);
    }
    if (nfull < ng) {
      const uint32_t wd[4] = {w.x, w.y, w.z, w.w};
      const uint32_t r = m - nfull * kL16Group;   // 1..7 entries left
      float4 xq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = 2u * u + lane;
        xq[u] = gather4(X, j < r ? list.decode(__builtin_amdgcn_ubfe(wd[u], sh, 16), p)
                                 : list.safe(p));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (2u * u + lane < r) f(0, xq[u]);
    }
  } else if constexpr (LANES == 1 && LI::kGrouped32) {
    // the grouped 32-bit list: 8 absolute indices in two 16-byte loads, the
    // next group's loads issued before the current group's gathers
    const uint4* __restrict__ g4 = reinterpret_cast<const uint4*>(list.l);
    const uint32_t ng = (m + 7u) / 8u, nfull = m / 8u;
    uint4 wa = make_uint4(0u, 0u, 0u, 0u), wb = wa;
    if (ng) { wa = g4[p]; wb = g4[(uint64_t)n + p]; }
    for (uint32_t g = 0; g < nfull; ++g) {
      const uint4 ca = wa, cb = wb;
      if (g + 1 < ng) {
        wa = g4[(uint64_t)(2u * g + 2u) * n + p];
        wb = g4[(uint64_t)(2u * g + 3u) * n + p];
      }
      const uint32_t e[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
      float4 xq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xq[u] = gather4(X, e[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u) f(u & 1, xq[u]);
    }
    if (nfull < ng) {
      const uint32_t e[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
      const uint32_t r = m - nfull * 8u;   // 1..7 entries left
      float4 xq[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xq[u] = gather4(X, (uint32_t)u < r ? e[u] : p);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if ((uint32_t)u < r) f(u & 1, xq[u]);
    }
  } else if constexpr (LANES == 2 && LI::kGrouped32) {
    // two lanes on the grouped 32-bit list: lane l takes entries k = l
    // (mod 2) of each group of 8 (4 gathers per group and lane)
    const uint4* __restrict__ g4 = reinterpret_cast<const uint4*>(list.l);
    const uint32_t ng = (m + 7u) / 8u, nfull = m / 8u;
    uint4 wa = make_uint4(0u, 0u, 0u, 0u), wb = wa;
    if (ng) { wa = g4[p]; wb = g4[(uint64_t)n + p]; }
    for (uint32_t g = 0; g < nfull; ++g) {
      const uint4 ca = wa, cb = wb;
      const uint32_t gn = 2u * min(g + 1, ng - 1);
      wa = g4[(uint64_t)gn * n + p];
      wb = g4[(uint64_t)(gn + 1u) * n + p];
      const uint32_t e[4] = {lane ? ca.y : ca.x, lane ? ca.w : ca.z, lane ? cb.y : cb.x,
                             lane ? cb.w : cb.z};
      float4 xq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) xq[u] = gather4(X, e[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u) f(0, xq[u]);
    }
    if (nfull < ng) {
      const uint32_t e[4] = {lane ? wa.y : wa.x, lane ? wa.w : wa.z, lane ? wb.y : wb.x,
                             lane ? wb.w : wb.z};
      const uint32_t r = m - nfull * 8u;   // 1..7 entries left
      float4 xq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) xq[u] = gather4(X, 2u * u + lane < r ? e[u] : p);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (2u * u + lane < r) f(0, xq[u]);
    }
  } else if constexpr (LANES == 1) {
    uint32_t k = 0;
    for (; k + U <= m; k += U) {
      uint32_t q[U];
      float4 xq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) q[u] = list.at(k + u, p);
#pragma unroll
      for (int u = 0; u < U; ++u) xq[u] = X[q[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) f(u & 1, xq[u]);
    }
    for (; k < m; ++k) f((int)(k & 1u), X[list.at(k, p)]);
  } else {
    uint32_t k = lane;
    for (; k + 2 * (U - 1) < m; k += 2 * U) {
      uint32_t q[U];
      float4 xq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) q[u] = list.at(k + 2 * u, p);
#pragma unroll
      for (int u = 0; u < U; ++u) xq[u] = X[q[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) f(0, xq[u]);
    }
    for (; k < m; k += 2) f(0, X[list.at(k, p)]);
  }
}

// Visit a list (List32 / List16 accessor) in order, gathering two vec4
// arrays per neighbour: f(xq, vq).  The grouped compact list brings 8
// indices per vector load (next group's load issued first); entries past m
// select a safe index and are not visited.
template <int U, class LI, class F>
__device__ __forceinline__ void for_pairs2(uint32_t p, uint32_t n, uint32_t m, const LI& list,
                                           const float4* __restrict__ X,
                                           const float4* __restrict__ V, F&& f) {
  if constexpr (LI::kGrouped) {
    const uint4* __restrict__ g4 = reinterpret_cast<const uint4*>(list.l);
    const uint32_t ng = (m + kL16Group - 1) / kL16Group;
    uint4 w = make_uint4(0u, 0u, 0u, 0u);
    if (ng) w = g4[p];
    // prefetch index clamped to the last group (an unconditional load)
    for (uint32_t g = 0; g < ng; ++g) {
      const uint4 cur = w;
      w = g4[(uint64_t)min(g + 1, ng - 1) * n + p];
      const uint32_t e[8] = {cur.x & 0xFFFFu, cur.x >> 16, cur.y & 0xFFFFu, cur.y >> 16,
                             cur.z & 0xFFFFu, cur.z >> 16, cur.w & 0xFFFFu, cur.w >> 16};
      const uint32_t k = g * kL16Group;
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        float4 xq[4], vq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t q = k + h + u < m ? list.decode(e[h + u], p) : list.safe(p);
          xq[u] = X[q];
          vq[u] = V[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (k + h + u < m) f(xq[u], vq[u]);
      }
    }
  } else {
    uint32_t k = 0;
    for (; k + U <= m; k += U) {
      uint32_t q[U];
      float4 xq[U], vq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) q[u] = list.at(k + u, p);
#pragma unroll
      for (int u = 0; u < U; ++u) { xq[u] = X[q[u]]; vq[u] = V[q[u]]; }
#pragma unroll
      for (int u = 0; u < U; ++u) f(xq[u], vq[u]);
    }
    for (; k < m; ++k) {
      const uint32_t q = list.at(k, p);
      f(X[q], V[q]);
    }
  }
}

// even + odd partial sums (see for_neighbors_split); both lanes of a pair
// get the same value
template <int LANES>
__device__ __forceinline__ float split_total(const float (&a)[2], uint32_t lane) {
  if (LANES == 1) return a[0] + a[1];
  const float other = __shfl_xor(a[0], 1, 64);
  return lane == 0 ? a[0] + other : other + a[0];
}

template <int N> __device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = N / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, N));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------
}  // namespace psb

struct psb_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // scratch for the standalone wrappers (grown on demand)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  uint32_t* d_flags = nullptr;     // device u32 flags for standalone calls
  uint32_t* h_pinned = nullptr;    // pinned host staging (4 KiB)
};

struct psb_grid_s {
  psb_context ctx = nullptr;
  psb_grid_desc desc{};
  uint64_t ncells = 0;
  uint32_t capacity = 0;
  uint32_t n = 0;                  // particles of the last store
  uint32_t* count = nullptr;       // [ncells], zero between builds
  uint32_t* start = nullptr;       // [ncells + 1]
  uint32_t* hash = nullptr;        // [cap]
  uint32_t* slot = nullptr;        // [cap]
  uint32_t* tmp = nullptr;         // [cap]
  uint32_t* perm = nullptr;        // [cap]
  uint32_t* bsum = nullptr;        // [scan blocks]: tile offsets
  uint32_t* tsum = nullptr;        // [16][scan blocks]: tile counts, zero between builds
  uint32_t* flags = nullptr;       // [1]
  uint32_t nscan = 0;
};

namespace psb {

psb_status ensure_scratch(psb_context ctx, size_t bytes);
GridDev grid_dev(const psb_grid_desc& d, float h);

// Particle state arrays moved by the stable reorder (src -> dst[rank]).
struct ReorderArrays {
  const float4* sx; const float4* sv; const float4* svh; const uint32_t* sid; const float* sp;
  float4* dx; float4* dv; float4* dvh; uint32_t* did; float* dp;
};

// grid.hip
struct StepReset;   // pcisph.h
// reset: the PCISPH per-step status reset, run by the hashing pass
psb_status grid_build(psb_grid g, const float4* pos, uint32_t n, const ReorderArrays* ra,
                      hipStream_t s, const StepReset* reset = nullptr);
psb_status launch_neighbors(psb_grid g, const float4* xs, uint32_t n, float r2, int R,
                            psb_neighbors nb, float* rho_sum_scaled, float mc,
                            uint32_t* flags, uint32_t* maxcnt, hipStream_t s,
                            const Compact16& c16 = Compact16{}, ErrSlots es = ErrSlots{},
                            const HaloTile* tiles = nullptr, const EpiArgs* ep = nullptr);
// halo tables of the tiles of n sorted particles xs (after grid_build):
// tiles[ceil(n / kHaloTile)], src[tiles * kHaloCap]
psb_status launch_halo(psb_grid g, const float4* xs, uint32_t n, int R, HaloTile* tiles,
                       uint32_t* src, hipStream_t s);
// batches: g spans the cells of every scene (scene s owns cells
// [cell_off, cell_off + its grid)); nblocks x bsize threads, scene-aligned
psb_status grid_build_batch(psb_grid g, const BatchMap& map, uint32_t nblocks, uint32_t bsize,
                            const float4* pos, uint32_t n, const ReorderArrays* ra,
                            hipStream_t s, const StepReset* reset = nullptr);
psb_status launch_neighbors_batch(psb_grid g, const BatchMap& map, uint32_t nblocks,
                                  uint32_t bsize, const float4* xs, uint32_t n, float r2, int R,
                                  psb_neighbors nb, float* rho, float mc, uint32_t* flags,
                                  uint32_t* maxcnt, hipStream_t s,
                                  const Compact16& c16 = Compact16{}, ErrSlots es = ErrSlots{},
                                  const EpiArgs* ep = nullptr);

}  // namespace psb
