// oracle.cpp -- CPU restatement of the PSB_Project SPH/PCISPH timestep.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg use it, as the checker.
//
// It restates the reference kernels (file:line cited per function) in plain
// C++ with OpenMP over particles, templated on the scalar type:
//   - R = float  : the canonical FP32 arithmetic.  Every operation is written
//                  in a fixed order (no contraction: built with
//                  -ffp-contract=off; fused multiply-adds only where std::fma
//                  is written), so an FP32 GPU implementation following the
//                  same order can be compared bit for bit;
//   - R = double : the same algorithm in FP64, for accuracy reports.
//
// Two semantic modes (psb_sph_params.mode):
//   PSB_MODE_REFERENCE  what the reference computes (linear EOS, exactly three
//                       PCISPH iterations with a per-particle delta, sigma
//                       forced to 0.0728 and g = -9.81 y inside PCISPH, the box
//                       collider ignoring restitution);
//   PSB_MODE_CONFIG     the configured semantics (Tait EOS, tolerance-driven
//                       PCISPH with a template delta and clamped pressure,
//                       restitution).
//
// The neighbour search replaces the reference's atomic 64-slot buckets by a
// stable counting sort: the particles of a cell appear in ascending index
// order, which is the order the reference's buckets hold when filled serially
// (UniformGridKernel.cu:43-55).  Neighbour lists are in ascending sorted index
// order (rows of the stencil visited z, then y, then x ascending).
//
// What pins each function (golden values, brute force, closed forms,
// invariants) is in tests/test_oracle_pinning.py, tests/test_wall_oracle.py
// and tests/test_mlp_oracle.py.  Config-mode extensions with no reference
// counterpart: the per-particle delta cap, the lattice walls and the learned
// warm start (features, MLP).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/psb_sph.h"

namespace {

template <class R> struct V4 { R x, y, z, w; };

// Loops over fewer items than this run on the calling thread: thread start-up
// and barriers would cost more than the work (small test scenes).
constexpr int64_t kOmpMinItems = 2048;

// 1/sqrt(x) of the config-mode pressure force (x normal, > 0).  FP32: the
// kernel's sequence, operation for operation -- the bit-pattern seed
// 0x5F375A86 - (bits(x) >> 1) (C. Lomont, "Fast inverse square root", 2003)
// and three Newton steps y <- y (3/2 - (x/2) y^2), each an fp32 multiply or
// fma -- so both give the same bits.  FP64: 1/sqrt(x).
inline float rsqrt_pair(float x) {
  int32_t i;
  std::memcpy(&i, &x, 4);
  i = 0x5F375A86 - (i >> 1);
  float y;
  std::memcpy(&y, &i, 4);
  const float hx = 0.5f * x;
  y = y * std::fma(-(hx * y), y, 1.5f);
  y = y * std::fma(-(hx * y), y, 1.5f);
  y = y * std::fma(-(hx * y), y, 1.5f);
  return y;
}
inline double rsqrt_pair(double x) { return 1.0 / std::sqrt(x); }
// squared distances below this are taken at this value (self entry: dx = 0)
constexpr float kMinPairD2 = 0x1p-96f;
// the corrector's divergence stop: error above this factor times its smallest so far
constexpr float kDivergeFactor = 2.0f;
constexpr float kDivergeFloor = 4.0f;   // ... and above this times the tolerance

// ---------------------------------------------------------------------------
// Derived constants.  Computed in double from the float parameters, then
// rounded once to R.  SphKernel.cu:19-20 (poly6), :59-60 (spiky gradient),
// :64 (viscosity Laplacian), :69-76 (surface Laplacian / gradient);
// PciSphSystem_Kernel.cu:44 (beta).
// ---------------------------------------------------------------------------
template <class R> struct Derived {
  R h, h2, h2x3, c_poly6, c_spiky, c_visc, c_surf_lap, c_surf_grad;
  R mc;           // m * c_poly6
  R beta;         // 2 (m dt / rho0)^2
  R delta;        // template delta
  R tait_b, inv_rho0, rho0, k, kp;   // kp = -m / rho0^2
  R kps;          // kp * c_spiky, rounded once
  R cv, cs, cn;   // mu m c_visc, m c_surf_lap, m c_surf_grad
  R m, dt, sigma, l, gx, gy, gz;
  R box_lo[3], box_hi[3], one_plus_e;
  int R_stencil;
  R pscale, inv_ps, h_over_ps, dt_over_h;   // warm-start feature scales
  bool wall;                                 // lattice box walls
  int wall_n[3], wall_layers;
  R wall_o[3], wall_s, wall_inv_s, wall_r;
};

double template_w2(double h, double s) {
  // Sum over a cubic lattice of spacing s of |grad W_spiky|^2 for all
  // offsets 0 < d < h (grad W = -45 (h-d)^2 / (pi h^6)).  Sum of grad W over
  // the full neighbourhood is 0 by symmetry, so W1.W1 = 0.
  const double c = 45.0 / (M_PI * std::pow(h, 6));
  int n = (int)std::ceil(h / s) + 1;
  double w2 = 0.0;
  for (int a = -n; a <= n; ++a)
    for (int b = -n; b <= n; ++b)
      for (int e = -n; e <= n; ++e) {
        double d = s * std::sqrt((double)(a * a + b * b + e * e));
        if (d > 0.0 && d < h) { double g = c * (h - d) * (h - d); w2 += g * g; }
      }
  return w2;
}

template <class R> Derived<R> derive(const psb_sph_params& p) {
  Derived<R> d;
  const double h = p.support_radius, m = p.mass, rho0 = p.rest_density, dt = p.dt;
  const double pi = M_PI;
  d.h = (R)h;
  d.h2 = (R)(h * h);
  d.h2x3 = (R)(3.0 * h * h);
  d.c_poly6 = (R)(315.0 / (64.0 * pi * std::pow(h, 9)));
  d.c_spiky = (R)(-45.0 / (pi * std::pow(h, 6)));
  d.c_visc = (R)(45.0 / (pi * std::pow(h, 6)));
  d.c_surf_lap = (R)(-945.0 / (32.0 * pi * std::pow(h, 9)));
  d.c_surf_grad = (R)(-945.0 / (32.0 * pi * std::pow(h, 9)));
  d.mc = (R)(m * 315.0 / (64.0 * pi * std::pow(h, 9)));
  d.beta = (R)(2.0 * std::pow(m * dt / rho0, 2));
  const double beta = 2.0 * std::pow(m * dt / rho0, 2);
  if (p.delta > 0.f) d.delta = (R)p.delta;
  else {
    double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    d.delta = (R)(1.0 / (beta * template_w2(h, s)));
  }
  // under-relaxed corrector (config mode); the feature scale keeps the template
  if (p.mode == PSB_MODE_CONFIG && p.relaxation > 0.f && p.relaxation != 1.0f)
    d.delta = d.delta * (R)p.relaxation;
  const double gamma = p.tait_gamma > 0.f ? p.tait_gamma : 7.0;
  d.tait_b = (R)(p.tait_b > 0.f ? (double)p.tait_b : (double)p.gas_stiffness * rho0 / gamma);
  d.inv_rho0 = (R)(1.0 / rho0);
  d.rho0 = (R)rho0;
  d.k = (R)p.gas_stiffness;
  d.kp = (R)(-m / (rho0 * rho0));
  d.kps = (R)((-m / (rho0 * rho0)) * (-45.0 / (pi * std::pow(h, 6))));
  d.cv = (R)((double)p.viscosity * m * 45.0 / (pi * std::pow(h, 6)));
  d.cs = (R)(m * -945.0 / (32.0 * pi * std::pow(h, 9)));
  d.cn = (R)(m * -945.0 / (32.0 * pi * std::pow(h, 9)));
  d.m = (R)m;
  d.dt = (R)dt;
  d.sigma = (R)p.surface_tension;
  d.l = (R)p.threshold;
  d.gx = (R)p.gravity[0]; d.gy = (R)p.gravity[1]; d.gz = (R)p.gravity[2];
  for (int a = 0; a < 3; ++a) {
    // BoxCollision_Kernel.cu:17-23 compares against Max - r and Min + r,
    // both evaluated in float.
    d.box_lo[a] = (R)(float)(p.box_min[a] + p.particle_radius);
    d.box_hi[a] = (R)(float)(p.box_max[a] - p.particle_radius);
  }
  d.one_plus_e = (R)(p.mode == PSB_MODE_REFERENCE ? 1.0f : 1.0f + p.elast);
  double rr = std::ceil((double)h / (double)p.grid.cell_size);
  if ((rr - 1.0) * (double)p.grid.cell_size >= (double)h) rr -= 1.0;
  d.R_stencil = (int)rr;
  {
    double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    const double ps = p.delta > 0.f ? (double)p.delta * rho0 * 0.01
                                     : rho0 * 0.01 / (beta * template_w2(h, s));
    d.pscale = (R)(float)ps;
    d.inv_ps = (R)(float)(1.0 / ps);
    d.h_over_ps = (R)(float)(h / ps);
    d.dt_over_h = (R)(float)(dt / h);
  }
  d.wall = p.mode == PSB_MODE_CONFIG && p.solver == PSB_SOLVER_PCISPH &&
           p.wall_model == PSB_WALL_LATTICE && p.box_enabled;
  {
    const double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    for (int a = 0; a < 3; ++a) {
      d.wall_o[a] = (R)p.box_min[a];
      d.wall_n[a] = (int)std::floor(((double)p.box_max[a] - p.box_min[a]) / s + 1e-3);
    }
    d.wall_s = (R)(float)s;
    d.wall_inv_s = (R)(float)(1.0 / s);
    d.wall_r = (R)(float)(h / s);
    d.wall_layers = (int)std::ceil(h / s) + 1;
  }
  return d;
}


template <class R> inline R fmaR(R a, R b, R c) { return std::fma(a, b, c); }
template <class R> inline R dot3(R ax, R ay, R az, R bx, R by, R bz) {
  return fmaR(ax, bx, fmaR(ay, by, az * bz));
}

// Lattice box walls (PSB_WALL_LATTICE): implicit boundary particles of mass
// m at wall_o + (i + 1/2) s for each lattice index outside [0, wall_n) on
// some axis, within wall_layers of the box.  wall_range gives the candidate
// index range along one axis (the distance test decides); wall_sums adds the
// poly6 numerators (h^2 - d^2)^3 and the spiky gradients of the sites within
// h, visited x-major, then y, then z ascending.
template <class R>
inline void wall_range(const Derived<R>& d, int a, R x, int& lo, int& hi) {
  lo = 0; hi = -1;
  const R u = (x - d.wall_o[a]) * d.wall_inv_s - (R)0.5;
  R l = std::ceil(u - d.wall_r), r = std::floor(u + d.wall_r);
  if (!(l <= r)) return;
  l = std::max(l, (R)-d.wall_layers);
  r = std::min(r, (R)(d.wall_n[a] + d.wall_layers - 1));
  if (l <= r) { lo = (int)l; hi = (int)r; }
}

template <class R>
inline void wall_sums(const Derived<R>& d, bool dens, bool grad, R x, R y, R z, R& S, R& gx,
                      R& gy, R& gz) {
  S = 0; gx = 0; gy = 0; gz = 0;
  int lx, hx, ly, hy, lz, hz;
  wall_range(d, 0, x, lx, hx);
  wall_range(d, 1, y, ly, hy);
  wall_range(d, 2, z, lz, hz);
  const int nx = d.wall_n[0], ny = d.wall_n[1], nz = d.wall_n[2];
  if (!(lx < 0 || hx >= nx || ly < 0 || hy >= ny || lz < 0 || hz >= nz)) return;
  for (int ix = lx; ix <= hx; ++ix) {
    const bool xin = ix >= 0 && ix < nx;
    const R dx = x - (d.wall_o[0] + ((R)ix + (R)0.5) * d.wall_s);
    for (int iy = ly; iy <= hy; ++iy) {
      const bool yin = iy >= 0 && iy < ny;
      const R dy = y - (d.wall_o[1] + ((R)iy + (R)0.5) * d.wall_s);
      for (int iz = lz; iz <= hz; ++iz) {
        if (xin && yin && iz >= 0 && iz < nz) continue;   // inside the box
        const R dz = z - (d.wall_o[2] + ((R)iz + (R)0.5) * d.wall_s);
        const R d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (d2 < d.h2) {
          if (dens) {
            const R w = d.h2 - d2;
            S = fmaR(w * w, w, S);
          }
          if (grad && d2 > (R)0) {
            const R dd = std::sqrt(d2);
            const R invd = (R)1 / dd;
            const R hd = d.h - dd;
            const R gw = d.c_spiky * (hd * hd);
            gx = gx + (dx * invd) * gw;
            gy = gy + (dy * invd) * gw;
            gz = gz + (dz * invd) * gw;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Grid hash: UniformGridKernel.cu:48-51 (cell = floor((x - Min)/cell),
// flat index i + j nx + k nx ny).  Out-of-grid coordinates are clamped to the
// edge cell and flagged (the reference writes out of bounds there).
// ---------------------------------------------------------------------------
inline int cell_coord(float x, float mn, float cell, int n, uint32_t& flags) {
  float t = std::floor((x - mn) / cell);
  if (!(t >= 0.0f && t < (float)n)) {
    flags |= PSB_FLAG_OUT_OF_GRID;
    t = (t >= (float)n) ? (float)(n - 1) : 0.0f;
  }
  return (int)t;
}

struct Grid {
  float min[3], cell;
  int nx, ny, nz;
  uint64_t ncells() const { return (uint64_t)nx * ny * nz; }
};

Grid grid_from(const psb_grid_desc& g) {
  Grid r;
  for (int a = 0; a < 3; ++a) r.min[a] = g.min[a];
  r.cell = g.cell_size; r.nx = g.dims[0]; r.ny = g.dims[1]; r.nz = g.dims[2];
  return r;
}

inline void cell_of(const Grid& g, float x, float y, float z, int& cx, int& cy, int& cz,
                    uint32_t& flags) {
  cx = cell_coord(x, g.min[0], g.cell, g.nx, flags);
  cy = cell_coord(y, g.min[1], g.cell, g.ny, flags);
  cz = cell_coord(z, g.min[2], g.cell, g.nz, flags);
}

// Stable counting sort by hash: perm lists particle indices in (hash, index)
// order, start[c] = #particles with hash < c.
void grid_build(const Grid& g, uint32_t n, const float* x4, uint32_t* hash, uint32_t* perm,
                uint32_t* start, uint32_t& flags) {
  const uint64_t nc = g.ncells();
  for (uint32_t i = 0; i < n; ++i) {
    int cx, cy, cz;
    cell_of(g, x4[4 * i], x4[4 * i + 1], x4[4 * i + 2], cx, cy, cz, flags);
    if (!std::isfinite(x4[4 * i]) || !std::isfinite(x4[4 * i + 1]) || !std::isfinite(x4[4 * i + 2]))
      flags |= PSB_FLAG_NONFINITE;
    hash[i] = (uint32_t)cx + (uint32_t)g.nx * ((uint32_t)cy + (uint32_t)g.ny * (uint32_t)cz);
  }
  std::vector<uint32_t> cnt(nc + 1, 0);
  for (uint32_t i = 0; i < n; ++i) cnt[hash[i]]++;
  uint32_t run = 0;
  for (uint64_t c = 0; c < nc; ++c) { start[c] = run; run += cnt[c]; }
  start[nc] = run;
  std::vector<uint32_t> fill(start, start + nc);
  for (uint32_t i = 0; i < n; ++i) perm[fill[hash[i]]++] = i;
}

// searchNeighbooring restated (UniformGridKernel.cu:58-95): all j with
// |x_i - x_j|^2 <= radius^2, self included, over (2R+1)^3 cells.  Rows of
// the stencil are visited z-major, then y, then x ascending, i.e. ascending
// sorted index.  Density of the same pairs (SphKernel.cu:9-27) is accumulated
// when rho_sum != nullptr: S = sum (h^2 - d^2)^3.
template <class R>
uint32_t neighbors(const Grid& g, const uint32_t* start, uint32_t n, const V4<R>* xs,
                   R radius2, int Rs, uint32_t cap, uint32_t* count, uint32_t* list,
                   R* rho_sum, uint32_t* maxcnt) {
  uint32_t flags = 0, mx = 0;
#pragma omp parallel for schedule(static) reduction(| : flags) reduction(max : mx) if(n >= kOmpMinItems)
  for (int64_t pi = 0; pi < (int64_t)n; ++pi) {
    const uint32_t p = (uint32_t)pi;
    const V4<R> xi = xs[p];
    int cx, cy, cz;
    uint32_t f = 0;
    cell_of(g, (float)xi.x, (float)xi.y, (float)xi.z, cx, cy, cz, f);
    uint32_t cnt = 0;
    R S = 0;
    for (int dz = -Rs; dz <= Rs; ++dz) {
      int z = cz + dz;
      if (z < 0 || z >= g.nz) continue;
      for (int dy = -Rs; dy <= Rs; ++dy) {
        int y = cy + dy;
        if (y < 0 || y >= g.ny) continue;
        int x0 = std::max(cx - Rs, 0), x1 = std::min(cx + Rs, g.nx - 1);
        uint64_t base = (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * z);
        uint32_t q0 = start[base + x0], q1 = start[base + x1 + 1];
        for (uint32_t q = q0; q < q1; ++q) {
          const V4<R> xj = xs[q];
          R dx = xi.x - xj.x, dy_ = xi.y - xj.y, dz_ = xi.z - xj.z;
          R d2 = dot3(dx, dy_, dz_, dx, dy_, dz_);
          if (d2 <= radius2) {
            if (cnt < cap) list[(uint64_t)cnt * n + p] = q;
            cnt++;
            R w = radius2 - d2;
            S = fmaR(w * w, w, S);
          }
        }
      }
    }
    if (cnt > cap) f |= PSB_FLAG_NEIGHBOR_OVERFLOW;
    count[p] = std::min(cnt, cap);
    if (rho_sum) rho_sum[p] = S;
    flags |= (f & PSB_FLAG_NEIGHBOR_OVERFLOW);
    mx = std::max(mx, cnt);
  }
  if (maxcnt) *maxcnt = mx;
  return flags;
}

// EOS: SphKernel.cu:25 (linear); Tait B((rho/rho0)^7 - 1) for PSB_EOS_TAIT.
// PSB_MODE_CONFIG clamps negative pressure to zero.
template <class R> inline R eos(const Derived<R>& d, const psb_sph_params& p, R rho) {
  R pr;
  if (p.eos == PSB_EOS_TAIT) {
    R x = rho * d.inv_rho0;
    R x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x7 = x6 * x;
    pr = d.tait_b * (x7 - (R)1);
  } else {
    pr = d.k * (rho - d.rho0);
  }
  if (p.mode == PSB_MODE_CONFIG && pr < (R)0) pr = (R)0;
  return pr;
}

// BoxCollision_Kernel.cu:7-34.  Clamp per axis (max side first, min side
// overrides), normalise the summed face normals, remove (1+e) of the normal
// velocity component.
template <class R> inline void box_collide(const Derived<R>& d, V4<R>& x, V4<R>& v) {
  R px = x.x, py = x.y, pz = x.z, nx = 0, ny = 0, nz = 0;
  if (x.x >= d.box_hi[0]) { px = d.box_hi[0]; nx = -1; }
  if (x.y >= d.box_hi[1]) { py = d.box_hi[1]; ny = -1; }
  if (x.z >= d.box_hi[2]) { pz = d.box_hi[2]; nz = -1; }
  if (x.x <= d.box_lo[0]) { px = d.box_lo[0]; nx = 1; }
  if (x.y <= d.box_lo[1]) { py = d.box_lo[1]; ny = 1; }
  if (x.z <= d.box_lo[2]) { pz = d.box_lo[2]; nz = 1; }
  R ln2 = dot3(nx, ny, nz, nx, ny, nz);
  if (ln2 > (R)0) {
    R ln = std::sqrt(ln2);
    nx = nx / ln; ny = ny / ln; nz = nz / ln;
    x.x = px; x.y = py; x.z = pz;
    R vn = d.one_plus_e * dot3(v.x, v.y, v.z, nx, ny, nz);
    v.x = v.x - vn * nx; v.y = v.y - vn * ny; v.z = v.z - vn * nz;
  }
}

// Colliders after the box (psb_collider).  SphereCollision_Kernel.cu:5-80 and
// PlaneCollision_Kernel.cu:6-80 restated with the documented differences:
// the obstacle sphere pushes particles out along their direction from the
// centre (the reference's branch is dead, :28-45), the plane's bounds test
// ignores the normal's axis through 1 - |N| (1 - N in the reference, which
// misses every particle of a plane facing -axis), a zero-length or
// plane-parallel step projects onto the plane (the reference divides by
// zero), and each contact normal is normalised once.
struct OColl {
  int type = 0, container = 0;
  float elast = 0, radius = 0, o[3] = {0, 0, 0}, a[3] = {0, 0, 0};
  float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
};

inline OColl make_ocoll(const psb_collider& in) {
  OColl k;
  k.type = in.type;
  k.container = in.container ? 1 : 0;
  k.elast = in.elast;
  k.radius = in.radius;
  for (int a = 0; a < 3; ++a) {
    k.o[a] = in.origin[a];
    // corner bounds as detectionPlan forms them (PlaneCollision_Kernel.cu:9-19)
    float mn = 100000.f, mx = -100000.f;
    for (int v = 0; v < 4; ++v) {
      if (mn >= in.corners[v][a]) mn = in.corners[v][a];
      if (mx <= in.corners[v][a]) mx = in.corners[v][a];
    }
    k.mn[a] = mn; k.mx[a] = mx;
    k.a[a] = in.corners[0][a];
    k.nrm[a] = in.normal[a];
  }
  return k;
}

// r = elast |x - x_int| / (dt |v|) (|v| > 0, elast > 0); v -= ((1 + r) v.n) n
template <class R> inline void collider_response(R elast, R dt, V4<R>& x, V4<R>& v, R ix, R iy,
                                                 R iz, R nx, R ny, R nz) {
  const R vl = std::sqrt(dot3(v.x, v.y, v.z, v.x, v.y, v.z));
  R r = 0;
  if (vl > (R)0 && elast > (R)0) {
    const R ex = x.x - ix, ey = x.y - iy, ez = x.z - iz;
    r = (elast * std::sqrt(dot3(ex, ey, ez, ex, ey, ez))) / (dt * vl);
  }
  const R nl = std::sqrt(dot3(nx, ny, nz, nx, ny, nz));
  nx = nx / nl; ny = ny / nl; nz = nz / nl;
  const R vn = ((R)1 + r) * dot3(v.x, v.y, v.z, nx, ny, nz);
  v.x = v.x - vn * nx; v.y = v.y - vn * ny; v.z = v.z - vn * nz;
  x.x = ix; x.y = iy; x.z = iz;
}

template <class R> inline bool sphere_hit(const OColl& k, const V4<R>& x, R& ix, R& iy, R& iz,
                                          R& nx, R& ny, R& nz) {
  const R rad = (R)k.radius;
  const R ox = x.x - (R)k.o[0], oy = x.y - (R)k.o[1], oz = x.z - (R)k.o[2];
  const R d = std::sqrt(dot3(ox, oy, oz, ox, oy, oz));
  if (k.container ? !(d >= rad) : !(d <= rad)) return false;
  if (d == rad) {
    ix = x.x; iy = x.y; iz = x.z;
    nx = -ox / d; ny = -oy / d; nz = -oz / d;
    return true;
  }
  R ux = 0, uy = 1, uz = 0;
  if (d > (R)0) { ux = ox / d; uy = oy / d; uz = oz / d; }
  ix = (R)k.o[0] + ux * rad; iy = (R)k.o[1] + uy * rad; iz = (R)k.o[2] + uz * rad;
  if (k.container) { nx = (R)k.o[0] - ix; ny = (R)k.o[1] - iy; nz = (R)k.o[2] - iz; }
  else { nx = ix - (R)k.o[0]; ny = iy - (R)k.o[1]; nz = iz - (R)k.o[2]; }
  return true;
}

template <class R> inline bool plane_hit(const OColl& k, const V4<R>& x0, const V4<R>& x, R& ix,
                                         R& iy, R& iz, R& nx, R& ny, R& nz) {
  const R N0 = (R)k.nrm[0], N1 = (R)k.nrm[1], N2 = (R)k.nrm[2];
  const R fx = (R)1 - std::fabs(N0), fy = (R)1 - std::fabs(N1), fz = (R)1 - std::fabs(N2);
  const R px = x.x * fx, py = x.y * fy, pz = x.z * fz;
  if (!(px <= (R)k.mx[0] * fx && px >= (R)k.mn[0] * fx && py <= (R)k.mx[1] * fy &&
        py >= (R)k.mn[1] * fy && pz <= (R)k.mx[2] * fz && pz >= (R)k.mn[2] * fz))
    return false;
  const R d1 = dot3(x.x - (R)k.a[0], x.y - (R)k.a[1], x.z - (R)k.a[2], N0, N1, N2);
  if (!(d1 <= (R)0)) return false;
  const R aon = dot3((R)k.a[0] - x0.x, (R)k.a[1] - x0.y, (R)k.a[2] - x0.z, N0, N1, N2);
  R dx = x.x - x0.x, dy = x.y - x0.y, dz = x.z - x0.z;
  const R dl = std::sqrt(dot3(dx, dy, dz, dx, dy, dz));
  R dn = 0;
  if (dl > (R)0) {
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    dn = dot3(dx, dy, dz, N0, N1, N2);
  }
  if (dn != (R)0) {
    const R t = aon / dn;
    ix = x0.x + t * dx; iy = x0.y + t * dy; iz = x0.z + t * dz;
  } else {
    ix = x.x - d1 * N0; iy = x.y - d1 * N1; iz = x.z - d1 * N2;
  }
  nx = N0; ny = N1; nz = N2;
  return true;
}

template <class R> inline void collide_one(const OColl& k, R dt, const V4<R>& x0, V4<R>& x,
                                           V4<R>& v) {
  R ix, iy, iz, nx, ny, nz;
  const bool hit = k.type == PSB_COLLIDER_SPHERE ? sphere_hit(k, x, ix, iy, iz, nx, ny, nz)
                                                 : plane_hit(k, x0, x, ix, iy, iz, nx, ny, nz);
  if (hit) collider_response((R)k.elast, dt, x, v, ix, iy, iz, nx, ny, nz);
}

// Pair sums of internalForces (SphKernel.cu:47-78) and of
// compute_Viscosity_SurfaceTension_Force_PCI (PciSphSystem_Kernel.cu:134-158)
// for particle p; xs[q].w = p_q / rho_q^2 (pressure term, only when PRESS),
// vs[q].w = 1 / rho_q.
template <class R> struct PairSums { R fp[3], fv[3], fs, nn[3]; };

template <class R, bool PRESS, bool SURF>
inline PairSums<R> internal_sums(const Derived<R>& d, uint32_t p, uint32_t n,
                                 const uint32_t* count, const uint32_t* list,
                                 const V4<R>* xs, const V4<R>* vs) {
  PairSums<R> s{};
  const V4<R> xi = xs[p], vi = vs[p];
  const R pvi = xi.w;
  for (uint32_t k = 0; k < count[p]; ++k) {
    uint32_t q = list[(uint64_t)k * n + p];
    const V4<R> xj = xs[q], vj = vs[q];
    R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    R d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (d2 > (R)0) {
      R dd = std::sqrt(d2);
      R hd = d.h - dd;
      R rinvj = vj.w;
      if (PRESS) {
        R invd = (R)1 / dd;
        R gw = d.c_spiky * (hd * hd);
        R c = gw * (pvi + xj.w);
        s.fp[0] = fmaR(dx * invd, c, s.fp[0]);
        s.fp[1] = fmaR(dy * invd, c, s.fp[1]);
        s.fp[2] = fmaR(dz * invd, c, s.fp[2]);
      }
      R lv = hd * rinvj;
      s.fv[0] = fmaR(vj.x - vi.x, lv, s.fv[0]);
      s.fv[1] = fmaR(vj.y - vi.y, lv, s.fv[1]);
      s.fv[2] = fmaR(vj.z - vi.z, lv, s.fv[2]);
      if (SURF) {
        R a = d.h2 - d2;
        R b = d.h2x3 - (R)7 * d2;
        s.fs = fmaR(a * b, rinvj, s.fs);
        R an = (a * a) * rinvj;
        s.nn[0] = fmaR(dx, an, s.nn[0]);
        s.nn[1] = fmaR(dy, an, s.nn[1]);
        s.nn[2] = fmaR(dz, an, s.nn[2]);
      }
    }
  }
  return s;
}

// Surface-tension finish (SphKernel.cu:80-88): n = sum * m c_grad;
// if |n| >= l: f_S = (n * (fs * -sigma)) * (1/|n|), n <- n/|n|; else 0.
template <class R>
inline void surface_finish(const Derived<R>& d, const PairSums<R>& s, R sigma, R fS[3], R nrm[3]) {
  R fsum = s.fs * d.cs;
  R n0 = s.nn[0] * d.cn, n1 = s.nn[1] * d.cn, n2 = s.nn[2] * d.cn;
  R ln = std::sqrt(dot3(n0, n1, n2, n0, n1, n2));
  if (ln >= d.l) {
    R tt = fsum * (-sigma);
    R r = (R)1 / ln;
    fS[0] = (n0 * tt) * r; fS[1] = (n1 * tt) * r; fS[2] = (n2 * tt) * r;
    nrm[0] = n0 / ln; nrm[1] = n1 / ln; nrm[2] = n2 / ln;
  } else {
    fS[0] = fS[1] = fS[2] = 0;
    nrm[0] = nrm[1] = nrm[2] = 0;
  }
}

// ---------------------------------------------------------------------------
// Pressure MLP 8 -> 64 -> 64 -> 1 (SiLU), parameters W1[64][8] b1[64]
// W2[64][64] b2[64] W3[64] b3 (W[out][in]).  Computed in double; with
// emulate_bf16 the operands of the GPU's BF16 matrix products are rounded to
// BF16 (round to nearest even) at the same points: normalised inputs, W1,
// layer-1 activations, W2, W3, and in the backward pass dZ2, dZ1 and the
// activations they multiply.  Without it, plain FP64 (for finite differences).
// ---------------------------------------------------------------------------
constexpr int kIn = 8, kHid = 64, kNP = 4801;
constexpr int oW1 = 0, oB1 = 512, oW2 = 576, oB2 = 4672, oW3 = 4736, oB3 = 4800;

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) {
    u = (u & 0xffff0000u) | ((u & 0xffffu) ? 0x400000u : 0u);
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
  }
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

struct Mlp {
  const float* w;
  float mu[kIn], inv_sd[kIn];
  bool bf;
  double q(double v) const { return bf ? (double)bf16_round((float)v) : v; }
  double W1(int j, int i) const { return q(w[oW1 + j * kIn + i]); }
  double W2(int j, int k) const { return q(w[oW2 + j * kHid + k]); }
  double W3(int k) const { return q(w[oW3 + k]); }
  static double sg(double z) { return 1.0 / (1.0 + std::exp(-z)); }
  // forward of one sample; returns y, fills activations
  double forward(const double x[kIn], double z1[kHid], double h1[kHid], double z2[kHid],
                 double h2[kHid]) const {
    for (int j = 0; j < kHid; ++j) {
      double s = 0;
      for (int i = 0; i < kIn; ++i) s += q(x[i]) * W1(j, i);
      z1[j] = s + w[oB1 + j];
      h1[j] = z1[j] * sg(z1[j]);
    }
    for (int j = 0; j < kHid; ++j) {
      double s = 0;
      for (int k = 0; k < kHid; ++k) s += q(h1[k]) * W2(j, k);
      z2[j] = s + w[oB2 + j];
      h2[j] = z2[j] * sg(z2[j]);
    }
    double y = 0;
    for (int k = 0; k < kHid; ++k) y += h2[k] * W3(k);
    return y + w[oB3];
  }
  void input(const float* feat, uint64_t stride, uint64_t s, double x[kIn]) const {
    for (int i = 0; i < kIn; ++i)
      x[i] = bf ? (double)((feat[i * stride + s] - mu[i]) * inv_sd[i])
                : ((double)feat[i * stride + s] - mu[i]) * inv_sd[i];
  }
};

Mlp make_mlp(const float* w, const float* mu, const float* sd, bool bf) {
  Mlp m;
  m.w = w;
  m.bf = bf;
  for (int i = 0; i < kIn; ++i) { m.mu[i] = mu[i]; m.inv_sd[i] = 1.0f / sd[i]; }
  return m;
}

// ---------------------------------------------------------------------------
// System: sorted state of one scene, advanced by WCSPH or PCISPH steps.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Per-particle materials (reference mode): SphSystem.cpp:157-167 and
// PciSphSystem.cpp:103-113 pack each particle's material into per-particle
// arrays; the kernels mix them per pair.  WCSPH takes the neighbour's radius
// and mass (SphKernel.cu:17-21, 51-71), PCISPH max(h_i, h_j) and the
// neighbour's mass (PciSphSystem_Kernel.cu:51, 79, 101, 138), both the mean
// viscosity (SphKernel.cu:62, PciSphSystem_Kernel.cu:144); the search keeps
// j when |x_i - x_j| <= h_i (UniformGridKernel.cu:82).  Constants per
// (material of i, material of j), computed in double and rounded once.
// ---------------------------------------------------------------------------
constexpr int kMaxMat = PSB_MAX_MATERIALS;
template <class R> struct MatOwnR { R h2, rho0, inv_rho0, k, l, sigma, beta; };
template <class R> struct MatPairR { R h, h2, h2x3, mc, cs, csm, cvm, csf, mu; };
template <class R> struct MatTabR {
  MatOwnR<R> own[kMaxMat];
  MatPairR<R> pair[kMaxMat * kMaxMat];
};

// 0 on success, 1 on an invalid material.
template <class R>
int make_mat_tab(const psb_sph_params& sys, const psb_sph_params* mats, uint32_t count,
                 MatTabR<R>& t) {
  if (!mats || count < 1 || count > (uint32_t)kMaxMat) return 1;
  const double pi = 3.14159265358979323846, dt = sys.dt;
  for (uint32_t a = 0; a < count; ++a) {
    const psb_sph_params& q = mats[a];
    if (!(q.support_radius > 0.f && q.support_radius <= sys.support_radius)) return 1;
    if (!(q.mass > 0.f && q.rest_density > 0.f)) return 1;
    const double h = q.support_radius, m = q.mass, rho0 = q.rest_density;
    MatOwnR<R>& o = t.own[a];
    o.h2 = (R)(h * h);
    o.rho0 = (R)rho0;
    o.inv_rho0 = (R)(1.0 / rho0);
    o.k = (R)q.gas_stiffness;
    o.l = (R)q.threshold;
    o.sigma = (R)q.surface_tension;
    o.beta = (R)(2.0 * std::pow(m * dt / rho0, 2));
  }
  for (uint32_t a = 0; a < count; ++a)
    for (uint32_t b = 0; b < count; ++b) {
      const psb_sph_params &qa = mats[a], &qb = mats[b];
      const double hk = sys.solver == PSB_SOLVER_PCISPH
                            ? (double)std::max(qa.support_radius, qb.support_radius)
                            : (double)qb.support_radius;
      const double mb = qb.mass;
      MatPairR<R>& P = t.pair[a * kMaxMat + b];
      P.h = (R)hk;
      P.h2 = (R)(hk * hk);
      P.h2x3 = (R)(3.0 * hk * hk);
      P.mc = (R)(mb * 315.0 / (64.0 * pi * std::pow(hk, 9)));
      P.cs = (R)(-45.0 / (pi * std::pow(hk, 6)));
      P.csm = (R)(mb * (-45.0 / (pi * std::pow(hk, 6))));
      P.cvm = (R)(mb * 45.0 / (pi * std::pow(hk, 6)));
      P.csf = (R)(mb * (-945.0 / (32.0 * pi * std::pow(hk, 9))));
      P.mu = (R)(((double)qa.viscosity + (double)qb.viscosity) / 2.0);
    }
  return 0;
}

template <class R> struct System {
  psb_sph_params prm;
  Derived<R> d;
  Grid g;
  uint32_t cap = 0, n = 0;
  // A = start-of-step (sorted), B = end-of-step (same order)
  std::vector<V4<R>> Ax, Av, Avh, Bx, Bv, Bvh;
  std::vector<uint32_t> Aid, Bid, hash, perm, start, count, list;
  std::vector<R> Ap, Bp, rho, pres, rho_pred, dl;
  bool feat_on = false;
  std::vector<R> feat;
  std::vector<float> mlp_w;
  float mlp_mu[kIn] = {0}, mlp_sd[kIn] = {1, 1, 1, 1, 1, 1, 1, 1};
  std::vector<V4<R>> fa, fs4, anp, xp0, xp1;
  uint32_t flags = 0;
  int last_iters = 0;
  R last_err = 0;
  int64_t total_iters = 0;
  R worst_err = 0;
  int retried_steps = 0, substeps = 0;   // retries of the last update call
  int steps_done = 0;
  uint32_t max_nbr_seen = 0;
  std::vector<int> iter_hist;
  std::vector<R> err_hist;
  // rigid DEM spheres: boundary particles with ids >= n_fluid
  static constexpr uint32_t kNone = 0xFFFFFFFFu;
  uint32_t n_fluid = kNone;
  uint32_t n_spheres = 0, per_sphere = 0;
  psb_sphere_params sphere{};
  std::vector<float> sph_off;                       // [per][4]
  std::vector<V4<R>> sph_c, sph_v, sph_f, fb;       // centres, velocities, fluid forces
  std::vector<OColl> colls;                         // colliders after the box

  bool bnd(uint32_t p) const { return n_fluid != kNone && Aid[p] >= n_fluid; }

  void init(const psb_sph_params& p, uint32_t capacity) {
    prm = p; d = derive<R>(p); g = grid_from(p.grid); cap = capacity;
    auto rs = [&](auto& v, size_t k) { v.assign(k, {}); };
    rs(Ax, cap); rs(Av, cap); rs(Avh, cap); rs(Bx, cap); rs(Bv, cap); rs(Bvh, cap);
    rs(fa, cap); rs(fs4, cap); rs(anp, cap); rs(xp0, cap); rs(xp1, cap);
    Aid.assign(cap, 0); Bid.assign(cap, 0); hash.assign(cap, 0); perm.assign(cap, 0);
    start.assign(g.ncells() + 1, 0); count.assign(cap, 0);
    list.assign((size_t)cap * p.max_neighbors, 0);
    Ap.assign(cap, 0); Bp.assign(cap, 0); rho.assign(cap, 0); pres.assign(cap, 0);
    rho_pred.assign(cap, 0);
    dl.assign(cap, 0);
    feat.assign(8 * (size_t)cap, 0);
  }

  // storeParticles + reorder (B -> A by the stable order).
  void build_grid() {
    std::vector<float> x4(4 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
      x4[4 * i] = (float)Bx[i].x; x4[4 * i + 1] = (float)Bx[i].y; x4[4 * i + 2] = (float)Bx[i].z;
      x4[4 * i + 3] = 0.f;
    }
    grid_build(g, n, x4.data(), hash.data(), perm.data(), start.data(), flags);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t k = 0; k < (int64_t)n; ++k) {
      uint32_t i = perm[k];
      Ax[k] = Bx[i]; Av[k] = Bv[i]; Avh[k] = Bvh[i]; Aid[k] = Bid[i]; Ap[k] = Bp[i];
    }
  }

  void neighbors_density() {
    std::vector<R> S(n);
    uint32_t mx = 0;
    flags |= neighbors<R>(g, start.data(), n, Ax.data(), d.h2, d.R_stencil,
                          (uint32_t)prm.max_neighbors, count.data(), list.data(), S.data(), &mx);
    max_nbr_seen = std::max(max_nbr_seen, mx);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t k = 0; k < (int64_t)n; ++k) rho[k] = S[k] * d.mc;
  }

  void finish_particle(uint32_t p, V4<R> x, V4<R> vh) {
    // collide (SphSystem.cpp:323-328) then interpolate (SphKernel.cu:128-135)
    if (prm.box_enabled) box_collide(d, x, vh);
    for (const OColl& k : colls) collide_one(k, (R)d.dt, Ax[p], x, vh);
    Bx[p] = {x.x, x.y, x.z, 0};
    Bvh[p] = {vh.x, vh.y, vh.z, 0};
    const V4<R> va = Avh[p];
    Bv[p] = {(va.x + vh.x) * (R)0.5, (va.y + vh.y) * (R)0.5, (va.z + vh.z) * (R)0.5, 0};
    Bid[p] = Aid[p];
  }

  // SPHSystem::update (SphSystem.cpp:273-297): grid, densities+forces
  // (SphKernel.cu:9-91), gravity as g*rho (SphSystem.cpp:312,
  // ForceExt_Constante.cu:15-24), leapfrog (SphKernel.cu:112-124), box,
  // interpolation.
  void wcsph_step() {
    build_grid();
    neighbors_density();
    const bool surf = prm.surface_tension != 0.f;
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t k = 0; k < (int64_t)n; ++k) {
      R r = rho[k];
      R pr = eos(d, prm, r);
      pres[k] = pr;
      Ax[k].w = pr / (r * r);
      Av[k].w = (R)1 / r;
    }
    uint32_t zf = 0;
#pragma omp parallel for schedule(static) reduction(| : zf) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      PairSums<R> s = surf ? internal_sums<R, true, true>(d, p, n, count.data(), list.data(), Ax.data(), Av.data())
                           : internal_sums<R, true, false>(d, p, n, count.data(), list.data(), Ax.data(), Av.data());
      const R ri = rho[p];
      if (!(ri > (R)0)) zf |= PSB_FLAG_ZERO_DENSITY;
      R cpf = -(d.m * ri);
      R fP[3] = {s.fp[0] * cpf, s.fp[1] * cpf, s.fp[2] * cpf};
      R fV[3] = {s.fv[0] * d.cv, s.fv[1] * d.cv, s.fv[2] * d.cv};
      R fS[3] = {0, 0, 0}, nrm[3];
      if (surf) surface_finish(d, s, d.sigma, fS, nrm);
      R F[3];
      F[0] = (fV[0] + fS[0]) + fP[0];
      F[1] = (fV[1] + fS[1]) + fP[1];
      F[2] = (fV[2] + fS[2]) + fP[2];
      F[0] = F[0] + d.gx * ri; F[1] = F[1] + d.gy * ri; F[2] = F[2] + d.gz * ri;
      R inv = (R)1 / ri;
      V4<R> vh = Avh[p], x = Ax[p];
      vh.x = vh.x + (F[0] * inv) * d.dt;
      vh.y = vh.y + (F[1] * inv) * d.dt;
      vh.z = vh.z + (F[2] * inv) * d.dt;
      x.x = x.x + vh.x * d.dt; x.y = x.y + vh.y * d.dt; x.z = x.z + vh.z * d.dt;
      finish_particle(p, x, vh);
      Bp[p] = pres[p];
    }
    flags |= zf;
    last_iters = 0; last_err = 0;
  }

  // PCI_SPHSystem::update (PciSphSystem.cpp:163-204) -> pci_SPH_update
  // (PciSphSystem.cu:8-68).
  void pcisph_step() { pcisph_advance(0, false); }

  // Constants of a retry substep at level L: the step size dt / 2^L and
  // everything derived from it (beta, the template delta, feature scales),
  // with delta also relaxed by 2^-L (on top of params.relaxation).
  Derived<R> level_derived(int level) const {
    psb_sph_params p = prm;
    const float scale = std::ldexp(1.0f, -level);
    p.dt = prm.dt * scale;
    if (p.delta > 0.f) p.delta = prm.delta / (scale * scale);
    p.relaxation = (prm.relaxation > 0.f ? prm.relaxation : 1.0f) * scale;
    return derive<R>(p);
  }

  // One step of dt / 2^level.  A config-mode step whose corrector ends above
  // the tolerance is redone from its start state A as two steps of half the
  // size, down to params.retry_levels, each restarting from p = 0 without
  // warm start; the first half reuses A's grid, lists and densities.
  void pcisph_advance(int level, bool from_a) {
    const Derived<R> saved = d;
    if (level > 0) d = level_derived(level);
    const bool ref = prm.mode == PSB_MODE_REFERENCE;
    const bool retry_ok = !ref && level < prm.retry_levels;
    const bool plain = level > 0;   // substeps run the plain corrector
    pci_prologue(from_a, plain);
    const bool warm = !ref && !plain && prm.warm_start == PSB_WARM_PREVIOUS;
    const bool mlp = !ref && !plain && prm.warm_start == PSB_WARM_MLP;
    const bool feat_now = !ref && !plain && (mlp || feat_on);
    std::vector<V4<R>>* cur = &xp0;
    std::vector<V4<R>>* nxt = &xp1;
    if (warm) {
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
      for (int64_t k = 0; k < (int64_t)n; ++k) (*cur)[k].w = pres[k];
      pci_force(*cur, *nxt, false);
      std::swap(cur, nxt);
    }
    const int min_it = ref ? 3 : std::max(1, prm.min_iterations);
    int max_it = ref ? 3 : std::max(min_it, prm.max_iterations);
    // the deepest retry level may iterate longer: halving dt does not help a
    // particle whose density barely responds to pressure (a near-coincident
    // pair), which only more iterations separate
    if (!ref && level > 0 && level == prm.retry_levels)
      max_it = std::min(4 * prm.max_iterations, 255);
    int it = 0;
    R err = 0;
    // smallest error of the earlier iterations: a retry substep's corrector
    // that grows past twice it and past 4x the tolerance (after min_it) stops,
    // failed
    R min_prev = std::numeric_limits<R>::infinity();
    while (true) {
      err = pci_density(*cur, it == 0 && feat_now, it == 0 && mlp);
      if (it == 0 && mlp) mlp_pressure(*cur);
      it++;
      const bool diverged = !ref && level > 0 && it >= min_it &&
                            err > (R)kDivergeFloor * (R)prm.tolerance &&
                            err > (R)kDivergeFactor * min_prev;
      bool fin = (it >= min_it && (ref || err <= (R)prm.tolerance)) || it >= max_it || diverged;
      min_prev = std::min(min_prev, err);
      pci_force(*cur, *nxt, fin);
      if (fin) break;
      std::swap(cur, nxt);
    }
    const bool failed = !ref && err > (R)prm.tolerance;
    last_iters = it; last_err = err;
    total_iters += it;
    if (failed && retry_ok) {
      if (level == 0) retried_steps++;
      substeps += 2;
      d = saved;
      pcisph_advance(level + 1, true);
      pcisph_advance(level + 1, false);
      return;
    }
    if (failed) flags |= PSB_FLAG_NOT_CONVERGED;
    worst_err = std::max(worst_err, err);
    iter_hist.push_back(it); err_hist.push_back(err);
    if (!ref && n_spheres) bodies_step(*cur);
    d = saved;
  }

  // The PCISPH step up to its corrector loop: grid build, neighbours and
  // densities, initial pressure, non-pressure forces and the first
  // prediction (xp0).  The corrector probe of the C API stops here.
  // from_a: a retry from the step's start state A, whose grid, lists and
  // densities (walls included) are kept; the pressure starts at 0.  plain:
  // a retry substep, without warm start or features.
  void pci_prologue(bool from_a = false, bool plain = false) {
    if (!from_a) {
      build_grid();
      neighbors_density();
    }
    const bool ref = prm.mode == PSB_MODE_REFERENCE;
    const R sigma = ref ? (R)0.0728f : d.sigma;
    const bool surf = ref ? true : (prm.surface_tension != 0.f);
    const R gx = ref ? (R)0 : d.gx, gz = ref ? (R)0 : d.gz;
    const R gy = ref ? (R)-9.81f : d.gy;
    const bool warm = !ref && !plain && prm.warm_start == PSB_WARM_PREVIOUS;
    const bool mlp = !ref && !plain && prm.warm_start == PSB_WARM_MLP;
    const bool feat_now = !ref && !plain && (mlp || feat_on);
    // initial pressure (PciSphSystem_Kernel.cu:131 resets it to zero)
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t k = 0; k < (int64_t)n; ++k) {
      if (!ref && bnd((uint32_t)k)) {      // sphere boundary particle: p = 0, zero volume
        pres[k] = 0;
        if (feat_now) Ax[k].w = 0;
        Av[k].w = 0;
        continue;
      }
      pres[k] = warm ? Ap[k] : (R)0;
      if (feat_now) Ax[k].w = Ap[k];
      if (d.wall && !from_a) {
        R Sw, gx_, gy_, gz_;
        wall_sums(d, true, false, Ax[k].x, Ax[k].y, Ax[k].z, Sw, gx_, gy_, gz_);
        rho[k] = rho[k] + Sw * d.mc;
      }
      Av[k].w = (R)1 / rho[k];
    }
    // non-pressure forces + first prediction (PciSphSystem_Kernel.cu:117-170,
    // SphKernel.cu:95-109)
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      if (!ref && bnd(p)) {               // static during the corrector
        anp[p] = {0, 0, 0, 0};
        dl[p] = 0;
        xp0[p] = {Ax[p].x, Ax[p].y, Ax[p].z, 0};
        if (feat_now)
          for (int k = 1; k < 8; ++k) feat[(size_t)k * n + p] = 0;
        continue;
      }
      PairSums<R> s = surf ? internal_sums<R, false, true>(d, p, n, count.data(), list.data(), Ax.data(), Av.data())
                           : internal_sums<R, false, false>(d, p, n, count.data(), list.data(), Ax.data(), Av.data());
      const R ri = rho[p];
      R fV[3] = {s.fv[0] * d.cv, s.fv[1] * d.cv, s.fv[2] * d.cv};
      R fS[3] = {0, 0, 0}, nrm[3];
      if (surf) surface_finish(d, s, sigma, fS, nrm);
      const V4<R> x = Ax[p], vh0 = Avh[p];
      if (!ref) dl[p] = feat_now ? delta_and_features(p) : local_delta(p);
      V4<R> xp;
      if (ref) {
        // forcesAccum = (0, -9.81 rho, 0); F = ((Facc + fV) + fP) + fS
        R fa0 = (R)0 + fV[0], fa1 = (gy * ri) + fV[1], fa2 = (R)0 + fV[2];
        fa[p] = {fa0, fa1, fa2, 0};
        fs4[p] = {fS[0], fS[1], fS[2], 0};
        R F0 = (fa0 + (R)0) + fS[0], F1 = (fa1 + (R)0) + fS[1], F2 = (fa2 + (R)0) + fS[2];
        R inv = (R)1 / ri;
        R v0 = vh0.x + (F0 * inv) * d.dt, v1 = vh0.y + (F1 * inv) * d.dt, v2 = vh0.z + (F2 * inv) * d.dt;
        xp = {x.x + v0 * d.dt, x.y + v1 * d.dt, x.z + v2 * d.dt, 0};
      } else {
        R inv = (R)1 / ri;
        R a0 = (fV[0] + fS[0]) * inv + gx, a1 = (fV[1] + fS[1]) * inv + gy, a2 = (fV[2] + fS[2]) * inv + gz;
        anp[p] = {a0, a1, a2, 0};
        if (warm) {
          xp = {x.x, x.y, x.z, 0};
        } else {
          R v0 = vh0.x + a0 * d.dt, v1 = vh0.y + a1 * d.dt, v2 = vh0.z + a2 * d.dt;
          xp = {x.x + v0 * d.dt, x.y + v1 * d.dt, x.z + v2 * d.dt, 0};
        }
      }
      xp0[p] = xp;
    }
  }

  // Rigid spheres after the corrector, X = the final iteration's predicted
  // positions (w = final pressure).  (1) the opposite of the pressure force
  // each boundary particle exerted on its fluid neighbours; (2) per sphere:
  // fluid force in boundary-id order, gravity, config-law DEM contacts with
  // every other sphere in index order and the six walls (signed distance,
  // inward normal), symplectic Euler; (3) boundary particles follow.
  void bodies_step(const std::vector<V4<R>>& X) {
    const uint32_t nb = n_spheres * per_sphere;
    std::vector<V4<R>> fbb(nb);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      if (!bnd(p)) continue;
      const V4<R> xb = X[p];
      R sx = 0, sy = 0, sz = 0;
      for (uint32_t k = 0; k < count[p]; ++k) {
        const V4<R> xq = X[list[(uint64_t)k * n + p]];
        R dx = xq.x - xb.x, dy = xq.y - xb.y, dz = xq.z - xb.z;
        R d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (d2 > (R)0 && d2 < d.h2) {
          R dd = std::sqrt(d2);
          R invd = (R)1 / dd;
          R hd = d.h - dd;
          R gw = d.c_spiky * (hd * hd);
          R cc = (gw * invd) * xq.w;
          sx = fmaR(dx, cc, sx); sy = fmaR(dy, cc, sy); sz = fmaR(dz, cc, sz);
        }
      }
      const R sc = -(d.m * d.kp);
      fbb[Aid[p] - n_fluid] = {sx * sc, sy * sc, sz * sc, 0};
    }
    const R Rr = (R)sphere.radius;
    const R mass = (R)(float)(sphere.density * 4.0 / 3.0 * M_PI * std::pow((double)sphere.radius, 3));
    const R k = (R)sphere.spring, c = (R)sphere.damping;
    std::vector<V4<R>> nc(n_spheres), nv(n_spheres);
    for (uint32_t j = 0; j < n_spheres; ++j) {
      R fx = 0, fy = 0, fz = 0;
      for (uint32_t l = 0; l < per_sphere; ++l) {
        const V4<R>& f = fbb[(size_t)j * per_sphere + l];
        fx = fx + f.x; fy = fy + f.y; fz = fz + f.z;
      }
      sph_f[j] = {fx, fy, fz, 0};
      fx = fx + mass * d.gx; fy = fy + mass * d.gy; fz = fz + mass * d.gz;
      const V4<R> cj = sph_c[j], vj = sph_v[j];
      for (uint32_t i = 0; i < n_spheres; ++i) {
        if (i == j) continue;
        // config-law DEM pair, the same expressions as dem_pair
        R dx = sph_c[i].x - cj.x, dy = sph_c[i].y - cj.y, dz = sph_c[i].z - cj.z;
        R dist = std::sqrt(dot3(dx, dy, dz, dx, dy, dz));
        if (!(dist > (R)0)) continue;
        R RR = Rr + Rr;
        if (!(dist < RR)) continue;
        R inv = (R)1 / dist;
        R nx = dx * inv, ny = dy * inv, nz = dz * inv;
        R sdem = -k * (RR - dist);
        R vn = dot3(sph_v[i].x - vj.x, sph_v[i].y - vj.y, sph_v[i].z - vj.z, nx, ny, nz);
        sdem = sdem + c * vn;
        fx = fx + sdem * nx; fy = fy + sdem * ny; fz = fz + sdem * nz;
      }
      const R pc[3] = {cj.x, cj.y, cj.z}, pv[3] = {vj.x, vj.y, vj.z};
      R fw[3] = {0, 0, 0};
      for (int ax = 0; ax < 3; ++ax)
        for (int side = 0; side < 2; ++side) {
          const R dd = side == 0 ? pc[ax] - (R)prm.box_min[ax] : (R)prm.box_max[ax] - pc[ax];
          if (dd < Rr) {
            const R nin = side == 0 ? (R)1 : (R)-1;
            const R sw = k * (Rr - dd) - c * (pv[ax] * nin);
            fw[ax] = fw[ax] + sw * nin;
          }
        }
      fx = fx + fw[0]; fy = fy + fw[1]; fz = fz + fw[2];
      const R inv_m = (R)1 / mass;
      const R vx = vj.x + (fx * inv_m) * d.dt, vy = vj.y + (fy * inv_m) * d.dt,
              vz = vj.z + (fz * inv_m) * d.dt;
      nv[j] = {vx, vy, vz, 0};
      nc[j] = {cj.x + vx * d.dt, cj.y + vy * d.dt, cj.z + vz * d.dt, 0};
    }
    sph_c = nc;
    sph_v = nv;
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      if (!bnd(p)) continue;
      const uint32_t kb = Aid[p] - n_fluid;
      const uint32_t j = kb / per_sphere, l = kb - j * per_sphere;
      Bx[p] = {sph_c[j].x + (R)sph_off[4 * l], sph_c[j].y + (R)sph_off[4 * l + 1],
               sph_c[j].z + (R)sph_off[4 * l + 2], 0};
      Bv[p] = sph_v[j];
      Bvh[p] = sph_v[j];
      Bid[p] = Aid[p];
      Bp[p] = 0;
    }
  }

  // Pressure scaling of particle p (config mode): the template delta, capped
  // by the per-particle value 1/(beta (W1.W1 + W2)) of compute_Pressure_PCI
  // (PciSphSystem_Kernel.cu:44-63) evaluated once at the start positions.
  R local_delta(uint32_t p) const {
    if (!prm.delta_local_cap) return d.delta;
    const V4<R> xi = Ax[p];
    R W1x = 0, W1y = 0, W1z = 0, W2 = 0;
    for (uint32_t k = 0; k < count[p]; ++k) {
      uint32_t q = list[(uint64_t)k * n + p];
      const V4<R> xj = Ax[q];
      R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      R d2 = dot3(dx, dy, dz, dx, dy, dz);
      if (d2 > (R)0 && d2 < d.h2) {
        R dd = std::sqrt(d2);
        R invd = (R)1 / dd;
        R hd = d.h - dd;
        R gw = d.c_spiky * (hd * hd);
        R ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
        W1x = W1x + ux; W1y = W1y + uy; W1z = W1z + uz;
        // sphere boundary particles (v.w = 0) do not move: W1 only
        if (n_fluid == kNone || Av[q].w > (R)0) W2 = W2 + dot3(ux, uy, uz, ux, uy, uz);
      }
    }
    if (d.wall) {
      R Sw, wgx, wgy, wgz;
      wall_sums(d, false, true, xi.x, xi.y, xi.z, Sw, wgx, wgy, wgz);
      W1x = W1x + wgx; W1y = W1y + wgy; W1z = W1z + wgz;
    }
    R den = d.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2);
    if (!(den > (R)0)) return d.delta;
    R loc = (R)1 / den;
    return loc < d.delta ? loc : d.delta;
  }

  // compute_Density_PCI + compute_Pressure_PCI (PciSphSystem_Kernel.cu:36-85)
  // in reference mode; template-delta pressure update in config mode.
  // Returns max_i max(0, rho_i - rho0)/rho0.
  // Warm-start features 1..7 and the pressure scaling of particle p (config
  // mode), from its start-of-step neighbours; Ax[q].w = previous pressure.
  R delta_and_features(uint32_t p) {
    const V4<R> xi = Ax[p], vi = Av[p];
    R W1x = 0, W1y = 0, W1z = 0, W2 = 0, div = 0, gx = 0, gy = 0, gz = 0;
    for (uint32_t k = 0; k < count[p]; ++k) {
      uint32_t q = list[(uint64_t)k * n + p];
      const V4<R> xj = Ax[q], vj = Av[q];
      R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      R d2 = dot3(dx, dy, dz, dx, dy, dz);
      if (d2 > (R)0 && d2 < d.h2) {
        R dd = std::sqrt(d2);
        R invd = (R)1 / dd;
        R hd = d.h - dd;
        R gw = d.c_spiky * (hd * hd);
        R ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
        W1x = W1x + ux; W1y = W1y + uy; W1z = W1z + uz;
        if (vj.w > (R)0) W2 = W2 + dot3(ux, uy, uz, ux, uy, uz);   // boundary: W1 only
        R vol = d.m * vj.w;
        R t = dot3(vj.x - vi.x, vj.y - vi.y, vj.z - vi.z, ux, uy, uz);
        div = fmaR(t, vol, div);
        R dp = (xj.w - xi.w) * vol;
        gx = fmaR(ux, dp, gx); gy = fmaR(uy, dp, gy); gz = fmaR(uz, dp, gz);
      }
    }
    feat[1 * (size_t)n + p] = div * d.dt;
    feat[2 * (size_t)n + p] = (R)(float)count[p] * (R)0.03125f;
    feat[3 * (size_t)n + p] = xi.w * d.inv_ps;
    feat[4 * (size_t)n + p] = std::sqrt(dot3(vi.x, vi.y, vi.z, vi.x, vi.y, vi.z)) * d.dt_over_h;
    feat[5 * (size_t)n + p] = gx * d.h_over_ps;
    feat[6 * (size_t)n + p] = gy * d.h_over_ps;
    feat[7 * (size_t)n + p] = gz * d.h_over_ps;
    if (!prm.delta_local_cap) return d.delta;
    if (d.wall) {
      R Sw, wgx, wgy, wgz;
      wall_sums(d, false, true, xi.x, xi.y, xi.z, Sw, wgx, wgy, wgz);
      W1x = W1x + wgx; W1y = W1y + wgy; W1z = W1z + wgz;
    }
    R den = d.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2);
    if (!(den > (R)0)) return d.delta;
    R loc = (R)1 / den;
    return loc < d.delta ? loc : d.delta;
  }

  // MLP warm start: p = max(0, pressure_scale * net(features)), also into X.w
  void mlp_pressure(std::vector<V4<R>>& X) {
    std::vector<float> f(8 * (size_t)n);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)feat[i];
    Mlp m = make_mlp(mlp_w.data(), mlp_mu, mlp_sd, true);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      double x[kIn], z1[kHid], h1[kHid], z2[kHid], h2[kHid];
      m.input(f.data(), n, (uint64_t)kk, x);
      float y = (float)m.forward(x, z1, h1, z2, h2);
      float v = y * (float)d.pscale;
      if (v < 0.f) v = 0.f;
      if (bnd((uint32_t)kk)) v = 0.f;
      pres[kk] = (R)v;
      X[kk].w = (R)v;
    }
  }

  R pci_density(std::vector<V4<R>>& X, bool write_feature = false, bool skip_update = false) {
    const bool ref = prm.mode == PSB_MODE_REFERENCE;
    R emax = 0;
#pragma omp parallel for schedule(static) reduction(max : emax) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      if (!ref && bnd(p)) {               // sphere boundary particle: no pressure, no error
        X[p].w = 0;
        pres[p] = 0;
        rho_pred[p] = d.rho0;
        if (write_feature) feat[p] = 0;
        continue;
      }
      const V4<R> xi = X[p];
      // two partial sums, over even and over odd list positions, added as
      // even + odd (the canonical order of the corrector passes)
      R S2[2] = {0, 0}, W1x2[2] = {0, 0}, W1y2[2] = {0, 0}, W1z2[2] = {0, 0}, W22[2] = {0, 0};
      for (uint32_t k = 0; k < count[p]; ++k) {
        const int a = (int)(k & 1u);
        uint32_t q = list[(uint64_t)k * n + p];
        const V4<R> xj = X[q];
        R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        R d2 = dot3(dx, dy, dz, dx, dy, dz);
        R w = d.h2 - d2;
        if (!ref) w = std::max(w, (R)0);
        S2[a] = fmaR(w * w, w, S2[a]);
        if (ref && d2 > (R)0) {
          R dd = std::sqrt(d2);
          R invd = (R)1 / dd;
          R hd = d.h - dd;
          R gw = d.c_spiky * (hd * hd);
          R ux = (dx * invd) * gw, uy = (dy * invd) * gw, uz = (dz * invd) * gw;
          W1x2[a] = W1x2[a] + ux; W1y2[a] = W1y2[a] + uy; W1z2[a] = W1z2[a] + uz;
          W22[a] = W22[a] + dot3(ux, uy, uz, ux, uy, uz);
        }
      }
      const R S = S2[0] + S2[1];
      const R W1x = W1x2[0] + W1x2[1], W1y = W1y2[0] + W1y2[1], W1z = W1z2[0] + W1z2[1];
      const R W2 = W22[0] + W22[1];
      R r = S * d.mc;
      if (!ref && d.wall) {
        R Sw, gx_, gy_, gz_;
        wall_sums(d, true, false, xi.x, xi.y, xi.z, Sw, gx_, gy_, gz_);
        r = r + Sw * d.mc;
      }
      R pr = pres[p];
      R e = r - d.rho0;
      if (ref) {
        if (r > d.rho0 && W2 > (R)0)
          pr = pr + e / (d.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2));
        X[p].w = pr / (r * r);
        rho_pred[p] = r;
      } else {
        if (write_feature) feat[p] = (e * d.inv_rho0) * (R)100;
        if (!skip_update) {
          pr = pr + dl[p] * e;
          if (pr < (R)0) pr = (R)0;
        }
        X[p].w = pr;
        rho_pred[p] = r;
      }
      pres[p] = pr;
      R err = (e > (R)0 ? e : (R)0) * d.inv_rho0;
      emax = std::max(emax, err);
    }
    return emax;
  }

  // compute_Pressure_Force_PCI (PciSphSystem_Kernel.cu:88-114) +
  // integrateSPH_LeapFrog_Forces (SphKernel.cu:95-109) from the restored
  // start state (restore_Pos_Vel_PCI, :190-202).
  void pci_force(std::vector<V4<R>>& X, std::vector<V4<R>>& Xn, bool fin) {
    const bool ref = prm.mode == PSB_MODE_REFERENCE;
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const V4<R> xi = X[p];
      if (!ref && bnd(p)) {               // static; the sphere update writes its end state
        if (!fin) Xn[p] = {xi.x, xi.y, xi.z, 0};
        continue;
      }
      R sx2[2] = {0, 0}, sy2[2] = {0, 0}, sz2[2] = {0, 0};   // even / odd list positions
      for (uint32_t k = 0; k < count[p]; ++k) {
        const int a = (int)(k & 1u);
        uint32_t q = list[(uint64_t)k * n + p];
        const V4<R> xj = X[q];
        R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        R d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (ref) {
          if (d2 > (R)0) {
            R dd = std::sqrt(d2);
            R invd = (R)1 / dd;
            R hd = d.h - dd;
            R gw = d.c_spiky * (hd * hd);
            R c = gw * (xi.w + xj.w);
            sx2[a] = fmaR(dx * invd, c, sx2[a]); sy2[a] = fmaR(dy * invd, c, sy2[a]);
            sz2[a] = fmaR(dz * invd, c, sz2[a]);
          }
        } else {
          // (h - d)^2 / d (p_i + p_j), r = 1/d, d = d2 r; zero past h through
          // the clamp, zero for the self entry (dx = 0); c_spiky kp applied to
          // the sum
          const R d2c = std::max(d2, (R)kMinPairD2);
          const R r = rsqrt_pair(d2c);
          const R hd = std::max(fmaR(-d2c, r, d.h), (R)0);
          const R c = ((hd * hd) * r) * (xi.w + xj.w);
          sx2[a] = fmaR(dx, c, sx2[a]); sy2[a] = fmaR(dy, c, sy2[a]); sz2[a] = fmaR(dz, c, sz2[a]);
        }
      }
      R sx = sx2[0] + sx2[1], sy = sy2[0] + sy2[1], sz = sz2[0] + sz2[1];
      const V4<R> x0 = Ax[p], vh0 = Avh[p];
      V4<R> vh, x;
      if (ref) {
        const R rp = rho_pred[p];
        R cpf = -(d.m * rp);
        R fP0 = sx * cpf, fP1 = sy * cpf, fP2 = sz * cpf;
        const V4<R> a = fa[p], s = fs4[p];
        R F0 = (a.x + fP0) + s.x, F1 = (a.y + fP1) + s.y, F2 = (a.z + fP2) + s.z;
        R inv = (R)1 / rp;
        vh = {vh0.x + (F0 * inv) * d.dt, vh0.y + (F1 * inv) * d.dt, vh0.z + (F2 * inv) * d.dt, 0};
      } else {
        const V4<R> a = anp[p];
        R fP0 = sx * d.kps, fP1 = sy * d.kps, fP2 = sz * d.kps;
        if (d.wall) {   // wall particles push with p_i only (p_b = 0)
          R Sw, gx_, gy_, gz_;
          wall_sums(d, false, true, xi.x, xi.y, xi.z, Sw, gx_, gy_, gz_);
          const R kpw = d.kp * xi.w;
          fP0 = fmaR(gx_, kpw, fP0); fP1 = fmaR(gy_, kpw, fP1); fP2 = fmaR(gz_, kpw, fP2);
        }
        R a0 = a.x + fP0, a1 = a.y + fP1, a2 = a.z + fP2;
        vh = {vh0.x + a0 * d.dt, vh0.y + a1 * d.dt, vh0.z + a2 * d.dt, 0};
      }
      x = {x0.x + vh.x * d.dt, x0.y + vh.y * d.dt, x0.z + vh.z * d.dt, 0};
      if (!fin) {
        Xn[p].x = x.x; Xn[p].y = x.y; Xn[p].z = x.z;
      } else {
        finish_particle(p, x, vh);
        Bp[p] = pres[p];
      }
    }
  }

  // ---- per-particle materials ----------------------------------------------
  uint32_t nmat = 0;               // 0: the system's one material
  MatTabR<R> mt{};
  std::vector<uint8_t> mat_by_id, ms;

  const MatPairR<R>& mpair(uint32_t a, uint32_t b) const { return mt.pair[a * kMaxMat + b]; }

  // Lists over the system radius, then each particle's cut to its own
  // material's radius at the start positions (order kept).
  void mixed_lists() {
    build_grid();
    neighbors_density();
    ms.assign(n, 0);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const uint32_t a = mat_by_id[Aid[p]];
      ms[p] = (uint8_t)a;
      const V4<R> xi = Ax[p];
      uint32_t kept = 0;
      for (uint32_t k = 0; k < count[p]; ++k) {
        const uint32_t j = list[(uint64_t)k * n + p];
        const V4<R> xj = Ax[j];
        R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        if (dot3(dx, dy, dz, dx, dy, dz) <= mt.own[a].h2) list[(uint64_t)(kept++) * n + p] = j;
      }
      count[p] = kept;
    }
  }

  // lanes = 4: the sums split as in mix_total (the PCISPH passes); 1: one
  // sum in list order (WCSPH)
  R mixed_density(uint32_t p, const std::vector<V4<R>>& X, uint32_t lanes = 1) const {
    const V4<R> xi = X[p];
    R S4[4] = {};
    for (uint32_t k = 0; k < count[p]; ++k) {
      const uint32_t a = k % lanes;
      const uint32_t j = list[(uint64_t)k * n + p];
      const MatPairR<R>& P = mpair(ms[p], ms[j]);
      const V4<R> xj = X[j];
      R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      R w = P.h2 - dot3(dx, dy, dz, dx, dy, dz);
      S4[a] = fmaR((w * w) * w, P.mc, S4[a]);
    }
    return lanes == 1 ? S4[0] : mix_total(S4);
  }

  // internalForces / compute_Viscosity_SurfaceTension_Force_PCI pair sums on
  // the start positions (Ax.w = p / rho^2 when press, Av.w = 1 / rho).
  PairSums<R> mixed_sums(uint32_t p, bool press, uint32_t lanes = 1) const {
    PairSums<R> q[4] = {};
    const V4<R> xi = Ax[p], vi = Av[p];
    for (uint32_t k = 0; k < count[p]; ++k) {
      PairSums<R>& s = q[k % lanes];
      const uint32_t j = list[(uint64_t)k * n + p];
      const V4<R> xj = Ax[j], vj = Av[j];
      R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      R d2 = dot3(dx, dy, dz, dx, dy, dz);
      if (!(d2 > (R)0)) continue;
      const MatPairR<R>& P = mpair(ms[p], ms[j]);
      R dd = std::sqrt(d2);
      R invd = (R)1 / dd;
      R hd = P.h - dd;
      R rinvj = vj.w;
      if (press) {
        R c = (P.csm * (hd * hd)) * (xi.w + xj.w);
        s.fp[0] = fmaR(dx * invd, c, s.fp[0]);
        s.fp[1] = fmaR(dy * invd, c, s.fp[1]);
        s.fp[2] = fmaR(dz * invd, c, s.fp[2]);
      }
      R lv = ((P.cvm * hd) * rinvj) * P.mu;
      s.fv[0] = fmaR(vj.x - vi.x, lv, s.fv[0]);
      s.fv[1] = fmaR(vj.y - vi.y, lv, s.fv[1]);
      s.fv[2] = fmaR(vj.z - vi.z, lv, s.fv[2]);
      R aa = P.h2 - d2;
      R bb = P.h2x3 - (R)7 * d2;
      R cj = P.csf * rinvj;
      s.fs = fmaR(aa * bb, cj, s.fs);
      R an = (aa * aa) * cj;
      s.nn[0] = fmaR(dx, an, s.nn[0]);
      s.nn[1] = fmaR(dy, an, s.nn[1]);
      s.nn[2] = fmaR(dz, an, s.nn[2]);
    }
    if (lanes == 1) return q[0];
    PairSums<R> s;
    auto mix = [&](R PairSums<R>::*f) {
      const R a[4] = {q[0].*f, q[1].*f, q[2].*f, q[3].*f};
      return mix_total(a);
    };
    auto mix3 = [&](R (PairSums<R>::*f)[3], R out[3]) {
      for (int c = 0; c < 3; ++c) {
        const R a[4] = {(q[0].*f)[c], (q[1].*f)[c], (q[2].*f)[c], (q[3].*f)[c]};
        out[c] = mix_total(a);
      }
    };
    mix3(&PairSums<R>::fp, s.fp);
    mix3(&PairSums<R>::fv, s.fv);
    s.fs = mix(&PairSums<R>::fs);
    mix3(&PairSums<R>::nn, s.nn);
    return s;
  }

  static void mixed_surface(const PairSums<R>& s, R l, R sigma, R fS[3]) {
    R ln = std::sqrt(dot3(s.nn[0], s.nn[1], s.nn[2], s.nn[0], s.nn[1], s.nn[2]));
    if (ln >= l) {
      R tt = s.fs * (-sigma);
      R r = (R)1 / ln;
      fS[0] = (s.nn[0] * tt) * r; fS[1] = (s.nn[1] * tt) * r; fS[2] = (s.nn[2] * tt) * r;
    } else {
      fS[0] = fS[1] = fS[2] = 0;
    }
  }

  // SPHSystem::update with per-particle materials: densityEvaluation with
  // the neighbour's radius and mass, p = k_i (rho - rho0_i)
  // (SphKernel.cu:9-27), internalForces (:32-91), gravity, leapfrog, box,
  // interpolation.
  void mixed_wcsph_step() {
    mixed_lists();
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const MatOwnR<R>& o = mt.own[ms[p]];
      const R r = mixed_density(p, Ax);
      const R pr = o.k * (r - o.rho0);
      rho[p] = r;
      pres[p] = pr;
      Ax[p].w = pr / (r * r);
      Av[p].w = (R)1 / r;
    }
    uint32_t zf = 0;
#pragma omp parallel for schedule(static) reduction(| : zf) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const MatOwnR<R>& o = mt.own[ms[p]];
      const PairSums<R> s = mixed_sums(p, true);
      const R ri = rho[p];
      const R cpf = -ri;
      R fS[3];
      mixed_surface(s, o.l, o.sigma, fS);
      R F0 = (s.fv[0] + fS[0]) + s.fp[0] * cpf;
      R F1 = (s.fv[1] + fS[1]) + s.fp[1] * cpf;
      R F2 = (s.fv[2] + fS[2]) + s.fp[2] * cpf;
      if (!(ri > (R)0)) zf |= PSB_FLAG_ZERO_DENSITY;
      F0 = F0 + d.gx * ri; F1 = F1 + d.gy * ri; F2 = F2 + d.gz * ri;
      const R inv = (R)1 / ri;
      V4<R> vh = Avh[p], x = Ax[p];
      vh.x = vh.x + (F0 * inv) * d.dt;
      vh.y = vh.y + (F1 * inv) * d.dt;
      vh.z = vh.z + (F2 * inv) * d.dt;
      x.x = x.x + vh.x * d.dt; x.y = x.y + vh.y * d.dt; x.z = x.z + vh.z * d.dt;
      finish_particle(p, x, vh);
      Bp[p] = pres[p];
    }
    flags |= zf;
    last_iters = 0; last_err = 0;
  }

  // PCI_SPHSystem::update with per-particle materials: pci_SPH_update
  // (PciSphSystem.cu:8-68) with compute_Density_PCI, compute_Pressure_PCI,
  // compute_Pressure_Force_PCI and compute_Viscosity_SurfaceTension_Force_PCI
  // (PciSphSystem_Kernel.cu:36-170) on max(h_i, h_j), three iterations.
  void mixed_pcisph_step() {
    mixed_lists();
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const R r = mixed_density(p, Ax, 4);
      rho[p] = r;
      pres[p] = 0;
      Av[p].w = (R)1 / r;
    }
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const MatOwnR<R>& o = mt.own[ms[p]];
      const PairSums<R> s = mixed_sums(p, false, 4);
      R fS[3];
      mixed_surface(s, o.l, (R)0.0728f, fS);
      const R ri = rho[p];
      const R fa0 = (R)0 + s.fv[0], fa1 = ((R)-9.81f * ri) + s.fv[1], fa2 = (R)0 + s.fv[2];
      fa[p] = {fa0, fa1, fa2, 0};
      fs4[p] = {fS[0], fS[1], fS[2], 0};
      const R F0 = (fa0 + (R)0) + fS[0], F1 = (fa1 + (R)0) + fS[1], F2 = (fa2 + (R)0) + fS[2];
      const R inv = (R)1 / ri;
      const V4<R> x0 = Ax[p], vh0 = Avh[p];
      const R v0 = vh0.x + (F0 * inv) * d.dt, v1 = vh0.y + (F1 * inv) * d.dt,
              v2 = vh0.z + (F2 * inv) * d.dt;
      xp0[p] = {x0.x + v0 * d.dt, x0.y + v1 * d.dt, x0.z + v2 * d.dt, 0};
    }
    std::vector<V4<R>>* cur = &xp0;
    std::vector<V4<R>>* nxt = &xp1;
    constexpr int kIters = 3;
    R err = 0;
    for (int it = 0; it < kIters; ++it) {
      err = mixed_pci_density(*cur);
      mixed_pci_force(*cur, *nxt, it == kIters - 1);
      std::swap(cur, nxt);
    }
    last_iters = kIters; last_err = err;
    total_iters += kIters;
    worst_err = std::max(worst_err, err);
    iter_hist.push_back(kIters); err_hist.push_back(err);
  }

  // The mixed corrector passes' sums: four partial sums over the list
  // entries k = 0, 1, 2, 3 (mod 4), each in increasing k, added as
  // (s0 + s1) + (s2 + s3) (the kernels' four lanes per particle).
  static R mix_total(const R (&a)[4]) { return (a[0] + a[1]) + (a[2] + a[3]); }

  R mixed_pci_density(std::vector<V4<R>>& X) {
    R emax = 0;
#pragma omp parallel for schedule(static) reduction(max : emax) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const V4<R> xi = X[p];
      R S4[4] = {}, X4[4] = {}, Y4[4] = {}, Z4[4] = {}, V4s[4] = {};
      for (uint32_t k = 0; k < count[p]; ++k) {
        const uint32_t a = k & 3u;
        const uint32_t j = list[(uint64_t)k * n + p];
        const V4<R> xj = X[j];
        const MatPairR<R>& P = mpair(ms[p], ms[j]);
        R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        R d2 = dot3(dx, dy, dz, dx, dy, dz);
        R w = P.h2 - d2;
        S4[a] = fmaR((w * w) * w, P.mc, S4[a]);
        if (d2 > (R)0) {
          R dd = std::sqrt(d2);
          R invd = (R)1 / dd;
          R hd = P.h - dd;
          R g = P.cs * (hd * hd);
          R ux = (dx * invd) * g, uy = (dy * invd) * g, uz = (dz * invd) * g;
          X4[a] = X4[a] + ux; Y4[a] = Y4[a] + uy; Z4[a] = Z4[a] + uz;
          V4s[a] = V4s[a] + dot3(ux, uy, uz, ux, uy, uz);
        }
      }
      const R S = mix_total(S4), W1x = mix_total(X4), W1y = mix_total(Y4), W1z = mix_total(Z4),
              W2 = mix_total(V4s);
      const MatOwnR<R>& o = mt.own[ms[p]];
      const R r = S;
      const R e = r - o.rho0;
      R pr = pres[p];
      if (r > o.rho0 && W2 > (R)0) pr = pr + e / (o.beta * (dot3(W1x, W1y, W1z, W1x, W1y, W1z) + W2));
      pres[p] = pr;
      rho_pred[p] = r;
      X[p].w = pr / (r * r);
      emax = std::max(emax, (e > (R)0 ? e : (R)0) * o.inv_rho0);
    }
    return emax;
  }

  void mixed_pci_force(const std::vector<V4<R>>& X, std::vector<V4<R>>& Xn, bool fin) {
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
      const uint32_t p = (uint32_t)kk;
      const V4<R> xi = X[p];
      R sx4[4] = {}, sy4[4] = {}, sz4[4] = {};
      for (uint32_t k = 0; k < count[p]; ++k) {
        const uint32_t a = k & 3u;
        const uint32_t j = list[(uint64_t)k * n + p];
        const V4<R> xj = X[j];
        R dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        R d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (!(d2 > (R)0)) continue;
        const MatPairR<R>& P = mpair(ms[p], ms[j]);
        R dd = std::sqrt(d2);
        R invd = (R)1 / dd;
        R hd = P.h - dd;
        R c = (P.csm * (hd * hd)) * (xi.w + xj.w);
        sx4[a] = fmaR(dx * invd, c, sx4[a]);
        sy4[a] = fmaR(dy * invd, c, sy4[a]);
        sz4[a] = fmaR(dz * invd, c, sz4[a]);
      }
      const R sx = mix_total(sx4), sy = mix_total(sy4), sz = mix_total(sz4);
      const R rp = rho_pred[p];
      const R cpf = -rp;
      const V4<R> f = fa[p], s = fs4[p];
      const R F0 = (f.x + sx * cpf) + s.x, F1 = (f.y + sy * cpf) + s.y, F2 = (f.z + sz * cpf) + s.z;
      const R inv = (R)1 / rp;
      const V4<R> x0 = Ax[p], vh0 = Avh[p];
      V4<R> vh = {vh0.x + (F0 * inv) * d.dt, vh0.y + (F1 * inv) * d.dt, vh0.z + (F2 * inv) * d.dt, 0};
      V4<R> x = {x0.x + vh.x * d.dt, x0.y + vh.y * d.dt, x0.z + vh.z * d.dt, 0};
      if (!fin) {
        Xn[p] = x;
      } else {
        finish_particle(p, x, vh);
        Bp[p] = pres[p];
      }
    }
  }

  void update(int steps) {
    flags = 0; total_iters = 0; worst_err = 0; steps_done = 0; max_nbr_seen = 0;
    retried_steps = 0; substeps = 0;
    for (int s = 0; s < steps; ++s) {
      if (nmat) {
        if (prm.solver == PSB_SOLVER_PCISPH) mixed_pcisph_step(); else mixed_wcsph_step();
      } else if (prm.solver == PSB_SOLVER_PCISPH) {
        pcisph_step();
      } else {
        wcsph_step();
      }
      steps_done++;
    }
  }
};

struct Handle {
  int prec;  // 32 or 64
  System<float>* f = nullptr;
  System<double>* d = nullptr;
};

template <class R>
int set_particles(System<R>* s, uint32_t n, const float* pos4, const float* vel4, const float* vh4) {
  const uint32_t nb = s->n_spheres * s->per_sphere;
  if ((uint64_t)n + nb > s->cap) return 3;
  s->n = n;
  s->nmat = 0;   // per-particle materials refer to the previous particles
  for (uint32_t i = 0; i < n; ++i) {
    s->Bx[i] = {(R)pos4[4 * i], (R)pos4[4 * i + 1], (R)pos4[4 * i + 2], 0};
    s->Bv[i] = vel4 ? V4<R>{(R)vel4[4 * i], (R)vel4[4 * i + 1], (R)vel4[4 * i + 2], 0} : V4<R>{0, 0, 0, 0};
    s->Bvh[i] = vh4 ? V4<R>{(R)vh4[4 * i], (R)vh4[4 * i + 1], (R)vh4[4 * i + 2], 0} : V4<R>{0, 0, 0, 0};
    s->Bid[i] = i;
    s->Bp[i] = 0;
  }
  // sphere boundary particles after the fluid, at rest relative to their sphere
  s->n_fluid = s->n_spheres ? n : System<R>::kNone;
  for (uint32_t j = 0; j < s->n_spheres; ++j)
    for (uint32_t l = 0; l < s->per_sphere; ++l) {
      const uint32_t k = n + j * s->per_sphere + l;
      s->Bx[k] = {s->sph_c[j].x + (R)s->sph_off[4 * l], s->sph_c[j].y + (R)s->sph_off[4 * l + 1],
                  s->sph_c[j].z + (R)s->sph_off[4 * l + 2], 0};
      s->Bv[k] = s->sph_v[j];
      s->Bvh[k] = s->sph_v[j];
      s->Bid[k] = k;
      s->Bp[k] = 0;
    }
  s->n = n + nb;
  s->Ax = s->Bx; s->Av = s->Bv; s->Avh = s->Bvh; s->Aid = s->Bid;
  return 0;
}

template <class R>
void set_spheres(System<R>* s, uint32_t count, const float* c3, const float* v3,
                 const psb_sphere_params* sp) {
  s->n_spheres = count;
  s->per_sphere = 0;
  s->sph_off.clear();
  if (!count) return;
  s->sphere = *sp;
  const double lat = s->prm.lattice_spacing > 0.f ? (double)s->prm.lattice_spacing
                                                  : 0.5 * (double)s->prm.support_radius;
  const double r = sp->radius;
  const int m = (int)std::ceil(r / lat);
  for (int k = -m; k <= m; ++k)
    for (int j = -m; j <= m; ++j)
      for (int i = -m; i <= m; ++i) {
        const double d2 = ((double)i * i + (double)j * j + (double)k * k) * lat * lat;
        if (d2 < r * r) {
          s->sph_off.push_back((float)(i * lat)); s->sph_off.push_back((float)(j * lat));
          s->sph_off.push_back((float)(k * lat)); s->sph_off.push_back(0.f);
        }
      }
  s->per_sphere = (uint32_t)(s->sph_off.size() / 4);
  s->sph_c.assign(count, {});
  s->sph_v.assign(count, {});
  s->sph_f.assign(count, {});
  for (uint32_t j = 0; j < count; ++j) {
    s->sph_c[j] = {(R)c3[3 * j], (R)c3[3 * j + 1], (R)c3[3 * j + 2], 0};
    if (v3) s->sph_v[j] = {(R)v3[3 * j], (R)v3[3 * j + 1], (R)v3[3 * j + 2], 0};
  }
}

template <class R>
void get_particles(System<R>* s, float* pos4, float* vel4, float* vh4, float* old_pos4,
                   float* old_vel4, float* vav4, float* rho, float* pres, uint32_t* ids) {
  for (uint32_t i = 0; i < s->n; ++i) {
    auto put = [&](float* dst, const V4<R>& v) {
      if (!dst) return;
      dst[4 * i] = (float)v.x; dst[4 * i + 1] = (float)v.y; dst[4 * i + 2] = (float)v.z; dst[4 * i + 3] = 0.f;
    };
    put(pos4, s->Bx[i]); put(vel4, s->Bv[i]); put(vh4, s->Bvh[i]);
    put(old_pos4, s->Ax[i]); put(old_vel4, s->Av[i]); put(vav4, s->Avh[i]);
    if (rho) rho[i] = (float)(s->prm.solver == PSB_SOLVER_PCISPH ? s->rho_pred[i] : s->rho[i]);
    if (pres) pres[i] = (float)s->pres[i];
    if (ids) ids[i] = s->Bid[i];
  }
}

// ---------------------------------------------------------------------------
// Surface extraction (SurfaceSPH.cpp:29-190, evaluateIso_Kernel
// UniformGridKernel.cu:98-138, MarchingCube_Kernel.cu:14-158), restated with
// the documented differences: the field sums every particle within
// r = h / scale, one vertex per crossed node-grid edge (numbered by node,
// x fastest, then axis x, y, z), normals along -grad f, and a case table built
// from the contours: on each face (corners counter-clockwise seen from
// outside) every maximal run of corners <= iso is cut off by a segment from
// the crossed edge entering the run to the one leaving it; segments chain
// into loops, a loop starts at its smallest edge, loops go in that order and
// are fanned from their start, wound towards the low side.
struct McTab {
  int ntri[256];
  int tri[256][15];
};

const int kMcCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                             {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
const int kMcEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                            {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// faces, corners counter-clockwise seen from outside: -x +x -y +y -z +z
const int kMcFace[6][4] = {{3, 7, 4, 0}, {1, 5, 6, 2}, {1, 2, 3, 0},
                           {4, 7, 6, 5}, {4, 5, 1, 0}, {3, 2, 6, 7}};

int mc_edge_between(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((kMcEdge[e][0] == a && kMcEdge[e][1] == b) || (kMcEdge[e][0] == b && kMcEdge[e][1] == a)) return e;
  return -1;
}

McTab mc_table() {
  McTab t{};
  for (int cs = 0; cs < 256; ++cs) {
    int next[12];
    for (int e = 0; e < 12; ++e) next[e] = -1;
    for (const auto& f : kMcFace)
      for (int i = 0; i < 4; ++i) {
        const int c = f[i], prev = f[(i + 3) % 4];
        const bool bc = (cs >> c) & 1, bp = (cs >> prev) & 1;
        if (!bc || bp) continue;
        int j = i;
        while ((cs >> f[(j + 1) % 4]) & 1) j = (j + 1) % 4;
        next[mc_edge_between(prev, c)] = mc_edge_between(f[j], f[(j + 1) % 4]);
      }
    bool seen[12] = {false};
    int nt = 0;
    for (int e0 = 0; e0 < 12; ++e0) {
      if (next[e0] < 0 || seen[e0]) continue;
      int loop[12], m = 0;
      for (int e = e0; !seen[e]; e = next[e]) { seen[e] = true; loop[m++] = e; }
      for (int k = 1; k + 1 < m; ++k) {
        // low side: with corner 0 alone below (case 1) the loop e0 -> e3 -> e8
        // runs clockwise seen from corner 0, so triangles take (0, k+1, k)
        t.tri[cs][3 * nt] = loop[0];
        t.tri[cs][3 * nt + 1] = loop[k + 1];
        t.tri[cs][3 * nt + 2] = loop[k];
        ++nt;
      }
    }
    for (int k = 3 * nt; k < 15; ++k) t.tri[cs][k] = -1;
    t.ntri[cs] = nt;
  }
  return t;
}

struct SurfaceOut {
  std::vector<float> field, verts, norms;
  std::vector<uint32_t> tris;
  uint32_t dims[3] = {0, 0, 0};
  float origin[3] = {0, 0, 0}, cell = 0;
};

inline int clamp_cell(float x, float mn, float cell, int n) {
  float t = std::floor((x - mn) / cell);
  if (!(t >= 0.0f && t < (float)n)) t = (t >= (float)n) ? (float)(n - 1) : 0.0f;
  return (int)t;
}

void surface_extract(const float* x4, const uint32_t* ids, uint32_t n, uint32_t limit, float h,
                     const psb_surface_params& prm, SurfaceOut& o) {
  auto kept = [&](uint32_t p) { return !ids || ids[p] < limit; };
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool any = false;
  for (uint32_t p = 0; p < n; ++p) {
    if (!kept(p)) continue;
    for (int a = 0; a < 3; ++a) {
      const float v = x4[4 * p + a];
      if (!any || v < lo[a]) lo[a] = v;
      if (!any || v > hi[a]) hi[a] = v;
    }
    any = true;
  }
  if (!any) return;
  const float r = h / prm.scale;
  const double margin = (double)prm.margin * (double)r;
  double mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = (double)lo[a] - margin;
    mx[a] = (double)hi[a] + margin;
    const double len = std::fabs(mx[a] - mn[a]);
    o.dims[a] = (uint32_t)std::ceil(len / (double)prm.cell_size);
    o.origin[a] = (float)((mn[a] + mx[a]) / 2 - len / 2 + (double)prm.cell_size / 2);
  }
  o.cell = prm.cell_size;
  const int NX = (int)o.dims[0], NY = (int)o.dims[1], NZ = (int)o.dims[2];
  // particle grid of cell r
  Grid g;
  for (int a = 0; a < 3; ++a) g.min[a] = (float)mn[a];
  g.cell = r;
  g.nx = (int)std::ceil((mx[0] - mn[0]) / (double)r) + 1;
  g.ny = (int)std::ceil((mx[1] - mn[1]) / (double)r) + 1;
  g.nz = (int)std::ceil((mx[2] - mn[2]) / (double)r) + 1;
  std::vector<uint32_t> hash(n), perm(n), start(g.ncells() + 1);
  uint32_t fl = 0;
  grid_build(g, n, x4, hash.data(), perm.data(), start.data(), fl);
  auto near = [&](float x, float y, float z, auto&& f) {
    const int cx = clamp_cell(x, g.min[0], g.cell, g.nx);
    const int cy = clamp_cell(y, g.min[1], g.cell, g.ny);
    const int cz = clamp_cell(z, g.min[2], g.cell, g.nz);
    for (int zz = std::max(cz - 1, 0); zz <= std::min(cz + 1, g.nz - 1); ++zz)
      for (int yy = std::max(cy - 1, 0); yy <= std::min(cy + 1, g.ny - 1); ++yy) {
        const uint64_t row = (uint64_t)g.nx * ((uint64_t)yy + (uint64_t)g.ny * zz);
        const uint32_t q0 = start[row + std::max(cx - 1, 0)], q1 = start[row + std::min(cx + 1, g.nx - 1) + 1];
        for (uint32_t q = q0; q < q1; ++q) f(perm[q]);
      }
  };
  const float r2 = r * r;
  const float inv_r6 = 1.0f / (r2 * r2 * r2);
  const float iso = prm.iso_level;
  auto node = [&](int i, int j, int k, float& x, float& y, float& z) {
    x = o.origin[0] + (float)i * o.cell;
    y = o.origin[1] + (float)j * o.cell;
    z = o.origin[2] + (float)k * o.cell;
  };
  const uint64_t NN = (uint64_t)NX * NY * NZ;
  o.field.assign(NN, 0.f);
#pragma omp parallel for schedule(static) if(NN >= kOmpMinItems)
  for (int64_t t = 0; t < (int64_t)NN; ++t) {
    const int i = (int)(t % NX), j = (int)((t / NX) % NY), k = (int)(t / ((int64_t)NX * NY));
    float x, y, z;
    node(i, j, k, x, y, z);
    float acc = 0.f;
    near(x, y, z, [&](uint32_t p) {
      if (!kept(p)) return;
      const float dx = x - x4[4 * p], dy = y - x4[4 * p + 1], dz = z - x4[4 * p + 2];
      const float d2 = dot3(dx, dy, dz, dx, dy, dz);
      if (d2 <= r2) {
        const float w = r2 - d2;
        acc = std::fma(w * w, w, acc);
      }
    });
    o.field[t] = acc * inv_r6;
  }
  // crossed edges, vertex numbers
  std::vector<uint8_t> ef(NN);
  std::vector<uint32_t> vbase(NN + 1);
  uint32_t nv = 0;
  const uint64_t SY = (uint64_t)NX, SZ = (uint64_t)NX * NY;
  for (uint64_t t = 0; t < NN; ++t) {
    const int i = (int)(t % NX), j = (int)((t / NX) % NY), k = (int)(t / SZ);
    const bool b0 = o.field[t] <= iso;
    uint8_t f = 0;
    if (i + 1 < NX && (o.field[t + 1] <= iso) != b0) f |= 1;
    if (j + 1 < NY && (o.field[t + SY] <= iso) != b0) f |= 2;
    if (k + 1 < NZ && (o.field[t + SZ] <= iso) != b0) f |= 4;
    ef[t] = f;
    vbase[t] = nv;
    nv += (f & 1) + ((f >> 1) & 1) + ((f >> 2) & 1);
  }
  vbase[NN] = nv;
  o.verts.assign(3 * (size_t)nv, 0.f);
  o.norms.assign(3 * (size_t)nv, 0.f);
#pragma omp parallel for schedule(static) if(NN >= kOmpMinItems)
  for (int64_t t = 0; t < (int64_t)NN; ++t) {
    if (!ef[t]) continue;
    const int i = (int)(t % NX), j = (int)((t / NX) % NY), k = (int)(t / SZ);
    float ax, ay, az;
    node(i, j, k, ax, ay, az);
    const float va = o.field[t];
    uint32_t v = vbase[t];
    for (int axis = 0; axis < 3; ++axis) {
      if (!((ef[t] >> axis) & 1)) continue;
      const uint64_t tb = t + (axis == 0 ? 1 : axis == 1 ? SY : SZ);
      float bx, by, bz;
      node(i + (axis == 0), j + (axis == 1), k + (axis == 2), bx, by, bz);
      const float vb = o.field[tb];
      float px, py, pz;
      if (std::fabs(iso - va) < 0.000001f) { px = ax; py = ay; pz = az; }
      else if (std::fabs(iso - vb) < 0.000001f) { px = bx; py = by; pz = bz; }
      else if (std::fabs(va - vb) < 0.000001f) { px = ax; py = ay; pz = az; }
      else {
        const float mu = (iso - va) / (vb - va);
        px = ax + mu * (bx - ax); py = ay + mu * (by - ay); pz = az + mu * (bz - az);
      }
      float gx = 0.f, gy = 0.f, gz = 0.f;
      near(px, py, pz, [&](uint32_t p) {
        if (!kept(p)) return;
        const float dx = px - x4[4 * p], dy = py - x4[4 * p + 1], dz = pz - x4[4 * p + 2];
        const float d2 = dot3(dx, dy, dz, dx, dy, dz);
        if (d2 < r2) {
          const float w = r2 - d2;
          const float w2 = w * w;
          gx = std::fma(dx, w2, gx); gy = std::fma(dy, w2, gy); gz = std::fma(dz, w2, gz);
        }
      });
      const float gl = std::sqrt(dot3(gx, gy, gz, gx, gy, gz));
      if (gl > 0.f) { gx = gx / gl; gy = gy / gl; gz = gz / gl; }
      o.verts[3 * v] = px; o.verts[3 * v + 1] = py; o.verts[3 * v + 2] = pz;
      o.norms[3 * v] = gx; o.norms[3 * v + 1] = gy; o.norms[3 * v + 2] = gz;
      ++v;
    }
  }
  // cubes, x fastest
  static const McTab tab = mc_table();
  // edge -> node offset and axis
  const int em[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 2}, {0, 0, 1, 0}, {0, 0, 0, 2},
                         {0, 1, 0, 0}, {1, 1, 0, 2}, {0, 1, 1, 0}, {0, 1, 0, 2},
                         {0, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 1, 1}, {0, 0, 1, 1}};
  o.tris.clear();
  for (int k = 0; k + 1 < NZ; ++k)
    for (int j = 0; j + 1 < NY; ++j)
      for (int i = 0; i + 1 < NX; ++i) {
        const uint64_t t = (uint64_t)i + SY * j + SZ * k;
        int cs = 0;
        for (int c = 0; c < 8; ++c) {
          const uint64_t tc = t + kMcCorner[c][0] + SY * kMcCorner[c][1] + SZ * kMcCorner[c][2];
          if (o.field[tc] <= iso) cs |= 1 << c;
        }
        for (int q = 0; q < 3 * tab.ntri[cs]; ++q) {
          const int e = tab.tri[cs][q];
          const uint64_t nd = t + em[e][0] + SY * em[e][1] + SZ * em[e][2];
          const int ax = em[e][3];
          uint32_t rank = 0;
          for (int b = 0; b < ax; ++b) rank += (ef[nd] >> b) & 1;
          o.tris.push_back(vbase[nd] + rank);
        }
      }
}

}  // namespace

// ===========================================================================
// C entry points (ctypes).  All arrays are host memory.
// ===========================================================================
extern "C" {

// rsqrt_pair on n inputs (the pressure force's 1/sqrt)
void oracle_rsqrt_pair(const float* x, float* out, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) out[i] = rsqrt_pair(x[i]);
}

int oracle_derive(const psb_sph_params* p, double* out /* 16 */) {
  Derived<double> d = derive<double>(*p);
  Derived<float> f = derive<float>(*p);
  double v[16] = {f.h2, f.c_poly6, f.c_spiky, f.c_visc, f.c_surf_lap, f.c_surf_grad, f.mc,
                  f.beta, f.delta, f.tait_b, (double)f.R_stencil, d.delta, d.beta, d.c_poly6,
                  d.c_spiky, d.mc};
  std::memcpy(out, v, sizeof(v));
  return 0;
}

uint32_t oracle_grid_build(uint32_t n, const float* x4, const psb_grid_desc* gd, uint32_t* hash,
                           uint32_t* perm, uint32_t* start) {
  uint32_t flags = 0;
  grid_build(grid_from(*gd), n, x4, hash, perm, start, flags);
  return flags;
}

uint32_t oracle_neighbors(uint32_t n, const float* sorted_x4, const psb_grid_desc* gd,
                          const uint32_t* start, float radius, uint32_t cap, uint32_t* count,
                          uint32_t* list) {
  Grid g = grid_from(*gd);
  double rr = std::ceil((double)radius / (double)g.cell);
  if ((rr - 1.0) * (double)g.cell >= (double)radius) rr -= 1.0;
  return neighbors<float>(g, start, n, reinterpret_cast<const V4<float>*>(sorted_x4),
                          radius * radius, (int)rr, cap, count, list, nullptr, nullptr);
}

// evaluate_densities_forces restated on a given neighbour list
// (SphKernel.cu:9-91).  Arrays in the list's index space.
int oracle_evaluate_densities_forces(const psb_sph_params* prm, uint32_t n, const uint32_t* count,
                                     const uint32_t* list, const float* pos4, const float* vel4,
                                     float* density, float* pressure, float* normals4,
                                     float* fP4, float* fV4, float* fS4, float* F4) {
  Derived<float> d = derive<float>(*prm);
  std::vector<V4<float>> xs(n), vs(n);
  std::memcpy(xs.data(), pos4, 16 * (size_t)n);
  std::memcpy(vs.data(), vel4, 16 * (size_t)n);
  for (uint32_t p = 0; p < n; ++p) {
    float S = 0;
    for (uint32_t k = 0; k < count[p]; ++k) {
      uint32_t q = list[(uint64_t)k * n + p];
      float dx = xs[p].x - xs[q].x, dy = xs[p].y - xs[q].y, dz = xs[p].z - xs[q].z;
      float d2 = dot3(dx, dy, dz, dx, dy, dz);
      float w = d.h2 - d2;
      S = std::fma(w * w, w, S);
    }
    float r = S * d.mc;
    density[p] = r;
    float pr = eos(d, *prm, r);
    pressure[p] = pr;
  }
  for (uint32_t p = 0; p < n; ++p) {
    xs[p].w = pressure[p] / (density[p] * density[p]);
    vs[p].w = 1.0f / density[p];
  }
  const bool surf = prm->surface_tension != 0.f;
  for (uint32_t p = 0; p < n; ++p) {
    PairSums<float> s = surf ? internal_sums<float, true, true>(d, p, n, count, list, xs.data(), vs.data())
                             : internal_sums<float, true, false>(d, p, n, count, list, xs.data(), vs.data());
    float ri = density[p];
    float cpf = -(d.m * ri);
    float fP[3] = {s.fp[0] * cpf, s.fp[1] * cpf, s.fp[2] * cpf};
    float fV[3] = {s.fv[0] * d.cv, s.fv[1] * d.cv, s.fv[2] * d.cv};
    float fS[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
    if (surf) surface_finish(d, s, d.sigma, fS, nrm);
    for (int a = 0; a < 3; ++a) {
      if (normals4) normals4[4 * p + a] = nrm[a];
      if (fP4) fP4[4 * p + a] = fP[a];
      if (fV4) fV4[4 * p + a] = fV[a];
      if (fS4) fS4[4 * p + a] = fS[a];
      if (F4) F4[4 * p + a] = (fV[a] + fS[a]) + fP[a];
    }
  }
  return 0;
}

int oracle_collision_box(const psb_sph_params* prm, uint32_t n, float* pos4, float* vel4) {
  Derived<float> d = derive<float>(*prm);
  for (uint32_t i = 0; i < n; ++i) {
    V4<float> x = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], pos4[4 * i + 3]};
    V4<float> v = {vel4[4 * i], vel4[4 * i + 1], vel4[4 * i + 2], vel4[4 * i + 3]};
    box_collide(d, x, v);
    pos4[4 * i] = x.x; pos4[4 * i + 1] = x.y; pos4[4 * i + 2] = x.z;
    vel4[4 * i] = v.x; vel4[4 * i + 1] = v.y; vel4[4 * i + 2] = v.z;
  }
  return 0;
}

// DEM contact force on A from B (collideParticle, CudaParticle_Kernel.cu:10-38)
// in FP32: d = |x_B - x_A|, n = (x_B - x_A) * (1/d), R = r_A + r_B.  Reference
// mode: -k (R - d) n for 0 < d < 2R (damping / shear / attraction are dropped
// by the reference's non-mutating double3 +=, double3.h:115-123).  Config
// mode: (-k (R - d) + c (v_B - v_A).n) n in contact only (0 < d < R).
static void dem_pair(bool ref, const float* xa, const float* xb, const float* vb,
                     const float* va, float ra, float rb, float k, float c, float* f) {
  f[0] = 0.f; f[1] = 0.f; f[2] = 0.f;
  const float dx = xb[0] - xa[0], dy = xb[1] - xa[1], dz = xb[2] - xa[2];
  const float dist = std::sqrt(dot3(dx, dy, dz, dx, dy, dz));
  if (!(dist > 0.f)) return;
  const float R = ra + rb;
  if (ref ? !(dist < R * 2.0f) : !(dist < R)) return;
  const float inv = 1.0f / dist;
  const float nx = dx * inv, ny = dy * inv, nz = dz * inv;
  float s = -k * (R - dist);
  if (!ref) {
    const float vn = dot3(vb[0] - va[0], vb[1] - va[1], vb[2] - va[2], nx, ny, nz);
    s = s + c * vn;
  }
  f[0] = s * nx; f[1] = s * ny; f[2] = s * nz;
}

// interactionSystem (CudaSystem.cu:6-14 -> collideParticles,
// CudaParticle_Kernel.cu:42-55): forces[A] += contact of each listed B, in
// list order, with A's parameters.
int oracle_interaction_system(int mode, uint32_t n, const float* pos4, const float* vel4,
                              float* forces4, const float* radius, const float* spring,
                              const float* damping, const uint32_t* count, const uint32_t* list) {
  const bool ref = mode == PSB_MODE_REFERENCE;
  const float zero[4] = {0.f, 0.f, 0.f, 0.f};
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
  for (int64_t aa = 0; aa < (int64_t)n; ++aa) {
    const uint32_t a = (uint32_t)aa;
    const float* xa = pos4 + 4 * (size_t)a;
    const float* va = ref ? zero : vel4 + 4 * (size_t)a;
    float fx = forces4[4 * (size_t)a], fy = forces4[4 * (size_t)a + 1], fz = forces4[4 * (size_t)a + 2];
    for (uint32_t j = 0; j < count[a]; ++j) {
      const uint32_t b = list[(uint64_t)j * n + a];
      float f[3];
      dem_pair(ref, xa, pos4 + 4 * (size_t)b, ref ? zero : vel4 + 4 * (size_t)b, va, radius[a],
               radius[b], spring[a], ref ? 0.f : damping[a], f);
      fx = fx + f[0]; fy = fy + f[1]; fz = fz + f[2];
    }
    forces4[4 * (size_t)a] = fx; forces4[4 * (size_t)a + 1] = fy; forces4[4 * (size_t)a + 2] = fz;
  }
  return 0;
}

// Lattice-wall sums at n points (FP32): out[4 i] = sum (h^2 - d^2)^3,
// out[4 i + 1 .. 3] = sum of spiky gradients gradW(x_i - x_b).
int oracle_wall_sums(const psb_sph_params* prm, uint32_t n, const float* xyz, float* out) {
  Derived<float> d = derive<float>(*prm);
  if (!d.wall) return 1;
  for (uint32_t i = 0; i < n; ++i)
    wall_sums(d, true, true, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], out[4 * i],
              out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
  return 0;
}

void* oracle_system_create(const psb_sph_params* prm, uint32_t capacity, int precision) {
  Handle* h = new Handle();
  h->prec = precision;
  if (precision == 64) { h->d = new System<double>(); h->d->init(*prm, capacity); }
  else { h->f = new System<float>(); h->f->init(*prm, capacity); }
  return h;
}

void oracle_system_destroy(void* hp) {
  Handle* h = (Handle*)hp;
  delete h->f; delete h->d; delete h;
}

int oracle_system_set_particles(void* hp, uint32_t n, const float* pos4, const float* vel4,
                                const float* vh4) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) return set_particles(h->d, n, pos4, vel4, vh4);
  return set_particles(h->f, n, pos4, vel4, vh4);
}

// Per-particle materials (psb_system_set_materials): 0 on success, 1 on an
// invalid argument (as the library refuses it).
int oracle_system_set_materials(void* hp, uint32_t count, const psb_sph_params* mats,
                                const uint8_t* index, uint32_t n) {
  Handle* h = (Handle*)hp;
  auto set = [&](auto* s) -> int {
    if (count == 0) { s->nmat = 0; return 0; }
    if (s->prm.mode != PSB_MODE_REFERENCE) return 1;
    if (s->prm.solver != PSB_SOLVER_PCISPH && s->prm.eos != PSB_EOS_LINEAR) return 1;
    if (!index || n != s->n || s->n_spheres) return 1;
    for (uint32_t i = 0; i < n; ++i)
      if (index[i] >= count) return 1;
    if (make_mat_tab(s->prm, mats, count, s->mt)) return 1;
    s->mat_by_id.assign(index, index + n);
    s->nmat = count;
    return 0;
  };
  return h->prec == 64 ? set(h->d) : set(h->f);
}

uint32_t oracle_system_count(void* hp) {
  Handle* h = (Handle*)hp;
  return h->prec == 64 ? h->d->n : h->f->n;
}

// rigid DEM spheres (before set_particles, which appends their boundary particles)
int oracle_collision_collider(const psb_collider* k, float dt, uint32_t n, const float* old_pos4,
                              float* new_pos4, float* new_vel4) {
  const OColl c = make_ocoll(*k);
  for (uint32_t i = 0; i < n; ++i) {
    const V4<float> x0 = {old_pos4[4 * i], old_pos4[4 * i + 1], old_pos4[4 * i + 2], 0.f};
    V4<float> x = {new_pos4[4 * i], new_pos4[4 * i + 1], new_pos4[4 * i + 2], new_pos4[4 * i + 3]};
    V4<float> v = {new_vel4[4 * i], new_vel4[4 * i + 1], new_vel4[4 * i + 2], new_vel4[4 * i + 3]};
    collide_one(c, dt, x0, x, v);
    new_pos4[4 * i] = x.x; new_pos4[4 * i + 1] = x.y; new_pos4[4 * i + 2] = x.z;
    new_vel4[4 * i] = v.x; new_vel4[4 * i + 1] = v.y; new_vel4[4 * i + 2] = v.z;
  }
  return 0;
}

int oracle_system_add_collider(void* hp, const psb_collider* k) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) h->d->colls.push_back(make_ocoll(*k));
  else h->f->colls.push_back(make_ocoll(*k));
  return 0;
}

int oracle_system_set_spheres(void* hp, uint32_t count, const float* centers3,
                              const float* velocities3, const psb_sphere_params* sp) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) set_spheres(h->d, count, centers3, velocities3, sp);
  else set_spheres(h->f, count, centers3, velocities3, sp);
  return 0;
}

int oracle_system_get_spheres(void* hp, float* c3, float* v3, float* f3) {
  Handle* h = (Handle*)hp;
  auto get = [&](auto* s) {
    for (uint32_t j = 0; j < s->n_spheres; ++j) {
      const float cc[3] = {(float)s->sph_c[j].x, (float)s->sph_c[j].y, (float)s->sph_c[j].z};
      const float vv[3] = {(float)s->sph_v[j].x, (float)s->sph_v[j].y, (float)s->sph_v[j].z};
      const float ff[3] = {(float)s->sph_f[j].x, (float)s->sph_f[j].y, (float)s->sph_f[j].z};
      for (int a = 0; a < 3; ++a) {
        if (c3) c3[3 * j + a] = cc[a];
        if (v3) v3[3 * j + a] = vv[a];
        if (f3) f3[3 * j + a] = ff[a];
      }
    }
    return s->n_spheres;
  };
  return (int)(h->prec == 64 ? get(h->d) : get(h->f));
}

int oracle_system_update(void* hp, int steps) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) h->d->update(steps); else h->f->update(steps);
  return 0;
}

int oracle_system_get(void* hp, float* pos4, float* vel4, float* vh4, float* old_pos4,
                      float* old_vel4, float* vav4, float* rho, float* pres, uint32_t* ids) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) get_particles(h->d, pos4, vel4, vh4, old_pos4, old_vel4, vav4, rho, pres, ids);
  else get_particles(h->f, pos4, vel4, vh4, old_pos4, old_vel4, vav4, rho, pres, ids);
  return 0;
}

// stats: flags, last_iters, last_err, total_iters, worst_err, steps, max_nbr,
// retried_steps, substeps
int oracle_system_stats(void* hp, double* out /* 9 */) {
  Handle* h = (Handle*)hp;
  auto fill = [&](auto* s) {
    out[0] = s->flags; out[1] = s->last_iters; out[2] = (double)s->last_err;
    out[3] = (double)s->total_iters; out[4] = (double)s->worst_err; out[5] = s->steps_done;
    out[6] = s->max_nbr_seen; out[7] = s->retried_steps; out[8] = s->substeps;
  };
  if (h->prec == 64) fill(h->d); else fill(h->f);
  return 0;
}

// Corrector probe: the PCISPH step of the particles set last, stopped after
// its non-pressure pass (pci_prologue), then single density and force passes
// of the corrector on caller-given predicted states.  All arrays are in the
// sorted order of the prologue (ids gives the caller's particle index); vec4
// arrays are n x 4 floats.  Any pointer may be null.
int oracle_pci_prologue(void* hp) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) h->d->pci_prologue(); else h->f->pci_prologue();
  return 0;
}

extern "C++" {
template <class R> static void put4(float* dst, const std::vector<V4<R>>& v, uint32_t n) {
  if (!dst) return;
  for (uint32_t i = 0; i < n; ++i) {
    dst[4 * i] = (float)v[i].x; dst[4 * i + 1] = (float)v[i].y;
    dst[4 * i + 2] = (float)v[i].z; dst[4 * i + 3] = (float)v[i].w;
  }
}
template <class R> static void get4(std::vector<V4<R>>& v, const float* src, uint32_t n) {
  if (!src) return;
  for (uint32_t i = 0; i < n; ++i)
    v[i] = {(R)src[4 * i], (R)src[4 * i + 1], (R)src[4 * i + 2], (R)src[4 * i + 3]};
}
template <class R> static void put1(float* dst, const std::vector<R>& v, uint32_t n) {
  if (dst) for (uint32_t i = 0; i < n; ++i) dst[i] = (float)v[i];
}
template <class R> static void get1(std::vector<R>& v, const float* src, uint32_t n) {
  if (src) for (uint32_t i = 0; i < n; ++i) v[i] = (R)src[i];
}
}  // extern "C++"

// x0 / vh0: start-of-step positions and half-step velocities; xp: first
// prediction (w = 0); anp: config-mode non-pressure acceleration; fa / fs:
// reference-mode forcesAccum (gravity + viscosity) and surface force;
// delta: config-mode per-particle pressure scaling; pres, rho_pred: the
// corrector's pressure and predicted density; rho: start-of-step density.
int oracle_pci_get(void* hp, float* x0, float* vh0, float* xp, float* anp, float* fa, float* fs,
                   float* delta, float* pres, float* rho_pred, float* rho, uint32_t* ids) {
  Handle* h = (Handle*)hp;
  auto run = [&](auto* s) {
    const uint32_t n = s->n;
    put4(x0, s->Ax, n); put4(vh0, s->Avh, n); put4(xp, s->xp0, n); put4(anp, s->anp, n);
    put4(fa, s->fa, n); put4(fs, s->fs4, n);
    put1(delta, s->dl, n); put1(pres, s->pres, n); put1(rho_pred, s->rho_pred, n);
    put1(rho, s->rho, n);
    if (ids) for (uint32_t i = 0; i < n; ++i) ids[i] = s->Aid[i];
  };
  if (h->prec == 64) run(h->d); else run(h->f);
  return 0;
}

int oracle_pci_set(void* hp, const float* x0, const float* vh0, const float* anp, const float* fa,
                   const float* fs, const float* delta, const float* pres, const float* rho_pred) {
  Handle* h = (Handle*)hp;
  auto run = [&](auto* s) {
    const uint32_t n = s->n;
    get4(s->Ax, x0, n); get4(s->Avh, vh0, n); get4(s->anp, anp, n); get4(s->fa, fa, n);
    get4(s->fs4, fs, n);
    get1(s->dl, delta, n); get1(s->pres, pres, n); get1(s->rho_pred, rho_pred, n);
  };
  if (h->prec == 64) run(h->d); else run(h->f);
  return 0;
}

// One density pass (compute_Density_PCI + compute_Pressure_PCI) on the
// predicted positions xp (xyz in; w out: the pressure term the force pass
// gathers); returns the max compression error.
double oracle_pci_density(void* hp, float* xp) {
  Handle* h = (Handle*)hp;
  auto run = [&](auto* s) {
    using R = typename std::remove_reference<decltype(s->rho[0])>::type;
    std::vector<V4<R>> X(s->n);
    get4(X, xp, s->n);
    double e = (double)s->pci_density(X);
    put4(xp, X, s->n);
    return e;
  };
  return h->prec == 64 ? run(h->d) : run(h->f);
}

// One non-final force pass (compute_Pressure_Force_PCI + restore +
// integrateSPH_LeapFrog_Forces) on xp (w = pressure term); the next
// prediction goes to xn (w = 0).
int oracle_pci_force(void* hp, const float* xp, float* xn) {
  Handle* h = (Handle*)hp;
  auto run = [&](auto* s) {
    using R = typename std::remove_reference<decltype(s->rho[0])>::type;
    std::vector<V4<R>> X(s->n), Xn(s->n);
    get4(X, xp, s->n);
    s->pci_force(X, Xn, false);
    put4(xn, Xn, s->n);
  };
  if (h->prec == 64) run(h->d); else run(h->f);
  return 0;
}

int oracle_system_enable_features(void* hp, int on) {
  Handle* h = (Handle*)hp;
  if (h->prec == 64) h->d->feat_on = on != 0; else h->f->feat_on = on != 0;
  return 0;
}

int oracle_system_features(void* hp, float* out) {
  Handle* h = (Handle*)hp;
  auto fill = [&](auto* s) { for (size_t i = 0; i < 8 * (size_t)s->n; ++i) out[i] = (float)s->feat[i]; };
  if (h->prec == 64) fill(h->d); else fill(h->f);
  return 0;
}

int oracle_system_set_mlp(void* hp, const float* w, const float* mu, const float* sd) {
  Handle* h = (Handle*)hp;
  auto set = [&](auto* s) {
    s->mlp_w.assign(w, w + kNP);
    for (int i = 0; i < kIn; ++i) { s->mlp_mu[i] = mu[i]; s->mlp_sd[i] = sd[i]; }
  };
  if (h->prec == 64) set(h->d); else set(h->f);
  return 0;
}

double oracle_system_pressure_scale(void* hp) {
  Handle* h = (Handle*)hp;
  return h->prec == 64 ? (double)h->d->d.pscale : (double)h->f->d.pscale;
}

// MLP forward over a feature SoA [8][n] (unscaled network output).
int oracle_mlp_forward(const float* w, const float* mu, const float* sd, const float* feat,
                       uint64_t n, float* out, int emulate_bf16) {
  Mlp m = make_mlp(w, mu, sd, emulate_bf16 != 0);
#pragma omp parallel for schedule(static) if(n >= kOmpMinItems)
  for (int64_t s = 0; s < (int64_t)n; ++s) {
    double x[kIn], z1[kHid], h1[kHid], z2[kHid], h2[kHid];
    m.input(feat, n, (uint64_t)s, x);
    out[s] = (float)m.forward(x, z1, h1, z2, h2);
  }
  return 0;
}

// Gradient of the minibatch MSE (1/B) sum (y - t)^2 over samples idx[0..nb)
// of a dataset (feature SoA [8][ns], target[ns]); grad[4801], *loss = mean.
int oracle_mlp_grad(const float* w, const float* mu, const float* sd, const float* feat,
                    const float* target, uint64_t ns, const uint32_t* idx, uint32_t nb,
                    double* grad, double* loss, int emulate_bf16) {
  Mlp m = make_mlp(w, mu, sd, emulate_bf16 != 0);
  for (int i = 0; i < kNP; ++i) grad[i] = 0;
  double L = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint64_t s = idx[b];
    double x[kIn], z1[kHid], h1[kHid], z2[kHid], h2[kHid];
    m.input(feat, ns, s, x);
    const double y = m.forward(x, z1, h1, z2, h2);
    const double e = y - target[s];
    L += e * e;
    const double dy = 2.0 * e / nb;
    double dz2[kHid], dz1[kHid];
    for (int k = 0; k < kHid; ++k) {
      grad[oW3 + k] += h2[k] * dy;
      const double sg = Mlp::sg(z2[k]);
      dz2[k] = dy * m.W3(k) * (sg * (1.0 + z2[k] * (1.0 - sg)));
      grad[oB2 + k] += dz2[k];
    }
    grad[oB3] += dy;
    for (int j = 0; j < kHid; ++j)
      for (int k = 0; k < kHid; ++k) grad[oW2 + j * kHid + k] += m.q(dz2[j]) * m.q(h1[k]);
    for (int k = 0; k < kHid; ++k) {
      double dh = 0;
      for (int j = 0; j < kHid; ++j) dh += m.q(dz2[j]) * m.W2(j, k);
      const double sg = Mlp::sg(z1[k]);
      dz1[k] = dh * (sg * (1.0 + z1[k] * (1.0 - sg)));
      grad[oB1 + k] += dz1[k];
    }
    for (int j = 0; j < kHid; ++j)
      for (int i = 0; i < kIn; ++i) grad[oW1 + j * kIn + i] += m.q(dz1[j]) * m.q(x[i]);
  }
  *loss = L / nb;
  return 0;
}

// Sorted-order grid/neighbour arrays of the last step (for kernel-level parity).
int oracle_system_arrays(void* hp, uint32_t* hash, uint32_t* perm, uint32_t* start,
                         uint32_t* count, uint32_t* list) {
  Handle* h = (Handle*)hp;
  auto fill = [&](auto* s) {
    if (hash) std::memcpy(hash, s->hash.data(), 4 * (size_t)s->n);
    if (perm) std::memcpy(perm, s->perm.data(), 4 * (size_t)s->n);
    if (start) std::memcpy(start, s->start.data(), 4 * (size_t)(s->g.ncells() + 1));
    if (count) std::memcpy(count, s->count.data(), 4 * (size_t)s->n);
    if (list) std::memcpy(list, s->list.data(), 4 * (size_t)s->n * s->prm.max_neighbors);
  };
  if (h->prec == 64) fill(h->d); else fill(h->f);
  return 0;
}

int oracle_surface_table(int* ntri, int* tri15) {
  const McTab t = mc_table();
  for (int c = 0; c < 256; ++c) {
    ntri[c] = t.ntri[c];
    for (int k = 0; k < 15; ++k) tri15[15 * c + k] = t.tri[c][k];
  }
  return 0;
}

void* oracle_surface_extract(const float* pos4, const uint32_t* ids, uint32_t n, uint32_t limit,
                             float h, const psb_surface_params* prm, uint32_t* nv, uint32_t* nt) {
  auto* o = new SurfaceOut;
  surface_extract(pos4, ids, n, limit, h, *prm, *o);
  *nv = (uint32_t)(o->verts.size() / 3);
  *nt = (uint32_t)(o->tris.size() / 3);
  return o;
}

int oracle_surface_get(void* hp, float* v3, float* n3, uint32_t* t3, float* field, uint32_t* dims,
                       float* origin) {
  auto* o = (SurfaceOut*)hp;
  if (v3) std::copy(o->verts.begin(), o->verts.end(), v3);
  if (n3) std::copy(o->norms.begin(), o->norms.end(), n3);
  if (t3) std::copy(o->tris.begin(), o->tris.end(), t3);
  if (field) std::copy(o->field.begin(), o->field.end(), field);
  if (dims) for (int a = 0; a < 3; ++a) dims[a] = o->dims[a];
  if (origin) for (int a = 0; a < 3; ++a) origin[a] = o->origin[a];
  return 0;
}

void oracle_surface_free(void* hp) { delete (SurfaceOut*)hp; }

}  // extern "C"
