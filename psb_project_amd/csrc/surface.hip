// surface.hip -- iso-surface extraction from particle positions.
//
// SurfaceSPH::extract (SurfaceSPH.cpp:29-190) lays a node grid over the
// particles' bounds plus a margin, evaluates the field of evaluateIso_Kernel
// (UniformGridKernel.cu:98-138) at every node, and polygonises it with
// marching cubes (MarchingCube_Kernel.cu:14-158).  Here, on the device:
//
//   k_bounds     min / max of the particle positions (order-independent)
//   k_iso        field at each node, f = sum_j (r^2 - d_j^2)^3 / r^6 over
//                the particles within r = h / scale, read through a
//                uniform grid of cell r (27 cells), in sorted order
//   k_edges      crossed grid edges (one end <= iso, the other not), three
//                per node (+x, +y, +z); exclusive scan -> vertex numbers
//   k_vertices   one vertex per crossed edge, vertexInterpolate_MarchingCube
//                (MarchingCube_Kernel.cu:145-155), and its normal: the
//                direction of -grad f
//   k_cubes      case index and triangle count per cube; exclusive scan
//   k_emit       triangles as vertex-index triples, cube by cube (x fastest)
//
// Vertices are shared through their grid edge, so the mesh is indexed
// without the host-side float-equality merge of SurfaceSPH.cpp:118-150.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mc_cases.h"
#include "psb_internal.h"

namespace psb {

__constant__ McCases c_mc;
// cube edge -> (node offset x, y, z, axis of the grid edge from that node)
// for the corner numbering of polygonize_MarchingCube_Kernel
__constant__ int8_t c_edge_node[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 2}, {0, 0, 1, 0}, {0, 0, 0, 2},
                                          {0, 1, 0, 0}, {1, 1, 0, 2}, {0, 1, 1, 0}, {0, 1, 0, 2},
                                          {0, 0, 0, 1}, {1, 0, 0, 1}, {1, 0, 1, 1}, {0, 0, 1, 1}};

// float -> unsigned key with the same order (for atomicMin / atomicMax)
__device__ __forceinline__ uint32_t order_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static float key_float(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ bool kept(const uint32_t* ids, uint32_t limit, uint32_t p) {
  return !ids || ids[p] < limit;
}

__global__ __launch_bounds__(256) void k_bounds_init(uint32_t* out) {
  if (threadIdx.x < 6) out[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ __launch_bounds__(256) void k_bounds(uint32_t n, const float4* __restrict__ pos,
                                                const uint32_t* __restrict__ ids, uint32_t limit,
                                                uint32_t* __restrict__ out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  if (p < n && kept(ids, limit, p)) {
    const float4 x = pos[p];
    const float c[3] = {x.x, x.y, x.z};
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = order_key(c[a]);
  }
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], o, 64));
      hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], o, 64));
    }
  }
  if ((threadIdx.x & 63u) == 0) {
    for (int a = 0; a < 3; ++a) {
      if (lo[a] != 0xFFFFFFFFu) atomicMin(out + a, lo[a]);
      if (hi[a] != 0u) atomicMax(out + 3 + a, hi[a]);
    }
  }
}

struct NodeGrid {
  int nx, ny, nz;
  float ox, oy, oz, cell;
  float iso, r2, inv_r6;
};

__device__ __forceinline__ void node_pos(const NodeGrid& G, int i, int j, int k, float& x, float& y,
                                         float& z) {
  x = G.ox + (float)i * G.cell;
  y = G.oy + (float)j * G.cell;
  z = G.oz + (float)k * G.cell;
}

// particles of the 3 x 3 x 3 cells around (x, y, z), rows of x-consecutive
// cells in ascending (z, y), each row in sorted order
template <class F>
__device__ __forceinline__ void for_near(const GridDev& g, const uint32_t* __restrict__ start,
                                         const uint32_t* __restrict__ perm, float x, float y,
                                         float z, F&& f) {
  uint32_t bad = 0;
  const int cx = cell_coord(x, g.minx, g.cell, g.nx, bad);
  const int cy = cell_coord(y, g.miny, g.cell, g.ny, bad);
  const int cz = cell_coord(z, g.minz, g.cell, g.nz, bad);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, g.nx - 1);
  for (int zz = max(cz - 1, 0); zz <= min(cz + 1, g.nz - 1); ++zz)
    for (int yy = max(cy - 1, 0); yy <= min(cy + 1, g.ny - 1); ++yy) {
      const uint64_t base = (uint64_t)g.nx * ((uint64_t)yy + (uint64_t)g.ny * (uint64_t)zz);
      const uint32_t q1 = start[base + x1 + 1];
      for (uint32_t q = start[base + x0]; q < q1; ++q) f(perm[q]);
    }
}

__global__ __launch_bounds__(256) void k_iso(NodeGrid G, GridDev g, const uint32_t* __restrict__ start,
                                             const uint32_t* __restrict__ perm,
                                             const float4* __restrict__ pos,
                                             const uint32_t* __restrict__ ids, uint32_t limit,
                                             float* __restrict__ iso) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nn = (uint64_t)G.nx * G.ny * G.nz;
  if (t >= nn) return;
  const int i = (int)(t % G.nx), j = (int)((t / G.nx) % G.ny), k = (int)(t / ((uint64_t)G.nx * G.ny));
  float x, y, z;
  node_pos(G, i, j, k, x, y, z);
  float acc = 0.f;
  for_near(g, start, perm, x, y, z, [&](uint32_t p) {
    if (!kept(ids, limit, p)) return;
    const float4 q = pos[p];
    const float dx = x - q.x, dy = y - q.y, dz = z - q.z;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    if (d2 <= G.r2) {
      const float w = G.r2 - d2;
      acc = __builtin_fmaf(w * w, w, acc);
    }
  });
  iso[t] = acc * G.inv_r6;
}

__global__ __launch_bounds__(256) void k_edges(NodeGrid G, const float* __restrict__ iso,
                                               uint8_t* __restrict__ eflags,
                                               uint32_t* __restrict__ vcount) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nn = (uint64_t)G.nx * G.ny * G.nz;
  if (t >= nn) return;
  const int i = (int)(t % G.nx), j = (int)((t / G.nx) % G.ny), k = (int)(t / ((uint64_t)G.nx * G.ny));
  const bool b0 = iso[t] <= G.iso;
  uint32_t f = 0;
  if (i + 1 < G.nx && (iso[t + 1] <= G.iso) != b0) f |= 1u;
  if (j + 1 < G.ny && (iso[t + G.nx] <= G.iso) != b0) f |= 2u;
  if (k + 1 < G.nz && (iso[t + (uint64_t)G.nx * G.ny] <= G.iso) != b0) f |= 4u;
  eflags[t] = (uint8_t)f;
  vcount[t] = __builtin_popcount(f);
}

__global__ __launch_bounds__(256) void k_vertices(NodeGrid G, GridDev g,
                                                  const uint32_t* __restrict__ start,
                                                  const uint32_t* __restrict__ perm,
                                                  const float4* __restrict__ pos,
                                                  const uint32_t* __restrict__ ids, uint32_t limit,
                                                  const float* __restrict__ iso,
                                                  const uint8_t* __restrict__ eflags,
                                                  const uint32_t* __restrict__ vbase,
                                                  float* __restrict__ verts,
                                                  float* __restrict__ norms) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nn = (uint64_t)G.nx * G.ny * G.nz;
  if (t >= nn) return;
  const uint32_t f = eflags[t];
  if (!f) return;
  const int i = (int)(t % G.nx), j = (int)((t / G.nx) % G.ny), k = (int)(t / ((uint64_t)G.nx * G.ny));
  float ax, ay, az;
  node_pos(G, i, j, k, ax, ay, az);
  const float va = iso[t];
  uint32_t v = vbase[t];
  for (int axis = 0; axis < 3; ++axis) {
    if (!(f >> axis & 1u)) continue;
    const uint64_t tb = t + (axis == 0 ? 1 : axis == 1 ? (uint64_t)G.nx : (uint64_t)G.nx * G.ny);
    float bx, by, bz;
    node_pos(G, i + (axis == 0), j + (axis == 1), k + (axis == 2), bx, by, bz);
    const float vb = iso[tb];
    float px, py, pz;
    // vertexInterpolate_MarchingCube (MarchingCube_Kernel.cu:145-155)
    if (fabsf(G.iso - va) < 0.000001f) { px = ax; py = ay; pz = az; }
    else if (fabsf(G.iso - vb) < 0.000001f) { px = bx; py = by; pz = bz; }
    else if (fabsf(va - vb) < 0.000001f) { px = ax; py = ay; pz = az; }
    else {
      const float mu = (G.iso - va) / (vb - va);
      px = ax + mu * (bx - ax); py = ay + mu * (by - ay); pz = az + mu * (bz - az);
    }
    // normal: -grad f ~ sum_j (x - x_j)(r^2 - d_j^2)^2, normalised
    float gx = 0.f, gy = 0.f, gz = 0.f;
    for_near(g, start, perm, px, py, pz, [&](uint32_t p) {
      if (!kept(ids, limit, p)) return;
      const float4 q = pos[p];
      const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
      const float d2 = dot3(dx, dy, dz, dx, dy, dz);
      if (d2 < G.r2) {
        const float w = G.r2 - d2;
        const float w2 = w * w;
        gx = __builtin_fmaf(dx, w2, gx);
        gy = __builtin_fmaf(dy, w2, gy);
        gz = __builtin_fmaf(dz, w2, gz);
      }
    });
    const float gl = sqrtf(dot3(gx, gy, gz, gx, gy, gz));
    if (gl > 0.f) { gx = gx / gl; gy = gy / gl; gz = gz / gl; }
    verts[3 * (size_t)v] = px; verts[3 * (size_t)v + 1] = py; verts[3 * (size_t)v + 2] = pz;
    norms[3 * (size_t)v] = gx; norms[3 * (size_t)v + 1] = gy; norms[3 * (size_t)v + 2] = gz;
    ++v;
  }
}

__device__ __forceinline__ uint32_t cube_case(const NodeGrid& G, const float* __restrict__ iso,
                                              uint64_t t) {
  const uint64_t sx = 1, sy = (uint64_t)G.nx, sz = (uint64_t)G.nx * G.ny;
  // corner order of polygonize_MarchingCube_Kernel (MarchingCube_Kernel.cu:45-53)
  const uint64_t c[8] = {t, t + sx, t + sx + sz, t + sz, t + sy, t + sx + sy, t + sx + sy + sz, t + sy + sz};
  uint32_t cs = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (iso[c[q]] <= G.iso) cs |= 1u << q;
  return cs;
}

__device__ __forceinline__ void cube_of(const NodeGrid& G, uint64_t u, int& i, int& j, int& k) {
  const uint64_t cx = G.nx - 1, cy = G.ny - 1;
  i = (int)(u % cx); j = (int)((u / cx) % cy); k = (int)(u / (cx * cy));
}

__global__ __launch_bounds__(256) void k_cubes(NodeGrid G, const float* __restrict__ iso,
                                               uint32_t* __restrict__ tcount) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nc = (uint64_t)(G.nx - 1) * (G.ny - 1) * (G.nz - 1);
  if (u >= nc) return;
  int i, j, k;
  cube_of(G, u, i, j, k);
  const uint64_t t = (uint64_t)i + (uint64_t)G.nx * ((uint64_t)j + (uint64_t)G.ny * k);
  tcount[u] = c_mc.ntri[cube_case(G, iso, t)];
}

__global__ __launch_bounds__(256) void k_emit(NodeGrid G, const float* __restrict__ iso,
                                              const uint8_t* __restrict__ eflags,
                                              const uint32_t* __restrict__ vbase,
                                              const uint32_t* __restrict__ tbase,
                                              uint32_t* __restrict__ tris) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nc = (uint64_t)(G.nx - 1) * (G.ny - 1) * (G.nz - 1);
  if (u >= nc) return;
  const uint32_t b = tbase[u], e = tbase[u + 1];
  if (b == e) return;
  int i, j, k;
  cube_of(G, u, i, j, k);
  const uint64_t t = (uint64_t)i + (uint64_t)G.nx * ((uint64_t)j + (uint64_t)G.ny * k);
  const uint32_t cs = cube_case(G, iso, t);
  for (uint32_t q = 0; q < 3 * (e - b); ++q) {
    const int ed = c_mc.tri[cs][q];
    const uint64_t nd = t + (uint64_t)c_edge_node[ed][0] +
                        (uint64_t)G.nx * ((uint64_t)c_edge_node[ed][1] + (uint64_t)G.ny * c_edge_node[ed][2]);
    const uint32_t below = eflags[nd] & ((1u << c_edge_node[ed][3]) - 1u);
    tris[3 * (size_t)b + q] = vbase[nd] + __builtin_popcount(below);
  }
}

psb_status scan_exclusive(uint32_t* count, uint64_t n, uint32_t* bsum, uint32_t* start,
                          hipStream_t s);

}  // namespace psb

using namespace psb;

struct psb_surface_s {
  psb_context ctx = nullptr;
  psb_grid grid = nullptr;
  psb_grid_desc gdesc{};
  uint32_t grid_cap = 0;
  NodeGrid G{};
  uint64_t nodes = 0, cubes = 0;
  // element capacities of the device arrays below
  size_t cap_iso = 0, cap_ef = 0, cap_vc = 0, cap_vb = 0, cap_tc = 0, cap_tb = 0, cap_bs = 0;
  size_t cap_v = 0, cap_n = 0, cap_t = 0;
  float* iso = nullptr;
  uint8_t* eflags = nullptr;
  uint32_t *vcount = nullptr, *vbase = nullptr, *tcount = nullptr, *tbase = nullptr;
  uint32_t* bsum = nullptr;
  uint32_t* bounds = nullptr;
  uint32_t* h_counts = nullptr;   // pinned
  float *verts = nullptr, *norms = nullptr;
  uint32_t* tris = nullptr;
  uint32_t nv = 0, nt = 0;
};

namespace {
std::vector<bool> g_table_loaded;   // per device

psb_status load_table(int dev) {
  if ((int)g_table_loaded.size() <= dev) g_table_loaded.resize(dev + 1, false);
  if (g_table_loaded[dev]) return PSB_OK;
  static const McCases tab = mc_generate();
  PSB_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_mc), &tab, sizeof(tab)));
  g_table_loaded[dev] = true;
  return PSB_OK;
}

template <class T> psb_status grow(T*& p, size_t& cap, size_t want) {
  if (want <= cap && p) return PSB_OK;
  if (p) PSB_HIP(hipFree(p));
  p = nullptr;
  const size_t c = std::max<size_t>(want, 1);
  PSB_HIP(hipMalloc((void**)&p, sizeof(T) * c));
  cap = c;
  return PSB_OK;
}
}  // namespace

extern "C" {

psb_status psb_surface_default_params(psb_surface_params* p) {
  PSB_ARG(p, "null argument");
  p->cell_size = 0.03f;   // SurfaceSPH.cpp:31
  p->scale = 0.65f;       // SurfaceSPH.cpp:32
  p->iso_level = 0.5f;
  p->margin = 2.0f;       // SurfaceSPH.cpp:43: 2 h / scale
  return PSB_OK;
}

psb_status psb_surface_case_table(uint8_t* triangle_counts, int8_t* triangle_edges) {
  McCases t;
  try {
    t = mc_generate();
  } catch (const std::exception& e) {
    set_error(e.what());
    return PSB_ERR_STATE;
  }
  if (triangle_counts) std::memcpy(triangle_counts, t.ntri, sizeof(t.ntri));
  if (triangle_edges)
    for (int c = 0; c < 256; ++c)
      for (int k = 0; k < 16; ++k) triangle_edges[16 * c + k] = k < 3 * kMcMaxTri ? t.tri[c][k] : -1;
  return PSB_OK;
}

psb_status psb_surface_create(psb_context ctx, psb_surface* out) {
  PSB_ARG(ctx && out, "null argument");
  PSB_HIP(hipSetDevice(ctx->device));
  auto* s = new psb_surface_s;
  s->ctx = ctx;
  psb_status r = PSB_OK;
  if (hipMalloc((void**)&s->bounds, 64) != hipSuccess ||
      hipHostMalloc((void**)&s->h_counts, 64, hipHostMallocDefault) != hipSuccess) {
    set_error("surface allocation failed");
    r = PSB_ERR_OUT_OF_MEMORY;
  }
  if (r == PSB_OK) r = load_table(ctx->device);
  if (r != PSB_OK) { psb_surface_destroy(s); return r; }
  *out = s;
  return PSB_OK;
}

psb_status psb_surface_destroy(psb_surface s) {
  if (!s) return PSB_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->grid) psb_finalize_grid(s->grid);
  for (void* p : {(void*)s->iso, (void*)s->eflags, (void*)s->vcount, (void*)s->vbase,
                  (void*)s->tcount, (void*)s->tbase, (void*)s->bsum, (void*)s->bounds,
                  (void*)s->verts, (void*)s->norms, (void*)s->tris})
    if (p) (void)hipFree(p);
  if (s->h_counts) (void)hipHostFree(s->h_counts);
  delete s;
  return PSB_OK;
}

psb_status psb_surface_extract_points(psb_surface s, const float* pos4, const uint32_t* ids,
                                      uint32_t n, uint32_t id_limit, float support_radius,
                                      const psb_surface_params* prm, uint32_t* num_vertices,
                                      uint32_t* num_triangles) {
  PSB_ARG(s && prm && (n == 0 || pos4), "null argument");
  PSB_ARG(std::isfinite(support_radius) && support_radius > 0.f, "support radius must be > 0");
  PSB_ARG(std::isfinite(prm->scale) && prm->scale > 0.f && std::isfinite(prm->cell_size) &&
              prm->cell_size > 0.f && std::isfinite(prm->margin) && prm->margin >= 0.f &&
              std::isfinite(prm->iso_level),
          "surface parameters: scale, cell_size > 0, margin >= 0, finite iso_level");
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const float4* P = reinterpret_cast<const float4*>(pos4);
  s->nv = s->nt = 0;
  s->nodes = s->cubes = 0;
  if (num_vertices) *num_vertices = 0;
  if (num_triangles) *num_triangles = 0;
  if (!n) return PSB_OK;
  // bounds of the kept particles
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, s->bounds);
  hipLaunchKernelGGL(k_bounds, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, P, ids, id_limit,
                     s->bounds);
  PSB_HIP(hipGetLastError());
  PSB_HIP(hipMemcpyAsync(s->h_counts, s->bounds, 24, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  if (s->h_counts[0] == 0xFFFFFFFFu) return PSB_OK;   // no particle kept
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { lo[a] = key_float(s->h_counts[a]); hi[a] = key_float(s->h_counts[3 + a]); }
  for (int a = 0; a < 3; ++a)
    PSB_ARG(std::isfinite(lo[a]) && std::isfinite(hi[a]), "non-finite particle positions");
  // node grid as SurfaceSPH::extract lays it (SurfaceSPH.cpp:38-60), in double
  const float r = support_radius / prm->scale;
  const double margin = (double)prm->margin * (double)r;
  double mn[3], mx[3];
  uint32_t dims[3];
  float org[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = (double)lo[a] - margin;
    mx[a] = (double)hi[a] + margin;
    const double len = std::fabs(mx[a] - mn[a]);
    const double nd = std::ceil(len / (double)prm->cell_size);
    PSB_ARG(nd >= 2 && nd <= 4096, "surface node grid must have 2..4096 nodes per axis");
    dims[a] = (uint32_t)nd;
    const double center = (mn[a] + mx[a]) / 2;
    org[a] = (float)(center - len / 2 + (double)prm->cell_size / 2);
  }
  NodeGrid& G = s->G;
  G.nx = (int)dims[0]; G.ny = (int)dims[1]; G.nz = (int)dims[2];
  G.ox = org[0]; G.oy = org[1]; G.oz = org[2];
  G.cell = prm->cell_size;
  G.iso = prm->iso_level;
  G.r2 = r * r;
  G.inv_r6 = 1.0f / (G.r2 * G.r2 * G.r2);
  s->nodes = (uint64_t)dims[0] * dims[1] * dims[2];
  s->cubes = (uint64_t)(dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1);
  // particle grid of cell r over the margin-extended bounds
  psb_grid_desc gd{};
  for (int a = 0; a < 3; ++a) {
    gd.min[a] = (float)mn[a];
    gd.dims[a] = (int32_t)std::ceil((mx[a] - mn[a]) / (double)r) + 1;
  }
  gd.cell_size = r;
  PSB_ARG((uint64_t)gd.dims[0] * gd.dims[1] * gd.dims[2] <= (1ull << 31),
          "particle grid too fine for the surface radius");
  if (!s->grid || std::memcmp(&gd, &s->gdesc, sizeof gd) != 0 || n > s->grid_cap) {
    if (s->grid) { psb_finalize_grid(s->grid); s->grid = nullptr; }
    PSB_TRY(psb_create_grid_desc(s->ctx, &gd, n, &s->grid));
    s->gdesc = gd;
    s->grid_cap = n;
  }
  PSB_TRY(grid_build(s->grid, P, n, nullptr, st));
  const GridDev g = grid_dev(gd, r);
  PSB_TRY(grow(s->iso, s->cap_iso, s->nodes));
  PSB_TRY(grow(s->eflags, s->cap_ef, s->nodes));
  PSB_TRY(grow(s->vcount, s->cap_vc, s->nodes));
  PSB_TRY(grow(s->vbase, s->cap_vb, s->nodes + 1));
  PSB_TRY(grow(s->tcount, s->cap_tc, s->cubes));
  PSB_TRY(grow(s->tbase, s->cap_tb, s->cubes + 1));
  const uint64_t scan_tiles = (std::max(s->nodes, s->cubes) + kScanTile - 1) / kScanTile;
  PSB_TRY(grow(s->bsum, s->cap_bs, scan_tiles + 1));
  const uint32_t nbn = blocks_for(s->nodes), nbc = blocks_for(s->cubes);
  hipLaunchKernelGGL(k_iso, dim3(nbn), dim3(kBlock), 0, st, G, g, s->grid->start, s->grid->perm, P,
                     ids, id_limit, s->iso);
  hipLaunchKernelGGL(k_edges, dim3(nbn), dim3(kBlock), 0, st, G, s->iso, s->eflags, s->vcount);
  PSB_HIP(hipGetLastError());
  PSB_TRY(scan_exclusive(s->vcount, s->nodes, s->bsum, s->vbase, st));
  if (s->cubes) {
    hipLaunchKernelGGL(k_cubes, dim3(nbc), dim3(kBlock), 0, st, G, s->iso, s->tcount);
    PSB_HIP(hipGetLastError());
  }
  PSB_TRY(scan_exclusive(s->tcount, s->cubes, s->bsum, s->tbase, st));
  PSB_HIP(hipMemcpyAsync(s->h_counts, s->vbase + s->nodes, 4, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipMemcpyAsync(s->h_counts + 1, s->tbase + s->cubes, 4, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  s->nv = s->h_counts[0];
  s->nt = s->h_counts[1];
  PSB_TRY(grow(s->verts, s->cap_v, 3 * (size_t)s->nv));
  PSB_TRY(grow(s->norms, s->cap_n, 3 * (size_t)s->nv));
  PSB_TRY(grow(s->tris, s->cap_t, 3 * (size_t)s->nt));
  hipLaunchKernelGGL(k_vertices, dim3(nbn), dim3(kBlock), 0, st, G, g, s->grid->start,
                     s->grid->perm, P, ids, id_limit, s->iso, s->eflags, s->vbase, s->verts,
                     s->norms);
  if (s->cubes)
    hipLaunchKernelGGL(k_emit, dim3(nbc), dim3(kBlock), 0, st, G, s->iso, s->eflags, s->vbase,
                       s->tbase, s->tris);
  PSB_HIP(hipGetLastError());
  PSB_HIP(hipStreamSynchronize(st));
  if (num_vertices) *num_vertices = s->nv;
  if (num_triangles) *num_triangles = s->nt;
  return PSB_OK;
}

psb_status psb_surface_get_mesh(psb_surface s, float* vertices3, float* normals3,
                                uint32_t* triangles3) {
  PSB_ARG(s, "null surface");
  PSB_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  if (vertices3 && s->nv)
    PSB_HIP(hipMemcpyAsync(vertices3, s->verts, 12 * (size_t)s->nv, hipMemcpyDeviceToHost, st));
  if (normals3 && s->nv)
    PSB_HIP(hipMemcpyAsync(normals3, s->norms, 12 * (size_t)s->nv, hipMemcpyDeviceToHost, st));
  if (triangles3 && s->nt)
    PSB_HIP(hipMemcpyAsync(triangles3, s->tris, 12 * (size_t)s->nt, hipMemcpyDeviceToHost, st));
  PSB_HIP(hipStreamSynchronize(st));
  return PSB_OK;
}

psb_status psb_surface_get_field(psb_surface s, float* field, uint32_t dims[3], float origin[3],
                                 float* cell) {
  PSB_ARG(s, "null surface");
  if (dims) { dims[0] = s->nodes ? s->G.nx : 0; dims[1] = s->nodes ? s->G.ny : 0; dims[2] = s->nodes ? s->G.nz : 0; }
  if (origin) { origin[0] = s->G.ox; origin[1] = s->G.oy; origin[2] = s->G.oz; }
  if (cell) *cell = s->G.cell;
  if (field && s->nodes) {
    PSB_HIP(hipSetDevice(s->ctx->device));
    PSB_HIP(hipMemcpyAsync(field, s->iso, 4 * s->nodes, hipMemcpyDeviceToHost, s->ctx->stream));
    PSB_HIP(hipStreamSynchronize(s->ctx->stream));
  }
  return PSB_OK;
}

}  // extern "C"
