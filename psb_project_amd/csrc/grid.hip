// grid.hip -- uniform grid: cell hash, stable counting sort, cell ranges,
// reorder and neighbour search for gfx950.
//
// Replaces createGrid / reinitCells / storeParticles / searchNeighbooring
// (UniformGrid.cu:17-99, UniformGridKernel.cu:12-95).  The reference inserts
// particles into 64-slot cell buckets with atomicAdd (order depends on the
// race, no overflow check) and writes a 200-wide neighbour list with a global
// read-modify-write per hit.  Here:
//
//   1. k_hash_count   hash[i] = cell of x_i;  slot[i] = atomicAdd(count[c])
//   2. k_scan_*       start = exclusive prefix sum of count over all cells
//                     (start[c] = #particles with hash < c, start[ncells] = n);
//                     the per-tile sums come from step 1 (tsum), so the scan
//                     reads count once; k_rank_reorder re-zeroes the occupied
//                     cells' counts for the next build
//   3. k_scatter      tmp[start[c] + slot[i]] = i      (order within a cell
//                     still depends on the atomics)
//   4. k_rank_reorder rank of i among its cell's tmp entries by index ->
//                     dst = start[c] + rank: the stable (hash, index) order;
//                     perm[dst] = i and the particle state is gathered here
//
// The dense lower-bound table start[] makes every stencil row of cells one
// contiguous particle range [start[c0], start[c1 + 1]) because the hash is
// x-fastest, so the neighbour kernel does two table loads per row.
#include "psb_internal.h"
#include "pcisph.h"

namespace psb {

constexpr int kTileSlots = 16;   // copies of the per-tile count table (atomic spreading)

// M: SingleMap or BatchMap (a batch hashes into its scene's cell range).
template <class M>
__global__ __launch_bounds__(M::kMaxThreads) void k_hash_count(M map, const float4* __restrict__ pos,
                                                               uint32_t n, GridDev g0,
                                                               uint32_t* __restrict__ hash,
                                                               uint32_t* __restrict__ slot,
                                                               uint32_t* __restrict__ count,
                                                               uint32_t* __restrict__ flags,
                                                               uint32_t* __restrict__ tsum,
                                                               uint32_t nscan, StepReset rs) {
  if (rs.st && blockIdx.x < rs.scenes) pci_reset_scene(rs.st, blockIdx.x, rs.done, rs.es);
  uint32_t i, sc;
  bool lead;
  if (!map.locate(n, i, sc, lead)) return;
  uint32_t off;
  const GridDev g = map.grid(g0, sc, off);
  float4 p = pos[i];
  uint32_t bad = 0;
  int cx = cell_coord(p.x, g.minx, g.cell, g.nx, bad);
  int cy = cell_coord(p.y, g.miny, g.cell, g.ny, bad);
  int cz = cell_coord(p.z, g.minz, g.cell, g.nz, bad);
  uint32_t c = off + (uint32_t)cx + (uint32_t)g.nx * ((uint32_t)cy + (uint32_t)g.ny * (uint32_t)cz);
  hash[i] = c;
  slot[i] = atomicAdd(&count[c], 1u);
  // per-scan-tile particle counts (the scan's first pass), one atomic per
  // distinct tile of the wave (the particles arrive in last step's sorted
  // order, so a wave's cells mostly share a tile), spread over kTileSlots
  // copies of the table by workgroup
  const uint32_t t = c / kScanTile;
  uint32_t* __restrict__ ts = tsum + (size_t)(blockIdx.x % kTileSlots) * nscan;
  uint64_t todo = __ballot(1);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const uint32_t tl = (uint32_t)__shfl((int)t, leader, 64);
    const uint64_t same = __ballot(t == tl) & todo;
    if ((int)(threadIdx.x & 63u) == leader) atomicAdd(&ts[tl], (uint32_t)__popcll(same));
    todo &= ~same;
  }
  uint32_t f = 0;
  if (bad) f |= PSB_FLAG_OUT_OF_GRID;
  if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) f |= PSB_FLAG_NONFINITE;
  if (f) atomicOr(flags, f);
}

// exclusive scan of one value per thread over a 256-thread block
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    uint32_t t = lds[w];
    if (w < wid) woff += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return woff + x - v;
}

// per-tile sums of count
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ count,
                                                     uint64_t ncells, uint32_t* __restrict__ bsum) {
  __shared__ uint32_t lds[4];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  if (base + kScanTile <= ncells) {
    const uint4* c4 = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint4 v = c4[threadIdx.x + 256 * j];
      s += v.x + v.y + v.z + v.w;
    }
  } else {
    for (uint64_t k = base + threadIdx.x; k < ncells; k += 256) s += count[k];
  }
  uint32_t tot;
  block_excl_scan(s, lds, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// exclusive scan of the tile sums in one workgroup: dst[k] = sum of src[<k]
// (dst may be src).  CONSUME: src holds kTileSlots tables of nb partial sums
// (added here) and is zeroed for the next build.  start[ncells] = total.
template <bool CONSUME>
__global__ __launch_bounds__(256) void k_scan_tiles(uint32_t* src, uint32_t* dst, uint32_t nb,
                                                    uint32_t* __restrict__ start, uint64_t ncells) {
  __shared__ uint32_t lds[4];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
    uint32_t k = b0 + threadIdx.x;
    uint32_t v = 0u;
    if (k < nb) {
      if (CONSUME) {
#pragma unroll
        for (int j = 0; j < kTileSlots; ++j) {
          v += src[(size_t)j * nb + k];
          src[(size_t)j * nb + k] = 0u;
        }
      } else {
        v = src[k];
      }
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, lds, tot);
    if (k < nb) dst[k] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) start[ncells] = carry;
}

// start[c] = tile offset + in-tile exclusive prefix; ZERO: count[c] = 0
template <bool ZERO>
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* __restrict__ count, uint64_t ncells,
                                                    const uint32_t* __restrict__ bsum,
                                                    uint32_t* __restrict__ start) {
  __shared__ uint32_t lds[4];
  const uint32_t toff = bsum[blockIdx.x];   // loaded with the counts, not after the block scan
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * 16;
  uint32_t v[16];
  const bool full = (uint64_t)(blockIdx.x + 1) * kScanTile <= ncells;
  if (full) {
    const uint4* c4 = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint4 t = c4[j];
      v[4 * j] = t.x; v[4 * j + 1] = t.y; v[4 * j + 2] = t.z; v[4 * j + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = (base + j < ncells) ? count[base + j] : 0u;
  }
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) { uint32_t t = v[j]; v[j] = s; s += t; }
  uint32_t tot;
  uint32_t off = block_excl_scan(s, lds, tot) + toff;
  if (full) {
    uint4* s4 = reinterpret_cast<uint4*>(start + base);
    uint4* c4 = reinterpret_cast<uint4*>(count + base);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      s4[j] = make_uint4(v[4 * j] + off, v[4 * j + 1] + off, v[4 * j + 2] + off, v[4 * j + 3] + off);
      if (ZERO) c4[j] = make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (base + j < ncells) {
        start[base + j] = v[j] + off;
        if (ZERO) count[base + j] = 0u;
      }
  }
}

__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ hash,
                                                 const uint32_t* __restrict__ slot, uint32_t n,
                                                 const uint32_t* __restrict__ start,
                                                 uint32_t* __restrict__ tmp) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tmp[start[hash[i]] + slot[i]] = i;
}

// Stable rank within the cell + gather of the particle state.
// The cell's first particle re-zeroes its count for the next build.
template <bool STATE>
__global__ __launch_bounds__(256) void k_rank_reorder(const uint32_t* __restrict__ hash, uint32_t n,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ tmp,
                                                      uint32_t* __restrict__ perm, ReorderArrays ra,
                                                      uint32_t* __restrict__ count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = hash[i];
  const uint32_t s = start[c], e = start[c + 1];
  uint32_t rank = 0;
  if (e - s > 1)
    for (uint32_t q = s; q < e; ++q) rank += (tmp[q] < i) ? 1u : 0u;
  const uint32_t dst = s + rank;
  perm[dst] = i;
  if (rank == 0) count[c] = 0u;
  if (STATE) {
    ra.dx[dst] = ra.sx[i];
    ra.dv[dst] = ra.sv[i];
    ra.dvh[dst] = ra.svh[i];
    ra.did[dst] = ra.sid[i];
    ra.dp[dst] = ra.sp[i];
  }
}

// searchNeighbooring (UniformGridKernel.cu:58-95) on sorted positions; the
// density sum of the same pairs rides along (SphKernel.cu:9-27).
// RT > 0: the stencil half-width is a compile-time constant; the (2RT+1)^2
// row ranges are loaded first (independent table loads) and the candidates of
// a row are gathered kRowSlots at a time before being tested in order.  RT =
// 0: any stencil width, rows visited one after the other.  With ep.on the
// density epilogue (density_epilogue_one) runs on the particle's sum.
constexpr int kRowSlots = 8;   // candidate loads in flight per stencil row

template <bool DENS, int RT, class M>
__global__ __launch_bounds__(M::kMaxThreads) void k_neighbors(M map, const float4* xs, uint32_t n,
                                                              GridDev g0,
                                                              const uint32_t* __restrict__ start0,
                                                              float r2, uint32_t cap,
                                                              uint32_t* __restrict__ ncount,
                                                              uint32_t* __restrict__ nlist,
                                                              float* __restrict__ rho, float mc,
                                                              uint32_t* __restrict__ flags,
                                                              uint32_t* __restrict__ maxcnt,
                                                              Compact16 c16, ErrSlots es,
                                                              const HaloTile* __restrict__ tiles,
                                                              EpiArgs ep) {
  uint32_t p, sc;
  bool lead;
  const bool valid = map.locate(n, p, sc, lead);
  uint32_t cnt = 0;
  if (valid) {
    uint32_t off;
    const GridDev g = map.grid(g0, sc, off);
    const uint32_t* __restrict__ start = start0 + off;
    const float4 xi = xs[p];
    uint32_t bad = 0;
    const int cx = cell_coord(xi.x, g.minx, g.cell, g.nx, bad);
    const int cy = cell_coord(xi.y, g.miny, g.cell, g.ny, bad);
    const int cz = cell_coord(xi.z, g.minz, g.cell, g.nz, bad);
    const int R = RT > 0 ? RT : g.R;
    const int x0 = max(cx - R, 0), x1 = min(cx + R, g.nx - 1);
    float S = 0.f;
    bool wide = false;
    // grouped 32-bit list of a system whose compact offsets have overflowed
    // (then in place of the compact list, which its readers no longer take)
    uint32_t* const l32g =
        (!tiles && c16.list32g && c16.want32 && *c16.want32) ? c16.list32g : nullptr;
    // staged tiles: the compact entry is the LDS slot q + sb of the row
    const HaloTile* T = tiles ? tiles + p / kHaloTile : nullptr;
    const bool hok = T && T->ok;
    int32_t sb = 0;
    auto visit = [&](uint32_t q, const float4& xj) {
      const float dx = xi.x - xj.x, dy_ = xi.y - xj.y, dz_ = xi.z - xj.z;
      const float d2 = dot3(dx, dy_, dz_, dx, dy_, dz_);
      const bool acc = d2 <= r2;
      if (DENS) {
        // a rejected candidate adds +0, which leaves S (never -0) unchanged
        const float w = acc ? r2 - d2 : 0.f;
        S = __builtin_fmaf(w * w, w, S);
      }
      if (acc && cnt < cap) {
        nlist[(uint64_t)cnt * n + p] = q;
        if (l32g) l32g[l32g_at(cnt, p, n)] = q;
        else if (c16.list16) {
          const int d = (int)(q - p);
          const bool fits = d >= -kCompactReach && d <= kCompactReach;
          const uint16_t e = hok ? (uint16_t)((int32_t)q + sb) : (uint16_t)(d + 32768);
          if (hok || fits) c16.list16[l16_at(cnt, p, n)] = e;
          wide = wide || (!hok && !fits);
        }
      }
      cnt += acc ? 1u : 0u;
    };
    // a row's candidates 8 at a time: one round of loads covers a typical
    // row (slots past the row's end load nothing and are not visited)
    auto row = [&](uint32_t q0, uint32_t q1) {
      for (uint32_t q = q0; q < q1; q += kRowSlots) {
        float4 xq[kRowSlots];
#pragma unroll
        for (int u = 0; u < kRowSlots; ++u)
          if (q + u < q1) xq[u] = xs[q + u];
#pragma unroll
        for (int u = 0; u < kRowSlots; ++u)
          if (q + u < q1) visit(q + u, xq[u]);
      }
    };
    if (RT > 0) {
      constexpr int W = 2 * RT + 1;
      uint32_t q0s[W * W], q1s[W * W];
#pragma unroll
      for (int r = 0; r < W * W; ++r) {
        const int z = cz + r / W - RT, y = cy + r % W - RT;
        const bool ok = z >= 0 && z < g.nz && y >= 0 && y < g.ny;
        const uint64_t base = ok ? (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z) : 0;
        q0s[r] = ok ? start[base + x0] : 0u;
        q1s[r] = ok ? start[base + x1 + 1] : 0u;
      }
      for (int r = 0; r < W * W; ++r) {
        if (hok && q1s[r] > q0s[r]) sb = halo_slot_base(*T, cy + r % W - RT, cz + r / W - RT);
        row(q0s[r], q1s[r]);
      }
    } else {
      for (int dz = -R; dz <= R; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -R; dy <= R; ++dy) {
          const int y = cy + dy;
          if (y < 0 || y >= g.ny) continue;
          const uint64_t base = (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
          if (hok) sb = halo_slot_base(*T, y, z);
          row(start[base + x0], start[base + x1 + 1]);
        }
      }
    }
    ncount[p] = cnt < cap ? cnt : cap;
    if (DENS) {
      if (ep.on) {
        // the density epilogue, fused (x.w written here is not read by the
        // search: it tests positions only)
        Consts cc = ep.c;
        map.patch(cc, sc);
        density_epilogue_one(cc, ep, p, S * mc, rho);
      } else {
        rho[p] = S * mc;
      }
    }
    if (cnt > cap) atomicOr(flags, PSB_FLAG_NEIGHBOR_OVERFLOW);
    if (wide) atomicOr(c16.wide + (size_t)sc * c16.wide_stride, 1u);
  }
  if (maxcnt || es.base) {
    // workgroup maximum, then one atomic per workgroup on a spread slot
    __shared__ uint32_t red[M::kMaxThreads / 64];
    const uint32_t m = wave_max_u(cnt);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t b = 0;
      for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) b = max(b, red[w]);
      if (b) {
        uint32_t* dst = es.base ? es.base + (size_t)(blockIdx.x % kErrSlots) * es.stride +
                                      (size_t)map.scene_of_block() * kMaxIterSlots + kCountEntry
                                : maxcnt;
        atomicMax(dst, b);
      }
    }
  }
}

template <int W>
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ src, float* __restrict__ dst,
                                                const uint32_t* __restrict__ perm, uint32_t n) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (W == 4) reinterpret_cast<float4*>(dst)[k] = reinterpret_cast<const float4*>(src)[perm[k]];
  else dst[k] = src[perm[k]];
}

// counting sort of the hashed particles + cell ranges + reorder
static psb_status grid_sort(psb_grid g, uint32_t n, const ReorderArrays* ra, hipStream_t s) {
  // the tile sums came with the hashes (tsum); the scan's second pass
  // leaves count alone, the reorder re-zeroes the occupied cells
  const uint32_t nb = blocks_for(n);
  hipLaunchKernelGGL(k_scan_tiles<true>, dim3(1), dim3(kBlock), 0, s, g->tsum, g->bsum, g->nscan,
                     g->start, g->ncells);
  hipLaunchKernelGGL(k_scan_apply<false>, dim3(g->nscan), dim3(kBlock), 0, s, g->count, g->ncells,
                     g->bsum, g->start);
  if (n) {
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, s, g->hash, g->slot, n, g->start, g->tmp);
    if (ra) {
      hipLaunchKernelGGL(k_rank_reorder<true>, dim3(nb), dim3(kBlock), 0, s, g->hash, n, g->start,
                         g->tmp, g->perm, *ra, g->count);
    } else {
      ReorderArrays none{};
      hipLaunchKernelGGL(k_rank_reorder<false>, dim3(nb), dim3(kBlock), 0, s, g->hash, n, g->start,
                         g->tmp, g->perm, none, g->count);
    }
  }
  PSB_HIP(hipGetLastError());
  g->n = n;
  return PSB_OK;
}

// One workgroup per tile of kHaloTile sorted particles: the tile's halo rows
// (HaloTile) and, per LDS slot, the sorted index staged there.
__global__ __launch_bounds__(kHaloTile) void k_halo(const float4* __restrict__ xs, uint32_t n,
                                                    GridDev g, const uint32_t* __restrict__ start,
                                                    HaloTile* __restrict__ tiles,
                                                    uint32_t* __restrict__ src) {
  __shared__ int s_x0, s_x1, s_z0, s_z1, s_ok, s_nrows, s_total;
  __shared__ int s_ylo[kHaloMaxZ], s_yhi[kHaloMaxZ], s_rowbase[kHaloMaxZ + 1];
  __shared__ uint32_t s_gs[kHaloMaxRows], s_lb[kHaloMaxRows + 1];
  const uint32_t t = blockIdx.x;
  const uint32_t p = t * kHaloTile + threadIdx.x;
  const int R = g.R;
  if (threadIdx.x == 0) { s_x0 = 0x7FFFFFFF; s_x1 = -1; s_z0 = 0x7FFFFFFF; s_z1 = -1; }
  if (threadIdx.x < kHaloMaxZ) { s_ylo[threadIdx.x] = 0x7FFFFFFF; s_yhi[threadIdx.x] = -1; }
  if (threadIdx.x < kHaloMaxZ + 1) s_rowbase[threadIdx.x] = 0;
  __syncthreads();
  int cx = 0, cy = 0, cz = 0;
  const bool v = p < n;
  if (v) {
    uint32_t bad = 0;
    const float4 x = xs[p];
    cx = cell_coord(x.x, g.minx, g.cell, g.nx, bad);
    cy = cell_coord(x.y, g.miny, g.cell, g.ny, bad);
    cz = cell_coord(x.z, g.minz, g.cell, g.nz, bad);
    atomicMin(&s_x0, cx); atomicMax(&s_x1, cx);
    atomicMin(&s_z0, cz); atomicMax(&s_z1, cz);
  }
  __syncthreads();
  const int zlo = max(s_z0 - R, 0), zhi = min(s_z1 + R, g.nz - 1);
  const bool zok = s_z1 >= 0 && zhi - zlo + 1 <= kHaloMaxZ;
  if (v && zok)
    for (int z = max(cz - R, zlo); z <= min(cz + R, zhi); ++z) {
      atomicMin(&s_ylo[z - zlo], max(cy - R, 0));
      atomicMax(&s_yhi[z - zlo], min(cy + R, g.ny - 1));
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = zok ? 1 : 0, rows = 0;
    const int nzr = zok ? zhi - zlo + 1 : 0;
    for (int i = 0; i < nzr; ++i) {
      s_rowbase[i] = rows;
      if (s_yhi[i] >= s_ylo[i]) rows += s_yhi[i] - s_ylo[i] + 1;
      else s_ylo[i] = 0;   // no row in this layer
    }
    s_rowbase[nzr] = rows;
    if (rows > kHaloMaxRows) ok = 0;
    s_ok = ok;
    s_nrows = ok ? rows : 0;
  }
  __syncthreads();
  const int x0 = max(s_x0 - R, 0), x1 = min(s_x1 + R, g.nx - 1);
  const int nrows = s_nrows;
  for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
    int i = 0;
    while (s_rowbase[i + 1] <= r) ++i;
    const int y = s_ylo[i] + (r - s_rowbase[i]), z = zlo + i;
    const uint64_t base = (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
    const uint32_t gs = start[base + x0];
    s_gs[r] = gs;
    s_lb[r] = start[base + x1 + 1] - gs;   // count, scanned below
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int r = 0; r < nrows; ++r) { const uint32_t c = s_lb[r]; s_lb[r] = tot; tot += c; }
    s_lb[nrows] = tot;
    if (tot > (uint32_t)kHaloCap) s_ok = 0;
    s_total = (int)tot;
  }
  __syncthreads();
  HaloTile& T = tiles[t];
  const int ok = s_ok;
  if (threadIdx.x == 0) {
    T.ok = ok; T.nrows = nrows; T.total = ok ? s_total : 0;
    T.xlo = x0; T.xhi = x1; T.zlo = zlo; T.nz = zok ? zhi - zlo + 1 : 0; T.pad = 0;
  }
  if (threadIdx.x < kHaloMaxZ) T.ylo[threadIdx.x] = s_ylo[threadIdx.x];
  if (threadIdx.x < kHaloMaxZ + 1) T.rowbase[threadIdx.x] = s_rowbase[threadIdx.x];
  for (int r = threadIdx.x; r < nrows; r += blockDim.x) T.gstart[r] = s_gs[r];
  for (int r = threadIdx.x; r <= nrows; r += blockDim.x) T.lbase[r] = s_lb[r];
  if (!ok) return;
  uint32_t* sr = src + (size_t)t * kHaloCap;
  for (int r = 0; r < nrows; ++r) {
    const uint32_t lb = s_lb[r], c = s_lb[r + 1] - lb, gs = s_gs[r];
    for (uint32_t k = threadIdx.x; k < c; k += blockDim.x) sr[lb + k] = gs + k;
  }
}

psb_status launch_halo(psb_grid g, const float4* xs, uint32_t n, int R, HaloTile* tiles,
                       uint32_t* src, hipStream_t s) {
  if (!n) return PSB_OK;
  GridDev gd = grid_dev(g->desc, g->desc.cell_size);
  gd.R = R;
  const uint32_t nt = (n + kHaloTile - 1) / kHaloTile;
  hipLaunchKernelGGL(k_halo, dim3(nt), dim3(kHaloTile), 0, s, xs, n, gd, g->start, tiles, src);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

// start[0..n] = exclusive prefix sums of count[0..n) (start[n] = total);
// count is zeroed; bsum holds one entry per kScanTile counts
psb_status scan_exclusive(uint32_t* count, uint64_t n, uint32_t* bsum, uint32_t* start,
                          hipStream_t s) {
  const uint32_t nt = (uint32_t)((n + kScanTile - 1) / kScanTile);
  if (nt) hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, s, count, n, bsum);
  hipLaunchKernelGGL(k_scan_tiles<false>, dim3(1), dim3(kBlock), 0, s, bsum, bsum, nt, start, n);
  if (nt) hipLaunchKernelGGL(k_scan_apply<true>, dim3(nt), dim3(kBlock), 0, s, count, n, bsum, start);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status grid_build(psb_grid g, const float4* pos, uint32_t n, const ReorderArrays* ra,
                      hipStream_t s, const StepReset* reset) {
  GridDev gd = grid_dev(g->desc, g->desc.cell_size);
  StepReset rs{};
  if (reset) {
    if (n && blocks_for(n) >= reset->scenes) rs = *reset;   // rides on the hashing pass
    else PSB_TRY(launch_pci_reset(reset->st, false, s, reset->scenes, reset->done, reset->es));
  }
  if (n) {
    hipLaunchKernelGGL(k_hash_count<SingleMap>, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                       SingleMap{}, pos, n, gd, g->hash, g->slot, g->count, g->flags, g->tsum,
                       g->nscan, rs);
  }
  return grid_sort(g, n, ra, s);
}

psb_status grid_build_batch(psb_grid g, const BatchMap& map, uint32_t nblocks, uint32_t bsize,
                            const float4* pos, uint32_t n, const ReorderArrays* ra,
                            hipStream_t s, const StepReset* reset) {
  StepReset rs{};
  if (reset) {
    if (n && nblocks >= reset->scenes) rs = *reset;   // rides on the hashing pass
    else PSB_TRY(launch_pci_reset(reset->st, false, s, reset->scenes, reset->done, reset->es));
  }
  if (n) {
    GridDev none{};
    hipLaunchKernelGGL(k_hash_count<BatchMap>, dim3(nblocks), dim3(bsize), 0, s, map, pos, n,
                       none, g->hash, g->slot, g->count, g->flags, g->tsum, g->nscan, rs);
  }
  return grid_sort(g, n, ra, s);
}

psb_status launch_neighbors_batch(psb_grid g, const BatchMap& map, uint32_t nblocks,
                                  uint32_t bsize, const float4* xs, uint32_t n, float r2, int R,
                                  psb_neighbors nb, float* rho, float mc, uint32_t* flags,
                                  uint32_t* maxcnt, hipStream_t s, const Compact16& c16,
                                  ErrSlots es, const EpiArgs* ep) {
  if (!n) return PSB_OK;
  GridDev none{};
  dim3 gb(nblocks), tb(bsize);
  EpiArgs e{};
  if (ep && rho) e = *ep;
#define PSB_NB(D_, R_)                                                                       \
  hipLaunchKernelGGL((k_neighbors<D_, R_, BatchMap>), gb, tb, 0, s, map, xs, n, none, g->start, \
                     r2, nb.capacity, nb.count, nb.list, rho, mc, flags, maxcnt, c16, es, \
                     nullptr, e)
  if (rho) {
    if (R == 1) PSB_NB(true, 1); else if (R == 2) PSB_NB(true, 2); else PSB_NB(true, 0);
  } else {
    if (R == 1) PSB_NB(false, 1); else if (R == 2) PSB_NB(false, 2); else PSB_NB(false, 0);
  }
#undef PSB_NB
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status launch_neighbors(psb_grid g, const float4* xs, uint32_t n, float r2, int R,
                            psb_neighbors nb, float* rho, float mc, uint32_t* flags,
                            uint32_t* maxcnt, hipStream_t s, const Compact16& c16,
                            ErrSlots es, const HaloTile* tiles, const EpiArgs* ep) {
  if (!n) return PSB_OK;
  GridDev gd = grid_dev(g->desc, g->desc.cell_size);
  gd.R = R;
  dim3 gb(blocks_for(n)), tb(kBlock);
  EpiArgs e{};
  if (ep && rho) e = *ep;
#define PSB_NB(D_, R_)                                                                       \
  hipLaunchKernelGGL((k_neighbors<D_, R_, SingleMap>), gb, tb, 0, s, SingleMap{}, xs, n, gd,    \
                     g->start, r2, nb.capacity, nb.count, nb.list, rho, mc, flags, maxcnt, c16, es, \
                     tiles, e)
  if (rho) {
    if (R == 1) PSB_NB(true, 1); else if (R == 2) PSB_NB(true, 2); else PSB_NB(true, 0);
  } else {
    if (R == 1) PSB_NB(false, 1); else if (R == 2) PSB_NB(false, 2); else PSB_NB(false, 0);
  }
#undef PSB_NB
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

static psb_status grid_alloc(psb_grid g) {
  g->ncells = (uint64_t)g->desc.dims[0] * g->desc.dims[1] * g->desc.dims[2];
  g->nscan = (uint32_t)((g->ncells + kScanTile - 1) / kScanTile);
  const size_t cap = g->capacity ? g->capacity : 1;
  PSB_HIP(hipMalloc(&g->count, 4 * (g->nscan * (size_t)kScanTile)));
  PSB_HIP(hipMalloc(&g->start, 4 * (g->ncells + 1)));
  PSB_HIP(hipMalloc(&g->hash, 4 * cap));
  PSB_HIP(hipMalloc(&g->slot, 4 * cap));
  PSB_HIP(hipMalloc(&g->tmp, 4 * cap));
  PSB_HIP(hipMalloc(&g->perm, 4 * cap));
  PSB_HIP(hipMalloc(&g->bsum, 4 * (size_t)g->nscan));
  PSB_HIP(hipMalloc(&g->tsum, 4 * (size_t)kTileSlots * g->nscan));
  PSB_HIP(hipMemsetAsync(g->tsum, 0, 4 * (size_t)kTileSlots * g->nscan, g->ctx->stream));
  PSB_HIP(hipMalloc(&g->flags, 64));
  PSB_HIP(hipMemsetAsync(g->count, 0, 4 * (g->nscan * (size_t)kScanTile), g->ctx->stream));
  PSB_HIP(hipMemsetAsync(g->flags, 0, 64, g->ctx->stream));
  PSB_HIP(hipStreamSynchronize(g->ctx->stream));
  return PSB_OK;
}

psb_status grid_create(psb_context ctx, const psb_grid_desc& d, uint32_t capacity, psb_grid* out) {
  PSB_ARG(ctx && out, "null argument");
  PSB_ARG(d.cell_size > 0.f && d.dims[0] > 0 && d.dims[1] > 0 && d.dims[2] > 0, "bad grid");
  PSB_ARG((uint64_t)d.dims[0] * d.dims[1] * d.dims[2] < (1ull << 32) - 1, "grid too large");
  PSB_HIP(hipSetDevice(ctx->device));
  psb_grid g = new psb_grid_s();
  g->ctx = ctx;
  g->desc = d;
  g->capacity = capacity;
  psb_status st = grid_alloc(g);
  if (st != PSB_OK) { psb_finalize_grid(g); return st; }
  *out = g;
  return PSB_OK;
}

}  // namespace psb

using namespace psb;

extern "C" {

psb_status psb_create_grid(psb_context ctx, double xc, double yc, double zc, double length,
                           double width, double depth, float cell_size, uint32_t capacity,
                           psb_grid* out) {
  PSB_ARG(cell_size > 0.f && length > 0 && width > 0 && depth > 0, "bad grid extent");
  psb_grid_desc d{};
  // createGrid (UniformGrid.cu:19-29): Min = C - extent/2 (stored as float),
  // nbCells = ceil(extent / sizeCell)
  d.min[0] = (float)(xc - length / 2);
  d.min[1] = (float)(yc - width / 2);
  d.min[2] = (float)(zc - depth / 2);
  d.cell_size = cell_size;
  d.dims[0] = (int32_t)std::ceil(length / cell_size);
  d.dims[1] = (int32_t)std::ceil(width / cell_size);
  d.dims[2] = (int32_t)std::ceil(depth / cell_size);
  return grid_create(ctx, d, capacity, out);
}

psb_status psb_create_grid_desc(psb_context ctx, const psb_grid_desc* desc, uint32_t capacity,
                                psb_grid* out) {
  PSB_ARG(desc, "null desc");
  return grid_create(ctx, *desc, capacity, out);
}

psb_status psb_finalize_grid(psb_grid g) {
  if (!g) return PSB_OK;
  if (g->ctx) { hipSetDevice(g->ctx->device); hipStreamSynchronize(g->ctx->stream); }
  hipFree(g->count); hipFree(g->start); hipFree(g->hash); hipFree(g->slot);
  hipFree(g->tmp); hipFree(g->perm); hipFree(g->bsum); hipFree(g->tsum); hipFree(g->flags);
  delete g;
  return PSB_OK;
}

psb_status psb_grid_get_desc(psb_grid g, psb_grid_desc* out) {
  PSB_ARG(g && out, "null argument");
  *out = g->desc;
  return PSB_OK;
}

psb_status psb_store_particles(psb_grid g, const float* pos4, uint32_t n) {
  PSB_ARG(g && (pos4 || !n), "null argument");
  if (n > g->capacity) { set_error("n exceeds grid capacity"); return PSB_ERR_CAPACITY; }
  PSB_HIP(hipSetDevice(g->ctx->device));
  PSB_HIP(hipMemsetAsync(g->flags, 0, 4, g->ctx->stream));
  return grid_build(g, reinterpret_cast<const float4*>(pos4), n, nullptr, g->ctx->stream);
}

psb_status psb_grid_get_arrays(psb_grid g, const uint32_t** hash, const uint32_t** perm,
                               const uint32_t** cell_start, uint64_t* ncells) {
  PSB_ARG(g, "null grid");
  if (hash) *hash = g->hash;
  if (perm) *perm = g->perm;
  if (cell_start) *cell_start = g->start;
  if (ncells) *ncells = g->ncells;
  return PSB_OK;
}

psb_status psb_grid_reorder(psb_grid g, const float* src, float* dst, uint32_t n,
                            int32_t elem_floats) {
  PSB_ARG(g && src && dst, "null argument");
  PSB_ARG(elem_floats == 1 || elem_floats == 4, "elem_floats must be 1 or 4");
  PSB_ARG(n == g->n, "n differs from the last psb_store_particles");
  if (!n) return PSB_OK;
  if (elem_floats == 4)
    hipLaunchKernelGGL(k_gather<4>, dim3(blocks_for(n)), dim3(kBlock), 0, g->ctx->stream, src, dst,
                       g->perm, n);
  else
    hipLaunchKernelGGL(k_gather<1>, dim3(blocks_for(n)), dim3(kBlock), 0, g->ctx->stream, src, dst,
                       g->perm, n);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status psb_search_neighbors(psb_grid g, const float* sorted_pos4, uint32_t n, float radius,
                                psb_neighbors nb, uint32_t* flags_out) {
  PSB_ARG(g && (sorted_pos4 || !n), "null argument");
  PSB_ARG(nb.count && nb.list && nb.capacity > 0, "bad neighbour buffers");
  PSB_ARG(nb.n == n, "neighbors.n must equal n (ELL stride)");
  PSB_ARG(n == g->n, "n differs from the last psb_store_particles");
  PSB_ARG(radius > 0.f, "radius must be > 0");
  psb_context ctx = g->ctx;
  PSB_HIP(hipSetDevice(ctx->device));
  PSB_HIP(hipMemsetAsync(ctx->d_flags, 0, 8, ctx->stream));
  GridDev gd = grid_dev(g->desc, radius);
  PSB_TRY(launch_neighbors(g, reinterpret_cast<const float4*>(sorted_pos4), n, radius * radius,
                           gd.R, nb, nullptr, 0.f, ctx->d_flags, ctx->d_flags + 1, ctx->stream));
  PSB_HIP(hipMemcpyAsync(ctx->h_pinned, ctx->d_flags, 8, hipMemcpyDeviceToHost, ctx->stream));
  PSB_HIP(hipMemcpyAsync(ctx->h_pinned + 2, g->flags, 4, hipMemcpyDeviceToHost, ctx->stream));
  PSB_HIP(hipStreamSynchronize(ctx->stream));
  if (flags_out) *flags_out = ctx->h_pinned[0] | ctx->h_pinned[2];
  return PSB_OK;
}

}  // extern "C"
