// mlp.hip -- per-particle pressure MLP for the PCISPH warm start, on BF16
// MFMA with FP32 accumulation (gfx950: v_mfma_f32_16x16x16_bf16).
//
// Network: 8 features -> 64 -> 64 -> 1, SiLU, 4,801 parameters.  FP32 master
// parameters, flat:  W1[64][8] @0, b1[64] @512, W2[64][64] @576, b2[64] @4672,
// W3[64] @4736, b3 @4800  (W[out][in]).  The layer-1 and layer-2 products
// and every weight-gradient product run on MFMA with BF16 operands; biases,
// activations and the 64-term output dot product are FP32.
//
// MFMA 16x16x16 fragment layouts (one wave, lane l, g = l / 16, r = l % 16):
//   A (16 x 16, [i][k])   a[j] = A[r][4g + j]
//   B (16 x 16, [k][n])   b[j] = B[4g + j][r]
//   D (16 x 16, [i][n])   d[j] = D[4g + j][r]
// A tile is 16 samples (particles).  Layer outputs come out in D layout
// (lane holds 4 samples of one unit); the next layer needs A layout (lane
// holds 4 units of one sample), so activations pass through a per-wave LDS
// tile.  In the backward pass the D-layout tiles are directly the A operand
// (dZ^T) and B operand (H) of the weight-gradient products, so those need no
// transpose.
#include <cmath>
#include <cstring>
#include <vector>

#include "psb_internal.h"
#include "mlp.h"

namespace psb {

typedef bf16x4_t bf16x4;
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = (short)f2bf(a); v[1] = (short)f2bf(b); v[2] = (short)f2bf(c); v[3] = (short)f2bf(d);
  return v;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---- fragment packing of the FP32 master parameters --------------------------
// W1f[nt][l]: b[j] = W1[nt*16 + r][4g + j] (zero for inputs >= 8)
// W2f[kt][nt][l]: b[j] = W2[nt*16 + r][kt*16 + 4g + j]    (forward: Z2 = H1 W2^T)
// W2b[kt][nt][l]: b[j] = W2[kt*16 + 4g + j][nt*16 + r]    (backward: dH1 = dZ2 W2)
__global__ void k_mlp_pack(const float* __restrict__ w, MlpFrags f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;   // 0 .. 36*64-1
  if (t >= 36 * 64) return;
  const int frag = t / 64, l = t % 64, g = l / 16, r = l % 16;
  const float* W1 = w + kOffW1;
  const float* W2 = w + kOffW2;
  float v[4];
  if (frag < 4) {
    const int nt = frag;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * g + j;
      v[j] = k < kIn ? W1[(nt * 16 + r) * kIn + k] : 0.f;
    }
    f.w1[frag * 64 + l] = pack4(v[0], v[1], v[2], v[3]);
  } else if (frag < 20) {
    const int kt = (frag - 4) / 4, nt = (frag - 4) % 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = W2[(nt * 16 + r) * kHid + kt * 16 + 4 * g + j];
    f.w2f[(frag - 4) * 64 + l] = pack4(v[0], v[1], v[2], v[3]);
  } else {
    const int kt = (frag - 20) / 4, nt = (frag - 20) % 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = W2[(kt * 16 + 4 * g + j) * kHid + nt * 16 + r];
    f.w2b[(frag - 20) * 64 + l] = pack4(v[0], v[1], v[2], v[3]);
  }
  if (t < kHid) f.w3[t] = bf2f(f2bf(w[kOffW3 + t]));
}

// Forward of one 16-sample tile.  xin(r, k) gives normalised input k of
// sample r of the tile (0 for padding).  Outputs in D layout: z1/h1, z2/h2
// [nt][j] = unit nt*16 + r of sample 4g + j; returns y for samples 4g + j in
// lanes with r == 0 (other lanes hold partial sums).
struct TileAct {
  f32x4 z1[4], z2[4], h1[4], h2[4];
  float y[4];
};

template <class XIN>
__device__ __forceinline__ void tile_forward(const MlpFrags& f, const float* __restrict__ w,
                                             XIN xin, uint16_t* lds, TileAct& t) {
  const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  // layer 1: A1[r][4g+j] = x[r][4g+j]
  float a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) a[j] = xin(r, 4 * g + j);
  const bf16x4 A1 = pack4(a[0], a[1], a[2], a[3]);
  const float* b1 = w + kOffB1;
  const float* b2 = w + kOffB2;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 z = mfma(A1, f.w1[nt * 64 + l], f32x4{0.f, 0.f, 0.f, 0.f});
    const float bias = b1[nt * 16 + r];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float zz = z[j] + bias;
      t.z1[nt][j] = zz;
      t.h1[nt][j] = zz * sigm(zz);
      lds[(4 * g + j) * kLdsRow + nt * 16 + r] = f2bf(t.h1[nt][j]);
    }
  }
  __builtin_amdgcn_wave_barrier();
  // layer 2: A2[kt][r][4g+j] = h1[r][kt*16 + 4g + j]
  bf16x4 A2[4];
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) {
    const uint16_t* src = lds + r * kLdsRow + kt * 16 + 4 * g;
    A2[kt] = bf16x4{(short)src[0], (short)src[1], (short)src[2], (short)src[3]};
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) z = mfma(A2[kt], f.w2f[(kt * 4 + nt) * 64 + l], z);
    const float bias = b2[nt * 16 + r];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float zz = z[j] + bias;
      t.z2[nt][j] = zz;
      t.h2[nt][j] = zz * sigm(zz);
    }
  }
  // layer 3 (FP32 VALU): y[s] = sum_k h2[s][k] W3[k] + b3
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float s = 0.f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) s = __builtin_fmaf(t.h2[nt][j], f.w3[nt * 16 + r], s);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    t.y[j] = s + w[kOffB3];
  }
}

// Inference over n samples of a feature SoA [8][n]; y = max(0, out * pscale)
// written to pres[] and, when xw != nullptr, to the w lane of xw[].  With
// ids, samples whose id >= n_fluid (rigid-sphere boundary particles) get 0.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ feat, uint32_t n,
                                                   const float* __restrict__ w, MlpFrags f,
                                                   MlpNorm nm, float* __restrict__ out,
                                                   float4* xw, float pscale, int clamp_relu,
                                                   const uint32_t* __restrict__ ids,
                                                   uint32_t n_fluid) {
  __shared__ uint16_t lds_all[4 * 16 * kLdsRow];
  const int wave = threadIdx.x >> 6;
  uint16_t* lds = lds_all + wave * 16 * kLdsRow;
  const uint64_t tile = (uint64_t)blockIdx.x * 4 + wave;
  const uint64_t p0 = tile * 16;
  if (p0 >= n) return;   // whole wave exits together
  auto xin = [&](int r, int k) -> float {
    const uint64_t p = p0 + r;
    if (k >= kIn || p >= n) return 0.f;
    return (feat[(uint64_t)k * n + p] - nm.mu[k]) * nm.inv_sd[k];
  };
  TileAct t;
  tile_forward(f, w, xin, lds, t);
  const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  if (r == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t p = p0 + 4 * g + j;
      if (p < n) {
        float v = t.y[j] * pscale;
        if (clamp_relu && v < 0.f) v = 0.f;
        if (ids && ids[p] >= n_fluid) v = 0.f;
        out[p] = v;
        if (xw) reinterpret_cast<float*>(xw)[4 * p + 3] = v;
      }
    }
  }
}

// Per-wave partial gradients: row w of part holds wave w's 4801 gradient
// entries and its loss sum at kOffLoss; k_mlp_reduce adds the rows in wave
// order, so a training run gives the same network every time.
constexpr int kOffLoss = kNumParams;
constexpr int kPartRow = kNumParams + 1;

// Forward + backward over a minibatch; each wave stores its gradient sums
// (and loss sum) into its row of part.  Samples idx[b0 .. b0 + nb) of the
// dataset (feature SoA [8][ns] + target[ns]); MSE on normalised targets.
__global__ __launch_bounds__(256) void k_mlp_train(const float* __restrict__ feat,
                                                   const float* __restrict__ target, uint64_t ns,
                                                   const uint32_t* __restrict__ idx, uint32_t nb,
                                                   const float* __restrict__ w, MlpFrags f,
                                                   MlpNorm nm, float* __restrict__ part) {
  __shared__ uint16_t lds_all[4 * 16 * kLdsRow];
  const int wave = threadIdx.x >> 6;
  uint16_t* lds = lds_all + wave * 16 * kLdsRow;
  const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
  const uint32_t ntiles = (nb + 15) / 16;
  const uint32_t wave_id = blockIdx.x * 4 + wave, nwaves = gridDim.x * 4;
  f32x4 gW2[4][4], gW1[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f}, gw3[4] = {0.f, 0.f, 0.f, 0.f};
  float gb3 = 0.f, lsum = 0.f;
  const float inv_b = 1.0f / (float)nb;
  for (uint32_t tile = wave_id; tile < ntiles; tile += nwaves) {
    const uint32_t s0 = tile * 16;
    auto sample = [&](int rr) -> int64_t {
      const uint32_t b = s0 + rr;
      return b < nb ? (int64_t)idx[b] : -1;
    };
    auto xin = [&](int rr, int k) -> float {
      const int64_t s = sample(rr);
      if (k >= kIn || s < 0) return 0.f;
      return (feat[(uint64_t)k * ns + s] - nm.mu[k]) * nm.inv_sd[k];
    };
    TileAct t;
    tile_forward(f, w, xin, lds, t);
    // dy for the 4 samples 4g + j, broadcast from lane 16g
    float dy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t s = sample(4 * g + j);
      float d = 0.f;
      if (s >= 0) {
        const float e = t.y[j] - target[s];
        d = 2.0f * e * inv_b;
        if (r == 0) lsum += e * e;
      }
      dy[j] = __shfl(d, 16 * g, 64);
    }
    // layer 3
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) gw3[nt] = __builtin_fmaf(t.h2[nt][j], dy[j], gw3[nt]);
    if (r == 0) gb3 += dy[0] + dy[1] + dy[2] + dy[3];
    // dZ2 = dy W3 * silu'(z2)   (D layout)
    f32x4 dz2[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const float w3 = f.w3[nt * 16 + r];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float sg = sigm(t.z2[nt][j]);
        const float dsilu = sg * (1.0f + t.z2[nt][j] * (1.0f - sg));
        dz2[nt][j] = dy[j] * w3 * dsilu;
        gb2[nt] += dz2[nt][j];
      }
    }
    // dW2[out][in] += sum_s dZ2[s][out] H1[s][in]:  A = dZ2^T (D tile), B = H1 (D tile)
    bf16x4 Ad[4], Bh[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      Ad[a] = pack4(dz2[a][0], dz2[a][1], dz2[a][2], dz2[a][3]);
      Bh[a] = pack4(t.h1[a][0], t.h1[a][1], t.h1[a][2], t.h1[a][3]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nk = 0; nk < 4; ++nk) gW2[mt][nk] = mfma(Ad[mt], Bh[nk], gW2[mt][nk]);
    // dH1 = dZ2 W2: dZ2 to A layout through LDS
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) lds[(4 * g + j) * kLdsRow + nt * 16 + r] = f2bf(dz2[nt][j]);
    __builtin_amdgcn_wave_barrier();
    bf16x4 A2[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const uint16_t* src = lds + r * kLdsRow + kt * 16 + 4 * g;
      A2[kt] = bf16x4{(short)src[0], (short)src[1], (short)src[2], (short)src[3]};
    }
    __builtin_amdgcn_wave_barrier();
    f32x4 dz1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 dh = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) dh = mfma(A2[kt], f.w2b[(kt * 4 + nt) * 64 + l], dh);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float sg = sigm(t.z1[nt][j]);
        const float dsilu = sg * (1.0f + t.z1[nt][j] * (1.0f - sg));
        dz1[nt][j] = dh[j] * dsilu;
        gb1[nt] += dz1[nt][j];
      }
    }
    // dW1[out][in] += sum_s dZ1[s][out] X[s][in]:  A = dZ1^T (D tile), B[s][in] = x
    float xb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xb[j] = xin(4 * g + j, r);
    const bf16x4 Bx = pack4(xb[0], xb[1], xb[2], xb[3]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
      gW1[mt] = mfma(pack4(dz1[mt][0], dz1[mt][1], dz1[mt][2], dz1[mt][3]), Bx, gW1[mt]);
  }
  // ---- flush: reduce over the 4 lane groups where needed; every entry of
  // the wave's row is written by exactly one lane ----
  float* row = part + (size_t)wave_id * kPartRow;
  float* gw1p = row + kOffW1;
  float* gw2p = row + kOffW2;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int out = mt * 16 + 4 * g + j;
      if (r < kIn) gw1p[out * kIn + r] = gW1[mt][j];
#pragma unroll
      for (int nk = 0; nk < 4; ++nk) gw2p[out * kHid + nk * 16 + r] = gW2[mt][nk][j];
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float a = gb1[nt], b = gb2[nt], c = gw3[nt];
    a += __shfl_xor(a, 16, 64); a += __shfl_xor(a, 32, 64);
    b += __shfl_xor(b, 16, 64); b += __shfl_xor(b, 32, 64);
    c += __shfl_xor(c, 16, 64); c += __shfl_xor(c, 32, 64);
    if (g == 0) {
      row[kOffB1 + nt * 16 + r] = a;
      row[kOffB2 + nt * 16 + r] = b;
      row[kOffW3 + nt * 16 + r] = c;
    }
  }
  gb3 += __shfl_xor(gb3, 16, 64); gb3 += __shfl_xor(gb3, 32, 64);
  lsum += __shfl_xor(lsum, 16, 64); lsum += __shfl_xor(lsum, 32, 64);
  if (l == 0) {
    row[kOffB3] = gb3;
    row[kOffLoss] = lsum;
  }
}

// grad[k] += sum over the waves' rows of part[w][k]; loss = the loss sum.
// Fixed order: row w goes to accumulator w % 8 (eight loads in flight), the
// accumulators are added pairwise.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, uint32_t rows,
                                                    float* __restrict__ grad,
                                                    float* __restrict__ loss) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (uint32_t)kPartRow) return;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  uint32_t w = 0;
  for (; w + 8 <= rows; w += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += part[(size_t)(w + u) * kPartRow + k];
  }
  for (; w < rows; ++w) acc[w & 7u] += part[(size_t)w * kPartRow + k];
  const float sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  if (k < (uint32_t)kNumParams) grad[k] = grad[k] + sum;
  else *loss = sum;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; zeroes grad.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  const float gr = grad[i];
  const float m = b1 * m1[i] + (1.0f - b1) * gr;
  const float v = b2 * m2[i] + (1.0f - b2) * gr * gr;
  m1[i] = m;
  m2[i] = v;
  w[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  grad[i] = 0.f;
}

}  // namespace psb

using namespace psb;

struct psb_mlp_s {
  psb_context ctx = nullptr;
  float* w = nullptr;      // [kNumParams] FP32 master
  float* grad = nullptr;   // [kNumParams]
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* loss = nullptr;   // [1]
  float* part = nullptr;   // per-wave gradient rows of a training step
  uint32_t part_rows = 0;
  MlpFrags frags{};
  void* frag_mem = nullptr;
  MlpNorm norm{};
  float pscale = 1.f;
  int64_t adam_t = 0;
};

namespace psb {

psb_status mlp_pack(psb_mlp m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((36 * 64 + 255) / 256), dim3(256), 0, m->ctx->stream, m->w,
                     m->frags);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

psb_status mlp_infer(psb_mlp m, const float* feat, uint32_t n, float* out, float4* xw, int relu,
                     float scale, hipStream_t s, const uint32_t* ids, uint32_t n_fluid) {
  if (!n) return PSB_OK;
  const uint32_t tiles = (n + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer, dim3((tiles + 3) / 4), dim3(256), 0, s, feat, n, m->w, m->frags,
                     m->norm, out, xw, scale, relu, ids, n_fluid);
  PSB_HIP(hipGetLastError());
  return PSB_OK;
}

float mlp_pscale(psb_mlp m) { return m->pscale; }

}  // namespace psb

extern "C" {

psb_status psb_mlp_create(psb_context ctx, uint64_t seed, psb_mlp* out) {
  PSB_ARG(ctx && out, "null argument");
  PSB_HIP(hipSetDevice(ctx->device));
  psb_mlp m = new psb_mlp_s();
  m->ctx = ctx;
  auto fail = [&](hipError_t e) { psb_mlp_destroy(m); return hip_fail(e, "mlp alloc", __FILE__, __LINE__); };
  hipError_t e;
  if ((e = hipMalloc(&m->w, 4 * kNumParams)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->grad, 4 * kNumParams)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->m1, 4 * kNumParams)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->m2, 4 * kNumParams)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&m->loss, 64)) != hipSuccess) return fail(e);
  const size_t fb = sizeof(bf16x4) * 64 * 36 + 4 * kHid;
  if ((e = hipMalloc(&m->frag_mem, fb)) != hipSuccess) return fail(e);
  bf16x4* fm = (bf16x4*)m->frag_mem;
  m->frags.w1 = fm;
  m->frags.w2f = fm + 4 * 64;
  m->frags.w2b = fm + 20 * 64;
  m->frags.w3 = (float*)(fm + 36 * 64);
  for (int k = 0; k < kIn; ++k) { m->norm.mu[k] = 0.f; m->norm.inv_sd[k] = 1.f; }
  // He-uniform initialisation from a splitmix64 stream (deterministic)
  std::vector<float> w(kNumParams, 0.f);
  uint64_t st = seed ? seed : 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int cnt, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < cnt; ++i) w[off + i] = (float)((2.0 * next() - 1.0) * a);
  };
  fill(kOffW1, kHid * kIn, kIn);
  fill(kOffW2, kHid * kHid, kHid);
  fill(kOffW3, kHid, kHid);
  return psb_mlp_set_params(m, w.data(), kNumParams) == PSB_OK ? (*out = m, PSB_OK)
                                                               : (psb_mlp_destroy(m), PSB_ERR_HIP);
}

psb_status psb_mlp_destroy(psb_mlp m) {
  if (!m) return PSB_OK;
  hipSetDevice(m->ctx->device);
  hipStreamSynchronize(m->ctx->stream);
  hipFree(m->w); hipFree(m->grad); hipFree(m->m1); hipFree(m->m2); hipFree(m->loss);
  hipFree(m->frag_mem);
  if (m->part) hipFree(m->part);
  delete m;
  return PSB_OK;
}

int32_t psb_mlp_num_params(void) { return kNumParams; }

psb_status psb_mlp_set_params(psb_mlp m, const float* host_w, int32_t count) {
  PSB_ARG(m && host_w && count == kNumParams, "psb_mlp_set_params: need 4801 parameters");
  hipStream_t s = m->ctx->stream;
  PSB_HIP(hipMemcpyAsync(m->w, host_w, 4 * kNumParams, hipMemcpyHostToDevice, s));
  PSB_HIP(hipMemsetAsync(m->grad, 0, 4 * kNumParams, s));
  PSB_HIP(hipMemsetAsync(m->m1, 0, 4 * kNumParams, s));
  PSB_HIP(hipMemsetAsync(m->m2, 0, 4 * kNumParams, s));
  m->adam_t = 0;
  PSB_TRY(mlp_pack(m));
  PSB_HIP(hipStreamSynchronize(s));
  return PSB_OK;
}

psb_status psb_mlp_get_params(psb_mlp m, float* host_w, int32_t count) {
  PSB_ARG(m && host_w && count == kNumParams, "psb_mlp_get_params: need 4801 parameters");
  PSB_HIP(hipMemcpyAsync(host_w, m->w, 4 * kNumParams, hipMemcpyDeviceToHost, m->ctx->stream));
  PSB_HIP(hipStreamSynchronize(m->ctx->stream));
  return PSB_OK;
}

psb_status psb_mlp_set_normalization(psb_mlp m, const float mean[8], const float stddev[8],
                                     float pressure_scale) {
  PSB_ARG(m && mean && stddev && pressure_scale > 0.f, "bad normalisation");
  for (int k = 0; k < kIn; ++k) {
    PSB_ARG(stddev[k] > 0.f, "stddev must be > 0");
    m->norm.mu[k] = mean[k];
    m->norm.inv_sd[k] = 1.0f / stddev[k];
  }
  m->pscale = pressure_scale;
  return PSB_OK;
}

psb_status psb_mlp_forward(psb_mlp m, const float* features, uint32_t n, float* out) {
  PSB_ARG(m && (n == 0 || (features && out)), "null argument");
  PSB_TRY(mlp_infer(m, features, n, out, nullptr, 0, 1.0f, m->ctx->stream));
  return PSB_OK;
}

psb_status psb_mlp_train_step(psb_mlp m, const float* features, const float* targets,
                              uint64_t num_samples, const uint32_t* batch_idx, uint32_t batch,
                              float lr, int32_t apply_adam, float* loss_out) {
  PSB_ARG(m && features && targets && batch_idx && batch > 0, "null argument");
  hipStream_t s = m->ctx->stream;
  const uint32_t tiles = (batch + 15) / 16;
  uint32_t blocks = (tiles + 3) / 4;
  if (blocks > 512) blocks = 512;
  const uint32_t rows = 4 * blocks;
  if (rows > m->part_rows) {
    PSB_HIP(hipStreamSynchronize(s));
    if (m->part) PSB_HIP(hipFree(m->part));
    m->part = nullptr;
    m->part_rows = 0;
    PSB_HIP(hipMalloc((void**)&m->part, (size_t)rows * kPartRow * sizeof(float)));
    m->part_rows = rows;
  }
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, s, features, targets, num_samples,
                     batch_idx, batch, m->w, m->frags, m->norm, m->part);
  PSB_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kPartRow + 255) / 256), dim3(256), 0, s, m->part, rows,
                     m->grad, m->loss);
  PSB_HIP(hipGetLastError());
  if (apply_adam) {
    m->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)m->adam_t));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)m->adam_t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, s, m->w, m->grad,
                       m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2);
    PSB_HIP(hipGetLastError());
    PSB_TRY(mlp_pack(m));
  }
  if (loss_out) {
    float l = 0.f;
    PSB_HIP(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, s));
    PSB_HIP(hipStreamSynchronize(s));
    *loss_out = l / (float)batch;
  }
  return PSB_OK;
}

psb_status psb_mlp_get_grad(psb_mlp m, float* host_grad, int32_t count) {
  PSB_ARG(m && host_grad && count == kNumParams, "need 4801 entries");
  PSB_HIP(hipMemcpyAsync(host_grad, m->grad, 4 * kNumParams, hipMemcpyDeviceToHost, m->ctx->stream));
  PSB_HIP(hipStreamSynchronize(m->ctx->stream));
  return PSB_OK;
}

psb_status psb_mlp_zero_grad(psb_mlp m) {
  PSB_ARG(m, "null mlp");
  PSB_HIP(hipMemsetAsync(m->grad, 0, 4 * kNumParams, m->ctx->stream));
  return PSB_OK;
}

}  // extern "C"
