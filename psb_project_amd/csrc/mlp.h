// mlp.h -- pressure MLP (internal declarations).
#pragma once
#include "psb_internal.h"

namespace psb {

constexpr int kIn = 8, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kHid * kIn, kOffW2 = kOffB1 + kHid,
              kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid;
constexpr int kNumParams = kOffB3 + 1;   // 4801
static_assert(kNumParams == 4801, "8-64-64-1 parameter count");
constexpr int kLdsRow = 72;              // padded row (bf16 elements) of a 16x64 LDS tile

typedef short bf16x4_t __attribute__((ext_vector_type(4)));

// BF16 fragments of the weights in MFMA lane order (see mlp.hip).
struct MlpFrags {
  bf16x4_t* w1;   // [4][64]
  bf16x4_t* w2f;  // [16][64]
  bf16x4_t* w2b;  // [16][64]
  float* w3;      // [64] BF16-rounded values
};

struct MlpNorm {
  float mu[kIn];
  float inv_sd[kIn];
};

psb_status mlp_infer(psb_mlp m, const float* feat, uint32_t n, float* out, float4* xw, int relu,
                     float scale, hipStream_t s, const uint32_t* ids = nullptr,
                     uint32_t n_fluid = kNoBodies);
float mlp_pscale(psb_mlp m);

}  // namespace psb
