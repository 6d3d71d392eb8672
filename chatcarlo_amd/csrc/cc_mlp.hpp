// Learned importance map: a 7 -> 64 -> 64 -> 1 MLP on BF16 MFMA with FP32 accumulation.
//
// f(x) = w3 . relu(relu(X W1) W2 + b2) + b3, where X holds the 7 phase-space
// features of a photon in the frame of the scoring region (position relative to
// the region centre over a length scale; direction in the basis whose first axis
// points from the photon to the region centre, so that "heading for the region"
// is one coordinate; ln(E / 50 keV) / 2), a constant 1 in column 7 (so W1's row
// 7 is the first-layer bias) and zeros up to the MFMA depth of 16.  exp(f)
// estimates the expected future H*(10) contribution of the photon to the region.
//
// The GEMMs of forward and backward passes run on v_mfma_f32_16x16x16_bf16 with
// operands in registers, a wave's 64 photons or rows at a time.  Fragment layout of
// the 16x16x16 BF16 MFMA (wave64): lane l holds A[l % 16][4 (l / 16) + e],
// B[4 (l / 16) + e][l % 16] for e = 0..3 and returns D[4 (l / 16) + e][l % 16].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cc {

constexpr int kMlpIn = 16;     // padded input width (7 features + bias column + zeros)
constexpr int kMlpH = 64;      // hidden width
// flat parameter vector: W1 [16][64], W2 [64][64], b2 [64], w3 [64], b3 [1]
constexpr int kOffW1 = 0;
constexpr int kOffW2 = kOffW1 + kMlpIn * kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpH;
constexpr int kMlpParams = kOffB3 + 1;   // 5249 slots; rows 8..15 of W1 are held at zero

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float x) {
  // gfx950's v_cvt_pk_bf16_f32: round to nearest even, denormals kept (the f32 mode of HIP code),
  // the same bits as adding 0x7FFF plus the kept lsb to the float's bits for every finite input
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Feature row of one photon state (7 features, bias 1, zeros).
struct MlpFrame {
  float c[3];      // frame centre (the scoring region's centre)
  float inv_l;     // 1 / length scale
};
__device__ __forceinline__ void mlp_features(const MlpFrame& F, const float p[3], const float d[3], float E,
                                             float x[kMlpIn]) {
  const float r[3] = {F.c[0] - p[0], F.c[1] - p[1], F.c[2] - p[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] = -r[k] * F.inv_l;
  const float rl = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  float u[3] = {0.f, 0.f, 1.f};
  if (rl > 0.f) {
    const float ir = 1.0f / rl;
    u[0] = r[0] * ir; u[1] = r[1] * ir; u[2] = r[2] * ir;
  }
  // e2 = unit(z x u), or unit(u x x) when u is within 2.6 degrees of z; e3 = u x e2
  float e2[3];
  if (fabsf(u[2]) < 0.999f) {
    const float in = 1.0f / sqrtf(u[0] * u[0] + u[1] * u[1]);
    e2[0] = -u[1] * in; e2[1] = u[0] * in; e2[2] = 0.f;
  } else {
    const float in = 1.0f / sqrtf(u[1] * u[1] + u[2] * u[2]);
    e2[0] = 0.f; e2[1] = u[2] * in; e2[2] = -u[1] * in;
  }
  const float e3[3] = {u[1] * e2[2] - u[2] * e2[1], u[2] * e2[0] - u[0] * e2[2], u[0] * e2[1] - u[1] * e2[0]};
  x[3] = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
  x[4] = d[0] * e2[0] + d[1] * e2[1] + d[2] * e2[2];
  x[5] = d[0] * e3[0] + d[1] * e3[1] + d[2] * e3[2];
  x[6] = 0.5f * logf(E / 50.0f);
  x[7] = 1.0f;
#pragma unroll
  for (int k = 8; k < kMlpIn; ++k) x[k] = 0.0f;
}

// ---------------------------------------------------------------------------
// Wave-level inference: one wavefront evaluates the 64 photons held by its lanes.
// Every product is computed transposed, weights as the A operand and photons as
// the B operand: Z1^T = W1^T X^T, then Z2^T = W2^T relu(Z1)^T.  The D layout of a
// 16x16x16 MFMA (row 4 (l / 16) + e, column l % 16) is then exactly the B layout
// the next MFMA wants, so the hidden layer never leaves registers.  W1 and W2 are
// held in registers for the whole kernel, b2 and w3 in a 512-byte LDS block
// (lanes of one g read one address: broadcast).
// ---------------------------------------------------------------------------
struct MlpWave {
  bf16x4 a1[4];        // a1[tm][e] = W1[4g + e][16 tm + r]          (r = lane % 16, g = lane / 16)
  bf16x4 a2[4][4];     // a2[tm][kb][e] = W2[16 kb + 4g + e][16 tm + r]
  const float* bw;     // LDS: b2 [64], w3 [64]
  float b3;
};

// Block-wide (call before any divergence): registers of every wave and the LDS block sbw[128].
__device__ __forceinline__ void mlp_wave_load(const uint16_t* wb, const float* w, MlpWave& M, float* sbw) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  for (int k = threadIdx.x; k < 2 * kMlpH; k += blockDim.x) sbw[k] = w[kOffB2 + k];   // b2 then w3
#pragma unroll
  for (int tm = 0; tm < 4; ++tm) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      M.a1[tm][e] = (short)wb[(4 * g + e) * kMlpH + 16 * tm + r];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) M.a2[tm][kb][e] = (short)wb[kOffW2 + (16 * kb + 4 * g + e) * kMlpH + 16 * tm + r];
    }
  }
  M.bw = sbw;
  M.b3 = w[kOffB3];
  __syncthreads();
}

// f of the photon whose 16 features this lane holds.  Wave-uniform call.
__device__ __forceinline__ float mlp_wave_eval(const MlpWave& M, const float x[kMlpIn]) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  int pk[kMlpIn / 2];   // bf16 pairs (2c, 2c + 1) of this lane's features
#pragma unroll
  for (int c = 0; c < kMlpIn / 2; ++c) pk[c] = (int)f2bf(x[2 * c]) | ((int)f2bf(x[2 * c + 1]) << 16);
  float s[4];
#pragma unroll
  for (int tn = 0; tn < 4; ++tn) {
    // B operand of layer 1: features 4g .. 4g + 3 of photon 16 tn + r
    int lo = 0, hi = 0;
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const int t0 = __shfl(pk[2 * gg], 16 * tn + r);
      const int t1 = __shfl(pk[2 * gg + 1], 16 * tn + r);
      if (gg == g) { lo = t0; hi = t1; }
    }
    bf16x4 b;
    b[0] = (short)(lo & 0xFFFF); b[1] = (short)((unsigned)lo >> 16);
    b[2] = (short)(hi & 0xFFFF); b[3] = (short)((unsigned)hi >> 16);
    bf16x4 h[4];
#pragma unroll
    for (int tm = 0; tm < 4; ++tm) {
      const f32x4 z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(M.a1[tm], b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
      for (int e = 0; e < 4; ++e) h[tm][e] = (short)f2bf(fmaxf(z[e], 0.0f));
    }
    float acc = 0.f;
#pragma unroll
    for (int tm = 0; tm < 4; ++tm) {
      f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(M.a2[tm][kb], h[kb], z, 0, 0, 0);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int h2 = 16 * tm + 4 * g + e;
        acc += fmaxf(z[e] + M.bw[h2], 0.0f) * M.bw[kMlpH + h2];
      }
    }
    acc += __shfl_xor(acc, 16);
    acc += __shfl_xor(acc, 32);
    s[tn] = acc;                    // f - b3 of photon 16 tn + r
  }
  float f = 0.f;
#pragma unroll
  for (int tn = 0; tn < 4; ++tn) {
    const float t = __shfl(s[tn], r);
    if (g == tn) f = t;
  }
  return f + M.b3;
}

}  // namespace cc
