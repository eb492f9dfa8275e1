// Device side of the learned step-size predictor (layout and conventions in
// predictor.h): LDS staging of the parameters, the raw inputs of a neuron,
// and the wavefront inference on the matrix cores.
#pragma once
#include <hip/hip_runtime.h>

#include "predictor.h"

namespace ngpu {

// Predictor parameters staged in LDS by every block of a kernel that infers.
struct PredLDS {
  float mu[kPI], is[kPI];
  unsigned short w1[kPH * kPI];   // bfloat16, [hidden][input]
  unsigned short w2[kPH * kPH];   // bfloat16, [out][in]
  float b1[kPH], b2[kPH], w3[kPH];
  float b3;
};

// All threads of the block call this, then __syncthreads().
__device__ inline void pred_stage(PredLDS& s, const float* P) {
  for (int k = threadIdx.x; k < kPH * kPH; k += blockDim.x) {
    s.w2[k] = (unsigned short)bf16_bits(P[kPOffW2 + k]);
    if (k < kPH * kPI) s.w1[k] = (unsigned short)bf16_bits(P[kPOffW1 + k]);
    if (k < kPH) {
      s.b1[k] = P[kPOffB1 + k];
      s.b2[k] = P[kPOffB2 + k];
      s.w3[k] = P[kPOffW3 + k];
    }
    if (k < kPI) {
      s.mu[k] = P[kPOffMu + k];
      s.is[k] = P[kPOffIs + k];
    }
    if (k == 0) s.b3 = P[kPOffB3];
  }
}

// Raw inputs of a neuron of aeif_cond_beta_multisynapse with NP ports
// (state [V, w, g_0, g1_0, g_1, g1_1, ...]).
template <int NP>
__device__ inline void pred_raw_inputs(const float* y, float h_ctrl, float err_last, float dt,
                                       float* r) {
  r[0] = y[0];
  r[1] = y[1];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float v = 0.0f;
    if (k < NP) v = y[2 + 2 * k];
    else if (k - NP < NP) v = y[3 + 2 * (k - NP)];
    r[2 + k] = v;
  }
  r[6] = __log2f(fmaxf(h_ctrl, 1.0e-30f) / dt);
  r[7] = __log2f(fmaxf(err_last, 1.0e-12f));
}

typedef short pred_s4 __attribute__((ext_vector_type(4)));
typedef float pred_f16 __attribute__((ext_vector_type(16)));
typedef __bf16 pred_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int pred_u4 __attribute__((ext_vector_type(4)));

// eight bfloat16 from four packed pairs
__device__ inline pred_bf8 pred_pack8(unsigned int p0, unsigned int p1, unsigned int p2,
                                      unsigned int p3) {
  pred_u4 v;
  v[0] = p0;
  v[1] = p1;
  v[2] = p2;
  v[3] = p3;
  return __builtin_bit_cast(pred_bf8, v);
}

__device__ inline pred_s4 pred_pack4(unsigned int lo, unsigned int hi) {
  pred_s4 v;
  v[0] = (short)(lo & 0xffffu);
  v[1] = (short)(lo >> 16);
  v[2] = (short)(hi & 0xffffu);
  v[3] = (short)(hi >> 16);
  return v;
}
// two floats rounded to bfloat16 (nearest even, as bf16_bits) and packed,
// one v_cvt_pk_bf16_f32 on gfx950
typedef __bf16 pred_bf2 __attribute__((ext_vector_type(2)));
__device__ inline unsigned int pred_pair(float a, float b) {
  pred_bf2 v;
  v[0] = (__bf16)a;
  v[1] = (__bf16)b;
  return __builtin_bit_cast(unsigned int, v);
}

// Output log2(h/dt) for the neuron held by this lane.  Every lane of the
// wave must call it together (the matrix instructions read all 64 lanes);
// lanes without a neuron pass any finite inputs and ignore the result.
__device__ inline float pred_infer_wave(const PredLDS& s_in, const float* raw) {
  // the parameters are read from LDS at every call: an opaque zero offset
  // keeps the compiler from hoisting ~70 loop-invariant values out of a
  // caller's loop into registers for the whole kernel
  int z = 0;
  asm volatile("" : "+s"(z));
  const PredLDS& s = *(&s_in + z);
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5, col = lane & 31;
  float x[kPI];
#pragma unroll
  for (int i = 0; i < kPI; ++i) x[i] = (raw[i] - s.mu[i]) * s.is[i];
  const unsigned int f01 = pred_pair(x[0], x[1]), f23 = pred_pair(x[2], x[3]),
                     f45 = pred_pair(x[4], x[5]), f67 = pred_pair(x[6], x[7]);
  const unsigned int g01 = __shfl_xor(f01, 32), g23 = __shfl_xor(f23, 32),
                     g45 = __shfl_xor(f45, 32), g67 = __shfl_xor(f67, 32);
  // B operand, k = 4*half + j, column = neuron: block 0 holds neurons 0..31,
  // block 1 neurons 32..63 (inputs 0..3 come from lanes < 32, 4..7 from >= 32)
  const pred_s4 b0 = half ? pred_pack4(g45, g67) : pred_pack4(f01, f23);
  const pred_s4 b1 = half ? pred_pack4(f45, f67) : pred_pack4(g01, g23);
  // A operand W1[row = col][k = 4*half + j]
  const unsigned short* w1 = s.w1 + col * kPI + 4 * half;
  const pred_s4 a1 = pred_pack4(w1[0] | ((unsigned int)w1[1] << 16),
                                w1[2] | ((unsigned int)w1[3] << 16));
  pred_f16 c1, c2;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 8 * (i >> 2) + 4 * half + (i & 3);
    c1[i] = s.b1[row];
    c2[i] = s.b2[row];
  }
  float out[2];
#pragma unroll
  for (int blk = 0; blk < 2; ++blk) {
    pred_f16 h1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1, blk ? b1 : b0, c1, 0, 0, 0);
    pred_f16 h2 = c2;
    // layer 2 on gfx950's v_mfma_f32_32x32x16_bf16, two k-blocks of 16
    // hidden units: lane (half, slot j) of k-block kb carries hidden unit
    // 16 kb + 8 (j / 4) + 4 half + j % 4 in both operands -- its own layer-1
    // results h1[8 kb + j] as B, the matching W2 columns as A (the sum over
    // k does not depend on how the units are numbered within a block)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const unsigned short* w2 = s.w2 + col * kPH + 16 * kb + 4 * half;
      const pred_bf8 a2 = pred_pack8(w2[0] | ((unsigned int)w2[1] << 16), w2[2] | ((unsigned int)w2[3] << 16),
                                     w2[8] | ((unsigned int)w2[9] << 16), w2[10] | ((unsigned int)w2[11] << 16));
      const int i0 = 8 * kb;
      const pred_bf8 bb = pred_pack8(pred_pair(fmaxf(h1[i0], 0.0f), fmaxf(h1[i0 + 1], 0.0f)),
                                     pred_pair(fmaxf(h1[i0 + 2], 0.0f), fmaxf(h1[i0 + 3], 0.0f)),
                                     pred_pair(fmaxf(h1[i0 + 4], 0.0f), fmaxf(h1[i0 + 5], 0.0f)),
                                     pred_pair(fmaxf(h1[i0 + 6], 0.0f), fmaxf(h1[i0 + 7], 0.0f)));
      h2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, bb, h2, 0, 0, 0);
    }
    float part = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = 8 * (i >> 2) + 4 * half + (i & 3);
      part += s.w3[row] * fmaxf(h2[i], 0.0f);
    }
    out[blk] = part + __shfl_xor(part, 32);
  }
  return (half ? out[1] : out[0]) + s.b3;
}

}  // namespace ngpu
