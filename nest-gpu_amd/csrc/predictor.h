// Learned step-size predictor: an MLP 8 -> 32 -> 32 -> 1 (ReLU) that proposes
// each RK5 substep's step size from the neuron's state; the step tried is the
// smaller of its proposal and the RK5 controller's (src/rk5.h:135-153).  The
// embedded error estimate still accepts or rejects every attempt, and a
// rejected attempt falls back to the controller's shrink rule, so the plain
// integrator's error test decides every step.  No counterpart exists in the
// reference.
//
// Parameter vector (floats; the C ABI uses the same layout):
//   mu[8], inv_sigma[8]          input normalisation (from the training log)
//   W1[32][8], b1[32]            hidden layer 1
//   W2[32][32], b2[32]           hidden layer 2
//   W3[32], b3                   output: log2(h / dt)
// Raw inputs: V, w, four conductance variables (the first four of g_0..g_P-1,
// g1_0..g1_P-1), log2(h_ctrl / dt) with h_ctrl the controller's proposal, and
// log2 of the last accepted error ratio.
//
// Inference runs per wavefront on the matrix cores: with the 64 neurons of a
// wave as the N dimension in two blocks of 32, layer 1 is one
// v_mfma_f32_32x32x8_bf16 per block (M = 32 hidden units, K = 8 inputs) and
// layer 2 two v_mfma_f32_32x32x16_bf16 (K = 32 in k-blocks of 16); a lane's
// layer-1 results (rows 8j + 4*(lane/32) + r, column lane%32) serve as its
// layer-2 B operand as they are, with the W2 columns of the A operand
// numbered to match, so the hidden activations never leave registers.
// Inputs and W1, W2 and the layer-1 activations are bfloat16 (round to
// nearest even); accumulation, biases, layer 3 and the output are FP32.
#pragma once
#include <cstdint>
#include <cstring>

#ifndef NGPU_HD
#ifdef __HIPCC__
#define NGPU_HD __host__ __device__
#else
#define NGPU_HD
#endif
#endif

namespace ngpu {

constexpr int kPI = 8, kPH = 32;
constexpr int kPOffMu = 0, kPOffIs = 8, kPOffW1 = 16, kPOffB1 = kPOffW1 + kPH * kPI,
              kPOffW2 = kPOffB1 + kPH, kPOffB2 = kPOffW2 + kPH * kPH, kPOffW3 = kPOffB2 + kPH,
              kPOffB3 = kPOffW3 + kPH, kPNPar = kPOffB3 + 1;
constexpr int kPNTrain = kPNPar - kPOffW1;   // trained parameters (W1 .. b3)
constexpr int kLogCols = 10;                 // log record: 8 raw inputs, target, censored

// float -> bfloat16 bits, round to nearest even (finite inputs)
NGPU_HD inline unsigned int bf16_bits(float f) {
  unsigned int u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
NGPU_HD inline float bf16_round(float f) {
  unsigned int u = bf16_bits(f) << 16;
  float r;
  __builtin_memcpy(&r, &u, 4);
  return r;
}

}  // namespace ngpu
