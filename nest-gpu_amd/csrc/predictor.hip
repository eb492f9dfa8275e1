// Training and test kernels of the learned step-size predictor (predictor.h).
//
// Training minimises the mean censored squared error of log2(h/dt) over
// (state, accepted h) records logged by the plain RK5 kernel.  The forward
// pass evaluates what inference computes (bfloat16 inputs, W1, W2 and
// layer-1 activations, FP32 accumulation); gradients pass straight through
// the rounding to the FP32 master weights, which Adam updates.  Gradient
// sums are deterministic: every block (one wavefront) stages the per-sample
// vectors of 64 samples in LDS and lane L accumulates parameters L, L+64, ...
// over the samples in a fixed order, blocks write rows of partial sums, and
// the update kernel adds the rows in block order.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "philox.h"
#include "predictor_device.h"
#include "predictor_kernels.h"

namespace ngpu {

// ---- normalisation: per-block sums of x and x^2 of the 8 raw inputs --------
__global__ void __launch_bounds__(256) pred_stats_kernel(const float* log, long long n,
                                                         double* part) {
  __shared__ double red[256];
  double s[kPI], q[kPI];
#pragma unroll
  for (int k = 0; k < kPI; ++k) s[k] = q[k] = 0.0;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < kPI; ++k) {
      const double v = log[r * kLogCols + k];
      s[k] += v;
      q[k] += v * v;
    }
  }
  for (int k = 0; k < 2 * kPI; ++k) {
    red[threadIdx.x] = k < kPI ? s[k] : q[k - kPI];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 2 * kPI + k] = red[0];
    __syncthreads();
  }
}

void launch_pred_stats(const float* log, long long n, double* part, int grid, hipStream_t s) {
  hipLaunchKernelGGL(pred_stats_kernel, dim3(grid), dim3(256), 0, s, log, n, part);
}

// ---- forward + backward, one wavefront per block ----------------------------
struct GradLDS {
  float w1q[kPH * kPI], w2q[kPH * kPH], b1[kPH], b2[kPH], w3[kPH], mu[kPI], is[kPI];
  float b3;
  float xq[64][kPI], h1q[64][kPH], dz1[64][kPH], dz2[64][kPH], h2[64][kPH], dout[64];
};

__global__ void __launch_bounds__(64) pred_grad_kernel(PredGradArgs a) {
  __shared__ GradLDS L;
  const int lane = threadIdx.x;
  const float* P = a.params;
  for (int k = lane; k < kPH * kPH; k += 64) {
    L.w2q[k] = bf16_round(P[kPOffW2 + k]);
    if (k < kPH * kPI) L.w1q[k] = bf16_round(P[kPOffW1 + k]);
    if (k < kPH) {
      L.b1[k] = P[kPOffB1 + k];
      L.b2[k] = P[kPOffB2 + k];
      L.w3[k] = P[kPOffW3 + k];
    }
    if (k < kPI) {
      L.mu[k] = P[kPOffMu + k];
      L.is[k] = P[kPOffIs + k];
    }
  }
  if (lane == 0) L.b3 = P[kPOffB3];
  __syncthreads();

  constexpr int kOwn = (kPNTrain + 63) / 64;
  float acc[kOwn];
#pragma unroll
  for (int k = 0; k < kOwn; ++k) acc[k] = 0.0f;
  float loss = 0.0f;
  const float inv_b = 1.0f / (float)a.batch;

  for (int rd = 0; rd < a.rounds; ++rd) {
    const long long slot = ((long long)blockIdx.x * a.rounds + rd) * 64 + lane;
    long long src = -1;
    if (a.raw) {
      src = slot < a.n_log ? slot : -1;
    } else if (a.idx) {
      src = slot < a.n_idx ? a.idx[slot] : -1;
    } else {
      const U32x4 ctr = {(uint32_t)slot, (uint32_t)(slot >> 32), (uint32_t)a.step,
                         (uint32_t)(a.step >> 32)};
      const U32x4 o = philox4x32_10(ctr, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
      const unsigned long long r64 = ((unsigned long long)o.x << 32) | o.y;
      src = (long long)(r64 % (unsigned long long)a.n_log);
    }
    float x[kPI], z1[kPH], h1[kPH], z2[kPH], dz2[kPH], dout = 0.0f;
    float t = 0.0f;
    bool cens = false;
    if (src >= 0 && a.raw) {
      const float* rec = a.raw + src * kPI;
#pragma unroll
      for (int i = 0; i < kPI; ++i) x[i] = bf16_round((rec[i] - L.mu[i]) * L.is[i]);
      t = a.target[src];
      cens = a.cens[src] != 0;
    } else if (src >= 0) {
      const float* rec = a.log + src * kLogCols;
#pragma unroll
      for (int i = 0; i < kPI; ++i) x[i] = bf16_round((rec[i] - L.mu[i]) * L.is[i]);
      t = rec[kPI];
      cens = rec[kPI + 1] != 0.0f;
    } else {
#pragma unroll
      for (int i = 0; i < kPI; ++i) x[i] = 0.0f;
    }
#pragma unroll
    for (int o = 0; o < kPH; ++o) {
      float s = L.b1[o];
#pragma unroll
      for (int i = 0; i < kPI; ++i) s += L.w1q[o * kPI + i] * x[i];
      z1[o] = s;
      h1[o] = bf16_round(s > 0.0f ? s : 0.0f);
    }
    float out = L.b3;
#pragma unroll 4
    for (int o = 0; o < kPH; ++o) {
      float s = L.b2[o];
#pragma unroll
      for (int i = 0; i < kPH; ++i) s += L.w2q[o * kPH + i] * h1[i];
      z2[o] = s;
      out += L.w3[o] * (s > 0.0f ? s : 0.0f);
    }
    if (src >= 0) {
      const float e = out - t;
      if (!(cens && e > 0.0f)) {
        loss += e * e;
        dout = 2.0f * e * inv_b;
      }
    }
#pragma unroll
    for (int o = 0; o < kPH; ++o) {
      dz2[o] = z2[o] > 0.0f ? dout * L.w3[o] : 0.0f;
      L.dz2[lane][o] = dz2[o];
      L.h2[lane][o] = z2[o] > 0.0f ? z2[o] : 0.0f;
      L.h1q[lane][o] = h1[o];
    }
#pragma unroll 4
    for (int i = 0; i < kPH; ++i) {
      float s = 0.0f;
#pragma unroll
      for (int o = 0; o < kPH; ++o) s += dz2[o] * L.w2q[o * kPH + i];
      L.dz1[lane][i] = z1[i] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kPI; ++i) L.xq[lane][i] = x[i];
    L.dout[lane] = dout;
    __syncthreads();
    // lane accumulates parameters p = lane + 64 k over the 64 samples
#pragma unroll
    for (int k = 0; k < kOwn; ++k) {
      const int p = lane + 64 * k;
      if (p >= kPNTrain) break;
      float sum = 0.0f;
      if (p < kPH * kPI) {
        const int o = p / kPI, i = p % kPI;
        for (int s = 0; s < 64; ++s) sum += L.dz1[s][o] * L.xq[s][i];
      } else if (p < kPH * kPI + kPH) {
        const int o = p - kPH * kPI;
        for (int s = 0; s < 64; ++s) sum += L.dz1[s][o];
      } else if (p < kPH * kPI + kPH + kPH * kPH) {
        const int q = p - kPH * kPI - kPH, o = q / kPH, i = q % kPH;
        for (int s = 0; s < 64; ++s) sum += L.dz2[s][o] * L.h1q[s][i];
      } else if (p < kPH * kPI + 2 * kPH + kPH * kPH) {
        const int o = p - kPH * kPI - kPH - kPH * kPH;
        for (int s = 0; s < 64; ++s) sum += L.dz2[s][o];
      } else if (p < kPNTrain - 1) {
        const int i = p - (kPH * kPI + 2 * kPH + kPH * kPH);
        for (int s = 0; s < 64; ++s) sum += L.dout[s] * L.h2[s][i];
      } else {
        for (int s = 0; s < 64; ++s) sum += L.dout[s];
      }
      acc[k] += sum;
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kOwn; ++k) {
    const int p = lane + 64 * k;
    if (p < kPNTrain) a.gpart[(size_t)blockIdx.x * kPNTrain + p] = acc[k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
  if (lane == 0) a.lpart[blockIdx.x] = loss;
}

void launch_pred_grad(const PredGradArgs& a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(pred_grad_kernel, dim3(blocks), dim3(64), 0, s, a);
}

// ---- sum of the block rows, then Adam (or only the sum) ---------------------
__global__ void __launch_bounds__(256) pred_update_kernel(PredUpdateArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kPNTrain) {
    float g = 0.0f;
    for (int b = 0; b < a.blocks; ++b) g += a.gpart[(size_t)b * kPNTrain + p];
    if (a.grad_out) a.grad_out[p] = g;
    if (a.lr > 0.0f) {
      const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
      const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
      a.m[p] = m;
      a.v[p] = v;
      const float mh = m / a.bc1, vh = v / a.bc2;
      a.params[kPOffW1 + p] -= a.lr * mh / (sqrtf(vh) + a.eps);
    }
  }
  if (p == 0) {
    float l = 0.0f;
    for (int b = 0; b < a.blocks; ++b) l += a.lpart[b];
    a.loss_out[0] = l / (float)a.batch;
  }
}

void launch_pred_update(const PredUpdateArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(pred_update_kernel, dim3((kPNTrain + 255) / 256), dim3(256), 0, s, a);
}

// ---- inference through the matrix-core path (the one the RK5 kernel uses) ---
__global__ void __launch_bounds__(256) pred_infer_kernel(const float* P, const float* raw,
                                                         long long n, float* out) {
  __shared__ PredLDS s;
  pred_stage(s, P);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x / 64);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long base = wave * 64; base < n; base += stride) {   // uniform per wave
    const long long r = base + lane;
    float in[kPI];
#pragma unroll
    for (int k = 0; k < kPI; ++k) in[k] = r < n ? raw[r * kPI + k] : 0.0f;
    const float o = pred_infer_wave(s, in);
    if (r < n) out[r] = o;
  }
}

void launch_pred_infer(const float* P, const float* raw, long long n, float* out, hipStream_t s) {
  long long waves = (n + 63) / 64;
  int grid = (int)((waves + 3) / 4);
  if (grid > 1024) grid = 1024;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(pred_infer_kernel, dim3(grid), dim3(256), 0, s, P, raw, n, out);
}

// ---- loss and gradient without allocation ------------------------------------
// The partial gradient rows and loss sums of pred_grad_kernel live in this
// fixed array (1.4 MB per GPU, held until the process ends): rows
// [kPredScratchMaxBlocks][kPNTrain], then kPredScratchMaxBlocks loss sums.
// Every block writes its whole row and loss sum and the update kernel reads
// only the rows of the current call, so nothing needs clearing between calls.
constexpr size_t kPredScratchFloats =
    (size_t)kPredScratchMaxBlocks * kPNTrain + kPredScratchMaxBlocks;
static __device__ __align__(256) float scratch[kPredScratchFloats];

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address; the
// symbol's name in host code is not that address).
static float* scratch_ptr() {
  constexpr int kMaxDev = 64;
  static float* ptr[kMaxDev] = {};
  static std::mutex mu;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDev) throw std::runtime_error("device index out of range for the predictor scratch");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptr[dev] = static_cast<float*>(p);
  }
  return ptr[dev];
}

void pred_loss_grad_noalloc(const float* params, const float* raw, const float* target,
                            const unsigned char* cens, int n, float* loss, float* grad,
                            hipStream_t s) {
  if (n < 1 || n > kPredScratchMaxN)
    throw std::runtime_error("predictor loss/gradient without allocation: batch of " +
                             std::to_string(n) + " records outside [1, " +
                             std::to_string(kPredScratchMaxN) + "]");
  float* base = scratch_ptr();
  const int blocks = (n + 63) / 64;
  PredGradArgs ga{};
  ga.params = params;
  ga.n_log = n;
  ga.raw = raw;
  ga.target = target;
  ga.cens = cens;
  ga.rounds = 1;
  ga.batch = n;
  ga.gpart = base;
  ga.lpart = base + (size_t)kPredScratchMaxBlocks * kPNTrain;
  launch_pred_grad(ga, blocks, s);
  PredUpdateArgs ua{};
  ua.params = const_cast<float*>(params);   // lr = 0: read only
  ua.gpart = ga.gpart;
  ua.lpart = ga.lpart;
  ua.blocks = blocks;
  ua.batch = n;
  ua.lr = 0.0f;
  ua.grad_out = grad;
  ua.loss_out = loss;
  launch_pred_update(ua, s);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("predictor loss/gradient launch: ") + hipGetErrorString(e));
}

}  // namespace ngpu
