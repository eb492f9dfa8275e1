// The adaptive RK5 update of src/rk5.h:69-207 as HIP kernel templates over
// a neuron model (rk5_models.h): the Cash-Karp attempt, the lane-per-neuron
// kernel, the persistent lane-refill kernel with hot-first scheduling and
// split steps, and the launch helpers.  Each model family instantiates them
// in a file of its own (rk5_*.hip), so the families compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "common.h"
#include "device_common.h"
#include "kernels.h"
#include "predictor_device.h"
#include "rk5_models.h"
#include "stepfactor.h"

namespace ngpu {

// ---------------------------------------------------------------------------
// Cash-Karp RK5(4) coefficients (published tableau, same float rounding as
// the reference's __constant__ float table, src/rk5.cu:35-72).
namespace ck {
constexpr float a21 = (float)0.2;
constexpr float a31 = (float)(3.0 / 40.0), a32 = (float)(9.0 / 40.0);
constexpr float a41 = (float)0.3, a42 = (float)-0.9, a43 = (float)1.2;
constexpr float a51 = (float)(-11.0 / 54.0), a52 = (float)2.5,
                a53 = (float)(-70.0 / 27.0), a54 = (float)(35.0 / 27.0);
constexpr float a61 = (float)(1631.0 / 55296.0), a62 = (float)(175.0 / 512.0),
                a63 = (float)(575.0 / 13824.0), a64 = (float)(44275.0 / 110592.0),
                a65 = (float)(253.0 / 4096.0);
constexpr float b1 = (float)(37.0 / 378.0), b3 = (float)(250.0 / 621.0),
                b4 = (float)(125.0 / 594.0), b6 = (float)(512.0 / 1771.0);
constexpr float e1 = (float)(37.0 / 378.0 - 2825.0 / 27648.0),
                e3 = (float)(250.0 / 621.0 - 18575.0 / 48384.0),
                e4 = (float)(125.0 / 594.0 - 13525.0 / 55296.0),
                e5 = (float)(-277.00 / 14336.0), e6 = (float)(512.0 / 1771.0 - 0.25);
// step factors err^-0.25 (reject) and err^-0.2 (accept): stepfactor.h
constexpr float eps = 1.0e-6f, safety = 0.9f, err_min = 1.889568e-4f, scal_min = 0.1f;
// Bound on RK5 attempts per neuron-step.  Finite runs need at most
// (dt/h_min + 1) substeps x ~67 rejections each (h shrinks by >= 1/0.9 per
// rejection down to h_min), i.e. < 70,000 at h_min = dt/1000; the bound only
// ends the loop of a neuron whose state became non-finite.
constexpr int kMaxAttempts = 100000;
}  // namespace ck
using ck::kMaxAttempts;

// ---------------------------------------------------------------------------
// Neuron models (right-hand sides and post-step updates).
// Parameter and variable indices follow the reference's name tables.  Each
// product-plus-sum of the reference's expressions (a*b + c) is one fma: the
// reference's CUDA build contracts by default, though which products its
// compiler fuses is its own choice; this build fixes this pattern (fewer
// instructions on the attempt chain) and the oracle evaluates the same one.
//
// Every division of a right-hand side is by a parameter (M::divisor(j)).
// dv.div(x, p, j) returns x / p[j].  DivIEEE divides.  DivFast, for
// populations whose divisor parameters are uniform with a usable
// reciprocal, multiplies by r[j] = RN(1/p[j]) and applies Markstein's
// remainder correction, 3 multiply/fma giving the IEEE quotient for
// dividends with magnitude in [2^-64, 2^64) and +0 (div_rn); any other
// dividend (tiny, huge, inf, NaN, -0) only sets `bad`, and rhs() evaluates
// the whole right-hand side again with DivIEEE when a lane of the wave has
// it set: one wave-uniform branch, normally not taken, per evaluation
// instead of a branch per division.  A negated dividend -x/d is written
// -(x/d) (the same IEEE result) so that x = +0 stays on the fast path.
struct DivIEEE {
  __device__ inline float div(float x, const float* p, int j) const { return x / p[j]; }
};
struct DivFast {
  const float* r;
  bool bad;
  __device__ inline float div(float x, const float* p, int j) {
    unsigned int u;
    __builtin_memcpy(&u, &x, sizeof(u));
    bad |= ((u & 0x7fffffffu) - 0x1f800000u >= 0x40000000u) & (u != 0u);
    const float q = x * r[j];
    const float rem = fmaf(-q, p[j], x);
    return fmaf(rem, r[j], q);
  }
};

// aeif_cond_beta_multisynapse, src/aeif_cond_beta_multisynapse_kernel.h:165-226

// ---------------------------------------------------------------------------
// Right-hand side of a population: with uniform divisor parameters (UNI)
// through DivFast, evaluated again with IEEE divisions when a lane of the
// wave had a dividend outside DivFast's exact range.
template <class M, bool UNI>
__device__ inline void rhs(const float* y, float* dy, const float* p, const float* r) {
  if constexpr (UNI) {
    DivFast dv{r, false};
    M::deriv(y, dy, p, dv);
    if (__builtin_expect(__any(dv.bad), 0)) {
      DivIEEE di;
      M::deriv(y, dy, p, di);
    }
  } else {
    DivIEEE di;
    M::deriv(y, dy, p, di);
  }
}

__device__ inline bool lagged_spike(int raw) { return ((unsigned int)raw >> 31) != 0; }

// ---------------------------------------------------------------------------
// One Cash-Karp attempt of step h from (y, k1), src/rk5.h:88-131: the five
// further stages, the 5th-order solution yn and the error ratio
// max_j |h sum_i e_i k_i[j]| / ys[j] (before the division by eps).  Each
// stage's sum_i a_i k_i is accumulated with fma and added to y with one more
// (y + h*(...)), the contracted pattern this build fixes for kernel and oracle.
template <class M, class F>
__device__ inline float rk5_try(const float* y, const float* k1, const float* ys, float h,
                                float* yn, F&& deriv) {
  float k2[M::NV], k3[M::NV], k4[M::NV], k5[M::NV], k6[M::NV];
  const float h21 = h * ck::a21;
#pragma unroll
  for (int j = 0; j < M::NV; ++j) yn[j] = fmaf(h21, k1[j], y[j]);
  deriv(yn, k2);
#pragma unroll
  for (int j = 0; j < M::NV; ++j) yn[j] = fmaf(h, fmaf(ck::a32, k2[j], ck::a31 * k1[j]), y[j]);
  deriv(yn, k3);
#pragma unroll
  for (int j = 0; j < M::NV; ++j) {
    float t = fmaf(ck::a42, k2[j], ck::a41 * k1[j]);
    t = fmaf(ck::a43, k3[j], t);
    yn[j] = fmaf(h, t, y[j]);
  }
  deriv(yn, k4);
#pragma unroll
  for (int j = 0; j < M::NV; ++j) {
    float t = fmaf(ck::a52, k2[j], ck::a51 * k1[j]);
    t = fmaf(ck::a53, k3[j], t);
    t = fmaf(ck::a54, k4[j], t);
    yn[j] = fmaf(h, t, y[j]);
  }
  deriv(yn, k5);
#pragma unroll
  for (int j = 0; j < M::NV; ++j) {
    float t = fmaf(ck::a62, k2[j], ck::a61 * k1[j]);
    t = fmaf(ck::a63, k3[j], t);
    t = fmaf(ck::a64, k4[j], t);
    t = fmaf(ck::a65, k5[j], t);
    yn[j] = fmaf(h, t, y[j]);
  }
  deriv(yn, k6);
  float err = 0.0f;
#pragma unroll
  for (int j = 0; j < M::NV; ++j) {
    float t = fmaf(ck::b3, k3[j], ck::b1 * k1[j]);
    t = fmaf(ck::b4, k4[j], t);
    t = fmaf(ck::b6, k6[j], t);
    yn[j] = fmaf(h, t, y[j]);
    float e = fmaf(ck::e3, k3[j], ck::e1 * k1[j]);
    e = fmaf(ck::e4, k4[j], e);
    e = fmaf(ck::e5, k5[j], e);
    e = fmaf(ck::e6, k6[j], e);
    const float v = (h * e) / ys[j];
    err = fmaxf(err, fabsf(v));
  }
  return err;
}

// ---------------------------------------------------------------------------
// Adaptive RK5 update of one population, src/rk5.h:69-207: one lane per
// neuron, IEEE divisions (the reference's arithmetic operation by operation).
template <class M>
__global__ void __launch_bounds__(256) rk5_kernel(RK5Args a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < a.n;
  unsigned long long n_att = 0, n_rej = 0, n_sub = 0;
  if (active) {
    const int n = a.n;
    float y[M::NV], p[M::NPAR];
#pragma unroll
    for (int j = 0; j < M::NV; ++j) y[j] = a.var[(size_t)j * n + i];
#pragma unroll
    for (int j = 0; j < M::NPAR; ++j)
      p[j] = a.tab.is_per_neuron(j) ? a.param[(size_t)j * n + i] : a.tab.val[j];
    double x = a.x[i];
    float h = a.h[i];
    const double x1 = a.ctx.t1;
    const long long t_idx = a.ctx.t_idx;
    const float h_min = a.h_min;
    const float ref_in = p[M::i_ref];

    bool end_step = false;
    int att_i = 0;
    while (!end_step) {
      const float h_max = (float)(x1 - x);
      // ---- RK5Step ----
      float k1[M::NV];
      float yn[M::NV], ys[M::NV];
      rhs<M, false>(y, k1, p, nullptr);
#pragma unroll
      for (int j = 0; j < M::NV; ++j) ys[j] = fabsf(y[j]) + fabsf(k1[j] * h) + ck::scal_min;
      float err;
      for (;;) {
        if (h > h_max) h = h_max;
        if (h < h_min) h = h_min;
        err = rk5_try<M>(y, k1, ys, h, yn, [&](const float* yy, float* dd) {
          rhs<M, false>(yy, dd, p, nullptr);
        });
        err /= ck::eps;
        ++n_att;
        if (++att_i > kMaxAttempts) {   // non-finite state: end the step (never hit by finite runs)
          x = x1 - h;
          break;
        }
        if (err <= 1.0f || (double)h <= (double)h_min * (1.0 + 1.0e-5)) break;
        ++n_rej;
        const float h_new = h * ck::safety * pow_m025(err);
        h = (float)fmax((double)h_new, 0.1 * (double)h);
      }
      ++n_sub;
      x += h;
      h = (err > ck::err_min) ? h * ck::safety * pow_m02(err) : 5.0f * h;
#pragma unroll
      for (int j = 0; j < M::NV; ++j) y[j] = yn[j];
      // ---- ExternalUpdate ----
      end_step = (x >= x1 - (double)h_min);
      if (y[M::iV] < -1.0e3) {
        y[M::iV] = p[M::i_reset];
        y[M::iw] = 0.0f;
        continue;
      }
      if (y[M::iw] < -1.0e6 || y[M::iw] > 1.0e6) {
        y[M::iV] = p[M::i_reset];
        y[M::iw] = 0.0f;
        continue;
      }
      if (p[M::i_ref] > 0.0f) {
        y[M::iV] = p[M::i_reset];
        if (end_step) p[M::i_ref] -= 1.0f;
      } else if (y[M::iV] >= p[M::i_thr]) {
        emit_spike(a.sl, a.sb, a.ctx.lst, a.i0 + i, i, 1.0f, t_idx, x1);
        y[M::iV] = p[M::i_reset];
        y[M::iw] += p[M::i_jump];
        float r = (float)(int)roundf(p[M::i_tref] / a.dt);
        p[M::i_ref] = r < 0.0f ? 0.0f : r;
      }
    }
#pragma unroll
    for (int j = 0; j < M::NV; ++j) a.var[(size_t)j * n + i] = y[j];
    a.x[i] = x;
    a.h[i] = h;
    if (p[M::i_ref] != ref_in) a.param[(size_t)M::i_ref * n + i] = p[M::i_ref];
  }
  add_wave_counters(a.ctr, (blockIdx.x * blockDim.x + threadIdx.x) / kWave,
                    threadIdx.x & (kWave - 1), n_att, n_rej, n_sub);
}

// ---------------------------------------------------------------------------
// Persistent RK5 update with lane refill (wavefront-ballot compaction).
//
// 99.7% of neuron-steps take a single RK5 attempt, a spiking neuron's step
// ~40.  With one lane per neuron a wave holding one spiking neuron runs ~40
// attempts with 63 idle lanes.  Here each wave owns a contiguous chunk of
// neurons; every iteration is one attempt for all busy lanes (preceded by
// the k1 evaluation for lanes that start a new substep), and lanes whose
// neuron finished the step store it and take the next neurons of the chunk
// (ballot + lane rank; no atomics).  Per neuron the arithmetic is exactly
// that of rk5_kernel.  UNI: every parameter except refractory_step is
// uniform over the population, so parameters come from the kernel-argument
// table and divisions by them use reciprocals (DivRecip, bit-identical).
//
// Step values of one persistent RK5 launch, read where they are used from
// the launch's own kernel arguments (so the compiler reloads them instead of
// holding them in registers): a single-population launch takes them from its
// RK5Args; the multi-population batch kernel from its batch arguments, its
// per-population RK5Args staying in device memory unchanged from step to step.
struct StepFromArgs {
  const RK5Args& a;
  __device__ const StepCtx& ctx() const { return a.ctx; }
  __device__ const StepCtx& pctx() const { return a.pf.ctx; }   // the previous step
  __device__ int prefold() const { return a.prefold; }
  __device__ int defer_cap() const { return a.defer_cap; }
};

// The persistent loop of one wave over its population's share: wave / n_waves
// index the population's waves, ctr_row is the wave's counter row (unique in
// the launch), my_hot / my_spk its LDS buffers of kHotLds / kSpkLds entries,
// s_pred the staged predictor (MODE 2).  Synapses delivered by the wave are
// added to n_dev.
template <class M, bool UNI, int MODE, class SV>
__device__ __attribute__((always_inline)) inline void rk5_persistent_body(
    const RK5Args& a, const SV& sv, const int wave, const int n_waves,
    const int ctr_row, int* my_hot, int* my_spk, PredLDS* s_pred_p, unsigned long long& n_dev) {
  static_assert(MODE == 0 || PredCapable<M>::value, "predictor modes need the AdEx beta model");
  const int lane = threadIdx.x & (kWave - 1);
  const int n = a.n;
  const double x1 = sv.ctx().t1;
  const long long t_idx = sv.ctx().t_idx;
  const int parity = sv.ctx().parity;
  const int lst = sv.ctx().lst;
  // end of the previous step (the end_step_kernel formula one step back): a
  // lagged neuron integrates up to it before starting this step
  const double x1_prev = sv.ctx().t0 + (double)a.dt * (double)sv.ctx().it;
  // the wave's own hot list first (neurons of its chunk that ended the last
  // step about to spike, or split; ~1 per wave, so each takes a lane of its
  // own and the long attempt chains start together in the first iteration),
  // then its chunk, skipping the listed neurons; indices below are relative
  // to the wave's range
  const unsigned char* flag_in = nullptr;
  unsigned char* flag_out = nullptr;
  const int* dense = nullptr;
  int n_hot = 0, q_out = 0;
  if (a.hot_flag) {
    const int q_in = (int)(t_idx % 3);
    q_out = q_in == 2 ? 0 : q_in + 1;
    const int q_clr = q_out == 2 ? 0 : q_out + 1;   // read by the previous launch
    n_hot = a.hot_count[q_in];
    if (n_hot > n) n_hot = n;
    dense = a.hot_list + (size_t)q_in * n;
    flag_in = a.hot_flag + (size_t)parity * n;
    flag_out = a.hot_flag + (size_t)(parity ^ 1) * n;
    if (wave == 0 && lane == 0) a.hot_count[q_clr] = 0;
  }
  // wave w takes hot entries [w*hp, w*hp + hp) (one per lane, so the long
  // attempt chains start together in the first iteration, packed into few
  // waves), then its equal share of all neurons
  const int hl = a.hot_lanes < 1 ? 1 : (a.hot_lanes > kWave ? kWave : a.hot_lanes);
  const int hp = n_hot <= (long long)n_waves * hl
                     ? hl
                     : (int)((((long long)n_hot + n_waves - 1) / n_waves + kWave - 1) & ~(long long)(kWave - 1));
  const long long hs = (long long)wave * hp;
  const int n_hw = hs < n_hot ? (int)((hs + hp < n_hot ? hs + hp : n_hot) - hs) : 0;
  // while the hot entries fill less than half the grid, the waves holding
  // them take no share of the bulk (a wave's bulk refills would stall its
  // long chains on memory); the other waves split all neurons
  const int w_hot = (int)(((long long)n_hot + hp - 1) / hp);
  const bool hot_only = a.hot_only && w_hot > 0 && 2 * w_hot <= n_waves;
  const int n_bulk_waves = hot_only ? n_waves - w_hot : n_waves;
  const long long bwave = hot_only ? (long long)wave - w_hot : (long long)wave;
  const long long bchunk =
      (((long long)n + n_bulk_waves - 1) / n_bulk_waves + kWave - 1) & ~(long long)(kWave - 1);
  const long long b0 = bwave * bchunk;
  const int n_bw = (bwave >= 0 && b0 < n) ? (int)((b0 + bchunk < n ? b0 + bchunk : n) - b0) : 0;
  long long next = 0;
  const long long end = (long long)n_hw + n_bw;
  // the waves holding the chains issue first; the others fill the free slots
  if (n_hw > 0) __builtin_amdgcn_s_setprio(3);
  // spikes of the wave, appended to the step's list with one atomic when the
  // wave exits (a returning atomic per spike would stall the long chains)
  int scnt = 0;         // spikes of this wave (wave-uniform)
  bool spk = false;     // the lane's neuron spiked in this iteration
  int spk_src = 0;
  int hcnt = 0;         // entries of this wave's hot list for the next step
  int pend = 0;         // bit 0: neuron finished this iteration, bit 1: it is hot
  int pend_i = 0;
  const float h_min = a.h_min;
  const float r_eps = recip_rn(ck::eps);
  float y[M::NV], p[M::NPAR], r[M::NPAR], k1[M::NV], ys[M::NV];
  double x = 0.0;
  double cx1 = 0.0;     // end of the step the lane's neuron is in (x1, or x1_prev if lagged)
  bool lagged = false;
  // a lagged neuron's stashed input, loaded when the lane takes it, so
  // resuming a split step needs no memory round trip inside the chain
  double pend_v[M::NPORT];
  int rec_k = 0;        // spike-time fill count of the lane's neuron (recording on)
  float h = 0.0f, ref_in = 0.0f;
  int i = -1;
  int att_i = 0;   // attempts of the current neuron in this step
  bool need_k1 = false;
  float el = 1.0f;      // last accepted error ratio (predictor input)
  float h_sub = 0.0f;   // controller proposal at the start of the substep
  bool clip = false;    // accepted attempt was cut by the end of the step
  bool rej_sub = false; // the current substep had a rejected attempt
  unsigned long long n_att = 0, n_rej = 0, n_sub = 0;
  unsigned int n_split = 0;
#pragma unroll
  for (int j = 0; j < M::NPAR; ++j) {
    p[j] = a.tab.val[j];
    r[j] = UNI ? recip_rn(p[j]) : 0.0f;
  }
  for (;;) {
    // ---- refill idle lanes from the wave's chunk ----
    const bool want = (i < 0) && next < end;
    const unsigned long long m = __ballot(want);
    if (m) {
      if (want) {
        const long long v = next + __popcll(m & ((1ull << lane) - 1ull));
        // (written without a short-circuit on the flag: that form was
        // compiled with the candidate register overwritten by the flag address)
        int cand = -1;
        if (v < end) {
          if (v < n_hw) {
            cand = dense[hs + v];
          } else {
            const int b = (int)(b0 + (v - n_hw));
            const int f = flag_in ? (int)flag_in[b] : 0;
            cand = f ? -1 : b;
          }
        }
        if (cand >= 0) {
          i = cand;
#pragma unroll
          for (int j = 0; j < M::NV; ++j) y[j] = a.var[(size_t)j * n + i];
          if (UNI) {
            p[M::i_ref] = a.param[(size_t)M::i_ref * n + i];
          } else {
#pragma unroll
            for (int j = 0; j < M::NPAR; ++j)
              if (a.tab.is_per_neuron(j)) p[j] = a.param[(size_t)j * n + i];
          }
          ref_in = p[M::i_ref];
          x = a.x[i];
          h = a.h[i];
          if constexpr (MODE != 0) el = a.err_last[i];
          lagged = a.lag ? a.lag[i] != 0 : false;
          if (sv.prefold()) {
            // the previous step's fold of this neuron's channels (the fold
            // kernel's arithmetic): added to the input variables, or kept
            // as the stashed input of a neuron still inside that step
            const FoldArgs& f = a.pf;
#pragma unroll
            for (int k = 0; k < M::NPORT; ++k) {
              const long long lo = (long long)k * n;
              double* rp = f.ring + (size_t)sv.pctx().slot * f.n_ch + (size_t)(f.ch_base + lo + i);
              double acc = poisson_input(f, lo + i, *rp, sv.pctx().t1, sv.pctx().t_idx,
                                         f.pu_b < lo + n && f.pu_e > lo);
              *rp = 0.0;
              if (lagged) {
                pend_v[k] = acc;
              } else if (acc != 0.0) {
                const int wj = M::in_w(k);
                const float w = wj >= 0 ? p[wj < 0 ? 0 : wj] : 1.0f;
                y[M::in_y(k)] = fold_add(y[M::in_y(k)], acc, w);
              }
            }
          } else if (lagged) {
#pragma unroll
            for (int k = 0; k < M::NPORT; ++k) pend_v[k] = a.pend[(size_t)k * n + i];
          }
          if (a.sb.max_rec > 0) rec_k = a.sb.n_rec[i];
          cx1 = lagged ? x1_prev : x1;
          need_k1 = true;
          att_i = 0;
          // the loads complete here: left pending, the compiler waits for them
          // where the paths join (the k1 evaluation of every later
          // iteration), and that wait, counting stores too, would hold the
          // chain on the stores of finished neurons
#pragma unroll
          for (int j = 0; j < M::NV; ++j) asm volatile("" ::"v"(y[j]));
          if (UNI) {
            asm volatile("" ::"v"(p[M::i_ref]));
          } else {
#pragma unroll
            for (int j = 0; j < M::NPAR; ++j)
              if (a.tab.is_per_neuron(j)) asm volatile("" ::"v"(p[j]));
          }
          asm volatile("" ::"v"(x), "v"(h));
          if constexpr (MODE != 0) asm volatile("" ::"v"(el));
          if (lagged) {
#pragma unroll
            for (int k = 0; k < M::NPORT; ++k) asm volatile("" ::"v"(pend_v[k]));
          }
          if (a.sb.max_rec > 0) asm volatile("" ::"v"(rec_k));
        }
      }
      next += __popcll(m);
    }
    // all lanes idle with the chunk not exhausted (every candidate was on
    // the hot list): the masked iteration below does nothing, refill again
    if (!__any(i >= 0) && next >= end) break;
    if constexpr (MODE == 2) {
      // learned proposal for every lane starting a substep whose controller
      // proposal ends inside the step (one cut by the step end, the quiet
      // neuron's usual single substep, is left alone); whole wave on the
      // matrix cores, lanes without a neuron feed zeros
      const bool start = (i >= 0) && need_k1 && h < (float)(cx1 - x);
      if (__any(start)) {
        float raw[kPI];
        if (i >= 0) {
          pred_raw_inputs<M::NPORT>(y, h, el, a.dt, raw);
        } else {
#pragma unroll
          for (int k = 0; k < kPI; ++k) raw[k] = 0.0f;
        }
        const float o = pred_infer_wave(*s_pred_p, raw);
        // the learned proposal only shortens the controller's (the error
        // scale |y| + |k1 h| + 0.1 grows with h, so a longer step than the
        // controller's could pass the test less accurately); floored at
        // 2 h_min, since a step at h_min is accepted whatever its error
        if (start && o > -60.0f && o < 60.0f)
          h = fminf(h, fmaxf(a.dt * exp2f(o), 2.0f * h_min));
      }
    }
    if (i >= 0) {
      // ---- start of a substep: k1 and the error scale (src/rk5.h:83-86) ----
      if (need_k1) {
        if constexpr (MODE == 1) {
          h_sub = h;
          rej_sub = false;
        }
        rhs<M, UNI>(y, k1, p, r);
#pragma unroll
        for (int j = 0; j < M::NV; ++j) ys[j] = fabsf(y[j]) + fabsf(k1[j] * h) + ck::scal_min;
        need_k1 = false;
      }
      // ---- one attempt (src/rk5.h:88-131) ----
      const float h_max = (float)(cx1 - x);
      clip = false;
      if (h > h_max) {
        h = h_max;
        clip = true;
      }
      if (h < h_min) h = h_min;
      float yn[M::NV];
      float err = rk5_try<M>(y, k1, ys, h, yn, [&](const float* yy, float* dd) {
        rhs<M, UNI>(yy, dd, p, r);
      });
      err = div_rn(err, ck::eps, r_eps);
      ++n_att;
      if (++att_i > kMaxAttempts) {   // non-finite state: end the step (never hit by finite runs)
        x = cx1;
        err = 0.0f;
      }
      // the step factor of either branch: with a predictor (whose kernel
      // already holds more registers) evaluated once for the wave; the plain
      // kernel keeps one power per branch (the merged form costs it a
      // register spill in the refill)
      float pw = 0.0f;
      if constexpr (MODE != 0)
        pw = pow_step(err, err <= 1.0f || (double)h <= (double)h_min * (1.0 + 1.0e-5));
      if (!(err <= 1.0f || (double)h <= (double)h_min * (1.0 + 1.0e-5))) {
        ++n_rej;   // rejected: shrink h, retry with the same k1 and scale
        if constexpr (MODE == 1) rej_sub = true;
        const float h_new = h * ck::safety * (MODE == 0 ? pow_m025(err) : pw);
        h = (float)fmax((double)h_new, 0.1 * (double)h);
      } else {
        ++n_sub;   // accepted (src/rk5.h:135-159), then ExternalUpdate
        if constexpr (MODE == 1) {
          if (i % a.log_stride == 0) {
            float raw[kPI];
            pred_raw_inputs<M::NPORT>(y, h_sub, el, a.dt, raw);
            const unsigned long long pos = atomicAdd(a.log_count, 1ull);
            if (pos < (unsigned long long)a.log_cap) {
              float* rec = a.log_buf + pos * kLogCols;
#pragma unroll
              for (int k = 0; k < kPI; ++k) rec[k] = raw[k];
              rec[kPI] = __log2f(h / a.dt);
              // lower bound only (censored): h cut by the end of the step,
              // or accepted at the first try (a longer step might have
              // passed too); exact after a rejection, where the controller
              // overshot and the accepted h is what should have been proposed
              rec[kPI + 1] = (clip || !rej_sub) ? 1.0f : 0.0f;
            }
          }
        }
        if constexpr (MODE != 0) el = err;
        if (att_i <= kMaxAttempts) x += h;
        h = (err > ck::err_min) ? h * ck::safety * (MODE == 0 ? pow_m02(err) : pw) : 5.0f * h;
#pragma unroll
        for (int j = 0; j < M::NV; ++j) y[j] = yn[j];
        need_k1 = true;
        const bool end_step = (x >= cx1 - (double)h_min);
        if (y[M::iV] < -1.0e3 || y[M::iw] < -1.0e6 || y[M::iw] > 1.0e6) {
          y[M::iV] = p[M::i_reset];
          y[M::iw] = 0.0f;
        } else if (p[M::i_ref] > 0.0f) {
          y[M::iV] = p[M::i_reset];
          if (end_step) p[M::i_ref] -= 1.0f;
        } else if (y[M::iV] >= p[M::i_thr]) {
          spike_book(a.sb, a.i0 + i, i, 1.0f, lagged ? t_idx - 1 : t_idx, cx1, &rec_k);
          spk = true;
          spk_src = lagged ? (int)((unsigned int)(a.i0 + i) | kSpikeLagBit) : a.i0 + i;
          y[M::iV] = p[M::i_reset];
          y[M::iw] += p[M::i_jump];
          const float rr = (float)(int)roundf(p[M::i_tref] / a.dt);
          p[M::i_ref] = rr < 0.0f ? 0.0f : rr;
        }
        if (end_step && lagged) {
          // the previous step is complete: add its input (computed when the
          // lane took the neuron, or stashed by that step's fold launch; the
          // fold's arithmetic), then go on with this step
#pragma unroll
          for (int k = 0; k < M::NPORT; ++k) {
            const double acc = pend_v[k];
            if (acc != 0.0) {
              const int wj = M::in_w(k);
              const float w = wj >= 0 ? p[wj < 0 ? 0 : wj] : 1.0f;
              y[M::in_y(k)] = fold_add(y[M::in_y(k)], acc, w);
            }
          }
          a.lag[i] = 0;
          lagged = false;
          cx1 = x1;
          att_i = 0;
        } else if (end_step || (sv.defer_cap() > 0 && !lagged && att_i >= sv.defer_cap())) {
          // end of the step, or split: the neuron is stored mid-step with
          // lag = 1 (and listed hot, so the next launch resumes it first)
          if (!end_step) {
            a.lag[i] = 1;
            ++n_split;
          }
#pragma unroll
          for (int j = 0; j < M::NV; ++j) a.var[(size_t)j * n + i] = y[j];
          a.x[i] = x;
          a.h[i] = h;
          if constexpr (MODE != 0) a.err_last[i] = el;
          if (p[M::i_ref] != ref_in) a.param[(size_t)M::i_ref * n + i] = p[M::i_ref];
          pend = 1 | ((!end_step || (p[M::i_ref] <= 0.0f && y[M::iV] >= M::hot_v(p, a.hot_margin))) ? 2 : 0);
          pend_i = i;
          i = -1;
        }
      }
    }
    // spikes of this iteration into the wave's buffer (the list itself when
    // the buffer is full)
    {
      const unsigned long long ms = __ballot(spk);
      if (spk) {
        const int pos = scnt + __popcll(ms & ((1ull << lane) - 1ull));
        if (pos < kSpkLds) {
          my_spk[pos] = spk_src;
        } else {
          spike_list_add(a.sl, lst, spk_src, 1.0f);
          if (a.fd.lag_d0_at_exit && lagged_spike(spk_src)) {
            // its d_int = 0 synapses now (the next launch folds their slot)
            const int src = (int)((unsigned int)spk_src & ~kSpikeLagBit);
            n_dev += deliver_segment(a.fd.st, spike_base_slot((unsigned int)spk_src, t_idx,
                                                              a.fd.st.n_slot),
                                     1.0f, a.fd.st.row_ptr[src], a.fd.st.row_ptr[src + 1], 0, 1, 1);
          }
        }
        spk = false;
      }
      scnt += __popcll(ms);
    }
    // neurons that finished the step: next step's hot list and flags
    if (a.hot_flag) {
      const unsigned long long mh = __ballot(pend == 3);
      if (pend) {
        bool listed = false;
        if (pend == 3) {
          const int pos = hcnt + __popcll(mh & ((1ull << lane) - 1ull));
          if (pos < kHotLds) {
            my_hot[pos] = pend_i;
            listed = true;
          }
        }
        flag_out[pend_i] = listed ? 1 : 0;
        pend = 0;
      }
      hcnt += __popcll(mh);
    }
  }
  {
    // the wave's spikes into the step's list (order within a list does not
    // matter: delivery sums exactly)
    const int listed = scnt < kSpkLds ? scnt : kSpkLds;
    if (listed > 0) {
      int base = 0;
      if (lane == 0) base = atomicAdd(a.sl.count + lst, listed);
      base = __shfl(base, 0, kWave);
      for (int e = lane; e < listed; e += kWave) {
        const int pos = base + e;
        if (pos < a.sl.cap) {
          a.sl.src[(size_t)lst * a.sl.cap + pos] = my_spk[e];
          a.sl.height[(size_t)lst * a.sl.cap + pos] = 1.0f;
        } else {
          atomicAdd(a.sl.overflow, 1);
        }
      }
      if (a.fd.lag_d0_at_exit) {
        // the d_int = 0 synapses of the wave's lagged spikes: their slot
        // (t_idx) is folded by the next launch; the other synapses are
        // delivered with the list by the next launch's idle waves
        for (int e = 0; e < listed; ++e) {
          const int raw = my_spk[e];
          if (!lagged_spike(raw)) continue;
          const int src = (int)((unsigned int)raw & ~kSpikeLagBit);
          n_dev += deliver_segment(a.fd.st, spike_base_slot((unsigned int)raw, t_idx,
                                                            a.fd.st.n_slot),
                                   1.0f, a.fd.st.row_ptr[src], a.fd.st.row_ptr[src + 1], lane,
                                   kWave, 1);
        }
      }
    }
  }
  if (a.hot_flag) {
    // append the wave's list to the next step's (one atomic per wave; the
    // order of the entries only affects scheduling)
    const int listed = hcnt < kHotLds ? hcnt : kHotLds;
    if (listed > 0) {
      int base = 0;
      if (lane == 0) base = atomicAdd(a.hot_count + q_out, listed);
      base = __shfl(base, 0, kWave);
      int* out = a.hot_list + (size_t)q_out * n + base;
      for (int e = lane; e < listed; e += kWave) out[e] = my_hot[e];
      if (lane == 0 && a.ctr) a.ctr[(size_t)ctr_row * kCtrCols + 4] += (unsigned long long)listed;
    }
  }
  if (a.lag && a.ctr) {
    const unsigned long long ns = wave_sum(n_split);
    if (lane == 0 && ns) a.ctr[(size_t)ctr_row * kCtrCols + 5] += ns;
  }
  add_wave_counters(a.ctr, ctr_row, lane, n_att, n_rej, n_sub);
}

// Fused delivery: the previous step's spike list, one spike per wave at a
// time, taken from the last wave of the launch down so that the first waves
// (the hot ones, which end last) take none while there are fewer spikes than
// waves.  Returns the synapses delivered by the lane.
__device__ inline unsigned long long deliver_prev_list(const FusedDeliver& fd, const SpikeList& sl,
                                                       int wave, int n_waves, int lane) {
  unsigned long long n_dev = 0;
  const int pl = fd.prev_lst;
  int n_prev = sl.count[pl];
  if (n_prev > sl.cap) n_prev = sl.cap;
  const int* srcs = sl.src + (size_t)pl * sl.cap;
  const float* hts = sl.height + (size_t)pl * sl.cap;
  for (int sp = n_waves - 1 - wave; sp < n_prev; sp += n_waves) {
    const unsigned int raw = (unsigned int)srcs[sp];
    const int src = (int)(raw & ~kSpikeLagBit);
    const int mode = (fd.skip_lagged_d0 && (raw >> 31)) ? 2 : 0;
    n_dev += deliver_segment(fd.st, spike_base_slot(raw, fd.prev_t_idx, fd.st.n_slot), hts[sp],
                             fd.st.row_ptr[src], fd.st.row_ptr[src + 1], lane, kWave, mode);
  }
  if (wave == 0 && lane == 0 && fd.spikes && n_prev) *fd.spikes += (unsigned long long)n_prev;
  return n_dev;
}

// Clears the count of the spike list the next step writes (with fused
// delivery no delivery launch does: it held the list of step t_idx - 2,
// delivered by the previous launch).
__device__ inline void clear_next_spike_list(const SpikeList& sl, int lst) {
  sl.count[lst == kSpikeLists - 1 ? 0 : lst + 1] = 0;
}

// Register budget of a small model's kernel (at most 6 state variables): the
// plain kernel is held to 128 VGPRs (4 waves per SIMD) and the
// learned-proposal kernel to 168 (3 waves); the compiler would otherwise take
// a few more and drop a wave per SIMD.
template <class M, bool UNI, int MODE>
__global__ void __launch_bounds__(kRK5Block)
__attribute__((amdgpu_waves_per_eu(M::NV > 6 ? 1 : MODE == 2 ? 3 : MODE == 0 ? 4 : 1)))
rk5_persistent_kernel(RK5Args a) {
  __shared__ int s_hot[kRK5Block / kWave][kHotLds];
  __shared__ int s_spk[kRK5Block / kWave][kSpkLds];
  __shared__ PredLDS s_pred;
  if constexpr (MODE == 2) {
    pred_stage(s_pred, a.pred_params);
    __syncthreads();
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * (blockDim.x / kWave) +
                   __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int n_waves = gridDim.x * (blockDim.x / kWave);
  if (a.clear_next_list && blockIdx.x == 0 && threadIdx.x == 0) clear_next_spike_list(a.sl, a.ctx.lst);
  const StepFromArgs sv{a};
  unsigned long long n_dev = 0;   // synapses delivered by this lane (fused delivery)
  rk5_persistent_body<M, UNI, MODE>(a, sv, wave, n_waves, wave, s_hot[threadIdx.x / kWave],
                                    s_spk[threadIdx.x / kWave], &s_pred, n_dev);
  if (a.fd.enable) n_dev += deliver_prev_list(a.fd, a.sl, wave, n_waves, lane);
  if (a.ctr) {
    n_dev = wave_sum(n_dev);
    if (lane == 0 && n_dev) a.ctr[(size_t)wave * kCtrCols + 3] += n_dev;
  }
}

inline int persistent_grid(int n, int max_blocks) {
  const int waves_needed = (n + kWave - 1) / kWave;
  int grid = (waves_needed + kRK5Block / kWave - 1) / (kRK5Block / kWave);
  if (max_blocks > 0 && grid > max_blocks) grid = max_blocks;
  return grid;
}


// Blocks of one persistent RK5 instantiation that are resident on the device
// at once (its register and LDS use decide; every block must be resident,
// since each wave owns a static chunk of the population).
template <class M, bool UNI, int MODE>
static int resident_blocks() {
  static int blocks = 0;
  if (blocks == 0) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rk5_persistent_kernel<M, UNI, MODE>,
                                                     kRK5Block, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu <= 0 || cus <= 0)
      throw std::runtime_error("cannot query the occupancy of the persistent RK5 kernel");
    blocks = per_cu * cus;
  }
  return blocks;
}

template <class M, bool UNI, int MODE>
static void launch_persistent(const RK5Args& a, hipStream_t s) {
  // enough waves to fill the GPU with every block resident, at most one
  // lane per neuron
  int cap = resident_blocks<M, UNI, MODE>();
  if (a.max_blocks > 0 && a.max_blocks < cap) cap = a.max_blocks;
  const int grid = persistent_grid(a.n, cap);
  hipLaunchKernelGGL((rk5_persistent_kernel<M, UNI, MODE>), dim3(grid), dim3(kRK5Block), 0, s, a);
}

template <class M>
static void launch_rk5_t(const RK5Args& a, hipStream_t s) {
  const int block = kRK5Block;
  if (a.persistent) {
    if constexpr (M::NPORT > kFastPort) {
      // many ports: one persistent variant (parameters per lane, IEEE
      // divisions), the same arithmetic as the others
      launch_persistent<M, false, 0>(a, s);
      return;
    } else {
    const uint64_t rest = a.tab.per_neuron[0] & ~(1ull << M::i_ref);
    bool uni = rest == 0 && a.tab.per_neuron[1] == 0;
    for (int j = 0; j < M::NPAR && uni; ++j)
      if (M::divisor(j) && recip_rn(a.tab.val[j]) == 0.0f) uni = false;
    if constexpr (PredCapable<M>::value) {
      if (uni && a.pred_mode == 1) {
        launch_persistent<M, true, 1>(a, s);
        return;
      }
      if (uni && a.pred_mode == 2) {
        launch_persistent<M, true, 2>(a, s);
        return;
      }
    }
    if (uni)
      launch_persistent<M, true, 0>(a, s);
    else
      launch_persistent<M, false, 0>(a, s);
    return;
    }
  }
  const int grid = (a.n + block - 1) / block;
  hipLaunchKernelGGL(rk5_kernel<M>, dim3(grid), dim3(block), 0, s, a);
}

// Launch for a model templated on its port count, NP = 1..kMaxPort.
template <template <int> class M, int NP = 1>
inline void dispatch_ports(int n_port, const RK5Args& a, hipStream_t s) {
  if constexpr (NP <= kMaxPort) {
    if (n_port == NP) {
      launch_rk5_t<M<NP>>(a, s);
      return;
    }
    dispatch_ports<M, NP + 1>(n_port, a, s);
  } else {
    throw std::runtime_error("RK5: port count " + std::to_string(n_port) + " out of range");
  }
}

// Whether the compile-time input layout of model M (input variable and
// weight parameter of each port) is the fold's layout.
template <class M>
inline bool fold_layout_is(int in_var, int in_stride, int w_param, int w_stride) {
  for (int k = 0; k < M::NPORT; ++k) {
    if (M::in_y(k) != in_var + k * in_stride) return false;
    if (M::in_w(k) != (w_param >= 0 ? w_param + k * w_stride : -1)) return false;
  }
  return true;
}

template <template <int> class M, int NP = 1>
inline bool fold_layout_ports(int n_port, int iv, int is, int wp, int ws) {
  if constexpr (NP <= kMaxPort) {
    if (n_port == NP) return fold_layout_is<M<NP>>(iv, is, wp, ws);
    return fold_layout_ports<M, NP + 1>(n_port, iv, is, wp, ws);
  } else {
    return false;
  }
}

}  // namespace ngpu
