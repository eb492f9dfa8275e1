// HIP kernels of the simulation step for gfx950 (MI355X).
//
// One lane per neuron; state and parameters SoA (coalesced 64-lane loads);
// uniform population parameters come from the kernel-argument table; the
// adaptive RK5 loop, the exact iaf_psc_exp update, spike emission into a
// per-step spike list, delivery of the spike list over the source-major
// delay-sorted synapse table into a FP64 input ring, and the end-of-step
// fold of ring + Poisson input into the receiving state variables.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "philox.h"
#include "predictor_device.h"
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
// product-plus-sum of the reference's expressions (a*b + c) is one fma, as
// the reference's compiler contracts them; the oracle evaluates the same
// sequence.
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
template <int NP>
struct AeifCondBeta {
  static constexpr int NV = 2 + 2 * NP, NPAR = 14 + 4 * NP, NPORT = NP;
  enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
         t_ref, ref, den };
  static constexpr int iV = 0, iw = 1, i_reset = V_reset, i_thr = V_peak,
                       i_jump = b, i_tref = t_ref, i_ref = ref;
  static constexpr bool divisor(int j) {
    return j == Delta_T || j == C_m || j == tau_w || (j >= 14 && (j - 14) % 4 >= 1 && (j - 14) % 4 <= 2);
  }
  // membrane potential from which a neuron is scheduled first (spike
  // upswing): V_th + m Delta_T
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[Delta_T], p[V_th]); }
  // input of port k: variable g1_k, weight g0_k (src/aeif_cond_beta_multisynapse.cu:151-159)
  static constexpr int in_y(int k) { return 3 + 2 * k; }
  static constexpr int in_w(int k) { return 14 + 4 * k + 3; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn = 0.0f;
    const float V = (p[ref] > 0) ? p[V_reset] : fminf(y[0], p[V_peak]);
#pragma unroll
    for (int i = 0; i < NP; ++i) I_syn = fmaf(y[2 + 2 * i], p[14 + 4 * i] - V, I_syn);
    const float V_spike = p[Delta_T] * exp_fma(dv.div(V - p[V_th], p, Delta_T));
    dy[0] = (p[ref] > 0)
                ? 0.0f
                : dv.div(fmaf(-p[g_L], V - p[E_L] - V_spike, I_syn) - y[1] + p[I_e], p, C_m);
    dy[1] = dv.div(fmaf(p[a], V - p[E_L], -y[1]), p, tau_w);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      dy[3 + 2 * i] = -dv.div(y[3 + 2 * i], p, 14 + 4 * i + 1);
      dy[2 + 2 * i] = y[3 + 2 * i] - dv.div(y[2 + 2 * i], p, 14 + 4 * i + 2);
    }
  }
};

// aeif_cond_alpha_multisynapse, src/aeif_cond_alpha_multisynapse_kernel.h:163-226
template <int NP>
struct AeifCondAlpha {
  static constexpr int NV = 2 + 2 * NP, NPAR = 14 + 3 * NP, NPORT = NP;
  enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
         t_ref, ref, den };
  static constexpr int iV = 0, iw = 1, i_reset = V_reset, i_thr = V_peak,
                       i_jump = b, i_tref = t_ref, i_ref = ref;
  static constexpr bool divisor(int j) {
    return j == Delta_T || j == C_m || j == tau_w || (j >= 14 && (j - 14) % 3 == 1);
  }
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[Delta_T], p[V_th]); }
  // input of port k: variable g1_k, weight g0_k
  static constexpr int in_y(int k) { return 3 + 2 * k; }
  static constexpr int in_w(int k) { return 14 + 3 * k + 2; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn = 0.0f;
    const float V = (p[ref] > 0) ? p[V_reset] : fminf(y[0], p[V_peak]);
#pragma unroll
    for (int i = 0; i < NP; ++i) I_syn = fmaf(y[2 + 2 * i], p[14 + 3 * i] - V, I_syn);
    const float V_spike = p[Delta_T] * exp_fma(dv.div(V - p[V_th], p, Delta_T));
    dy[0] = (p[ref] > 0)
                ? 0.0f
                : dv.div(fmaf(-p[g_L], V - p[E_L] - V_spike, I_syn) - y[1] + p[I_e], p, C_m);
    dy[1] = dv.div(fmaf(p[a], V - p[E_L], -y[1]), p, tau_w);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      dy[3 + 2 * i] = -dv.div(y[3 + 2 * i], p, 14 + 3 * i + 1);
      dy[2 + 2 * i] = y[3 + 2 * i] - dv.div(y[2 + 2 * i], p, 14 + 3 * i + 1);
    }
  }
};

// aeif_psc_exp, src/aeif_psc_exp_kernel.h:145-201 (2 fixed ports, I_ex - I_in)
struct AeifPscExp {
  static constexpr int NV = 4, NPAR = 16, NPORT = 2;
  enum { I_ex, I_in, V_m, w };
  // input of port k: I_syn_ex / I_syn_in, weight 1
  static constexpr int in_y(int k) { return k; }
  static constexpr int in_w(int) { return -1; }
  enum { tau_ex, tau_in, V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak,
         V_reset, t_ref, ref, den };
  static constexpr int iV = V_m, iw = w, i_reset = V_reset, i_thr = V_peak,
                       i_jump = b, i_tref = t_ref, i_ref = ref;
  static constexpr bool divisor(int j) {
    return j == tau_ex || j == tau_in || j == Delta_T || j == C_m || j == tau_w;
  }
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[Delta_T], p[V_th]); }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn_tot = 0.0f;
    I_syn_tot += y[I_ex] - y[I_in];
    const float V = (p[ref] > 0) ? p[V_reset] : fminf(y[V_m], p[V_peak]);
    const float V_spike = p[Delta_T] == 0.0f
                              ? 0.0f
                              : p[Delta_T] * exp_fma(dv.div(V - p[V_th], p, Delta_T));
    dy[V_m] = (p[ref] > 0)
                  ? 0.0f
                  : dv.div(fmaf(-p[g_L], V - p[E_L] - V_spike, I_syn_tot) - y[w] + p[I_e], p, C_m);
    dy[w] = dv.div(fmaf(p[a], V - p[E_L], -y[w]), p, tau_w);
    dy[I_ex] = -dv.div(y[I_ex], p, tau_ex);
    dy[I_in] = -dv.div(y[I_in], p, tau_in);
  }
};

// izhikevich_cond_beta, src/izhikevich_cond_beta_kernel.h:152-213
template <int NP>
struct IzhCondBeta {
  static constexpr int NV = 2 + 2 * NP, NPAR = 9 + 4 * NP, NPORT = NP;
  enum { V_th, a, b, c, d, I_e, t_ref, ref, den };
  // input of port k: variable g1_k, weight g0_k
  static constexpr int in_y(int k) { return 3 + 2 * k; }
  static constexpr int in_w(int k) { return 9 + 4 * k + 3; }
  static constexpr int iV = 0, iw = 1, i_reset = c, i_thr = V_th, i_jump = d,
                       i_tref = t_ref, i_ref = ref;
  static constexpr bool divisor(int j) { return j >= 9 && (j - 9) % 4 >= 1 && (j - 9) % 4 <= 2; }
  __device__ static inline float hot_v(const float*, float) { return -45.0f; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn = 0.0f;
    const float V = (p[ref] > 0) ? p[c] : y[0];
#pragma unroll
    for (int i = 0; i < NP; ++i) I_syn = fmaf(y[2 + 2 * i], p[9 + 4 * i] - V, I_syn);
    // evaluated in double, as the reference's double literals make it
    dy[0] = (p[ref] > 0) ? 0.0f
                         : (float)(0.04 * V * V + 5.0 * V + 140.0 - y[1] + I_syn + p[I_e]);
    dy[1] = p[a] * fmaf(p[b], V, -y[1]);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      dy[3 + 2 * i] = -dv.div(y[3 + 2 * i], p, 9 + 4 * i + 1);
      dy[2 + 2 * i] = y[3 + 2 * i] - dv.div(y[2 + 2 * i], p, 9 + 4 * i + 2);
    }
  }
};

// Models the step-size predictor reads (its raw inputs are V, w and the
// conductance variables of aeif_cond_beta_multisynapse).
template <class M> struct PredCapable : std::false_type {};
template <int NP> struct PredCapable<AeifCondBeta<NP>> : std::true_type {};

// ---------------------------------------------------------------------------
// Spike emission: append to this step's spike list and do the per-node
// bookkeeping of src/spike_buffer.cu:162-217.
// Spike bookkeeping of the emitting node: plain stores and a counter add
// whose result is not used, so nothing here waits for memory.  rec_k: the
// node's spike-time fill count, held by the caller (read when it took the
// node); null: read here.
__device__ inline void spike_book(const SpikeBook& sb, int node, int rel, float height,
                                  long long t_idx, double t1, int* rec_k = nullptr) {
  sb.last_idx[node] = t_idx;
  sb.last_height[node] = height;
  if (sb.spike_count) atomicAdd(&sb.spike_count[rel], 1);
  if (sb.max_rec > 0) {
    const int k = rec_k ? *rec_k : sb.n_rec[rel];
    if (k < sb.max_rec - 1) {
      sb.rec_times[(size_t)rel * sb.max_rec + k] = (float)t1;
      sb.n_rec[rel] = k + 1;
      if (rec_k) *rec_k = k + 1;
    }
  }
}

// Appends one source (node | kSpikeLagBit for a lagged spike) to the spike list.
__device__ inline void spike_list_add(const SpikeList& sl, int parity, int src, float height) {
  int pos = atomicAdd(&sl.count[parity], 1);
  if (pos < sl.cap) {
    sl.src[(size_t)parity * sl.cap + pos] = src;
    sl.height[(size_t)parity * sl.cap + pos] = height;
  } else {
    atomicAdd(sl.overflow, 1);
  }
}

__device__ inline void emit_spike(const SpikeList& sl, const SpikeBook& sb,
                                  int parity, int node, int rel, float height,
                                  long long t_idx, double t1, bool lag = false) {
  spike_list_add(sl, parity, lag ? (int)((unsigned int)node | kSpikeLagBit) : node, height);
  spike_book(sb, node, rel, height, t_idx, t1);
}

// Poisson input of channel idx (= port * n + local neuron) for the step with
// time index t_idx ending at t1, added to acc (the channel's ring value) edge
// by edge, src/poiss_gen.cu:55-86.  Poisson edges are sorted by channel, so
// each lane sums its own edges (no atomics).  Contraction is off: the double
// expressions round exactly as the reference's.
__device__ inline double poisson_input(const FoldArgs& f, long long idx, double acc, double t1,
                                       long long t_idx, bool run_here) {
#pragma clang fp contract(off)
  if (f.pu_e > f.pu_b) {
    if (!run_here) return acc;
    // uniform run: nothing to look up but the generator's own table, read
    // by every lane (a lane outside the run reads the first generator's) so
    // the loads issue together with the caller's, without a branch between
    const bool in_run = idx >= f.pu_b && idx < f.pu_e;
    const int r = in_run ? (int)(idx - f.pu_b) : 0;
    const int g = f.pu_gen0 + f.pu_gstep * r;
    const float origin = f.gen_origin[g], start = f.gen_start[g], stop = f.gen_stop[g];
    const int chunks = f.gen_chunks[g];
    const double e_chunk = f.gen_exp[g];
    // one wait for all five (the compiler would otherwise sink each load to
    // its first use and wait for them one by one)
    asm volatile("" ::"v"(origin), "v"(start), "v"(stop), "v"(chunks), "v"(e_chunk));
    const double t_rel = t1 - origin - f.pu_delay;
    if (in_run & (t_rel >= start) & (t_rel <= stop)) {
      const int cnt = poisson_draw(f.seed, f.pu_stream0 + (unsigned int)r, t_idx, chunks, e_chunk);
      if (cnt > 0) acc += (double)(f.pu_w * (float)cnt);
    }
  } else if (f.pe_off) {
    const int b = f.pe_off[idx], e = f.pe_off[idx + 1];
    for (int k = b; k < e; ++k) {
      const int g = f.pe_gen[k];
      const double t_rel = t1 - f.gen_origin[g] - f.pe_delay[k];
      if (t_rel >= f.gen_start[g] && t_rel <= f.gen_stop[g]) {
        const int cnt = poisson_draw(f.seed, f.pe_stream[k], t_idx, f.gen_chunks[g], f.gen_exp[g]);
        if (cnt > 0) acc += (double)(f.pe_w[k] * (float)cnt);
      }
    }
  }
  return acc;
}

// input = (float)((double)input + acc * port_weight), skipped for acc == 0
// (input + 0*w == input exactly)
__device__ inline float fold_add(float input, double acc, float w) {
#pragma clang fp contract(off)
  return (float)((double)input + acc * (double)w);
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Adds a wave's counters into its own slot of the per-wave counter array
// (slot = global wave index of the launch; no two waves of a launch share a
// slot and launches are stream-ordered, so plain read-add-write suffices and
// no atomic contends on a single address).
__device__ inline void add_wave_counters(unsigned long long* ctr, int slot, int lane,
                                         unsigned long long c0, unsigned long long c1,
                                         unsigned long long c2) {
  if (!ctr) return;
  c0 = wave_sum(c0);
  c1 = wave_sum(c1);
  c2 = wave_sum(c2);
  if (lane == 0 && c0) {
    unsigned long long* s = ctr + (size_t)slot * kCtrCols;
    s[0] += c0;
    s[1] += c1;
    s[2] += c2;
  }
}

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

// ---------------------------------------------------------------------------
// One Cash-Karp attempt of step h from (y, k1), src/rk5.h:88-131: the five
// further stages, the 5th-order solution yn and the error ratio
// max_j |h sum_i e_i k_i[j]| / ys[j] (before the division by eps).  Each
// stage's sum_i a_i k_i is accumulated with fma and added to y with one more
// (y + h*(...) as the reference's compiler contracts it).
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
    const int parity = a.ctx.parity;
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
        emit_spike(a.sl, a.sb, parity, a.i0 + i, i, 1.0f, t_idx, x1);
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
// Register budget of a small model's kernel (at most 6 state variables): the
// plain kernel is held to 128 VGPRs (4 waves per SIMD) and the
// learned-proposal kernel to 168 (3 waves); the compiler would otherwise take
// a few more and drop a wave per SIMD.
template <class M, bool UNI, int MODE>
__global__ void __launch_bounds__(kRK5Block)
__attribute__((amdgpu_waves_per_eu(M::NV > 6 ? 1 : MODE == 2 ? 3 : MODE == 0 ? 4 : 1)))
rk5_persistent_kernel(RK5Args a) {
  static_assert(MODE == 0 || PredCapable<M>::value, "predictor modes need the AdEx beta model");
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * (blockDim.x / kWave) +
                   __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int n_waves = gridDim.x * (blockDim.x / kWave);
  const int n = a.n;
  const double x1 = a.ctx.t1;
  const long long t_idx = a.ctx.t_idx;
  const int parity = a.ctx.parity;
  // end of the previous step (the end_step_kernel formula one step back): a
  // lagged neuron integrates up to it before starting this step
  const double x1_prev = a.ctx.t0 + (double)a.dt * (double)a.ctx.it;
  // the wave's own hot list first (neurons of its chunk that ended the last
  // step about to spike, or split; ~1 per wave, so each takes a lane of its
  // own and the long attempt chains start together in the first iteration),
  // then its chunk, skipping the listed neurons; indices below are relative
  // to the wave's range
  const unsigned char* flag_in = nullptr;
  unsigned char* flag_out = nullptr;
  const int* dense = nullptr;
  int n_hot = 0, q_out = 0;
  __shared__ int s_hot[kRK5Block / kWave][kHotLds];
  int* my_hot = s_hot[threadIdx.x / kWave];
  if (a.hot_flag) {
    const int q_in = (int)(t_idx % 3);
    q_out = q_in == 2 ? 0 : q_in + 1;
    const int q_clr = q_out == 2 ? 0 : q_out + 1;   // read by the previous launch
    n_hot = a.hot_count[q_in];
    if (n_hot > n) n_hot = n;
    dense = a.hot_list + (size_t)q_in * n;
    flag_in = a.hot_flag + (size_t)parity * n;
    flag_out = a.hot_flag + (size_t)(parity ^ 1) * n;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.hot_count[q_clr] = 0;
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
  __shared__ int s_spk[kRK5Block / kWave][kSpkLds];
  int* my_spk = s_spk[threadIdx.x / kWave];
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
  __shared__ PredLDS s_pred;
  if constexpr (MODE == 2) {
    pred_stage(s_pred, a.pred_params);
    __syncthreads();
  }
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
          if (a.prefold) {
            // the previous step's fold of this neuron's channels (the fold
            // kernel's arithmetic): added to the input variables, or kept
            // as the stashed input of a neuron still inside that step
            const FoldArgs& f = a.pf;
#pragma unroll
            for (int k = 0; k < M::NPORT; ++k) {
              const long long lo = (long long)k * n;
              double* rp = f.ring + (size_t)f.ctx.slot * f.n_ch + (size_t)(f.ch_base + lo + i);
              double acc = poisson_input(f, lo + i, *rp, f.ctx.t1, f.ctx.t_idx,
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
        const float o = pred_infer_wave(s_pred, raw);
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
      if (!(err <= 1.0f || (double)h <= (double)h_min * (1.0 + 1.0e-5))) {
        ++n_rej;   // rejected: shrink h, retry with the same k1 and scale
        if constexpr (MODE == 1) rej_sub = true;
        const float h_new = h * ck::safety * pow_m025(err);
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
        h = (err > ck::err_min) ? h * ck::safety * pow_m02(err) : 5.0f * h;
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
        } else if (end_step || (a.defer_cap > 0 && !lagged && att_i >= a.defer_cap)) {
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
        if (pos < kSpkLds) my_spk[pos] = spk_src;
        else spike_list_add(a.sl, parity, spk_src, 1.0f);
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
      if (lane == 0) base = atomicAdd(a.sl.count + parity, listed);
      base = __shfl(base, 0, kWave);
      for (int e = lane; e < listed; e += kWave) {
        const int pos = base + e;
        if (pos < a.sl.cap) {
          a.sl.src[(size_t)parity * a.sl.cap + pos] = my_spk[e];
          a.sl.height[(size_t)parity * a.sl.cap + pos] = 1.0f;
        } else {
          atomicAdd(a.sl.overflow, 1);
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
      if (lane == 0 && a.ctr) a.ctr[(size_t)wave * kCtrCols + 4] += (unsigned long long)listed;
    }
  }
  if (a.lag && a.ctr) {
    const unsigned long long ns = wave_sum(n_split);
    if (lane == 0 && ns) a.ctr[(size_t)wave * kCtrCols + 5] += ns;
  }
  add_wave_counters(a.ctr, wave, lane, n_att, n_rej, n_sub);
}

static int persistent_grid(int n, int max_blocks) {
  const int waves_needed = (n + kWave - 1) / kWave;
  int grid = (waves_needed + kRK5Block / kWave - 1) / (kRK5Block / kWave);
  if (max_blocks > 0 && grid > max_blocks) grid = max_blocks;
  return grid;
}

long long rk5_waves(int n, int persistent, int max_blocks) {
  if (n <= 0) return 0;
  return persistent ? (long long)(kRK5Block / kWave) * persistent_grid(n, max_blocks)
                    : 4LL * ((n + 255) / 256);
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
  const int grid = (a.n + block - 1) / block;
  hipLaunchKernelGGL(rk5_kernel<M>, dim3(grid), dim3(block), 0, s, a);
}

template <template <int> class M>
static void dispatch_ports(int n_port, const RK5Args& a, hipStream_t s) {
  switch (n_port) {
    case 1: launch_rk5_t<M<1>>(a, s); break;
    case 2: launch_rk5_t<M<2>>(a, s); break;
    case 3: launch_rk5_t<M<3>>(a, s); break;
    case 4: launch_rk5_t<M<4>>(a, s); break;
    case 5: launch_rk5_t<M<5>>(a, s); break;
    case 6: launch_rk5_t<M<6>>(a, s); break;
    case 7: launch_rk5_t<M<7>>(a, s); break;
    case 8: launch_rk5_t<M<8>>(a, s); break;
    default: break;
  }
}

int rk5_predictor_capable(int model, int n_port) {
  return model == MODEL_AEIF_COND_BETA_MS && n_port >= 1 && n_port <= kMaxPort;
}

int rk5_supported(int model, int n_port) {
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS:
    case MODEL_AEIF_COND_ALPHA_MS:
    case MODEL_IZH_COND_BETA:
      return n_port >= 1 && n_port <= kMaxPort;
    case MODEL_AEIF_PSC_EXP:
      return n_port == 2;
    default:
      return 0;
  }
}

template <class M>
static bool fold_layout_is(int in_var, int in_stride, int w_param, int w_stride) {
  for (int k = 0; k < M::NPORT; ++k) {
    if (M::in_y(k) != in_var + k * in_stride) return false;
    if (M::in_w(k) != (w_param >= 0 ? w_param + k * w_stride : -1)) return false;
  }
  return true;
}

template <template <int> class M>
static bool fold_layout_ports(int n_port, int iv, int is, int wp, int ws) {
  switch (n_port) {
    case 1: return fold_layout_is<M<1>>(iv, is, wp, ws);
    case 2: return fold_layout_is<M<2>>(iv, is, wp, ws);
    case 3: return fold_layout_is<M<3>>(iv, is, wp, ws);
    case 4: return fold_layout_is<M<4>>(iv, is, wp, ws);
    case 5: return fold_layout_is<M<5>>(iv, is, wp, ws);
    case 6: return fold_layout_is<M<6>>(iv, is, wp, ws);
    case 7: return fold_layout_is<M<7>>(iv, is, wp, ws);
    case 8: return fold_layout_is<M<8>>(iv, is, wp, ws);
    default: return false;
  }
}

bool rk5_fold_layout_matches(int model, int n_port, int in_var, int in_port_stride,
                             int w_param, int w_port_stride) {
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS:
      return fold_layout_ports<AeifCondBeta>(n_port, in_var, in_port_stride, w_param, w_port_stride);
    case MODEL_AEIF_COND_ALPHA_MS:
      return fold_layout_ports<AeifCondAlpha>(n_port, in_var, in_port_stride, w_param, w_port_stride);
    case MODEL_IZH_COND_BETA:
      return fold_layout_ports<IzhCondBeta>(n_port, in_var, in_port_stride, w_param, w_port_stride);
    case MODEL_AEIF_PSC_EXP:
      return n_port == 2 && fold_layout_is<AeifPscExp>(in_var, in_port_stride, w_param, w_port_stride);
    default:
      return false;
  }
}

void launch_rk5(int model, int n_port, const RK5Args& a, hipStream_t s) {
  if (a.n <= 0) return;
  if (a.defer_cap > 0 && (!a.persistent || !a.lag || !a.pend))
    throw std::runtime_error("RK5 split steps need the persistent kernel and the lag/pend buffers");
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS: dispatch_ports<AeifCondBeta>(n_port, a, s); break;
    case MODEL_AEIF_COND_ALPHA_MS: dispatch_ports<AeifCondAlpha>(n_port, a, s); break;
    case MODEL_IZH_COND_BETA: dispatch_ports<IzhCondBeta>(n_port, a, s); break;
    case MODEL_AEIF_PSC_EXP: launch_rk5_t<AeifPscExp>(a, s); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// iaf_psc_exp exact update, src/iaf_psc_exp.cu:85-110.  HBM-bound: 16 B of
// state read and written per neuron-step, coefficients from the uniform
// table.  Contraction is off so the float arithmetic matches the reference
// operation by operation.
__global__ void __launch_bounds__(256) iaf_kernel(IafArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int n = a.n;
  enum { I_ex, I_in, V, REF };
  enum { tau_m, C_m, E_L, I_e, Theta, V_reset, tau_ex, tau_in, t_ref, den,
         P20, P11ex, P11in, P21ex, P21in, P22 };
  auto P = [&](int j) {
    return a.tab.is_per_neuron(j) ? a.param[(size_t)j * n + i] : a.tab.val[j];
  };
  float iex = a.var[(size_t)I_ex * n + i];
  float iin = a.var[(size_t)I_in * n + i];
  float v = a.var[(size_t)V * n + i];
  float ref = a.var[(size_t)REF * n + i];
  if (ref > 0.0f) {
    ref -= 1.0f;
  } else {
    v = v * P(P22) + iex * P(P21ex) + iin * P(P21in) + P(I_e) * P(P20);
  }
  iex *= P(P11ex);
  iin *= P(P11in);
  if (v >= P(Theta)) {
    emit_spike(a.sl, a.sb, a.ctx.parity, a.i0 + i, i, 1.0f, a.ctx.t_idx, a.ctx.t1);
    v = P(V_reset);
    ref = (float)(int)roundf(P(t_ref) / a.dt);
  }
  a.var[(size_t)I_ex * n + i] = iex;
  a.var[(size_t)I_in * n + i] = iin;
  a.var[(size_t)V * n + i] = v;
  a.var[(size_t)REF * n + i] = ref;
}

void launch_iaf(const IafArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(iaf_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// spike_generator, src/spike_generator.cu:49-67 (step index relative to the
// start of the current Simulate call, as the reference passes it_).
__global__ void spike_gen_kernel(SpikeGenArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int b = a.off[i], e = a.off[i + 1];
  if (b == e) return;
  const int is = b + a.next[i];
  if (is < e && (long long)a.t_idx[is] == a.ctx.it) {
    emit_spike(a.sl, a.sb, a.ctx.parity, a.i0 + i, i, a.height[is], a.ctx.t_idx, a.ctx.t1);
    a.next[i]++;
  }
}

void launch_spike_gen(const SpikeGenArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(spike_gen_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

// spike_detector, src/spike_detector.cu:58-79
__global__ void detector_kernel(DetectorArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float sh = a.var[i];                 // input_spike_height
  if (sh != 0.0f) {
    if (a.param[i] == 0.0f) sh = 1.0f;  // hold_spike_height
    a.var[i] = 0.0f;
  }
  a.var[(size_t)a.n + i] = sh;         // spike_height
}

void launch_detector(const DetectorArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(detector_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Delivery: every spike of this step is expanded over its source's synapse
// segment (source-major, delay groups ascending) by one 256-lane block; each
// synapse adds height*weight into ring slot (t_idx+1+d_int) mod n_slot at its
// channel.  Replaces the reference's per-node spike ring + send list +
// per-(spike, group) blocks (src/spike_buffer.cu:247-317,
// src/get_spike.cu:60-108): the arrival step t_idx+1+d_int is the step whose
// SpikeBufferUpdate would send the spike in the reference.  FP64 hardware
// atomics accumulate (double)(height*weight), the reference's accumulator
// type, so sums of float products are exact and order-independent.  Each
// lane loads the synapse records of up to 4 synapses before issuing their
// atomics, so a 1000-synapse segment is one round of loads and adds.
constexpr int kDeliverUnroll = 4;

__global__ void __launch_bounds__(256) deliver_kernel(DeliverArgs a) {
  const int parity = a.ctx.parity;
  int n_spk = a.sl.count[parity];
  if (n_spk > a.sl.cap) n_spk = a.sl.cap;
  const int base = (int)((a.ctx.t_idx + 1) % a.n_slot);
  const int* srcs = a.sl.src + (size_t)parity * a.sl.cap;
  const float* hts = a.sl.height + (size_t)parity * a.sl.cap;
  const int stride = blockDim.x * kDeliverUnroll;
  unsigned long long n_ev = 0;
  for (int s = blockIdx.x; s < n_spk; s += gridDim.x) {
    // a lagged spike (emitted in the previous step) arrives one slot earlier
    const unsigned int raw = (unsigned int)srcs[s];
    const int src = (int)(raw & ~kSpikeLagBit);
    int sbase = base - (int)(raw >> 31);
    if (sbase < 0) sbase += a.n_slot;
    const float hgt = hts[s];
    const long long b = a.row_ptr[src], e = a.row_ptr[src + 1];
    n_ev += (unsigned long long)(e - b);
    for (long long k0 = b + threadIdx.x; k0 < e; k0 += stride) {
      unsigned int ch[kDeliverUnroll];
      float w[kDeliverUnroll];
      int d[kDeliverUnroll];
#pragma unroll
      for (int u = 0; u < kDeliverUnroll; ++u) {
        const long long k = k0 + (long long)u * blockDim.x;
        if (k < e) {
          ch[u] = a.syn_ch[k];
          w[u] = a.syn_w[k];
          d[u] = a.syn_d[k];
        } else {
          d[u] = -1;
        }
      }
#pragma unroll
      for (int u = 0; u < kDeliverUnroll; ++u) {
        if (d[u] < 0) continue;
        int slot = sbase + d[u];
        if (slot >= a.n_slot) slot -= a.n_slot;
        unsafeAtomicAdd(&a.ring[(size_t)slot * a.n_ch + ch[u]], (double)(hgt * w[u]));
      }
    }
  }
  if (threadIdx.x == 0) {
    if (a.ctr && n_ev)
      a.ctr[(size_t)blockIdx.x * (blockDim.x / kWave) * kCtrCols + 3] += n_ev;
    if (blockIdx.x == 0) {
      if (a.spikes && n_spk) *a.spikes += (unsigned long long)n_spk;
      // the next step writes the other list: clear its counter now (it was
      // consumed by the previous step's delivery)
      a.sl.count[parity ^ 1] = 0;
    }
  }
}

void launch_deliver(const DeliverArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(deliver_kernel, dim3(kDeliverBlocks), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Fold, src/get_spike.cu:142-166 plus poisson_generator input
// (src/poiss_gen.cu:55-86): for each (port, neuron) channel,
//   acc   = ring[slot][ch] + sum_e (double)(w_e * Poisson(lambda_e))
//   input = (float)((double)input + acc * port_weight)
// and the ring slot is cleared for reuse.  One lane per neuron handles its
// NP ports, so the port and its weight are uniform over the wave (scalar
// loads) and the split flag is read once.  Poisson edges are sorted by target
// channel, so each lane sums its own edges (no atomics).  Contraction is off
// so the double expression rounds exactly as the reference's.
// split flag read by populations without split steps
__device__ const unsigned char kNoLag = 0;

template <int NP>
__global__ void __launch_bounds__(256) fold_kernel(FoldArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  // every load that depends on nothing else is issued first (ring slots,
  // split flag, input variables), and the slots are cleared last, so one
  // wait covers them all; an input is read whatever acc turns out to be: a
  // 128-byte line of it is almost always needed by some lane of the wave
  const bool lagged = *(a.lag ? a.lag + i : &kNoLag) != 0;
  double* rp[NP];
  double acc[NP];
  float* in[NP];
  float in0[NP], w[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    rp[p] = a.ring + (size_t)a.ctx.slot * a.n_ch + (size_t)(a.ch_base + (long long)p * a.n + i);
    acc[p] = *rp[p];
    in[p] = a.var + (size_t)(a.in_var + p * a.in_port_stride) * a.n + i;
    in0[p] = *in[p];
    w[p] = 1.0f;
    if (a.w_param >= 0) {
      const int j = a.w_param + p * a.w_port_stride;
      w[p] = a.tab.is_per_neuron(j) ? a.param[(size_t)j * a.n + i] : a.tab.val[j];
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const long long lo = (long long)p * a.n;
    acc[p] = poisson_input(a, lo + i, acc[p], a.ctx.t1, a.ctx.t_idx,
                           a.pu_b < lo + a.n && a.pu_e > lo);
    *rp[p] = 0.0;
  }
  if (lagged) {   // still inside this step: its RK5 launch adds the input later
#pragma unroll
    for (int p = 0; p < NP; ++p) a.pend[(size_t)p * a.n + i] = acc[p];
    return;
  }
#pragma unroll
  for (int p = 0; p < NP; ++p)
    if (acc[p] != 0.0) *in[p] = fold_add(in0[p], acc[p], w[p]);   // input + 0*w == input exactly
}

void launch_fold(const FoldArgs& a, hipStream_t s) {
  if (a.n <= 0 || a.n_port <= 0) return;
  const dim3 grid((unsigned)((a.n + 255) / 256)), block(256);
  switch (a.n_port) {
    case 1: hipLaunchKernelGGL(fold_kernel<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(fold_kernel<2>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(fold_kernel<3>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(fold_kernel<4>, grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL(fold_kernel<5>, grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL(fold_kernel<6>, grid, block, 0, s, a); break;
    case 7: hipLaunchKernelGGL(fold_kernel<7>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(fold_kernel<8>, grid, block, 0, s, a); break;
    default: throw std::runtime_error("fold: more than " + std::to_string(kMaxPort) + " ports");
  }
}

// ---------------------------------------------------------------------------
// Multimeter row, src/multimeter.cu:76-110: time column then one value per
// column; "spike" reads the height emitted in this step (nodes without
// outgoing connections keep no spike buffer in the reference and read 0,
// src/base_neuron.cu:1205-1234).
__global__ void record_kernel(RecordArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > a.ncol) return;
  const long long row = a.ctx.it - a.row_base;
  if (row < 0 || row >= a.max_rows) return;
  float v;
  if (c == 0) {
    v = (float)a.ctx.t1;
  } else {
    const RecCol rc = a.cols[c - 1];
    if (rc.ptr) {
      v = *rc.ptr;
    } else {
      v = (rc.has_out && a.last_idx[rc.node] == a.ctx.t_idx) ? a.last_height[rc.node] : 0.0f;
    }
  }
  a.buf[(size_t)row * (a.ncol + 1) + c] = v;
}

void launch_record(const RecordArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(record_kernel, dim3((a.ncol + 1 + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace ngpu
