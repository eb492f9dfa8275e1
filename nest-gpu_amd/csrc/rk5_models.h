// Right-hand sides and post-step constants of the RK5 neuron models.
//
// Parameter and variable indices follow the reference's name tables, so
// Set/GetStatus names map to the same indices.  Each product-plus-sum of the
// reference's expressions (a*b + c, a*b - c) is one fma, as the reference's
// compiler contracts them; the oracle evaluates the same sequence.
//
// Every division of a right-hand side is by a parameter (M::divisor(j)).
// dv.div(x, p, j) returns x / p[j] (DivIEEE or DivFast, rk5_device.h).
// A negated dividend -x/d is written -(x/d) (the same IEEE result) so that
// x = +0 stays on DivFast's exact path.
//
// Per model: NV state variables, NPAR parameters, NPORT receptor ports;
// iV/iw the membrane potential and adaptation variable, i_reset/i_thr/
// i_jump/i_tref/i_ref the reset potential, spike threshold, spike-driven
// adaptation jump, refractory time and refractory counter parameters;
// in_y(k)/in_w(k) the variable that receives port k's input and the
// parameter that weights it (-1: weight 1), src/*.cu Init
// (port_input_arr_, port_weight_arr_); hot_v the potential from which a
// neuron is scheduled first in the next step (spike upswing).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "stepfactor.h"

namespace ngpu {

// Parameter layout shared by the AdEx models after their model-specific
// leading block (src/aeif_*_kernel.h ScalParamIndexes).
enum AdexTail { aV_th, aDelta_T, ag_L, aE_L, aC_m, aa, ab, atau_w, aI_e, aV_peak, aV_reset,
                at_ref, aref, aden, kAdexTail };

// dV/dt and dw/dt of the AdEx membrane, with the reference's term order:
// (-g_L (V - E_L - V_spike) + I_1 + I_2 - w + I_e) / C_m where I_1 is the
// first summed input term (fused into the g_L product) and I_2 the rest
// (has_i2 false: "- w" follows the product directly and is fused instead).
// psc: the models that skip the exponential when Delta_T == 0.
template <bool PSC, class D>
__device__ inline void adex_vw(const float* p, float V_m, float w, float i1, float i2, bool has_i1,
                               bool has_i2, float* dV, float* dw, D& dv) {
  const float* q = p;
  const float V = (q[aref] > 0) ? q[aV_reset] : fminf(V_m, q[aV_peak]);
  float V_spike;
  if (PSC) {
    V_spike = q[aDelta_T] == 0.0f ? 0.0f
                                  : q[aDelta_T] * exp_fma(dv.div(V - q[aV_th], q, aDelta_T));
  } else {
    V_spike = q[aDelta_T] * exp_fma(dv.div(V - q[aV_th], q, aDelta_T));
  }
  float num;
  if (has_i1) {
    num = fmaf(-q[ag_L], V - q[aE_L] - V_spike, i1);
    if (has_i2) num = num + i2;
    num = num - w + q[aI_e];
  } else {
    num = fmaf(-q[ag_L], V - q[aE_L] - V_spike, -w) + q[aI_e];
  }
  *dV = (q[aref] > 0) ? 0.0f : dv.div(num, q, aC_m);
  *dw = dv.div(fmaf(q[aa], V - q[aE_L], -w), q, atau_w);
}

// Membrane potential of the AdEx models as their Derivatives see it.
__device__ inline float adex_v(const float* q, float V_m) {
  return (q[aref] > 0) ? q[aV_reset] : fminf(V_m, q[aV_peak]);
}

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

// Two-port conductance models with scalar ex/in names: aeif_cond_alpha
// (src/aeif_cond_alpha_kernel.h:42-196; BETA false: one time constant per
// port) and aeif_cond_beta (src/aeif_cond_beta_kernel.h:42-200; BETA true:
// rise and decay time constants).  State [g_ex, g_in, g1_ex, g1_in, V_m, w];
// port k's input goes to g1_ex/g1_in weighted by g0_ex/g0_in.
template <bool BETA>
struct AeifCond2 {
  static constexpr int NV = 6, NPORT = 2;
  static constexpr int LEAD = BETA ? 8 : 6;   // model-specific leading parameters
  static constexpr int NPAR = LEAD + kAdexTail;
  enum { g_ex, g_in, g1_ex, g1_in, V_m, w };
  // leading parameters: g0_ex, g0_in, E_rev_ex, E_rev_in, then
  // tau_syn_ex, tau_syn_in (alpha) or tau_rise_ex, tau_rise_in,
  // tau_decay_ex, tau_decay_in (beta)
  enum { g0_ex, g0_in, E_rev_ex, E_rev_in, tau_a_ex, tau_a_in };
  static constexpr int tau_r(int k) { return 4 + k; }
  static constexpr int tau_d(int k) { return (BETA ? 6 : 4) + k; }
  static constexpr int iV = V_m, iw = w, i_reset = LEAD + aV_reset, i_thr = LEAD + aV_peak,
                       i_jump = LEAD + ab, i_tref = LEAD + at_ref, i_ref = LEAD + aref;
  static constexpr bool divisor(int j) {
    return (j >= 4 && j < LEAD) || j == LEAD + aDelta_T || j == LEAD + aC_m || j == LEAD + atau_w;
  }
  __device__ static inline float hot_v(const float* p, float m) {
    return fmaf(m, p[LEAD + aDelta_T], p[LEAD + aV_th]);
  }
  static constexpr int in_y(int k) { return g1_ex + k; }
  static constexpr int in_w(int k) { return g0_ex + k; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    const float* q = p + LEAD;
    struct Shift {   // divisions by q[j] = p[LEAD + j]
      D& d;
      const float* p;
      __device__ float div(float x, const float*, int j) { return d.div(x, p, LEAD + j); }
    } sh{dv, p};
    const float V = adex_v(q, y[V_m]);
    const float I_ex = fmaf(y[g_ex], p[E_rev_ex] - V, 0.0f);
    const float I_in = fmaf(y[g_in], p[E_rev_in] - V, 0.0f);
    adex_vw<false>(q, y[V_m], y[w], I_ex, I_in, true, true, &dy[V_m], &dy[w], sh);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      dy[g1_ex + k] = -dv.div(y[g1_ex + k], p, tau_r(k));
      dy[g_ex + k] = y[g1_ex + k] - dv.div(y[g_ex + k], p, tau_d(k));
    }
  }
};
using AeifCondAlpha2 = AeifCond2<false>;
using AeifCondBeta2 = AeifCond2<true>;

// aeif_psc_alpha, src/aeif_psc_alpha_kernel.h:38-200 (2 fixed ports):
// state [I_syn_ex, I_syn_in, I1_syn_ex, I1_syn_in, V_m, w], leading
// parameters tau_syn_ex, tau_syn_in, I0_ex, I0_in; the inhibitory current
// enters with a minus sign (positive weights); input to I1 weighted by I0.
struct AeifPscAlpha {
  static constexpr int NV = 6, NPORT = 2, LEAD = 4, NPAR = LEAD + kAdexTail;
  enum { I_ex, I_in, I1_ex, I1_in, V_m, w };
  enum { tau_ex, tau_in, I0_ex, I0_in };
  static constexpr int iV = V_m, iw = w, i_reset = LEAD + aV_reset, i_thr = LEAD + aV_peak,
                       i_jump = LEAD + ab, i_tref = LEAD + at_ref, i_ref = LEAD + aref;
  static constexpr bool divisor(int j) {
    return j == tau_ex || j == tau_in || j == LEAD + aDelta_T || j == LEAD + aC_m ||
           j == LEAD + atau_w;
  }
  __device__ static inline float hot_v(const float* p, float m) {
    return fmaf(m, p[LEAD + aDelta_T], p[LEAD + aV_th]);
  }
  static constexpr int in_y(int k) { return I1_ex + k; }
  static constexpr int in_w(int k) { return I0_ex + k; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    const float* q = p + LEAD;
    struct Shift {
      D& d;
      const float* p;
      __device__ float div(float x, const float*, int j) { return d.div(x, p, LEAD + j); }
    } sh{dv, p};
    float I_syn_tot = 0.0f;
    I_syn_tot += y[I_ex] - y[I_in];
    adex_vw<true>(q, y[V_m], y[w], I_syn_tot, 0.0f, true, false, &dy[V_m], &dy[w], sh);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      dy[I1_ex + k] = -dv.div(y[I1_ex + k], p, tau_ex + k);
      dy[I_ex + k] = y[I1_ex + k] - dv.div(y[I_ex + k], p, tau_ex + k);
    }
  }
};

// aeif_psc_alpha_multisynapse, src/aeif_psc_alpha_multisynapse_kernel.h:150-220:
// state [V_m, w, (I_syn, I1_syn) x P], parameters 14 + (tau_syn, I0) x P;
// every port's current adds (signed weights); input to I1_syn weighted by I0.
template <int NP>
struct AeifPscAlphaMS {
  static constexpr int NV = 2 + 2 * NP, NPAR = kAdexTail + 2 * NP, NPORT = NP;
  static constexpr int iV = 0, iw = 1, i_reset = aV_reset, i_thr = aV_peak, i_jump = ab,
                       i_tref = at_ref, i_ref = aref;
  static constexpr bool divisor(int j) {
    return j == aDelta_T || j == aC_m || j == atau_w || (j >= kAdexTail && (j - kAdexTail) % 2 == 0);
  }
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[aDelta_T], p[aV_th]); }
  static constexpr int in_y(int k) { return 3 + 2 * k; }
  static constexpr int in_w(int k) { return kAdexTail + 2 * k + 1; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn_tot = 0.0f;
#pragma unroll
    for (int i = 0; i < NP; ++i) I_syn_tot += y[2 + 2 * i];
    adex_vw<true>(p, y[0], y[1], I_syn_tot, 0.0f, true, false, &dy[0], &dy[1], dv);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      dy[3 + 2 * i] = -dv.div(y[3 + 2 * i], p, kAdexTail + 2 * i);
      dy[2 + 2 * i] = y[3 + 2 * i] - dv.div(y[2 + 2 * i], p, kAdexTail + 2 * i);
    }
  }
};

// aeif_psc_exp_multisynapse, src/aeif_psc_exp_multisynapse_kernel.h:150-210:
// state [V_m, w, I_syn x P], parameters 14 + tau_syn x P; input to I_syn,
// weight 1.
template <int NP>
struct AeifPscExpMS {
  static constexpr int NV = 2 + NP, NPAR = kAdexTail + NP, NPORT = NP;
  static constexpr int iV = 0, iw = 1, i_reset = aV_reset, i_thr = aV_peak, i_jump = ab,
                       i_tref = at_ref, i_ref = aref;
  static constexpr bool divisor(int j) {
    return j == aDelta_T || j == aC_m || j == atau_w || j >= kAdexTail;
  }
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[aDelta_T], p[aV_th]); }
  static constexpr int in_y(int k) { return 2 + k; }
  static constexpr int in_w(int) { return -1; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    float I_syn_tot = 0.0f;
#pragma unroll
    for (int i = 0; i < NP; ++i) I_syn_tot += y[2 + i];
    adex_vw<true>(p, y[0], y[1], I_syn_tot, 0.0f, true, false, &dy[0], &dy[1], dv);
#pragma unroll
    for (int i = 0; i < NP; ++i) dy[2 + i] = -dv.div(y[2 + i], p, kAdexTail + i);
  }
};

// aeif_psc_delta, src/aeif_psc_delta_kernel.h:130-200: state [V_m, w]; every
// port's input jumps V_m directly (weight 1), applied port after port.
template <int NP>
struct AeifPscDelta {
  static constexpr int NV = 2, NPAR = kAdexTail, NPORT = NP;
  static constexpr int iV = 0, iw = 1, i_reset = aV_reset, i_thr = aV_peak, i_jump = ab,
                       i_tref = at_ref, i_ref = aref;
  static constexpr bool divisor(int j) { return j == aDelta_T || j == aC_m || j == atau_w; }
  __device__ static inline float hot_v(const float* p, float m) { return fmaf(m, p[aDelta_T], p[aV_th]); }
  static constexpr int in_y(int) { return 0; }
  static constexpr int in_w(int) { return -1; }
  template <class D>
  __device__ static inline void deriv(const float* y, float* dy, const float* p, D& dv) {
    adex_vw<true>(p, y[0], y[1], 0.0f, 0.0f, false, false, &dy[0], &dy[1], dv);
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

}  // namespace ngpu
