// TEST INFRASTRUCTURE ONLY.  This file is the CPU oracle: a from-scratch host
// C++ restatement of the NEST GPU (v1.0.8) single-GPU simulation step.  It is
// used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
// leg as the checker and as the reported CPU baseline; the product library
// never links, loads or calls it.
//
// It is templated on the state type: T = float restates the reference float
// path (including its float/double mixing), T = double is the FP64 host
// integrator used to state the error of the float path.
//
// What each part follows (reference paths relative to the nest-gpu tree):
//   RK5 step / update loop ...... src/rk5.h:69-207, tableau src/rk5.cu:35-72
//   aeif_cond_beta_multisynapse . src/aeif_cond_beta_multisynapse_kernel.h:42-226,
//                                 src/aeif_cond_beta_multisynapse.cu:37-171
//   aeif_cond_alpha_multisynapse  src/aeif_cond_alpha_multisynapse_kernel.h:163-226,
//                                 src/aeif_cond_alpha_multisynapse.cu:37-80
//   aeif_psc_exp ................ src/aeif_psc_exp_kernel.h:145-201, src/aeif_psc_exp.cu:37-80
//   izhikevich_cond_beta ........ src/izhikevich_cond_beta_kernel.h:152-213,
//                                 src/izhikevich_cond_beta.cu:37-100
//   iaf_psc_exp ................. src/iaf_psc_exp.cu:68-175
//   propagator_32 ............... src/propagator_stability.cu:31-52
//   spike buffer / send list .... src/spike_buffer.cu:162-317, src/send_spike.cu:45-58
//   delivery + fold ............. src/get_spike.cu:60-166
//   poisson_generator ........... src/poiss_gen.cu:55-86, src/nestgpu.cu:1190-1226
//   spike_generator ............. src/spike_generator.cu:49-67,230-269
//   spike_detector .............. src/spike_detector.cu:58-79
//   connection store ............ src/connect.cu:38-74
//   step order / time ........... src/nestgpu.cu:294-352,359-400,469-635
//   multimeter rows ............. src/multimeter.cu:76-110, src/base_neuron.cu:1205-1234
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "philox_ref.h"

namespace orc {

enum Model {
  M_AEIF_COND_BETA_MS = 0,
  M_AEIF_COND_ALPHA_MS,
  M_AEIF_PSC_EXP,
  M_IZH_COND_BETA,
  M_AEIF_COND_ALPHA,
  M_AEIF_COND_BETA,
  M_AEIF_PSC_ALPHA,
  M_AEIF_PSC_ALPHA_MS,
  M_AEIF_PSC_DELTA,
  M_AEIF_PSC_EXP_MS,
  M_IAF_PSC_EXP,
  M_POISSON_GEN,
  M_SPIKE_GEN,
  M_SPIKE_DETECTOR,
  M_PARROT,
  N_MODEL
};

struct ModelInfo {
  const char* name;
  std::vector<std::string> scal_var, port_var, scal_param, port_param;
  int forced_port;       // -1: user chooses
  bool rk5;
  int input_var;         // index of the receiving var (port block var for
                         // multisynapse models, scalar var otherwise)
  bool input_in_port_block;
  int weight_param;      // port param index of the input weight, -1 => 1.0
  int input_scal_step = 1;    // scalar input var: index step between ports (0: shared)
  bool weight_in_scal = false; // weight_param indexes the scalar block, one per port
  int weight_unit_ports = 1 << 30;  // weight_param < 0: weight 1 below this port, 0 from it
};

static std::vector<std::string> adex_tail() {
  return {"V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w", "I_e",
          "V_peak", "V_reset", "t_ref", "refractory_step", "den_delay"};
}
static std::vector<std::string> join(std::vector<std::string> a, const std::vector<std::string>& b) {
  a.insert(a.end(), b.begin(), b.end());
  return a;
}

static const ModelInfo& info(int m) {
  static const std::vector<ModelInfo> tab = {
      {"aeif_cond_beta_multisynapse",
       {"V_m", "w"},
       {"g", "g1"},
       {"V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w", "I_e",
        "V_peak", "V_reset", "t_ref", "refractory_step", "den_delay"},
       {"E_rev", "tau_rise", "tau_decay", "g0"},
       -1, true, 1, true, 3},
      {"aeif_cond_alpha_multisynapse",
       {"V_m", "w"},
       {"g", "g1"},
       {"V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w", "I_e",
        "V_peak", "V_reset", "t_ref", "refractory_step", "den_delay"},
       {"E_rev", "tau_syn", "g0"},
       -1, true, 1, true, 2},
      {"aeif_psc_exp",
       {"I_syn_ex", "I_syn_in", "V_m", "w"},
       {},
       {"tau_syn_ex", "tau_syn_in", "V_th", "Delta_T", "g_L", "E_L", "C_m",
        "a", "b", "tau_w", "I_e", "V_peak", "V_reset", "t_ref",
        "refractory_step", "den_delay"},
       {},
       2, true, 0, false, -1},
      {"izhikevich_cond_beta",
       {"V_m", "u"},
       {"g", "g1"},
       {"V_th", "a", "b", "c", "d", "I_e", "t_ref", "refractory_step",
        "den_delay"},
       {"E_rev", "tau_rise", "tau_decay", "g0"},
       -1, true, 1, true, 3},
      // src/aeif_cond_alpha_kernel.h:42-110: port k -> g1_ex/g1_in, weight g0_ex/g0_in
      {"aeif_cond_alpha",
       {"g_ex", "g_in", "g1_ex", "g1_in", "V_m", "w"},
       {},
       join({"g0_ex", "g0_in", "E_rev_ex", "E_rev_in", "tau_syn_ex", "tau_syn_in"}, adex_tail()),
       {},
       2, true, 2, false, 0, 1, true},
      // src/aeif_cond_beta_kernel.h:42-110
      {"aeif_cond_beta",
       {"g_ex", "g_in", "g1_ex", "g1_in", "V_m", "w"},
       {},
       join({"g0_ex", "g0_in", "E_rev_ex", "E_rev_in", "tau_rise_ex", "tau_rise_in",
             "tau_decay_ex", "tau_decay_in"}, adex_tail()),
       {},
       2, true, 2, false, 0, 1, true},
      // src/aeif_psc_alpha_kernel.h:42-110: port k -> I1_syn_ex/in, weight I0_ex/in
      {"aeif_psc_alpha",
       {"I_syn_ex", "I_syn_in", "I1_syn_ex", "I1_syn_in", "V_m", "w"},
       {},
       join({"tau_syn_ex", "tau_syn_in", "I0_ex", "I0_in"}, adex_tail()),
       {},
       2, true, 2, false, 2, 1, true},
      // src/aeif_psc_alpha_multisynapse_kernel.h:42-110: input I1_syn, weight I0
      {"aeif_psc_alpha_multisynapse", {"V_m", "w"}, {"I_syn", "I1_syn"}, adex_tail(),
       {"tau_syn", "I0"}, -1, true, 1, true, 1},
      // src/aeif_psc_delta.cu:123-133: every port into V_m, weight 1
      {"aeif_psc_delta", {"V_m", "w"}, {}, adex_tail(), {}, -1, true, 0, false, -1, 0, false},
      // src/aeif_psc_exp_multisynapse_kernel.h:42-100: input I_syn, weight 1
      {"aeif_psc_exp_multisynapse", {"V_m", "w"}, {"I_syn"}, adex_tail(), {"tau_syn"},
       -1, true, 0, true, -1},
      {"iaf_psc_exp",
       {"I_syn_ex", "I_syn_in", "V_m_rel", "refractory_step"},
       {},
       {"tau_m", "C_m", "E_L", "I_e", "Theta_rel", "V_reset_rel", "tau_ex",
        "tau_in", "t_ref", "den_delay", "P20", "P11ex", "P11in", "P21ex",
        "P21in", "P22"},
       {},
       2, false, 0, false, -1},
      {"poisson_generator", {}, {}, {"rate", "origin", "start", "stop"}, {},
       0, false, -1, false, -1},
      {"spike_generator", {}, {}, {}, {}, 0, false, -1, false, -1},
      {"spike_detector", {"input_spike_height", "spike_height"}, {},
       {"hold_spike_height"}, {}, 1, false, 0, false, -1},
      // src/parrot_neuron.cu:32-100: port 0 -> input_spike_height (weight 1),
      // port 1 -> dummy_input (weight 0)
      {"parrot_neuron", {"input_spike_height", "dummy_input", "V"}, {},
       {"hold_spike_height", "den_delay"}, {}, 2, false, 0, false, -1, 1, false, 1},
  };
  return tab.at(m);
}

static int model_by_name(const std::string& s) {
  for (int m = 0; m < N_MODEL; ++m)
    if (s == info(m).name) return m;
  return -1;
}

// ---------------------------------------------------------------------------
// Cash-Karp RK5(4) tableau, src/rk5.cu:35-72.  For T=float each constant is
// the double expression rounded to float, exactly as the reference's
// `__constant__ float` initialisers.
template <class T>
struct Tab {
  static constexpr T c2 = T(0.2), c3 = T(0.3), c4 = T(0.6), c5 = T(1.0),
                     c6 = T(0.875);
  static constexpr T a21 = T(0.2);
  static constexpr T a31 = T(3.0 / 40.0), a32 = T(9.0 / 40.0);
  static constexpr T a41 = T(0.3), a42 = T(-0.9), a43 = T(1.2);
  static constexpr T a51 = T(-11.0 / 54.0), a52 = T(2.5),
                     a53 = T(-70.0 / 27.0), a54 = T(35.0 / 27.0);
  static constexpr T a61 = T(1631.0 / 55296.0), a62 = T(175.0 / 512.0),
                     a63 = T(575.0 / 13824.0), a64 = T(44275.0 / 110592.0),
                     a65 = T(253.0 / 4096.0);
  static constexpr T a71 = T(37.0 / 378.0), a73 = T(250.0 / 621.0),
                     a74 = T(125.0 / 594.0), a76 = T(512.0 / 1771.0);
  static constexpr T e1 = T(37.0 / 378.0 - 2825.0 / 27648.0),
                     e3 = T(250.0 / 621.0 - 18575.0 / 48384.0),
                     e4 = T(125.0 / 594.0 - 13525.0 / 55296.0),
                     e5 = T(-277.00 / 14336.0), e6 = T(512.0 / 1771.0 - 0.25);
  static constexpr T eps = T(1.0e-6), coeff = T(0.9), exp_inc = T(-0.2),
                     exp_dec = T(-0.25), err_min = T(1.889568e-4),
                     scal_min = T(1.0e-1);
};

inline float exp_model(float x);
inline double exp_model(double x);

// ---------------------------------------------------------------------------
// Model right-hand sides.  `y` and `p` follow the reference index layout
// (scalar block, then `port_var`/`port_param` interleaved per port).  A
// product plus a sum (a*b + c) of the reference's expressions is one fused
// multiply-add, as the reference's CUDA compiler contracts them by default
// (no -fmad=false in its build); the CUDA compiler's exact choice of which
// products to fuse is not observable here, so this restatement fixes one.
template <class T>
struct AeifCondBeta {  // src/aeif_cond_beta_multisynapse_kernel.h:165-188
  static void deriv(int np, const T* y, T* dy, const T* p) {
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T I_syn = 0;
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[0], p[V_peak]);
    for (int i = 0; i < np; ++i) I_syn = std::fma(y[2 + 2 * i], p[14 + 4 * i] - V, I_syn);
    T V_spike = p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[0] = (p[ref] > 0) ? T(0)
                         : (std::fma(-p[g_L], V - p[E_L] - V_spike, I_syn) - y[1] +
                            p[I_e]) / p[C_m];
    dy[1] = std::fma(p[a], V - p[E_L], -y[1]) / p[tau_w];
    for (int i = 0; i < np; ++i) {
      dy[3 + 2 * i] = -y[3 + 2 * i] / p[14 + 4 * i + 1];
      dy[2 + 2 * i] = y[3 + 2 * i] - y[2 + 2 * i] / p[14 + 4 * i + 2];
    }
  }
};

template <class T>
struct AeifCondAlpha {  // src/aeif_cond_alpha_multisynapse_kernel.h:163-186
  static void deriv(int np, const T* y, T* dy, const T* p) {
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T I_syn = 0;
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[0], p[V_peak]);
    for (int i = 0; i < np; ++i) I_syn = std::fma(y[2 + 2 * i], p[14 + 3 * i] - V, I_syn);
    T V_spike = p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[0] = (p[ref] > 0) ? T(0)
                         : (std::fma(-p[g_L], V - p[E_L] - V_spike, I_syn) - y[1] +
                            p[I_e]) / p[C_m];
    dy[1] = std::fma(p[a], V - p[E_L], -y[1]) / p[tau_w];
    for (int i = 0; i < np; ++i) {
      dy[3 + 2 * i] = -y[3 + 2 * i] / p[14 + 3 * i + 1];
      dy[2 + 2 * i] = y[3 + 2 * i] - y[2 + 2 * i] / p[14 + 3 * i + 1];
    }
  }
};

template <class T>
struct AeifPscExp {  // src/aeif_psc_exp_kernel.h:145-162
  static void deriv(int, const T* y, T* dy, const T* p) {
    enum { I_ex, I_in, V_m, w };
    enum { tau_ex, tau_in, V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e,
           V_peak, V_reset, t_ref, ref, den };
    T I_syn_tot = 0;
    I_syn_tot += y[I_ex] - y[I_in];
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[V_m], p[V_peak]);
    T V_spike = p[Delta_T] == 0. ? T(0.)
                                 : p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[V_m] = (p[ref] > 0) ? T(0)
                           : (std::fma(-p[g_L], V - p[E_L] - V_spike, I_syn_tot) -
                              y[w] + p[I_e]) / p[C_m];
    dy[w] = std::fma(p[a], V - p[E_L], -y[w]) / p[tau_w];
    dy[I_ex] = -y[I_ex] / p[tau_ex];
    dy[I_in] = -y[I_in] / p[tau_in];
  }
};

template <class T>
struct IzhCondBeta {  // src/izhikevich_cond_beta_kernel.h:152-170
  static void deriv(int np, const T* y, T* dy, const T* p) {
    enum { V_th, a, b, c, d, I_e, t_ref, ref, den };
    T I_syn = 0;
    T V = (p[ref] > 0) ? p[c] : y[0];
    for (int i = 0; i < np; ++i) I_syn = std::fma(y[2 + 2 * i], p[9 + 4 * i] - V, I_syn);
    // The reference evaluates this polynomial in double (double literals)
    // and stores the result in the state type.
    dy[0] = (p[ref] > 0) ? T(0)
                         : T(0.04 * V * V + 5.0 * V + 140.0 - y[1] + I_syn +
                             p[I_e]);
    dy[1] = p[a] * std::fma(p[b], V, -y[1]);
    for (int i = 0; i < np; ++i) {
      dy[3 + 2 * i] = -y[3 + 2 * i] / p[9 + 4 * i + 1];
      dy[2 + 2 * i] = y[3 + 2 * i] - y[2 + 2 * i] / p[9 + 4 * i + 2];
    }
  }
};

// aeif_cond_alpha (src/aeif_cond_alpha_kernel.h:158-176) and aeif_cond_beta
// (src/aeif_cond_beta_kernel.h:160-180): state g_ex, g_in, g1_ex, g1_in,
// V_m, w; leading parameters g0_ex, g0_in, E_rev_ex, E_rev_in and the
// synaptic time constants (alpha: tau_syn_ex, tau_syn_in; beta: tau_rise_ex,
// tau_rise_in, tau_decay_ex, tau_decay_in), then the AdEx block.
template <class T, bool BETA>
struct AeifCond2 {
  static void deriv(int, const T* y, T* dy, const T* p) {
    enum { g_ex, g_in, g1_ex, g1_in, V_m, w };
    const T* q = p + (BETA ? 8 : 6);
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    const T tr_ex = p[4], tr_in = p[5];
    const T td_ex = BETA ? p[6] : p[4], td_in = BETA ? p[7] : p[5];
    T I_syn_ex = 0, I_syn_in = 0;
    T V = (q[ref] > 0) ? q[V_reset] : std::min(y[V_m], q[V_peak]);
    I_syn_ex = std::fma(y[g_ex], p[2] - V, I_syn_ex);
    I_syn_in = std::fma(y[g_in], p[3] - V, I_syn_in);
    T V_spike = q[Delta_T] * exp_model((V - q[V_th]) / q[Delta_T]);
    dy[V_m] = (q[ref] > 0) ? T(0)
                           : (std::fma(-q[g_L], V - q[E_L] - V_spike, I_syn_ex) + I_syn_in -
                              y[w] + q[I_e]) / q[C_m];
    dy[w] = std::fma(q[a], V - q[E_L], -y[w]) / q[tau_w];
    dy[g1_ex] = -y[g1_ex] / tr_ex;
    dy[g_ex] = y[g1_ex] - y[g_ex] / td_ex;
    dy[g1_in] = -y[g1_in] / tr_in;
    dy[g_in] = y[g1_in] - y[g_in] / td_in;
  }
};

// aeif_psc_alpha, src/aeif_psc_alpha_kernel.h:160-180
template <class T>
struct AeifPscAlpha {
  static void deriv(int, const T* y, T* dy, const T* p) {
    enum { I_ex, I_in, I1_ex, I1_in, V_m, w };
    const T* q = p + 4;   // after tau_syn_ex, tau_syn_in, I0_ex, I0_in
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T I_syn_tot = 0;
    T V = (q[ref] > 0) ? q[V_reset] : std::min(y[V_m], q[V_peak]);
    I_syn_tot += y[I_ex] - y[I_in];
    T V_spike = q[Delta_T] == 0. ? T(0.)
                                 : q[Delta_T] * exp_model((V - q[V_th]) / q[Delta_T]);
    dy[V_m] = (q[ref] > 0) ? T(0)
                           : (std::fma(-q[g_L], V - q[E_L] - V_spike, I_syn_tot) - y[w] +
                              q[I_e]) / q[C_m];
    dy[w] = std::fma(q[a], V - q[E_L], -y[w]) / q[tau_w];
    dy[I1_ex] = -y[I1_ex] / p[0];
    dy[I1_in] = -y[I1_in] / p[1];
    dy[I_ex] = y[I1_ex] - y[I_ex] / p[0];
    dy[I_in] = y[I1_in] - y[I_in] / p[1];
  }
};

// aeif_psc_alpha_multisynapse, src/aeif_psc_alpha_multisynapse_kernel.h:157-181
// (state V_m, w, then I_syn, I1_syn per port; port params tau_syn, I0)
template <class T>
struct AeifPscAlphaMS {
  static void deriv(int np, const T* y, T* dy, const T* p) {
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T I_syn_tot = 0;
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[0], p[V_peak]);
    for (int i = 0; i < np; ++i) I_syn_tot += y[2 + 2 * i];
    T V_spike = p[Delta_T] == 0. ? T(0.)
                                 : p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[0] = (p[ref] > 0) ? T(0)
                         : (std::fma(-p[g_L], V - p[E_L] - V_spike, I_syn_tot) - y[1] +
                            p[I_e]) / p[C_m];
    dy[1] = std::fma(p[a], V - p[E_L], -y[1]) / p[tau_w];
    for (int i = 0; i < np; ++i) {
      const T tau = p[14 + 2 * i];
      dy[3 + 2 * i] = -y[3 + 2 * i] / tau;
      dy[2 + 2 * i] = y[3 + 2 * i] - y[2 + 2 * i] / tau;
    }
  }
};

// aeif_psc_exp_multisynapse, src/aeif_psc_exp_multisynapse_kernel.h:157-177
template <class T>
struct AeifPscExpMS {
  static void deriv(int np, const T* y, T* dy, const T* p) {
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T I_syn_tot = 0;
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[0], p[V_peak]);
    for (int i = 0; i < np; ++i) I_syn_tot += y[2 + i];
    T V_spike = p[Delta_T] == 0. ? T(0.)
                                 : p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[0] = (p[ref] > 0) ? T(0)
                         : (std::fma(-p[g_L], V - p[E_L] - V_spike, I_syn_tot) - y[1] +
                            p[I_e]) / p[C_m];
    dy[1] = std::fma(p[a], V - p[E_L], -y[1]) / p[tau_w];
    for (int i = 0; i < np; ++i) dy[2 + i] = -y[2 + i] / p[14 + i];
  }
};

// aeif_psc_delta, src/aeif_psc_delta_kernel.h:166-178 (no synaptic state:
// inputs jump V_m at the fold)
template <class T>
struct AeifPscDelta {
  static void deriv(int, const T* y, T* dy, const T* p) {
    enum { V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
           t_ref, ref, den };
    T V = (p[ref] > 0) ? p[V_reset] : std::min(y[0], p[V_peak]);
    T V_spike = p[Delta_T] == 0. ? T(0.)
                                 : p[Delta_T] * exp_model((V - p[V_th]) / p[Delta_T]);
    dy[0] = (p[ref] > 0) ? T(0)
                         : (std::fma(-p[g_L], V - p[E_L] - V_spike, -y[1]) + p[I_e]) / p[C_m];
    dy[1] = std::fma(p[a], V - p[E_L], -y[1]) / p[tau_w];
  }
};

// Test-only linear model y' = lambda*y (param[0] = lambda).
template <class T>
struct LinearModel {
  static void deriv(int, const T* y, T* dy, const T* p) { dy[0] = p[0] * y[0]; }
};

constexpr int kMaxVar = 2 + 2 * 32;

struct StepStats {
  long long attempts = 0;   // accepted + rejected RK5 attempts
  long long rejected = 0;
  long long substeps = 0;   // accepted attempts
};

// Step factors err^(-1/4) and err^(-1/5) of src/rk5.h:135-153.  The float
// path evaluates them in double from +,-,* only and rounds to float: with
// x = m 2^(nq+r), m in [1,2), x^(-1/n) = 2^(-q) 2^(-r/n) m^(-1/n), where
// m^(-1/n) comes from a quadratic start and three Newton steps
// z <- z (n+1 - m z^n) / n of the inverse root.  The result is the correctly
// rounded float power (pinned against pow in tests/test_oracle.py); the
// FP64 path uses the library pow.
inline double inv_root_ieee(float x, int n) {
  int e;
  double m = std::frexp((double)x, &e);
  m *= 2.0;
  e -= 1;
  int q = e / n, r = e % n;
  if (r < 0) { r += n; q -= 1; }
  static const double c4[4] = {1.0, 0.8408964152537145, 0.7071067811865476, 0.5946035575013605};
  static const double c5[5] = {1.0, 0.8705505632961241, 0.757858283255199, 0.6597539553864471,
                               0.5743491774985174};
  double z;
  if (n == 4) {
    z = 1.2852 + m * (-0.35326 + m * 0.06590);
    for (int k = 0; k < 3; ++k) {
      double z2 = z * z, z4 = z2 * z2;
      z = z * ((5.0 - m * z4) * 0.25);
    }
    return std::ldexp(z * c4[r], -q);
  }
  z = 1.2284 + m * (-0.28173 + m * 0.05166);
  for (int k = 0; k < 3; ++k) {
    double z2 = z * z, z4 = z2 * z2, z5 = z4 * z;
    z = z * ((6.0 - m * z5) * 0.2);
  }
  return std::ldexp(z * c5[r], -q);
}
// e^x of the AdEx exponential term, float path: Cody-Waite reduction
// x = k ln2 + r with ln2 as a 15-bit float plus a float correction, the
// degree-7 Taylor polynomial of e^r in Horner form with fma, and the exact
// scaling by 2^k (one rounding where the result is subnormal).  Within one
// ulp of exp (pinned against the library exp in tests/test_oracle.py).  The
// FP64 path uses the library exp.
inline float exp_model(float x) {
  if (!(x == x)) return x;
  if (x > 100.0f) return (float)INFINITY;
  if (x < -110.0f) return 0.0f;
  const float k = std::rint(x * 1.44269504f);
  float r = std::fma(-k, 0.693145752f, x);
  r = std::fma(-k, 1.42860677e-06f, r);
  const float c[7] = {1.0f / 720.0f, 1.0f / 120.0f, 1.0f / 24.0f, 1.0f / 6.0f, 0.5f, 1.0f, 1.0f};
  float p = 1.0f / 5040.0f;
  for (int i = 0; i < 7; ++i) p = std::fma(p, r, c[i]);
  return std::ldexp(p, (int)k);
}
inline double exp_model(double x) { return std::exp(x); }

inline float step_factor_dec(float err) {
  if (!(err == err)) return err;
  if (err == INFINITY) return 0.0f;
  if (err == 0.0f) return INFINITY;
  return (float)inv_root_ieee(err, 4);
}
inline float step_factor_inc(float err) {
  if (!(err == err)) return err;
  if (err == INFINITY) return 0.0f;
  if (err == 0.0f) return INFINITY;
  return (float)inv_root_ieee(err, 5);
}
inline double step_factor_dec(double err) { return std::pow(err, -0.25); }
inline double step_factor_inc(double err) { return std::pow(err, -0.2); }

// src/rk5.h:69-159.  `nv` state variables, k1 outside the retry loop,
// y_scal computed once from the pre-loop h.
template <class T, class M>
void rk5_step(int np, int nv, double& x, T* y, T& h, T h_min, T h_max,
              const T* p, StepStats* st) {
  using K = Tab<T>;
  T y_new[kMaxVar], k1[kMaxVar], k2[kMaxVar], k3[kMaxVar], k4[kMaxVar],
      k5[kMaxVar], k6[kMaxVar], y_scal[kMaxVar];
  M::deriv(np, y, k1, p);
  for (int i = 0; i < nv; ++i)
    y_scal[i] = std::fabs(y[i]) + std::fabs(k1[i] * h) + K::scal_min;
  T err;
  for (;;) {
    if (h > h_max) h = h_max;
    if (h < h_min) h = h_min;
    // y + h * (sum_i a_i k_i): the sum accumulated left to right with fma
    // from the first product, then one fma for y + h * sum
    const T h21 = h * K::a21;
    for (int i = 0; i < nv; ++i) y_new[i] = std::fma(h21, k1[i], y[i]);
    M::deriv(np, y_new, k2, p);
    for (int i = 0; i < nv; ++i)
      y_new[i] = std::fma(h, std::fma(K::a32, k2[i], K::a31 * k1[i]), y[i]);
    M::deriv(np, y_new, k3, p);
    for (int i = 0; i < nv; ++i) {
      T s = std::fma(K::a42, k2[i], K::a41 * k1[i]);
      s = std::fma(K::a43, k3[i], s);
      y_new[i] = std::fma(h, s, y[i]);
    }
    M::deriv(np, y_new, k4, p);
    for (int i = 0; i < nv; ++i) {
      T s = std::fma(K::a52, k2[i], K::a51 * k1[i]);
      s = std::fma(K::a53, k3[i], s);
      s = std::fma(K::a54, k4[i], s);
      y_new[i] = std::fma(h, s, y[i]);
    }
    M::deriv(np, y_new, k5, p);
    for (int i = 0; i < nv; ++i) {
      T s = std::fma(K::a62, k2[i], K::a61 * k1[i]);
      s = std::fma(K::a63, k3[i], s);
      s = std::fma(K::a64, k4[i], s);
      s = std::fma(K::a65, k5[i], s);
      y_new[i] = std::fma(h, s, y[i]);
    }
    M::deriv(np, y_new, k6, p);
    err = 0.0;
    for (int i = 0; i < nv; ++i) {
      T s = std::fma(K::a73, k3[i], K::a71 * k1[i]);
      s = std::fma(K::a74, k4[i], s);
      s = std::fma(K::a76, k6[i], s);
      y_new[i] = std::fma(h, s, y[i]);
      T e = std::fma(K::e3, k3[i], K::e1 * k1[i]);
      e = std::fma(K::e4, k4[i], e);
      e = std::fma(K::e5, k5[i], e);
      e = std::fma(K::e6, k6[i], e);
      T val = (h * e) / y_scal[i];
      err = std::max(err, (T)std::fabs(val));
    }
    err /= K::eps;
    if (st) st->attempts++;
    // double comparison as in the reference (1.0e-5 is a double literal)
    if (err <= 1.0 || (double)h <= (double)h_min * (1.0 + 1.0e-5)) break;
    if (st) st->rejected++;
    T h_new = h * K::coeff * step_factor_dec(err);
    h = (T)std::max((double)h_new, 0.1 * (double)h);
  }
  if (st) st->substeps++;
  x += h;
  if (err > K::err_min)
    h = h * K::coeff * step_factor_inc(err);
  else
    h = (T)(5.0 * (double)h);
  for (int i = 0; i < nv; ++i) y[i] = y_new[i];
}

// src/propagator_stability.cu:31-52 (FP64): P32 with the singular-case switch
static double propagator_32(double tau_syn, double tau, double C, double h) {
  const double P32_linear = 1.0 / (2.0 * C * tau * tau) * h * h *
                            (tau_syn - tau) * std::exp(-h / tau);
  const double P32_singular = h / C * std::exp(-h / tau);
  const double P32 = -tau / (C * (1.0 - tau / tau_syn)) * std::exp(-h / tau_syn) *
                     std::expm1(h * (1.0 / tau_syn - 1.0 / tau));
  const double dev_P32 = std::fabs(P32 - P32_singular);
  if (tau == tau_syn ||
      (std::fabs(tau - tau_syn) < 0.1 && dev_P32 > 2.0 * std::fabs(P32_linear)))
    return P32_singular;
  return P32;
}

// ---------------------------------------------------------------------------
template <class T>
struct Group {
  int model = 0;
  int i0 = 0, n = 0, n_port = 0;
  int nv = 0, np_ = 0;               // vars / params per node
  std::vector<T> y;                  // AoS [n][nv]
  std::vector<T> p;                  // AoS [n][np_]
  std::vector<double> x;             // RK5 time per node
  std::vector<T> h;                  // RK5 step per node
  float h_min_rel = 1.0e-3f, h0_rel = 1.0e-2f;
  float h_min = 0, h0 = 0;
  std::vector<double> acc;           // get_spike_array [port][n]
  // spike generator
  std::vector<std::vector<float>> sg_times, sg_heights;
  std::vector<std::vector<int>> sg_idx;
  std::vector<int> sg_next;
  // poisson generator direct connections (src/dir_connect.h:30-37)
  struct DirConn { int irel; int target; int port; float weight; float delay;
                   uint32_t stream; };
  std::vector<DirConn> dir;
  std::vector<int> pg_chunks;      // per generator: Poisson chunks
  std::vector<double> pg_exp;      // per generator: exp(-lambda/chunks)
};

struct Record { int node; int group; int rel; std::string var; int port;
                std::vector<float> t, v; };

template <class T>
class Engine {
 public:
  Engine(float dt, uint64_t seed) : dt_(dt), seed_(seed) {}

  int create(const std::string& name, int n, int n_port) {
    if (calibrated_) throw std::runtime_error("Nodes cannot be created after calibration");
    int m = model_by_name(name);
    if (m < 0) throw std::runtime_error("Unknown neuron model name: " + name);
    const ModelInfo& mi = info(m);
    if (mi.forced_port >= 0) n_port = mi.forced_port;
    Group<T> g;
    g.model = m;
    g.i0 = n_nodes_;
    g.n = n;
    g.n_port = n_port;
    g.nv = mi.scal_var.size() + mi.port_var.size() * n_port;
    g.np_ = mi.scal_param.size() + mi.port_param.size() * n_port;
    g.y.assign((size_t)n * g.nv, T(0));
    g.p.assign((size_t)n * g.np_, T(0));
    init_defaults(g);
    if (m == M_SPIKE_GEN) {
      g.sg_times.assign(n, {});
      g.sg_heights.assign(n, {});
      g.sg_idx.assign(n, {});
      g.sg_next.assign(n, 0);
    }
    groups_.push_back(std::move(g));
    for (int i = 0; i < n; ++i) node_group_.push_back(groups_.size() - 1);
    n_nodes_ += n;
    return n_nodes_ - n;
  }

  // Index of a named parameter / variable of group g, or -1.
  int param_index(const Group<T>& g, const std::string& name, int port) const {
    const ModelInfo& mi = info(g.model);
    for (size_t i = 0; i < mi.scal_param.size(); ++i)
      if (mi.scal_param[i] == name) return i;
    for (size_t i = 0; i < mi.port_param.size(); ++i)
      if (mi.port_param[i] == name)
        return mi.scal_param.size() + port * mi.port_param.size() + i;
    return -1;
  }
  int var_index(const Group<T>& g, const std::string& name, int port) const {
    const ModelInfo& mi = info(g.model);
    for (size_t i = 0; i < mi.scal_var.size(); ++i)
      if (mi.scal_var[i] == name) return i;
    for (size_t i = 0; i < mi.port_var.size(); ++i)
      if (mi.port_var[i] == name)
        return mi.scal_var.size() + port * mi.port_var.size() + i;
    return -1;
  }

  void set(int node0, int n, const std::string& name, int port,
           const float* vals) {
    Group<T>& g = groups_.at(node_group_.at(node0));
    if (name == "h_min_rel") { g.h_min_rel = vals[0]; return; }
    if (name == "h0_rel") { g.h0_rel = vals[0]; return; }
    int ip = param_index(g, name, port);
    int iv = var_index(g, name, port);
    if (ip < 0 && iv < 0) throw std::runtime_error("Unrecognized name " + name);
    for (int k = 0; k < n; ++k) {
      int rel = node0 + k - g.i0;
      if (ip >= 0) g.p[(size_t)rel * g.np_ + ip] = (T)vals[k];
      else g.y[(size_t)rel * g.nv + iv] = (T)vals[k];
    }
  }

  double get(int node, const std::string& name, int port) const {
    const Group<T>& g = groups_.at(node_group_.at(node));
    int rel = node - g.i0;
    int ip = param_index(g, name, port);
    if (ip >= 0) return g.p[(size_t)rel * g.np_ + ip];
    int iv = var_index(g, name, port);
    if (iv >= 0) return g.y[(size_t)rel * g.nv + iv];
    throw std::runtime_error("Unrecognized name " + name);
  }

  void set_spike_times(int node, int n, const float* t, const float* hgt) {
    Group<T>& g = groups_.at(node_group_.at(node));
    int rel = node - g.i0;
    g.sg_times[rel].assign(t, t + n);
    if (hgt) g.sg_heights[rel].assign(hgt, hgt + n);
    else g.sg_heights[rel].clear();
  }

  // src/connect.cu:38-54: d_int = round(delay/dt) - 1; delay < dt rejected.
  // stream: for a poisson_generator's connection, the Poisson stream id to
  // use instead of the connection's index in the sorted store (null: the
  // index); lets a test replay part of a larger network with its draws
  void connect(long long n, const int* src, const int* tgt,
               const unsigned char* port, const float* w, const float* delay,
               const long long* stream = nullptr) {
    if (calibrated_) throw std::runtime_error("Connections cannot be created after calibration");
    for (long long i = 0; i < n; ++i) {
      if (delay[i] < dt_) throw std::runtime_error("Delay must be >= time resolution");
      int d_int = (int)std::round(delay[i] / dt_) - 1;
      c_src_.push_back(src[i]);
      c_tgt_.push_back(tgt[i]);
      c_port_.push_back(port[i]);
      c_w_.push_back(w[i]);
      c_d_.push_back(d_int);
      c_stream_.push_back(stream ? stream[i] : -1);
    }
  }

  int record(int node, const std::string& var, int port) {
    Record r;
    r.node = node;
    r.group = node_group_.at(node);
    r.rel = node - groups_[r.group].i0;
    r.var = var;
    r.port = port;
    records_.push_back(r);
    return records_.size() - 1;
  }

  void simulate(float sim_time) {
    if (!calibrated_) calibrate();
    if (first_sim_) {
      write_records(neural_time_);
      first_sim_ = false;
    }
    neur_t0_ = neural_time_;
    long long Nt = (long long)std::round(sim_time / dt_);
    update_poisson_lambda();
    for (long long it = 0; it < Nt; ++it) step(it);
  }

  // public results
  std::vector<Record> records_;
  std::vector<long long> ev_t_;   // spike events: time index
  std::vector<int> ev_node_;      // spike events: node
  StepStats stats_;
  long long n_syn_events_ = 0;    // delivered (spike, synapse) pairs
  long long n_conn() const { return c_src_.size(); }

 private:
  // ---- defaults, src/*.cu NodeInit / Init -------------------------------
  void init_defaults(Group<T>& g) {
    const ModelInfo& mi = info(g.model);
    // defaults are the reference's float constants
    auto P = [&](int rel, const char* nm, int port, double v) {
      g.p[(size_t)rel * g.np_ + param_index(g, nm, port)] = (T)(float)v;
    };
    auto V = [&](int rel, const char* nm, int port, double v) {
      g.y[(size_t)rel * g.nv + var_index(g, nm, port)] = (T)(float)v;
    };
    for (int r = 0; r < g.n; ++r) {
      switch (g.model) {
        case M_AEIF_COND_BETA_MS:
        case M_AEIF_COND_ALPHA_MS:
        case M_AEIF_PSC_EXP:
        case M_AEIF_COND_ALPHA:
        case M_AEIF_COND_BETA:
        case M_AEIF_PSC_ALPHA:
        case M_AEIF_PSC_ALPHA_MS:
        case M_AEIF_PSC_DELTA:
        case M_AEIF_PSC_EXP_MS:
          P(r, "V_th", 0, -50.4); P(r, "Delta_T", 0, 2.0); P(r, "g_L", 0, 30.0);
          P(r, "E_L", 0, -70.6); P(r, "C_m", 0, 281.0); P(r, "a", 0, 4.0);
          P(r, "b", 0, 80.5); P(r, "tau_w", 0, 144.0); P(r, "I_e", 0, 0.0);
          P(r, "V_peak", 0, 0.0); P(r, "V_reset", 0, -60.0); P(r, "t_ref", 0, 0.0);
          P(r, "den_delay", 0, 0.0); P(r, "refractory_step", 0, 0.0);
          V(r, "V_m", 0, -70.6); V(r, "w", 0, 0.0);  // V_m = E_L
          if (g.model == M_AEIF_PSC_EXP || g.model == M_AEIF_PSC_ALPHA) {
            P(r, "tau_syn_ex", 0, 0.2); P(r, "tau_syn_in", 0, 2.0);
          } else if (g.model == M_AEIF_COND_ALPHA) {  // src/aeif_cond_alpha.cu:55-58
            P(r, "E_rev_ex", 0, 0.0); P(r, "E_rev_in", 0, -85.0);
            P(r, "tau_syn_ex", 0, 0.2); P(r, "tau_syn_in", 0, 2.0);
          } else if (g.model == M_AEIF_COND_BETA) {   // src/aeif_cond_beta.cu:56-61
            P(r, "E_rev_ex", 0, 0.0); P(r, "E_rev_in", 0, 0.0);
            P(r, "tau_decay_ex", 0, 20.0); P(r, "tau_decay_in", 0, 20.0);
            P(r, "tau_rise_ex", 0, 2.0); P(r, "tau_rise_in", 0, 2.0);
          } else if (g.model == M_AEIF_PSC_ALPHA_MS || g.model == M_AEIF_PSC_EXP_MS) {
            for (int q = 0; q < g.n_port; ++q) P(r, "tau_syn", q, 0.2);
          } else if (g.model == M_AEIF_COND_BETA_MS || g.model == M_AEIF_COND_ALPHA_MS) {
            for (int q = 0; q < g.n_port; ++q) {
              P(r, "E_rev", q, 0.0);
              if (g.model == M_AEIF_COND_BETA_MS) {
                P(r, "tau_decay", q, 20.0); P(r, "tau_rise", q, 2.0);
              } else {
                P(r, "tau_syn", q, 2.0);
              }
            }
          }
          break;
        case M_IZH_COND_BETA:
          P(r, "V_th", 0, 30.0); P(r, "a", 0, 0.02); P(r, "b", 0, 0.2);
          P(r, "c", 0, -65.0); P(r, "d", 0, 8.0); P(r, "I_e", 0, 0.0);
          P(r, "t_ref", 0, 0.0); P(r, "den_delay", 0, 0.0);
          V(r, "V_m", 0, -70.0);
          // u = b*V_m in float, src/izhikevich_cond_beta.cu:57
          V(r, "u", 0, (double)((float)0.2f * (float)-70.0f));
          for (int q = 0; q < g.n_port; ++q) {
            P(r, "E_rev", q, 0.0); P(r, "tau_decay", q, 20.0); P(r, "tau_rise", q, 2.0);
          }
          break;
        case M_IAF_PSC_EXP:  // src/iaf_psc_exp.cu:134-155
          P(r, "tau_m", 0, 10.0); P(r, "C_m", 0, 250.0); P(r, "E_L", 0, -70.0);
          P(r, "I_e", 0, 0.0); P(r, "Theta_rel", 0, 15.0); P(r, "V_reset_rel", 0, 0.0);
          P(r, "tau_ex", 0, 2.0); P(r, "tau_in", 0, 2.0); P(r, "t_ref", 0, 2.0);
          P(r, "den_delay", 0, 0.0);
          break;
        case M_POISSON_GEN:
          P(r, "rate", 0, 0.0); P(r, "origin", 0, 0.0); P(r, "start", 0, 0.0);
          P(r, "stop", 0, 1.0e30);
          break;
        case M_SPIKE_DETECTOR:
          P(r, "hold_spike_height", 0, 1.0);
          break;
        default:
          break;
      }
    }
    (void)mi;
  }

  // ---- calibrate, src/nestgpu.cu:294-352 ----------------------------------
  void calibrate() {
    calibrated_ = true;
    neural_time_ = t_min_;
    // connection store sorted like NetConnection::Insert: per source,
    // delay ascending, insertion order within a delay.
    size_t nc = c_src_.size();
    std::vector<long long> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) {
      if (c_src_[a] != c_src_[b]) return c_src_[a] < c_src_[b];
      return c_d_[a] < c_d_[b];
    });
    s_tgt_.resize(nc); s_port_.resize(nc); s_w_.resize(nc); s_stream_.resize(nc);
    grp_off_.assign(n_nodes_ + 1, 0);
    std::vector<int> src_sorted(nc), d_sorted(nc);
    for (size_t k = 0; k < nc; ++k) {
      long long o = order[k];
      s_tgt_[k] = c_tgt_[o]; s_port_[k] = c_port_[o]; s_w_[k] = c_w_[o];
      s_stream_[k] = c_stream_[o];
      src_sorted[k] = c_src_[o]; d_sorted[k] = c_d_[o];
    }
    // delay groups per source
    grp_start_.clear(); grp_delay_.clear();
    size_t k = 0;
    for (int s = 0; s < n_nodes_; ++s) {
      grp_off_[s] = grp_start_.size();
      while (k < nc && src_sorted[k] == s) {
        int d = d_sorted[k];
        grp_start_.push_back(k);
        grp_delay_.push_back(d);
        while (k < nc && src_sorted[k] == s && d_sorted[k] == d) ++k;
      }
    }
    grp_off_[n_nodes_] = grp_start_.size();
    grp_start_.push_back(nc);
    // BuildDirectConnections, src/nestgpu.cu:1190-1226.  The stream id of
    // each direct connection is its index in the sorted connection store.
    for (auto& g : groups_) {
      if (g.model != M_POISSON_GEN) continue;
      for (int s = g.i0; s < g.i0 + g.n; ++s) {
        for (int gi = grp_off_[s]; gi < grp_off_[s + 1]; ++gi) {
          for (long long c = grp_start_[gi]; c < grp_start_[gi + 1]; ++c) {
            typename Group<T>::DirConn dc;
            dc.irel = s - g.i0;
            dc.target = s_tgt_[c];
            dc.port = s_port_[c];
            dc.weight = s_w_[c];
            dc.delay = dt_ * (grp_delay_[gi] + 1);
            dc.stream = s_stream_[c] >= 0 ? (uint32_t)s_stream_[c] : (uint32_t)c;
            g.dir.push_back(dc);
          }
        }
      }
    }
    // spike buffers, src/spike_buffer.cu:331-553
    sb_size_.assign(n_nodes_, 0);
    sb_idx0_.assign(n_nodes_, 0);
    sb_time_.assign((size_t)n_nodes_ * max_sb_, 0);
    sb_conn_.assign((size_t)n_nodes_ * max_sb_, 0);
    sb_height_.assign((size_t)n_nodes_ * max_sb_, 0.f);
    last_spike_idx_.assign(n_nodes_, 0x70000000LL);
    // group calibrate
    for (auto& g : groups_) {
      g.acc.assign((size_t)g.n * g.n_port, 0.0);
      const ModelInfo& mi = info(g.model);
      if (mi.rk5) {
        g.h_min = g.h_min_rel * dt_;
        g.h0 = g.h0_rel * dt_;
        g.x.assign(g.n, t_min_);
        g.h.assign(g.n, (T)g.h0);
        for (int r = 0; r < g.n; ++r) node_calibrate(g, r);
      } else if (g.model == M_IAF_PSC_EXP) {
        for (int r = 0; r < g.n; ++r) iaf_calibrate(g, r);
      } else if (g.model == M_SPIKE_GEN) {
        for (int r = 0; r < g.n; ++r) sg_calibrate(g, r);
      }
    }
  }

  // src/aeif_cond_beta_multisynapse.cu:72-99 (and the izhikevich /
  // alpha / psc_exp counterparts)
  void node_calibrate(Group<T>& g, int r) {
    T* p = &g.p[(size_t)r * g.np_];
    int ref = param_index(g, "refractory_step", 0);
    p[ref] = 0;
    int ns = info(g.model).scal_param.size();
    // beta peak normalisation of a rise/decay pair into g0
    auto beta_norm = [](T tr, T td) -> T {
      T denom1 = td - tr;
      T denom2 = 0;
      if (denom1 != 0) {
        T t_p = td * tr * std::log(td / tr) / denom1;
        denom2 = std::exp(-t_p / td) - std::exp(-t_p / tr);
      }
      if (denom2 == 0) return (T)(M_E / td);
      return (T)((1. / tr - 1. / td) / denom2);
    };
    auto PI = [&](const char* nm, int port) { return param_index(g, nm, port); };
    if (g.model == M_AEIF_COND_BETA_MS || g.model == M_IZH_COND_BETA) {
      for (int i = 0; i < g.n_port; ++i) {
        T* pp = p + ns + 4 * i;  // E_rev, tau_rise, tau_decay, g0
        pp[3] = beta_norm(pp[1], pp[2]);
      }
    } else if (g.model == M_AEIF_COND_ALPHA_MS) {
      for (int i = 0; i < g.n_port; ++i) {
        T* pp = p + ns + 3 * i;  // E_rev, tau_syn, g0
        pp[2] = (T)(M_E / pp[1]);
      }
    } else if (g.model == M_AEIF_COND_BETA) {   // src/aeif_cond_beta.cu:81-113
      p[PI("g0_ex", 0)] = beta_norm(p[PI("tau_rise_ex", 0)], p[PI("tau_decay_ex", 0)]);
      p[PI("g0_in", 0)] = beta_norm(p[PI("tau_rise_in", 0)], p[PI("tau_decay_in", 0)]);
    } else if (g.model == M_AEIF_COND_ALPHA) {  // src/aeif_cond_alpha.cu:79-80
      p[PI("g0_ex", 0)] = (T)(M_E / p[PI("tau_syn_ex", 0)]);
      p[PI("g0_in", 0)] = (T)(M_E / p[PI("tau_syn_in", 0)]);
    } else if (g.model == M_AEIF_PSC_ALPHA) {   // src/aeif_psc_alpha.cu:78-79
      p[PI("I0_ex", 0)] = (T)(M_E / p[PI("tau_syn_ex", 0)]);
      p[PI("I0_in", 0)] = (T)(M_E / p[PI("tau_syn_in", 0)]);
    } else if (g.model == M_AEIF_PSC_ALPHA_MS) {  // src/aeif_psc_alpha_multisynapse.cu:80-82
      for (int i = 0; i < g.n_port; ++i) p[PI("I0", i)] = (T)(M_E / p[PI("tau_syn", i)]);
    }
    if (g.model == M_AEIF_PSC_EXP || g.model == M_AEIF_PSC_ALPHA ||
        g.model == M_AEIF_PSC_ALPHA_MS || g.model == M_AEIF_PSC_DELTA ||
        g.model == M_AEIF_PSC_EXP_MS) {
      int iDT = param_index(g, "Delta_T", 0);
      if (p[iDT] <= 0.0) p[param_index(g, "V_peak", 0)] = p[param_index(g, "V_th", 0)];
    }
  }

  // src/iaf_psc_exp.cu:68-82 (coefficients in float as the reference)
  void iaf_calibrate(Group<T>& g, int r) {
    T* p = &g.p[(size_t)r * g.np_];
    enum { tau_m, C_m, E_L, I_e, Theta, V_reset, tau_ex, tau_in, t_ref, den,
           P20, P11ex, P11in, P21ex, P21in, P22 };
    float h = dt_;
    float tm = (float)p[tau_m], cm = (float)p[C_m], te = (float)p[tau_ex],
          ti = (float)p[tau_in];
    p[P11ex] = (T)std::exp(-h / te);
    p[P11in] = (T)std::exp(-h / ti);
    float p22 = std::exp(-h / tm);
    p[P22] = (T)p22;
    p[P21ex] = (T)(float)propagator_32(te, tm, cm, h);
    p[P21in] = (T)(float)propagator_32(ti, tm, cm, h);
    p[P20] = (T)(float)(tm / cm * (1.0 - p22));
  }

  // src/spike_generator.cu:230-269
  void sg_calibrate(Group<T>& g, int r) {
    auto& t = g.sg_times[r];
    auto& h = g.sg_heights[r];
    if (t.empty()) return;
    if (h.empty()) h.assign(t.size(), 1.0f);
    else if (h.size() != t.size())
      throw std::runtime_error("spike time array and spike height array must have the same size");
    g.sg_idx[r].resize(t.size());
    float tmin = (float)t_min_;
    for (size_t i = 0; i < t.size(); ++i) {
      g.sg_idx[r][i] = (int)std::round((t[i] - tmin) / dt_);
      if (i > 0 && g.sg_idx[r][i] <= g.sg_idx[r][i - 1])
        throw std::runtime_error("Spike times must be ordered");
    }
  }

  // ---- one simulation step, src/nestgpu.cu:469-635 -------------------------
  void step(long long it) {
    // 1. SpikeBufferUpdate -> send list
    send_src_.clear(); send_grp_.clear(); send_h_.clear();
    spike_buffer_update();
    // 2. time
    neural_time_ = neur_t0_ + (double)dt_ * (it + 1);
    time_idx_ = (long long)(int)std::round(neur_t0_ / dt_) + it + 1;
    // 3. node updates, in creation order
    for (auto& g : groups_) update_group(g, it);
    // 4. multimeter
    write_records(neural_time_);
    // 5. clear accumulators, collect, poisson, fold
    for (auto& g : groups_) std::fill(g.acc.begin(), g.acc.end(), 0.0);
    collect();
    for (auto& g : groups_)
      if (g.model == M_POISSON_GEN) poisson_send(g);
    for (auto& g : groups_) fold(g);
  }

  void spike_buffer_update() {
    // src/spike_buffer.cu:247-317 (no dendritic delay: den_delay_idx = 0)
    for (int s = 0; s < n_nodes_; ++s) {
      int is0 = sb_idx0_[s];
      int Ns = sb_size_[s];
      int ngrp = grp_off_[s + 1] - grp_off_[s];
      for (int is = 0; is < Ns; ++is) {
        int is1 = (is0 + is) % max_sb_;
        size_t ia = (size_t)is1 * n_nodes_ + s;
        int i_conn = sb_conn_[ia];
        int st = sb_time_[ia];
        if (i_conn < ngrp && st == grp_delay_[grp_off_[s] + i_conn]) {
          send_src_.push_back(s);
          send_grp_.push_back(grp_off_[s] + i_conn);
          send_h_.push_back(sb_height_[ia]);
          i_conn++;
          sb_conn_[ia] = i_conn;
        }
        if (is == Ns - 1 && i_conn >= ngrp && st + 1 >= 0)
          sb_size_[s]--;
        else
          sb_time_[ia]++;
      }
    }
  }

  // src/spike_buffer.cu:162-242 (recording parts kept to what the tests read)
  void push_spike(int node, float height) {
    last_spike_idx_[node] = time_idx_;
#pragma omp critical(orc_events)
    {
      ev_t_.push_back(time_idx_);
      ev_node_.push_back(node);
    }
    if (grp_off_[node + 1] - grp_off_[node] > 0) {
      int Ns = sb_size_[node];
      if (Ns >= max_sb_) return;  // buffer overflow: spike dropped
      sb_size_[node]++;
      int is0 = (sb_idx0_[node] + max_sb_ - 1) % max_sb_;
      sb_idx0_[node] = is0;
      size_t ia = (size_t)is0 * n_nodes_ + node;
      sb_time_[ia] = 0;
      sb_conn_[ia] = 0;
      sb_height_[ia] = height;
    }
  }

  void update_group(Group<T>& g, long long it) {
    const ModelInfo& mi = info(g.model);
    if (mi.rk5) {
      StepStats st_tot;
#pragma omp parallel
      {
        StepStats st;
#pragma omp for schedule(static)
        for (int r = 0; r < g.n; ++r) rk5_update(g, r, &st);
#pragma omp critical(orc_stats)
        {
          st_tot.attempts += st.attempts;
          st_tot.rejected += st.rejected;
          st_tot.substeps += st.substeps;
        }
      }
      stats_.attempts += st_tot.attempts;
      stats_.rejected += st_tot.rejected;
      stats_.substeps += st_tot.substeps;
    } else if (g.model == M_IAF_PSC_EXP) {
#pragma omp parallel for schedule(static)
      for (int r = 0; r < g.n; ++r) iaf_update(g, r);
    } else if (g.model == M_SPIKE_GEN) {
      for (int r = 0; r < g.n; ++r) {
        auto& idx = g.sg_idx[r];
        int is = g.sg_next[r];
        if (!idx.empty() && is < (int)idx.size() && idx[is] == it) {
          push_spike(g.i0 + r, g.sg_heights[r][is]);
          g.sg_next[r]++;
        }
      }
    } else if (g.model == M_SPIKE_DETECTOR) {
      for (int r = 0; r < g.n; ++r) {
        T* y = &g.y[(size_t)r * g.nv];
        T* p = &g.p[(size_t)r * g.np_];
        T sh = y[0];
        if (sh != 0.0) {
          if (p[0] == 0.0) sh = 1.0;
          y[0] = 0;
        }
        y[1] = sh;
      }
    } else if (g.model == M_PARROT) {   // src/parrot_neuron.cu:55-81
      for (int r = 0; r < g.n; ++r) {
        T* y = &g.y[(size_t)r * g.nv];
        const T* p = &g.p[(size_t)r * g.np_];
        T sh = y[0];      // input_spike_height
        y[2] = sh;        // V
        if (sh != 0.0) {
          if (p[0] == 0.0) sh = 1.0;   // hold_spike_height
          y[0] = 0;
          push_spike(g.i0 + r, (float)sh);
        }
      }
    }
  }

  // src/rk5.h:161-207 with the model's ExternalUpdate
  void rk5_update(Group<T>& g, int r, StepStats* st) {
    T* y = &g.y[(size_t)r * g.nv];
    T* p = &g.p[(size_t)r * g.np_];
    double& x = g.x[r];
    T& h = g.h[r];
    double x1 = neural_time_;
    T h_min = (T)g.h_min;
    bool end = false;
    while (!end) {
      T hmax = (T)(x1 - x);
      switch (g.model) {
        case M_AEIF_COND_BETA_MS:
          rk5_step<T, AeifCondBeta<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_COND_ALPHA_MS:
          rk5_step<T, AeifCondAlpha<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_PSC_EXP:
          rk5_step<T, AeifPscExp<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_IZH_COND_BETA:
          rk5_step<T, IzhCondBeta<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_COND_ALPHA:
          rk5_step<T, AeifCond2<T, false>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_COND_BETA:
          rk5_step<T, AeifCond2<T, true>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_PSC_ALPHA:
          rk5_step<T, AeifPscAlpha<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_PSC_ALPHA_MS:
          rk5_step<T, AeifPscAlphaMS<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_PSC_DELTA:
          rk5_step<T, AeifPscDelta<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
        case M_AEIF_PSC_EXP_MS:
          rk5_step<T, AeifPscExpMS<T>>(g.n_port, g.nv, x, y, h, h_min, hmax, p, st); break;
      }
      end = (x >= x1 - (double)h_min);
      external_update(g, r, y, p, end);
    }
  }

  // ExternalUpdate of the RK5 models (src/aeif_*_kernel.h, src/izhikevich_cond_beta_kernel.h):
  // the same rule on each model's own names
  void external_update(Group<T>& g, int r, T* y, T* p, bool end) {
    const bool izh = g.model == M_IZH_COND_BETA;
    int iV = var_index(g, "V_m", 0);
    int iw = var_index(g, izh ? "u" : "w", 0);
    int ref = param_index(g, "refractory_step", 0);
    int i_reset = param_index(g, izh ? "c" : "V_reset", 0);
    int i_thr = param_index(g, izh ? "V_th" : "V_peak", 0);
    int i_jump = param_index(g, izh ? "d" : "b", 0);
    int i_tref = param_index(g, "t_ref", 0);
    if (y[iV] < -1.0e3) { y[iV] = p[i_reset]; y[iw] = 0; return; }
    if (y[iw] < -1.0e6 || y[iw] > 1.0e6) { y[iV] = p[i_reset]; y[iw] = 0; return; }
    if (p[ref] > 0.0) {
      y[iV] = p[i_reset];
      if (end) p[ref] -= 1.0;
    } else if (y[iV] >= p[i_thr]) {
      push_spike(g.i0 + r, 1.0f);
      y[iV] = p[i_reset];
      y[iw] += p[i_jump];
      // (int)round(t_ref/dt) in float
      p[ref] = (T)(int)std::round((float)p[i_tref] / dt_);
      if (p[ref] < 0) p[ref] = 0;
    }
  }

  // src/iaf_psc_exp.cu:85-110
  void iaf_update(Group<T>& g, int r) {
    T* y = &g.y[(size_t)r * g.nv];
    T* p = &g.p[(size_t)r * g.np_];
    enum { I_ex, I_in, V, ref };
    enum { tau_m, C_m, E_L, I_e, Theta, V_reset, tau_ex, tau_in, t_ref, den,
           P20, P11ex, P11in, P21ex, P21in, P22 };
    if (y[ref] > 0.0) {
      y[ref] -= 1.0;
    } else {
      y[V] = y[V] * p[P22] + y[I_ex] * p[P21ex] + y[I_in] * p[P21in] + p[I_e] * p[P20];
    }
    y[I_ex] *= p[P11ex];
    y[I_in] *= p[P11in];
    if (y[V] >= p[Theta]) {
      push_spike(g.i0 + r, 1.0f);
      y[V] = p[V_reset];
      y[ref] = (T)(int)std::round((float)p[t_ref] / dt_);
    }
  }

  // src/get_spike.cu:60-108: acc += (double)(height*weight)
  void collect() {
    long long nsend = send_src_.size();
    for (long long k = 0; k < nsend; ++k) {
      int gi = send_grp_[k];
      float hgt = send_h_[k];
      for (long long c = grp_start_[gi]; c < grp_start_[gi + 1]; ++c) {
        int tgt = s_tgt_[c];
        Group<T>& tg = groups_[node_group_[tgt]];
        size_t i = (size_t)s_port_[c] * tg.n + (tgt - tg.i0);
        tg.acc[i] += (double)(hgt * s_w_[c]);
        ++n_syn_events_;
      }
    }
  }

  void update_poisson_lambda() {
    // time_step*rate in float, src/poiss_gen.cu:74 and src/nestgpu.cu:560
    float time_step = (float)(dt_ / 1000.0);
    for (auto& g : groups_) {
      if (g.model != M_POISSON_GEN) continue;
      g.pg_chunks.resize(g.n);
      g.pg_exp.resize(g.n);
      for (int r = 0; r < g.n; ++r) {
        float lambda = time_step * (float)g.p[(size_t)r * g.np_ + 0];
        int m = poisson_chunks(lambda);
        g.pg_chunks[r] = m;
        g.pg_exp[r] = m > 0 ? std::exp(-(double)lambda / m) : 0.0;
      }
    }
  }

  // src/poiss_gen.cu:55-86
  void poisson_send(Group<T>& g) {
    for (auto& dc : g.dir) {
      const T* p = &g.p[(size_t)dc.irel * g.np_];
      double t_rel = neural_time_ - (float)p[1] - dc.delay;
      if (t_rel >= (float)p[2] && t_rel <= (float)p[3]) {
        int n = poisson_draw(seed_, dc.stream, time_idx_, g.pg_chunks[dc.irel],
                             g.pg_exp[dc.irel]);
        if (n > 0) {
          Group<T>& tg = groups_[node_group_[dc.target]];
          size_t i = (size_t)dc.port * tg.n + (dc.target - tg.i0);
          tg.acc[i] += (double)(dc.weight * n);
        }
      }
    }
  }

  // src/get_spike.cu:142-166: input = (float)((double)input + acc*weight)
  void fold(Group<T>& g) {
    if (g.n_port == 0) return;
    const ModelInfo& mi = info(g.model);
    int nsv = mi.scal_var.size(), npv = mi.port_var.size();
    int nsp = mi.scal_param.size(), npp = mi.port_param.size();
    // ports in order (a model whose ports share one variable adds them one after another)
    for (int q = 0; q < g.n_port; ++q) {
      for (int r = 0; r < g.n; ++r) {
        double a = g.acc[(size_t)q * g.n + r];
        int iv = mi.input_in_port_block ? nsv + q * npv + mi.input_var
                                        : mi.input_var + q * mi.input_scal_step;
        T* in = &g.y[(size_t)r * g.nv + iv];
        double wgt = q < mi.weight_unit_ports ? 1.0 : 0.0;
        if (mi.weight_param >= 0) {
          int ip = mi.weight_in_scal ? mi.weight_param + q : nsp + q * npp + mi.weight_param;
          wgt = (double)g.p[(size_t)r * g.np_ + ip];
        }
        *in = (T)((double)*in + a * wgt);
      }
    }
  }

  // src/multimeter.cu:76-110
  void write_records(double t) {
    for (auto& r : records_) {
      Group<T>& g = groups_[r.group];
      float v;
      if (r.var == "spike") {
        // GetSpikeActivity: most recent buffered spike of age 0
        int s = r.node;
        v = 0.f;
        if (sb_size_[s] > 0) {
          size_t ia = (size_t)sb_idx0_[s] * n_nodes_ + s;
          if (sb_time_[ia] == 0) v = sb_height_[ia];
        }
      } else {
        int iv = var_index(g, r.var, r.port);
        if (iv < 0) throw std::runtime_error("Unrecognized variable " + r.var);
        v = (float)g.y[(size_t)r.rel * g.nv + iv];
      }
      r.t.push_back((float)t);
      r.v.push_back(v);
    }
  }

  float dt_;
  uint64_t seed_;
  bool calibrated_ = false, first_sim_ = true;
  double t_min_ = 0.0, neural_time_ = 0.0, neur_t0_ = 0.0;
  long long time_idx_ = 0;
  int n_nodes_ = 0;
  int max_sb_ = 20;
  std::vector<Group<T>> groups_;
  std::vector<int> node_group_;
  // connections (insertion order), then sorted store
  std::vector<int> c_src_, c_tgt_, c_d_;
  std::vector<unsigned char> c_port_;
  std::vector<float> c_w_;
  std::vector<long long> c_stream_, s_stream_;
  std::vector<int> s_tgt_;
  std::vector<unsigned char> s_port_;
  std::vector<float> s_w_;
  std::vector<int> grp_off_;          // per source: first delay group
  std::vector<long long> grp_start_;  // per delay group: first connection
  std::vector<int> grp_delay_;
  // spike buffers
  std::vector<int> sb_size_, sb_idx0_, sb_time_, sb_conn_;
  std::vector<float> sb_height_;
  std::vector<long long> last_spike_idx_;
  // send list
  std::vector<int> send_src_, send_grp_;
  std::vector<float> send_h_;
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C API used by the tests (ctypes).  Handles carry their precision.
using namespace orc;

struct OrcHandle {
  int prec;  // 0 float, 1 double
  Engine<float>* f;
  Engine<double>* d;
  std::string err;
};

template <class F>
static int guard(OrcHandle* h, F fn) {
  try {
    fn();
    return 0;
  } catch (std::exception& e) {
    h->err = e.what();
    return -1;
  }
}

#define DISPATCH(h, call) ((h)->prec == 0 ? (h)->f->call : (h)->d->call)

extern "C" {

void* orc_new(float dt, unsigned long long seed, int prec) {
  OrcHandle* h = new OrcHandle{prec, nullptr, nullptr, ""};
  if (prec == 0) h->f = new Engine<float>(dt, seed);
  else h->d = new Engine<double>(dt, seed);
  return h;
}

void orc_free(void* hv) {
  OrcHandle* h = (OrcHandle*)hv;
  delete h->f;
  delete h->d;
  delete h;
}

const char* orc_error(void* hv) { return ((OrcHandle*)hv)->err.c_str(); }

int orc_create(void* hv, const char* model, int n, int n_port) {
  OrcHandle* h = (OrcHandle*)hv;
  int ret = -1;
  if (guard(h, [&] { ret = DISPATCH(h, create(model, n, n_port)); }) < 0) return -1;
  return ret;
}

int orc_set(void* hv, int node0, int n, const char* name, int port, const float* vals) {
  OrcHandle* h = (OrcHandle*)hv;
  return guard(h, [&] { DISPATCH(h, set(node0, n, name, port, vals)); });
}

double orc_get(void* hv, int node, const char* name, int port) {
  OrcHandle* h = (OrcHandle*)hv;
  double v = NAN;
  guard(h, [&] { v = DISPATCH(h, get(node, name, port)); });
  return v;
}

int orc_set_spike_times(void* hv, int node, int n, const float* t, const float* hg) {
  OrcHandle* h = (OrcHandle*)hv;
  return guard(h, [&] { DISPATCH(h, set_spike_times(node, n, t, hg)); });
}

int orc_connect(void* hv, long long n, const int* src, const int* tgt,
                const unsigned char* port, const float* w, const float* delay) {
  OrcHandle* h = (OrcHandle*)hv;
  return guard(h, [&] { DISPATCH(h, connect(n, src, tgt, port, w, delay)); });
}

int orc_connect_stream(void* hv, long long n, const int* src, const int* tgt,
                       const unsigned char* port, const float* w, const float* delay,
                       const long long* stream) {
  OrcHandle* h = (OrcHandle*)hv;
  return guard(h, [&] { DISPATCH(h, connect(n, src, tgt, port, w, delay, stream)); });
}

int orc_record(void* hv, int node, const char* var, int port) {
  OrcHandle* h = (OrcHandle*)hv;
  int ret = -1;
  if (guard(h, [&] { ret = DISPATCH(h, record(node, var, port)); }) < 0) return -1;
  return ret;
}

int orc_simulate(void* hv, float sim_time) {
  OrcHandle* h = (OrcHandle*)hv;
  return guard(h, [&] { DISPATCH(h, simulate(sim_time)); });
}

long long orc_record_rows(void* hv, int rec) {
  OrcHandle* h = (OrcHandle*)hv;
  return h->prec == 0 ? h->f->records_.at(rec).t.size() : h->d->records_.at(rec).t.size();
}

void orc_record_copy(void* hv, int rec, float* t, float* v) {
  OrcHandle* h = (OrcHandle*)hv;
  const Record& r = h->prec == 0 ? h->f->records_.at(rec) : h->d->records_.at(rec);
  std::copy(r.t.begin(), r.t.end(), t);
  std::copy(r.v.begin(), r.v.end(), v);
}

long long orc_n_events(void* hv) {
  OrcHandle* h = (OrcHandle*)hv;
  return h->prec == 0 ? h->f->ev_t_.size() : h->d->ev_t_.size();
}

void orc_events(void* hv, long long* t_idx, int* node) {
  OrcHandle* h = (OrcHandle*)hv;
  if (h->prec == 0) {
    std::copy(h->f->ev_t_.begin(), h->f->ev_t_.end(), t_idx);
    std::copy(h->f->ev_node_.begin(), h->f->ev_node_.end(), node);
  } else {
    std::copy(h->d->ev_t_.begin(), h->d->ev_t_.end(), t_idx);
    std::copy(h->d->ev_node_.begin(), h->d->ev_node_.end(), node);
  }
}

// stats[0]=attempts, [1]=rejected, [2]=accepted substeps, [3]=synaptic events
void orc_stats(void* hv, long long* stats) {
  OrcHandle* h = (OrcHandle*)hv;
  const StepStats& s = h->prec == 0 ? h->f->stats_ : h->d->stats_;
  stats[0] = s.attempts;
  stats[1] = s.rejected;
  stats[2] = s.substeps;
  stats[3] = h->prec == 0 ? h->f->n_syn_events_ : h->d->n_syn_events_;
}

int orc_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// One RK5 step of y' = lambda*y from (x0, y0) with step h (h_min = h_max =
// h so the step is forced).  out = {y1, err_estimate_scaled, h_next}.
void orc_rk5_linear(int prec, double lambda, double y0, double h, double* out) {
  if (prec == 0) {
    float y = (float)y0, hh = (float)h, p = (float)lambda;
    double x = 0.0;
    rk5_step<float, LinearModel<float>>(0, 1, x, &y, hh, (float)h, (float)h, &p, nullptr);
    out[0] = y; out[1] = x; out[2] = hh;
  } else {
    double y = y0, hh = h, p = lambda;
    double x = 0.0;
    rk5_step<double, LinearModel<double>>(0, 1, x, &y, hh, h, h, &p, nullptr);
    out[0] = y; out[1] = x; out[2] = hh;
  }
}

// Adaptive integration of y' = lambda*y over [0, x1] with the reference
// controller (h_min, h0).  out = {y(x1), x_final, n_attempts, n_rejected}.
void orc_rk5_linear_adaptive(int prec, double lambda, double y0, double x1,
                             double h0, double h_min, double* out) {
  StepStats st;
  if (prec == 0) {
    float y = (float)y0, h = (float)h0, p = (float)lambda;
    double x = 0.0;
    while (!(x >= x1 - (double)(float)h_min)) {
      float hmax = (float)(x1 - x);
      rk5_step<float, LinearModel<float>>(0, 1, x, &y, h, (float)h_min, hmax, &p, &st);
    }
    out[0] = y; out[1] = x;
  } else {
    double y = y0, h = h0, p = lambda;
    double x = 0.0;
    while (!(x >= x1 - h_min)) {
      double hmax = x1 - x;
      rk5_step<double, LinearModel<double>>(0, 1, x, &y, h, h_min, hmax, &p, &st);
    }
    out[0] = y; out[1] = x;
  }
  out[2] = st.attempts;
  out[3] = st.rejected;
}

void orc_philox(unsigned int* ctr, unsigned int k0, unsigned int k1, unsigned int* out) {
  Philox4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
  Philox4 r = philox4x32_10(c, k0, k1);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

int orc_poisson(unsigned long long seed, unsigned int stream, long long t_idx, double lambda) {
  int m = poisson_chunks(lambda);
  return poisson_draw(seed, stream, t_idx, m, m > 0 ? std::exp(-lambda / m) : 0.0);
}

// test hooks for the step factors
float orc_step_factor_dec(float e) { return step_factor_dec(e); }
float orc_exp_model(float x) { return exp_model(x); }
void orc_exp_model_n(const float* x, float* out, long long n) {
  for (long long i = 0; i < n; ++i) out[i] = exp_model(x[i]);
}
float orc_step_factor_inc(float e) { return step_factor_inc(e); }
void orc_step_factors(const float* x, long long n, float* dec, float* inc) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) {
    dec[i] = step_factor_dec(x[i]);
    inc[i] = step_factor_inc(x[i]);
  }
}

double orc_propagator_32(double tau_syn, double tau, double C, double h) {
  return propagator_32(tau_syn, tau, C, h);
}

}  // extern "C"
