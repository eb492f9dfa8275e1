// Host engine of the MI355X simulation step (see engine.h).
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "connect_device.h"
#include "philox.h"
#include "predictor_host.h"

namespace ngpu {

#define HIPCHK(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw Error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " \
                  + #call);                                                   \
  } while (0)

static double wall_time() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
void DBuf<T>::alloc(size_t count) {
  free();
  n = count;
  if (count == 0) return;
  HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
}

template <class T>
void DBuf<T>::free() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}

template <class T>
static void upload(DBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(h.size());
  if (!h.empty()) {
    HIPCHK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
}

// Blocking copy ordered after all work queued on the simulator stream (a
// plain hipMemcpy runs on the null stream, which does not wait for a
// non-blocking stream's kernels).
static void copy_on(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, s));
  HIPCHK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Model name tables.  Names and their order are the reference's interface
// (src/*_kernel.h enums and name arrays) so Set/GetStatus indices match.
static const std::vector<std::string> kAeifScal = {
    "V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b", "tau_w", "I_e",
    "V_peak", "V_reset", "t_ref", "refractory_step", "den_delay"};

static std::vector<std::string> cat(std::vector<std::string> a, const std::vector<std::string>& b) {
  a.insert(a.end(), b.begin(), b.end());
  return a;
}

static const std::vector<ModelDesc>& model_table() {
  static const std::vector<ModelDesc> t = {
      {"aeif_cond_beta_multisynapse", MODEL_AEIF_COND_BETA_MS, {"V_m", "w"}, {"g", "g1"},
       kAeifScal, {"E_rev", "tau_rise", "tau_decay", "g0"}, {"h_min_rel", "h0_rel"}, {},
       -1, true, 1, true, 3},
      {"aeif_cond_alpha_multisynapse", MODEL_AEIF_COND_ALPHA_MS, {"V_m", "w"}, {"g", "g1"},
       kAeifScal, {"E_rev", "tau_syn", "g0"}, {"h_min_rel", "h0_rel"}, {},
       -1, true, 1, true, 2},
      {"aeif_psc_exp", MODEL_AEIF_PSC_EXP, {"I_syn_ex", "I_syn_in", "V_m", "w"}, {},
       {"tau_syn_ex", "tau_syn_in", "V_th", "Delta_T", "g_L", "E_L", "C_m", "a", "b",
        "tau_w", "I_e", "V_peak", "V_reset", "t_ref", "refractory_step", "den_delay"},
       {}, {"h_min_rel", "h0_rel"}, {}, 2, true, 0, false, -1},
      {"izhikevich_cond_beta", MODEL_IZH_COND_BETA, {"V_m", "u"}, {"g", "g1"},
       {"V_th", "a", "b", "c", "d", "I_e", "t_ref", "refractory_step", "den_delay"},
       {"E_rev", "tau_rise", "tau_decay", "g0"}, {"h_min_rel", "h0_rel"}, {},
       -1, true, 1, true, 3},
      // two fixed ports with ex/in names (src/aeif_cond_alpha_kernel.h:42-110,
      // src/aeif_cond_beta_kernel.h:42-110, src/aeif_psc_alpha_kernel.h:42-110)
      {"aeif_cond_alpha", MODEL_AEIF_COND_ALPHA, {"g_ex", "g_in", "g1_ex", "g1_in", "V_m", "w"}, {},
       cat({"g0_ex", "g0_in", "E_rev_ex", "E_rev_in", "tau_syn_ex", "tau_syn_in"}, kAeifScal),
       {}, {"h_min_rel", "h0_rel"}, {}, 2, true, 2, false, 0, 1, true},
      {"aeif_cond_beta", MODEL_AEIF_COND_BETA, {"g_ex", "g_in", "g1_ex", "g1_in", "V_m", "w"}, {},
       cat({"g0_ex", "g0_in", "E_rev_ex", "E_rev_in", "tau_rise_ex", "tau_rise_in",
            "tau_decay_ex", "tau_decay_in"}, kAeifScal),
       {}, {"h_min_rel", "h0_rel"}, {}, 2, true, 2, false, 0, 1, true},
      {"aeif_psc_alpha", MODEL_AEIF_PSC_ALPHA,
       {"I_syn_ex", "I_syn_in", "I1_syn_ex", "I1_syn_in", "V_m", "w"}, {},
       cat({"tau_syn_ex", "tau_syn_in", "I0_ex", "I0_in"}, kAeifScal),
       {}, {"h_min_rel", "h0_rel"}, {}, 2, true, 2, false, 2, 1, true},
      // src/aeif_psc_alpha_multisynapse_kernel.h:42-110
      {"aeif_psc_alpha_multisynapse", MODEL_AEIF_PSC_ALPHA_MS, {"V_m", "w"}, {"I_syn", "I1_syn"},
       kAeifScal, {"tau_syn", "I0"}, {"h_min_rel", "h0_rel"}, {}, -1, true, 1, true, 1},
      // src/aeif_psc_delta_kernel.h:42-100: every port jumps V_m, weight 1
      {"aeif_psc_delta", MODEL_AEIF_PSC_DELTA, {"V_m", "w"}, {}, kAeifScal, {},
       {"h_min_rel", "h0_rel"}, {}, -1, true, 0, false, -1, 0, false},
      // src/aeif_psc_exp_multisynapse_kernel.h:42-100
      {"aeif_psc_exp_multisynapse", MODEL_AEIF_PSC_EXP_MS, {"V_m", "w"}, {"I_syn"}, kAeifScal,
       {"tau_syn"}, {"h_min_rel", "h0_rel"}, {}, -1, true, 0, true, -1},
      {"iaf_psc_exp", MODEL_IAF_PSC_EXP, {"I_syn_ex", "I_syn_in", "V_m_rel", "refractory_step"},
       {},
       {"tau_m", "C_m", "E_L", "I_e", "Theta_rel", "V_reset_rel", "tau_ex", "tau_in",
        "t_ref", "den_delay", "P20", "P11ex", "P11in", "P21ex", "P21in", "P22"},
       {}, {}, {}, 2, false, 0, false, -1},
      {"poisson_generator", MODEL_POISSON_GEN, {}, {}, {"rate", "origin", "start", "stop"},
       {}, {}, {}, 0, false, -1, false, -1},
      {"spike_generator", MODEL_SPIKE_GEN, {}, {}, {}, {}, {}, {"spike_times", "spike_heights"},
       0, false, -1, false, -1},
      {"spike_detector", MODEL_SPIKE_DETECTOR, {"input_spike_height", "spike_height"}, {},
       {"hold_spike_height"}, {}, {}, {}, 1, false, 0, false, -1},
      // src/parrot_neuron.cu:32-100: port 0 -> input_spike_height (weight 1),
      // port 1 -> dummy_input (weight 0)
      {"parrot_neuron", MODEL_PARROT, {"input_spike_height", "dummy_input", "V"}, {},
       {"hold_spike_height", "den_delay"}, {}, {}, {}, 2, false, 0, false, -1, 1, false, 1},
  };
  return t;
}

// Reference model names outside this build's scope (src/neuron_models.h:60-86).
static const char* kOtherModels[] = {
    "iaf_psc_exp_g", "iaf_psc_exp_hc", "iaf_psc_alpha", "ext_neuron", "izhikevich", "izhikevich_psc_exp_5s", "izhikevich_psc_exp_2s", "izhikevich_psc_exp",
    "user_m1", "user_m2"};

// Where a model's port input lands: variable index of port 0's input and the
// stride between ports, parameter index of port 0's weight (-1: weight 1).
struct FoldLayout {
  int in_var, in_port_stride, w_param, w_port_stride, w_unit_ports;
};
static FoldLayout fold_layout(const ModelDesc& m) {
  FoldLayout f;
  if (m.in_port_block) {
    f.in_var = (int)m.scal_var.size() + m.in_var;
    f.in_port_stride = (int)m.port_var.size();
  } else {
    f.in_var = m.in_var;
    f.in_port_stride = m.in_scal_stride;
  }
  f.w_unit_ports = m.w_unit_ports;
  if (m.w_param < 0) {
    f.w_param = -1;
    f.w_port_stride = 0;
  } else if (m.w_in_scal) {
    f.w_param = m.w_param;
    f.w_port_stride = 1;
  } else {
    f.w_param = (int)m.scal_param.size() + m.w_param;
    f.w_port_stride = (int)m.port_param.size();
  }
  return f;
}

const ModelDesc* find_model(const std::string& name) {
  for (auto& m : model_table())
    if (name == m.name) return &m;
  return nullptr;
}

// ---------------------------------------------------------------------------
// Host-side staging: before calibration a population's state lives in host
// vectors (no GPU needed); calibration uploads it and from then on the
// device columns are authoritative.
Engine::Engine() { start_real_time_ = wall_time(); }

Engine::~Engine() {
  batch_graph_destroy(batch_graph_);
  for (auto* g : groups_) delete g;
  for (auto& r : records_)
    if (r.fp) fclose(r.fp);
  for (auto& p : ev_pending_) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : ev_free_) (void)hipEventDestroy(e);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Host waits for all queued work of the simulator's stream.
void Engine::sync_streams() {
  if (stream_) HIPCHK(hipStreamSynchronize(stream_));
}

// Time of step `it` of the current Simulate() call, src/nestgpu.cu:486-490
// (the same double expressions as before each step of the reference).
StepCtx Engine::step_ctx(long long it) const {
  StepCtx c;
  c.it = it;
  c.t_idx = (long long)(int)std::round(neur_t0_ / dt_) + it + 1;
  c.t0 = neur_t0_;
  c.t1 = neur_t0_ + (double)dt_ * (double)(it + 1);
  c.slot = (int)(c.t_idx % n_slot_);
  c.parity = (int)(c.t_idx & 1);
  c.lst = (int)(c.t_idx % kSpikeLists);
  return c;
}

void Engine::check_uncalibrated(const char* msg) const {
  if (calibrated_) throw Error(msg);
}

int Engine::SetRandomSeed(unsigned long long seed) {
  seed_ = seed;
  kernel_seed_ = seed + 12345ull;   // src/nestgpu.cu:184
  rng_ctr_ = 0;
  return 0;
}

int Engine::SetTimeResolution(float dt) {
  dt_ = dt;
  return 0;
}

// ---- kernel parameters ------------------------------------------------------
std::vector<std::string> Engine::FloatParamNames() const {
  return {"time_resolution", "max_spike_num_fact", "max_spike_per_host_fact"};
}
std::vector<std::string> Engine::IntParamNames() const {
  return {"rnd_seed", "verbosity_level", "max_spike_buffer_size", "remote_spike_height_flag"};
}
std::vector<std::string> Engine::BoolParamNames() const { return {"print_time"}; }

float Engine::GetFloatParam(const std::string& n) const {
  if (n == "time_resolution") return dt_;
  if (n == "max_spike_num_fact") return max_spike_num_fact_;
  if (n == "max_spike_per_host_fact") return max_spike_per_host_fact_;
  throw Error("Unrecognized kernel float parameter " + n);
}
int Engine::SetFloatParam(const std::string& n, float v) {
  if (n == "time_resolution") return SetTimeResolution(v);
  if (n == "max_spike_num_fact") { max_spike_num_fact_ = v; return 0; }
  if (n == "max_spike_per_host_fact") { max_spike_per_host_fact_ = v; return 0; }
  throw Error("Unrecognized kernel float parameter " + n);
}
int Engine::GetIntParam(const std::string& n) const {
  if (n == "rnd_seed") return (int)seed_;
  if (n == "verbosity_level") return verbosity_;
  if (n == "max_spike_buffer_size") return max_spike_buffer_size_;
  if (n == "remote_spike_height_flag") return remote_spike_height_;
  throw Error("Unrecognized kernel int parameter " + n);
}
int Engine::SetIntParam(const std::string& n, int v) {
  if (n == "rnd_seed") return SetRandomSeed((unsigned long long)v);
  if (n == "verbosity_level") {
    if (v < 0) throw Error("Verbosity level must be >= 0");
    verbosity_ = v;
    return 0;
  }
  if (n == "max_spike_buffer_size") { max_spike_buffer_size_ = v; return 0; }
  if (n == "remote_spike_height_flag") {
    if (v != 0) throw Error("remote spikes (MPI) are not supported in this build");
    return 0;
  }
  throw Error("Unrecognized kernel int parameter " + n);
}
bool Engine::GetBoolParam(const std::string& n) const {
  if (n == "print_time") return print_time_;
  throw Error("Unrecognized kernel boolean parameter " + n);
}
int Engine::SetBoolParam(const std::string& n, bool v) {
  if (n == "print_time") { print_time_ = v; return 0; }
  throw Error("Unrecognized kernel boolean parameter " + n);
}

// ---- nodes ------------------------------------------------------------------
int Engine::param_idx(const Group& g, const std::string& name, int port) const {
  const ModelDesc& m = *g.md;
  for (size_t i = 0; i < m.scal_param.size(); ++i)
    if (m.scal_param[i] == name) return (int)i;
  for (size_t i = 0; i < m.port_param.size(); ++i)
    if (m.port_param[i] == name) return (int)(m.scal_param.size() + port * m.port_param.size() + i);
  return -1;
}

int Engine::var_idx(const Group& g, const std::string& name, int port) const {
  const ModelDesc& m = *g.md;
  for (size_t i = 0; i < m.scal_var.size(); ++i)
    if (m.scal_var[i] == name) return (int)i;
  for (size_t i = 0; i < m.port_var.size(); ++i)
    if (m.port_var[i] == name) return (int)(m.scal_var.size() + port * m.port_var.size() + i);
  return -1;
}

static int group_index(const std::vector<Group*>& gs, const Group* g) {
  for (size_t k = 0; k < gs.size(); ++k)
    if (gs[k] == g) return (int)k;
  return -1;
}

int Engine::Create(const std::string& model, int n, int n_port) {
  check_uncalibrated("Nodes cannot be created after calibration");
  if (n <= 0) throw Error("Number of nodes must be greater than zero.");
  if (n_port < 0) throw Error("Number of ports must be >= zero.");
  const ModelDesc* md = find_model(model);
  if (!md) {
    for (const char* o : kOtherModels)
      if (model == o)
        throw Error("Neuron model " + model + " is outside the scope of this build");
    throw Error("Unknown neuron model name: " + model);
  }
  if (md->forced_port >= 0) n_port = md->forced_port;
  if (md->rk5 && md->forced_port < 0 && (n_port < 1 || n_port > kMaxPort))
    throw Error("Number of ports of " + model + " must be between 1 and " +
                std::to_string(kMaxPort) + " in this build");
  if ((long long)n_nodes_ + n > (1ll << 24))
    throw Error("Maximum number of nodes (16777216) exceeded");
  Group* g = new Group;
  g->md = md;
  g->i0 = n_nodes_;
  g->n = n;
  g->n_port = n_port;
  g->nv = (int)(md->scal_var.size() + md->port_var.size() * n_port);
  g->np = (int)(md->scal_param.size() + md->port_param.size() * n_port);
  g->uni.assign(g->np, 1);
  g->uni_val.assign(g->np, 0.f);
  std::vector<float> var((size_t)g->nv * n, 0.f);
  std::vector<float> dflt_var(g->nv, 0.f);
  auto P = [&](const char* nm, float v, int port = 0) { g->uni_val[param_idx(*g, nm, port)] = v; };
  auto V = [&](const char* nm, float v, int port = 0) { dflt_var[var_idx(*g, nm, port)] = v; };
  switch (md->id) {  // defaults of NodeInit / Init in the reference model files
    case MODEL_AEIF_COND_BETA_MS:
    case MODEL_AEIF_COND_ALPHA_MS:
    case MODEL_AEIF_PSC_EXP:
    case MODEL_AEIF_COND_ALPHA:
    case MODEL_AEIF_COND_BETA:
    case MODEL_AEIF_PSC_ALPHA:
    case MODEL_AEIF_PSC_ALPHA_MS:
    case MODEL_AEIF_PSC_DELTA:
    case MODEL_AEIF_PSC_EXP_MS:
      P("V_th", -50.4f); P("Delta_T", 2.0f); P("g_L", 30.0f); P("E_L", -70.6f);
      P("C_m", 281.0f); P("a", 4.0f); P("b", 80.5f); P("tau_w", 144.0f); P("I_e", 0.f);
      P("V_peak", 0.f); P("V_reset", -60.0f); P("t_ref", 0.f); P("den_delay", 0.f);
      P("refractory_step", 0.f);
      V("V_m", -70.6f); V("w", 0.f);
      switch (md->id) {
        case MODEL_AEIF_PSC_EXP:
        case MODEL_AEIF_PSC_ALPHA:   // src/aeif_psc_alpha.cu:64-65
          P("tau_syn_ex", 0.2f); P("tau_syn_in", 2.0f);
          break;
        case MODEL_AEIF_COND_ALPHA:  // src/aeif_cond_alpha.cu:55-58
          P("E_rev_ex", 0.f); P("E_rev_in", -85.f); P("tau_syn_ex", 0.2f); P("tau_syn_in", 2.0f);
          break;
        case MODEL_AEIF_COND_BETA:   // src/aeif_cond_beta.cu:56-61
          P("E_rev_ex", 0.f); P("E_rev_in", 0.f); P("tau_decay_ex", 20.f); P("tau_decay_in", 20.f);
          P("tau_rise_ex", 2.f); P("tau_rise_in", 2.f);
          break;
        case MODEL_AEIF_PSC_ALPHA_MS:   // src/aeif_psc_alpha_multisynapse.cu:61-65
        case MODEL_AEIF_PSC_EXP_MS:     // src/aeif_psc_exp_multisynapse.cu:60-63
          for (int q = 0; q < n_port; ++q) P("tau_syn", 0.2f, q);
          break;
        case MODEL_AEIF_COND_BETA_MS:
        case MODEL_AEIF_COND_ALPHA_MS:
          for (int q = 0; q < n_port; ++q) {
            P("E_rev", 0.f, q);
            if (md->id == MODEL_AEIF_COND_BETA_MS) { P("tau_decay", 20.f, q); P("tau_rise", 2.f, q); }
            else P("tau_syn", 2.f, q);
          }
          break;
        default:
          break;
      }
      break;
    case MODEL_IZH_COND_BETA:
      P("V_th", 30.f); P("a", 0.02f); P("b", 0.2f); P("c", -65.f); P("d", 8.f); P("I_e", 0.f);
      P("t_ref", 0.f); P("den_delay", 0.f); P("refractory_step", 0.f);
      V("V_m", -70.f); V("u", 0.2f * -70.f);
      for (int q = 0; q < n_port; ++q) { P("E_rev", 0.f, q); P("tau_decay", 20.f, q); P("tau_rise", 2.f, q); }
      break;
    case MODEL_IAF_PSC_EXP:
      P("tau_m", 10.f); P("C_m", 250.f); P("E_L", -70.f); P("I_e", 0.f);
      P("Theta_rel", -55.0f - (-70.0f)); P("V_reset_rel", 0.f); P("tau_ex", 2.f); P("tau_in", 2.f);
      P("t_ref", 2.f); P("den_delay", 0.f);
      break;
    case MODEL_POISSON_GEN:
      P("rate", 0.f); P("origin", 0.f); P("start", 0.f); P("stop", 1.0e30f);
      break;
    case MODEL_SPIKE_DETECTOR:
      P("hold_spike_height", 1.f);
      break;
    case MODEL_PARROT:   // src/parrot_neuron.cu:84-88
      P("hold_spike_height", 0.f); P("den_delay", 0.f);
      break;
    default:
      break;
  }
  for (int j = 0; j < g->nv; ++j)
    std::fill(var.begin() + (size_t)j * n, var.begin() + (size_t)(j + 1) * n, dflt_var[j]);
  std::vector<float> par((size_t)g->np * n, 0.f);
  for (int j = 0; j < g->np; ++j)
    std::fill(par.begin() + (size_t)j * n, par.begin() + (size_t)(j + 1) * n, g->uni_val[j]);
  if (md->rk5) {  // refractory_step is per-neuron mutable state
    int r = param_idx(*g, "refractory_step", 0);
    g->uni[r] = 0;
  }
  if (md->id == MODEL_SPIKE_GEN) {
    g->sg_times.assign(n, {});
    g->sg_heights.assign(n, {});
  }
  groups_.push_back(g);
  stage_var_.push_back(std::move(var));
  stage_param_.push_back(std::move(par));
  node_group_.insert(node_group_.end(), n, (int)groups_.size() - 1);
  n_nodes_ += n;
  return g->i0;
}

Group& Engine::GroupOf(int node, int n_node, int* rel) {
  if (node < 0 || n_node < 1 || node + n_node > n_nodes_)
    throw Error("Unrecognized node in getting node sequence offset");
  int gi = node_group_[node];
  if (node_group_[node + n_node - 1] != gi)
    throw Error("Nodes belong to different node groups in setting parameter");
  *rel = node - groups_[gi]->i0;
  return *groups_[gi];
}

std::vector<int> Engine::NodeListRel(const int* nodes, int n, Group** g) {
  if (n < 1) throw Error("Empty node list");
  int in0 = nodes[0];
  if (in0 < 0 || in0 >= n_nodes_) throw Error("Unrecognized node in setting parameter");
  int gi = node_group_[in0];
  std::vector<int> rel(n);
  for (int k = 0; k < n; ++k) {
    int in = nodes[k];
    if (in < 0 || in >= n_nodes_) throw Error("Unrecognized node in setting parameter");
    if (node_group_[in] != gi) throw Error("Nodes belong to different node groups in setting parameter");
    rel[k] = in - groups_[gi]->i0;
  }
  *g = groups_[gi];
  return rel;
}

static Group* group_at(const std::vector<Group*>& gs, const std::vector<int>& ng, int node, int n_nodes) {
  if (node < 0 || node >= n_nodes) throw Error("Unrecognized node");
  return gs[ng[node]];
}

bool Engine::IsScalParam(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->scal_param;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsPortParam(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->port_param;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsArrayParam(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->array_param;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsGroupParam(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->group_param;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsIntVar(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  return std::find(g->int_var.begin(), g->int_var.end(), n) != g->int_var.end();
}
bool Engine::IsScalVar(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->scal_var;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsPortVar(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  auto& v = g->md->port_var;
  return std::find(v.begin(), v.end(), n) != v.end();
}
bool Engine::IsArrayVar(int, const std::string&) { return false; }

int Engine::GetParamSize(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  if (IsArrayParam(node, n)) {
    int rel = node - g->i0;
    if (n == "spike_times") return (int)g->sg_times[rel].size();
    return (int)g->sg_heights[rel].size();
  }
  if (IsScalParam(node, n)) return 1;
  if (IsPortParam(node, n)) return g->n_port;
  throw Error("Unrecognized parameter " + n);
}

int Engine::GetVarSize(int node, const std::string& n) {
  Group* g = group_at(groups_, node_group_, node, n_nodes_);
  if (IsIntVar(node, n) || IsScalVar(node, n)) return 1;
  if (IsPortVar(node, n)) return g->n_port;
  throw Error("Unrecognized variable " + n);
}

// Column j of a population: host staging before calibration, device after.
static void col_read(hipStream_t stream_, bool cal, Group& g, std::vector<float>& stage,
                     DBuf<float>& dev, int j, int rel0, int cnt, float* out) {
  if (!cal) {
    std::memcpy(out, stage.data() + (size_t)j * g.n + rel0, cnt * sizeof(float));
  } else {
    copy_on(stream_, out, dev.p + (size_t)j * g.n + rel0, cnt * sizeof(float), hipMemcpyDeviceToHost);
  }
}
static void col_write(hipStream_t stream_, bool cal, Group& g, std::vector<float>& stage,
                      DBuf<float>& dev, int j, int rel0, int cnt, const float* in) {
  if (!cal) {
    std::memcpy(stage.data() + (size_t)j * g.n + rel0, in, cnt * sizeof(float));
  } else {
    copy_on(stream_, dev.p + (size_t)j * g.n + rel0, in, cnt * sizeof(float), hipMemcpyHostToDevice);
  }
}

// Writes vals[k] at rows rel[k] of column j: into the host staging before
// calibration (O(entries)), else to the device (one copy for a contiguous
// ascending range, one read-modify-write of the column otherwise).
static void col_write_entries(hipStream_t stream_, bool cal, Group& g, std::vector<float>& stage,
                              DBuf<float>& dev, int j, const std::vector<int>& rel,
                              const float* vals) {
  if (rel.empty()) return;
  if (!cal) {
    float* c = stage.data() + (size_t)j * g.n;
    for (size_t k = 0; k < rel.size(); ++k) c[rel[k]] = vals[k];
    return;
  }
  bool contiguous = true;
  for (size_t k = 1; k < rel.size() && contiguous; ++k) contiguous = rel[k] == rel[0] + (int)k;
  if (contiguous) {
    copy_on(stream_, dev.p + (size_t)j * g.n + rel[0], vals, rel.size() * sizeof(float),
                     hipMemcpyHostToDevice);
    return;
  }
  std::vector<float> col(g.n);
  copy_on(stream_, col.data(), dev.p + (size_t)j * g.n, g.n * sizeof(float), hipMemcpyDeviceToHost);
  for (size_t k = 0; k < rel.size(); ++k) col[rel[k]] = vals[k];
  copy_on(stream_, dev.p + (size_t)j * g.n, col.data(), g.n * sizeof(float), hipMemcpyHostToDevice);
}

void Engine::set_param_index(Group& g, int j, const std::vector<int>& rel, const float* vals,
                             bool whole) {
  int gi = group_index(groups_, &g);
  auto& stage = stage_param_[gi];
  if (whole) {
    std::vector<float> col(g.n, vals[0]);
    col_write(stream_, calibrated_, g, stage, g.param, j, 0, g.n, col.data());
    if (!(g.md->rk5 && j == param_idx(g, "refractory_step", 0))) {
      g.uni[j] = 1;
      g.uni_val[j] = vals[0];
    }
    return;
  }
  col_write_entries(stream_, calibrated_, g, stage, g.param, j, rel, vals);
  // stays uniform only if every written value equals the uniform value (a
  // column made equal again by partial writes is just read per neuron)
  if (g.uni[j]) {
    for (size_t k = 0; k < rel.size(); ++k)
      if (vals[k] != g.uni_val[j]) {
        g.uni[j] = 0;
        break;
      }
  }
}

void Engine::set_var_index(Group& g, int j, const std::vector<int>& rel, const float* vals) {
  int gi = group_index(groups_, &g);
  auto& stage = stage_var_[gi];
  col_write_entries(stream_, calibrated_, g, stage, g.var, j, rel, vals);
}

static std::vector<int> iota_rel(int rel0, int n) {
  std::vector<int> r(n);
  std::iota(r.begin(), r.end(), rel0);
  return r;
}

void Engine::SetParam(int i_node, int n, const std::string& name, float val) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  auto& sp = g.md->scal_param;
  if (std::find(sp.begin(), sp.end(), name) == sp.end())
    throw Error("Unrecognized scalar parameter " + name);
  int j = param_idx(g, name, 0);
  std::vector<float> v(n, val);
  set_param_index(g, j, iota_rel(rel0, n), v.data(), rel0 == 0 && n == g.n);
}

void Engine::SetParam(const int* nodes, int n, const std::string& name, float val) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  auto& sp = g->md->scal_param;
  if (std::find(sp.begin(), sp.end(), name) == sp.end())
    throw Error("Unrecognized scalar parameter " + name);
  std::vector<float> v(n, val);
  set_param_index(*g, param_idx(*g, name, 0), rel, v.data(), false);
}

void Engine::SetParamArr(int i_node, int n, const std::string& name, const float* v, int size) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  auto& pp = g.md->port_param;
  if (std::find(pp.begin(), pp.end(), name) != pp.end()) {
    if (size != g.n_port) throw Error("Parameter array size must be equal to the number of ports.");
    for (int q = 0; q < size; ++q) {
      std::vector<float> vv(n, v[q]);
      set_param_index(g, param_idx(g, name, q), iota_rel(rel0, n), vv.data(), rel0 == 0 && n == g.n);
    }
    return;
  }
  auto& ap = g.md->array_param;
  if (std::find(ap.begin(), ap.end(), name) != ap.end()) {
    check_uncalibrated("Spike generator arrays must be set before calibration");
    for (int k = rel0; k < rel0 + n; ++k) {
      if (name == "spike_times") g.sg_times[k].assign(v, v + size);
      else g.sg_heights[k].assign(v, v + size);
    }
    return;
  }
  throw Error("Unrecognized array parameter " + name);
}

void Engine::SetParamArr(const int* nodes, int n, const std::string& name, const float* v, int size) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  auto& pp = g->md->port_param;
  if (std::find(pp.begin(), pp.end(), name) != pp.end()) {
    if (size != g->n_port) throw Error("Parameter array size must be equal to the number of ports.");
    for (int q = 0; q < size; ++q) {
      std::vector<float> vv(n, v[q]);
      set_param_index(*g, param_idx(*g, name, q), rel, vv.data(), false);
    }
    return;
  }
  auto& ap = g->md->array_param;
  if (std::find(ap.begin(), ap.end(), name) != ap.end()) {
    check_uncalibrated("Spike generator arrays must be set before calibration");
    for (int k : rel) {
      if (name == "spike_times") g->sg_times[k].assign(v, v + size);
      else g->sg_heights[k].assign(v, v + size);
    }
    return;
  }
  throw Error("Unrecognized array parameter " + name);
}

void Engine::SetVar(int i_node, int n, const std::string& name, float val) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  auto& sv = g.md->scal_var;
  if (std::find(sv.begin(), sv.end(), name) == sv.end())
    throw Error("Unrecognized scalar variable " + name);
  std::vector<float> v(n, val);
  set_var_index(g, var_idx(g, name, 0), iota_rel(rel0, n), v.data());
}

void Engine::SetVar(const int* nodes, int n, const std::string& name, float val) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  auto& sv = g->md->scal_var;
  if (std::find(sv.begin(), sv.end(), name) == sv.end())
    throw Error("Unrecognized scalar variable " + name);
  std::vector<float> v(n, val);
  set_var_index(*g, var_idx(*g, name, 0), rel, v.data());
}

void Engine::SetVarArr(int i_node, int n, const std::string& name, const float* v, int size) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  auto& pv = g.md->port_var;
  if (std::find(pv.begin(), pv.end(), name) == pv.end())
    throw Error("Unrecognized port variable " + name);
  if (size != g.n_port) throw Error("Variable array size must be equal to the number of ports.");
  for (int q = 0; q < size; ++q) {
    std::vector<float> vv(n, v[q]);
    set_var_index(g, var_idx(g, name, q), iota_rel(rel0, n), vv.data());
  }
}

void Engine::SetVarArr(const int* nodes, int n, const std::string& name, const float* v, int size) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  auto& pv = g->md->port_var;
  if (std::find(pv.begin(), pv.end(), name) == pv.end())
    throw Error("Unrecognized port variable " + name);
  if (size != g->n_port) throw Error("Variable array size must be equal to the number of ports.");
  for (int q = 0; q < size; ++q) {
    std::vector<float> vv(n, v[q]);
    set_var_index(*g, var_idx(*g, name, q), rel, vv.data());
  }
}

void Engine::SetIntVar(int i_node, int n, const std::string& name, int val) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (std::find(g.int_var.begin(), g.int_var.end(), name) == g.int_var.end())
    throw Error("Unrecognized integer variable " + name);
  if (!calibrated_) return;  // counters are created zeroed at calibration
  DBuf<int>& d = (name == "spike_count") ? g.spike_count : g.n_rec;
  std::vector<int> v(n, val);
  copy_on(stream_, d.p + rel0, v.data(), n * sizeof(int), hipMemcpyHostToDevice);
}

void Engine::SetIntVar(const int* nodes, int n, const std::string& name, int val) {
  for (int k = 0; k < n; ++k) SetIntVar(nodes[k], 1, name, val);
}

// Returned arrays stay valid until the next call of the same getter.
static std::vector<float>& ret_float(int which) {
  static std::vector<float> b[4];
  return b[which];
}
static std::vector<int>& ret_int() {
  static std::vector<int> b;
  return b;
}

float* Engine::GetParam(int i_node, int n, const std::string& name) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  int gi = group_index(groups_, &g);
  auto& sp = g.md->scal_param;
  auto& pp = g.md->port_param;
  std::vector<float>& out = ret_float(0);
  if (std::find(sp.begin(), sp.end(), name) != sp.end()) {
    out.resize(n);
    col_read(stream_, calibrated_, g, stage_param_[gi], g.param, param_idx(g, name, 0), rel0, n, out.data());
    return out.data();
  }
  if (std::find(pp.begin(), pp.end(), name) != pp.end()) {
    out.assign((size_t)n * g.n_port, 0.f);
    std::vector<float> col(n);
    for (int q = 0; q < g.n_port; ++q) {
      col_read(stream_, calibrated_, g, stage_param_[gi], g.param, param_idx(g, name, q), rel0, n, col.data());
      for (int k = 0; k < n; ++k) out[(size_t)k * g.n_port + q] = col[k];
    }
    return out.data();
  }
  auto& ap = g.md->array_param;
  if (std::find(ap.begin(), ap.end(), name) != ap.end()) {
    if (n != 1) throw Error("Cannot get array parameters for more than one node at a time");
    return GetArrayParam(i_node, name);
  }
  throw Error("Unrecognized parameter " + name);
}

float* Engine::GetParam(const int* nodes, int n, const std::string& name) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  auto& sp = g->md->scal_param;
  auto& pp = g->md->port_param;
  int gi = group_index(groups_, g);
  std::vector<float> col(g->n);
  std::vector<float>& out = ret_float(0);
  if (std::find(sp.begin(), sp.end(), name) != sp.end()) {
    col_read(stream_, calibrated_, *g, stage_param_[gi], g->param, param_idx(*g, name, 0), 0, g->n, col.data());
    out.resize(n);
    for (int k = 0; k < n; ++k) out[k] = col[rel[k]];
    return out.data();
  }
  if (std::find(pp.begin(), pp.end(), name) != pp.end()) {
    out.assign((size_t)n * g->n_port, 0.f);
    for (int q = 0; q < g->n_port; ++q) {
      col_read(stream_, calibrated_, *g, stage_param_[gi], g->param, param_idx(*g, name, q), 0, g->n, col.data());
      for (int k = 0; k < n; ++k) out[(size_t)k * g->n_port + q] = col[rel[k]];
    }
    return out.data();
  }
  auto& ap = g->md->array_param;
  if (std::find(ap.begin(), ap.end(), name) != ap.end()) {
    if (n != 1) throw Error("Cannot get array parameters for more than one node at a time");
    return GetArrayParam(nodes[0], name);
  }
  throw Error("Unrecognized parameter " + name);
}

float* Engine::GetArrayParam(int node, const std::string& name) {
  int rel;
  Group& g = GroupOf(node, 1, &rel);
  if (name == "spike_times" && g.md->id == MODEL_SPIKE_GEN) return g.sg_times[rel].data();
  if (name == "spike_heights" && g.md->id == MODEL_SPIKE_GEN) return g.sg_heights[rel].data();
  throw Error("Unrecognized parameter " + name);
}

float* Engine::GetVar(int i_node, int n, const std::string& name) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  int gi = group_index(groups_, &g);
  auto& sv = g.md->scal_var;
  auto& pv = g.md->port_var;
  std::vector<float>& out = ret_float(1);
  if (std::find(sv.begin(), sv.end(), name) != sv.end()) {
    out.resize(n);
    col_read(stream_, calibrated_, g, stage_var_[gi], g.var, var_idx(g, name, 0), rel0, n, out.data());
    return out.data();
  }
  if (std::find(pv.begin(), pv.end(), name) != pv.end()) {
    out.assign((size_t)n * g.n_port, 0.f);
    std::vector<float> col(n);
    for (int q = 0; q < g.n_port; ++q) {
      col_read(stream_, calibrated_, g, stage_var_[gi], g.var, var_idx(g, name, q), rel0, n, col.data());
      for (int k = 0; k < n; ++k) out[(size_t)k * g.n_port + q] = col[k];
    }
    return out.data();
  }
  throw Error("Unrecognized variable " + name);
}

float* Engine::GetVar(const int* nodes, int n, const std::string& name) {
  Group* g;
  auto rel = NodeListRel(nodes, n, &g);
  int gi = group_index(groups_, g);
  auto& sv = g->md->scal_var;
  auto& pv = g->md->port_var;
  std::vector<float> col(g->n);
  std::vector<float>& out = ret_float(1);
  if (std::find(sv.begin(), sv.end(), name) != sv.end()) {
    col_read(stream_, calibrated_, *g, stage_var_[gi], g->var, var_idx(*g, name, 0), 0, g->n, col.data());
    out.resize(n);
    for (int k = 0; k < n; ++k) out[k] = col[rel[k]];
    return out.data();
  }
  if (std::find(pv.begin(), pv.end(), name) != pv.end()) {
    out.assign((size_t)n * g->n_port, 0.f);
    for (int q = 0; q < g->n_port; ++q) {
      col_read(stream_, calibrated_, *g, stage_var_[gi], g->var, var_idx(*g, name, q), 0, g->n, col.data());
      for (int k = 0; k < n; ++k) out[(size_t)k * g->n_port + q] = col[rel[k]];
    }
    return out.data();
  }
  throw Error("Unrecognized variable " + name);
}

float* Engine::GetArrayVar(int, const std::string& name) {
  throw Error("Unrecognized array variable " + name);
}

int* Engine::GetIntVar(int i_node, int n, const std::string& name) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (std::find(g.int_var.begin(), g.int_var.end(), name) == g.int_var.end())
    throw Error("Unrecognized integer variable " + name);
  std::vector<int>& out = ret_int();
  out.assign(n, 0);
  if (calibrated_) {
    DBuf<int>& d = (name == "spike_count") ? g.spike_count : g.n_rec;
    sync_streams();
    copy_on(stream_, out.data(), d.p + rel0, n * sizeof(int), hipMemcpyDeviceToHost);
  }
  return out.data();
}

int* Engine::GetIntVar(const int* nodes, int n, const std::string& name) {
  std::vector<int> tmp(n);
  for (int k = 0; k < n; ++k) tmp[k] = GetIntVar(nodes[k], 1, name)[0];
  std::vector<int>& out = ret_int();
  out = tmp;
  return out.data();
}

int Engine::SetGroupParam(int i_node, int n, const std::string& name, float val) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (rel0 != 0 || n != g.n)
    throw Error("Group parameter must be set for all and only the nodes of the same group");
  if (name == "h_min_rel") g.h_min_rel = val;
  else if (name == "h0_rel") g.h0_rel = val;
  else throw Error("Unrecognized neuron group parameter " + name);
  if (calibrated_ && g.md->rk5) {
    g.h_min = g.h_min_rel * dt_;
    g.h0 = g.h0_rel * dt_;
  }
  return 0;
}

float Engine::GetGroupParam(int node, const std::string& name) {
  int rel;
  Group& g = GroupOf(node, 1, &rel);
  if (!g.md->rk5) throw Error("Unrecognized neuron group parameter " + name);
  if (name == "h_min_rel") return g.h_min_rel;
  if (name == "h0_rel") return g.h0_rel;
  throw Error("Unrecognized neuron group parameter " + name);
}

std::vector<std::string> Engine::ScalVarNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->scal_var; }
std::vector<std::string> Engine::PortVarNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->port_var; }
std::vector<std::string> Engine::ScalParamNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->scal_param; }
std::vector<std::string> Engine::PortParamNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->port_param; }
std::vector<std::string> Engine::GroupParamNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->group_param; }
std::vector<std::string> Engine::ArrayParamNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->md->array_param; }
std::vector<std::string> Engine::ArrayVarNames(int) { return {}; }
std::vector<std::string> Engine::IntVarNames(int node) { return group_at(groups_, node_group_, node, n_nodes_)->int_var; }

// ---- host random numbers ----------------------------------------------------
void Engine::random_block(uint32_t out[4]) {
  U32x4 c{(uint32_t)rng_ctr_, (uint32_t)(rng_ctr_ >> 32), 0x686f7374u, 0u};
  ++rng_ctr_;
  U32x4 r = philox4x32_10(c, (uint32_t)seed_, (uint32_t)(seed_ >> 32));
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

unsigned int* Engine::RandomInt(size_t n) {
  unsigned int* a = new unsigned int[n];
  uint32_t b[4];
  for (size_t i = 0; i < n; i += 4) {
    random_block(b);
    for (size_t k = 0; k < 4 && i + k < n; ++k) a[i + k] = b[k];
  }
  return a;
}

// uniform on (0, 1]: ((u >> 8) + 1) * 2^-24, exact in float
float* Engine::RandomUniform(size_t n) {
  float* a = new float[n];
  uint32_t b[4];
  for (size_t i = 0; i < n; i += 4) {
    random_block(b);
    for (size_t k = 0; k < 4 && i + k < n; ++k) a[i + k] = ((b[k] >> 8) + 1) * (1.0f / 16777216.0f);
  }
  return a;
}

// Box-Muller on pairs of uniforms (Box & Muller 1958)
float* Engine::RandomNormal(size_t n, float mean, float sd) {
  float* a = new float[n];
  uint32_t b[4];
  for (size_t i = 0; i < n; i += 4) {
    random_block(b);
    for (int k = 0; k < 4; k += 2) {
      double u1 = ((b[k] >> 8) + 1) * (1.0 / 16777216.0);
      double u2 = (b[k + 1] >> 8) * (1.0 / 16777216.0);
      double rr = std::sqrt(-2.0 * std::log(u1));
      double z0 = rr * std::cos(2.0 * M_PI * u2), z1 = rr * std::sin(2.0 * M_PI * u2);
      if (i + k < n) a[i + k] = (float)(mean + sd * z0);
      if (i + k + 1 < n) a[i + k + 1] = (float)(mean + sd * z1);
    }
  }
  return a;
}

// Normal draws redrawn until inside [vmin, vmax], then optionally rounded to
// the grid vmin + k*vstep (the reference's rule, src/nestgpu.cu:1150-1188).
float* Engine::RandomNormalClipped(size_t n, float mean, float sd, float vmin, float vmax, float vstep) {
  float* a = RandomNormal(n, mean, sd);
  for (size_t i = 0; i < n; ++i) {
    int tries = 0;
    while (a[i] < vmin || a[i] > vmax) {
      float* e = RandomNormal(2, mean, sd);
      a[i] = e[0];
      delete[] e;
      if (++tries > 1000000) throw Error("normal_clipped: interval has negligible probability");
    }
  }
  if (vstep > sd * 1.0e-6f)
    for (size_t i = 0; i < n; ++i) a[i] = vmin + vstep * std::round((a[i] - vmin) / vstep);
  return a;
}

// ---- synapse groups, src/syn_model.cu:136-238 -------------------------------
// parameter names and defaults: src/stdp.cu:64-79, src/test_syn_model.cu:43-52
static const std::vector<std::string> kStdpParams = {"tau_plus", "tau_minus", "lambda", "alpha",
                                                     "mu_plus", "mu_minus", "Wmax"};
static const std::vector<std::string> kTestSynParams = {"fact", "offset"};

static const std::vector<std::string>& syn_param_names(int model) {
  return model == kSynModelStdp ? kStdpParams : kTestSynParams;
}

int Engine::CreateSynGroup(const std::string& model) {
  check_uncalibrated("Nodes cannot be created after calibration");
  SynGroupS g;
  if (model == "stdp") {
    g.model = kSynModelStdp;
    g.p = {20.0f, 20.0f, 1.0e-4f, 1.0f, 1.0f, 1.0f, 100.0f};
  } else if (model == "test_syn_model") {
    g.model = kSynModelTest;
    g.p = {0.1f, 0.0f};
  } else {
    throw Error("Unknown synapse model name: " + model);
  }
  if (syn_groups_.size() >= 256) throw Error("Too many synapse groups");
  syn_groups_.push_back(g);
  return (int)syn_groups_.size() - 1;
}

const Engine::SynGroupS& Engine::syn_group_at(int g) const {
  if (g < 1 || g >= (int)syn_groups_.size()) throw Error("Unrecognized synapse group");
  return syn_groups_[g];
}

int Engine::GetSynGroupNParam(int g) { return (int)syn_group_at(g).p.size(); }

std::vector<std::string> Engine::GetSynGroupParamNames(int g) {
  return syn_param_names(syn_group_at(g).model);
}

bool Engine::IsSynGroupParam(int g, const std::string& name) {
  const auto& v = syn_param_names(syn_group_at(g).model);
  return std::find(v.begin(), v.end(), name) != v.end();
}

int Engine::GetSynGroupParamIdx(int g, const std::string& name) {
  const auto& v = syn_param_names(syn_group_at(g).model);
  auto it = std::find(v.begin(), v.end(), name);
  if (it == v.end()) throw Error("Unrecognized parameter " + name);
  return (int)(it - v.begin());
}

float Engine::GetSynGroupParam(int g, const std::string& name) {
  if (!IsSynGroupParam(g, name)) throw Error("Unrecognized synapse parameter " + name);
  return syn_groups_[g].p[GetSynGroupParamIdx(g, name)];
}

void Engine::SetSynGroupParam(int g, const std::string& name, float v) {
  if (!IsSynGroupParam(g, name)) throw Error("Unrecognized synapse parameter " + name);
  syn_groups_[g].p[GetSynGroupParamIdx(g, name)] = v;
  if (calibrated_ && plastic_) {
    std::vector<float> tab(syn_groups_.size() * kSynGroupParams, 0.0f);
    for (size_t i = 1; i < syn_groups_.size(); ++i) {
      tab[i * kSynGroupParams] = (float)syn_groups_[i].model;
      for (size_t j = 0; j < syn_groups_[i].p.size(); ++j) tab[i * kSynGroupParams + 1 + j] = syn_groups_[i].p[j];
    }
    upload(d_sg_param_, tab, stream_);
    sync_streams();
  }
}

// ---- connections ------------------------------------------------------------
int Engine::Connect(int src, int tgt, int port, int syn_group, float w, float delay) {
  check_uncalibrated("Connections cannot be created after calibration");
  if (src < 0 || src >= n_nodes_ || tgt < 0 || tgt >= n_nodes_)
    throw Error("Unrecognized node in connection");
  if (delay < dt_) throw Error("Delay must be >= time resolution");
  if (syn_group < 0 || syn_group >= (int)syn_groups_.size())
    throw Error("Unrecognized synapse group");
  if (port < 0 || port > 255) throw Error("Port index must be >=0 and < 256");
  int d_int = (int)std::round(delay / dt_) - 1;   // src/connect.cu:45
  if (d_int > 65535) throw Error("Delay too large (max 65536 steps)");
  c_src_.push_back(src);
  c_tgt_.push_back(tgt);
  c_port_.push_back((unsigned char)port);
  c_syn_.push_back((unsigned char)syn_group);
  c_w_.push_back(w);
  c_d_.push_back(d_int);
  sorted_ = false;
  return 0;
}

int Engine::ConnectRule(const int* src_list, int i_src0, int n_src, const int* tgt_list,
                        int i_tgt0, int n_tgt, const ConnSpecS& cs, const SynSpecS& ss) {
  check_uncalibrated("Connections cannot be created after calibration");
  if (ss.syn_group < 0 || ss.syn_group >= (int)syn_groups_.size())
    throw Error("Unrecognized synapse group");
  auto S = [&](int k) { return src_list ? src_list[k] : i_src0 + k; };
  auto T = [&](int k) { return tgt_list ? tgt_list[k] : i_tgt0 + k; };
  for (int k = 0; k < n_src; ++k)
    if (S(k) < 0 || S(k) >= n_nodes_) throw Error("Unrecognized source node");
  for (int k = 0; k < n_tgt; ++k)
    if (T(k) < 0 || T(k) >= n_nodes_) throw Error("Unrecognized target node");
  // Connections are appended in the reference's order: for rule connection
  // k the source index, target index and synapse-array index i_array are
  //   one_to_one: (k, k, k); all_to_all: (k % n_src, k / n_src, k)
  //   (target-major, src/connect_rules.h:173-185); fixed_total_number:
  //   (r[2k] % n_src, r[2k+1] % n_tgt, k); fixed_indegree: (pick[k], k / deg,
  //   k); fixed_outdegree: (k / deg, pick[k], k).
  long long n = 0;
  int deg = 0;
  std::vector<int> chosen;
  unsigned int* rnd = nullptr;
  switch (cs.rule) {
    case 0:
      if (n_src != n_tgt)
        throw Error("Number of source and target nodes must be equal for the one-to-one connection rule");
      n = n_src;
      break;
    case 1:
      n = (long long)n_src * n_tgt;
      break;
    case 2:
      n = cs.total_num;
      if (device_connect(n)) {
        // the host generator's next 2n words, drawn on the GPU
        rnd = new unsigned int[2 * (size_t)n];
        device_random_words(rng_ctr_, seed_, 2 * (size_t)n, rnd, stream_);
        rng_ctr_ += (2 * (uint64_t)n + 3) / 4;
      } else {
        rnd = RandomInt(2 * (size_t)cs.total_num);
      }
      break;
    case 3:    // fixed_indegree: distinct sources per target, drawn per target
    case 4: {  // fixed_outdegree: distinct targets per source, drawn per source
      const bool in = cs.rule == 3;
      deg = in ? cs.indegree : cs.outdegree;
      const int n_pool = in ? n_src : n_tgt;
      const int n_own = in ? n_tgt : n_src;
      if (deg > n_pool)
        throw Error(in ? "Indegree larger than number of source nodes"
                       : "Outdegree larger than number of target nodes");
      n = (long long)n_own * deg;
      const uint64_t key = seed_ ^ ((uint64_t)(++rng_ctr_) << 40);
      chosen.resize(n);
      if (device_connect(n)) {
        device_pick_distinct(n_own, deg, n_pool, key, chosen.data(), stream_);
        break;
      }
#pragma omp parallel
      {
        std::vector<char> used(n_pool, 0);
        std::vector<int> perm;
#pragma omp for schedule(dynamic, 64)
        for (int o = 0; o < n_own; ++o) {
          uint32_t blk = 0;
          U32x4 r{};
          int avail = 0;
          auto next = [&]() -> uint32_t {
            if (avail == 0) {
              r = philox4x32_10(U32x4{blk++, (uint32_t)o, 0x72756c65u, 0u}, (uint32_t)key,
                                (uint32_t)(key >> 32));
              avail = 4;
            }
            return pick(r, 4 - avail--);
          };
          int* out = chosen.data() + (size_t)o * deg;
          if (n_pool < 5 * deg) {  // partial Fisher-Yates over the pool
            perm.resize(n_pool);
            std::iota(perm.begin(), perm.end(), 0);
            for (int i = 0; i < deg; ++i) {
              int j = i + (int)(next() % (uint32_t)(n_pool - i));
              std::swap(perm[i], perm[j]);
              out[i] = perm[i];
            }
          } else {  // rejection with a marker array
            for (int i = 0; i < deg; ++i) {
              int j;
              do { j = (int)(next() % (uint32_t)n_pool); } while (used[j]);
              used[j] = 1;
              out[i] = j;
            }
            for (int i = 0; i < deg; ++i) used[out[i]] = 0;
          }
        }
      }
      break;
    }
    default:
      throw Error("Unknown connection rule");
  }
  const int rule = cs.rule;
  auto pair_of = [&](long long k, int* is, int* it) {
    switch (rule) {
      case 0: *is = (int)k; *it = (int)k; break;
      case 1: *is = (int)(k % n_src); *it = (int)(k / n_src); break;
      case 2: *is = (int)(rnd[2 * k] % (unsigned)n_src); *it = (int)(rnd[2 * k + 1] % (unsigned)n_tgt); break;
      case 3: *is = chosen[k]; *it = (int)(k / deg); break;
      default: *is = (int)(k / deg); *it = chosen[k]; break;
    }
  };
  const size_t base = c_src_.size();
  c_src_.resize(base + n); c_tgt_.resize(base + n); c_port_.resize(base + n);
  c_syn_.resize(base + n); c_w_.resize(base + n); c_d_.resize(base + n);
  int bad_delay = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_delay)
  for (long long k = 0; k < n; ++k) {
    int is, it;
    pair_of(k, &is, &it);
    const float w = ss.weight_array ? ss.weight_array[k] : ss.weight;
    const float d = ss.delay_array ? ss.delay_array[k] : ss.delay;
    if (d < dt_) bad_delay |= 1;
    const int d_int = (int)std::round(d / dt_) - 1;
    if (d_int > 65535) bad_delay |= 2;
    c_src_[base + k] = S(is);
    c_tgt_[base + k] = T(it);
    c_port_[base + k] = (unsigned char)ss.port;
    c_syn_[base + k] = (unsigned char)ss.syn_group;
    c_w_[base + k] = w;
    c_d_[base + k] = d_int;
  }
  delete[] rnd;
  if (bad_delay) {  // leave the store as it was
    c_src_.resize(base); c_tgt_.resize(base); c_port_.resize(base);
    c_syn_.resize(base); c_w_.resize(base); c_d_.resize(base);
    if (bad_delay & 1) throw Error("Delay must be >= time resolution");
    throw Error("Delay too large (max 65536 steps)");
  }
  if (n) sorted_ = false;
  return 0;
}

// Random connection rules draw on the GPU (connect_device.hip, the same
// Philox words as the host loops) from 2^20 connections on;
// NGPU_DEVICE_CONNECT=1 always, 0 never.
bool Engine::device_connect(long long n) const {
  if (const char* env = getenv("NGPU_DEVICE_CONNECT")) return atoi(env) != 0 && n > 0;
  return n >= (1ll << 20);
}

// Stable sort of the store by (source, d_int): counting sort by source, then
// a stable sort by delay inside each source segment.
void Engine::sort_connections() {
  const size_t nc = c_src_.size();
  if (sorted_ && row_ptr_.size() == (size_t)n_nodes_ + 1) return;
  if (!sort_connections_device()) sort_connections_host();
  // delay groups: count per source, prefix, fill (sources in parallel)
  grp_ptr_.assign(n_nodes_ + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int s = 0; s < n_nodes_; ++s) {
    long long g = 0;
    for (long long k = row_ptr_[s]; k < row_ptr_[s + 1]; ++k)
      if (k == row_ptr_[s] || c_d_[k] != c_d_[k - 1]) ++g;
    grp_ptr_[s + 1] = g;
  }
  for (int s = 0; s < n_nodes_; ++s) grp_ptr_[s + 1] += grp_ptr_[s];
  grp_start_.assign((size_t)grp_ptr_[n_nodes_] + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int s = 0; s < n_nodes_; ++s) {
    long long g = grp_ptr_[s];
    for (long long k = row_ptr_[s]; k < row_ptr_[s + 1]; ++k)
      if (k == row_ptr_[s] || c_d_[k] != c_d_[k - 1]) grp_start_[g++] = k;
  }
  grp_start_[grp_ptr_[n_nodes_]] = (long long)nc;
  sorted_ = true;
}

// The same order on the GPU (device_sort_connections: a stable radix sort of
// (source << 16 | d_int) keys and the gathers on the device).  Used
// from 2^20 connections on (NGPU_DEVICE_SORT=1 always, 0 never); false when
// not used.
bool Engine::sort_connections_device() {
  const size_t nc = c_src_.size();
  int mode = -1;
  if (const char* env = getenv("NGPU_DEVICE_SORT")) mode = atoi(env) != 0 ? 1 : 0;
  if (mode == 0 || nc == 0 || nc >= (1ull << 32) || (mode < 0 && nc < (1u << 20))) return false;
  std::vector<long long> rp;
  const auto t0 = std::chrono::steady_clock::now();
  if (!device_sort_connections(c_src_.data(), c_tgt_.data(), c_port_.data(), c_w_.data(), c_d_.data(), nc,
                               n_nodes_, rp, stream_))
    return false;
  row_ptr_.swap(rp);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int s = 0; s < n_nodes_; ++s)
    for (long long k = row_ptr_[s]; k < row_ptr_[s + 1]; ++k) c_src_[k] = s;
  // c_syn_ holds synapse group 0 everywhere (no other group is accepted)
  if (getenv("NGPU_CALIB_LOG") && atoi(getenv("NGPU_CALIB_LOG")) != 0)
    fprintf(stderr, "calibrate:   device sort and gathers %.3f s\n",
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return true;
}

// Counting sort by source (creation order kept), then a stable sort by
// delay inside each source segment.
void Engine::sort_connections_host() {
  const size_t nc = c_src_.size();
  std::vector<long long> cnt(n_nodes_ + 1, 0);
  for (size_t k = 0; k < nc; ++k) cnt[c_src_[k] + 1]++;
  for (int s = 0; s < n_nodes_; ++s) cnt[s + 1] += cnt[s];
  row_ptr_ = cnt;
  std::vector<long long> dest(nc);
  {
    std::vector<long long> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < nc; ++k) dest[k] = pos[c_src_[k]]++;
  }
  auto permute = [&](auto& v) {
    using T = typename std::decay<decltype(v)>::type::value_type;
    std::vector<T> o(nc);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < (long long)nc; ++k) o[dest[k]] = v[k];
    v.swap(o);
  };
  permute(c_src_); permute(c_tgt_); permute(c_port_); permute(c_syn_); permute(c_w_); permute(c_d_);
  dest.clear();
  dest.shrink_to_fit();
  // stable sort by delay within each source segment
#pragma omp parallel
  {
    std::vector<long long> idx;
    std::vector<int> t_tgt, t_d;
    std::vector<unsigned char> t_port, t_syn;
    std::vector<float> t_w;
#pragma omp for schedule(dynamic, 256)
    for (int s = 0; s < n_nodes_; ++s) {
      long long b = row_ptr_[s], e = row_ptr_[s + 1];
      bool ok = true;
      for (long long k = b + 1; k < e && ok; ++k) ok = c_d_[k - 1] <= c_d_[k];
      if (ok) continue;
      idx.resize(e - b);
      std::iota(idx.begin(), idx.end(), b);
      std::stable_sort(idx.begin(), idx.end(), [&](long long x, long long y) { return c_d_[x] < c_d_[y]; });
      t_tgt.resize(e - b); t_d.resize(e - b); t_port.resize(e - b); t_syn.resize(e - b); t_w.resize(e - b);
      for (long long k = 0; k < e - b; ++k) {
        t_tgt[k] = c_tgt_[idx[k]]; t_d[k] = c_d_[idx[k]]; t_port[k] = c_port_[idx[k]];
        t_syn[k] = c_syn_[idx[k]]; t_w[k] = c_w_[idx[k]];
      }
      for (long long k = 0; k < e - b; ++k) {
        c_tgt_[b + k] = t_tgt[k]; c_d_[b + k] = t_d[k]; c_port_[b + k] = t_port[k];
        c_syn_[b + k] = t_syn[k]; c_w_[b + k] = t_w[k];
      }
    }
  }
}

std::vector<long long> Engine::GetConnections(const int* src_list, int i_src0, int n_src,
                                              const int* tgt_list, int i_tgt0, int n_tgt,
                                              int syn_group) {
  sort_connections();
  if (!src_list && n_src <= 0) { i_src0 = 0; n_src = n_nodes_; }
  if (!tgt_list && n_tgt <= 0) { i_tgt0 = 0; n_tgt = n_nodes_; }
  std::vector<int> tv;
  if (tgt_list) { tv.assign(tgt_list, tgt_list + n_tgt); std::sort(tv.begin(), tv.end()); }
  std::vector<long long> out;
  for (int k = 0; k < n_src; ++k) {
    int s = src_list ? src_list[k] : i_src0 + k;
    if (s < 0 || s >= n_nodes_) throw Error("Unrecognized source node");
    for (long long gi = grp_ptr_[s]; gi < grp_ptr_[s + 1]; ++gi) {
      for (long long c = grp_start_[gi]; c < grp_start_[gi + 1]; ++c) {
        int t = c_tgt_[c];
        bool hit = tgt_list ? std::binary_search(tv.begin(), tv.end(), t)
                            : (t >= i_tgt0 && t < i_tgt0 + n_tgt);
        if (syn_group >= 0 && c_syn_[c] != syn_group) hit = false;
        if (hit) {
          out.push_back(s);
          out.push_back(gi - grp_ptr_[s]);
          out.push_back(c - grp_start_[gi]);
        }
      }
    }
  }
  return out;
}

void Engine::GetConnectionStatus(int src, int grp, int conn, int* tgt, unsigned char* port,
                                 unsigned char* syn, float* delay, float* w) {
  sort_connections();
  if (src < 0 || src >= n_nodes_) throw Error("Unrecognized source node");
  long long gi = grp_ptr_[src] + grp;
  if (grp < 0 || gi >= grp_ptr_[src + 1]) throw Error("Unrecognized connection group");
  long long c = grp_start_[gi] + conn;
  if (conn < 0 || c >= grp_start_[gi + 1]) throw Error("Unrecognized connection index");
  *tgt = c_tgt_[c];
  *port = c_port_[c];
  *syn = c_syn_[c];
  *delay = dt_ * (c_d_[c] + 1);   // src/connect.cu:168
  *w = c_w_[c];
  if (plastic_ && calibrated_ && c_syn_[c] != 0)   // its weight changes on the device
    copy_on(stream_, w, d_syn_w_.p + c, sizeof(float), hipMemcpyDeviceToHost);
}

void Engine::OutDegrees(int i_node, int n, long long* out) {
  sort_connections();
  if (i_node < 0 || n < 0 || i_node + n > n_nodes_) throw Error("Unrecognized node");
  for (int k = 0; k < n; ++k) out[k] = row_ptr_[i_node + k + 1] - row_ptr_[i_node + k];
}

// Per-thread ranges of the source-major table holding synapses into the
// marked targets, filled in table order.
static std::vector<char> target_mask(const int* targets, int n, int n_nodes) {
  std::vector<char> m(n_nodes, 0);
  for (int k = 0; k < n; ++k) {
    if (targets[k] < 0 || targets[k] >= n_nodes) throw Error("Unrecognized target node");
    m[targets[k]] = 1;
  }
  return m;
}

long long Engine::CountIncoming(const int* targets, int n) {
  sort_connections();
  const std::vector<char> m = target_mask(targets, n, n_nodes_);
  long long cnt = 0;
  const long long nc = (long long)c_tgt_.size();
#pragma omp parallel for schedule(static) reduction(+ : cnt)
  for (long long c = 0; c < nc; ++c) cnt += m[c_tgt_[c]];
  return cnt;
}

void Engine::GetIncoming(const int* targets, int n, int* src, int* tgt, unsigned char* port,
                         float* w, float* delay, long long* index) {
  sort_connections();
  const std::vector<char> m = target_mask(targets, n, n_nodes_);
  const long long nc = (long long)c_tgt_.size();
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  std::vector<long long> per(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    const long long b = nc * t / nt, e = nc * (t + 1) / nt;
    long long c0 = 0;
    for (long long c = b; c < e; ++c) c0 += m[c_tgt_[c]];
    per[t + 1] = c0;
#pragma omp barrier
#pragma omp single
    for (int k = 0; k < nt; ++k) per[k + 1] += per[k];
    long long o = per[t];
    for (long long c = b; c < e; ++c) {
      if (!m[c_tgt_[c]]) continue;
      src[o] = c_src_[c];
      tgt[o] = c_tgt_[c];
      port[o] = c_port_[c];
      w[o] = c_w_[c];
      delay[o] = dt_ * (c_d_[c] + 1);   // src/connect.cu:168
      index[o] = c;
      ++o;
    }
  }
}

// ---- recording ----------------------------------------------------------------
int Engine::CreateRecord(const std::string& file, char** var_names, const int* nodes,
                         const int* ports, int n) {
  RecordDef r;
  r.file = file;
  for (int k = 0; k < n; ++k) {
    int rel;
    Group& g = GroupOf(nodes[k], 1, &rel);
    std::string v = var_names[k];
    if (v != "spike" && var_idx(g, v, 0) < 0) throw Error("Unrecognized variable " + v);
    if (ports[k] < 0 || (ports[k] != 0 && ports[k] >= g.n_port)) throw Error("Port index out of range");
    r.group.push_back(group_index(groups_, &g));
    r.rel.push_back(rel);
    r.port.push_back(ports[k]);
    r.node.push_back(nodes[k]);
    r.var.push_back(v);
  }
  if (!file.empty()) {
    r.fp = fopen(file.c_str(), "w");
    if (!r.fp) throw Error("Cannot open record file " + file);
  }
  records_.push_back(std::move(r));
  return (int)records_.size() - 1;
}

RecordDef& Engine::Record(int i) {
  if (i < 0 || i >= (int)records_.size()) throw Error("Record does not exist.");
  RecordDef& r = records_[i];
  r.rows.resize(r.data.size());
  for (size_t k = 0; k < r.data.size(); ++k) r.rows[k] = r.data[k].data();
  return r;
}

int Engine::ActivateSpikeCount(int i_node, int n) {
  check_uncalibrated("Spike count must be activated before calibration");
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (rel0 != 0 || n != g.n)
    throw Error("Spike count must be activated for all and only  the nodes of the same group");
  if (std::find(g.int_var.begin(), g.int_var.end(), "spike_count") != g.int_var.end())
    throw Error("Spike count already activated");
  g.int_var.push_back("spike_count");
  return 0;
}

int Engine::ActivateRecSpikeTimes(int i_node, int n, int max_n) {
  check_uncalibrated("Spike time recording must be activated before calibration");
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (rel0 != 0 || n != g.n)
    throw Error("Spike count must be activated for all and only  the nodes of the same group");
  if (max_n <= 0) throw Error("Maximum number of recorded spike times must be greater than 0");
  if (g.max_rec > 0) throw Error("Spike times recording already activated");
  g.int_var.push_back("n_rec_spike_times");
  g.max_rec = max_n;
  g.rec_host.assign(g.n, {});
  return 0;
}

int Engine::SetRecSpikeTimesStep(int i_node, int n, int step) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (rel0 != 0 || n != g.n)
    throw Error("Time step for buffering spike time recording must be set for all and only the nodes of the same group");
  g.rec_step = step;
  return 0;
}

int Engine::GetNRecSpikeTimes(int node) {
  int rel;
  Group& g = GroupOf(node, 1, &rel);
  if (g.max_rec <= 0) throw Error("Spike times recording was not activated");
  if (!calibrated_) return 0;
  int v;
  sync_streams();
  copy_on(stream_, &v, g.n_rec.p + rel, sizeof(int), hipMemcpyDeviceToHost);
  return v;
}

// Move the device spike-time buffers to the host and reset the counters
// (src/base_neuron.cu:1488-1526).
void Engine::buffer_rec_spike_times(Group& g) {
  if (!calibrated_ || g.max_rec <= 0) return;
  sync_streams();
  std::vector<int> cnt(g.n);
  copy_on(stream_, cnt.data(), g.n_rec.p, g.n * sizeof(int), hipMemcpyDeviceToHost);
  long long tot = 0;
  for (int c : cnt) tot += c;
  if (tot == 0) return;
  std::vector<float> t((size_t)g.n * g.max_rec);
  copy_on(stream_, t.data(), g.rec_times.p, t.size() * sizeof(float), hipMemcpyDeviceToHost);
  for (int i = 0; i < g.n; ++i)
    g.rec_host[i].insert(g.rec_host[i].end(), t.begin() + (size_t)i * g.max_rec,
                         t.begin() + (size_t)i * g.max_rec + cnt[i]);
  // stream-ordered: the next step's kernels read the fill counts
  HIPCHK(hipMemsetAsync(g.n_rec.p, 0, g.n * sizeof(int), stream_));
}

int Engine::GetRecSpikeTimes(int i_node, int n, int** n_pt, float*** t_pt) {
  int rel0;
  Group& g = GroupOf(i_node, n, &rel0);
  if (rel0 != 0 || n != g.n)
    throw Error("Spike times must be extracted for all and only  the nodes of the same group");
  if (g.max_rec <= 0) throw Error("Spike times recording was not activated");
  buffer_rec_spike_times(g);
  g.rec_n_host.resize(g.n);
  g.rec_pt_host.resize(g.n);
  for (int i = 0; i < g.n; ++i) {
    g.rec_n_host[i] = (int)g.rec_host[i].size();
    g.rec_pt_host[i] = g.rec_host[i].data();
  }
  *n_pt = g.rec_n_host.data();
  *t_pt = g.rec_pt_host.data();
  return 0;
}

// ---- calibration --------------------------------------------------------------
void Engine::calibrate_group(Group& g) {
  int gi = group_index(groups_, &g);
  std::vector<float>& par = stage_param_[gi];
  const int n = g.n;
  auto col = [&](int j) { return par.data() + (size_t)j * n; };
  const ModelDesc& m = *g.md;
  const int ns = (int)m.scal_param.size(), npp = (int)m.port_param.size();
  if (m.rk5) {
    g.h_min = g.h_min_rel * dt_;
    g.h0 = g.h0_rel * dt_;
    float* ref = col(param_idx(g, "refractory_step", 0));
    std::fill(ref, ref + n, 0.f);
    // beta normalisation, src/aeif_cond_beta_multisynapse.cu:72-99
    // (src/aeif_cond_beta.cu:81-113 per ex/in port)
    auto beta_g0 = [&](int jr, int jd, int jg) {
      auto g0_of = [](float tr, float td) -> float {
        float denom1 = td - tr, denom2 = 0;
        if (denom1 != 0) {
          float t_p = td * tr * std::log(td / tr) / denom1;
          denom2 = std::exp(-t_p / td) - std::exp(-t_p / tr);
        }
        if (denom2 == 0) return (float)(M_E / td);
        return (float)((1. / tr - 1. / td) / denom2);
      };
      float* gc = col(jg);
      const float* trc = col(jr);
      const float* tdc = col(jd);
      for (int i = 0; i < n; ++i) gc[i] = g0_of(trc[i], tdc[i]);
      g.uni[jg] = g.uni[jr] && g.uni[jd];
      g.uni_val[jg] = gc[0];
    };
    // alpha normalisation e / tau (src/aeif_cond_alpha_multisynapse.cu:79,
    // src/aeif_cond_alpha.cu, src/aeif_psc_alpha.cu:78-79)
    auto alpha_g0 = [&](int js, int jg) {
      float* gc = col(jg);
      const float* ts = col(js);
      for (int i = 0; i < n; ++i) gc[i] = (float)(M_E / ts[i]);
      g.uni[jg] = g.uni[js];
      g.uni_val[jg] = gc[0];
    };
    auto P = [&](const char* nm, int port = 0) { return param_idx(g, nm, port); };
    switch (m.id) {
      case MODEL_AEIF_COND_BETA_MS:
      case MODEL_IZH_COND_BETA:
        for (int q = 0; q < g.n_port; ++q)
          beta_g0(ns + q * npp + 1, ns + q * npp + 2, ns + q * npp + 3);
        break;
      case MODEL_AEIF_COND_ALPHA_MS:
        for (int q = 0; q < g.n_port; ++q) alpha_g0(ns + q * npp + 1, ns + q * npp + 2);
        break;
      case MODEL_AEIF_COND_BETA:
        beta_g0(P("tau_rise_ex"), P("tau_decay_ex"), P("g0_ex"));
        beta_g0(P("tau_rise_in"), P("tau_decay_in"), P("g0_in"));
        break;
      case MODEL_AEIF_COND_ALPHA:
        alpha_g0(P("tau_syn_ex"), P("g0_ex"));
        alpha_g0(P("tau_syn_in"), P("g0_in"));
        break;
      case MODEL_AEIF_PSC_ALPHA:
        alpha_g0(P("tau_syn_ex"), P("I0_ex"));
        alpha_g0(P("tau_syn_in"), P("I0_in"));
        break;
      case MODEL_AEIF_PSC_ALPHA_MS:
        for (int q = 0; q < g.n_port; ++q) alpha_g0(P("tau_syn", q), P("I0", q));
        break;
      default:
        break;
    }
    if (m.id == MODEL_AEIF_PSC_EXP || m.id == MODEL_AEIF_PSC_ALPHA ||
        m.id == MODEL_AEIF_PSC_ALPHA_MS || m.id == MODEL_AEIF_PSC_DELTA ||
        m.id == MODEL_AEIF_PSC_EXP_MS) {
      // Delta_T <= 0: threshold at V_th (src/aeif_psc_exp.cu, src/aeif_psc_alpha.cu:75-77)
      int jdt = P("Delta_T"), jth = P("V_th"), jpk = P("V_peak");
      bool changed = false;
      for (int i = 0; i < n; ++i)
        if (col(jdt)[i] <= 0.0f) { col(jpk)[i] = col(jth)[i]; changed = true; }
      if (changed) {
        bool same = std::all_of(col(jpk), col(jpk) + n, [&](float v) { return v == col(jpk)[0]; });
        g.uni[jpk] = same;
        g.uni_val[jpk] = col(jpk)[0];
      }
    }
  } else if (m.id == MODEL_IAF_PSC_EXP) {
    // propagators, src/iaf_psc_exp.cu:68-82 (FP64 propagator_32)
    auto p32 = [](double tau_syn, double tau, double C, double h) {
      const double lin = 1.0 / (2.0 * C * tau * tau) * h * h * (tau_syn - tau) * std::exp(-h / tau);
      const double sing = h / C * std::exp(-h / tau);
      const double p = -tau / (C * (1.0 - tau / tau_syn)) * std::exp(-h / tau_syn) *
                       std::expm1(h * (1.0 / tau_syn - 1.0 / tau));
      const double dev = std::fabs(p - sing);
      if (tau == tau_syn || (std::fabs(tau - tau_syn) < 0.1 && dev > 2.0 * std::fabs(lin))) return sing;
      return p;
    };
    enum { tau_m, C_m, E_L, I_e, Theta, V_reset, tau_ex, tau_in, t_ref, den, P20, P11ex, P11in,
           P21ex, P21in, P22 };
    const float h = dt_;
    for (int i = 0; i < n; ++i) {
      float tm = col(tau_m)[i], cm = col(C_m)[i], te = col(tau_ex)[i], ti = col(tau_in)[i];
      col(P11ex)[i] = std::exp(-h / te);
      col(P11in)[i] = std::exp(-h / ti);
      float p22 = std::exp(-h / tm);
      col(P22)[i] = p22;
      col(P21ex)[i] = (float)p32(te, tm, cm, h);
      col(P21in)[i] = (float)p32(ti, tm, cm, h);
      col(P20)[i] = (float)(tm / cm * (1.0 - p22));
    }
    for (int j : {P20, P11ex, P11in, P21ex, P21in, P22}) {
      bool same = std::all_of(col(j), col(j) + n, [&](float v) { return v == col(j)[0]; });
      g.uni[j] = same;
      g.uni_val[j] = col(j)[0];
    }
  } else if (m.id == MODEL_SPIKE_GEN) {
    // src/spike_generator.cu:230-269
    std::vector<int> off(n + 1, 0), tidx;
    std::vector<float> hh;
    for (int i = 0; i < n; ++i) {
      auto& t = g.sg_times[i];
      auto& hg = g.sg_heights[i];
      if (!t.empty()) {
        if (hg.empty()) hg.assign(t.size(), 1.0f);
        else if (hg.size() != t.size())
          throw Error("spike time array and spike height array must have the same size in spike generator");
        for (size_t k = 0; k < t.size(); ++k) {
          int ti = (int)std::round((t[k] - (float)t_min_) / dt_);
          if (k > 0 && ti <= tidx.back())
            throw Error("Spike times must be ordered, and the difference between\nconsecutive spikes must be >= the time resolution");
          tidx.push_back(ti);
          hh.push_back(hg[k]);
        }
      }
      off[i + 1] = (int)tidx.size();
    }
    upload(g.sg_off, off, stream_);
    upload(g.sg_tidx, tidx, stream_);
    upload(g.sg_h, hh, stream_);
    std::vector<int> zero(n, 0);
    upload(g.sg_next, zero, stream_);
  }
}

int Engine::Calibrate() {
  check_uncalibrated("Calibration can be made only once");
  if (!stream_) HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  // NGPU_CALIB_LOG=1: the phases' wall times on stderr
  const bool clog = getenv("NGPU_CALIB_LOG") && atoi(getenv("NGPU_CALIB_LOG")) != 0;
  auto t_last = std::chrono::steady_clock::now();
  auto phase = [&](const char* what) {
    if (!clog) return;
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "calibrate: %-24s %8.3f s\n", what, std::chrono::duration<double>(t - t_last).count());
    t_last = t;
  };
  sort_connections();
  phase("sort");
  const size_t nc = c_src_.size();
  // channels and port checks
  n_ch_ = 0;
  for (auto* g : groups_) {
    g->ch_base = n_ch_;
    n_ch_ += (long long)g->n * g->n_port;
  }
  if (n_ch_ >= (1ll << 32)) throw Error("Too many input channels");
  max_d_int_ = 0;
  has_out_.assign(n_nodes_, 0);
  std::vector<unsigned int> ch(nc);
  std::vector<unsigned short> dd(nc);
  bool bad_port = false;
#pragma omp parallel for schedule(static) reduction(max : max_d_int_) reduction(|| : bad_port)
  for (long long k = 0; k < (long long)nc; ++k) {
    const Group* g = groups_[node_group_[c_tgt_[k]]];
    int port = c_port_[k];
    if (port >= g->n_port) bad_port = true;
    ch[k] = (unsigned int)(g->ch_base + (long long)port * g->n + (c_tgt_[k] - g->i0));
    dd[k] = (unsigned short)c_d_[k];
    max_d_int_ = std::max(max_d_int_, c_d_[k]);
  }
  if (bad_port) throw Error("Connection port index larger than the number of ports of the target");
  for (int s = 0; s < n_nodes_; ++s) has_out_[s] = row_ptr_[s + 1] > row_ptr_[s];
  n_slot_ = max_d_int_ + 2;
  phase("channels");
  // device synapse table
  upload(d_row_ptr_, row_ptr_, stream_);
  upload(d_syn_ch_, ch, stream_);
  upload(d_syn_w_, c_w_, stream_);
  upload(d_syn_d_, dd, stream_);
  sync_streams();
  phase("synapse table upload");
  if (const char* env = getenv("NGPU_TILED")) tiled_env_ = atoi(env);
  plastic_ = false;
  for (size_t k = 0; k < nc && !plastic_; ++k) plastic_ = c_syn_[k] != 0;
  if (plastic_) tiled_env_ = 0;   // the tiled copy holds the weights of one moment
  build_tiled(ch, dd);
  // (before the populations' host staging is released: it reads den_delay)
  if (plastic_) build_plastic();
  phase("tiled table");
  ch.clear(); ch.shrink_to_fit(); dd.clear(); dd.shrink_to_fit();
  d_ring_.alloc((size_t)n_slot_ * std::max<long long>(n_ch_, 1));
  HIPCHK(hipMemsetAsync(d_ring_.p, 0, d_ring_.n * sizeof(double), stream_));
  // spike list: every node may emit in a step; headroom for repeated spikes
  spk_cap_ = std::max(2 * n_nodes_, 1024);
  d_spk_count_.alloc(kSpikeLists);
  d_overflow_.alloc(1);
  d_spk_src_.alloc((size_t)kSpikeLists * spk_cap_);
  d_spk_h_.alloc((size_t)kSpikeLists * spk_cap_);
  HIPCHK(hipMemsetAsync(d_spk_count_.p, 0, kSpikeLists * sizeof(int), stream_));
  HIPCHK(hipMemsetAsync(d_overflow_.p, 0, sizeof(int), stream_));
  std::vector<long long> guard(n_nodes_, 0x70000000LL);   // src/spike_buffer.cu:37
  upload(d_last_idx_, guard, stream_);
  d_last_h_.alloc(std::max(n_nodes_, 1));
  HIPCHK(hipMemsetAsync(d_last_h_.p, 0, d_last_h_.n * sizeof(float), stream_));
  {
    // persistent RK5 grid: 4 waves per SIMD on every CU
    int dev = 0, cus = 256;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 4;
    if (const char* env = getenv("NGPU_RK5_BLOCKS_PER_CU")) per_cu = std::min(4, std::max(1, atoi(env)));
    rk5_max_blocks_ = cus * per_cu;
  }
  if (const char* env = getenv("NGPU_RK5_PERSISTENT")) rk5_persistent_ = atoi(env) != 0;
  if (const char* env = getenv("NGPU_RK5_HOT_MARGIN")) rk5_hot_margin_ = (float)atof(env);
  if (const char* env = getenv("NGPU_RK5_HOT_ONLY")) rk5_hot_only_ = atoi(env) != 0;
  if (const char* env = getenv("NGPU_POISSON_RUNS")) poisson_runs_ = atoi(env) != 0;
  if (const char* env = getenv("NGPU_FOLD_IN_RK5")) fold_in_rk5_ = atoi(env) != 0;
  if (const char* env = getenv("NGPU_FUSED_DELIVER")) fused_deliver_env_ = atoi(env) != 0;
  if (const char* env = getenv("NGPU_LAG_D0")) lag_d0_mode_ = atoi(env) == 1 ? 1 : 0;
  if (const char* env = getenv("NGPU_RK5_HOT_LANES")) rk5_hot_lanes_ = std::min(64, std::max(1, atoi(env)));
  if (const char* env = getenv("NGPU_RK5_LEARNED_CAP")) rk5_learned_cap_ = std::max(0, atoi(env));
  if (const char* env = getenv("NGPU_RK5_DEFER")) {
    rk5_defer_cap_ = std::max(0, atoi(env));
    rk5_defer_env_ = true;
  }
  // plastic synapses need every spike in its own step (a post-synaptic spike
  // time is its step) and the delivery launch (which queues them for their
  // arrival): no split steps, no fused delivery, no batch launch
  if (plastic_) {
    rk5_defer_cap_ = 0;
    fused_deliver_env_ = false;
  }
  // per-wave counter rows: as many as the widest launch that adds to them
  long long rows = kCtrMinRows;
  for (auto* g : groups_)
    if (g->md->rk5)
      rows = std::max(rows, rk5_waves(g->n, rk5_persistent_ ? 1 : 0, rk5_max_blocks_));
  d_ctr_.alloc((size_t)rows * kCtrCols);
  HIPCHK(hipMemsetAsync(d_ctr_.p, 0, d_ctr_.n * sizeof(unsigned long long), stream_));
  d_stats_.alloc(1);
  HIPCHK(hipMemsetAsync(d_stats_.p, 0, sizeof(unsigned long long), stream_));
  // poisson generator tables and direct edges, src/nestgpu.cu:1190-1226
  n_gen_ = 0;
  for (auto* g : groups_)
    if (g->md->id == MODEL_POISSON_GEN) { g->gen_base = n_gen_; n_gen_ += g->n; }
  struct Edge { long long ch; int gen; float w, delay; unsigned int stream; };
  std::vector<Edge> edges;
  for (auto* g : groups_) {
    if (g->md->id != MODEL_POISSON_GEN) continue;
    for (int s = g->i0; s < g->i0 + g->n; ++s)
      for (long long gi = grp_ptr_[s]; gi < grp_ptr_[s + 1]; ++gi)
        for (long long c = grp_start_[gi]; c < grp_start_[gi + 1]; ++c) {
          const Group* tg = groups_[node_group_[c_tgt_[c]]];
          Edge e;
          e.ch = tg->ch_base + (long long)c_port_[c] * tg->n + (c_tgt_[c] - tg->i0);
          e.gen = g->gen_base + (s - g->i0);
          e.w = c_w_[c];
          e.delay = dt_ * (c_d_[c] + 1);
          e.stream = (unsigned int)c;
          edges.push_back(e);
        }
  }
  std::stable_sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) { return a.ch < b.ch; });
  for (auto* g : groups_) {
    if (g->n_port == 0) continue;
    long long lo = g->ch_base, hi = g->ch_base + (long long)g->n * g->n_port;
    auto b = std::lower_bound(edges.begin(), edges.end(), lo, [](const Edge& e, long long v) { return e.ch < v; });
    auto e = std::lower_bound(edges.begin(), edges.end(), hi, [](const Edge& e, long long v) { return e.ch < v; });
    g->has_pe = false;
    g->pe_run = false;
    if (b == e) continue;
    g->has_pe = true;
    size_t ne = e - b;
    // one uniform run: consecutive channels with one edge each, one weight
    // and delay, generator constant or stepping by one, stream stepping by one
    // (an all-to-all or one-to-one connection from generators)
    if (poisson_runs_) {
      const Edge& f = *b;
      const int gstep = ne > 1 ? (b + 1)->gen - f.gen : 0;
      bool run = gstep == 0 || gstep == 1;
      for (size_t k = 0; run && k < ne; ++k) {
        const Edge& x = *(b + k);
        run = x.ch == f.ch + (long long)k && x.gen == f.gen + gstep * (int)k && x.w == f.w &&
              x.delay == f.delay && x.stream == f.stream + (unsigned int)k;
      }
      if (run) {
        g->has_pe = false;
        g->pe_run = true;
        g->pu_b = f.ch - lo;
        g->pu_e = f.ch - lo + (long long)ne;
        g->pu_gen0 = f.gen;
        g->pu_gstep = gstep;
        g->pu_w = f.w;
        g->pu_delay = f.delay;
        g->pu_stream0 = f.stream;
        continue;
      }
    }
    std::vector<int> off((size_t)g->n * g->n_port + 1, 0), gen(ne);
    std::vector<float> w(ne), dl(ne);
    std::vector<unsigned int> st(ne);
    for (size_t k = 0; k < ne; ++k) {
      const Edge& x = *(b + k);
      off[x.ch - lo + 1]++;
      gen[k] = x.gen; w[k] = x.w; dl[k] = x.delay; st[k] = x.stream;
    }
    for (size_t k = 1; k < off.size(); ++k) off[k] += off[k - 1];
    upload(g->pe_off, off, stream_);
    upload(g->pe_gen, gen, stream_);
    upload(g->pe_w, w, stream_);
    upload(g->pe_delay, dl, stream_);
    upload(g->pe_stream, st, stream_);
  }
  phase("poisson edges");
  d_gen_chunks_.alloc(std::max(n_gen_, 1));
  d_gen_exp_.alloc(std::max(n_gen_, 1));
  d_gen_origin_.alloc(std::max(n_gen_, 1));
  d_gen_start_.alloc(std::max(n_gen_, 1));
  d_gen_stop_.alloc(std::max(n_gen_, 1));
  // populations: calibrate on the host staging, then upload
  for (auto* g : groups_) {
    calibrate_group(*g);
    int gi = group_index(groups_, g);
    upload(g->var, stage_var_[gi], stream_);
    upload(g->param, stage_param_[gi], stream_);
    if (g->md->rk5) {
      std::vector<double> x(g->n, t_min_);
      std::vector<float> h(g->n, g->h0);
      upload(g->x, x, stream_);
      upload(g->h, h, stream_);
      if (rk5_predictor_capable(g->md->id, g->n_port)) {
        std::vector<float> e(g->n, 1.0f);   // last accepted error ratio, predictor input
        upload(g->err_last, e, stream_);
      }
      const char* hot_env = getenv("NGPU_RK5_HOT");
      if (rk5_persistent_ && !(hot_env && atoi(hot_env) == 0)) {
        // hot-first lists (flags by parity, lists by t_idx mod 3)
        g->hot_flag.alloc((size_t)2 * g->n);
        g->hot_list.alloc((size_t)3 * g->n);
        g->hot_count.alloc(3);
        HIPCHK(hipMemsetAsync(g->hot_flag.p, 0, g->hot_flag.n, stream_));
        HIPCHK(hipMemsetAsync(g->hot_count.p, 0, 3 * sizeof(int), stream_));
      }
      if (rk5_persistent_ && rk5_defer_cap_ > 0) {
        // split steps: the kernel adds a lagged neuron's stashed input itself
        const FoldLayout fl = fold_layout(*g->md);
        if (!rk5_fold_layout_matches(g->md->id, g->n_port, fl.in_var, fl.in_port_stride,
                                     fl.w_param, fl.w_port_stride))
          throw Error(std::string("RK5 input layout of ") + g->md->name + " differs from the fold's");
        g->lag.alloc(g->n);
        g->pend.alloc((size_t)std::max(g->n_port, 1) * g->n);
        HIPCHK(hipMemsetAsync(g->lag.p, 0, g->lag.n, stream_));
      }
    }
    for (auto& iv : g->int_var) {
      std::vector<int> z(g->n, 0);
      if (iv == "spike_count") upload(g->spike_count, z, stream_);
      else upload(g->n_rec, z, stream_);
    }
    if (g->max_rec > 0) g->rec_times.alloc((size_t)g->n * g->max_rec);
  }
  sync_streams();
  phase("populations");
  for (size_t k = 0; k < groups_.size(); ++k) {
    stage_var_[k].clear(); stage_var_[k].shrink_to_fit();
    stage_param_[k].clear(); stage_param_[k].shrink_to_fit();
  }
  // fused delivery: the only spike-emitting population is a persistent RK5
  // population with split steps, whose launch can deliver the previous
  // step's list (spike generators, iaf_psc_exp and further RK5 populations
  // keep the delivery launch)
  {
    int emitters = 0;
    Group* rk = nullptr;
    for (auto* g : groups_) {
      const int id = g->md->id;
      if (id == MODEL_POISSON_GEN || id == MODEL_SPIKE_DETECTOR) continue;
      ++emitters;
      if (g->md->rk5 && g->lag.p) rk = g;
    }
    fused_deliver_ = fused_deliver_env_ && emitters == 1 && rk != nullptr;
    prev_list_pending_ = false;
  }
  if (const char* env = getenv("NGPU_BATCH")) batch_env_ = atoi(env) != 0;
  if (plastic_) batch_env_ = false;
  plan_batch();
  phase("launch plan");
  calibrated_ = true;
  neural_time_ = t_min_;
  return 0;
}

void Engine::upload_poisson_tables() {
  if (n_gen_ == 0) return;
  // lambda = time_step*rate in float, src/poiss_gen.cu:74, src/nestgpu.cu:560
  const float time_step = (float)(dt_ / 1000.0);
  std::vector<int> chunks(n_gen_);
  std::vector<double> ex(n_gen_);
  std::vector<float> org(n_gen_), sta(n_gen_), sto(n_gen_);
  for (auto* g : groups_) {
    if (g->md->id != MODEL_POISSON_GEN) continue;
    std::vector<float> p((size_t)g->np * g->n);
    copy_on(stream_, p.data(), g->param.p, p.size() * sizeof(float), hipMemcpyDeviceToHost);
    for (int i = 0; i < g->n; ++i) {
      float lambda = time_step * p[(size_t)0 * g->n + i];
      int m = poisson_chunks_host(lambda);
      chunks[g->gen_base + i] = m;
      ex[g->gen_base + i] = m > 0 ? std::exp(-(double)lambda / m) : 0.0;
      org[g->gen_base + i] = p[(size_t)1 * g->n + i];
      sta[g->gen_base + i] = p[(size_t)2 * g->n + i];
      sto[g->gen_base + i] = p[(size_t)3 * g->n + i];
    }
  }
  HIPCHK(hipMemcpyAsync(d_gen_chunks_.p, chunks.data(), n_gen_ * sizeof(int), hipMemcpyHostToDevice, stream_));
  HIPCHK(hipMemcpyAsync(d_gen_exp_.p, ex.data(), n_gen_ * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIPCHK(hipMemcpyAsync(d_gen_origin_.p, org.data(), n_gen_ * sizeof(float), hipMemcpyHostToDevice, stream_));
  HIPCHK(hipMemcpyAsync(d_gen_start_.p, sta.data(), n_gen_ * sizeof(float), hipMemcpyHostToDevice, stream_));
  HIPCHK(hipMemcpyAsync(d_gen_stop_.p, sto.data(), n_gen_ * sizeof(float), hipMemcpyHostToDevice, stream_));
  sync_streams();
}

static ParamTable make_table(const Group& g) {
  ParamTable t;
  std::memset(&t, 0, sizeof(t));
  for (int j = 0; j < g.np && j < kMaxParam; ++j) {
    t.val[j] = g.uni_val[j];
    if (!g.uni[j]) t.per_neuron[j >> 6] |= 1ull << (j & 63);
  }
  return t;
}

FoldArgs Engine::make_fold_args(const Group& g) {
  FoldArgs fa;
  std::memset(&fa, 0, sizeof(fa));
  fa.n = g.n; fa.n_port = g.n_port; fa.ch_base = g.ch_base; fa.var = g.var.p;
  const FoldLayout fl = fold_layout(*g.md);
  fa.in_var = fl.in_var;
  fa.in_port_stride = fl.in_port_stride;
  fa.param = g.param.p; fa.tab = make_table(g);
  fa.w_param = fl.w_param;
  fa.w_port_stride = fl.w_port_stride;
  fa.w_unit_ports = fl.w_unit_ports;
  fa.ring = d_ring_.p; fa.n_ch = n_ch_;
  if (g.has_pe) {
    fa.pe_off = g.pe_off.p; fa.pe_gen = g.pe_gen.p; fa.pe_w = g.pe_w.p; fa.pe_delay = g.pe_delay.p;
    fa.pe_stream = g.pe_stream.p;
  } else if (g.pe_run) {
    fa.pu_b = g.pu_b; fa.pu_e = g.pu_e; fa.pu_gen0 = g.pu_gen0; fa.pu_gstep = g.pu_gstep;
    fa.pu_w = g.pu_w; fa.pu_delay = g.pu_delay; fa.pu_stream0 = g.pu_stream0;
  }
  if (g.has_pe || g.pe_run) {
    fa.gen_chunks = d_gen_chunks_.p; fa.gen_exp = d_gen_exp_.p;
    fa.gen_origin = d_gen_origin_.p; fa.gen_start = d_gen_start_.p; fa.gen_stop = d_gen_stop_.p;
    fa.seed = kernel_seed_;
  }
  return fa;
}

RK5Args Engine::rk5_args(Group& gr, const StepCtx& sc, const SpikeList& sl, const SpikeBook& sb) {
  Group* g = &gr;
  RK5Args a{};
  a.n = g->n; a.i0 = g->i0; a.var = g->var.p; a.param = g->param.p; a.x = g->x.p; a.h = g->h.p;
  a.tab = make_table(*g); a.h_min = g->h_min; a.dt = dt_; a.ctx = sc; a.sl = sl; a.sb = sb;
  a.ctr = d_ctr_.p;
  a.persistent = rk5_persistent_ ? 1 : 0;
  a.max_blocks = rk5_max_blocks_;
  if (g->err_last.p && rk5_persistent_) {
    Predictor& pr = Predictor::get();
    a.pred_mode = pr.mode();
    a.err_last = g->err_last.p;
    a.pred_params = pr.d_params();
    a.log_buf = pr.d_log();
    a.log_count = pr.d_log_count();
    a.log_cap = pr.log_cap();
    a.log_stride = pr.log_stride();
  }
  if (g->hot_flag.p) {
    a.hot_flag = g->hot_flag.p;
    a.hot_list = g->hot_list.p;
    a.hot_count = g->hot_count.p;
    a.hot_margin = rk5_hot_margin_;
    a.hot_only = rk5_hot_only_ ? 1 : 0;
    a.hot_lanes = rk5_hot_lanes_;
  }
  if (g->lag.p) {
    a.lag = g->lag.p;
    a.pend = g->pend.p;
    // no split in a population a multimeter reads (it would see
    // mid-step states), nor in the last step of a Simulate() call
    // (state is complete between calls)
    // learned proposals shorten the spike upswing's attempt chain, so
    // the split point that balances two launches comes earlier
    const int cap = rk5_defer_env_ ? rk5_defer_cap_ : (a.pred_mode == 2 ? rk5_learned_cap_ : rk5_defer_cap_);
    a.defer_cap = (defer_ok_ && !g->recorded) ? cap : 0;
    if (g->fold_pending) {
      a.prefold = 1;
      a.pf = make_fold_args(*g);
      a.pf.ctx = step_ctx(it_ - 1);
    }
  }
  if (fused()) {
    a.fd.lag_d0_at_exit = lag_d0_mode_;
    a.fd.skip_lagged_d0 = 1;
    a.fd.st = syn_table();
    a.fd.spikes = d_stats_.p;
    a.clear_next_list = 1;
    if (prev_list_pending_) {
      a.fd.enable = 1;
      a.fd.prev_t_idx = sc.t_idx - 1;
      a.fd.prev_lst = (int)((sc.t_idx - 1) % kSpikeLists);
    }
  }
  return a;
}

void Engine::launch_step() {
  const StepCtx sc = step_ctx(it_);
  SpikeList sl{d_spk_count_.p, d_spk_src_.p, d_spk_h_.p, spk_cap_, d_overflow_.p};
  if (batch_active_) {
    batch_step(sc, sl);
    return;
  }
  for (size_t gi = 0; gi < groups_.size(); ++gi) {
    Group* g = groups_[gi];
    SpikeBook sb{d_last_idx_.p, d_last_h_.p, g->spike_count.p, g->rec_times.p, g->n_rec.p,
                 g->max_rec};
    switch (g->md->id) {
      case MODEL_AEIF_COND_BETA_MS:
      case MODEL_AEIF_COND_ALPHA_MS:
      case MODEL_AEIF_PSC_EXP:
      case MODEL_IZH_COND_BETA:
      case MODEL_AEIF_COND_ALPHA:
      case MODEL_AEIF_COND_BETA:
      case MODEL_AEIF_PSC_ALPHA:
      case MODEL_AEIF_PSC_ALPHA_MS:
      case MODEL_AEIF_PSC_DELTA:
      case MODEL_AEIF_PSC_EXP_MS: {
        RK5Args a = rk5_args(*g, sc, sl, sb);
        timed(KT_RK5, [&] { launch_rk5(g->md->id, g->n_port, a, stream_); });
        break;
      }
      case MODEL_IAF_PSC_EXP: {
        IafArgs a{};
        a.n = g->n; a.i0 = g->i0; a.var = g->var.p; a.param = g->param.p; a.tab = make_table(*g);
        a.dt = dt_; a.ctx = sc; a.sl = sl; a.sb = sb;
        timed(KT_IAF, [&] { launch_iaf(a, stream_); });
        break;
      }
      case MODEL_SPIKE_GEN: {
        SpikeGenArgs a{};
        a.n = g->n; a.i0 = g->i0; a.off = g->sg_off.p; a.t_idx = g->sg_tidx.p; a.height = g->sg_h.p;
        a.next = g->sg_next.p; a.ctx = sc; a.sl = sl; a.sb = sb;
        timed(KT_OTHER, [&] { launch_spike_gen(a, stream_); });
        break;
      }
      case MODEL_SPIKE_DETECTOR: {
        DetectorArgs a{g->n, g->var.p, g->param.p};
        timed(KT_OTHER, [&] { launch_detector(a, stream_); });
        break;
      }
      case MODEL_PARROT: {
        ParrotArgs a{};
        a.n = g->n; a.i0 = g->i0; a.var = g->var.p; a.param = g->param.p; a.ctx = sc; a.sl = sl;
        a.sb = sb;
        timed(KT_OTHER, [&] { launch_parrot(a, stream_); });
        break;
      }
      default:
        break;
    }
  }
  if (rec_ncol_ > 0) {
    RecordArgs ra{d_rec_cols_.p, rec_ncol_, d_rec_buf_.p, sc, d_last_idx_.p, d_last_h_.p,
                  rec_row_base_, rec_buf_rows_};
    timed(KT_OTHER, [&] { launch_record(ra, stream_); });
  }
  // delivery of this step's list: a launch of its own, or (fused) by the
  // next step's RK5 launch, except after the last step of a Simulate call
  // (or a step driven by SimulationStep), which leaves nothing pending
  if (!fused() || !defer_ok_) {
    deliver_list(sc, sl, fused() && lag_d0_mode_ != 0 ? 1 : 0);
    if (plastic_) plastic_step(sc, sl);
    prev_list_pending_ = false;
  } else {
    if (lag_d0_mode_ == 0) {
      // the d_int = 0 synapses of this step's lagged spikes, before the next
      // launch folds their slot
      DeliverArgs da{};
      da.ctx = sc; da.sl = sl; da.st = syn_table();
      da.ctr = d_ctr_.p;
      if (d_syn_w_.n > 0) timed(KT_OTHER, [&] { launch_deliver_lagged_d0(da, stream_); });
    }
    prev_list_pending_ = true;
  }
  for (auto* g : groups_) {
    if (g->n_port == 0) continue;
    // a split-step RK5 population's fold is done by its next RK5 launch, as
    // each lane takes its neuron (that launch comes after this delivery, so
    // spikes of split steps arriving in this slot are in), except after the
    // last step of a Simulate() call: state is complete between calls
    g->fold_pending = fold_in_rk5_ && g->lag.p && rk5_persistent_ && defer_ok_;
    if (g->fold_pending) continue;
    FoldArgs fa = make_fold_args(*g);
    fa.ctx = sc;
    fa.lag = g->lag.p;
    fa.pend = g->pend.p;
    timed(KT_FOLD, [&] { launch_fold(fa, stream_); });
  }
  if (timing_ && ev_pending_.size() >= 4096) flush_timing();
}

// ---- batch kernel: every population in one launch per step ------------------
// The network qualifies when every population other than the Poisson
// generators is of a batch kind (kernels.h BatchKind): an RK5 model of the
// catalog with uniform parameters, hot-first lists and split steps, or
// iaf_psc_exp; between 2 and kMaxBatchPops of them.  Spikes are then
// delivered by the next step's launch (fused delivery), and every
// population's fold by its next launch, except after the last step of a
// Simulate call, where the delivery and fold launches leave the state
// complete, as without the batch.
void Engine::plan_batch() {
  batch_on_ = false;
  batch_active_ = false;
  batch_groups_.clear();
  batch_fold_pending_ = false;
  batch_graph_destroy(batch_graph_);
  batch_graph_ = nullptr;
  if (!batch_env_ || !rk5_persistent_ || rk5_defer_cap_ <= 0 || !fold_in_rk5_ || !fused_deliver_env_)
    return;
  std::vector<int> gs;
  for (size_t gi = 0; gi < groups_.size(); ++gi) {
    Group* g = groups_[gi];
    const int id = g->md->id;
    if (id == MODEL_POISSON_GEN) continue;
    const int kind = batch_kind_of(id, g->n_port, 1);
    if (kind < 0 || g->n <= 0) return;
    if (g->md->rk5) {
      if (!g->hot_flag.p || !g->lag.p || !batch_uniform_ok(kind, make_table(*g))) return;
    } else {
      const FoldLayout fl = fold_layout(*g->md);
      if (g->n_port != 2 || fl.in_var != 0 || fl.in_port_stride != 1) return;
    }
    gs.push_back((int)gi);
  }
  if (gs.size() < 2 || gs.size() > (size_t)kMaxBatchPops) return;
  // counter rows for every wave of the widest batch launch (allocated once:
  // the uploaded per-population arguments hold the address)
  size_t rows = 0;
  for (int v = 0; v < kBatchVariants; ++v)
    rows = std::max(rows, (size_t)std::max(batch_resident_blocks(v), (int)gs.size()) *
                              (kRK5Block / kWave));
  if (d_ctr_.n < rows * kCtrCols) {
    d_ctr_.alloc(rows * kCtrCols);
    HIPCHK(hipMemsetAsync(d_ctr_.p, 0, d_ctr_.n * sizeof(unsigned long long), stream_));
  }
  batch_groups_ = gs;
  batch_on_ = true;
  batch_partition();
}

// Blocks of the launch per population, proportional to n x cost with cost =
// batch_rk5_weight_ x state variables for an RK5 population and 1 for
// iaf_psc_exp (at least one block each), over every block the launch keeps
// resident.
void Engine::batch_partition() {
  const int np = (int)batch_groups_.size();
  int grid = batch_resident_blocks(batch_variant_);
  grid = std::max(grid, np);
  std::vector<double> w(np);
  double tot = 0.0;
  for (int k = 0; k < np; ++k) {
    const Group* g = groups_[batch_groups_[k]];
    w[k] = (double)g->n * (g->md->rk5 ? batch_rk5_weight_ * (double)g->nv : 1.0);
    tot += w[k];
  }
  // blocks beyond one lane per neuron are useless: cap each population
  std::vector<int> nb(np), cap(np);
  int used = 0;
  for (int k = 0; k < np; ++k) {
    const Group* g = groups_[batch_groups_[k]];
    cap[k] = std::max(1, (g->n + kRK5Block - 1) / kRK5Block);
    nb[k] = std::min(cap[k], std::max(1, (int)std::floor((double)grid * w[k] / tot)));
    used += nb[k];
  }
  // hand the remaining blocks out by largest deficit, within the caps
  while (used < grid) {
    int best = -1;
    double bd = -1e300;
    for (int k = 0; k < np; ++k) {
      if (nb[k] >= cap[k]) continue;
      const double d = (double)grid * w[k] / tot - nb[k];
      if (d > bd) { bd = d; best = k; }
    }
    if (best < 0) break;
    ++nb[best];
    ++used;
  }
  while (used > grid) {   // more populations than blocks would fit: never below one each
    int best = -1;
    for (int k = 0; k < np; ++k)
      if (nb[k] > 1 && (best < 0 || nb[k] > nb[best])) best = k;
    if (best < 0) break;
    --nb[best];
    --used;
  }
  BatchArgs b{};
  b.n_pop = np;
  b.blk0[0] = 0;
  for (int k = 0; k < np; ++k) {
    const Group* g = groups_[batch_groups_[k]];
    b.kind[k] = batch_kind_of(g->md->id, g->n_port, batch_iaf_unroll_);
    b.blk0[k + 1] = b.blk0[k] + nb[k];
  }
  batch_grid_ = b.blk0[np];
  batch_plan_ = b;
  if ((size_t)batch_grid_ * (kRK5Block / kWave) * kCtrCols > d_ctr_.n)
    throw Error("Batch kernel: more waves than counter rows");
  batch_graph_destroy(batch_graph_);
  batch_graph_ = nullptr;
}

void Engine::batch_upload_args() {
  const int np = (int)batch_groups_.size();
  std::vector<RK5Args> rk(np);
  std::vector<IafArgs> ia(np);
  std::vector<FoldArgs> fo(np);
  SpikeList sl{d_spk_count_.p, d_spk_src_.p, d_spk_h_.p, spk_cap_, d_overflow_.p};
  const StepCtx sc = step_ctx(0);
  const bool saved_defer = defer_ok_, saved_fused = fused_deliver_;
  defer_ok_ = true;        // the cap each population takes when splits are on
  fused_deliver_ = true;   // lagged spikes: d_int = 0 synapses at wave exit
  for (int k = 0; k < np; ++k) {
    Group* g = groups_[batch_groups_[k]];
    SpikeBook sb{d_last_idx_.p, d_last_h_.p, g->spike_count.p, g->rec_times.p, g->n_rec.p,
                 g->max_rec};
    std::memset(&rk[k], 0, sizeof(RK5Args));
    std::memset(&ia[k], 0, sizeof(IafArgs));
    std::memset(&fo[k], 0, sizeof(FoldArgs));
    if (g->md->rk5) {
      const bool fp = g->fold_pending;
      g->fold_pending = false;
      rk[k] = rk5_args(*g, sc, sl, sb);
      g->fold_pending = fp;
      rk[k].pred_mode = 0;
      rk[k].pf = make_fold_args(*g);   // the previous step's fold (step values per launch)
      rk[k].fd.enable = 0;
      rk[k].clear_next_list = 0;
    } else {
      IafArgs& a = ia[k];
      a.n = g->n; a.i0 = g->i0; a.var = g->var.p; a.param = g->param.p; a.tab = make_table(*g);
      a.dt = dt_; a.ctx = sc; a.sl = sl; a.sb = sb;
      fo[k] = make_fold_args(*g);
    }
  }
  defer_ok_ = saved_defer;
  fused_deliver_ = saved_fused;
  batch_upload(rk.data(), ia.data(), fo.data(), np, stream_);
}

void Engine::batch_step(const StepCtx& sc, const SpikeList& sl) {
  BatchArgs b = batch_plan_;
  b.ctx = sc;
  b.pctx = step_ctx(it_ - 1);
  b.prefold = batch_fold_pending_ ? 1 : 0;
  b.defer_on = defer_ok_ ? 1 : 0;
  b.clear_next_list = 1;
  b.sl = sl;
  const bool fu = !tiled_use_;   // the next launch delivers this step's list
  b.clear_next_list = fu ? 1 : 0;
  b.fd.lag_d0_at_exit = fu ? lag_d0_mode_ : 0;
  b.fd.skip_lagged_d0 = 1;
  b.fd.st = syn_table();
  b.fd.spikes = d_stats_.p;
  if (fu && prev_list_pending_) {
    b.fd.enable = 1;
    b.fd.prev_t_idx = sc.t_idx - 1;
    b.fd.prev_lst = (int)((sc.t_idx - 1) % kSpikeLists);
  }
  b.ctr = d_ctr_.p;
  timed(KT_RK5, [&] {
    if (batch_graph_on_) {
      if (!batch_graph_) batch_graph_ = batch_graph_create(batch_variant_, batch_grid_, b);
      batch_graph_launch(batch_graph_, b, stream_);
    } else {
      launch_batch(b, batch_grid_, batch_variant_, stream_);
    }
  });
  if (rec_ncol_ > 0) {
    RecordArgs ra{d_rec_cols_.p, rec_ncol_, d_rec_buf_.p, sc, d_last_idx_.p, d_last_h_.p,
                  rec_row_base_, rec_buf_rows_};
    timed(KT_OTHER, [&] { launch_record(ra, stream_); });
  }
  if (defer_ok_) {
    // this step's folds are done by the next launch; with fused delivery its
    // spikes too, except the d_int = 0 synapses of lagged spikes, whose slot
    // that launch folds
    if (!fu) {
      deliver_list(sc, sl, 0);
    } else if (lag_d0_mode_ == 0) {
      DeliverArgs da{};
      da.ctx = sc; da.sl = sl; da.st = syn_table();
      da.ctr = d_ctr_.p;
      if (d_syn_w_.n > 0) timed(KT_OTHER, [&] { launch_deliver_lagged_d0(da, stream_); });
    }
    prev_list_pending_ = fu;
    batch_fold_pending_ = true;
    for (int gi : batch_groups_) groups_[gi]->fold_pending = true;
  } else {
    // last step of a Simulate call: deliver this step's list and fold every
    // population, so the state is complete between calls
    deliver_list(sc, sl, fu && lag_d0_mode_ != 0 ? 1 : 0);
    prev_list_pending_ = false;
    batch_fold_pending_ = false;
    for (int gi : batch_groups_) {
      Group* g = groups_[gi];
      g->fold_pending = false;
      FoldArgs fa = make_fold_args(*g);
      fa.ctx = sc;
      fa.lag = g->lag.p;
      fa.pend = g->pend.p;
      timed(KT_FOLD, [&] { launch_fold(fa, stream_); });
    }
  }
  if (timing_ && ev_pending_.size() >= 4096) flush_timing();
}

void Engine::GetBatchConfig(int* cfg) {
  cfg[0] = batch_on_ && Predictor::get().mode() == 0 ? 1 : 0;
  cfg[1] = batch_variant_;
  cfg[2] = batch_iaf_unroll_;
  cfg[3] = (int)std::lround(batch_rk5_weight_ * 100.0f);
  cfg[4] = batch_graph_on_ ? 1 : 0;
  cfg[5] = batch_on_ ? batch_grid_ : 0;
  cfg[6] = (int)batch_groups_.size();
  cfg[7] = tiled_on_ ? 1 : 0;
  cfg[8] = tiled_use_ ? 1 : 0;
  cfg[9] = tiled_on_ ? tile_ch_ : 0;
  cfg[10] = tiled_on_ ? n_tiles_ : 0;
  cfg[11] = tiled_on_ ? tile_rows_ : 0;
}

void Engine::SetBatchConfig(int variant, int iaf_unroll, float rk5_weight, int graph, int tiled) {
  if (variant < 0 || variant >= kBatchVariants) throw Error("Batch kernel variant must be 0 or 1");
  if (iaf_unroll != 1 && iaf_unroll != 2 && iaf_unroll != 4)
    throw Error("Batch iaf_psc_exp unroll must be 1, 2 or 4");
  if (!(rk5_weight > 0.0f && rk5_weight < 1.0e6f)) throw Error("Batch RK5 weight must be positive");
  batch_variant_ = variant;
  batch_iaf_unroll_ = iaf_unroll;
  batch_rk5_weight_ = rk5_weight;
  batch_graph_on_ = graph != 0;
  if (tiled >= 0) {
    if (tiled && !tiled_on_) throw Error("Tiled delivery is not available for this network");
    tiled_use_ = tiled != 0;
  }
  if (batch_on_) batch_partition();
}

// Autotuner of the batch kernel: register budget (3 or 4 waves per SIMD),
// iaf_psc_exp unroll (1, 2, 4 neurons per lane per pass), the RK5 block
// weight (2, 4, 8 per state variable) and, where the network has the tiled
// synapse table, fused delivery or a dense delivery launch, each candidate timed with HIP events
// around `steps` consecutive steps of the running simulation (a Simulate
// call of that length); the fastest is kept.  Every candidate computes the
// same results, so tuning on live steps changes nothing but time.
double Engine::AutotuneBatch(int steps, int* n_candidates) {
  if (!calibrated_) Calibrate();
  if (!batch_on_) throw Error("Autotune: the network does not run as one batch launch");
  if (Predictor::get().mode() != 0)
    throw Error("Autotune: the batch launch runs with the step-size predictor off only");
  if (steps < 2) throw Error("Autotune: at least 2 steps per candidate");
  struct Cand { int v, u; float w; int t; };
  std::vector<Cand> cands;
  for (int t = 0; t <= (tiled_on_ ? 1 : 0); ++t)
    for (int v = 0; v < kBatchVariants; ++v)
      for (int u : {1, 2, 4})
        for (float w : {2.0f, 4.0f, 8.0f}) cands.push_back({v, u, w, t});
  const float saved_sim = sim_time_;
  const int saved_verb = verbosity_;
  verbosity_ = 0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best_ms = 1e300;
  Cand best = cands[0];
  for (const Cand& c : cands) {
    SetBatchConfig(c.v, c.u, c.w, batch_graph_on_ ? 1 : 0, tiled_on_ ? c.t : -1);
    sim_time_ = dt_ * steps;
    HIPCHK(hipEventRecord(e0, stream_));
    Simulate();
    HIPCHK(hipEventRecord(e1, stream_));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  sim_time_ = saved_sim;
  verbosity_ = saved_verb;
  SetBatchConfig(best.v, best.u, best.w, batch_graph_on_ ? 1 : 0, tiled_on_ ? best.t : -1);
  if (n_candidates) *n_candidates = (int)cands.size();
  return best_ms / steps;
}

// Tiled copy of the synapse table for dense delivery (TiledSyn).  Built when
// a step's events may outnumber the ring's entries (at least 4 synapses per
// ring entry; NGPU_TILED=1 builds it whenever it fits, 0 never), delay slots
// stay below 256 and the (source, tile) table stays below 2^28 entries.
// Within a source the records keep their order inside a tile.  The LDS tile
// keeps the relative slots r = d_int + 1 - lag below rows, rows the smallest
// count that holds the delays of all but 1% of the synapses (records beyond
// go to the ring by global atomics).  tile_ch (a multiple of 8, at most
// kTileChMax, the rows x tile_ch FP64 tile within kTileLdsBytes and the
// device's LDS per block) minimises rounds x (tile_ch + 32): a round is one
// tile block on every block slot of the chip (LDS per CU / tile bytes, at
// most 2 blocks of 16 waves per CU), and 32 channels stand for a tile's
// fixed cost (reading and scanning the step's (first, end) runs).
// NGPU_TILE_CH=c sets tile_ch, NGPU_TILE_ROWS=r the rows.
int Engine::choose_tile_ch(int rows) const {
  const int row = rows * (int)sizeof(double);
  int dev = 0, cus = 256, lds_cu = 160 * 1024, lds_block = kTileLdsBytes;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  HIPCHK(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const int budget = std::min(kTileLdsBytes, lds_block - kTileStaticLds);
  const int tc_max = std::min(kTileChMax, budget / row / 8 * 8);
  if (tc_max < 8) return 0;
  if (const char* env = getenv("NGPU_TILE_CH")) {
    const int c = atoi(env) / 8 * 8;
    if (c >= 8 && c <= tc_max) return c;
  }
  lds_cu = std::max(lds_cu, budget);
  int best = 0;
  long long best_cost = 0;
  for (int tc = 8; tc <= tc_max; tc += 8) {
    const long long nt = (n_ch_ + tc - 1) / tc;
    const long long per_cu = std::min(2, lds_cu / (tc * row + kTileStaticLds));
    const long long rounds = (nt + cus * per_cu - 1) / (cus * per_cu);
    const long long cost = rounds * (tc + 32);
    if (!best || cost < best_cost) { best = tc; best_cost = cost; }
  }
  return best;
}

void Engine::build_tiled(const std::vector<unsigned int>& ch, const std::vector<unsigned short>& dd) {
  tiled_on_ = tiled_use_ = false;
  d_toff_.free(); d_trec_.free(); d_tev_.free(); d_tplan_.free(); d_tsched_.free();
  const size_t nc = ch.size();
  if (tiled_env_ == 0 || nc == 0 || nc >= (1ull << 32) || n_ch_ <= 0 || max_d_int_ > 255) return;
  const bool dense = (double)nc >= 4.0 * (double)n_ch_ * (double)n_slot_;
  if (tiled_env_ < 0 && !dense) return;
  // even the widest tile gives too large an offset table: no tiled copy
  if ((long long)n_nodes_ * ((n_ch_ + kTileChMax - 1) / kTileChMax + 1) >= (1ll << 28)) return;
  std::vector<long long> hist(max_d_int_ + 1, 0);
  for (size_t k = 0; k < nc; ++k) hist[dd[k]]++;
  int rows = n_slot_;            // r = d + 1 <= rows - 1 for every synapse
  long long beyond = 0;          // synapses whose r = d + 1 is rows or more
  while (rows > 2 && (beyond + hist[rows - 2]) * 100 <= (long long)nc) beyond += hist[--rows - 1];
  if (const char* env = getenv("NGPU_TILE_SPLIT_MIN")) tile_split_min_ = std::max(1, atoi(env));
  if (const char* env = getenv("NGPU_TILE_ROWS")) {
    const int r = atoi(env);
    if (r >= 1 && r <= n_slot_) rows = r;
  }
  const int tc = choose_tile_ch(rows);
  if (tc == 0) return;
  const long long nt = (n_ch_ + tc - 1) / tc;
  const long long entries = (long long)n_nodes_ * nt;
  if (entries >= (1ll << 28)) return;
  // runs packed source by source, tile by tile (aligning each run to a
  // whole line was measured slower)
  std::vector<unsigned int> cnt((size_t)entries, 0u);
  std::vector<long long> base((size_t)n_nodes_ + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int s = 0; s < n_nodes_; ++s) {
    unsigned int* c = cnt.data() + (size_t)s * nt;
    for (long long k = row_ptr_[s]; k < row_ptr_[s + 1]; ++k) c[ch[k] / tc]++;
    long long sz = 0;
    for (long long t = 0; t < nt; ++t) sz += c[t];
    base[s + 1] = sz;
  }
  for (int s = 0; s < n_nodes_; ++s) base[s + 1] += base[s];
  if (base[n_nodes_] >= (1ll << 32)) return;
  std::vector<uint2> runs((size_t)entries);
  std::vector<TileRec> rec((size_t)base[n_nodes_]);
#pragma omp parallel
  {
    std::vector<unsigned int> pos(nt);
#pragma omp for schedule(dynamic, 64)
    for (int s = 0; s < n_nodes_; ++s) {
      const unsigned int* c = cnt.data() + (size_t)s * nt;
      uint2* o = runs.data() + (size_t)s * nt;
      long long at = base[s];
      for (long long t = 0; t < nt; ++t) {
        o[t] = make_uint2((unsigned int)at, (unsigned int)(at + c[t]));
        pos[t] = (unsigned int)at;   // next free position of tile t
        at += c[t];
      }
      for (long long k = row_ptr_[s]; k < row_ptr_[s + 1]; ++k) {
        const unsigned int t = ch[k] / tc;
        TileRec& r = rec[pos[t]++];
        r.ch = (unsigned short)(ch[k] - t * (unsigned int)tc);
        r.d = (unsigned char)dd[k];
        r.pad = 0;
        r.w = c_w_[k];
      }
    }
  }
  cnt.clear(); cnt.shrink_to_fit();
  upload(d_toff_, runs, stream_);
  upload(d_trec_, rec, stream_);
  // first step's plan: one item per tile; then the kernel plans each next step
  std::vector<uint2> plan((size_t)nt * kTileSplitMax, make_uint2(0u, 0u));
  for (long long t = 0; t < nt; ++t) plan[t] = make_uint2((unsigned int)t, 1u << 16);
  const std::vector<unsigned int> tev((size_t)nt, 0u);
  std::vector<int> sched{(int)nt, 0, 0, 0};   // items, next item (set below), finished blocks
  upload(d_tplan_, plan, stream_);
  upload(d_tev_, tev, stream_);
  {
    // persistent grid: every block slot of the chip (tile LDS per CU, at
    // most 2 blocks of 16 waves)
    int dev = 0, cus = 256, lds_cu = 160 * 1024;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    const int tile_bytes = rows * tc * (int)sizeof(double) + kTileStaticLds;
    const int per_cu = std::max(1, std::min(2, lds_cu / tile_bytes));
    tiled_blocks_ = cus * per_cu;
    sched[1] = tiled_blocks_;   // items below the grid size are taken by block index
  }
  upload(d_tsched_, sched, stream_);
  sync_streams();
  tile_ch_ = tc;
  tile_rows_ = rows;
  n_tiles_ = (int)nt;
  tiled_on_ = true;
  tiled_use_ = true;
}

// Delivery launch of this step's list: dense (tiled) or global atomics.
void Engine::deliver_list(const StepCtx& sc, const SpikeList& sl, int skip_lagged_d0) {
  if (tiled_use_) {
    TiledDeliverArgs ta{};
    ta.ctx = sc; ta.sl = sl;
    ta.ts = TiledSyn{d_toff_.p, d_trec_.p, n_tiles_, tile_ch_, tile_rows_};
    ta.ring = d_ring_.p; ta.n_ch = n_ch_; ta.n_slot = n_slot_;
    ta.skip_lagged_d0 = skip_lagged_d0;
    ta.ctr = d_ctr_.p; ta.spikes = d_stats_.p;
    ta.tile_ev = d_tev_.p; ta.plan = d_tplan_.p; ta.sched = d_tsched_.p;
    ta.blocks = tiled_blocks_;
    ta.split_min = tile_split_min_;
    timed(KT_DELIVER, [&] { launch_deliver_tiled(ta, stream_); });
    return;
  }
  DeliverArgs da{};
  da.ctx = sc; da.sl = sl; da.st = syn_table();
  da.skip_lagged_d0 = skip_lagged_d0;
  da.ctr = d_ctr_.p; da.spikes = d_stats_.p;
  timed(KT_DELIVER, [&] { launch_deliver(da, stream_); });
}

// Device state of plastic synapses (synapse group >= 1): the group of every
// synapse, the node of every input channel, the arrival step of each
// synapse's last spike and each node's last spike step (both "long ago" at
// first, as the reference's 0x8000 marks, src/rev_spike.cu:171-173), the
// incoming plastic synapses of every node (src/spike_buffer.cu:399-460
// builds the reference's TargetRevConnection), the groups' rules, and one
// arrival queue per ring slot.  A dendritic delay other than 0 on a node
// with incoming plastic synapses is not supported (error).
void Engine::build_plastic() {
  const size_t nc = c_src_.size();
  std::vector<unsigned char> grp(c_syn_.begin(), c_syn_.end());
  std::vector<long long> rptr((size_t)n_nodes_ + 1, 0);
  size_t n_pl = 0;
  for (size_t k = 0; k < nc; ++k)
    if (grp[k]) { rptr[c_tgt_[k] + 1]++; ++n_pl; }
  for (int i = 0; i < n_nodes_; ++i) rptr[i + 1] += rptr[i];
  std::vector<unsigned int> ridx(std::max<size_t>(n_pl, 1));
  {
    std::vector<long long> pos(rptr.begin(), rptr.end() - 1);
    for (size_t k = 0; k < nc; ++k)
      if (grp[k]) ridx[pos[c_tgt_[k]]++] = (unsigned int)k;
  }
  for (auto* g : groups_) {
    const int j = param_idx(*g, "den_delay", 0);
    if (j < 0) continue;
    const int gi = group_index(groups_, g);
    for (int i = 0; i < g->n; ++i)
      if (rptr[g->i0 + i + 1] > rptr[g->i0 + i] && stage_param_[gi][(size_t)j * g->n + i] != 0.0f)
        throw Error("A dendritic delay on a target of plastic synapses is not supported");
  }
  std::vector<int> ch_node((size_t)std::max<long long>(n_ch_, 1), 0);
  for (auto* g : groups_)
    for (int q = 0; q < g->n_port; ++q)
      for (int i = 0; i < g->n; ++i) ch_node[(size_t)(g->ch_base + (long long)q * g->n + i)] = g->i0 + i;
  std::vector<float> tab(syn_groups_.size() * kSynGroupParams, 0.0f);
  for (size_t i = 1; i < syn_groups_.size(); ++i) {
    tab[i * kSynGroupParams] = (float)syn_groups_[i].model;
    for (size_t j = 0; j < syn_groups_[i].p.size(); ++j) tab[i * kSynGroupParams + 1 + j] = syn_groups_[i].p[j];
  }
  // queue per slot: every plastic synapse once (a source fires at most once
  // per step, so a slot receives each synapse at most once)
  pq_cap_ = (int)std::min<size_t>(std::max<size_t>(n_pl, 1024), (size_t)1 << 30);
  if ((size_t)pq_cap_ * n_slot_ > ((size_t)1 << 28)) throw Error("Too many plastic synapses for the arrival queues");
  const std::vector<int> far_pre(nc, -(1 << 30)), far_post(n_nodes_, -(1 << 30));
  const std::vector<int> zeros((size_t)n_slot_ + 1, 0), zero1(1, 0);
  upload(d_syn_grp_, grp, stream_);
  upload(d_rev_ptr_, rptr, stream_);
  upload(d_rev_idx_, ridx, stream_);
  upload(d_ch_node_, ch_node, stream_);
  upload(d_sg_param_, tab, stream_);
  upload(d_last_pre_, far_pre, stream_);
  upload(d_last_post_, far_post, stream_);
  upload(d_pq_n_, zeros, stream_);
  upload(d_pq_over_, zero1, stream_);
  d_pq_.alloc((size_t)pq_cap_ * n_slot_);
  sync_streams();
}

void Engine::plastic_step(const StepCtx& sc, const SpikeList& sl) {
  StdpArgs a{};
  a.st = syn_table();
  a.w = d_syn_w_.p;
  a.ch_node = d_ch_node_.p;
  a.last_pre = d_last_pre_.p;
  a.last_post = d_last_post_.p;
  a.rev_ptr = d_rev_ptr_.p;
  a.rev_idx = d_rev_idx_.p;
  a.grp_param = d_sg_param_.p;
  a.ctx = sc;
  a.sl = sl;
  a.dt = (float)dt_;
  timed(KT_DELIVER, [&] { launch_stdp_step(a, stream_); });
}

SynTable Engine::syn_table() const {
  SynTable t;
  t.row_ptr = d_row_ptr_.p;
  t.syn_ch = d_syn_ch_.p;
  t.syn_w = d_syn_w_.p;
  t.syn_d = d_syn_d_.p;
  t.ring = d_ring_.p;
  t.n_ch = n_ch_;
  t.n_slot = n_slot_;
  t.grp = plastic_ ? d_syn_grp_.p : nullptr;
  t.q = d_pq_.p;
  t.q_n = d_pq_n_.p;
  t.q_cap = pq_cap_;
  t.q_over = d_pq_over_.p;
  return t;
}

void Engine::timed(int cat, const std::function<void()>& launch, hipStream_t st) {
  if (!st) st = stream_;
  if (!timing_ || (timing_rk5_only_ && cat != KT_RK5) ||
      (timing_stride_ > 1 && it_ % timing_stride_ != kTimingPhase)) {
    launch();
    HIPCHK(hipGetLastError());   // a launch that cannot run fails the step loudly
    return;
  }
  auto get = [&]() {
    if (!ev_free_.empty()) {
      hipEvent_t e = ev_free_.back();
      ev_free_.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  };
  EvPair p{get(), get(), cat};
  HIPCHK(hipEventRecord(p.a, st));
  launch();
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(p.b, st));
  ev_pending_.push_back(p);
}

void Engine::flush_timing() {
  if (ev_pending_.empty()) return;
  sync_streams();
  for (auto& p : ev_pending_) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    kt_ms_[p.cat] += ms;
    kt_n_[p.cat] += 1;
    ev_free_.push_back(p.a);
    ev_free_.push_back(p.b);
  }
  ev_pending_.clear();
}

void Engine::SetKernelTiming(int mode) {
  if (timing_) flush_timing();
  timing_ = mode != 0;
  timing_rk5_only_ = mode == 2 || mode == 4;
  timing_stride_ = mode >= 3 ? kTimingStride : 1;
  for (int k = 0; k < KT_N; ++k) { kt_ms_[k] = 0.0; kt_n_[k] = 0; }
}

void Engine::KernelTimes(double* ms, long long* launches) {
  flush_timing();
  for (int k = 0; k < KT_N; ++k) { ms[k] = kt_ms_[k]; launches[k] = kt_n_[k]; }
}

int Engine::StartSimulation() {
  if (!calibrated_) Calibrate();
  if (first_sim_) {
    // record row at the initial time (src/nestgpu.cu:383-388)
    sync_streams();
    for (auto& r : records_) {
      std::vector<float> row{(float)neural_time_};
      for (size_t k = 0; k < r.var.size(); ++k) {
        Group* g = groups_[r.group[k]];
        float v = 0.f;
        if (r.var[k] != "spike") {
          int j = var_idx(*g, r.var[k], r.port[k]);
          copy_on(stream_, &v, g->var.p + (size_t)j * g->n + r.rel[k], sizeof(float), hipMemcpyDeviceToHost);
        }
        row.push_back(v);
      }
      if (r.fp) {
        fprintf(r.fp, "%f", row[0]);
        for (size_t k = 1; k < row.size(); ++k) fprintf(r.fp, "\t%f", row[k]);
        fprintf(r.fp, "\n");
      }
      r.data.push_back(std::move(row));
    }
    build_real_time_ = wall_time();
    first_sim_ = false;
  }
  neur_t0_ = neural_time_;
  it_ = 0;
  nt_ = (long long)std::round(sim_time_ / dt_);
  upload_poisson_tables();
  // record columns and buffer for this call
  std::vector<RecCol> cols;
  for (auto* g : groups_) g->recorded = false;
  for (auto& r : records_)
    for (size_t k = 0; k < r.var.size(); ++k) {
      Group* g = groups_[r.group[k]];
      g->recorded = true;
      RecCol rc;
      rc.node = r.node[k];
      rc.has_out = has_out_[r.node[k]];
      rc.ptr = nullptr;
      if (r.var[k] != "spike") {
        int j = var_idx(*g, r.var[k], r.port[k]);
        rc.ptr = g->var.p + (size_t)j * g->n + r.rel[k];
      }
      cols.push_back(rc);
    }
  rec_ncol_ = (int)cols.size();
  if (rec_ncol_ > 0) {
    upload(d_rec_cols_, cols, stream_);
    rec_buf_rows_ = std::max<long long>(std::min<long long>(nt_, 1 << 20), 1);
    d_rec_buf_.alloc((size_t)rec_buf_rows_ * (rec_ncol_ + 1));
  }
  rec_rows_pending_ = 0;
  rec_row_base_ = 0;
  // the batch kernel's per-population arguments (parameters may have been
  // set since the last call, and which populations are recorded decides
  // their split caps)
  // (learned step-size proposals run in the single-population RK5 launches)
  batch_active_ = batch_on_ && Predictor::get().mode() == 0;
  if (batch_active_) batch_upload_args();
  return 0;
}

int Engine::SimulationStep() {
  if (first_sim_ || !calibrated_) StartSimulation();
  if (rec_ncol_ > 0 && rec_rows_pending_ >= rec_buf_rows_) {
    // record buffer full: move its rows to the host and restart it
    sync_streams();
    fetch_records(rec_rows_pending_);
    rec_rows_pending_ = 0;
    rec_row_base_ = it_;
  }
  launch_step();
  ++rec_rows_pending_;
  long long t_idx = (long long)(int)std::round(neur_t0_ / dt_) + it_ + 1;
  ++it_;
  neural_time_ = neur_t0_ + (double)dt_ * it_;
  for (auto* g : groups_)
    if (g->max_rec > 0 && g->rec_step > 0 && (t_idx % g->rec_step == g->rec_step - 1))
      buffer_rec_spike_times(*g);
  return 0;
}

void Engine::fetch_records(long long n_rows) {
  if (rec_ncol_ == 0 || n_rows == 0) return;
  std::vector<float> buf((size_t)n_rows * (rec_ncol_ + 1));
  copy_on(stream_, buf.data(), d_rec_buf_.p, buf.size() * sizeof(float), hipMemcpyDeviceToHost);
  int col0 = 1;
  for (auto& r : records_) {
    int nc = (int)r.var.size();
    for (long long row = 0; row < n_rows; ++row) {
      const float* src = buf.data() + (size_t)row * (rec_ncol_ + 1);
      std::vector<float> v;
      v.reserve(nc + 1);
      v.push_back(src[0]);
      for (int k = 0; k < nc; ++k) v.push_back(src[col0 + k]);
      if (r.fp) {
        fprintf(r.fp, "%f", v[0]);
        for (int k = 1; k <= nc; ++k) fprintf(r.fp, "\t%f", v[k]);
        fprintf(r.fp, "\n");
      }
      r.data.push_back(std::move(v));
    }
    col0 += nc;
  }
}

int Engine::EndSimulation() {
  sync_streams();
  if (timing_) flush_timing();
  fetch_records(rec_rows_pending_);
  rec_rows_pending_ = 0;
  double end = wall_time();
  if (verbosity_ >= 1) {
    printf("Building time: %g\n", build_real_time_ - start_real_time_);
    printf("Simulation time: %g\n", end - build_real_time_);
  }
  return 0;
}

int Engine::Simulate() {
  StartSimulation();
  if (verbosity_ >= 1) printf("Neural activity simulation time: %.3lf ms\n", (double)sim_time_);
  for (long long it = 0; it < nt_; ++it) {
    defer_ok_ = it + 1 < nt_;
    SimulationStep();
  }
  defer_ok_ = false;
  EndSimulation();
  return 0;
}

void Engine::Stats(unsigned long long* out) {
  for (int k = 0; k < 8; ++k) out[k] = 0;
  if (!calibrated_) return;
  sync_streams();
  std::vector<unsigned long long> rows(d_ctr_.n);
  copy_on(stream_, rows.data(), d_ctr_.p, rows.size() * sizeof(unsigned long long),
                   hipMemcpyDeviceToHost);
  for (size_t r = 0; r < rows.size(); r += kCtrCols) {
    for (int k = 0; k < 4; ++k) out[k] += rows[r + k];
    out[6] += rows[r + 4];
    out[7] += rows[r + 5];
  }
  copy_on(stream_, &out[4], d_stats_.p, sizeof(unsigned long long), hipMemcpyDeviceToHost);
  int ov = 0;
  copy_on(stream_, &ov, d_overflow_.p, sizeof(int), hipMemcpyDeviceToHost);
  out[5] = (unsigned long long)ov;
}

void Engine::Synchronize() {
  if (stream_) sync_streams();
}

template struct DBuf<float>;
template struct DBuf<double>;
template struct DBuf<int>;
template struct DBuf<long long>;
template struct DBuf<unsigned int>;
template struct DBuf<unsigned short>;
template struct DBuf<unsigned char>;
template struct DBuf<unsigned long long>;
template struct DBuf<StepCtx>;
template struct DBuf<RecCol>;
template struct DBuf<TileRec>;
template struct DBuf<uint2>;

}  // namespace ngpu
