// Batch kernel: all populations of a mixed network (RK5 models of the batch
// catalog with uniform parameters, and iaf_psc_exp) advance in one launch per
// step.  See the BatchArgs comment in kernels.h for the scheme.
//
// Each block belongs to one population (blocks [blk0[p], blk0[p+1])).  An
// RK5 population's waves run exactly the single-population persistent loop
// (rk5_persistent_body: hot-first lists, split steps, the previous step's
// fold as lanes take their neurons), with its arguments read from constant
// memory.  An iaf_psc_exp population's waves fold and update U neurons per
// lane per pass over a contiguous share.  Every wave then takes spikes of
// the previous step's list to deliver (fused delivery), so per step one
// launch does the work of the population launches, the delivery launch and
// the fold launches.  Scheduling never changes arithmetic: a neuron's update
// and every delivered sum are those of the per-population launches.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "rk5_device.h"

namespace ngpu {

__constant__ RK5Args c_batch_rk5[kMaxBatchPops];
__constant__ IafArgs c_batch_iaf[kMaxBatchPops];
__constant__ FoldArgs c_batch_fold[kMaxBatchPops];

// Step values of a batch launch: this step and the previous one from the
// launch's arguments, the split cap from the population's own arguments.
struct StepFromBatch {
  const BatchArgs& b;
  const RK5Args& a;
  __device__ const StepCtx& ctx() const { return b.ctx; }
  __device__ const StepCtx& pctx() const { return b.pctx; }
  __device__ int prefold() const { return b.prefold; }
  __device__ int defer_cap() const { return b.defer_on ? a.defer_cap : 0; }
};

// iaf_psc_exp over a wave's contiguous share, U neurons per lane per pass
// (loads of the U neurons issued together): the previous step's fold of the
// neuron's two channels (ring slot + Poisson input, the fold kernel's
// arithmetic), then the exact update of src/iaf_psc_exp.cu:85-110 (the
// iaf_kernel's arithmetic).  Contraction is off, as in both kernels.
template <int U>
__device__ inline void iaf_batch_body(const IafArgs& a, const FoldArgs& f, const BatchArgs& b,
                                      int wave, int n_waves, int lane) {
#pragma clang fp contract(off)
  enum { I_ex, I_in, V, REF };
  enum { tau_m, C_m, E_L, I_e, Theta, V_reset, tau_ex, tau_in, t_ref, den,
         P20, P11ex, P11in, P21ex, P21in, P22 };
  const int n = a.n;
  constexpr int kPass = kWave * U;
  const long long share = (((long long)n + n_waves - 1) / n_waves + kPass - 1) / kPass * kPass;
  const long long c0 = (long long)wave * share;
  const long long c1 = c0 + share < n ? c0 + share : n;
  for (long long base = c0; base < c1; base += kPass) {
    float iex[U], iin[U], v[U], ref[U];
    double acc[U][2];
    int idx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * kWave + lane;
      idx[u] = i < c1 ? (int)i : -1;
      if (idx[u] < 0) continue;
      iex[u] = a.var[(size_t)I_ex * n + i];
      iin[u] = a.var[(size_t)I_in * n + i];
      v[u] = a.var[(size_t)V * n + i];
      ref[u] = a.var[(size_t)REF * n + i];
      if (b.prefold) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
          acc[u][p] = f.ring[(size_t)b.pctx.slot * f.n_ch + (size_t)(f.ch_base + (long long)p * n + i)];
      }
    }
    auto P = [&](int j, int i) {
      return a.tab.is_per_neuron(j) ? a.param[(size_t)j * n + i] : a.tab.val[j];
    };
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = idx[u];
      if (i < 0) continue;
      if (b.prefold) {
        float in[2] = {iex[u], iin[u]};
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const long long lo = (long long)p * n;
          const double s = poisson_input(f, lo + i, acc[u][p], b.pctx.t1, b.pctx.t_idx,
                                         f.pu_b < lo + n && f.pu_e > lo);
          f.ring[(size_t)b.pctx.slot * f.n_ch + (size_t)(f.ch_base + lo + i)] = 0.0;
          float w = p < f.w_unit_ports ? 1.0f : 0.0f;
          if (f.w_param >= 0) {
            const int j = f.w_param + p * f.w_port_stride;
            w = f.tab.is_per_neuron(j) ? f.param[(size_t)j * n + i] : f.tab.val[j];
          }
          if (s != 0.0) in[p] = fold_add(in[p], s, w);
        }
        iex[u] = in[0];
        iin[u] = in[1];
      }
      if (ref[u] > 0.0f) {
        ref[u] -= 1.0f;
      } else {
        v[u] = v[u] * P(P22, i) + iex[u] * P(P21ex, i) + iin[u] * P(P21in, i) + P(I_e, i) * P(P20, i);
      }
      iex[u] *= P(P11ex, i);
      iin[u] *= P(P11in, i);
      if (v[u] >= P(Theta, i)) {
        emit_spike(a.sl, a.sb, b.ctx.lst, a.i0 + i, i, 1.0f, b.ctx.t_idx, b.ctx.t1);
        v[u] = P(V_reset, i);
        ref[u] = (float)(int)roundf(P(t_ref, i) / a.dt);
      }
      a.var[(size_t)I_ex * n + i] = iex[u];
      a.var[(size_t)I_in * n + i] = iin[u];
      a.var[(size_t)V * n + i] = v[u];
      a.var[(size_t)REF * n + i] = ref[u];
    }
  }
}

template <class M>
__device__ __attribute__((always_inline)) inline void batch_rk5(const BatchArgs& b, int p,
                                                                int wave, int n_waves, int gwave,
                                                                int* my_hot, int* my_spk,
                                                                unsigned long long& n_dev) {
  const RK5Args& a = c_batch_rk5[p];
  const StepFromBatch sv{b, a};
  rk5_persistent_body<M, true, 0>(a, sv, wave, n_waves, gwave, my_hot, my_spk, nullptr, n_dev);
}

template <int WPE>
__global__ void __launch_bounds__(kRK5Block) __attribute__((amdgpu_waves_per_eu(WPE)))
rk5_batch_kernel(BatchArgs b) {
  __shared__ int s_hot[kRK5Block / kWave][kHotLds];
  __shared__ int s_spk[kRK5Block / kWave][kSpkLds];
  const int lane = threadIdx.x & (kWave - 1);
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  constexpr int kWpb = kRK5Block / kWave;
  int p = 0;
  for (int q = 1; q < b.n_pop; ++q)
    if ((int)blockIdx.x >= b.blk0[q]) p = q;
  const int wave = ((int)blockIdx.x - b.blk0[p]) * kWpb + wib;
  const int n_waves = (b.blk0[p + 1] - b.blk0[p]) * kWpb;
  const int gwave = (int)blockIdx.x * kWpb + wib;
  const int g_waves = (int)gridDim.x * kWpb;
  if (b.clear_next_list && blockIdx.x == 0 && threadIdx.x == 0) clear_next_spike_list(b.sl, b.ctx.lst);
  unsigned long long n_dev = 0;
  int* my_hot = s_hot[threadIdx.x / kWave];
  int* my_spk = s_spk[threadIdx.x / kWave];
  switch (b.kind[p]) {
    case BK_COND_BETA2: batch_rk5<AeifCondBeta<2>>(b, p, wave, n_waves, gwave, my_hot, my_spk, n_dev); break;
    case BK_COND_ALPHA4: batch_rk5<AeifCondAlpha<4>>(b, p, wave, n_waves, gwave, my_hot, my_spk, n_dev); break;
    case BK_PSC_EXP: batch_rk5<AeifPscExp>(b, p, wave, n_waves, gwave, my_hot, my_spk, n_dev); break;
    case BK_IZH2: batch_rk5<IzhCondBeta<2>>(b, p, wave, n_waves, gwave, my_hot, my_spk, n_dev); break;
    case BK_IAF1: iaf_batch_body<1>(c_batch_iaf[p], c_batch_fold[p], b, wave, n_waves, lane); break;
    case BK_IAF2: iaf_batch_body<2>(c_batch_iaf[p], c_batch_fold[p], b, wave, n_waves, lane); break;
    case BK_IAF4: iaf_batch_body<4>(c_batch_iaf[p], c_batch_fold[p], b, wave, n_waves, lane); break;
    default: break;
  }
  if (b.fd.enable) n_dev += deliver_prev_list(b.fd, b.sl, gwave, g_waves, lane);
  if (b.ctr) {
    n_dev = wave_sum(n_dev);
    if (lane == 0 && n_dev) b.ctr[(size_t)gwave * kCtrCols + 3] += n_dev;
  }
}

int batch_kind_of(int model, int n_port, int iaf_unroll) {
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS: return n_port == 2 ? BK_COND_BETA2 : -1;
    case MODEL_AEIF_COND_ALPHA_MS: return n_port == 4 ? BK_COND_ALPHA4 : -1;
    case MODEL_AEIF_PSC_EXP: return n_port == 2 ? BK_PSC_EXP : -1;
    case MODEL_IZH_COND_BETA: return n_port == 2 ? BK_IZH2 : -1;
    case MODEL_IAF_PSC_EXP:
      return iaf_unroll == 1 ? BK_IAF1 : iaf_unroll == 2 ? BK_IAF2 : iaf_unroll == 4 ? BK_IAF4 : -1;
    default: return -1;
  }
}

// Whether a population of an RK5 kind qualifies for the uniform-parameter
// loop the batch kernel runs (the single-population launch's test): every
// parameter but the refractory counter uniform, and every divisor with a
// usable reciprocal.
template <class M>
static bool uniform_ok(const ParamTable& t) {
  const uint64_t rest = t.per_neuron[0] & ~(1ull << M::i_ref);
  if (rest != 0 || t.per_neuron[1] != 0) return false;
  for (int j = 0; j < M::NPAR; ++j)
    if (M::divisor(j) && recip_rn(t.val[j]) == 0.0f) return false;
  return true;
}

bool batch_uniform_ok(int kind, const ParamTable& t) {
  switch (kind) {
    case BK_COND_BETA2: return uniform_ok<AeifCondBeta<2>>(t);
    case BK_COND_ALPHA4: return uniform_ok<AeifCondAlpha<4>>(t);
    case BK_PSC_EXP: return uniform_ok<AeifPscExp>(t);
    case BK_IZH2: return uniform_ok<IzhCondBeta<2>>(t);
    default: return true;
  }
}

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void batch_upload(const RK5Args* rk5, const IafArgs* iaf, const FoldArgs* fold, int n_pop,
                  hipStream_t s) {
  if (n_pop < 1 || n_pop > kMaxBatchPops) throw std::runtime_error("batch: population count out of range");
  hip_check(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_batch_rk5), rk5, sizeof(RK5Args) * n_pop, 0,
                                   hipMemcpyHostToDevice, s), "batch upload");
  hip_check(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_batch_iaf), iaf, sizeof(IafArgs) * n_pop, 0,
                                   hipMemcpyHostToDevice, s), "batch upload");
  hip_check(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_batch_fold), fold, sizeof(FoldArgs) * n_pop, 0,
                                   hipMemcpyHostToDevice, s), "batch upload");
}

static const void* batch_kernel_ptr(int variant) {
  if (variant == 0) return reinterpret_cast<const void*>(&rk5_batch_kernel<3>);
  if (variant == 1) return reinterpret_cast<const void*>(&rk5_batch_kernel<4>);
  throw std::runtime_error("batch: unknown kernel variant");
}

int batch_resident_blocks(int variant) {
  static int blocks[kBatchVariants] = {};
  if (variant < 0 || variant >= kBatchVariants) throw std::runtime_error("batch: unknown kernel variant");
  if (blocks[variant] == 0) {
    int per_cu = 0, dev = 0, cus = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel_ptr(variant),
                                                           kRK5Block, 0), "batch occupancy");
    hip_check(hipGetDevice(&dev), "batch occupancy");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "batch occupancy");
    if (per_cu <= 0 || cus <= 0) throw std::runtime_error("batch: kernel cannot be resident");
    blocks[variant] = per_cu * cus;
  }
  return blocks[variant];
}

static void check_plan(const BatchArgs& b, int grid) {
  if (b.n_pop < 1 || b.n_pop > kMaxBatchPops || b.blk0[0] != 0 || b.blk0[b.n_pop] != grid)
    throw std::runtime_error("batch: block partition does not cover the grid");
  for (int p = 0; p < b.n_pop; ++p) {
    if (b.blk0[p + 1] <= b.blk0[p]) throw std::runtime_error("batch: population without blocks");
    if (b.kind[p] < 0 || b.kind[p] >= BK_N) throw std::runtime_error("batch: unknown population kind");
  }
}

void launch_batch(const BatchArgs& b, int grid, int variant, hipStream_t s) {
  check_plan(b, grid);
  if (variant == 0)
    hipLaunchKernelGGL(rk5_batch_kernel<3>, dim3(grid), dim3(kRK5Block), 0, s, b);
  else if (variant == 1)
    hipLaunchKernelGGL(rk5_batch_kernel<4>, dim3(grid), dim3(kRK5Block), 0, s, b);
  else
    throw std::runtime_error("batch: unknown kernel variant");
}

struct BatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t node = nullptr;
  int variant = 0, grid = 0;
};

static hipKernelNodeParams node_params(const BatchGraph& g, const BatchArgs* b) {
  hipKernelNodeParams kp{};
  kp.func = const_cast<void*>(batch_kernel_ptr(g.variant));
  kp.gridDim = dim3(g.grid);
  kp.blockDim = dim3(kRK5Block);
  kp.sharedMemBytes = 0;
  kp.kernelParams = reinterpret_cast<void**>(const_cast<BatchArgs**>(&b));
  kp.extra = nullptr;
  return kp;
}

BatchGraph* batch_graph_create(int variant, int grid, const BatchArgs& b) {
  check_plan(b, grid);
  BatchGraph* g = new BatchGraph();
  g->variant = variant;
  g->grid = grid;
  try {
    hip_check(hipGraphCreate(&g->graph, 0), "batch graph");
    void* argp[1] = {const_cast<BatchArgs*>(&b)};
    hipKernelNodeParams kp = node_params(*g, &b);
    kp.kernelParams = argp;
    hip_check(hipGraphAddKernelNode(&g->node, g->graph, nullptr, 0, &kp), "batch graph node");
    hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "batch graph instantiate");
  } catch (...) {
    batch_graph_destroy(g);
    throw;
  }
  return g;
}

void batch_graph_launch(BatchGraph* g, const BatchArgs& b, hipStream_t s) {
  check_plan(b, g->grid);
  void* argp[1] = {const_cast<BatchArgs*>(&b)};
  hipKernelNodeParams kp = node_params(*g, &b);
  kp.kernelParams = argp;
  hip_check(hipGraphExecKernelNodeSetParams(g->exec, g->node, &kp), "batch graph arguments");
  hip_check(hipGraphLaunch(g->exec, s), "batch graph launch");
}

void batch_graph_destroy(BatchGraph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

}  // namespace ngpu
