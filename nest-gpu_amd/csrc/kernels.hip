// HIP kernels of the simulation step for gfx950 (MI355X), apart from the
// RK5 update (rk5_device.h, instantiated per model family in rk5_*.hip):
// the exact iaf_psc_exp update, spike generators and detectors, delivery of
// the spike list over the source-major delay-sorted synapse table into a
// FP64 input ring, the end-of-step fold of ring + Poisson input into the
// receiving state variables, and multimeter rows.  One lane per neuron,
// state and parameters SoA (coalesced 64-lane loads).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "common.h"
#include "device_common.h"
#include "kernels.h"
#include "philox.h"

namespace ngpu {

// ---------------------------------------------------------------------------
// RK5 dispatch over models (the kernels live in rk5_*.hip).
static int persistent_grid_of(int n, int max_blocks) {
  const int waves_needed = (n + kWave - 1) / kWave;
  int grid = (waves_needed + kRK5Block / kWave - 1) / (kRK5Block / kWave);
  if (max_blocks > 0 && grid > max_blocks) grid = max_blocks;
  return grid;
}

long long rk5_waves(int n, int persistent, int max_blocks) {
  if (n <= 0) return 0;
  return persistent ? (long long)(kRK5Block / kWave) * persistent_grid_of(n, max_blocks)
                    : 4LL * ((n + 255) / 256);
}

int rk5_predictor_capable(int model, int n_port) {
  return model == MODEL_AEIF_COND_BETA_MS && n_port >= 1 && n_port <= kFastPort;
}

int rk5_supported(int model, int n_port) {
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS:
    case MODEL_AEIF_COND_ALPHA_MS:
    case MODEL_IZH_COND_BETA:
    case MODEL_AEIF_PSC_ALPHA_MS:
    case MODEL_AEIF_PSC_EXP_MS:
    case MODEL_AEIF_PSC_DELTA:
      return n_port >= 1 && n_port <= kMaxPort;
    case MODEL_AEIF_PSC_EXP:
    case MODEL_AEIF_COND_ALPHA:
    case MODEL_AEIF_COND_BETA:
    case MODEL_AEIF_PSC_ALPHA:
      return n_port == 2;
    default:
      return 0;
  }
}

bool rk5_fold_layout_matches(int model, int n_port, int in_var, int in_port_stride,
                             int w_param, int w_port_stride) {
  const int iv = in_var, is = in_port_stride, wp = w_param, ws = w_port_stride;
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS: return layout_cond_beta_ms(n_port, iv, is, wp, ws);
    case MODEL_AEIF_COND_ALPHA_MS: return layout_cond_alpha_ms(n_port, iv, is, wp, ws);
    case MODEL_IZH_COND_BETA: return layout_izh_cond_beta(n_port, iv, is, wp, ws);
    case MODEL_AEIF_PSC_ALPHA_MS: return layout_psc_alpha_ms(n_port, iv, is, wp, ws);
    case MODEL_AEIF_PSC_EXP_MS: return layout_psc_exp_ms(n_port, iv, is, wp, ws);
    case MODEL_AEIF_PSC_DELTA: return layout_psc_delta(n_port, iv, is, wp, ws);
    case MODEL_AEIF_PSC_EXP:
    case MODEL_AEIF_COND_ALPHA:
    case MODEL_AEIF_COND_BETA:
    case MODEL_AEIF_PSC_ALPHA:
      return n_port == 2 && layout_two_port(model, iv, is, wp, ws);
    default:
      return false;
  }
}

void launch_rk5(int model, int n_port, const RK5Args& a, hipStream_t s) {
  if (a.n <= 0) return;
  if (a.defer_cap > 0 && (!a.persistent || !a.lag || !a.pend))
    throw std::runtime_error("RK5 split steps need the persistent kernel and the lag/pend buffers");
  if (!rk5_supported(model, n_port))
    throw std::runtime_error("RK5: unsupported model / port count");
  switch (model) {
    case MODEL_AEIF_COND_BETA_MS: launch_rk5_cond_beta_ms(n_port, a, s); break;
    case MODEL_AEIF_COND_ALPHA_MS: launch_rk5_cond_alpha_ms(n_port, a, s); break;
    case MODEL_IZH_COND_BETA: launch_rk5_izh_cond_beta(n_port, a, s); break;
    case MODEL_AEIF_PSC_ALPHA_MS: launch_rk5_psc_alpha_ms(n_port, a, s); break;
    case MODEL_AEIF_PSC_EXP_MS: launch_rk5_psc_exp_ms(n_port, a, s); break;
    case MODEL_AEIF_PSC_DELTA: launch_rk5_psc_delta(n_port, a, s); break;
    default: launch_rk5_two_port(model, a, s); break;
  }
}

// ---------------------------------------------------------------------------
// iaf_psc_exp exact update, src/iaf_psc_exp.cu:85-110.  HBM-bound: 16 B of
// state read and written per neuron-step, coefficients from the uniform
// table.  Contraction is off here (an fma buys nothing in an HBM-bound
// kernel), operation by operation as the oracle's iaf_psc_exp; the RK5
// kernels fix a contracted pattern instead.
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
    emit_spike(a.sl, a.sb, a.ctx.lst, a.i0 + i, i, 1.0f, a.ctx.t_idx, a.ctx.t1);
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
    emit_spike(a.sl, a.sb, a.ctx.lst, a.i0 + i, i, a.height[is], a.ctx.t_idx, a.ctx.t1);
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

// parrot_neuron, src/parrot_neuron.cu:55-81: V shows the input of the last
// fold; a non-zero input is re-emitted (height 1 unless hold_spike_height)
// and cleared.
__global__ void parrot_kernel(ParrotArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float sh = a.var[i];                       // input_spike_height
  a.var[(size_t)2 * a.n + i] = sh;           // V
  if (sh != 0.0f) {
    if (a.param[i] == 0.0f) sh = 1.0f;       // hold_spike_height
    a.var[i] = 0.0f;
    emit_spike(a.sl, a.sb, a.ctx.lst, a.i0 + i, i, sh, a.ctx.t_idx, a.ctx.t1);
  }
}

void launch_parrot(const ParrotArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(parrot_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Delivery launch: every spike of this step's list is expanded over its
// source's synapse segment by one 256-lane block (deliver_segment).  Replaces
// the reference's per-node spike ring + send list + per-(spike, group)
// blocks (src/spike_buffer.cu:247-317, src/get_spike.cu:60-108).
__global__ void __launch_bounds__(256) deliver_kernel(DeliverArgs a) {
  const int lst = a.ctx.lst;
  int n_spk = a.sl.count[lst];
  if (n_spk > a.sl.cap) n_spk = a.sl.cap;
  const int* srcs = a.sl.src + (size_t)lst * a.sl.cap;
  const float* hts = a.sl.height + (size_t)lst * a.sl.cap;
  unsigned long long n_ev = 0;
  for (int s = blockIdx.x; s < n_spk; s += gridDim.x) {
    const unsigned int raw = (unsigned int)srcs[s];
    const int src = (int)(raw & ~kSpikeLagBit);
    const int mode = (a.skip_lagged_d0 && (raw >> 31)) ? 2 : 0;
    const int base = spike_base_slot(raw, a.ctx.t_idx, a.st.n_slot);
    if (a.st.grp)
      n_ev += deliver_segment<true>(a.st, base, hts[s], a.st.row_ptr[src], a.st.row_ptr[src + 1],
                                    threadIdx.x, blockDim.x, mode);
    else
      n_ev += deliver_segment(a.st, base, hts[s], a.st.row_ptr[src], a.st.row_ptr[src + 1],
                              threadIdx.x, blockDim.x, mode);
  }
  n_ev = wave_sum(n_ev);
  if ((threadIdx.x & (kWave - 1)) == 0 && a.ctr && n_ev)
    a.ctr[((size_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * kCtrCols + 3] += n_ev;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (a.spikes && n_spk) *a.spikes += (unsigned long long)n_spk;
    // the next step writes list (lst + 1) mod 3: clear its count now (it held
    // the list of step t_idx - 2, delivered two steps ago)
    a.sl.count[lst == kSpikeLists - 1 ? 0 : lst + 1] = 0;
  }
}

void launch_deliver(const DeliverArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(deliver_kernel, dim3(kDeliverBlocks), dim3(256), 0, s, a);
}

// Dense delivery (TiledDeliverArgs).  The work of a step is a list of items
// (tile, k, K): the spike groups k, k + K, k + 2K, ... of the step's list
// into one channel tile.  Persistent blocks take items from the list: block
// b first item b, then the next free one from an atomic counter.  For an item, the block's waves take the list's spikes 64
// at a time: lane l reads the (first, end) run of (spike l, tile) from its
// source's row of runs (one 8-byte load), a wave prefix sum turns the 64
// runs into one flat
// range of records, and the wave walks that range 64 records per pass, each
// lane finding its spike by a 6-step binary search over the prefix sums
// (lane shuffles) -- every lane busy, and one pass reads a few whole runs
// instead of one record of 64 different runs.  Every record adds
// (double)(height * weight) for ring slot t + r (r = d_int + 1 - lag) into
// the LDS tile acc[r][channel] with FP64 LDS atomics when r < rows, else
// straight into the ring with an FP64 global atomic; then the block adds the
// tile into the ring: a plain read-modify-write of rows r < rows when the
// tile is one item (no other block touches these entries, no atomic of this
// block targets them, and no other kernel runs meanwhile), FP64 global
// atomics when it is split over K > 1 items.  The sums are the exact FP64
// sums of the float products, as the global-atomic delivery's, so the ring
// receives the same values bit for bit whatever the split and order.
//
// The plan: each item adds its event count to tile_ev[tile]; the last block
// to finish (a counter of finished blocks) reads the step's counts and
// writes the next step's items: tile t split into K_t = ceil(ev_t / E)
// items (at most kTileSplitMax), E = max(events / blocks, split_min), so
// that no item holds much more than a block's share of the step; items
// ordered by size, largest first (log2 buckets), so the blocks that take
// the heavy items early and the light ones fill in behind.  A burst's load
// moves from tile to tile over several steps, which the one-step-old counts
// follow.
constexpr int kTiledThreads = 1024;
constexpr int kTiledUnroll = 4;
constexpr int kRingBatch = 2;

__device__ inline void tiled_plan(const TiledDeliverArgs& a) {
  __shared__ unsigned long long s_total;
  __shared__ unsigned int s_hist[32];
  const int nt = a.ts.n_tiles;
  const int lane = threadIdx.x & (kWave - 1);
  if (threadIdx.x == 0) s_total = 0;
  if (threadIdx.x < 32) s_hist[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long part = 0;
  for (int t = threadIdx.x; t < nt; t += blockDim.x)
    part += __hip_atomic_load(&a.tile_ev[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  part = wave_sum(part);
  if (lane == 0) atomicAdd(&s_total, part);
  __syncthreads();
  // a heuristic: float arithmetic is exact enough here
  const float inv_e = 1.0f / fmaxf((float)s_total / (float)gridDim.x, (float)a.split_min);
  auto split = [&](unsigned int ev, unsigned int& bucket) {
    const float q = ceilf((float)ev * inv_e);
    const unsigned int k = q < 1.0f ? 1u : (q > (float)kTileSplitMax ? (unsigned int)kTileSplitMax : (unsigned int)q);
    const unsigned int size = ev / k;
    bucket = 31u - (unsigned int)__builtin_clz(size | 1u);
    return k;
  };
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    unsigned int bk;
    const unsigned int k = split(__hip_atomic_load(&a.tile_ev[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), bk);
    atomicAdd(&s_hist[bk], k);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // largest bucket first: exclusive bases in descending bucket order
    unsigned int run = 0;
    for (int b = 31; b >= 0; --b) {
      const unsigned int c = s_hist[b];
      s_hist[b] = run;
      run += c;
    }
    a.sched[0] = (int)run;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    unsigned int bk;
    const unsigned int ev = __hip_atomic_load(&a.tile_ev[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int k = split(ev, bk);
    const unsigned int pos = atomicAdd(&s_hist[bk], k);
    for (unsigned int j = 0; j < k; ++j) a.plan[pos + j] = make_uint2((unsigned int)t, j | (k << 16));
    a.tile_ev[t] = 0;
  }
  if (threadIdx.x == 0) {
    a.sched[1] = (int)gridDim.x;
    a.sched[2] = 0;
  }
}

// at most 64 VGPRs: two 16-wave blocks per CU
__global__ void __launch_bounds__(kTiledThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
deliver_tiled_kernel(TiledDeliverArgs a) {
  extern __shared__ double acc[];
  __shared__ int s_item;
  __shared__ unsigned int s_ev;
  __shared__ int s_last;
  const int tc = a.ts.tile_ch, ns = a.n_slot, rows = a.ts.rows;
  const int n_acc = rows * tc;
  const int ts0 = (int)(a.ctx.t_idx % ns);   // ring slot of r = 0
  const int lst = a.ctx.lst;
  int n_spk = a.sl.count[lst];
  if (n_spk > a.sl.cap) n_spk = a.sl.cap;
  const int* srcs = a.sl.src + (size_t)lst * a.sl.cap;
  const float* hts = a.sl.height + (size_t)lst * a.sl.cap;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int n_wv = blockDim.x / kWave;
  const int n_items = a.sched[0];
  const size_t row = (size_t)a.ts.n_tiles;
  // tiles can outnumber the counter rows: the waves add their counts into
  // the first kCtrMinRows rows with atomics
  unsigned long long* ctr_row =
      a.ctr ? a.ctr + (size_t)(((long long)blockIdx.x * n_wv + wv) % kCtrMinRows) * kCtrCols + 3 : nullptr;
  // block b's first item is item b; the counter hands out the rest (it
  // starts at the grid size)
  for (bool first = true;; first = false) {
    if (threadIdx.x == 0) {
      s_item = first ? (int)blockIdx.x : atomicAdd(&a.sched[1], 1);
      s_ev = 0;
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= n_items) break;
    const uint2 it = a.plan[item];
    const int tile = __builtin_amdgcn_readfirstlane((int)it.x);
    const int kk = __builtin_amdgcn_readfirstlane((int)(it.y & 0xffffu));
    const int nk = __builtin_amdgcn_readfirstlane((int)(it.y >> 16));
    const long long c0 = (long long)tile * tc;
    const uint2* runs = a.ts.runs + tile;
    for (int k = threadIdx.x; k < n_acc; k += blockDim.x) acc[k] = 0.0;
    __syncthreads();
    unsigned int item_ev = 0;
    // the item's groups kk, kk + nk, ...; wave wv takes every n_wv-th of
    // them.  Every load is issued by every lane (a lane past the list's end
    // reads entry 0 and source 0's run, and masks them), so the compiler
    // counts the loads in flight and waits only for the ones it uses.
    const int gs = nk * n_wv * kWave;
    for (int s = (kk + nk * wv) * kWave + lane; s - lane < n_spk; s += gs) {
      // lane l: spike s (or an empty run past the list's end)
      const bool valid = s < n_spk;
      const unsigned int raw = (unsigned int)srcs[valid ? s : 0];
      const unsigned int r0 = valid ? raw : 0u;
      const float hgt = hts[valid ? s : 0];
      const uint2 run = runs[(size_t)(r0 & ~kSpikeLagBit) * row];
      const unsigned int beg = run.x;
      const unsigned int cnt = valid ? run.y - run.x : 0u;
      // bit 0: lagged spike (r = d_int), bit 1: its d_int = 0 synapses are
      // delivered already
      const unsigned int meta = (r0 >> 31) | ((a.skip_lagged_d0 && (r0 >> 31)) ? 2u : 0u);
      unsigned int incl = cnt;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const unsigned int t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
      }
      const unsigned int pre = incl - cnt;
      const unsigned int total = __shfl(incl, kWave - 1, kWave);
      for (unsigned int j0 = 0; j0 < total; j0 += kWave * kTiledUnroll) {
        unsigned int jj[kTiledUnroll], lo[kTiledUnroll];
#pragma unroll
        for (int u = 0; u < kTiledUnroll; ++u) {
          jj[u] = j0 + u * kWave + lane;
          lo[u] = 0;
        }
        // spike of flat record j: the last lane whose run starts at or
        // before j (a lane with an empty run shares its start with the next)
#pragma unroll
        for (int step = kWave / 2; step > 0; step >>= 1) {
#pragma unroll
          for (int u = 0; u < kTiledUnroll; ++u) {
            const unsigned int p = __shfl(pre, (int)lo[u] + step, kWave);
            if (p <= jj[u]) lo[u] += step;
          }
        }
        TileRec rec[kTiledUnroll];
        unsigned int m[kTiledUnroll];
        float h[kTiledUnroll];
#pragma unroll
        for (int u = 0; u < kTiledUnroll; ++u) {
          // record index beg + (j - pre), as one shuffled value (mod 2^32);
          // a lane past the range loads record 0 (no branch around the
          // loads, so all of them are in flight together)
          const unsigned int bp = __shfl(beg - pre, (int)lo[u], kWave);
          m[u] = __shfl(meta, (int)lo[u], kWave);
          h[u] = __shfl(hgt, (int)lo[u], kWave);
          rec[u] = a.ts.rec[jj[u] < total ? bp + jj[u] : 0u];
        }
#pragma unroll
        for (int u = 0; u < kTiledUnroll; ++u) {
          if (jj[u] >= total) continue;
          if ((m[u] & 2u) && rec[u].d == 0) continue;
          const int r = (int)rec[u].d + 1 - (int)(m[u] & 1u);
          const double v = (double)(h[u] * rec[u].w);
          if (r < rows) {
            atomicAdd(&acc[r * tc + rec[u].ch], v);
          } else {
            int slot = ts0 + r;
            if (slot >= ns) slot -= ns;
            atomicAdd(&a.ring[(size_t)slot * a.n_ch + (size_t)(c0 + rec[u].ch)], v);
          }
          ++item_ev;
        }
      }
    }
    {
      const unsigned long long w_ev = wave_sum((unsigned long long)item_ev);
      if (lane == 0 && w_ev) {
        atomicAdd(&s_ev, (unsigned int)w_ev);
        if (ctr_row) atomicAdd(ctr_row, w_ev);
      }
    }
    __syncthreads();
    // tile -> ring, kRingBatch entries per thread at a time: every ring load
    // of the batch is issued before its stores
    const long long left = a.n_ch - c0;
    const int w = left < tc ? (int)left : tc;
    if (nk == 1) {
      for (int k0 = threadIdx.x; k0 < n_acc; k0 += kRingBatch * blockDim.x) {
        double v[kRingBatch], r[kRingBatch];
        size_t gi[kRingBatch];
#pragma unroll
        for (int u = 0; u < kRingBatch; ++u) {
          const int k = k0 + u * (int)blockDim.x;
          const int rr = k / tc, c = k - rr * tc;
          v[u] = (k < n_acc && c < w) ? acc[k] : 0.0;
          const int slot = ts0 + rr >= ns ? ts0 + rr - ns : ts0 + rr;
          gi[u] = (size_t)slot * a.n_ch + (size_t)(c0 + c);
        }
#pragma unroll
        for (int u = 0; u < kRingBatch; ++u)
          if (v[u] != 0.0) r[u] = a.ring[gi[u]];
#pragma unroll
        for (int u = 0; u < kRingBatch; ++u)
          if (v[u] != 0.0) a.ring[gi[u]] = r[u] + v[u];
      }
    } else {
      for (int k = threadIdx.x; k < n_acc; k += blockDim.x) {
        const int rr = k / tc, c = k - rr * tc;
        const double v = c < w ? acc[k] : 0.0;
        const int slot = ts0 + rr >= ns ? ts0 + rr - ns : ts0 + rr;
        if (v != 0.0) atomicAdd(&a.ring[(size_t)slot * a.n_ch + (size_t)(c0 + c)], v);
      }
    }
    if (threadIdx.x == 0 && s_ev) atomicAdd(&a.tile_ev[tile], s_ev);
    __syncthreads();   // acc, s_item and s_ev are reused by the next item
  }
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (a.spikes && n_spk) *a.spikes += (unsigned long long)n_spk;
    // the next step writes list (lst + 1) mod 3: clear its count now (it held
    // the list of step t_idx - 2, delivered two steps ago)
    a.sl.count[lst == kSpikeLists - 1 ? 0 : lst + 1] = 0;
  }
  // The last block to finish plans the next step.  Its inputs are the
  // tile_ev counts, which every block's thread 0 adds with device-scope
  // atomics (performed beyond the per-XCD L2s) before its finished-block
  // increment: a workgroup-scope release makes that thread wait for its
  // atomics to complete, and the planner reads the counts with device-scope
  // atomic loads.  (A device-scope fence here would write back and
  // invalidate the whole L2 of the XCD under every other block still at
  // work.)
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    s_last = atomicAdd(&a.sched[2], 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) tiled_plan(a);
}

void launch_deliver_tiled(const TiledDeliverArgs& a, hipStream_t s) {
  const size_t lds = (size_t)a.ts.rows * a.ts.tile_ch * sizeof(double);
  if (lds > (size_t)kTileLdsBytes || a.ts.n_tiles <= 0 || a.ts.tile_ch > kTileChMax ||
      a.ts.rows < 1 || a.ts.rows > a.n_slot || a.n_slot > 257 || a.blocks <= 0 ||
      !a.tile_ev || !a.plan || !a.sched)
    throw std::runtime_error("tiled delivery: tile does not fit its LDS budget or no work list");
  hipLaunchKernelGGL(deliver_tiled_kernel, dim3((unsigned)a.blocks), dim3(kTiledThreads), lds, s, a);
}

// Fused delivery, right after the step's update launch: the d_int = 0
// synapses of the list's lagged spikes (their slot, t_emit + 1, is this
// step's, which the next launch folds as its lanes take their neurons); the
// rest of each lagged spike, and every other spike, is delivered with the
// list by the next launch.  A wave takes kD0Batch list entries: lanes
// 0..kD0Batch-1 load them together with the list's count, then the segment
// bounds of the lagged ones, then every lane loads one record of each (the
// d_int = 0 group heads the segment and has far fewer than 64 records) before
// any atomic is issued, so a wave waits for memory three times in all.
constexpr int kLagD0Blocks = 64;
constexpr int kD0Batch = 8;

__global__ void __launch_bounds__(256) deliver_lagged_d0_kernel(DeliverArgs a) {
  const int lst = a.ctx.lst;
  const int cap = a.sl.cap;
  const int* srcs = a.sl.src + (size_t)lst * cap;
  const float* hts = a.sl.height + (size_t)lst * cap;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int n_waves = gridDim.x * (blockDim.x / kWave);
  const SynTable& t = a.st;
  unsigned long long n_ev = 0;
  for (int e0 = wave * kD0Batch; e0 < cap; e0 += n_waves * kD0Batch) {
    // entries are read below the list's capacity whatever its count (the
    // count arrives with them; entries at or beyond it are masked after)
    const int e = e0 + lane;
    const bool inb = lane < kD0Batch && e < cap;
    unsigned int raw = 0u;
    float hgt = 0.0f;
    if (inb) {
      raw = (unsigned int)srcs[e];
      hgt = hts[e];
    }
    int n_spk = a.sl.count[lst];
    asm volatile("" ::"v"(raw), "v"(hgt));
    if (n_spk > cap) n_spk = cap;
    if (e0 >= n_spk) break;
    const bool lg = inb && e < n_spk && (raw >> 31);
    long long sb = 0, se = 0;
    if (lg) {
      const int src = (int)(raw & ~kSpikeLagBit);
      sb = t.row_ptr[src];
      se = t.row_ptr[src + 1];
    }
    if (!__any(lg)) continue;
    unsigned int ch[kD0Batch];
    float w[kD0Batch];
    int d[kD0Batch];
    long long kk[kD0Batch];
#pragma unroll
    for (int q = 0; q < kD0Batch; ++q) {
      const long long b = __shfl(sb, q, kWave), en = __shfl(se, q, kWave);
      kk[q] = b + lane;
      d[q] = -1;
      if (kk[q] < en) {   // en = 0 for an unlagged entry
        ch[q] = t.syn_ch[kk[q]];
        w[q] = t.syn_w[kk[q]];
        d[q] = t.syn_d[kk[q]];
      }
    }
#pragma unroll
    for (int q = 0; q < kD0Batch; ++q) {
      const int base = spike_base_slot(__shfl(raw, q, kWave), a.ctx.t_idx, t.n_slot);
      const float h = __shfl(hgt, q, kWave);
      if (d[q] == 0) {
        unsafeAtomicAdd(&t.ring[(size_t)base * t.n_ch + ch[q]], (double)(h * w[q]));
        ++n_ev;
      }
      // a d_int = 0 group longer than one record per lane: the rest of it by
      // the whole wave
      if (__shfl(d[q], kWave - 1, kWave) == 0)
        n_ev += deliver_segment(t, base, h, __shfl(kk[q], kWave - 1, kWave) + 1,
                                __shfl(se, q, kWave), lane, kWave, 1);
    }
  }
  n_ev = wave_sum(n_ev);
  if (lane == 0 && a.ctr && n_ev) a.ctr[(size_t)wave * kCtrCols + 3] += n_ev;
}

void launch_deliver_lagged_d0(const DeliverArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(deliver_lagged_d0_kernel, dim3(kLagD0Blocks), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------
// Plastic synapses (StdpArgs).  The weight rules are the reference's,
// evaluated in double as there (src/stdp.cu:36-61, src/test_syn_model.cu:
// 36-41); p[0] is the group's model, p[1..] its parameters in the
// reference's order.
__device__ inline float syn_rule(float w, float dt_ms, const float* p) {
  if ((int)p[0] == kSynModelTest) return w + (p[2] + p[1] * dt_ms);
  const double tau_plus = p[1], tau_minus = p[2], lambda = p[3], alpha = p[4];
  const double mu_plus = p[5], mu_minus = p[6], w_max = p[7];
  const double wd = w;
  double w1;
  if (dt_ms >= 0.0f) {
    const double fact = lambda * exp(-(double)dt_ms / tau_plus);
    w1 = wd + fact * w_max * pow(1.0 - wd / w_max, mu_plus);
  } else {
    const double fact = -alpha * lambda * exp((double)dt_ms / tau_minus);
    w1 = wd + fact * w_max * pow(wd / w_max, mu_minus);
  }
  w1 = w1 > 0.0 ? w1 : 0.0;
  w1 = w1 < w_max ? w1 : w_max;
  return (float)w1;
}

// the slot's arrivals, one thread per queued event (distinct synapses: a
// source fires at most once per step)
__global__ void stdp_arrive_kernel(StdpArgs a) {
  const int t = (int)a.ctx.t_idx;
  const int slot = (int)(a.ctx.t_idx % a.st.n_slot);
  int n = a.st.q_n[slot];
  if (n > a.st.q_cap) n = a.st.q_cap;
  const uint2* q = a.st.q + (size_t)slot * a.st.q_cap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint2 e = q[i];
    const unsigned int k = e.x;
    const float h = __uint_as_float(e.y);
    const unsigned int ch = a.st.syn_ch[k];
    const float w = a.w[k];
    unsafeAtomicAdd(&a.st.ring[(size_t)slot * a.st.n_ch + ch], (double)(h * w));
    a.last_pre[k] = t;
    const int dt_int = t - a.last_post[a.ch_node[ch]];
    if (dt_int > 0 && dt_int < kMaxSynDt)
      a.w[k] = syn_rule(w, -a.dt * (float)dt_int, a.grp_param + (size_t)a.st.grp[k] * kSynGroupParams);
  }
}

// the step's spiking nodes: one block each over its incoming plastic
// synapses; then the node's spike time.  Block 0 empties the slot's queue
// for the step that reuses it.
__global__ void stdp_post_kernel(StdpArgs a) {
  const int t = (int)a.ctx.t_idx;
  const int lst = a.ctx.lst;
  int n_spk = a.sl.count[lst];
  if (n_spk > a.sl.cap) n_spk = a.sl.cap;
  const int* srcs = a.sl.src + (size_t)lst * a.sl.cap;
  for (int s = blockIdx.x; s < n_spk; s += gridDim.x) {
    const int node = (int)((unsigned int)srcs[s] & ~kSpikeLagBit);
    for (long long r = a.rev_ptr[node] + threadIdx.x; r < a.rev_ptr[node + 1]; r += blockDim.x) {
      const unsigned int k = a.rev_idx[r];
      const int dt_int = t - a.last_pre[k];
      if (dt_int >= 0 && dt_int < kMaxSynDt)
        a.w[k] = syn_rule(a.w[k], a.dt * (float)dt_int, a.grp_param + (size_t)a.st.grp[k] * kSynGroupParams);
    }
    if (threadIdx.x == 0) a.last_post[node] = t;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.st.q_n[(int)(a.ctx.t_idx % a.st.n_slot)] = 0;
}

void launch_stdp_step(const StdpArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(stdp_arrive_kernel, dim3(256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(stdp_post_kernel, dim3(1024), dim3(256), 0, s, a);
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
    w[p] = p < a.w_unit_ports ? 1.0f : 0.0f;
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
  if (a.in_port_stride == 0) {
    // every port feeds one variable (aeif_psc_delta: V_m): added port
    // after port, each sum rounded as its own fold
    float v = in0[0];
    bool any = false;
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (acc[p] != 0.0) {
        v = fold_add(v, acc[p], w[p]);
        any = true;
      }
    if (any) *in[0] = v;
    return;
  }
#pragma unroll
  for (int p = 0; p < NP; ++p)
    if (acc[p] != 0.0) *in[p] = fold_add(in0[p], acc[p], w[p]);   // input + 0*w == input exactly
}

template <int NP = 1>
static void launch_fold_np(const FoldArgs& a, dim3 grid, dim3 block, hipStream_t s) {
  if constexpr (NP <= kMaxPort) {
    if (a.n_port == NP) {
      hipLaunchKernelGGL(fold_kernel<NP>, grid, block, 0, s, a);
      return;
    }
    launch_fold_np<NP + 1>(a, grid, block, s);
  } else {
    throw std::runtime_error("fold: more than " + std::to_string(kMaxPort) + " ports");
  }
}

void launch_fold(const FoldArgs& a, hipStream_t s) {
  if (a.n <= 0 || a.n_port <= 0) return;
  const dim3 grid((unsigned)((a.n + 255) / 256)), block(256);
  launch_fold_np(a, grid, block, s);
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