// Device helpers shared by the kernels of the simulation step: spike
// emission into the step's spike list with the per-node bookkeeping of
// src/spike_buffer.cu:162-217, the Poisson input of a channel
// (src/poiss_gen.cu:55-86), the fold arithmetic of src/get_spike.cu:142-166,
// and the per-wave counter rows.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"
#include "philox.h"

namespace ngpu {

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
__device__ inline void spike_list_add(const SpikeList& sl, int lst, int src, float height) {
  int pos = atomicAdd(&sl.count[lst], 1);
  if (pos < sl.cap) {
    sl.src[(size_t)lst * sl.cap + pos] = src;
    sl.height[(size_t)lst * sl.cap + pos] = height;
  } else {
    atomicAdd(sl.overflow, 1);
  }
}

__device__ inline void emit_spike(const SpikeList& sl, const SpikeBook& sb,
                                  int lst, int node, int rel, float height,
                                  long long t_idx, double t1, bool lag = false) {
  spike_list_add(sl, lst, lag ? (int)((unsigned int)node | kSpikeLagBit) : node, height);
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
// Delivery of one spike over its source's synapse segment [b, e) (delay
// groups ascending, src/connect.cu:56-74): each synapse adds height*weight
// into ring slot base_slot + d_int (mod n_slot), base_slot being the slot of
// d_int = 0, i.e. (emission time index + 1) mod n_slot.  The arrival step
// t_emit + 1 + d_int is the step whose SpikeBufferUpdate sends the spike in
// the reference (src/spike_buffer.cu:247-317, src/get_spike.cu:60-108).  FP64
// hardware atomics accumulate (double)(height*weight), the reference's
// accumulator type, so sums of float products are exact and independent of
// order.  Threads tid, tid + nthr, ... of the caller take the synapses, four
// records loaded before their atomics are issued.  mode 0: every synapse,
// 1: only d_int = 0, 2: only d_int >= 1.  Returns the synapses delivered.
constexpr int kDeliverUnroll = 4;

// kPlastic: the table may hold plastic synapses (t.grp set), queued for
// their arrival step instead of added into the ring (the delivery launch
// only; fused delivery runs for networks without them).
template <bool kPlastic = false>
__device__ inline unsigned long long deliver_segment(const SynTable& t, int base_slot, float hgt,
                                                      long long b, long long e, int tid,
                                                      int nthr, int mode) {
  unsigned long long n_ev = 0;
  const int stride = nthr * kDeliverUnroll;
  for (long long k0 = b + tid; k0 < e; k0 += stride) {
    unsigned int ch[kDeliverUnroll];
    float w[kDeliverUnroll];
    int d[kDeliverUnroll];
    int g[kDeliverUnroll];
#pragma unroll
    for (int u = 0; u < kDeliverUnroll; ++u) {
      const long long k = k0 + (long long)u * nthr;
      g[u] = 0;
      if (k < e) {
        ch[u] = t.syn_ch[k];
        w[u] = t.syn_w[k];
        d[u] = t.syn_d[k];
        if (kPlastic) g[u] = t.grp[k];
      } else {
        d[u] = -1;
      }
    }
#pragma unroll
    for (int u = 0; u < kDeliverUnroll; ++u) {
      if (d[u] < 0 || (mode == 1 && d[u] != 0) || (mode == 2 && d[u] == 0)) continue;
      int slot = base_slot + d[u];
      if (slot >= t.n_slot) slot -= t.n_slot;
      if (kPlastic && g[u]) {
        // a plastic synapse: its weight is read when the spike arrives
        const int pos = atomicAdd(&t.q_n[slot], 1);
        if (pos < t.q_cap)
          t.q[(size_t)slot * t.q_cap + pos] =
              make_uint2((unsigned int)(k0 + (long long)u * nthr), __float_as_uint(hgt));
        else
          atomicAdd(t.q_over, 1);
      } else {
        unsafeAtomicAdd(&t.ring[(size_t)slot * t.n_ch + ch[u]], (double)(hgt * w[u]));
      }
      ++n_ev;
    }
    // the d_int = 0 group is the segment's head: past it, nothing more
    if (mode == 1 && __all(k0 >= e || d[kDeliverUnroll - 1] > 0 || d[kDeliverUnroll - 1] < 0)) break;
  }
  return n_ev;
}

// Slot of d_int = 0 for a spike list entry (source | lag bit) of the list of
// the step with time index t_idx: (t_emit + 1) mod n_slot, t_emit = t_idx,
// or t_idx - 1 for a lagged spike.
__device__ inline int spike_base_slot(unsigned int raw, long long t_idx, int n_slot) {
  int b = (int)((t_idx + 1) % n_slot) - (int)(raw >> 31);
  return b < 0 ? b + n_slot : b;
}

}  // namespace ngpu
