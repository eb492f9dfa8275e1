// Shared host/device definitions of the MI355X simulation step.
//
// Data layout in HBM (one process, one GPU):
//   * neuron state and parameters are SoA, one lane per neuron:
//       var[j * n + i], param[j * n + i] for variable/parameter j of neuron i
//     with j following the reference's index order (scalar block, then the
//     per-port block), so names map to the same indices as the reference
//     (src/base_neuron.cu:1155-1203) while loads are coalesced.
//   * parameters that are equal over a whole population are also kept in a
//     per-population table passed by value to the kernels (kernel-argument
//     segment -> scalar registers) and are not read from HBM.
//   * every (node, port) input has a global "channel" index; the synapse
//     table stores channels, so delivery needs no node->group lookup.
//   * pending synaptic input lives in a FP64 ring  ring[slot][channel], slot
//     = arrival time index mod n_slot, filled by delivery and folded into
//     the receiving state variable at the end of each step.
#pragma once
#include <cstdint>

namespace ngpu {

// receptor ports per multisynapse population (the reference's MAX_PORT_NUM,
// src/aeif_cond_beta_multisynapse.h:165); populations with up to kFastPort
// ports get the persistent RK5 kernel's uniform-parameter and predictor
// variants
constexpr int kMaxPort = 20;
constexpr int kFastPort = 8;
constexpr int kMaxParam = 14 + 4 * kMaxPort;
constexpr int kWave = 64;
// persistent RK5 kernel: threads per block, and LDS entries per wave for the
// neurons it lists hot in one launch (~1 per wave at 5 Hz; a wave with more
// leaves the rest unlisted, which only delays them)
constexpr int kRK5Block = 256;
constexpr int kHotLds = 128;
// persistent RK5 kernel: LDS entries per wave for its spikes of one launch
// (a wave with more adds the rest to the spike list one atomic each)
constexpr int kSpkLds = 64;

// Time of the step being computed, kept in device memory so the step's
// kernels can be replayed (HIP graphs) without host arguments.
struct StepCtx {
  long long it;        // step index inside the current Simulate() call
  long long t_idx;     // global time index of t1 (src/nestgpu.cu:488)
  double t0;           // neural time at the start of Simulate()
  double t1;           // neural time at the end of this step
  int slot;            // ring slot folded at the end of this step (t_idx mod n_slot)
  int parity;          // t_idx & 1 (hot-first flags)
  int lst;             // spike list written this step (t_idx mod 3)
};

// Per-population parameter table: values of uniform parameters and the
// bitmask of parameters that are per-neuron (read from HBM).
struct ParamTable {
  float val[kMaxParam];
  uint64_t per_neuron[2];
  __host__ __device__ bool is_per_neuron(int j) const {
    return (per_neuron[j >> 6] >> (j & 63)) & 1ull;
  }
};

// Spike list of one step; three lists rotate by t_idx mod 3, so the list of
// step T can be delivered while step T+1 writes its own and the list of step
// T-1 (delivered) is cleared for step T+2.  A source with bit 31 set
// (kSpikeLagBit) was emitted in the previous step by a neuron whose step was
// split over two RK5 launches; it is delivered one slot earlier, i.e. with
// its own emission time.
constexpr unsigned int kSpikeLagBit = 0x80000000u;
constexpr int kSpikeLists = 3;
struct SpikeList {
  int* count;          // [3]
  int* src;            // [3][cap] global node id (| kSpikeLagBit)
  float* height;       // [3][cap]
  int cap;
  int* overflow;       // number of spikes dropped because the list was full
};

// Per-population spike bookkeeping done where a spike is emitted
// (src/spike_buffer.cu:162-217).
struct SpikeBook {
  long long* last_idx;    // [n_nodes] time index of the last spike
  float* last_height;     // [n_nodes]
  int* spike_count;       // [n] or null
  float* rec_times;       // [n][max_rec] or null
  int* n_rec;             // [n] or null
  int max_rec;            // 0 if recording disabled
};

}  // namespace ngpu
