// Host-side launchers of the HIP kernels (kernels.hip).  All launches are
// asynchronous on the given stream; none synchronises with the host.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace ngpu {

// Per-wave counter slots: each wave of a launch adds its counts into row
// (global wave index) with plain loads and stores, and the host sums the
// rows when asked.  Columns: RK5 attempts, rejected attempts, accepted
// substeps, delivered (spike, synapse) pairs, neurons put on a hot list,
// neuron-steps split over two RK5 launches.  The array holds at least
// kCtrMinRows rows; a launch with more waves (the dense delivery's) adds
// with atomics into row (wave mod kCtrMinRows).
constexpr int kCtrCols = 6;
constexpr int kDeliverBlocks = 1024;
constexpr int kCtrMinRows = 4 * kDeliverBlocks;

enum ModelId {
  MODEL_AEIF_COND_BETA_MS = 0,
  MODEL_AEIF_COND_ALPHA_MS,
  MODEL_AEIF_PSC_EXP,
  MODEL_IZH_COND_BETA,
  MODEL_AEIF_COND_ALPHA,
  MODEL_AEIF_COND_BETA,
  MODEL_AEIF_PSC_ALPHA,
  MODEL_AEIF_PSC_ALPHA_MS,
  MODEL_AEIF_PSC_DELTA,
  MODEL_AEIF_PSC_EXP_MS,
  MODEL_IAF_PSC_EXP,
  MODEL_POISSON_GEN,
  MODEL_SPIKE_GEN,
  MODEL_SPIKE_DETECTOR,
  MODEL_PARROT,
  N_MODEL_ID
};

// Fold, src/get_spike.cu:142-166 plus poisson_generator input: the ring
// slot of a step and the Poisson draws of that step are added to each
// channel's input variable.
struct FoldArgs {
  int n, n_port;
  long long ch_base;
  float* var;
  int in_var;            // var index of port 0's input
  int in_port_stride;    // var index stride between ports
  const float* param;
  ParamTable tab;
  int w_param;           // param index of port 0's weight, -1 => constant weights:
  int w_port_stride;     //   1.0 for ports below w_unit_ports, 0.0 above
  int w_unit_ports;
  double* ring;
  long long n_ch;
  StepCtx ctx;           // this step's time (by value)
  // poisson direct connections into this population, either one uniform
  // run (one edge per channel of [pu_b, pu_e), same weight and delay,
  // generator pu_gen0 + pu_gstep * r and stream pu_stream0 + r for the
  // channel r places into the run; pu_e > pu_b) or CSR by local channel
  long long pu_b, pu_e;
  int pu_gen0, pu_gstep;
  float pu_w, pu_delay;
  unsigned int pu_stream0;
  const int* pe_off;     // [n*n_port+1] or null
  const int* pe_gen;     // generator index
  const float* pe_w;
  const float* pe_delay;
  const unsigned int* pe_stream;
  const int* gen_chunks;
  const double* gen_exp;
  const float* gen_origin;
  const float* gen_start;
  const float* gen_stop;
  unsigned long long seed;
  // split RK5 steps with this fold launch: for a neuron still inside its
  // step (lag[i] = 1) the channel's input is stored in pend[idx] instead of
  // added; the neuron's next RK5 launch adds it when the step ends
  const unsigned char* lag;
  double* pend;
};

// Synapse table and input ring, as delivery reads them.
struct SynTable {
  const long long* row_ptr;    // [n_nodes+1] synapse segment per source
  const unsigned int* syn_ch;  // [n_syn] target channel
  const float* syn_w;          // [n_syn] weight
  const unsigned short* syn_d; // [n_syn] delay slot d_int (ascending within a segment)
  double* ring;                // [n_slot][n_ch]
  long long n_ch;
  int n_slot;
  // plastic synapses (synapse group >= 1; grp null when the network has
  // none): delivered at their arrival step, through a queue per ring slot
  const unsigned char* grp;    // [n_syn] synapse group, 0 = static
  uint2* q;                    // [n_slot][q_cap] (synapse index, height bits)
  int* q_n;                    // [n_slot] entries queued per slot
  int q_cap;
  int* q_over;                 // entries that found their slot's queue full
};

// Synapse groups' plasticity rules, src/syn_model.cu:41-56: "stdp"
// (src/stdp.cu:36-61) and "test_syn_model" (src/test_syn_model.cu:36-41).
constexpr int kSynModelTest = 1;
constexpr int kSynModelStdp = 2;
constexpr int kSynGroupParams = 8;   // model id, then up to 7 parameters
constexpr int kMaxSynDt = 16384;     // src/syn_model.h:33

// One step of plastic synapses (after the step's delivery, before the fold
// of its slot): stdp_arrive_kernel takes the slot's queue -- each event adds
// height * weight into the ring, records its arrival step, and depresses
// the weight if the target fired earlier (src/get_spike.cu:80-95) --, then
// stdp_post_kernel potentiates the incoming plastic synapses of every
// neuron that fired this step (src/rev_spike.cu:40-85) and records its
// spike time.
struct StdpArgs {
  SynTable st;
  float* w;                    // = st.syn_w, updated in place
  const int* ch_node;          // [n_ch] target node of each input channel
  int* last_pre;               // [n_syn] arrival step of the last spike
  int* last_post;              // [n_nodes] step of the node's last spike
  const long long* rev_ptr;    // [n_nodes+1] incoming plastic synapses per node
  const unsigned int* rev_idx; // their synapse indices
  const float* grp_param;      // [n_groups][kSynGroupParams]
  StepCtx ctx;
  SpikeList sl;
  float dt;
};
void launch_stdp_step(const StdpArgs& a, hipStream_t s);

// Delivery done by the persistent RK5 kernel (one spike-emitting population)
// or the batch kernel: its idle waves deliver the previous step's spike list
// (the step's fold of that list's earliest slot comes one launch later).  The
// d_int = 0 synapses of a lagged spike (emitted in the previous step's time)
// land in the slot the next launch folds, so they are delivered before it:
// by deliver_lagged_d0_kernel right after this launch (lag_d0_at_exit = 0),
// or by the emitting wave when it exits (lag_d0_at_exit = 1; that puts their
// loads and atomics at the end of the spike chains, which end the launch).
struct FusedDeliver {
  int enable;            // deliver the previous step's list in this launch
  int lag_d0_at_exit;    // lagged spikes' d_int = 0 synapses: 0 by a launch, 1 at wave exit
  int skip_lagged_d0;    // the previous list's lagged spikes had their d_int = 0 synapses delivered
  int prev_lst;          // list index of the previous step
  long long prev_t_idx;  // its time index
  SynTable st;
  unsigned long long* spikes;  // delivered spikes (one add per launch) or null
};

struct RK5Args {
  int n;                 // neurons in the population
  int i0;                // global id of the first neuron
  float* var;            // SoA [n_var][n]
  float* param;          // SoA [n_param][n]
  double* x;             // [n] integration time (FP64, src/rk5.h:180)
  float* h;              // [n] proposed step
  ParamTable tab;        // uniform parameters
  float h_min;
  float dt;
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  SpikeBook sb;
  unsigned long long* ctr;    // per-wave counters [slot][kCtrCols] or null
  int persistent;        // 1: persistent lane-refill kernel, 0: one lane per neuron
  int max_blocks;        // grid limit of the persistent kernel
  // step-size predictor (aeif_cond_beta_multisynapse with uniform parameters)
  int pred_mode;                   // 0 controller only, 1 log (state, accepted h), 2 learned proposal
  float* err_last;                 // [n] last accepted error ratio (modes 1, 2)
  const float* pred_params;        // predictor parameter vector (mode 2)
  float* log_buf;                  // [log_cap][kLogCols] (mode 1)
  unsigned long long* log_count;   // records written (may exceed log_cap)
  long long log_cap;
  int log_stride;                  // log neurons with index % log_stride == 0
  // hot-first scheduling of the persistent kernel: neurons that ended a
  // step about to spike (or split) are listed, and the next launch gives
  // them one lane each in the first waves.  Each wave collects its entries
  // in LDS and appends them with one atomic when it exits; lists by
  // t_idx mod 3 (read, written, cleared for the launch after next).
  unsigned char* hot_flag;         // [2][n] neuron listed hot for the step (by t_idx parity), or null
  int* hot_list;                   // [3][n]
  int* hot_count;                  // [3]
  float hot_margin;                // AdEx models: listed from V >= V_th + hot_margin * Delta_T
  int hot_only;                    // waves holding hot entries take no bulk share
  int hot_lanes;                   // hot entries per wave (1..64) while they fit the grid
  // split steps: a neuron still integrating after defer_cap attempts of a
  // step (at a substep boundary) is stored with lag = 1, finishes that step
  // first in the next launch, then adds that step's input (folded by that
  // launch itself, or stashed in pend by a fold launch) and goes on; its
  // spike is delivered with its own emission time
  unsigned char* lag;              // [n] or null (no split steps)
  double* pend;                    // [n_port][n]
  int defer_cap;                   // 0: no new split in this launch
  // the previous step's fold of this population, done by the lane that
  // takes each neuron (prefold = 1; pf.ctx is the previous step's): the
  // separate fold launch of that step was skipped
  int prefold;
  FoldArgs pf;
  FusedDeliver fd;
  int clear_next_list;   // clear the spike count of list (lst + 1) mod 3 (fused delivery)
};

// Whether the persistent RK5 kernel's compile-time input layout of a model
// (input variable and weight parameter per port) is the fold's layout, so a
// lagged neuron can apply its stashed input itself.
bool rk5_fold_layout_matches(int model_id, int n_port, int in_var, int in_port_stride,
                             int w_param, int w_port_stride);

struct IafArgs {
  int n, i0;
  float* var;
  float* param;
  ParamTable tab;
  float dt;
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  SpikeBook sb;
};

struct SpikeGenArgs {
  int n, i0;
  const int* off;        // [n+1] spikes per generator (CSR)
  const int* t_idx;      // spike step indexes (relative to Simulate start)
  const float* height;
  int* next;             // [n] next spike to emit
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  SpikeBook sb;
};

struct DetectorArgs {
  int n;
  float* var;
  float* param;
};

struct ParrotArgs {
  int n, i0;
  float* var;            // [input_spike_height, dummy_input, V] x n
  const float* param;    // [hold_spike_height, den_delay] x n
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  SpikeBook sb;
};

// Copy of the synapse table for dense delivery (deliver_tiled_kernel): input
// channels are cut into tiles of tile_ch consecutive channels; within each
// source's segment the synapses are grouped by tile, each as one 8-byte
// record (channel within the tile, delay slot, weight), and runs holds the
// (first, end) record of every (source, tile).  A spike of step t reaches
// ring slot t + r, r = d_int + 1 - lag (lag 1 for a spike emitted by a split
// step one step late); the LDS tile holds the rows r < rows (the delays of
// nearly all synapses), the rest go to the ring with FP64 global atomics.
struct TileRec {
  unsigned short ch;     // channel - tile * tile_ch
  unsigned char d;       // delay slot d_int
  unsigned char pad;
  float w;
};
struct TiledSyn {
  const uint2* runs;           // [n_nodes][n_tiles]
  const TileRec* rec;          // [n_syn]
  int n_tiles, tile_ch;
  int rows;                    // relative slots r kept in LDS: 0 .. rows - 1
};

// Dense delivery of one step's spike list (deliver_tiled_kernel): persistent
// blocks take work items (a channel tile, and every K-th group of 64 spikes
// of the list from group k) from a list planned from the previous step's
// events per tile; for an item, the block's waves read the spikes'
// (first, end) record runs for the tile from the source-major offset rows,
// walk 64 runs at a time as one flat range of records, add every record into
// the LDS tile acc[r][channel] (FP64), and the block adds the tile into the
// ring once.  Replaces global atomics by LDS atomics and one coalesced
// read-modify-write per ring entry that received input, for networks where a
// step's events outnumber the ring's entries.
struct TiledDeliverArgs {
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  TiledSyn ts;
  double* ring;          // [n_slot][n_ch]
  long long n_ch;
  int n_slot;
  int skip_lagged_d0;    // the d_int = 0 synapses of lagged spikes were delivered already
  unsigned long long* ctr;     // per-wave counter rows (column 3) or null
  unsigned long long* spikes;  // delivered spikes (one add per launch) or null
  unsigned int* tile_ev;       // [n_tiles] events of this step per tile (zeroed by the planner)
  uint2* plan;                 // [n_tiles * kTileSplitMax] items: (tile, k | K << 16)
  int* sched;                  // {items, next item, finished blocks}
  int blocks;                  // persistent grid
  int split_min;               // no split below this many events per item
};
void launch_deliver_tiled(const TiledDeliverArgs& a, hipStream_t s);
// a tile is split over up to kTileSplitMax items; by default no split below
// kTileSplitMin events per item
constexpr int kTileSplitMax = 16;
constexpr int kTileSplitMin = 16384;
// LDS bytes of one tile block at most (tile_ch * rows * 8): two blocks of
// 16 waves fill a CU's 160 KB; tile_ch a multiple of 8, at most
// kTileChMax; delay slots below 256 (the record's delay byte)
constexpr int kTileLdsBytes = 80 * 1024 - 512;
constexpr int kTileChMax = 2048;
constexpr int kTileStaticLds = 512;   // the kernel's own LDS beside the tile

struct DeliverArgs {
  StepCtx ctx;           // this step's time (by value)
  SpikeList sl;
  SynTable st;
  int skip_lagged_d0;          // the d_int = 0 synapses of lagged spikes were delivered already
  unsigned long long* ctr;     // per-wave counters [slot][kCtrCols] (column 3) or null
  unsigned long long* spikes;  // delivered spikes (one add per launch) or null
};


struct RecCol {
  const float* ptr;      // value to copy, or null for the "spike" pseudo-variable
  int node;              // global node (spike activity)
  int has_out;           // node has outgoing connections
};

struct RecordArgs {
  const RecCol* cols;
  int ncol;
  float* buf;            // [rows][ncol+1]
  StepCtx ctx;           // this step's time (by value)
  const long long* last_idx;
  const float* last_height;
  long long row_base;    // step index (ctx->it) stored in buffer row 0
  long long max_rows;    // rows in buf
};

// ---------------------------------------------------------------------------
// Batch kernel (rk5_batch.hip): every population of a mixed network advances
// in one launch per step.  The grid is split by blocks among the populations
// (blk0); each block runs the persistent RK5 loop of its population's model,
// or the iaf_psc_exp update, over the population's share of waves, folding
// the previous step's ring slot and Poisson input as lanes take their
// neurons; afterwards every wave of the launch helps deliver the previous
// step's spike list.  Per-step values travel by value in BatchArgs; the
// per-population arguments sit in the batch module's constant memory
// (batch_upload), unchanged from step to step, so the launch can also be
// replayed as a HIP graph node whose arguments are updated per step.
constexpr int kMaxBatchPops = 8;
enum BatchKind {
  BK_COND_BETA2 = 0,   // aeif_cond_beta_multisynapse, 2 ports
  BK_COND_ALPHA4,      // aeif_cond_alpha_multisynapse, 4 ports
  BK_PSC_EXP,          // aeif_psc_exp
  BK_IZH2,             // izhikevich_cond_beta, 2 ports
  BK_IAF1,             // iaf_psc_exp, 1 / 2 / 4 neurons per lane per pass
  BK_IAF2,
  BK_IAF4,
  BK_N
};
// register budget of the batch kernel: 3 (variant 0) or 4 (variant 1) waves per SIMD
constexpr int kBatchVariants = 2;

struct BatchArgs {
  StepCtx ctx;           // this step
  StepCtx pctx;          // the previous step (its fold, prefold = 1)
  int prefold;           // fold the previous step's slot as lanes take their neurons
  int defer_on;          // split steps allowed in this launch (cap per population in RK5Args)
  int clear_next_list;   // clear the count of the list the next step writes
  int n_pop;
  int kind[kMaxBatchPops];
  int blk0[kMaxBatchPops + 1];   // first block of each population; blk0[n_pop] = grid
  SpikeList sl;
  FusedDeliver fd;       // the previous step's list (fd.enable) and lagged spikes at exit
  unsigned long long* ctr;     // per-wave counter rows (global wave index) or null
};

// batch kind of a (model, ports) pair with uniform parameters, or -1
int batch_kind_of(int model, int n_port, int iaf_unroll);
// an RK5 population's parameters qualify for the batch kernel's
// uniform-parameter loop (iaf kinds: always)
bool batch_uniform_ok(int kind, const ParamTable& t);
// per-population arguments: rk5[k] for RK5 kinds, iaf[k] and fold[k] for iaf
void batch_upload(const RK5Args* rk5, const IafArgs* iaf, const FoldArgs* fold, int n_pop,
                  hipStream_t s);
int batch_resident_blocks(int variant);      // blocks of one launch resident at once
void launch_batch(const BatchArgs& b, int grid, int variant, hipStream_t s);
// The batch launch as a one-node HIP graph, instantiated once; each launch
// sets the node's arguments to this step's and replays it.
struct BatchGraph;
BatchGraph* batch_graph_create(int variant, int grid, const BatchArgs& b);
void batch_graph_launch(BatchGraph* g, const BatchArgs& b, hipStream_t s);
void batch_graph_destroy(BatchGraph* g);

// Per-family RK5 launchers and input-layout checks (rk5_*.hip), called by
// launch_rk5 / rk5_fold_layout_matches.
void launch_rk5_cond_beta_ms(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_cond_alpha_ms(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_izh_cond_beta(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_psc_alpha_ms(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_psc_exp_ms(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_psc_delta(int n_port, const RK5Args& a, hipStream_t s);
void launch_rk5_two_port(int model, const RK5Args& a, hipStream_t s);
bool layout_cond_beta_ms(int n_port, int iv, int is, int wp, int ws);
bool layout_cond_alpha_ms(int n_port, int iv, int is, int wp, int ws);
bool layout_izh_cond_beta(int n_port, int iv, int is, int wp, int ws);
bool layout_psc_alpha_ms(int n_port, int iv, int is, int wp, int ws);
bool layout_psc_exp_ms(int n_port, int iv, int is, int wp, int ws);
bool layout_psc_delta(int n_port, int iv, int is, int wp, int ws);
bool layout_two_port(int model, int iv, int is, int wp, int ws);

int rk5_supported(int model, int n_port);
void launch_rk5(int model, int n_port, const RK5Args& a, hipStream_t s);
long long rk5_waves(int n, int persistent, int max_blocks);   // waves of one launch
int rk5_predictor_capable(int model, int n_port);                // model supports pred_mode 1, 2
void launch_iaf(const IafArgs& a, hipStream_t s);
void launch_spike_gen(const SpikeGenArgs& a, hipStream_t s);
void launch_detector(const DetectorArgs& a, hipStream_t s);
void launch_parrot(const ParrotArgs& a, hipStream_t s);
void launch_deliver(const DeliverArgs& a, hipStream_t s);   // kDeliverBlocks x 256
// the d_int = 0 synapses of the lagged spikes of this step's list (fused delivery)
void launch_deliver_lagged_d0(const DeliverArgs& a, hipStream_t s);
void launch_fold(const FoldArgs& a, hipStream_t s);
void launch_record(const RecordArgs& a, hipStream_t s);

}  // namespace ngpu
