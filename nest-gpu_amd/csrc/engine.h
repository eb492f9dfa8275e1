// Host engine of the MI355X simulation step: node populations, connection
// store, calibration (CSR build, channel map, coefficient tables), and the
// per-step launch sequence.  It plays the role of the reference's NESTGPU
// class (src/nestgpu.h:106-718) behind the same C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <functional>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace ngpu {

// Library error, mapped to error code 2 by the C ABI (src/propagate_error.h).
class Error : public std::exception {
 public:
  explicit Error(std::string m) : msg_(std::move(m)) {}
  const char* what() const noexcept override { return msg_.c_str(); }

 private:
  std::string msg_;
};

struct ModelDesc {
  const char* name;
  int id;
  std::vector<std::string> scal_var, port_var, scal_param, port_param,
      group_param, array_param;
  int forced_port;        // -1: chosen by the caller
  bool rk5;
  int in_var;             // input variable (port-block index if in_port_block)
  bool in_port_block;
  int w_param;            // port-param index of the input weight, -1 => 1
  int in_scal_stride = 1; // scalar-block input: variable index step between ports (0: one shared variable)
  bool w_in_scal = false; // w_param is a scalar-param index, one per port in a row
  int w_unit_ports = 1 << 30;  // w_param < 0: ports below this have weight 1, the rest 0
};

const ModelDesc* find_model(const std::string& name);

// Device buffer owned by the engine.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count);
  void free();
  ~DBuf() { free(); }
};

struct Group {
  const ModelDesc* md = nullptr;
  int i0 = 0, n = 0, n_port = 0, nv = 0, np = 0;
  DBuf<float> var, param;
  DBuf<double> x;
  DBuf<float> h;
  DBuf<float> err_last;         // last accepted RK5 error ratio (step-size predictor input)
  // hot-first scheduling of the persistent RK5 kernel
  DBuf<unsigned char> hot_flag;
  DBuf<int> hot_list, hot_count;
  // split RK5 steps (persistent kernel): lagged flag, and the fold input
  // stashed for lagged neurons when the fold is a separate launch
  DBuf<unsigned char> lag;
  DBuf<double> pend;
  bool recorded = false;        // a multimeter column reads this population
  std::vector<float> uni_val;   // value of each parameter if uniform
  std::vector<char> uni;        // parameter is equal over the population
  float h_min_rel = 1.0e-3f, h0_rel = 1.0e-2f, h_min = 0.f, h0 = 0.f;
  long long ch_base = 0;        // first input channel
  // spike generator
  std::vector<std::vector<float>> sg_times, sg_heights;
  DBuf<int> sg_off, sg_tidx, sg_next;
  DBuf<float> sg_h;
  // spike count / spike-time recording
  std::vector<std::string> int_var;
  DBuf<int> spike_count, n_rec;
  DBuf<float> rec_times;
  int max_rec = 0, rec_step = 0;
  std::vector<std::vector<float>> rec_host;   // buffered spike times
  std::vector<int> rec_n_host;
  std::vector<float*> rec_pt_host;
  // poisson input edges into this population, CSR by local channel
  DBuf<int> pe_off, pe_gen;
  DBuf<float> pe_w, pe_delay;
  DBuf<unsigned int> pe_stream;
  bool has_pe = false;
  // or, when every edge fits one uniform run (see FoldArgs), that run
  bool pe_run = false;
  // the last step's fold was left to the next RK5 launch of this population
  bool fold_pending = false;
  long long pu_b = 0, pu_e = 0;
  int pu_gen0 = 0, pu_gstep = 0;
  float pu_w = 0.0f, pu_delay = 0.0f;
  unsigned int pu_stream0 = 0;
  // poisson generator: index of the first generator in the global tables
  int gen_base = 0;
};

struct RecordDef {
  std::string file;
  std::vector<int> group, rel, port, node;
  std::vector<std::string> var;
  std::vector<std::vector<float>> data;
  std::vector<float*> rows;     // returned by GetRecordData
  FILE* fp = nullptr;
};

struct ConnSpecS {
  int rule = 1;   // all_to_all
  int total_num = 0, indegree = 0, outdegree = 0;
};

struct SynSpecS {
  int syn_group = 0, port = 0;
  float weight = 0.f, delay = 0.f;
  float* weight_array = nullptr;
  float* delay_array = nullptr;
};

class Engine {
 public:
  Engine();
  ~Engine();

  // ---- kernel parameters (src/nestgpu.cu:64-99,1705-1947)
  int SetRandomSeed(unsigned long long seed);
  int SetTimeResolution(float dt);
  float GetTimeResolution() const { return dt_; }
  int SetMaxSpikeBufferSize(int m) { max_spike_buffer_size_ = m; return 0; }
  int GetMaxSpikeBufferSize() const { return max_spike_buffer_size_; }
  int SetSimTime(float t) { sim_time_ = t; return 0; }
  int SetVerbosityLevel(int v) { verbosity_ = v; return 0; }
  std::vector<std::string> FloatParamNames() const;
  std::vector<std::string> IntParamNames() const;
  std::vector<std::string> BoolParamNames() const;
  float GetFloatParam(const std::string& n) const;
  int SetFloatParam(const std::string& n, float v);
  int GetIntParam(const std::string& n) const;
  int SetIntParam(const std::string& n, int v);
  bool GetBoolParam(const std::string& n) const;
  int SetBoolParam(const std::string& n, bool v);

  // ---- nodes
  int Create(const std::string& model, int n, int n_port);
  int NNodes() const { return n_nodes_; }
  Group& GroupOf(int node, int n_node, int* rel);
  std::vector<int> NodeListRel(const int* nodes, int n, Group** g);

  // parameters / variables (src/base_neuron.cu:219-1203)
  bool IsScalParam(int node, const std::string& n);
  bool IsPortParam(int node, const std::string& n);
  bool IsArrayParam(int node, const std::string& n);
  bool IsGroupParam(int node, const std::string& n);
  bool IsIntVar(int node, const std::string& n);
  bool IsScalVar(int node, const std::string& n);
  bool IsPortVar(int node, const std::string& n);
  bool IsArrayVar(int node, const std::string& n);
  int GetParamSize(int node, const std::string& n);
  int GetVarSize(int node, const std::string& n);

  void SetParam(int i_node, int n, const std::string& name, float val);
  void SetParam(const int* nodes, int n, const std::string& name, float val);
  void SetParamArr(int i_node, int n, const std::string& name, const float* v, int size);
  void SetParamArr(const int* nodes, int n, const std::string& name, const float* v, int size);
  void SetVar(int i_node, int n, const std::string& name, float val);
  void SetVar(const int* nodes, int n, const std::string& name, float val);
  void SetVarArr(int i_node, int n, const std::string& name, const float* v, int size);
  void SetVarArr(const int* nodes, int n, const std::string& name, const float* v, int size);
  void SetIntVar(int i_node, int n, const std::string& name, int val);
  void SetIntVar(const int* nodes, int n, const std::string& name, int val);
  float* GetParam(int i_node, int n, const std::string& name);
  float* GetParam(const int* nodes, int n, const std::string& name);
  float* GetArrayParam(int node, const std::string& name);
  float* GetVar(int i_node, int n, const std::string& name);
  float* GetVar(const int* nodes, int n, const std::string& name);
  float* GetArrayVar(int node, const std::string& name);
  int* GetIntVar(int i_node, int n, const std::string& name);
  int* GetIntVar(const int* nodes, int n, const std::string& name);
  int SetGroupParam(int i_node, int n, const std::string& name, float val);
  float GetGroupParam(int node, const std::string& name);

  std::vector<std::string> ScalVarNames(int node);
  std::vector<std::string> PortVarNames(int node);
  std::vector<std::string> ScalParamNames(int node);
  std::vector<std::string> PortParamNames(int node);
  std::vector<std::string> GroupParamNames(int node);
  std::vector<std::string> ArrayParamNames(int node);
  std::vector<std::string> ArrayVarNames(int node);
  std::vector<std::string> IntVarNames(int node);

  // ---- connections (src/connect.cu:38-74, src/connect_rules.h:54-338)
  int Connect(int src, int tgt, int port, int syn_group, float w, float delay);
  // synapse groups, src/syn_model.cu:136-238 (group 0: static synapses)
  int CreateSynGroup(const std::string& model);
  int GetSynGroupNParam(int g);
  std::vector<std::string> GetSynGroupParamNames(int g);
  bool IsSynGroupParam(int g, const std::string& name);
  int GetSynGroupParamIdx(int g, const std::string& name);
  float GetSynGroupParam(int g, const std::string& name);
  void SetSynGroupParam(int g, const std::string& name, float v);
  int ConnectRule(const int* src_list, int i_src0, int n_src, const int* tgt_list,
                  int i_tgt0, int n_tgt, const ConnSpecS& cs, const SynSpecS& ss);
  // syn_group < 0: every synapse group
  std::vector<long long> GetConnections(const int* src_list, int i_src0, int n_src,
                                        const int* tgt_list, int i_tgt0, int n_tgt,
                                        int syn_group = -1);
  void GetConnectionStatus(int src, int grp, int conn, int* tgt, unsigned char* port,
                           unsigned char* syn, float* delay, float* w);
  long long NConnections() const { return (long long)c_src_.size(); }
  // export of the synapse table (not part of the reference API): outgoing
  // synapse counts, and the synapses into a set of targets in table order
  // with their index in the source-major table (the Poisson stream id of a
  // generator's synapse)
  void OutDegrees(int i_node, int n, long long* out);
  long long CountIncoming(const int* targets, int n);
  void GetIncoming(const int* targets, int n, int* src, int* tgt, unsigned char* port, float* w,
                   float* delay, long long* index);

  // ---- host random numbers (own Philox stream; src/random.cu counterpart)
  unsigned int* RandomInt(size_t n);
  float* RandomUniform(size_t n);
  float* RandomNormal(size_t n, float mean, float sd);
  float* RandomNormalClipped(size_t n, float mean, float sd, float vmin, float vmax, float vstep);

  // ---- recording
  int CreateRecord(const std::string& file, char** var_names, const int* nodes,
                   const int* ports, int n);
  RecordDef& Record(int i);
  int ActivateSpikeCount(int i_node, int n);
  int ActivateRecSpikeTimes(int i_node, int n, int max_n);
  int SetRecSpikeTimesStep(int i_node, int n, int step);
  int GetNRecSpikeTimes(int node);
  int GetRecSpikeTimes(int i_node, int n, int** n_pt, float*** t_pt);

  // ---- simulation (src/nestgpu.cu:294-635)
  int Calibrate();
  int Simulate();
  int StartSimulation();
  int SimulationStep();
  int EndSimulation();

  // ---- instrumentation (not part of the reference API)
  void Stats(unsigned long long* out);   // attempts, rejected, substeps, events, spikes, dropped, hot listed
  // Kernel timing with HIP events on the simulator stream, per category:
  // 0 RK5 update, 1 iaf_psc_exp update, 2 delivery, 3 fold+Poisson, 4 other.
  enum { KT_RK5 = 0, KT_IAF, KT_DELIVER, KT_FOLD, KT_OTHER, KT_N };
  // 0 off, 1 every launch, 2 RK5 launches only; 3 and 4 the same in the
  // steps it with it % kTimingStride == kTimingPhase only (events around a
  // launch lengthen its step)
  void SetKernelTiming(int mode);
  static constexpr int kTimingStride = 16;
  static constexpr int kTimingPhase = 7;
  // Batch kernel (all populations in one launch per step, replayed as a
  // one-node HIP graph): configuration in effect, a new one (takes effect
  // with the next step; results do not depend on it), and the autotuner,
  // which times every candidate over `steps` simulated steps each (so the
  // simulation advances by candidates x steps) and keeps the fastest.
  // cfg[12] = {on, variant, iaf_unroll, rk5_weight_x100, graph, grid, populations,
  //           tiled delivery available, tiled delivery used, channels per tile, tiles,
  //           delay rows per tile}
  void GetBatchConfig(int* cfg);
  void SetBatchConfig(int variant, int iaf_unroll, float rk5_weight, int graph, int tiled = -1);
  double AutotuneBatch(int steps, int* n_candidates);
  void KernelTimes(double* ms, long long* launches);
  void Synchronize();
  hipStream_t stream() const { return stream_; }

 private:
  void check_uncalibrated(const char* msg) const;
  void upload_param_column(Group& g, int j);
  void set_param_index(Group& g, int j, const std::vector<int>& rel, const float* vals, bool whole);
  void set_var_index(Group& g, int j, const std::vector<int>& rel, const float* vals);
  void sort_connections();
  bool sort_connections_device();
  bool device_connect(long long n) const;
  void sort_connections_host();
  void build_device_tables();
  void calibrate_group(Group& g);
  void upload_poisson_tables();
  void launch_step();
  RK5Args rk5_args(Group& g, const StepCtx& sc, const SpikeList& sl, const SpikeBook& sb);
  void plan_batch();
  void batch_partition();
  void batch_upload_args();
  void batch_step(const StepCtx& sc, const SpikeList& sl);
  void build_tiled(const std::vector<unsigned int>& ch, const std::vector<unsigned short>& dd);
  void deliver_list(const StepCtx& sc, const SpikeList& sl, int skip_lagged_d0);
  bool fused() const { return fused_deliver_ && !tiled_use_; }
  // plastic synapses: groups (model id + parameters, index 0 unused), and
  // the device state once a network has any (build_plastic)
  struct SynGroupS { int model; std::vector<float> p; };
  std::vector<SynGroupS> syn_groups_ = std::vector<SynGroupS>(1);
  const SynGroupS& syn_group_at(int g) const;
  bool plastic_ = false;
  DBuf<unsigned char> d_syn_grp_;
  DBuf<int> d_ch_node_, d_last_pre_, d_last_post_, d_pq_n_, d_pq_over_;
  DBuf<long long> d_rev_ptr_;
  DBuf<unsigned int> d_rev_idx_;
  DBuf<float> d_sg_param_;
  DBuf<uint2> d_pq_;
  int pq_cap_ = 0;
  void build_plastic();
  void plastic_step(const StepCtx& sc, const SpikeList& sl);
  FoldArgs make_fold_args(const Group& g);
  SynTable syn_table() const;
  void fetch_records(long long n_rows);
  void buffer_rec_spike_times(Group& g);
  int param_idx(const Group& g, const std::string& name, int port) const;
  int var_idx(const Group& g, const std::string& name, int port) const;
  void random_block(uint32_t out[4]);
  void timed(int cat, const std::function<void()>& launch, hipStream_t st = nullptr);
  void sync_streams();
  StepCtx step_ctx(long long it) const;
  void flush_timing();

  // kernel params
  float dt_ = 0.1f;
  float sim_time_ = 1000.f;
  float max_spike_num_fact_ = 1.f, max_spike_per_host_fact_ = 1.f;
  unsigned long long seed_ = 54321ull, kernel_seed_ = 54321ull + 12345ull;
  int verbosity_ = 4, max_spike_buffer_size_ = 20, remote_spike_height_ = 0;
  bool print_time_ = false;
  uint64_t rng_ctr_ = 0;

  std::vector<Group*> groups_;
  std::vector<std::vector<float>> stage_var_, stage_param_;   // pre-calibration state
  std::vector<int> node_group_;
  int n_nodes_ = 0;

  // connection store (insertion order until sorted; sorting is stable by
  // (source, d_int), the order NetConnection::Insert maintains)
  std::vector<int> c_src_, c_tgt_, c_d_;
  std::vector<unsigned char> c_port_, c_syn_;
  std::vector<float> c_w_;
  bool sorted_ = true;
  std::vector<long long> row_ptr_;      // [n_nodes+1]
  std::vector<long long> grp_ptr_;      // per source: first delay group index
  std::vector<long long> grp_start_;    // per delay group: first connection
  std::vector<int> has_out_;

  // device tables
  DBuf<long long> d_row_ptr_;
  DBuf<unsigned int> d_syn_ch_;
  DBuf<float> d_syn_w_;
  DBuf<unsigned short> d_syn_d_;
  DBuf<double> d_ring_;
  long long n_ch_ = 0;
  int n_slot_ = 2;
  int max_d_int_ = 0;
  DBuf<int> d_spk_count_, d_spk_src_, d_overflow_;
  DBuf<float> d_spk_h_;
  int spk_cap_ = 0;
  DBuf<long long> d_last_idx_;
  DBuf<float> d_last_h_;
  bool rk5_persistent_ = true;         // persistent lane-refill RK5 kernel (else lane per neuron)
  int rk5_learned_cap_ = 14;           // the same with learned step-size proposals
  int rk5_defer_cap_ = 23;             // attempts after which a step is split over two launches (0: never)
  bool rk5_defer_env_ = false;         // cap set by NGPU_RK5_DEFER (also for learned proposals)
  float rk5_hot_margin_ = 4.0f;        // hot list from V_th + margin * Delta_T (AdEx models)
  bool rk5_hot_only_ = true;           // waves holding hot entries take no bulk share
  bool poisson_runs_ = true;           // uniform poisson connections as runs, not CSR
  int rk5_hot_lanes_ = 32;             // hot entries per wave
  bool fold_in_rk5_ = true;            // split-step RK5 populations fold inside the next RK5 launch
  bool fused_deliver_env_ = true;      // allow the RK5 launch to deliver the previous step's spikes
  bool fused_deliver_ = false;         // ... and the network qualifies (one emitting RK5 population)
  bool prev_list_pending_ = false;     // the previous step's list awaits delivery by the next RK5 launch
  // lagged spikes' d_int = 0 synapses (fused delivery): 0 a launch after the
  // update launch, 1 by the emitting wave at its exit
  int lag_d0_mode_ = 0;
  // dense (tiled) delivery
  int tiled_env_ = -1;                 // NGPU_TILED: -1 when dense, 0 never, 1 whenever the table fits
  bool tiled_on_ = false;              // the tiled copy of the synapse table exists
  bool tiled_use_ = false;             // deliver with it (a batch autotuner knob)
  int tile_ch_ = 0, n_tiles_ = 0, tile_rows_ = 0;
  int choose_tile_ch(int rows) const;
  DBuf<uint2> d_toff_;                // (first, end) record of each (source, tile)
  DBuf<TileRec> d_trec_;
  DBuf<unsigned int> d_tev_;          // events per tile of the last step
  DBuf<uint2> d_tplan_;               // work items of the next step
  DBuf<int> d_tsched_;                // {items, next item, finished blocks, 0}
  int tiled_blocks_ = 0;              // persistent grid of the tiled delivery
  int tile_split_min_ = kTileSplitMin;  // NGPU_TILE_SPLIT_MIN
  bool defer_ok_ = false;              // inside Simulate() and not its last step
  int rk5_max_blocks_ = 1024;
  // batch kernel
  bool batch_env_ = true;              // NGPU_BATCH=0: one launch per population
  bool batch_on_ = false;              // the network qualifies (every emitter in the batch catalog)
  bool batch_active_ = false;          // ... and this Simulate call uses it (predictor off)
  bool batch_fold_pending_ = false;    // the previous step's fold is left to the next batch launch
  int batch_variant_ = 0;              // register budget: 0 = 3 waves per SIMD, 1 = 4
  int batch_iaf_unroll_ = 2;           // iaf_psc_exp neurons per lane per pass
  float batch_rk5_weight_ = 4.0f;      // blocks per RK5 neuron and state variable, relative to iaf
  bool batch_graph_on_ = true;         // replay the launch as a HIP graph node
  int batch_grid_ = 0;
  std::vector<int> batch_groups_;      // group index of each batch population
  BatchArgs batch_plan_{};             // kinds and block partition
  BatchGraph* batch_graph_ = nullptr;
  DBuf<unsigned long long> d_ctr_;     // per-wave counter rows [rows][kCtrCols]
  DBuf<unsigned long long> d_stats_;   // delivered spikes
  // poisson generator tables
  int n_gen_ = 0;
  DBuf<int> d_gen_chunks_;
  DBuf<double> d_gen_exp_;
  DBuf<float> d_gen_origin_, d_gen_start_, d_gen_stop_;
  // records
  std::vector<RecordDef> records_;
  DBuf<RecCol> d_rec_cols_;
  DBuf<float> d_rec_buf_;
  int rec_ncol_ = 0;

  hipStream_t stream_ = nullptr;
  bool timing_ = false;
  bool timing_rk5_only_ = false;
  int timing_stride_ = 1;
  std::vector<hipEvent_t> ev_free_;
  struct EvPair { hipEvent_t a, b; int cat; };
  std::vector<EvPair> ev_pending_;
  double kt_ms_[KT_N] = {0, 0, 0, 0, 0};
  long long kt_n_[KT_N] = {0, 0, 0, 0, 0};
  bool calibrated_ = false, first_sim_ = true;
  double t_min_ = 0.0, neural_time_ = 0.0, neur_t0_ = 0.0;
  long long it_ = 0, nt_ = 0;
  long long rec_rows_pending_ = 0;
  long long rec_row_base_ = 0, rec_buf_rows_ = 0;
  double build_real_time_ = 0.0, start_real_time_ = 0.0;
};

}  // namespace ngpu
