/*
 * C ABI of the MI355X-native NEST GPU simulation step.
 *
 * Every function keeps the name, signature and error convention of the
 * reference's C interface, src/nestgpu_C.h:34-353 (implementation
 * src/nestgpu_C.cpp:37-1194), so existing ctypes/C callers bind unchanged:
 *   - one process-global simulator, created on the first call;
 *   - each call clears the error state; on failure NESTGPU_GetErrorCode()
 *     returns 2 (library error), 1 (allocation error) or 255 (other) and
 *     NESTGPU_GetErrorMessage() the message (src/propagate_error.h:4-30);
 *   - with NESTGPU_SetOnException(0) (the default) a failing call prints the
 *     message and exits with the error code, with 1 it returns normally;
 *   - arrays returned by getters are owned by the library and stay valid
 *     until the next call of the same getter.
 * Functions of subsystems outside this build (MPI spike exchange, remote
 * connections, synapse groups/STDP, external neurons) exist and report
 * error 2 with a message saying so.
 *
 * Line numbers below cite the reference declaration each entry replaces.
 */
#ifndef NESTGPU_AMD_C_H
#define NESTGPU_AMD_C_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors and kernel settings (src/nestgpu_C.h:34-52) ---- */
char *NESTGPU_GetErrorMessage();                                   /* :34 */
unsigned char NESTGPU_GetErrorCode();                              /* :36 */
void NESTGPU_SetOnException(int on_exception);                     /* :38 */
int NESTGPU_SetRandomSeed(unsigned long long seed);                /* :40 */
int NESTGPU_SetTimeResolution(float time_res);                     /* :42 */
float NESTGPU_GetTimeResolution();                                 /* :44 */
int NESTGPU_SetMaxSpikeBufferSize(int max_size);                   /* :46 */
int NESTGPU_GetMaxSpikeBufferSize();                               /* :48 */
int NESTGPU_SetSimTime(float sim_time);                            /* :50 */
int NESTGPU_SetVerbosityLevel(int verbosity_level);                /* :52 */

/* ---- node creation and recording (src/nestgpu_C.h:54-68) ---- */
int NESTGPU_Create(char *model_name, int n_neuron, int n_port);    /* :54 */
int NESTGPU_CreatePoissonGenerator(int n_node, float rate);        /* :56 */
int NESTGPU_CreateRecord(char *file_name, char *var_name_arr[],    /* :58 */
                         int *i_node_arr, int *port_arr, int n_node);
int NESTGPU_GetRecordDataRows(int i_record);                       /* :62 */
int NESTGPU_GetRecordDataColumns(int i_record);                    /* :64 */
float **NESTGPU_GetRecordData(int i_record);                       /* :66 */

/* ---- parameters and variables (src/nestgpu_C.h:68-150) ---- */
int NESTGPU_SetNeuronScalParam(int i_node, int n_neuron, char *param_name, float val);
int NESTGPU_SetNeuronArrayParam(int i_node, int n_neuron, char *param_name,
                                float *param, int array_size);
int NESTGPU_SetNeuronPtScalParam(int *i_node, int n_neuron, char *param_name, float val);
int NESTGPU_SetNeuronPtArrayParam(int *i_node, int n_neuron, char *param_name,
                                  float *param, int array_size);
int NESTGPU_IsNeuronScalParam(int i_node, char *param_name);
int NESTGPU_IsNeuronPortParam(int i_node, char *param_name);
int NESTGPU_IsNeuronArrayParam(int i_node, char *param_name);
int NESTGPU_SetNeuronIntVar(int i_node, int n_neuron, char *var_name, int val);
int NESTGPU_SetNeuronScalVar(int i_node, int n_neuron, char *var_name, float val);
int NESTGPU_SetNeuronArrayVar(int i_node, int n_neuron, char *var_name,
                              float *var, int array_size);
int NESTGPU_SetNeuronPtIntVar(int *i_node, int n_neuron, char *var_name, int val);
int NESTGPU_SetNeuronPtScalVar(int *i_node, int n_neuron, char *var_name, float val);
int NESTGPU_SetNeuronPtArrayVar(int *i_node, int n_neuron, char *var_name,
                                float *var, int array_size);
int NESTGPU_IsNeuronIntVar(int i_node, char *var_name);
int NESTGPU_IsNeuronScalVar(int i_node, char *var_name);
int NESTGPU_IsNeuronPortVar(int i_node, char *var_name);
int NESTGPU_IsNeuronArrayVar(int i_node, char *var_name);
int NESTGPU_GetNeuronParamSize(int i_node, char *param_name);
int NESTGPU_GetNeuronVarSize(int i_node, char *var_name);
float *NESTGPU_GetNeuronParam(int i_node, int n_neuron, char *param_name);
float *NESTGPU_GetNeuronPtParam(int *i_node, int n_neuron, char *param_name);
float *NESTGPU_GetArrayParam(int i_node, char *param_name);
int *NESTGPU_GetNeuronIntVar(int i_node, int n_neuron, char *param_name);
int *NESTGPU_GetNeuronPtIntVar(int *i_node, int n_neuron, char *param_name);
float *NESTGPU_GetNeuronVar(int i_node, int n_neuron, char *param_name);
float *NESTGPU_GetNeuronPtVar(int *i_node, int n_neuron, char *param_name);
float *NESTGPU_GetArrayVar(int i_node, char *var_name);

/* ---- simulation (src/nestgpu_C.h:152-160) ---- */
int NESTGPU_Calibrate();
int NESTGPU_Simulate();
int NESTGPU_StartSimulation();
int NESTGPU_SimulationStep();
int NESTGPU_EndSimulation();

/* ---- MPI (src/nestgpu_C.h:162-170): not available in this build ---- */
int NESTGPU_ConnectMpiInit(int argc, char *argv[]);
int NESTGPU_MpiId();
int NESTGPU_MpiNp();
int NESTGPU_ProcMaster();
int NESTGPU_MpiFinalize();

/* ---- host random numbers (src/nestgpu_C.h:172-179) ---- */
unsigned int *NESTGPU_RandomInt(size_t n);
float *NESTGPU_RandomUniform(size_t n);
float *NESTGPU_RandomNormal(size_t n, float mean, float stddev);
float *NESTGPU_RandomNormalClipped(size_t n, float mean, float stddev,
                                   float vmin, float vmax, float vstep);

/* ---- connections (src/nestgpu_C.h:181-237) ---- */
int NESTGPU_Connect(int i_source_node, int i_target_node, unsigned char port,
                    unsigned char syn_group, float weight, float delay);
int NESTGPU_ConnSpecInit();
int NESTGPU_SetConnSpecParam(char *param_name, int value);
int NESTGPU_ConnSpecIsParam(char *param_name);
int NESTGPU_SynSpecInit();
int NESTGPU_SetSynSpecIntParam(char *param_name, int value);
int NESTGPU_SetSynSpecFloatParam(char *param_name, float value);
int NESTGPU_SetSynSpecFloatPtParam(char *param_name, float *array_pt);
int NESTGPU_SynSpecIsIntParam(char *param_name);
int NESTGPU_SynSpecIsFloatParam(char *param_name);
int NESTGPU_SynSpecIsFloatPtParam(char *param_name);
int NESTGPU_ConnectSeqSeq(int i_source, int n_source, int i_target, int n_target);
int NESTGPU_ConnectSeqGroup(int i_source, int n_source, int *i_target, int n_target);
int NESTGPU_ConnectGroupSeq(int *i_source, int n_source, int i_target, int n_target);
int NESTGPU_ConnectGroupGroup(int *i_source, int n_source, int *i_target, int n_target);
int NESTGPU_RemoteConnectSeqSeq(int i_source_host, int i_source, int n_source,
                                int i_target_host, int i_target, int n_target);
int NESTGPU_RemoteConnectSeqGroup(int i_source_host, int i_source, int n_source,
                                  int i_target_host, int *i_target, int n_target);
int NESTGPU_RemoteConnectGroupSeq(int i_source_host, int *i_source, int n_source,
                                  int i_target_host, int i_target, int n_target);
int NESTGPU_RemoteConnectGroupGroup(int i_source_host, int *i_source, int n_source,
                                    int i_target_host, int *i_target, int n_target);

/* ---- name lists (src/nestgpu_C.h:239-269) ---- */
char **NESTGPU_GetIntVarNames(int i_node);
char **NESTGPU_GetScalVarNames(int i_node);
int NESTGPU_GetNIntVar(int i_node);
int NESTGPU_GetNScalVar(int i_node);
char **NESTGPU_GetPortVarNames(int i_node);
int NESTGPU_GetNPortVar(int i_node);
char **NESTGPU_GetScalParamNames(int i_node);
int NESTGPU_GetNScalParam(int i_node);
char **NESTGPU_GetPortParamNames(int i_node);
int NESTGPU_GetNGroupParam(int i_node);
char **NESTGPU_GetGroupParamNames(int i_node);
int NESTGPU_GetNPortParam(int i_node);
char **NESTGPU_GetArrayParamNames(int i_node);
int NESTGPU_GetNArrayParam(int i_node);
char **NESTGPU_GetArrayVarNames(int i_node);
int NESTGPU_GetNArrayVar(int i_node);

/* ---- connection queries (src/nestgpu_C.h:271-290) ---- */
int *NESTGPU_GetSeqSeqConnections(int i_source, int n_source, int i_target,
                                  int n_target, int syn_group, int *n_conn);
int *NESTGPU_GetSeqGroupConnections(int i_source, int n_source, int *i_target,
                                    int n_target, int syn_group, int *n_conn);
int *NESTGPU_GetGroupSeqConnections(int *i_source, int n_source, int i_target,
                                    int n_target, int syn_group, int *n_conn);
int *NESTGPU_GetGroupGroupConnections(int *i_source, int n_source, int *i_target,
                                      int n_target, int syn_group, int *n_conn);
int NESTGPU_GetConnectionStatus(int i_source, int i_group, int i_conn, int *i_target,
                                unsigned char *port, unsigned char *syn_group,
                                float *delay, float *weight);

/* ---- synapse groups (src/nestgpu_C.h:292-306): models "stdp" and
 * "test_syn_model" (src/syn_model.cu:136-238); group 0 is the static synapse.
 * Plastic synapses are delivered at their arrival step; a network with any
 * runs without split RK5 steps, fused delivery or the batch launch, and a
 * dendritic delay on a target of plastic synapses is an error. ---- */
int NESTGPU_CreateSynGroup(char *model_name);
int NESTGPU_GetSynGroupNParam(int i_syn_group);
char **NESTGPU_GetSynGroupParamNames(int i_syn_group);
int NESTGPU_IsSynGroupParam(int i_syn_group, char *param_name);
int NESTGPU_GetSynGroupParamIdx(int i_syn_group, char *param_name);
float NESTGPU_GetSynGroupParam(int i_syn_group, char *param_name);
int NESTGPU_SetSynGroupParam(int i_syn_group, char *param_name, float val);

/* ---- spike counting / spike-time recording (src/nestgpu_C.h:308-320) ---- */
int NESTGPU_ActivateSpikeCount(int i_node, int n_node);
int NESTGPU_ActivateRecSpikeTimes(int i_node, int n_node, int max_n_rec_spike_times);
int NESTGPU_SetRecSpikeTimesStep(int i_node, int n_node, int rec_spike_times_step);
int NESTGPU_GetNRecSpikeTimes(int i_node);
int NESTGPU_GetRecSpikeTimes(int i_node, int n_node, int **n_spike_times_pt,
                             float ***spike_times_pt);

/* ---- external spikes (src/nestgpu_C.h:322-326): outside this build ---- */
int NESTGPU_PushSpikesToNodes(int n_spikes, int *node_id);
int NESTGPU_GetExtNeuronInputSpikes(int *n_spikes, int **node, int **port,
                                    float **spike_height, int include_zeros);

/* ---- group parameters and kernel parameters (src/nestgpu_C.h:328-350) ---- */
int NESTGPU_SetNeuronGroupParam(int i_node, int n_node, char *param_name, float val);
int NESTGPU_IsNeuronGroupParam(int i_node, char *param_name);
float NESTGPU_GetNeuronGroupParam(int i_node, char *param_name);
int NESTGPU_GetNBoolParam();
char **NESTGPU_GetBoolParamNames();
int NESTGPU_IsBoolParam(char *param_name);
int NESTGPU_GetBoolParamIdx(char *param_name);
bool NESTGPU_GetBoolParam(char *param_name);
int NESTGPU_SetBoolParam(char *param_name, bool val);
int NESTGPU_GetNFloatParam();
char **NESTGPU_GetFloatParamNames();
int NESTGPU_IsFloatParam(char *param_name);
int NESTGPU_GetFloatParamIdx(char *param_name);
float NESTGPU_GetFloatParam(char *param_name);
int NESTGPU_SetFloatParam(char *param_name, float val);
int NESTGPU_GetNIntParam();
char **NESTGPU_GetIntParamNames();
int NESTGPU_IsIntParam(char *param_name);
int NESTGPU_GetIntParamIdx(char *param_name);
int NESTGPU_GetIntParam(char *param_name);
int NESTGPU_SetIntParam(char *param_name, int val);
int NESTGPU_RemoteCreate(int i_host, char *model_name, int n_neuron, int n_port);

/* ---- instrumentation of this build (no counterpart in the reference) ----
 * out[0..7] = RK5 attempts, rejected attempts, accepted substeps, delivered
 * (spike, synapse) events, emitted spikes, spikes dropped by a full spike
 * list, neurons put on the RK5 kernel's hot-first lists, neuron-steps split
 * over two RK5 launches; all since calibration (zeros before). */
int NESTGPU_AMD_GetStepStats(unsigned long long *out);
/* Kernel timing with HIP events on the simulator's stream: on = 1 times
 * every launch, on = 2 only the RK5 launches (two events per step instead
 * of two per launch), on = 3 / 4 the same in one step of every 16 (step
 * index within the Simulate call = 7 mod 16; the events lengthen the steps
 * they bracket), 0 stops; a call zeroes the accumulation.
 * ms[5]/launches[5] receive the summed device
 * time and launch count of: RK5 neuron update, iaf_psc_exp update, spike
 * delivery, fold + Poisson input, other (generators, records, step end). */
int NESTGPU_AMD_SetKernelTiming(int on);
int NESTGPU_AMD_GetKernelTimes(double *ms, long long *launches);
/* Destroys the simulator (all nodes, connections, records, device memory);
 * the next call starts a new one with default settings. */
int NESTGPU_AMD_ResetKernel();
/* Blocks until all queued device work of the simulator is done. */
int NESTGPU_AMD_Synchronize();

/* Batch kernel (no counterpart in the reference, which launches one kernel
 * per node group, src/nestgpu.cu:501-503): when every population other than
 * Poisson generators is aeif_cond_beta_multisynapse (2 ports),
 * aeif_cond_alpha_multisynapse (4 ports), aeif_psc_exp,
 * izhikevich_cond_beta (2 ports) with uniform parameters, or iaf_psc_exp
 * (2 to 8 populations), each step is one launch, replayed as a one-node HIP
 * graph.  cfg[12] = {on, variant (0: 3 waves per SIMD, 1: 4), iaf_psc_exp
 * neurons per lane per pass (1, 2, 4), RK5 block weight x 100, graph,
 * grid blocks, populations, dense (tiled) delivery available, used,
 * channels per tile, tiles, delay rows per tile}.
 * tiled: 1 deliver each step's list by the dense tiled launch, 0 inside the
 * next step's launch, -1 unchanged.  A configuration changes scheduling
 * only, never results.  AutotuneBatch times every candidate configuration
 * over `steps` simulated steps (the simulation advances by candidates x
 * steps steps) and keeps the fastest; ms_per_step (or NULL) receives its
 * time. */
int NESTGPU_AMD_GetBatchConfig(int *cfg);
int NESTGPU_AMD_SetBatchConfig(int variant, int iaf_unroll, float rk5_weight, int graph,
                               int tiled);
int NESTGPU_AMD_AutotuneBatch(int steps, double *ms_per_step, int *n_candidates);

/* Synapse table export (no counterpart in the reference; used by tests to
 * replay a network's input): out[k] = outgoing synapses of node i_node + k;
 * the number of synapses into the listed targets; and those synapses in the
 * order of the source-major synapse table (source, target, port, weight,
 * delay in ms, and the synapse's index in that table, which is the Poisson
 * stream id of a poisson_generator's synapse).  Buffers are the caller's,
 * sized by NESTGPU_AMD_CountIncoming. */
int NESTGPU_AMD_GetOutDegrees(int i_node, int n_node, long long *out);
long long NESTGPU_AMD_CountIncoming(const int *targets, int n_target);
int NESTGPU_AMD_GetIncoming(const int *targets, int n_target, int *source, int *target,
                            unsigned char *port, float *weight, float *delay,
                            long long *index);

/* ---- learned step-size predictor (no counterpart in the reference) ----
 * An MLP 8 -> 32 -> 32 -> 1 proposes each RK5 substep's step size for
 * aeif_cond_beta_multisynapse populations with uniform parameters (persistent
 * RK5 kernel); the RK5 error estimate still accepts or rejects every attempt.
 * Parameter vector (1393 floats): mu[8], inv_sigma[8], W1[32][8], b1[32],
 * W2[32][32], b2[32], W3[32], b3; output log2(h/dt).  Raw inputs: V, w, four
 * conductance variables, log2(h_ctrl/dt), log2(last accepted error ratio).
 * The predictor outlives NESTGPU_AMD_ResetKernel.
 * mode: 0 controller only, 1 log (state, accepted h) records, 2 learned. */
int NESTGPU_AMD_SetStepPredictorMode(int mode);
/* Allocates (and clears) the log: capacity records of 10 floats (8 raw
 * inputs, target log2(h/dt), censored flag), every substep of the neurons
 * whose index within their population is a multiple of stride.  A record
 * is censored (its target only a lower bound of a good proposal) when the
 * accepted step was cut by the step end or accepted at the first attempt;
 * after a rejection the accepted step is the target itself. */
int NESTGPU_AMD_StartPredictorLog(long long capacity, int stride);
long long NESTGPU_AMD_GetPredictorLogCount();
int NESTGPU_AMD_GetPredictorLog(float *out, long long first, long long n);
int NESTGPU_AMD_AppendPredictorLog(const float *records, long long n);
/* Input normalisation from the log and He-normal weights (Philox seed). */
int NESTGPU_AMD_InitPredictor(unsigned long long seed);
/* steps Adam updates on minibatches of batch records (multiple of 64) drawn
 * from the log; loss[steps] (or NULL) receives each batch's mean loss. */
int NESTGPU_AMD_TrainPredictor(int steps, int batch, float lr, unsigned long long seed,
                               float *loss);
int NESTGPU_AMD_SetPredictorParams(const float *params, int n);
int NESTGPU_AMD_GetPredictorParams(float *params, int n);
/* Inference through the matrix-core path of the RK5 kernel: raw[n][8] ->
 * out[n] = log2(h/dt). */
int NESTGPU_AMD_PredictorInfer(const float *raw, long long n, float *out);
/* Mean censored loss of a batch and its gradient (1377 floats, the
 * parameters from W1 on) from the training kernels. */
int NESTGPU_AMD_PredictorLossGrad(const float *raw, const float *target,
                                  const unsigned char *censored, int n, float *loss,
                                  float *grad);
/* NESTGPU_AMD_PredictorLossGrad without any allocation (no counterpart in
 * the reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Every pointer is DEVICE memory:
 * raw[n][8], target[n], censored[n] in; loss[1] and grad[1377] out.  The
 * launches are queued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting.  The partial sums live in a fixed
 * array of the library (1.4 MB per GPU, held until the process ends), so
 * n must be in [1, 16384]: any other n launches nothing and sets error code
 * 2 with a message (NESTGPU_GetErrorCode/GetErrorMessage).  The
 * array is shared by every call on one GPU: calls that overlap on one GPU
 * race.  Results equal NESTGPU_AMD_PredictorLossGrad's bit for bit. */
int NESTGPU_AMD_PredictorLossGradNoAlloc(const float *raw, const float *target,
                                         const unsigned char *censored, int n, float *loss,
                                         float *grad, void *stream);

#ifdef __cplusplus
}
#endif

#endif
