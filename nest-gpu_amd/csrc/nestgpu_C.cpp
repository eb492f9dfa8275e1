// C ABI (include/nestgpu_C.h) over one process-global Engine, with the
// reference's error convention (src/propagate_error.h:4-30,
// src/nestgpu_C.cpp:37-1194).
#include "../../include/nestgpu_C.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <string>
#include <vector>

#include "engine.h"
#include "predictor_host.h"

using ngpu::Engine;
using ngpu::Error;

namespace {

Engine* g_engine = nullptr;
ngpu::ConnSpecS g_conn_spec;
ngpu::SynSpecS g_syn_spec;
std::string g_err_msg;
unsigned char g_err_code = 0;
int g_on_exception = 0;   // 0: exit on error, 1: return to the caller

Engine& E() {
  if (!g_engine) g_engine = new Engine();
  return *g_engine;
}

template <class R, class F>
R guarded(R dflt, F&& fn) {
  g_err_msg.clear();
  g_err_code = 0;
  R ret = dflt;
  try {
    ret = fn();
  } catch (ngpu::Error& e) {
    g_err_msg = e.what();
    g_err_code = 2;
  } catch (std::bad_alloc&) {
    g_err_msg = "Memory allocation error.";
    g_err_code = 1;
  } catch (std::exception& e) {
    g_err_msg = std::string("Error in NESTGPU function: ") + e.what();
    g_err_code = 255;
  } catch (...) {
    g_err_msg = "Error in NESTGPU function.";
    g_err_code = 255;
  }
  if (g_err_code != 0 && g_on_exception == 0) {
    std::cerr << g_err_msg;
    std::exit(g_err_code);
  }
  return ret;
}

// name lists are returned as malloc'ed arrays of malloc'ed strings, as the
// reference does (src/nestgpu_C.cpp:598-613)
char** name_array(const std::vector<std::string>& v) {
  char** a = (char**)std::malloc(std::max<size_t>(v.size(), 1) * sizeof(char*));
  for (size_t i = 0; i < v.size(); ++i) {
    a[i] = (char*)std::malloc(v[i].size() + 1);
    std::strcpy(a[i], v[i].c_str());
  }
  return a;
}

int* conn_array(const std::vector<long long>& ids, int* n_conn) {
  *n_conn = (int)(ids.size() / 3);
  static std::vector<int> out;
  out.resize(std::max<size_t>(ids.size(), 1));
  for (size_t k = 0; k < ids.size(); ++k) out[k] = (int)ids[k];
  return out.data();
}

[[noreturn]] void out_of_scope(const char* what) {
  throw Error(std::string(what) + " is outside the scope of this build");
}

int name_index(const std::vector<std::string>& v, const std::string& n, const char* kind) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == n) return (int)i;
  throw Error(std::string("Unrecognized ") + kind + " parameter " + n);
}

}  // namespace

extern "C" {

char* NESTGPU_GetErrorMessage() { return const_cast<char*>(g_err_msg.c_str()); }
unsigned char NESTGPU_GetErrorCode() { return g_err_code; }
void NESTGPU_SetOnException(int on_exception) { g_on_exception = on_exception; }

int NESTGPU_SetRandomSeed(unsigned long long seed) {
  return guarded(0, [&] { return E().SetRandomSeed(seed); });
}
int NESTGPU_SetTimeResolution(float time_res) {
  return guarded(0, [&] { return E().SetTimeResolution(time_res); });
}
float NESTGPU_GetTimeResolution() {
  return guarded(0.0f, [&] { return E().GetTimeResolution(); });
}
int NESTGPU_SetMaxSpikeBufferSize(int max_size) {
  return guarded(0, [&] { return E().SetMaxSpikeBufferSize(max_size); });
}
int NESTGPU_GetMaxSpikeBufferSize() {
  return guarded(0, [&] { return E().GetMaxSpikeBufferSize(); });
}
int NESTGPU_SetSimTime(float sim_time) {
  return guarded(0, [&] { return E().SetSimTime(sim_time); });
}
int NESTGPU_SetVerbosityLevel(int v) {
  return guarded(0, [&] { return E().SetVerbosityLevel(v); });
}

int NESTGPU_Create(char* model_name, int n_neuron, int n_port) {
  return guarded(0, [&] { return E().Create(model_name, n_neuron, n_port); });
}
int NESTGPU_CreatePoissonGenerator(int, float) {
  return guarded(0, [&]() -> int {
    out_of_scope("CreatePoissonGenerator (the legacy spike-buffer Poisson source; "
                 "use Create(\"poisson_generator\"))");
  });
}
int NESTGPU_CreateRecord(char* file_name, char* var_name_arr[], int* i_node_arr,
                         int* port_arr, int n_node) {
  return guarded(0, [&] {
    return E().CreateRecord(file_name, var_name_arr, i_node_arr, port_arr, n_node);
  });
}
int NESTGPU_GetRecordDataRows(int i_record) {
  return guarded(0, [&] { return (int)E().Record(i_record).data.size(); });
}
int NESTGPU_GetRecordDataColumns(int i_record) {
  return guarded(0, [&] {
    auto& r = E().Record(i_record);
    if (r.data.empty()) throw Error("Record has no data");
    return (int)r.data[0].size();
  });
}
float** NESTGPU_GetRecordData(int i_record) {
  return guarded((float**)nullptr, [&] { return E().Record(i_record).rows.data(); });
}

int NESTGPU_SetNeuronScalParam(int i_node, int n, char* name, float val) {
  return guarded(0, [&] { E().SetParam(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronArrayParam(int i_node, int n, char* name, float* p, int size) {
  return guarded(0, [&] { E().SetParamArr(i_node, n, name, p, size); return 0; });
}
int NESTGPU_SetNeuronPtScalParam(int* i_node, int n, char* name, float val) {
  return guarded(0, [&] { E().SetParam(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronPtArrayParam(int* i_node, int n, char* name, float* p, int size) {
  return guarded(0, [&] { E().SetParamArr(i_node, n, name, p, size); return 0; });
}
int NESTGPU_IsNeuronScalParam(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsScalParam(i_node, name); });
}
int NESTGPU_IsNeuronPortParam(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsPortParam(i_node, name); });
}
int NESTGPU_IsNeuronArrayParam(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsArrayParam(i_node, name); });
}
int NESTGPU_SetNeuronIntVar(int i_node, int n, char* name, int val) {
  return guarded(0, [&] { E().SetIntVar(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronScalVar(int i_node, int n, char* name, float val) {
  return guarded(0, [&] { E().SetVar(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronArrayVar(int i_node, int n, char* name, float* v, int size) {
  return guarded(0, [&] { E().SetVarArr(i_node, n, name, v, size); return 0; });
}
int NESTGPU_SetNeuronPtIntVar(int* i_node, int n, char* name, int val) {
  return guarded(0, [&] { E().SetIntVar(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronPtScalVar(int* i_node, int n, char* name, float val) {
  return guarded(0, [&] { E().SetVar(i_node, n, name, val); return 0; });
}
int NESTGPU_SetNeuronPtArrayVar(int* i_node, int n, char* name, float* v, int size) {
  return guarded(0, [&] { E().SetVarArr(i_node, n, name, v, size); return 0; });
}
int NESTGPU_IsNeuronIntVar(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsIntVar(i_node, name); });
}
int NESTGPU_IsNeuronScalVar(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsScalVar(i_node, name); });
}
int NESTGPU_IsNeuronPortVar(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsPortVar(i_node, name); });
}
int NESTGPU_IsNeuronArrayVar(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsArrayVar(i_node, name); });
}
int NESTGPU_GetNeuronParamSize(int i_node, char* name) {
  return guarded(0, [&] { return E().GetParamSize(i_node, name); });
}
int NESTGPU_GetNeuronVarSize(int i_node, char* name) {
  return guarded(0, [&] { return E().GetVarSize(i_node, name); });
}
float* NESTGPU_GetNeuronParam(int i_node, int n, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetParam(i_node, n, name); });
}
float* NESTGPU_GetNeuronPtParam(int* i_node, int n, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetParam(i_node, n, name); });
}
float* NESTGPU_GetArrayParam(int i_node, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetArrayParam(i_node, name); });
}
int* NESTGPU_GetNeuronIntVar(int i_node, int n, char* name) {
  return guarded((int*)nullptr, [&] { return E().GetIntVar(i_node, n, name); });
}
int* NESTGPU_GetNeuronPtIntVar(int* i_node, int n, char* name) {
  return guarded((int*)nullptr, [&] { return E().GetIntVar(i_node, n, name); });
}
float* NESTGPU_GetNeuronVar(int i_node, int n, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetVar(i_node, n, name); });
}
float* NESTGPU_GetNeuronPtVar(int* i_node, int n, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetVar(i_node, n, name); });
}
float* NESTGPU_GetArrayVar(int i_node, char* name) {
  return guarded((float*)nullptr, [&] { return E().GetArrayVar(i_node, name); });
}

int NESTGPU_Calibrate() { return guarded(0, [&] { return E().Calibrate(); }); }
int NESTGPU_Simulate() { return guarded(0, [&] { return E().Simulate(); }); }
int NESTGPU_StartSimulation() { return guarded(0, [&] { return E().StartSimulation(); }); }
int NESTGPU_SimulationStep() { return guarded(0, [&] { return E().SimulationStep(); }); }
int NESTGPU_EndSimulation() { return guarded(0, [&] { return E().EndSimulation(); }); }

int NESTGPU_ConnectMpiInit(int, char*[]) {
  return guarded(0, [&]() -> int { throw Error("MPI is not available in your build"); });
}
int NESTGPU_MpiId() {
  return guarded(0, [&]() -> int { throw Error("MPI is not available in your build"); });
}
int NESTGPU_MpiNp() {
  return guarded(0, [&]() -> int { throw Error("MPI is not available in your build"); });
}
int NESTGPU_ProcMaster() {
  return guarded(0, [&]() -> int { throw Error("MPI is not available in your build"); });
}
int NESTGPU_MpiFinalize() {
  return guarded(0, [&]() -> int { throw Error("MPI is not available in your build"); });
}

unsigned int* NESTGPU_RandomInt(size_t n) {
  return guarded((unsigned int*)nullptr, [&] { return E().RandomInt(n); });
}
float* NESTGPU_RandomUniform(size_t n) {
  return guarded((float*)nullptr, [&] { return E().RandomUniform(n); });
}
float* NESTGPU_RandomNormal(size_t n, float mean, float sd) {
  return guarded((float*)nullptr, [&] { return E().RandomNormal(n, mean, sd); });
}
float* NESTGPU_RandomNormalClipped(size_t n, float mean, float sd, float vmin, float vmax,
                                   float vstep) {
  return guarded((float*)nullptr,
                 [&] { return E().RandomNormalClipped(n, mean, sd, vmin, vmax, vstep); });
}

int NESTGPU_Connect(int src, int tgt, unsigned char port, unsigned char syn_group, float w,
                    float delay) {
  return guarded(0, [&] { return E().Connect(src, tgt, port, syn_group, w, delay); });
}

// ConnSpec / SynSpec staging, src/connect_rules.cpp:30-260
int NESTGPU_ConnSpecInit() {
  return guarded(0, [&] { E(); g_conn_spec = ngpu::ConnSpecS(); return 0; });
}
int NESTGPU_SetConnSpecParam(char* name, int value) {
  return guarded(0, [&] {
    std::string n = name;
    if (n == "rule") {
      if (value < 0 || value > 4) throw Error("Unknown connection rule");
      g_conn_spec.rule = value;
    } else if (n == "indegree") {
      if (value < 0) throw Error("Indegree must be >=0");
      g_conn_spec.indegree = value;
    } else if (n == "outdegree") {
      if (value < 0) throw Error("Outdegree must be >=0");
      g_conn_spec.outdegree = value;
    } else if (n == "total_num") {
      if (value < 0) throw Error("total_num must be >=0");
      g_conn_spec.total_num = value;
    } else {
      throw Error("Unknown connection int parameter");
    }
    return 0;
  });
}
int NESTGPU_ConnSpecIsParam(char* name) {
  return guarded(0, [&] {
    std::string n = name;
    return (int)(n == "rule" || n == "indegree" || n == "outdegree" || n == "total_num");
  });
}
int NESTGPU_SynSpecInit() {
  return guarded(0, [&] { E(); g_syn_spec = ngpu::SynSpecS(); return 0; });
}
int NESTGPU_SetSynSpecIntParam(char* name, int value) {
  return guarded(0, [&] {
    std::string n = name;
    if (n == "synapse_group") {
      if (value < 0) throw Error("Unknown synapse group");
      g_syn_spec.syn_group = value;
    } else if (n == "receptor") {
      if (value < 0) throw Error("Port index must be >=0");
      g_syn_spec.port = value;
    } else {
      throw Error("Unknown synapse int parameter");
    }
    return 0;
  });
}
int NESTGPU_SetSynSpecFloatParam(char* name, float value) {
  return guarded(0, [&] {
    std::string n = name;
    if (n == "weight") {
      g_syn_spec.weight = value;
    } else if (n == "delay") {
      if (value < 0) throw Error("Delay must be >=0");
      g_syn_spec.delay = value;
    } else {
      throw Error("Unknown synapse float parameter");
    }
    return 0;
  });
}
int NESTGPU_SetSynSpecFloatPtParam(char* name, float* arr) {
  return guarded(0, [&] {
    std::string n = name;
    if (n == "weight_array") g_syn_spec.weight_array = arr;
    else if (n == "delay_array") g_syn_spec.delay_array = arr;
    else throw Error("Unknown synapse array parameter");
    return 0;
  });
}
int NESTGPU_SynSpecIsIntParam(char* name) {
  return guarded(0, [&] {
    std::string n = name;
    return (int)(n == "synapse_group" || n == "receptor");
  });
}
int NESTGPU_SynSpecIsFloatParam(char* name) {
  return guarded(0, [&] {
    std::string n = name;
    return (int)(n == "weight" || n == "delay");
  });
}
int NESTGPU_SynSpecIsFloatPtParam(char* name) {
  return guarded(0, [&] {
    std::string n = name;
    return (int)(n == "weight_array" || n == "delay_array");
  });
}
int NESTGPU_ConnectSeqSeq(int i_source, int n_source, int i_target, int n_target) {
  return guarded(0, [&] {
    return E().ConnectRule(nullptr, i_source, n_source, nullptr, i_target, n_target,
                           g_conn_spec, g_syn_spec);
  });
}
int NESTGPU_ConnectSeqGroup(int i_source, int n_source, int* i_target, int n_target) {
  return guarded(0, [&] {
    return E().ConnectRule(nullptr, i_source, n_source, i_target, 0, n_target, g_conn_spec,
                           g_syn_spec);
  });
}
int NESTGPU_ConnectGroupSeq(int* i_source, int n_source, int i_target, int n_target) {
  return guarded(0, [&] {
    return E().ConnectRule(i_source, 0, n_source, nullptr, i_target, n_target, g_conn_spec,
                           g_syn_spec);
  });
}
int NESTGPU_ConnectGroupGroup(int* i_source, int n_source, int* i_target, int n_target) {
  return guarded(0, [&] {
    return E().ConnectRule(i_source, 0, n_source, i_target, 0, n_target, g_conn_spec,
                           g_syn_spec);
  });
}
int NESTGPU_RemoteConnectSeqSeq(int, int, int, int, int, int) {
  return guarded(0, [&]() -> int { out_of_scope("Remote (MPI) connection"); });
}
int NESTGPU_RemoteConnectSeqGroup(int, int, int, int, int*, int) {
  return guarded(0, [&]() -> int { out_of_scope("Remote (MPI) connection"); });
}
int NESTGPU_RemoteConnectGroupSeq(int, int*, int, int, int, int) {
  return guarded(0, [&]() -> int { out_of_scope("Remote (MPI) connection"); });
}
int NESTGPU_RemoteConnectGroupGroup(int, int*, int, int, int*, int) {
  return guarded(0, [&]() -> int { out_of_scope("Remote (MPI) connection"); });
}

char** NESTGPU_GetIntVarNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().IntVarNames(i_node)); });
}
char** NESTGPU_GetScalVarNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().ScalVarNames(i_node)); });
}
int NESTGPU_GetNIntVar(int i_node) {
  return guarded(0, [&] { return (int)E().IntVarNames(i_node).size(); });
}
int NESTGPU_GetNScalVar(int i_node) {
  return guarded(0, [&] { return (int)E().ScalVarNames(i_node).size(); });
}
char** NESTGPU_GetPortVarNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().PortVarNames(i_node)); });
}
int NESTGPU_GetNPortVar(int i_node) {
  return guarded(0, [&] { return (int)E().PortVarNames(i_node).size(); });
}
char** NESTGPU_GetScalParamNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().ScalParamNames(i_node)); });
}
int NESTGPU_GetNScalParam(int i_node) {
  return guarded(0, [&] { return (int)E().ScalParamNames(i_node).size(); });
}
char** NESTGPU_GetPortParamNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().PortParamNames(i_node)); });
}
int NESTGPU_GetNGroupParam(int i_node) {
  return guarded(0, [&] { return (int)E().GroupParamNames(i_node).size(); });
}
char** NESTGPU_GetGroupParamNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().GroupParamNames(i_node)); });
}
int NESTGPU_GetNPortParam(int i_node) {
  return guarded(0, [&] { return (int)E().PortParamNames(i_node).size(); });
}
char** NESTGPU_GetArrayParamNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().ArrayParamNames(i_node)); });
}
int NESTGPU_GetNArrayParam(int i_node) {
  return guarded(0, [&] { return (int)E().ArrayParamNames(i_node).size(); });
}
char** NESTGPU_GetArrayVarNames(int i_node) {
  return guarded((char**)nullptr, [&] { return name_array(E().ArrayVarNames(i_node)); });
}
int NESTGPU_GetNArrayVar(int i_node) {
  return guarded(0, [&] { return (int)E().ArrayVarNames(i_node).size(); });
}

int* NESTGPU_GetSeqSeqConnections(int i_source, int n_source, int i_target, int n_target,
                                  int syn_group, int* n_conn) {
  return guarded((int*)nullptr, [&] {
    return conn_array(E().GetConnections(nullptr, i_source, n_source, nullptr, i_target, n_target,
                                         syn_group),
                      n_conn);
  });
}
int* NESTGPU_GetSeqGroupConnections(int i_source, int n_source, int* i_target, int n_target,
                                    int syn_group, int* n_conn) {
  return guarded((int*)nullptr, [&] {
    return conn_array(E().GetConnections(nullptr, i_source, n_source, i_target, 0, n_target,
                                         syn_group),
                      n_conn);
  });
}
int* NESTGPU_GetGroupSeqConnections(int* i_source, int n_source, int i_target, int n_target,
                                    int syn_group, int* n_conn) {
  return guarded((int*)nullptr, [&] {
    return conn_array(E().GetConnections(i_source, 0, n_source, nullptr, i_target, n_target,
                                         syn_group),
                      n_conn);
  });
}
int* NESTGPU_GetGroupGroupConnections(int* i_source, int n_source, int* i_target, int n_target,
                                      int syn_group, int* n_conn) {
  return guarded((int*)nullptr, [&] {
    return conn_array(E().GetConnections(i_source, 0, n_source, i_target, 0, n_target, syn_group),
                      n_conn);
  });
}
int NESTGPU_GetConnectionStatus(int i_source, int i_group, int i_conn, int* i_target,
                                unsigned char* port, unsigned char* syn_group, float* delay,
                                float* weight) {
  return guarded(0, [&] {
    E().GetConnectionStatus(i_source, i_group, i_conn, i_target, port, syn_group, delay, weight);
    return 0;
  });
}

int NESTGPU_CreateSynGroup(char* model_name) {
  return guarded(0, [&] { return E().CreateSynGroup(model_name); });
}
int NESTGPU_GetSynGroupNParam(int i_syn_group) {
  return guarded(0, [&] { return E().GetSynGroupNParam(i_syn_group); });
}
char** NESTGPU_GetSynGroupParamNames(int i_syn_group) {
  return guarded((char**)nullptr, [&] { return name_array(E().GetSynGroupParamNames(i_syn_group)); });
}
int NESTGPU_IsSynGroupParam(int i_syn_group, char* param_name) {
  return guarded(0, [&] { return E().IsSynGroupParam(i_syn_group, param_name) ? 1 : 0; });
}
int NESTGPU_GetSynGroupParamIdx(int i_syn_group, char* param_name) {
  return guarded(0, [&] { return E().GetSynGroupParamIdx(i_syn_group, param_name); });
}
float NESTGPU_GetSynGroupParam(int i_syn_group, char* param_name) {
  return guarded(0.0f, [&] { return E().GetSynGroupParam(i_syn_group, param_name); });
}
int NESTGPU_SetSynGroupParam(int i_syn_group, char* param_name, float val) {
  return guarded(0, [&] {
    E().SetSynGroupParam(i_syn_group, param_name, val);
    return 0;
  });
}

int NESTGPU_ActivateSpikeCount(int i_node, int n_node) {
  return guarded(0, [&] { return E().ActivateSpikeCount(i_node, n_node); });
}
int NESTGPU_ActivateRecSpikeTimes(int i_node, int n_node, int max_n) {
  return guarded(0, [&] { return E().ActivateRecSpikeTimes(i_node, n_node, max_n); });
}
int NESTGPU_SetRecSpikeTimesStep(int i_node, int n_node, int step) {
  return guarded(0, [&] { return E().SetRecSpikeTimesStep(i_node, n_node, step); });
}
int NESTGPU_GetNRecSpikeTimes(int i_node) {
  return guarded(0, [&] { return E().GetNRecSpikeTimes(i_node); });
}
int NESTGPU_GetRecSpikeTimes(int i_node, int n_node, int** n_pt, float*** t_pt) {
  return guarded(0, [&] { return E().GetRecSpikeTimes(i_node, n_node, n_pt, t_pt); });
}

int NESTGPU_PushSpikesToNodes(int, int*) {
  return guarded(0, [&]() -> int { out_of_scope("External spike input (ext_neuron)"); });
}
int NESTGPU_GetExtNeuronInputSpikes(int*, int**, int**, float**, int) {
  return guarded(0, [&]() -> int { out_of_scope("External neurons (ext_neuron)"); });
}

int NESTGPU_SetNeuronGroupParam(int i_node, int n_node, char* name, float val) {
  return guarded(0, [&] { return E().SetGroupParam(i_node, n_node, name, val); });
}
int NESTGPU_IsNeuronGroupParam(int i_node, char* name) {
  return guarded(0, [&] { return (int)E().IsGroupParam(i_node, name); });
}
float NESTGPU_GetNeuronGroupParam(int i_node, char* name) {
  return guarded(0.0f, [&] { return E().GetGroupParam(i_node, name); });
}

int NESTGPU_GetNBoolParam() { return guarded(0, [&] { return (int)E().BoolParamNames().size(); }); }
char** NESTGPU_GetBoolParamNames() {
  return guarded((char**)nullptr, [&] { return name_array(E().BoolParamNames()); });
}
int NESTGPU_IsBoolParam(char* name) {
  return guarded(0, [&] {
    auto v = E().BoolParamNames();
    return (int)(std::find(v.begin(), v.end(), std::string(name)) != v.end());
  });
}
int NESTGPU_GetBoolParamIdx(char* name) {
  return guarded(0, [&] { return name_index(E().BoolParamNames(), name, "kernel bool"); });
}
bool NESTGPU_GetBoolParam(char* name) {
  return guarded(false, [&] { return E().GetBoolParam(name); });
}
int NESTGPU_SetBoolParam(char* name, bool val) {
  return guarded(0, [&] { return E().SetBoolParam(name, val); });
}
int NESTGPU_GetNFloatParam() { return guarded(0, [&] { return (int)E().FloatParamNames().size(); }); }
char** NESTGPU_GetFloatParamNames() {
  return guarded((char**)nullptr, [&] { return name_array(E().FloatParamNames()); });
}
int NESTGPU_IsFloatParam(char* name) {
  return guarded(0, [&] {
    auto v = E().FloatParamNames();
    return (int)(std::find(v.begin(), v.end(), std::string(name)) != v.end());
  });
}
int NESTGPU_GetFloatParamIdx(char* name) {
  return guarded(0, [&] { return name_index(E().FloatParamNames(), name, "kernel float"); });
}
float NESTGPU_GetFloatParam(char* name) {
  return guarded(0.0f, [&] { return E().GetFloatParam(name); });
}
int NESTGPU_SetFloatParam(char* name, float val) {
  return guarded(0, [&] { return E().SetFloatParam(name, val); });
}
int NESTGPU_GetNIntParam() { return guarded(0, [&] { return (int)E().IntParamNames().size(); }); }
char** NESTGPU_GetIntParamNames() {
  return guarded((char**)nullptr, [&] { return name_array(E().IntParamNames()); });
}
int NESTGPU_IsIntParam(char* name) {
  return guarded(0, [&] {
    auto v = E().IntParamNames();
    return (int)(std::find(v.begin(), v.end(), std::string(name)) != v.end());
  });
}
int NESTGPU_GetIntParamIdx(char* name) {
  return guarded(0, [&] { return name_index(E().IntParamNames(), name, "kernel int"); });
}
int NESTGPU_GetIntParam(char* name) {
  return guarded(0, [&] { return E().GetIntParam(name); });
}
int NESTGPU_SetIntParam(char* name, int val) {
  return guarded(0, [&] { return E().SetIntParam(name, val); });
}
int NESTGPU_RemoteCreate(int, char*, int, int) {
  return guarded(0, [&]() -> int { out_of_scope("Remote (MPI) node creation"); });
}

int NESTGPU_AMD_GetStepStats(unsigned long long* out) {
  return guarded(0, [&] {
    E().Stats(out);
    return 0;
  });
}
int NESTGPU_AMD_SetKernelTiming(int on) {
  return guarded(0, [&] {
    E().SetKernelTiming(on >= 2 && on <= 4 ? on : (on != 0 ? 1 : 0));
    return 0;
  });
}
int NESTGPU_AMD_GetKernelTimes(double* ms, long long* launches) {
  return guarded(0, [&] {
    E().KernelTimes(ms, launches);
    return 0;
  });
}
int NESTGPU_AMD_ResetKernel() {
  return guarded(0, [&] {
    delete g_engine;
    g_engine = nullptr;
    g_conn_spec = ngpu::ConnSpecS();
    g_syn_spec = ngpu::SynSpecS();
    return 0;
  });
}
int NESTGPU_AMD_Synchronize() {
  return guarded(0, [&] {
    E().Synchronize();
    return 0;
  });
}

int NESTGPU_AMD_GetOutDegrees(int i_node, int n_node, long long* out) {
  return guarded(0, [&] {
    E().OutDegrees(i_node, n_node, out);
    return 0;
  });
}
long long NESTGPU_AMD_CountIncoming(const int* targets, int n_target) {
  return guarded(0LL, [&] { return E().CountIncoming(targets, n_target); });
}
int NESTGPU_AMD_GetIncoming(const int* targets, int n_target, int* source, int* target,
                            unsigned char* port, float* weight, float* delay, long long* index) {
  return guarded(0, [&] {
    E().GetIncoming(targets, n_target, source, target, port, weight, delay, index);
    return 0;
  });
}

int NESTGPU_AMD_SetStepPredictorMode(int mode) {
  return guarded(0, [&] {
    ngpu::Predictor::get().SetMode(mode);
    return 0;
  });
}
int NESTGPU_AMD_StartPredictorLog(long long capacity, int stride) {
  return guarded(0, [&] {
    ngpu::Predictor::get().StartLog(capacity, stride);
    return 0;
  });
}
long long NESTGPU_AMD_GetPredictorLogCount() {
  return guarded(0LL, [&] { return ngpu::Predictor::get().LogCount(); });
}
int NESTGPU_AMD_GetPredictorLog(float* out, long long first, long long n) {
  return guarded(0, [&] {
    ngpu::Predictor::get().GetLog(out, first, n);
    return 0;
  });
}
int NESTGPU_AMD_AppendPredictorLog(const float* records, long long n) {
  return guarded(0, [&] {
    ngpu::Predictor::get().AppendLog(records, n);
    return 0;
  });
}
int NESTGPU_AMD_InitPredictor(unsigned long long seed) {
  return guarded(0, [&] {
    ngpu::Predictor::get().Init(seed);
    return 0;
  });
}
int NESTGPU_AMD_TrainPredictor(int steps, int batch, float lr, unsigned long long seed, float* loss) {
  return guarded(0, [&] {
    ngpu::Predictor::get().Train(steps, batch, lr, seed, loss);
    return 0;
  });
}
int NESTGPU_AMD_SetPredictorParams(const float* params, int n) {
  return guarded(0, [&] {
    ngpu::Predictor::get().SetParams(params, n);
    return 0;
  });
}
int NESTGPU_AMD_GetPredictorParams(float* params, int n) {
  return guarded(0, [&] {
    ngpu::Predictor::get().GetParams(params, n);
    return 0;
  });
}
int NESTGPU_AMD_PredictorInfer(const float* raw, long long n, float* out) {
  return guarded(0, [&] {
    ngpu::Predictor::get().Infer(raw, n, out);
    return 0;
  });
}
int NESTGPU_AMD_PredictorLossGrad(const float* raw, const float* target, const unsigned char* censored,
                                  int n, float* loss, float* grad) {
  return guarded(0, [&] {
    ngpu::Predictor::get().LossGrad(raw, target, censored, n, loss, grad);
    return 0;
  });
}
int NESTGPU_AMD_GetBatchConfig(int* cfg) {
  return guarded(0, [&] {
    E().GetBatchConfig(cfg);
    return 0;
  });
}
int NESTGPU_AMD_SetBatchConfig(int variant, int iaf_unroll, float rk5_weight, int graph,
                               int tiled) {
  return guarded(0, [&] {
    E().SetBatchConfig(variant, iaf_unroll, rk5_weight, graph, tiled);
    return 0;
  });
}
int NESTGPU_AMD_AutotuneBatch(int steps, double* ms_per_step, int* n_candidates) {
  return guarded(0, [&] {
    const double ms = E().AutotuneBatch(steps, n_candidates);
    if (ms_per_step) *ms_per_step = ms;
    return 0;
  });
}
int NESTGPU_AMD_PredictorLossGradNoAlloc(const float* raw, const float* target,
                                         const unsigned char* censored, int n, float* loss,
                                         float* grad, void* stream) {
  return guarded(0, [&] {
    ngpu::Predictor::get().LossGradNoAlloc(raw, target, censored, n, loss, grad,
                                           static_cast<hipStream_t>(stream));
    return 0;
  });
}

}  // extern "C"
