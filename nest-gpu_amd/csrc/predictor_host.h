// Host side of the learned step-size predictor (predictor.h): parameters,
// Adam state, the (state, accepted h) log written by the RK5 kernel, and the
// training loop.  One process-wide instance that outlives the simulator
// (NESTGPU_AMD_ResetKernel), so logs of several runs train one network.
// No counterpart in the reference.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"
#include "predictor.h"

namespace ngpu {

class Predictor {
 public:
  static Predictor& get();

  int mode() const { return mode_; }
  void SetMode(int mode);                          // 0 controller, 1 log, 2 learned
  void StartLog(long long capacity, int stride);   // (re)allocates and clears the log
  long long LogCount();                            // records held (<= capacity)
  void GetLog(float* out, long long first, long long n);
  void AppendLog(const float* recs, long long n);
  void SetParams(const float* p, int n);
  void GetParams(float* p, int n);
  void Init(unsigned long long seed);              // normalisation from the log, random weights
  void Train(int steps, int batch, float lr, unsigned long long seed, float* loss);
  void Infer(const float* raw, long long n, float* out);
  void LossGrad(const float* raw, const float* target, const unsigned char* cens, int n,
                float* loss, float* grad);
  // LossGrad on device buffers, asynchronous on `stream`, allocating nothing
  void LossGradNoAlloc(const float* raw, const float* target, const unsigned char* cens, int n,
                       float* loss, float* grad, hipStream_t stream);

  // buffers read by the RK5 kernel
  const float* d_params() const { return params_.p; }
  float* d_log() { return log_.p; }
  unsigned long long* d_log_count() { return log_count_.p; }
  long long log_cap() const { return log_cap_; }
  int log_stride() const { return log_stride_; }

 private:
  Predictor() = default;
  void ensure_stream();
  void ensure_params();
  int mode_ = 0;
  bool have_params_ = false;
  DBuf<float> params_, m_, v_, gpart_, lpart_, loss_, grad_;
  int adam_t_ = 0;
  DBuf<float> log_;
  DBuf<unsigned long long> log_count_;
  long long log_cap_ = 0;
  int log_stride_ = 1;
  hipStream_t stream_ = nullptr;
};

}  // namespace ngpu
