// Host-side launchers of the predictor training and test kernels
// (predictor.hip).  All launches are asynchronous on the given stream.
#pragma once
#include <hip/hip_runtime.h>

namespace ngpu {

struct PredGradArgs {
  const float* params;      // [kPNPar]
  const float* log;         // [n_log][kLogCols]
  long long n_log;
  const long long* idx;     // explicit sample indexes (-1: padding) or null: random draws
  long long n_idx;
  unsigned long long seed;  // random draws: Philox(seed; slot, step) mod n_log
  long long step;
  int rounds;               // 64-sample rounds per block
  int batch;                // samples of the mean (divisor of the loss)
  float* gpart;             // [blocks][kPNTrain] partial gradient rows
  float* lpart;             // [blocks] partial loss sums
};

struct PredUpdateArgs {
  float* params;
  float* m;
  float* v;
  const float* gpart;
  const float* lpart;
  int blocks;
  int batch;
  float lr;                 // 0: only sum the rows (gradient check)
  float beta1, beta2, eps;
  float bc1, bc2;           // bias corrections 1 - beta^t
  float* grad_out;          // [kPNTrain] summed gradient or null
  float* loss_out;          // [1] mean loss of the batch
};

void launch_pred_stats(const float* log, long long n, double* part, int grid, hipStream_t s);
void launch_pred_grad(const PredGradArgs& a, int blocks, hipStream_t s);
void launch_pred_update(const PredUpdateArgs& a, hipStream_t s);
void launch_pred_infer(const float* P, const float* raw, long long n, float* out, hipStream_t s);

}  // namespace ngpu
