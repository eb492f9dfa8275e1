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
  // columns instead of log records (log and idx unused): sample slot s < n_log
  // is record s, the slots after n_log are padding
  const float* raw;         // [n_log][kPI]
  const float* target;      // [n_log]
  const unsigned char* cens;  // [n_log]
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

// Loss and gradient of a batch held in device columns, with the partial rows
// in a fixed device array of this library instead of allocated memory: no
// allocation, asynchronous on stream s.  n must be in [1, kPredScratchMaxN];
// calls that overlap on one GPU share the array and race.
constexpr int kPredScratchMaxBlocks = 256;
constexpr int kPredScratchMaxN = kPredScratchMaxBlocks * 64;
void pred_loss_grad_noalloc(const float* params, const float* raw, const float* target,
                            const unsigned char* cens, int n, float* loss, float* grad,
                            hipStream_t s);

}  // namespace ngpu
