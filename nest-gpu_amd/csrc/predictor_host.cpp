// Host side of the learned step-size predictor (see predictor_host.h).
#include "predictor_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "philox.h"
#include "predictor_kernels.h"

namespace ngpu {

#define PCHK(call)                                                            \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw Error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " \
                  + #call);                                                   \
  } while (0)

Predictor& Predictor::get() {
  // never destroyed: device buffers stay valid until the process ends
  static Predictor* inst = new Predictor();
  return *inst;
}

void Predictor::ensure_stream() {
  if (!stream_) PCHK(hipStreamCreate(&stream_));
  // order after any simulation work queued on other streams
  PCHK(hipDeviceSynchronize());
}

void Predictor::ensure_params() {
  if (params_.p) return;
  params_.alloc(kPNPar);
  m_.alloc(kPNTrain);
  v_.alloc(kPNTrain);
  grad_.alloc(kPNTrain);
  ensure_stream();
  PCHK(hipMemsetAsync(params_.p, 0, kPNPar * sizeof(float), stream_));
  PCHK(hipMemsetAsync(m_.p, 0, kPNTrain * sizeof(float), stream_));
  PCHK(hipMemsetAsync(v_.p, 0, kPNTrain * sizeof(float), stream_));
  PCHK(hipStreamSynchronize(stream_));
}

void Predictor::SetMode(int mode) {
  if (mode < 0 || mode > 2) throw Error("Step predictor mode must be 0, 1 or 2");
  if (mode == 1 && !log_.p) throw Error("Step predictor log not started");
  if (mode == 2 && !have_params_) throw Error("Step predictor has no parameters");
  mode_ = mode;
}

void Predictor::StartLog(long long capacity, int stride) {
  if (capacity <= 0 || stride <= 0) throw Error("Step predictor log: capacity and stride must be > 0");
  ensure_stream();
  log_.alloc((size_t)capacity * kLogCols);
  log_count_.alloc(1);
  PCHK(hipMemsetAsync(log_count_.p, 0, sizeof(unsigned long long), stream_));
  PCHK(hipStreamSynchronize(stream_));
  log_cap_ = capacity;
  log_stride_ = stride;
}

long long Predictor::LogCount() {
  if (!log_count_.p) return 0;
  ensure_stream();
  unsigned long long c = 0;
  PCHK(hipMemcpy(&c, log_count_.p, sizeof(c), hipMemcpyDeviceToHost));
  return std::min<long long>((long long)c, log_cap_);
}

void Predictor::GetLog(float* out, long long first, long long n) {
  const long long have = LogCount();
  if (first < 0 || n < 0 || first + n > have) throw Error("Step predictor log: range out of bounds");
  if (n) PCHK(hipMemcpy(out, log_.p + first * kLogCols, (size_t)n * kLogCols * sizeof(float),
                        hipMemcpyDeviceToHost));
}

void Predictor::AppendLog(const float* recs, long long n) {
  if (!log_.p) throw Error("Step predictor log not started");
  const long long have = LogCount();
  if (n < 0 || have + n > log_cap_) throw Error("Step predictor log: capacity exceeded");
  if (n) PCHK(hipMemcpy(log_.p + have * kLogCols, recs, (size_t)n * kLogCols * sizeof(float),
                        hipMemcpyHostToDevice));
  const unsigned long long c = (unsigned long long)(have + n);
  PCHK(hipMemcpy(log_count_.p, &c, sizeof(c), hipMemcpyHostToDevice));
}

void Predictor::SetParams(const float* p, int n) {
  if (n != kPNPar) throw Error("Step predictor: parameter vector must have " + std::to_string(kPNPar) + " entries");
  ensure_stream();
  ensure_params();
  PCHK(hipMemcpy(params_.p, p, kPNPar * sizeof(float), hipMemcpyHostToDevice));
  PCHK(hipMemsetAsync(m_.p, 0, kPNTrain * sizeof(float), stream_));
  PCHK(hipMemsetAsync(v_.p, 0, kPNTrain * sizeof(float), stream_));
  PCHK(hipStreamSynchronize(stream_));
  adam_t_ = 0;
  have_params_ = true;
}

void Predictor::GetParams(float* p, int n) {
  if (n != kPNPar) throw Error("Step predictor: parameter vector must have " + std::to_string(kPNPar) + " entries");
  if (!have_params_) throw Error("Step predictor has no parameters");
  ensure_stream();
  PCHK(hipMemcpy(p, params_.p, kPNPar * sizeof(float), hipMemcpyDeviceToHost));
}

void Predictor::Init(unsigned long long seed) {
  const long long n = LogCount();
  if (n < 64) throw Error("Step predictor: the log holds fewer than 64 records");
  ensure_params();
  // mean and spread of the raw inputs and of the target over the log
  const int grid = 256;
  DBuf<double> part;
  part.alloc((size_t)grid * 2 * kPI);
  launch_pred_stats(log_.p, n, part.p, grid, stream_);
  PCHK(hipGetLastError());
  std::vector<double> hp((size_t)grid * 2 * kPI);
  PCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  PCHK(hipStreamSynchronize(stream_));
  std::vector<float> P(kPNPar, 0.0f);
  for (int k = 0; k < kPI; ++k) {
    double s = 0, q = 0;
    for (int b = 0; b < grid; ++b) {
      s += hp[(size_t)b * 2 * kPI + k];
      q += hp[(size_t)b * 2 * kPI + kPI + k];
    }
    const double mu = s / (double)n;
    const double var = std::max(q / (double)n - mu * mu, 1.0e-12);
    P[kPOffMu + k] = (float)mu;
    P[kPOffIs + k] = (float)(1.0 / std::sqrt(var));
  }
  // target mean from a host sample of the log (output bias)
  const long long ns = std::min<long long>(n, 65536);
  std::vector<float> smp((size_t)ns * kLogCols);
  PCHK(hipMemcpy(smp.data(), log_.p, smp.size() * sizeof(float), hipMemcpyDeviceToHost));
  double tm = 0;
  for (long long r = 0; r < ns; ++r) tm += smp[(size_t)r * kLogCols + kPI];
  // He-normal weights from Philox + Box-Muller
  long long ctr = 0;
  auto normal = [&]() {
    const U32x4 c = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0x5eedu, 0u};
    ++ctr;
    const U32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = ((double)r.x + 0.5) * 0x1p-32, u2 = ((double)r.y + 0.5) * 0x1p-32;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int k = 0; k < kPH * kPI; ++k) P[kPOffW1 + k] = (float)(normal() * std::sqrt(2.0 / kPI));
  for (int k = 0; k < kPH * kPH; ++k) P[kPOffW2 + k] = (float)(normal() * std::sqrt(2.0 / kPH));
  for (int k = 0; k < kPH; ++k) P[kPOffW3 + k] = (float)(normal() * std::sqrt(1.0 / kPH) * 0.1);
  P[kPOffB3] = (float)(tm / (double)ns);
  SetParams(P.data(), kPNPar);
}

void Predictor::Train(int steps, int batch, float lr, unsigned long long seed, float* loss) {
  if (!have_params_) throw Error("Step predictor has no parameters (InitPredictor first)");
  const long long n = LogCount();
  if (n < 1) throw Error("Step predictor: the log is empty");
  if (steps <= 0) return;
  if (batch < 64 || batch % 64) throw Error("Step predictor: batch must be a positive multiple of 64");
  int blocks = std::min(batch / 64, 2048);
  while ((batch / 64) % blocks) --blocks;
  const int rounds = batch / 64 / blocks;
  ensure_stream();
  gpart_.alloc((size_t)blocks * kPNTrain);
  lpart_.alloc(blocks);
  loss_.alloc(steps);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int st = 0; st < steps; ++st) {
    PredGradArgs ga{};
    ga.params = params_.p;
    ga.log = log_.p;
    ga.n_log = n;
    ga.seed = seed;
    ga.step = adam_t_;
    ga.rounds = rounds;
    ga.batch = batch;
    ga.gpart = gpart_.p;
    ga.lpart = lpart_.p;
    launch_pred_grad(ga, blocks, stream_);
    ++adam_t_;
    PredUpdateArgs ua{};
    ua.params = params_.p;
    ua.m = m_.p;
    ua.v = v_.p;
    ua.gpart = gpart_.p;
    ua.lpart = lpart_.p;
    ua.blocks = blocks;
    ua.batch = batch;
    ua.lr = lr;
    ua.beta1 = b1;
    ua.beta2 = b2;
    ua.eps = 1.0e-8f;
    ua.bc1 = 1.0f - (float)std::pow((double)b1, (double)adam_t_);
    ua.bc2 = 1.0f - (float)std::pow((double)b2, (double)adam_t_);
    ua.loss_out = loss_.p + st;
    launch_pred_update(ua, stream_);
  }
  PCHK(hipGetLastError());
  if (loss) PCHK(hipMemcpyAsync(loss, loss_.p, steps * sizeof(float), hipMemcpyDeviceToHost, stream_));
  PCHK(hipStreamSynchronize(stream_));
}

void Predictor::Infer(const float* raw, long long n, float* out) {
  if (!have_params_) throw Error("Step predictor has no parameters");
  if (n <= 0) return;
  ensure_stream();
  DBuf<float> d_in, d_out;
  d_in.alloc((size_t)n * kPI);
  d_out.alloc(n);
  PCHK(hipMemcpyAsync(d_in.p, raw, (size_t)n * kPI * sizeof(float), hipMemcpyHostToDevice, stream_));
  launch_pred_infer(params_.p, d_in.p, n, d_out.p, stream_);
  PCHK(hipGetLastError());
  PCHK(hipMemcpyAsync(out, d_out.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  PCHK(hipStreamSynchronize(stream_));
}

void Predictor::LossGrad(const float* raw, const float* target, const unsigned char* cens, int n,
                         float* loss, float* grad) {
  if (!have_params_) throw Error("Step predictor has no parameters");
  if (n <= 0) throw Error("Step predictor: empty batch");
  ensure_stream();
  std::vector<float> recs((size_t)n * kLogCols);
  for (int r = 0; r < n; ++r) {
    for (int k = 0; k < kPI; ++k) recs[(size_t)r * kLogCols + k] = raw[(size_t)r * kPI + k];
    recs[(size_t)r * kLogCols + kPI] = target[r];
    recs[(size_t)r * kLogCols + kPI + 1] = cens[r] ? 1.0f : 0.0f;
  }
  const int blocks = (n + 63) / 64;
  std::vector<long long> idx((size_t)blocks * 64, -1);
  for (int r = 0; r < n; ++r) idx[r] = r;
  DBuf<float> d_rec, d_gp, d_lp, d_loss;
  DBuf<long long> d_idx;
  d_rec.alloc(recs.size());
  d_idx.alloc(idx.size());
  d_gp.alloc((size_t)blocks * kPNTrain);
  d_lp.alloc(blocks);
  d_loss.alloc(1);
  PCHK(hipMemcpyAsync(d_rec.p, recs.data(), recs.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
  PCHK(hipMemcpyAsync(d_idx.p, idx.data(), idx.size() * sizeof(long long), hipMemcpyHostToDevice, stream_));
  PredGradArgs ga{};
  ga.params = params_.p;
  ga.log = d_rec.p;
  ga.n_log = n;
  ga.idx = d_idx.p;
  ga.n_idx = (long long)idx.size();
  ga.rounds = 1;
  ga.batch = n;
  ga.gpart = d_gp.p;
  ga.lpart = d_lp.p;
  launch_pred_grad(ga, blocks, stream_);
  PredUpdateArgs ua{};
  ua.params = params_.p;
  ua.gpart = d_gp.p;
  ua.lpart = d_lp.p;
  ua.blocks = blocks;
  ua.batch = n;
  ua.lr = 0.0f;
  ua.grad_out = grad_.p;
  ua.loss_out = d_loss.p;
  launch_pred_update(ua, stream_);
  PCHK(hipGetLastError());
  PCHK(hipMemcpyAsync(grad, grad_.p, kPNTrain * sizeof(float), hipMemcpyDeviceToHost, stream_));
  PCHK(hipMemcpyAsync(loss, d_loss.p, sizeof(float), hipMemcpyDeviceToHost, stream_));
  PCHK(hipStreamSynchronize(stream_));
}

void Predictor::LossGradNoAlloc(const float* raw, const float* target, const unsigned char* cens,
                                int n, float* loss, float* grad, hipStream_t stream) {
  if (!have_params_) throw Error("Step predictor has no parameters");
  if (!raw || !target || !cens || !loss || !grad) throw Error("Step predictor: null buffer");
  if (n < 1 || n > kPredScratchMaxN)
    throw Error("Step predictor: batch of " + std::to_string(n) +
                " records outside [1, " + std::to_string(kPredScratchMaxN) +
                "] for the loss/gradient without allocation");
  try {
    pred_loss_grad_noalloc(params_.p, raw, target, cens, n, loss, grad, stream);
  } catch (const std::exception& e) {
    throw Error(e.what());
  }
}

}  // namespace ngpu
