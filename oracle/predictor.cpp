// TEST INFRASTRUCTURE ONLY: host reference of the learned step-size
// predictor (an MLP 8 -> 32 -> 32 -> 1 with ReLU), used by tests/ as the
// checker of the device forward, backward and inference kernels.  It has no
// counterpart in the reference simulator: the predictor replaces the
// proposal of the RK5 step-size controller (src/rk5.h:135-153) while the
// embedded error estimate keeps accepting or rejecting every step.
//
// Parameter vector (floats), the layout the library's C ABI uses:
//   mu[8], inv_sigma[8]          input normalisation (not trained)
//   W1[32][8], b1[32]            hidden layer 1
//   W2[32][32], b2[32]           hidden layer 2
//   W3[32], b3                   output
// Output: log2(h / dt).  Loss per sample with target t and censoring flag c
// (c = 1: the logged step was cut by the end of the interval, so only a
// prediction below t is an error): e = out - t; L = (c && e > 0) ? 0 : e^2;
// the batch loss is the mean.
//
// quant = 1 evaluates what the device computes: normalised inputs, W1, the
// layer-1 activations and W2 rounded to bfloat16 (round to nearest even),
// FP32/FP64 accumulation, and gradients passed straight through the
// rounding.  quant = 0 is the plain real-valued network (finite-difference
// checks).  Everything here is in double.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {
constexpr int NI = 8, NH = 32;
constexpr int O_MU = 0, O_IS = 8, O_W1 = 16, O_B1 = O_W1 + NH * NI, O_W2 = O_B1 + NH,
              O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH, O_B3 = O_W3 + NH, N_PAR = O_B3 + 1;

double bf16(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return v;
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

struct Fwd {
  double x[NI], z1[NH], h1[NH], z2[NH], h2[NH], out;
};

void forward(const double* P, const float* raw, int quant, Fwd& f) {
  auto q = [&](double v) { return quant ? bf16(v) : v; };
  for (int i = 0; i < NI; ++i) f.x[i] = q(((double)raw[i] - P[O_MU + i]) * P[O_IS + i]);
  for (int o = 0; o < NH; ++o) {
    double s = P[O_B1 + o];
    for (int i = 0; i < NI; ++i) s += q(P[O_W1 + o * NI + i]) * f.x[i];
    f.z1[o] = s;
    f.h1[o] = q(s > 0 ? s : 0.0);
  }
  for (int o = 0; o < NH; ++o) {
    double s = P[O_B2 + o];
    for (int i = 0; i < NH; ++i) s += q(P[O_W2 + o * NH + i]) * f.h1[i];
    f.z2[o] = s;
    f.h2[o] = s > 0 ? s : 0.0;
  }
  double s = P[O_B3];
  for (int i = 0; i < NH; ++i) s += P[O_W3 + i] * f.h2[i];
  f.out = s;
}
}  // namespace

extern "C" {

int orc_mlp_n_params() { return N_PAR; }

void orc_mlp_forward(const double* P, const float* raw, long long n, int quant, double* out) {
#pragma omp parallel for schedule(static)
  for (long long s = 0; s < n; ++s) {
    Fwd f;
    forward(P, raw + s * NI, quant, f);
    out[s] = f.out;
  }
}

// Mean loss over the batch and its gradient with respect to every parameter
// (mu and inv_sigma get zero gradient).
void orc_mlp_loss_grad(const double* P, const float* raw, const float* target,
                       const unsigned char* cens, long long n, int quant, double* loss,
                       double* grad) {
  auto q = [&](double v) { return quant ? bf16(v) : v; };
  for (int k = 0; k < N_PAR; ++k) grad[k] = 0.0;
  double L = 0.0;
  for (long long s = 0; s < n; ++s) {
    Fwd f;
    forward(P, raw + s * NI, quant, f);
    const double e = f.out - (double)target[s];
    double dout;
    if (cens[s] && e > 0) {
      dout = 0.0;
    } else {
      L += e * e;
      dout = 2.0 * e / (double)n;
    }
    if (dout == 0.0) continue;
    grad[O_B3] += dout;
    double dz2[NH];
    for (int i = 0; i < NH; ++i) {
      grad[O_W3 + i] += dout * f.h2[i];
      dz2[i] = f.z2[i] > 0 ? dout * P[O_W3 + i] : 0.0;
    }
    double dh1[NH] = {0};
    for (int o = 0; o < NH; ++o) {
      if (dz2[o] == 0.0) continue;
      grad[O_B2 + o] += dz2[o];
      for (int i = 0; i < NH; ++i) {
        grad[O_W2 + o * NH + i] += dz2[o] * f.h1[i];
        dh1[i] += dz2[o] * q(P[O_W2 + o * NH + i]);
      }
    }
    for (int o = 0; o < NH; ++o) {
      const double dz1 = f.z1[o] > 0 ? dh1[o] : 0.0;
      if (dz1 == 0.0) continue;
      grad[O_B1 + o] += dz1;
      for (int i = 0; i < NI; ++i) grad[O_W1 + o * NI + i] += dz1 * f.x[i];
    }
  }
  *loss = L / (double)n;
}

}  // extern "C"
