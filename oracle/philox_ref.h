// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.
//
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and the Poisson
// sampler used for poisson_generator input.  The reference draws its Poisson
// counts with cuRAND (`curand_poisson`, src/poiss_gen.cu:74), a closed NVIDIA
// library that is not restated; both this oracle and the HIP build use the
// published Philox construction instead, so GPU and CPU see identical draws.
//
// Sampler: Knuth's multiplication method (TAOCP vol. 2, 3.4.1) evaluated in
// IEEE double with uniforms (k + 0.5) * 2^-32, applied to `m` equal chunks of
// the mean (the sum of independent Poisson variates is Poisson).  The
// threshold exp(-lambda/m) is precomputed on the host, so the per-draw
// arithmetic is a chain of correctly rounded double products and compares,
// reproducible bit for bit on any IEEE machine.
#pragma once
#include <cstdint>
#include <cmath>

namespace orc {

struct Philox4 { uint32_t v[4]; };

inline Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * ctr.v[0];
    uint64_t p1 = (uint64_t)M1 * ctr.v[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    Philox4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Number of equal chunks the mean is split into (keeps exp(-chunk) well
// inside the double range and the product chain short).
inline int poisson_chunks(double lambda) {
  const double kChunk = 16.0;
  if (!(lambda > 0.0)) return 0;
  return (int)std::ceil(lambda / kChunk);
}

// One Poisson(lambda) draw for stream (conn_id) at time index t_idx.
// exp_chunk = exp(-lambda / chunks), computed by the caller on the host.
inline int poisson_draw(uint64_t seed, uint32_t conn_id, int64_t t_idx,
                        int chunks, double exp_chunk) {
  if (chunks <= 0) return 0;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  Philox4 ctr;
  ctr.v[0] = 0;
  ctr.v[1] = (uint32_t)(uint64_t)t_idx;
  ctr.v[2] = (uint32_t)((uint64_t)t_idx >> 32);
  ctr.v[3] = conn_id;
  Philox4 r = philox4x32_10(ctr, k0, k1);
  int used = 0;
  int n = 0;
  for (int c = 0; c < chunks; ++c) {
    double p = 1.0;
    for (;;) {
      if (used == 4) {
        ctr.v[0] += 1;
        r = philox4x32_10(ctr, k0, k1);
        used = 0;
      }
      double u = ((double)r.v[used++] + 0.5) * (1.0 / 4294967296.0);
      p *= u;
      if (p <= exp_chunk) break;
      ++n;
    }
  }
  return n;
}

}  // namespace orc
