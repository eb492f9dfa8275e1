// Philox4x32-10 (Salmon et al., SC'11) and the Poisson sampler used for
// poisson_generator input, usable on host and device.
//
// The reference draws Poisson counts with cuRAND (src/poiss_gen.cu:74);
// this build uses the published Philox construction with a Knuth
// multiplication sampler (TAOCP 3.4.1) evaluated as a chain of IEEE double
// products, so every draw is reproducible bit for bit off the GPU.
//   stream  = index of the connection in the source-major connection table
//   counter = {block, t_idx lo, t_idx hi, stream}, key = 64-bit kernel seed
// The mean is split in `chunks` equal parts with threshold exp(-lambda/chunks)
// computed on the host, keeping the products inside the double range.
#pragma once
#include <cstdint>

#ifndef __HIPCC__
#define NGPU_HD
#else
#define NGPU_HD __host__ __device__
#endif

namespace ngpu {

struct U32x4 { uint32_t x, y, z, w; };

NGPU_HD inline U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
#ifdef __clang__
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

NGPU_HD inline uint32_t pick(const U32x4& r, int i) {
  return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}

// Chunk count for a Poisson mean (16 per chunk).
inline int poisson_chunks_host(double lambda) {
  if (!(lambda > 0.0)) return 0;
  int m = (int)(lambda / 16.0);
  if ((double)m * 16.0 < lambda) ++m;
  return m;
}

NGPU_HD inline int poisson_draw(uint64_t seed, uint32_t stream, long long t_idx,
                                int chunks, double exp_chunk) {
  if (chunks <= 0) return 0;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  U32x4 ctr{0u, (uint32_t)(uint64_t)t_idx, (uint32_t)((uint64_t)t_idx >> 32), stream};
  U32x4 r = philox4x32_10(ctr, k0, k1);
  int used = 0, n = 0;
  for (int c = 0; c < chunks; ++c) {
    double p = 1.0;
    for (;;) {
      if (used == 4) {
        ctr.x += 1u;
        r = philox4x32_10(ctr, k0, k1);
        used = 0;
      }
      double u = ((double)pick(r, used++) + 0.5) * (1.0 / 4294967296.0);
      p *= u;
      if (p <= exp_chunk) break;
      ++n;
    }
  }
  return n;
}

}  // namespace ngpu
