// Device-side connectivity: the random connection rules' draws, and the
// connection store flattening -- the stable order of the
// connections by (source, delay slot) that the synapse table needs (the
// reference flattens its per-source connection lists on the host,
// src/spike_buffer.cu:399-437, after NetConnection::Insert kept each
// source's list sorted by delay, src/connect.cu:56-74).  Keys
// (source << 16 | d_int) with the connection's index as value go through
// rocPRIM's LSD radix sort, which is stable: connections with equal keys
// keep their creation order, exactly as the host counting sort by source
// followed by a stable sort by delay inside each source segment.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "connect_device.h"
#include "philox.h"

namespace ngpu {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("device connection sort: ") + what + ": " + hipGetErrorString(e));
}

__global__ void sort_keys_kernel(const int* src, const int* d, size_t n, unsigned long long* keys,
                                 unsigned int* vals, unsigned long long* counts) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const unsigned int s = (unsigned int)src[i];
    keys[i] = ((unsigned long long)s << 16) | (unsigned int)d[i];
    vals[i] = (unsigned int)i;
    atomicAdd(&counts[s], 1ull);
  }
}

struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) {
      (void)hipDeviceSynchronize();   // an error path may leave work queued
      (void)hipFree(p);
    }
  }
  template <class T>
  T* alloc(size_t n, const char* what) {
    check(hipMalloc(&p, n * sizeof(T) + 1), what);
    return static_cast<T*>(p);
  }
  // only after the work that uses the buffer has completed
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

// fixed_indegree / fixed_outdegree: deg distinct pool indices per owner,
// drawn exactly as the host rule does (Engine::ConnectRule): the owner's
// Philox stream {block, owner, 'rule', 0} under the call's key, 32-bit words
// taken in order; a partial Fisher-Yates over the pool when the pool is
// smaller than 5 * deg, else rejection of indices already taken.  Each
// thread owns one scratch slot (a pool-sized permutation or bit set) and
// takes the owners o = thread, thread + threads, ...; a bit set is cleared
// again after each owner.
__global__ void pick_distinct_kernel(int n_own, int deg, int n_pool, uint32_t k0, uint32_t k1,
                                     int* out, unsigned int* scratch, size_t slot_words,
                                     int fisher_yates) {
  const int threads = gridDim.x * blockDim.x;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int* slot = scratch + (size_t)t * slot_words;
  for (int o = t; o < n_own; o += threads) {
    uint32_t blk = 0;
    U32x4 r{0u, 0u, 0u, 0u};
    int avail = 0;
    auto next = [&]() -> uint32_t {
      if (avail == 0) {
        r = philox4x32_10(U32x4{blk++, (uint32_t)o, 0x72756c65u, 0u}, k0, k1);
        avail = 4;
      }
      return pick(r, 4 - avail--);
    };
    int* dst = out + (size_t)o * deg;
    if (fisher_yates) {
      int* perm = reinterpret_cast<int*>(slot);
      for (int i = 0; i < n_pool; ++i) perm[i] = i;
      for (int i = 0; i < deg; ++i) {
        const int j = i + (int)(next() % (uint32_t)(n_pool - i));
        const int tmp = perm[i];
        perm[i] = perm[j];
        perm[j] = tmp;
        dst[i] = perm[i];
      }
    } else {
      for (int i = 0; i < deg; ++i) {
        int j;
        for (;;) {
          j = (int)(next() % (uint32_t)n_pool);
          const unsigned int bit = 1u << (j & 31);
          if (!(slot[j >> 5] & bit)) {
            slot[j >> 5] |= bit;
            break;
          }
        }
        dst[i] = j;
      }
      for (int i = 0; i < deg; ++i) slot[dst[i] >> 5] = 0u;
    }
  }
}

// The host generator's words (Engine::random_block): block b of the
// sequence is Philox {ctr0 + b (64-bit), 'host', 0} under the kernel seed.
__global__ void random_words_kernel(uint64_t ctr0, uint32_t k0, uint32_t k1, size_t n,
                                    unsigned int* out) {
  const size_t n_blk = (n + 3) / 4;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_blk; b += stride) {
    const uint64_t c = ctr0 + b;
    const U32x4 r = philox4x32_10(U32x4{(uint32_t)c, (uint32_t)(c >> 32), 0x686f7374u, 0u}, k0, k1);
    const size_t i = 4 * b;
    out[i] = r.x;
    if (i + 1 < n) out[i + 1] = r.y;
    if (i + 2 < n) out[i + 2] = r.z;
    if (i + 3 < n) out[i + 3] = r.w;
  }
}

}  // namespace

void device_pick_distinct(int n_own, int deg, int n_pool, uint64_t key, int* out,
                          hipStream_t stream) {
  if (n_own <= 0 || deg <= 0) return;
  if (deg > n_pool) throw std::runtime_error("device connection rule: degree larger than the pool");
  const int fy = n_pool < 5 * deg ? 1 : 0;
  const size_t slot_words = fy ? (size_t)n_pool : ((size_t)n_pool + 31) / 32;
  // scratch: at most 2 GiB of slots, at most 64K threads, whole blocks of 64
  size_t threads = ((size_t)2 << 30) / (slot_words * 4 + 4);
  threads = std::min<size_t>(threads, 65536);
  threads = std::min<size_t>(threads, ((size_t)n_own + 63) / 64 * 64);
  threads = threads / 64 * 64;
  if (threads == 0) threads = 64;
  const size_t n_out = (size_t)n_own * deg;
  DevMem m_out, m_scr;
  int* d_out = m_out.alloc<int>(n_out, "rule output");
  unsigned int* scr = m_scr.alloc<unsigned int>(threads * slot_words, "rule scratch");
  if (!fy) check(hipMemsetAsync(scr, 0, threads * slot_words * 4, stream), "clear");
  hipLaunchKernelGGL(pick_distinct_kernel, dim3((unsigned)(threads / 64)), dim3(64), 0, stream, n_own, deg,
                     n_pool, (uint32_t)key, (uint32_t)(key >> 32), d_out, scr, slot_words, fy);
  check(hipGetLastError(), "rule kernel");
  check(hipMemcpyAsync(out, d_out, n_out * sizeof(int), hipMemcpyDeviceToHost, stream), "download");
  check(hipStreamSynchronize(stream), "sync");
}

void device_random_words(uint64_t ctr0, uint64_t seed, size_t n, unsigned int* out,
                         hipStream_t stream) {
  if (n == 0) return;
  DevMem m_out;
  unsigned int* d_out = m_out.alloc<unsigned int>(n, "random words");
  hipLaunchKernelGGL(random_words_kernel, dim3(4096), dim3(256), 0, stream, ctr0, (uint32_t)seed,
                     (uint32_t)(seed >> 32), n, d_out);
  check(hipGetLastError(), "random words kernel");
  check(hipMemcpyAsync(out, d_out, n * sizeof(unsigned int), hipMemcpyDeviceToHost, stream), "download");
  check(hipStreamSynchronize(stream), "sync");
}

__global__ void gather_kernel(const unsigned int* perm, size_t n, const int* tgt, const unsigned char* port,
                              const float* w, const int* d, int* tgt_o, unsigned char* port_o, float* w_o,
                              int* d_o) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const unsigned int i = perm[k];
    tgt_o[k] = tgt[i];
    port_o[k] = port[i];
    w_o[k] = w[i];
    d_o[k] = d[i];
  }
}

bool device_sort_connections(int* src, int* tgt, unsigned char* port, float* w, int* d, size_t n,
                             int n_nodes, std::vector<long long>& row_ptr, hipStream_t stream) {
  if (n == 0 || n >= (1ull << 32) || n_nodes <= 0) return false;
  int end_bit = 16;
  while (end_bit < 64 && (1ull << (end_bit - 16)) < (unsigned long long)n_nodes) ++end_bit;
  std::vector<unsigned long long> counts((size_t)n_nodes);
  {
    DevMem m_src, m_d, m_k0, m_k1, m_v0, m_v1, m_cnt, m_tmp, m_tgt, m_port, m_w, m_to, m_po, m_wo;
    int* d_src = m_src.alloc<int>(n, "source upload buffer");
    int* d_d = m_d.alloc<int>(n, "delay upload buffer");
    unsigned long long* k0 = m_k0.alloc<unsigned long long>(n, "keys");
    unsigned long long* k1 = m_k1.alloc<unsigned long long>(n, "keys");
    unsigned int* v0 = m_v0.alloc<unsigned int>(n, "indices");
    unsigned int* v1 = m_v1.alloc<unsigned int>(n, "indices");
    unsigned long long* cnt = m_cnt.alloc<unsigned long long>((size_t)n_nodes, "counts");
    check(hipMemcpyAsync(d_src, src, n * sizeof(int), hipMemcpyHostToDevice, stream), "upload");
    check(hipMemcpyAsync(d_d, d, n * sizeof(int), hipMemcpyHostToDevice, stream), "upload");
    check(hipMemsetAsync(cnt, 0, (size_t)n_nodes * sizeof(unsigned long long), stream), "clear");
    hipLaunchKernelGGL(sort_keys_kernel, dim3(4096), dim3(256), 0, stream, d_src, d_d, n, k0, v0, cnt);
    check(hipGetLastError(), "key kernel");
    check(hipStreamSynchronize(stream), "sync");
    m_src.release();   // the sources follow from the counts
    hipcub::DoubleBuffer<unsigned long long> keys(k0, k1);
    hipcub::DoubleBuffer<unsigned int> vals(v0, v1);
    size_t tmp_bytes = 0;
    check(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, vals, n, 0, end_bit, stream),
          "sort size");
    void* tmp = m_tmp.alloc<unsigned char>(tmp_bytes, "sort scratch");
    check(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, vals, n, 0, end_bit, stream), "sort");
    check(hipStreamSynchronize(stream), "sync");
    m_tmp.release();
    m_k0.release();
    m_k1.release();
    int* d_tgt = m_tgt.alloc<int>(n, "target upload buffer");
    unsigned char* d_port = m_port.alloc<unsigned char>(n, "port upload buffer");
    float* d_w = m_w.alloc<float>(n, "weight upload buffer");
    check(hipMemcpyAsync(d_tgt, tgt, n * sizeof(int), hipMemcpyHostToDevice, stream), "upload");
    check(hipMemcpyAsync(d_port, port, n, hipMemcpyHostToDevice, stream), "upload");
    check(hipMemcpyAsync(d_w, w, n * sizeof(float), hipMemcpyHostToDevice, stream), "upload");
    int* t_o = m_to.alloc<int>(n, "sorted targets");
    unsigned char* p_o = m_po.alloc<unsigned char>(n, "sorted ports");
    float* w_o = m_wo.alloc<float>(n, "sorted weights");
    // the sorted delays reuse the index buffer that is not the sort's result
    int* d_o = reinterpret_cast<int*>(vals.Current() == v0 ? v1 : v0);
    hipLaunchKernelGGL(gather_kernel, dim3(4096), dim3(256), 0, stream, vals.Current(), n, d_tgt, d_port, d_w,
                       d_d, t_o, p_o, w_o, d_o);
    check(hipGetLastError(), "gather kernel");
    check(hipMemcpyAsync(tgt, t_o, n * sizeof(int), hipMemcpyDeviceToHost, stream), "download");
    check(hipMemcpyAsync(port, p_o, n, hipMemcpyDeviceToHost, stream), "download");
    check(hipMemcpyAsync(w, w_o, n * sizeof(float), hipMemcpyDeviceToHost, stream), "download");
    check(hipMemcpyAsync(d, d_o, n * sizeof(int), hipMemcpyDeviceToHost, stream), "download");
    check(hipMemcpyAsync(counts.data(), cnt, (size_t)n_nodes * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, stream),
          "download");
    check(hipStreamSynchronize(stream), "sync");
  }
  row_ptr.assign((size_t)n_nodes + 1, 0);
  for (int s = 0; s < n_nodes; ++s) row_ptr[s + 1] = row_ptr[s] + (long long)counts[s];
  return true;
}

}  // namespace ngpu
