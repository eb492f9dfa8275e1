// Device-side connectivity (connect_device.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace ngpu {

// Sorts the connection store (host arrays of n entries) in place into the
// stable order by (source, d_int): the sort and the gathers run on the GPU,
// tgt, port, w and d come back sorted (src is left to the caller to refill
// from row_ptr: source s at [row_ptr[s], row_ptr[s + 1])); row_ptr[s] =
// first position of source s (n_nodes + 1 entries).  src[k] in
// [0, n_nodes), d[k] in [0, 65535].  Returns false (nothing done) when n is
// 0 or at least 2^32; throws on a device error (the arrays are then
// unchanged).  Synchronous on `stream`.
bool device_sort_connections(int* src, int* tgt, unsigned char* port, float* w, int* d, size_t n,
                             int n_nodes, std::vector<long long>& row_ptr, hipStream_t stream);

// fixed_indegree / fixed_outdegree on the GPU: out[o * deg + i] = the i-th
// distinct pool index (of n_pool) drawn for owner o, the same draws as the
// host rule under the call's 64-bit key.  out is host memory of
// n_own * deg ints.  Synchronous on `stream`.
void device_pick_distinct(int n_own, int deg, int n_pool, uint64_t key, int* out,
                          hipStream_t stream);

// n 32-bit words of the host generator's sequence from block counter ctr0
// (4 words per block, the last block cut): the same values as
// Engine::RandomInt.  out is host memory.  Synchronous on `stream`.
void device_random_words(uint64_t ctr0, uint64_t seed, size_t n, unsigned int* out,
                         hipStream_t stream);

}  // namespace ngpu
