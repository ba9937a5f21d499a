// firex_audit.hip — the field checksum kernel behind firex_field_checksum
// (auditable captures: the manifest of an offline capture records, per
// frame, a checksum of every solver field).  The checksum's definition is
// fx_checksum_term in fx_common.h; this kernel streams the stored words once
// (HBM-bound), sums the terms per block and then the block sums (64-bit
// integer addition: the order does not matter, so the result is exact).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "firex_kernels.h"

namespace {

constexpr int kChecksumThreads = 256;
constexpr uint32_t kChecksumBlocks = 1024;  // partials of the first pass

__device__ inline uint64_t block_sum(uint64_t v) {
  __shared__ uint64_t part[kChecksumThreads];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int h = kChecksumThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  return part[0];
}

// First pass: each block sums the terms of a grid-strided share of the words.
__global__ __launch_bounds__(kChecksumThreads) void k_checksum(const uint32_t* __restrict__ words, uint64_t n,
                                                               uint32_t last_mask, int zero,
                                                               uint64_t* __restrict__ partial) {
  uint64_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kChecksumThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kChecksumThreads + threadIdx.x; i < n; i += stride) {
    uint32_t w = zero ? 0u : words[i];
    if (i == n - 1) w &= last_mask;
    acc += fx_checksum_term(i, w);
  }
  const uint64_t sum = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = sum;
}

// Second pass: one block adds the partials.
__global__ __launch_bounds__(kChecksumThreads) void k_checksum_final(const uint64_t* __restrict__ partial,
                                                                     uint32_t count, uint64_t* __restrict__ out) {
  uint64_t acc = 0;
  for (uint32_t i = threadIdx.x; i < count; i += kChecksumThreads) acc += partial[i];
  const uint64_t sum = block_sum(acc);
  if (threadIdx.x == 0) *out = sum;
}

}  // namespace

namespace fx {

void launch_checksum(const void* data, uint64_t bytes, uint64_t* partial, uint64_t* out, hipStream_t s) {
  const uint64_t n = (bytes + 3) / 4;
  const uint32_t rem = (uint32_t)(bytes % 4);
  const uint32_t last_mask = rem ? (1u << (8 * rem)) - 1u : 0xffffffffu;
  uint64_t blocks = (n + kChecksumThreads * 8 - 1) / (kChecksumThreads * 8);
  if (blocks > kChecksumBlocks) blocks = kChecksumBlocks;
  if (blocks == 0) blocks = 1;  // an empty field sums to 0
  k_checksum<<<(uint32_t)blocks, kChecksumThreads, 0, s>>>((const uint32_t*)data, n, last_mask, data == nullptr,
                                                           partial);
  k_checksum_final<<<1, kChecksumThreads, 0, s>>>(partial, (uint32_t)blocks, out);
}

}  // namespace fx
