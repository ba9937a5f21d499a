// firex_audit.hip — the field checksum kernel behind firex_field_checksum
// (auditable captures: the manifest of an offline capture records, per
// frame, a checksum of every solver field).  The checksum's definition is
// fx_checksum_term in fx_common.h; this kernel streams the stored words once
// (HBM-bound) and adds the terms with 64-bit integer atomics, whose order does
// not matter.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "firex_kernels.h"

namespace {

constexpr int kChecksumThreads = 256;

__global__ __launch_bounds__(kChecksumThreads) void k_checksum(const uint32_t* __restrict__ words, uint64_t n,
                                                               uint32_t last_mask, int zero, uint64_t* out) {
  uint64_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kChecksumThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kChecksumThreads + threadIdx.x; i < n; i += stride) {
    uint32_t w = zero ? 0u : words[i];
    if (i == n - 1) w &= last_mask;
    acc += fx_checksum_term(i, w);
  }
  __shared__ uint64_t part[kChecksumThreads];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kChecksumThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd((unsigned long long*)out, (unsigned long long)part[0]);
}

}  // namespace

namespace fx {

void launch_checksum(const void* data, uint64_t bytes, uint64_t* out, hipStream_t s) {
  launch_fill_u32(out, 2, 0, s);
  const uint64_t n = (bytes + 3) / 4;
  if (n == 0) return;
  const uint32_t rem = (uint32_t)(bytes % 4);
  const uint32_t last_mask = rem ? (1u << (8 * rem)) - 1u : 0xffffffffu;
  uint64_t blocks = (n + kChecksumThreads * 8 - 1) / (kChecksumThreads * 8);
  if (blocks > 4096) blocks = 4096;
  k_checksum<<<(uint32_t)blocks, kChecksumThreads, 0, s>>>((const uint32_t*)data, n, last_mask, data == nullptr,
                                                           out);
}

}  // namespace fx
