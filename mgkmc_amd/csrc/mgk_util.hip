// Measurement helper: a streaming copy that moves 16 bytes per lane.  Its achieved
// bandwidth is the practical HBM ceiling the sweeps are compared against
// (the published 8 TB/s peak is not reachable by a read + write stream).
#include "mgk_host.h"
#include "../../include/mgkmc_hip.h"
#include <algorithm>

namespace {
// one 16-byte element per lane, one pass over the grid (no grid-stride loop):
// measured the fastest of the usual copy shapes on MI355X (6.2 TB/s at 1-4 GiB,
// against 5.1-5.4 TB/s for grid-stride loops and 5.6-5.9 TB/s for 2-4
// elements per lane)
__global__ __launch_bounds__(MGK_NT) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       size_t n) {
  const size_t i = blockIdx.x * (size_t)MGK_NT + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
}  // namespace

extern "C" int mgk_stream_copy(const void* src, void* dst, size_t bytes, void* stream) {
  if (!src || !dst || bytes % 16 != 0 || ((uintptr_t)src % 16) != 0 || ((uintptr_t)dst % 16) != 0) {
    mgk_set_error("mgk_stream_copy: null pointer, or size / address not a multiple of 16 bytes", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  const size_t n = bytes / 16;
  if (n == 0) return MGK_OK;
  const size_t nb = (n + MGK_NT - 1) / MGK_NT;
  if (nb > 0x7fffffffULL) {
    mgk_set_error("mgk_stream_copy: copy too large for one launch", __FILE__, __LINE__);
    return MGK_ERR_ARG;
  }
  k_stream_copy<<<(unsigned)nb, MGK_NT, 0, (hipStream_t)stream>>>((const uint4*)src, (uint4*)dst, n);
  MGK_HIP(hipGetLastError());
  return MGK_OK;
}
