// The single-launch ragged batch kernel (welch and csd sums on real input at
// hop = N/2, every fused FFT size in one kernel).
#include "msig_launch.hpp"

namespace msig {

__global__ __launch_bounds__(BT, 2) void k_mean_ragged(const KArgs* __restrict__ jobs,
                                                       const WorkItem* __restrict__ queue,
                                                       int32_t nwork,
                                                       unsigned long long* ticket,
                                                       float* ws_base) {
  __shared__ cf pool[RAGGED_POOL];
  __shared__ float red[2 * (BT / 64)];
  __shared__ unsigned long long slot;
  const unsigned long long period = (unsigned long long)nwork + gridDim.x;
  int prev = -1;
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(ticket, 1ull) % period;
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane((int)slot);
    if (w >= nwork) break;
    const WorkItem item = queue[w];
    const KArgs& a = jobs[item.job];
    const int cls = a.cls;
    const bool refill = cls != prev;
    prev = cls;
    // threadIdx.x through an opaque move: the bodies' thread-index
    // arithmetic then depends on the item and stays inside its body (hoisted
    // out of this loop for all 18 bodies at once, it spilled in every one)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    switch (cls) {
#define MSIG_RAGGED_CASE(NV)                                                     \
  case ragged_class(NV, K_PSD):                                                  \
    ragged_item<NV, K_PSD>(a, item.blk, refill, pool, red, ws_base, tid);             \
    break;                                                                       \
  case ragged_class(NV, K_CSD):                                                  \
    ragged_item<NV, K_CSD>(a, item.blk, refill, pool, red, ws_base, tid);             \
    break;
      MSIG_RAGGED_CASE(16)
      MSIG_RAGGED_CASE(32)
      MSIG_RAGGED_CASE(64)
      MSIG_RAGGED_CASE(128)
      MSIG_RAGGED_CASE(256)
      MSIG_RAGGED_CASE(512)
      MSIG_RAGGED_CASE(1024)
      MSIG_RAGGED_CASE(2048)
      MSIG_RAGGED_CASE(4096)
#undef MSIG_RAGGED_CASE
      default:
        break;
    }
    __syncthreads();  // LDS and `slot` are free for the next item
  }
}

int ragged_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, reinterpret_cast<const void*>(&k_mean_ragged), BT, 0) == hipSuccess && nb > 0)
    return nb;
  (void)hipGetLastError();
  return 2;
}

hipError_t launch_mean_ragged(const KArgs* jobs, const WorkItem* queue, int32_t nwork,
                              unsigned long long* ticket, float* ws, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_mean_ragged, dim3(grid), dim3(BT), 0, s, jobs, queue, nwork, ticket, ws);
  return hipGetLastError();
}

}  // namespace msig
