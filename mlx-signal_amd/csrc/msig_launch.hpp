// Host-side launch helpers shared by the per-FFT-size translation units and
// the C ABI (msig_abi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <tuple>
#include <utility>

#include "msig_kernels.hpp"

namespace msig {

// Events that time the next fused transform launch made on this thread (the
// kernel-timing instrumentation; both null otherwise).  launch_timed passes
// them to hipExtLaunchKernel, which stamps them with the dispatch's own start
// and end, so the measured window holds the kernel and nothing queued
// around it.
struct LaunchTiming {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

template <class... P, class... A>
hipError_t launch_timed(void (*k)(P...), dim3 grid, dim3 block, hipStream_t s, A&&... args) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  auto held = std::tuple<P...>(std::forward<A>(args)...);
  void* argv[sizeof...(P) > 0 ? sizeof...(P) : 1];
  std::apply([&](auto&... v) {
    int i = 0;
    ((argv[i++] = static_cast<void*>(&v)), ...);
  }, held);
  return hipExtLaunchKernel(reinterpret_cast<const void*>(k), grid, block, argv, 0, s,
                            g_launch_timing.start, g_launch_timing.stop, 0);
}

// Segment-run decomposition of one job.  Depends only on the job's geometry,
// so a job's reduction order (hence its bits) is the same whether it runs
// alone or next to other work.
struct Geo {
  int32_t units;  // FFT units per row (segment pairs for real input)
  int32_t gpb;    // FFT groups per 256-thread block
  int32_t run;    // units walked by one group
  int32_t bpr;    // blocks per row
};

inline int elems_per_thread(int N) { return N < 16 ? N : 16; }

inline Geo geometry(int64_t rows, int64_t nseg, int N, bool cplx, int maxrun,
                    int64_t target_blocks = 1024) {
  Geo g;
  const int tpf = N / elems_per_thread(N);
  g.gpb = BT / tpf;
  g.units = (int32_t)(cplx ? nseg : (nseg + 1) / 2);
  int64_t bprt = (target_blocks + rows - 1) / rows;
  if (bprt < 1) bprt = 1;
  int64_t run = (g.units + g.gpb * bprt - 1) / (g.gpb * bprt);
  if (run < 1) run = 1;
  if (run > maxrun) run = maxrun;
  g.run = (int32_t)run;
  g.bpr = (int32_t)((g.units + (int64_t)g.gpb * run - 1) / ((int64_t)g.gpb * run));
  return g;
}

constexpr int MEAN_MAX_RUN = 32;    // <= 64 segments summed sequentially in FP32

// Mean-kernel plan sized to the blocks the GPU holds at once (`resident` =
// blocks per CU at the kernel's occupancy x CUs).  Blocks of equal work
// launched B at a time run in ceil(B / resident) rounds of `run` pairs per
// group, so the cost of a run is rounds x run.  The plan keeps the longest
// run (maxrun: fewest blocks, partials and block prologues) unless a shorter
// one costs at least 5% less; among equal costs the longer run wins.  A block
// count a little above the resident count would leave a second round with a
// quarter of the blocks at low occupancy: 64 x 2^20 at nperseg 1024 planned
// 1024 blocks for 768 resident ones (16 pairs per group, two rounds); this
// plan gives 768 blocks of 22 pairs per group in one round.  Jobs of many
// rounds keep maxrun, so a row of a large call is planned as it is alone.
inline Geo geometry_resident(int64_t rows, int64_t nseg, int N, bool cplx, int maxrun,
                             int64_t resident) {
  Geo g;
  const int tpf = N / elems_per_thread(N);
  g.gpb = BT / tpf;
  g.units = (int32_t)(cplx ? nseg : (nseg + 1) / 2);
  if (resident < 1) resident = 1;
  auto cost_of = [&](int64_t run) {
    const int64_t bpr = (g.units + g.gpb * run - 1) / (g.gpb * run);
    return ((rows * bpr + resident - 1) / resident) * run;
  };
  const int64_t top = g.units < maxrun ? (g.units > 1 ? g.units : 1) : maxrun;
  int64_t best_cost = cost_of(top), best_run = top;
  for (int64_t run = top - 1; run >= 1; --run) {
    const int64_t cost = cost_of(run);
    if (cost < best_cost) {
      best_cost = cost;
      best_run = run;
    }
  }
  if (20 * best_cost > 19 * cost_of(top)) best_run = top;
  g.run = (int32_t)best_run;
  g.bpr = (int32_t)((g.units + (int64_t)g.gpb * best_run - 1) / ((int64_t)g.gpb * best_run));
  return g;
}

constexpr int FRAMES_MAX_RUN = 8;
// Per-frame kernels: about 1536 blocks (stft 16 x 2^20, nperseg 1024, ran in
// 52 us at this target against 59 us at 1024; the power frames were flat).
constexpr int64_t FRAMES_TARGET_BLOCKS = 1536;
// Fused istft (hop = N/2): each group transforms one unit ahead of its own
// for the overlap carry, so longer runs waste less.
constexpr int ISTFT_OLA_MAX_RUN = 16;
constexpr int64_t ISTFT_OLA_TARGET_BLOCKS = 512;

// Split-precision matrix-core form (nfft = nperseg = 256, welch mean): one
// wave per segment pair, so a 256-thread block carries 4 groups.
enum : int { PREC_FP32 = 0, PREC_F16 = 1, PREC_BF16 = 2 };
constexpr int MFMA_NFFT = 256;

inline Geo geometry_mfma(int64_t rows, int64_t nseg, int maxrun, int64_t target_blocks = 2048) {
  Geo g;
  g.gpb = BT / 64;
  g.units = (int32_t)((nseg + 1) / 2);
  int64_t bprt = (target_blocks + rows - 1) / rows;
  if (bprt < 1) bprt = 1;
  int64_t run = (g.units + g.gpb * bprt - 1) / (g.gpb * bprt);
  if (run < 1) run = 1;
  if (run > maxrun) run = maxrun;
  g.run = (int32_t)run;
  g.bpr = (int32_t)((g.units + (int64_t)g.gpb * run - 1) / ((int64_t)g.gpb * run));
  return g;
}

hipError_t launch_mean_mfma256(int prec, const KArgs& a, int64_t rows, hipStream_t s);

// The single-launch ragged batch kernel (msig_ragged.hip).
int ragged_blocks_per_cu();
hipError_t launch_mean_ragged(const KArgs* jobs, const WorkItem* queue, int32_t nwork,
                              unsigned long long* ticket, float* ws, int grid, hipStream_t s);

// One pair of launchers per fused FFT size, defined in msig_n<N>.hip.
#define MSIG_DECLARE_LAUNCHERS(NV)                                                         \
  hipError_t launch_mean_##NV(int kind, bool cplx, bool half, const KArgs& a, int64_t rows,  \
                              hipStream_t s);                                              \
  hipError_t launch_frames_##NV(int kind, bool cplx, bool half, bool ext, const KArgs& a,    \
                                int64_t rows, hipStream_t s);                               \
  hipError_t launch_mean_batch_##NV(int kind, bool cplx, bool half, int sched,             \
                                    const KArgs* jobs, const WorkItem* queue, int32_t nwork, \
                                    float* ws, int grid, hipStream_t s);                    \
  hipError_t launch_istft_##NV(const KArgs& a, int64_t rows, hipStream_t s);                 \
  hipError_t launch_istft_ola_##NV(const KArgs& a, int64_t rows, hipStream_t s);             \
  int mean_blocks_per_cu_##NV(int kind, bool cplx, bool half);
MSIG_DECLARE_LAUNCHERS(16)
MSIG_DECLARE_LAUNCHERS(32)
MSIG_DECLARE_LAUNCHERS(64)
MSIG_DECLARE_LAUNCHERS(128)
MSIG_DECLARE_LAUNCHERS(256)
MSIG_DECLARE_LAUNCHERS(512)
MSIG_DECLARE_LAUNCHERS(1024)
MSIG_DECLARE_LAUNCHERS(2048)
MSIG_DECLARE_LAUNCHERS(4096)

}  // namespace msig

// Defines launch_mean_<N> / launch_frames_<N> in the translation unit that
// instantiates the kernels for FFT size N.  `half` selects the hop == N/2
// register-reuse kernels (real input, nperseg == N).
#define MSIG_LAUNCH_MEAN(NV, K, C, H) \
  le = launch_timed(&k_mean<NV, (NV < 16 ? NV : 16), K, C, H>, grid, block, s, a)
#define MSIG_LAUNCH_FRAMES(NV, K, C, H) \
  le = launch_timed(&k_frames<NV, (NV < 16 ? NV : 16), K, C, H>, grid, block, s, a)
// complex frames through the virtual zero extension (ext) or plain
#define MSIG_LAUNCH_FRAMES_COMPLEX(NV, C, H)                                                    \
  do {                                                                                        \
    if (ext)                                                                                  \
      le = launch_timed(&k_frames<NV, (NV < 16 ? NV : 16), F_COMPLEX, C, H, true>, grid, block, \
                        s, a);                                                                \
    else                                                                                      \
      MSIG_LAUNCH_FRAMES(NV, F_COMPLEX, C, H);                                                \
  } while (0)
#define MSIG_LAUNCH_MEAN_BATCH_S(NV, K, C, H, SCH)                                   \
  hipLaunchKernelGGL((k_mean_batch<NV, (NV < 16 ? NV : 16), K, C, H, SCH>), dim3(grid),  \
                     dim3(BT), 0, s, jobs, queue, nwork, ws)
// radix schedule 1 is a distinct kernel only where N is not a power of 16
#define MSIG_LAUNCH_MEAN_BATCH(NV, K, C, H)                                          \
  do {                                                                               \
    constexpr bool REM = Plan<NV, (NV < 16 ? NV : 16), 1>::REM_FIRST;                \
    if (sched == 1 && REM)                                                           \
      MSIG_LAUNCH_MEAN_BATCH_S(NV, K, C, H, (REM ? 1 : 0));                          \
    else                                                                             \
      MSIG_LAUNCH_MEAN_BATCH_S(NV, K, C, H, 0);                                      \
  } while (0)
#define MSIG_KIND_SWITCH(L, NV, K0, K1, K2, C, H) \
  if (kind == K0) L(NV, K0, C, H);                \
  else if (kind == K1) L(NV, K1, C, H);           \
  else L(NV, K2, C, H);

// Blocks per CU of one standalone mean kernel: the runtime's occupancy of the
// compiled kernel, or (no device, host-only planning) its launch bound.
#define MSIG_MEAN_OCC(NV, K, C, H)                                                        \
  do {                                                                                    \
    int nb = 0;                                                                           \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                     \
            &nb, reinterpret_cast<const void*>(&k_mean<NV, (NV < 16 ? NV : 16), K, C, H>), \
            BT, 0) == hipSuccess && nb > 0)                                               \
      return nb;                                                                          \
    (void)hipGetLastError();                                                              \
    return mean_waves<NV, K, H, C>();                                                     \
  } while (0)

#define MSIG_DEFINE_LAUNCHERS(NV)                                                          \
  namespace msig {                                                                         \
  int mean_blocks_per_cu_##NV(int kind, bool cplx, bool half) {                            \
    if (cplx) {                                                                            \
      MSIG_KIND_SWITCH(MSIG_MEAN_OCC, NV, K_PSD, K_CSD, K_COH, true, false)                \
    } else if (half) {                                                                     \
      MSIG_KIND_SWITCH(MSIG_MEAN_OCC, NV, K_PSD, K_CSD, K_COH, false, true)                \
    } else {                                                                               \
      MSIG_KIND_SWITCH(MSIG_MEAN_OCC, NV, K_PSD, K_CSD, K_COH, false, false)               \
    }                                                                                      \
  }                                                                                        \
  hipError_t launch_mean_##NV(int kind, bool cplx, bool half, const KArgs& a, int64_t rows, \
                              hipStream_t s) {                                             \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    hipError_t le = hipSuccess;                                                            \
    if (cplx) {                                                                            \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN, NV, K_PSD, K_CSD, K_COH, true, false)             \
    } else if (half) {                                                                     \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN, NV, K_PSD, K_CSD, K_COH, false, true)             \
    } else {                                                                               \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN, NV, K_PSD, K_CSD, K_COH, false, false)            \
    }                                                                                      \
    return le != hipSuccess ? le : hipGetLastError();                                      \
  }                                                                                        \
  hipError_t launch_frames_##NV(int kind, bool cplx, bool half, bool ext, const KArgs& a,  \
                                int64_t rows, hipStream_t s) {                             \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    hipError_t le = hipSuccess;                                                            \
    if (cplx) {                                                                            \
      if (kind == F_COMPLEX) MSIG_LAUNCH_FRAMES_COMPLEX(NV, true, false);                  \
      else MSIG_KIND_SWITCH(MSIG_LAUNCH_FRAMES, NV, F_POWER, F_CSD, F_COH, true, false)    \
    } else if (half) {                                                                     \
      if (kind == F_COMPLEX) MSIG_LAUNCH_FRAMES_COMPLEX(NV, false, true);                  \
      else MSIG_KIND_SWITCH(MSIG_LAUNCH_FRAMES, NV, F_POWER, F_CSD, F_COH, false, true)    \
    } else {                                                                               \
      if (kind == F_COMPLEX) MSIG_LAUNCH_FRAMES_COMPLEX(NV, false, false);                 \
      else MSIG_KIND_SWITCH(MSIG_LAUNCH_FRAMES, NV, F_POWER, F_CSD, F_COH, false, false)   \
    }                                                                                      \
    return le != hipSuccess ? le : hipGetLastError();                                      \
  }                                                                                        \
  hipError_t launch_mean_batch_##NV(int kind, bool cplx, bool half, int sched,             \
                                    const KArgs* jobs, const WorkItem* queue, int32_t nwork, \
                                    float* ws, int grid, hipStream_t s) {                   \
    if (cplx) {                                                                            \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN_BATCH, NV, K_PSD, K_CSD, K_COH, true, false)       \
    } else if (half) {                                                                     \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN_BATCH, NV, K_PSD, K_CSD, K_COH, false, true)       \
    } else {                                                                               \
      MSIG_KIND_SWITCH(MSIG_LAUNCH_MEAN_BATCH, NV, K_PSD, K_CSD, K_COH, false, false)      \
    }                                                                                      \
    return hipGetLastError();                                                              \
  }                                                                                        \
  hipError_t launch_istft_##NV(const KArgs& a, int64_t rows, hipStream_t s) {              \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    hipLaunchKernelGGL((k_istft<NV, (NV < 16 ? NV : 16)>), grid, block, 0, s, a);          \
    return hipGetLastError();                                                              \
  }                                                                                        \
  hipError_t launch_istft_ola_##NV(const KArgs& a, int64_t rows, hipStream_t s) {          \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    hipLaunchKernelGGL((k_istft_ola<NV, (NV < 16 ? NV : 16)>), grid, block, 0, s, a);      \
    return hipGetLastError();                                                              \
  }                                                                                        \
  }
