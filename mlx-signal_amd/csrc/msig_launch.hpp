// Host-side launch helpers shared by the per-FFT-size translation units and
// the C ABI (msig_abi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "msig_kernels.hpp"

namespace msig {

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

inline Geo geometry(int64_t rows, int64_t nseg, int N, bool cplx, int maxrun) {
  Geo g;
  const int tpf = N / elems_per_thread(N);
  g.gpb = BT / tpf;
  g.units = (int32_t)(cplx ? nseg : (nseg + 1) / 2);
  int64_t bprt = (1024 + rows - 1) / rows;
  if (bprt < 1) bprt = 1;
  int64_t run = (g.units + g.gpb * bprt - 1) / (g.gpb * bprt);
  if (run < 1) run = 1;
  if (run > maxrun) run = maxrun;
  g.run = (int32_t)run;
  g.bpr = (int32_t)((g.units + (int64_t)g.gpb * run - 1) / ((int64_t)g.gpb * run));
  return g;
}

constexpr int MEAN_MAX_RUN = 32;    // <= 64 segments summed sequentially in FP32
constexpr int FRAMES_MAX_RUN = 8;

// One pair of launchers per fused FFT size, defined in msig_n<N>.hip.
#define MSIG_DECLARE_LAUNCHERS(NV)                                                         \
  hipError_t launch_mean_##NV(int kind, bool cplx, const KArgs& a, int64_t rows,           \
                              hipStream_t s);                                              \
  hipError_t launch_frames_##NV(int ok, bool cplx, const KArgs& a, int64_t rows, hipStream_t s);
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
// instantiates the kernels for FFT size N.
#define MSIG_DEFINE_LAUNCHERS(NV)                                                          \
  namespace msig {                                                                         \
  hipError_t launch_mean_##NV(int kind, bool cplx, const KArgs& a, int64_t rows,            \
                             hipStream_t s) {                                              \
    constexpr int E = NV < 16 ? NV : 16;                                                   \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    if (cplx) {                                                                            \
      if (kind == K_PSD) hipLaunchKernelGGL((k_mean<NV, E, K_PSD, true>), grid, block, 0, s, a); \
      else if (kind == K_CSD) hipLaunchKernelGGL((k_mean<NV, E, K_CSD, true>), grid, block, 0, s, a); \
      else hipLaunchKernelGGL((k_mean<NV, E, K_COH, true>), grid, block, 0, s, a);         \
    } else {                                                                               \
      if (kind == K_PSD) hipLaunchKernelGGL((k_mean<NV, E, K_PSD, false>), grid, block, 0, s, a); \
      else if (kind == K_CSD) hipLaunchKernelGGL((k_mean<NV, E, K_CSD, false>), grid, block, 0, s, a); \
      else hipLaunchKernelGGL((k_mean<NV, E, K_COH, false>), grid, block, 0, s, a);        \
    }                                                                                      \
    return hipGetLastError();                                                              \
  }                                                                                        \
  hipError_t launch_frames_##NV(int ok, bool cplx, const KArgs& a, int64_t rows,            \
                               hipStream_t s) {                                            \
    constexpr int E = NV < 16 ? NV : 16;                                                   \
    const dim3 grid((unsigned)(rows * a.bpr)), block(BT);                                  \
    if (cplx) {                                                                            \
      if (ok == F_COMPLEX) hipLaunchKernelGGL((k_frames<NV, E, F_COMPLEX, true>), grid, block, 0, s, a); \
      else if (ok == F_POWER) hipLaunchKernelGGL((k_frames<NV, E, F_POWER, true>), grid, block, 0, s, a); \
      else if (ok == F_CSD) hipLaunchKernelGGL((k_frames<NV, E, F_CSD, true>), grid, block, 0, s, a); \
      else hipLaunchKernelGGL((k_frames<NV, E, F_COH, true>), grid, block, 0, s, a);       \
    } else {                                                                               \
      if (ok == F_COMPLEX) hipLaunchKernelGGL((k_frames<NV, E, F_COMPLEX, false>), grid, block, 0, s, a); \
      else if (ok == F_POWER) hipLaunchKernelGGL((k_frames<NV, E, F_POWER, false>), grid, block, 0, s, a); \
      else if (ok == F_CSD) hipLaunchKernelGGL((k_frames<NV, E, F_CSD, false>), grid, block, 0, s, a); \
      else hipLaunchKernelGGL((k_frames<NV, E, F_COH, false>), grid, block, 0, s, a);      \
    }                                                                                      \
    return hipGetLastError();                                                              \
  }                                                                                        \
  }
