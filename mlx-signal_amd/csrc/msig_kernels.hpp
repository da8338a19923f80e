// Fused spectral kernels for gfx950 (device code, included by the per-size
// translation units).
//
// Work decomposition.  A "unit" is one complex FFT: for real input it packs
// two consecutive segments of one row (z = x_s w + i x_{s+1} w), for complex
// input it is one segment.  A group of TPF threads walks a contiguous run of
// units of one row; a 256-thread block holds 256/TPF groups covering
// consecutive runs, so the overlapping samples of neighbouring segments are
// re-read by the same CU (L2/L1 hits) and leave HBM once.  Mean kinds keep
// per-bin sums in registers across the run, reduce the groups of a block in
// LDS in a fixed order, and a finalize kernel sums the blocks of a row in
// FP64, folds the conjugate mirror, and applies scale, 1/nseg and the
// onesided factor.  The reduction order depends only on the job's geometry.
//
// Reference counterparts: grouped mean kernels _stft_metal.py:458-505 and
// 620-667, per-frame kernels _stft_metal.py:104-147 and 318-351, group
// reduction _stft_metal.py:443-455.
#pragma once

#include <hip/hip_runtime.h>

#include "msig_fft.hpp"

namespace msig {

constexpr int BT = 256;  // threads per block

enum : int { K_PSD = 0, K_CSD = 1, K_COH = 2 };
enum : int { F_COMPLEX = 0, F_POWER = 1, F_CSD = 2, F_COH = 3 };

struct KArgs {
  const void* x;
  const void* y;
  const float* win;
  const cf* twt;     // exp(-2 pi i e / N), e < N
  float* ws;         // mean kinds: [rows][bpr][ncomp][N] block partials
  void* out0;        // frames kinds
  void* out1;
  void* out2;
  int64_t ld_x, ld_y;
  int64_t nseg;
  int32_t nperseg, hop;
  int32_t units, run, bpr;
  int32_t detrend;
  int32_t nbins;     // frames: bins written per frame
  int32_t apply_factor;
  float oscale;      // frames: power scale or sqrt(scale) for complex output
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier for the threads of one FFT group.  Groups that fit in a wavefront
// only need LDS ordering within the wave; larger groups span several waves
// and use the block barrier (every group of the block runs the same loop
// trip count, so all threads reach it).
template <int TPF>
__device__ __forceinline__ void group_sync() {
  if constexpr (TPF <= 64) {
    wave_lds_sync();
  } else {
    __syncthreads();
  }
}

// Sum a pair of floats over the TPF threads of a group.
template <int TPF>
__device__ __forceinline__ void group_sum2(float& a, float& b, float* red) {
  constexpr int W = TPF < 64 ? TPF : 64;
#pragma unroll
  for (int o = W / 2; o >= 1; o >>= 1) {
    a += __shfl_xor(a, o, W);
    b += __shfl_xor(b, o, W);
  }
  if constexpr (TPF > 64) {
    constexpr int NW = TPF / 64;  // waves per group
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    __syncthreads();
    if (lane == 0) {
      red[2 * wave] = a;
      red[2 * wave + 1] = b;
    }
    __syncthreads();
    const int w0 = (wave / NW) * NW;
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      sa += red[2 * (w0 + i)];
      sb += red[2 * (w0 + i) + 1];
    }
    a = sa;
    b = sb;
  }
}

// Forward FFT of the group's N points held E per thread.
template <int N, int E, int P = 1>
__device__ __forceinline__ void fft_passes(int t, cf* v, const cf* tw, cf* lds) {
  using PL = Plan<N, E>;
  if constexpr (P == 1) pass_compute<N, E, 0>(v, tw);
  if constexpr (P < PL::NP) {
    pass_store<N, E, P - 1>(t, v, lds);
    group_sync<PL::TPF>();
    pass_load<N, E, P>(t, v, lds);
    if constexpr (PL::TPF > 64) __syncthreads();  // reads done before next stores
    pass_compute<N, E, P>(v, tw);
    fft_passes<N, E, P + 1>(t, v, tw, lds);
  }
}

// Load one unit (two real segments, or one complex segment) of a row,
// detrend, window.  Units past the row's end load zeros.
template <int N, int E, bool CPLX>
__device__ __forceinline__ void load_unit(const void* xrow_v, int t, int unit, bool active,
                                          const KArgs& a, const float* w, float* red, cf* z) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  const bool full = (a.nperseg == N);
  if constexpr (!CPLX) {
    const float* xrow = static_cast<const float*>(xrow_v);
    const int64_t s0 = 2 * (int64_t)unit;
    const bool v0 = active;
    const bool v1 = active && (s0 + 1 < a.nseg);
    const float* p0 = xrow + s0 * a.hop;
    const float* p1 = p0 + a.hop;
    if (full && v1) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        z[e].re = p0[t + TPF * e];
        z[e].im = p1[t + TPF * e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = t + TPF * e;
        const bool in = n < a.nperseg;
        z[e].re = (v0 && in) ? p0[n] : 0.f;
        z[e].im = (v1 && in) ? p1[n] : 0.f;
      }
    }
  } else {
    const cf* xrow = static_cast<const cf*>(xrow_v);
    const cf* p0 = xrow + (int64_t)unit * a.hop;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = t + TPF * e;
      z[e] = (active && n < a.nperseg) ? p0[n] : cf{0.f, 0.f};
    }
  }
  if (a.detrend) {
    float sr = 0.f, si = 0.f;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      sr += z[e].re;
      si += z[e].im;
    }
    group_sum2<TPF>(sr, si, red);
    const float inv = 1.0f / (float)a.nperseg;
    const float mr = sr * inv, mi = si * inv;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      z[e].re = (z[e].re - mr) * w[e];
      z[e].im = (z[e].im - mi) * w[e];
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      z[e].re *= w[e];
      z[e].im *= w[e];
    }
  }
}

// Grouped mean over segments: per-bin sums of |Zx|^2 (PSD), conj(Zx) Zy
// (CSD) or both spectra and the cross term (COH), written per block.
template <int N, int E, int KIND, bool CPLX>
__global__ __launch_bounds__(BT) void k_mean(const KArgs a) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  constexpr int NC = KIND == K_PSD ? 1 : (KIND == K_CSD ? 2 : 4);
  __shared__ cf lds_all[GPB * PL::LDS];
  __shared__ float red[2 * (BT / 64)];
  const int tid = threadIdx.x;
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * PL::LDS;
  const int64_t row = blockIdx.x / a.bpr;
  const int bir = blockIdx.x % a.bpr;
  const int unit0 = (bir * GPB + g) * a.run;

  cf tw[PL::TWN];
  load_twiddles<N, E>(t, a.twt, tw);
  float w[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPF * e;
    w[e] = n < a.nperseg ? a.win[n] : 0.f;
  }
  const size_t esz = CPLX ? sizeof(cf) : sizeof(float);
  const char* xrow = static_cast<const char*>(a.x) + (size_t)row * a.ld_x * esz;
  const char* yrow = KIND == K_PSD ? nullptr : static_cast<const char*>(a.y) + (size_t)row * a.ld_y * esz;

  float acc[NC][E];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int e = 0; e < E; ++e) acc[c][e] = 0.f;

  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;
    cf zx[E];
    load_unit<N, E, CPLX>(xrow, t, unit, active, a, w, red, zx);
    fft_passes<N, E>(t, zx, tw, lds);
    if constexpr (KIND == K_PSD) {
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) acc[0][e] += zx[e].re * zx[e].re + zx[e].im * zx[e].im;
      }
    } else {
      cf zy[E];
      group_sync<TPF>();  // last-pass reads of zx done before zy's stores
      load_unit<N, E, CPLX>(yrow, t, unit, active, a, w, red, zy);
      fft_passes<N, E>(t, zy, tw, lds);
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const cf q = cmulc(zx[e], zy[e]);
          if constexpr (KIND == K_CSD) {
            acc[0][e] += q.re;
            acc[1][e] += q.im;
          } else {
            acc[0][e] += zx[e].re * zx[e].re + zx[e].im * zx[e].im;
            acc[1][e] += zy[e].re * zy[e].re + zy[e].im * zy[e].im;
            acc[2][e] += q.re;
            acc[3][e] += q.im;
          }
        }
      }
    }
    group_sync<TPF>();  // reads of this unit's last pass done before next stores
  }

  // block reduction: groups summed in a fixed order
  float* redf = reinterpret_cast<float*>(lds_all);
  float* dst = a.ws + ((size_t)row * a.bpr + bir) * NC * N;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) redf[g * N + PL::bin(t, e)] = acc[c][e];
    __syncthreads();
    for (int k = tid; k < N; k += BT) {
      float s = 0.f;
#pragma unroll
      for (int gg = 0; gg < GPB; ++gg) s += redf[gg * N + k];
      dst[c * N + k] = s;
    }
  }
}

// Per-frame spectra.  Real input: the two packed segments are separated with
// X_s = (Z[k] + conj Z[N-k]) / 2, X_{s+1} = -i (Z[k] - conj Z[N-k]) / 2.
template <int N, int E, int OK, bool CPLX>
__global__ __launch_bounds__(BT) void k_frames(const KArgs a) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  constexpr bool PAIR = (OK == F_CSD || OK == F_COH);
  __shared__ cf lds_all[GPB * PL::LDS];
  __shared__ float red[2 * (BT / 64)];
  const int tid = threadIdx.x;
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * PL::LDS;
  const int64_t row = blockIdx.x / a.bpr;
  const int bir = blockIdx.x % a.bpr;
  const int unit0 = (bir * GPB + g) * a.run;

  cf tw[PL::TWN];
  load_twiddles<N, E>(t, a.twt, tw);
  float w[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPF * e;
    w[e] = n < a.nperseg ? a.win[n] : 0.f;
  }
  const size_t esz = CPLX ? sizeof(cf) : sizeof(float);
  const char* xrow = static_cast<const char*>(a.x) + (size_t)row * a.ld_x * esz;
  const char* yrow = PAIR ? static_cast<const char*>(a.y) + (size_t)row * a.ld_y * esz : nullptr;
  const int nb = a.nbins;
  const float os = a.oscale;

  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;
    cf zx[E];
    load_unit<N, E, CPLX>(xrow, t, unit, active, a, w, red, zx);
    fft_passes<N, E>(t, zx, tw, lds);
    if constexpr (CPLX) {
      // one segment per unit; bins come out in registers, no separation
      cf zy[E];
      if constexpr (PAIR) {
        group_sync<TPF>();
        load_unit<N, E, CPLX>(yrow, t, unit, active, a, w, red, zy);
        fft_passes<N, E>(t, zy, tw, lds);
      }
      if (active) {
        const int64_t fr = row * a.nseg + unit;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int k = PL::bin(t, e);
          const size_t o = (size_t)fr * nb + k;
          if constexpr (OK == F_COMPLEX) {
            static_cast<cf*>(a.out0)[o] = cf{zx[e].re * os, zx[e].im * os};
          } else if constexpr (OK == F_POWER) {
            static_cast<float*>(a.out0)[o] = (zx[e].re * zx[e].re + zx[e].im * zx[e].im) * os;
          } else if constexpr (OK == F_CSD) {
            const cf q = cmulc(zx[e], zy[e]);
            static_cast<cf*>(a.out0)[o] = cf{q.re * os, q.im * os};
          } else {
            const cf q = cmulc(zx[e], zy[e]);
            static_cast<float*>(a.out0)[o] = (zx[e].re * zx[e].re + zx[e].im * zx[e].im) * os;
            static_cast<float*>(a.out1)[o] = (zy[e].re * zy[e].re + zy[e].im * zy[e].im) * os;
            static_cast<cf*>(a.out2)[o] = cf{q.re * os, q.im * os};
          }
        }
      }
    } else {
      // separate the two real segments through the conjugate mirror
      cf x0[E], x1[E];
      group_sync<TPF>();
#pragma unroll
      for (int e = 0; e < E; ++e) lds[PL::pad(PL::bin(t, e))] = zx[e];
      group_sync<TPF>();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = t + TPF * e;
        const cf p = lds[PL::pad(k)];
        const cf q = lds[PL::pad((N - k) & (N - 1))];
        x0[e] = cf{0.5f * (p.re + q.re), 0.5f * (p.im - q.im)};
        x1[e] = cf{0.5f * (p.im + q.im), -0.5f * (p.re - q.re)};
      }
      cf y0[E], y1[E];
      if constexpr (PAIR) {
        group_sync<TPF>();
        cf zy[E];
        load_unit<N, E, CPLX>(yrow, t, unit, active, a, w, red, zy);
        fft_passes<N, E>(t, zy, tw, lds);
        group_sync<TPF>();
#pragma unroll
        for (int e = 0; e < E; ++e) lds[PL::pad(PL::bin(t, e))] = zy[e];
        group_sync<TPF>();
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int k = t + TPF * e;
          const cf p = lds[PL::pad(k)];
          const cf q = lds[PL::pad((N - k) & (N - 1))];
          y0[e] = cf{0.5f * (p.re + q.re), 0.5f * (p.im - q.im)};
          y1[e] = cf{0.5f * (p.im + q.im), -0.5f * (p.re - q.re)};
        }
      }
      if (active) {
        const int64_t s0 = 2 * (int64_t)unit;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t s = s0 + h;
          if (s >= a.nseg) break;
          const int64_t fr = row * a.nseg + s;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int k = t + TPF * e;
            if (k >= nb) continue;
            const size_t o = (size_t)fr * nb + k;
            const cf X = h ? x1[e] : x0[e];
            float f = 1.f;
            if (a.apply_factor) f = (k == 0 || 2 * k == N) ? 1.f : 2.f;
            if constexpr (OK == F_COMPLEX) {
              static_cast<cf*>(a.out0)[o] = cf{X.re * os, X.im * os};
            } else if constexpr (OK == F_POWER) {
              static_cast<float*>(a.out0)[o] = (X.re * X.re + X.im * X.im) * (os * f);
            } else {
              const cf Y = h ? y1[e] : y0[e];
              const cf q = cmulc(X, Y);
              if constexpr (OK == F_CSD) {
                static_cast<cf*>(a.out0)[o] = cf{q.re * os * f, q.im * os * f};
              } else {
                static_cast<float*>(a.out0)[o] = (X.re * X.re + X.im * X.im) * (os * f);
                static_cast<float*>(a.out1)[o] = (Y.re * Y.re + Y.im * Y.im) * (os * f);
                static_cast<cf*>(a.out2)[o] = cf{q.re * os * f, q.im * os * f};
              }
            }
          }
        }
      }
    }
    group_sync<TPF>();
  }
}

}  // namespace msig
