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
#include "msig_stream.hpp"

namespace msig {

constexpr int BT = 256;  // threads per block

enum : int { K_PSD = 0, K_CSD = 1, K_COH = 2 };
enum : int { F_COMPLEX = 0, F_POWER = 1, F_CSD = 2, F_COH = 3 };


__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier for the threads of one FFT group.  Groups that fit in a wavefront
// only need LDS ordering within the wave; larger groups span several waves
// and use the block barrier, so every group of a block must run the same
// sequence of barriers (the mean, frames and istft loops give every group
// the same trip count and the same barrier-bearing steps per unit).
template <int TPF>
__device__ __forceinline__ void group_sync() {
  if constexpr (TPF <= 64) {
    wave_lds_sync();
  } else {
    __syncthreads();
  }
}

// Lane exchanges within a wave without LDS addressing.  dpp_f<CTRL> reads
// the lane given by a DPP control (quad permutes, half-row / row mirrors);
// swz16_f reads lane ^ 16 within each 32; lane_f reads one lane.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float swz16_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x401F));
}

__device__ __forceinline__ float lane_f(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Sum over aligned groups of W lanes (W a power of two <= 64); every lane of
// a group ends with the group's sum.  Each step adds the value of a partner
// lane: xor 1, xor 2 (quad permutes), the half-row mirror (i <-> 7 - i) and
// row mirror (i <-> 15 - i), which after the earlier steps pair up
// equal-sized partial sums, xor 16, and the two 32-lane halves.
template <int W>
__device__ __forceinline__ float lanes_sum(float v) {
  if constexpr (W >= 2) v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  if constexpr (W >= 4) v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  if constexpr (W >= 8) v += dpp_f<0x141>(v);  // row_half_mirror
  if constexpr (W >= 16) v += dpp_f<0x140>(v); // row_mirror
  if constexpr (W >= 32) v += swz16_f(v);
  if constexpr (W >= 64) v = lane_f(v, 0) + lane_f(v, 32);
  return v;
}

// Sums of two values over a whole wave in one reduction.  v_permlane32_swap
// exchanges a's upper 32 lanes with b's lower 32, so after one add lanes 0-31
// hold a[l] + a[l + 32] and lanes 32-63 hold b[l - 32] + b[l]; the five
// remaining steps stay inside each 32-lane half and reduce both values at
// once (9 VALU against 18 for two separate reductions).  Every lane ends
// with both sums.
__device__ __forceinline__ void wave_sum2(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  // element copies first: a bit cast applied to r[1] itself read element 0
  const unsigned r0 = r[0], r1 = r[1];
  float v = __uint_as_float(r0) + __uint_as_float(r1);
  v = lanes_sum<32>(v);
  a = lane_f(v, 0);
  b = lane_f(v, 32);
}

// Sum two values over the TPF threads of an FFT group; whole-wave groups
// reduce both in one chain (wave_sum2: same-box A/B on 64 x 2^20, median of
// four alternated processes: welch 88.1 -> 87.5 us, csd 176.5 -> 173.1,
// coherence 185.7 -> 181.5 per call).
template <int TPF>
__device__ __forceinline__ void group_sum2(float& a, float& b, float* red) {
  if constexpr (TPF >= 64) {
    wave_sum2(a, b);
  } else {
    a = lanes_sum<TPF>(a);
    b = lanes_sum<TPF>(b);
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

// Where a kernel keeps its Stockham twiddles: in the block's shared LDS
// table (Plan::twl_*, deduplicated across threads that share a twiddle) when
// the kernel wants it (WANT) and it fits in 32 KB, else in registers.  The
// table frees ~2*TWN VGPRs per thread, which buys occupancy.
template <int N, int E, int S = 0, bool WANT = true>
struct TwiddleHome {
  using PL = Plan<N, E, S>;
  static constexpr bool LDS = WANT && (PL::NP > 1) && (PL::TWL * 8 <= 32768);
  static constexpr int LDS_ELEMS = LDS ? PL::TWL : 1;
  static constexpr int REGS = LDS ? 1 : PL::TWN;
};

// Which kernels want the LDS table.  Up to N = 2048 (tables <= 16 KB) all
// do.  At N = 4096 the 32 KB table pays only where the VGPRs it frees let
// the kernel run two waves per SIMD instead of one, which is the two-signal
// hop == N/2 means (measured on 64 x 2^20: csd 358 -> 273 us, coherence
// 362 -> 251 us); the others stay at the same occupancy either way and only
// pay for the extra LDS reads (welch 125 -> 136 us, one-wave coherence
// 363 -> 413 us), so they keep register twiddles.
template <int N, int KIND, bool HALF>
constexpr bool mean_twl() {
  return N < 4096 || (KIND != K_PSD && HALF);
}
template <int N>
constexpr bool frames_twl() {
  return N < 4096;
}

// Fill the shared LDS twiddle table from the global FP64-derived table.  All
// of a thread's loads are issued before any LDS write, so the fill costs one
// L2 round trip per block rather than one per entry.  An all-wide table is
// column t = thread t's load_twiddles registers; a deduplicated one is
// filled entry by entry from Plan::twl_exp.
template <int N, int E, int S = 0>
__device__ __forceinline__ void fill_twiddle_lds(cf* twl, const cf* table) {
  using PL = Plan<N, E, S>;
  const MSIG_GLOBAL cf* g = as_global(table);
  if constexpr (!PL::TWL_DEDUP) {
    for (int t = threadIdx.x; t < PL::TPF; t += BT) {
      cf tmp[PL::TWN];
      load_twiddles<N, E, 1, const MSIG_GLOBAL cf*, S>(t, g, tmp);
#pragma unroll
      for (int s = 0; s < PL::TWN; ++s) twl[s * PL::TPF + t] = tmp[s];
    }
  } else {
    constexpr int ITER = (PL::TWL + BT - 1) / BT;
    cf tmp[ITER];
#pragma unroll
    for (int k = 0; k < ITER; ++k) {
      const int i = (int)threadIdx.x + k * BT;
      if (i < PL::TWL) tmp[k] = g[PL::twl_exp(i)];
    }
#pragma unroll
    for (int k = 0; k < ITER; ++k) {
      const int i = (int)threadIdx.x + k * BT;
      if (i < PL::TWL) twl[i] = tmp[k];
    }
  }
}

// Forward FFT of the group's N points held E per thread (register e holds
// sample t + TPF e); tw is this thread's twiddle registers (TWL = false) or
// the block's shared LDS table (TWL = true); radix schedule S (Plan).
// SPLIT: each exchange moves the real parts, then the imaginary parts,
// through Plan::LDS floats at `lds` (half the LDS of a complex exchange).
template <int N, int E, bool TWL, int P = 1, int S = 0, bool SPLIT = false>
__device__ __forceinline__ void fft_passes(int t, cf* v, const cf* tw, cf* lds) {
  using PL = Plan<N, E, S>;
  if constexpr (P == 1) {
    if constexpr (PL::REM_FIRST) {  // pass-0 register order (a renaming)
      cf tmp[E];
#pragma unroll
      for (int e = 0; e < E; ++e) tmp[e] = v[PL::load_order(e)];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = tmp[e];
    }
    pass_compute<N, E, 0, TWL, S>(v, tw, t);
  }
  if constexpr (P < PL::NP) {
    if constexpr (SPLIT) {
      float* lf = reinterpret_cast<float*>(lds);
      pass_store_part<N, E, P - 1, S, 0>(t, v, lf);
      group_sync<PL::TPF>();
      pass_load_part<N, E, P, S, 0>(t, v, lf);
      group_sync<PL::TPF>();  // real reads done before the imaginary stores
      pass_store_part<N, E, P - 1, S, 1>(t, v, lf);
      group_sync<PL::TPF>();
      pass_load_part<N, E, P, S, 1>(t, v, lf);
    } else {
      pass_store<N, E, P - 1, S>(t, v, lds);
      group_sync<PL::TPF>();
      pass_load<N, E, P, S>(t, v, lds);
    }
    if constexpr (PL::TPF > 64) __syncthreads();  // reads done before next stores
    pass_compute<N, E, P, TWL, S>(v, tw, t);
    fft_passes<N, E, TWL, P + 1, S, SPLIT>(t, v, tw, lds);
  }
}

// Detrend (segment mean over nperseg samples, per real/imag lane of the
// packed pair) and window, in place.  Samples n >= nperseg carry w = 0.
template <int N, int E>
__device__ __forceinline__ void prep_unit(cf* z, const float* w, int detrend, int nperseg,
                                          float* red) {
  constexpr int TPF = N / E;
  if (detrend) {
    float sr = z[0].re, si = z[0].im;
#pragma unroll
    for (int e = 1; e < E; ++e) {
      sr += z[e].re;
      si += z[e].im;
    }
    group_sum2<TPF>(sr, si, red);
    const float inv = 1.0f / (float)nperseg;
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

// A caller's side copy (cached constants returned beside the result), done
// by a kernel the call launches anyway: grid-strided over all its blocks
// (istft's time grid is as long as the signal).
__device__ __forceinline__ void side_copy(const float* src, float* dst, int64_t len) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += step)
    as_global_mut(dst)[i] = as_global(src)[i];
}

// Mean kernels whose FFT exchanges go through LDS one component at a time
// (fft_passes SPLIT): the block's exchange buffers then take half the LDS,
// which with the twiddle table lets four blocks share a CU.  The hop == N/2
// welch kernels up to N = 1024 fit four waves' VGPRs without spills this way
// (119-127); at 2048 the split form keeps three waves in fewer registers
// (168 -> 143).  The two-signal kinds take it from N = 512, at their own
// occupancy: the split exchange's linear, bank-conflict-free slots need no
// per-element address registers (csd 236 -> 200 VGPRs at N = 1024;
// coherence's 12-34 spilled VGPRs at 512-4096 go), where below 512 the
// complex exchange is already conflict free and the split form only doubles
// the LDS instructions.  Where same-box timing found no gain (csd at 2048
// and 4096) or a loss (welch at 4096, 106 -> 113 us at 64 x 2^20) the
// complex exchange stays.
template <int N, int KIND, bool HALF, bool CPLX = false>
constexpr bool mean_split() {
  constexpr int lo = KIND == K_PSD ? 32 : 512;
  constexpr int hi = KIND == K_PSD ? 2048 : (KIND == K_CSD ? 1024 : 4096);
  return HALF && !CPLX && N >= lo && N <= hi;
}
template <int N, int KIND, bool HALF, bool CPLX = false>
constexpr int mean_lds_cf() {
  using PL = Plan<N, (N < 16 ? N : 16)>;
  return (BT / PL::TPF) * (mean_split<N, KIND, HALF, CPLX>() ? PL::LDS / 2 : PL::LDS);
}

template <bool B>
struct RevTag {
  static constexpr bool value = B;
};

struct NoPrologue {
  __device__ void operator()() const {}
};

// Grouped mean over segments: per-bin sums of |Zx|^2 (PSD), conj(Zx) Zy
// (CSD) or both spectra and the cross term (COH), written per block.  `tid`
// is threadIdx.x (passed in so a kernel that runs several bodies can keep
// each body's thread-index arithmetic inside that body).
// `prologue` runs once in every thread before the twiddle table is read (k_mean
// fills the block's table there).  In the hop == N/2 stream it runs right
// after the group's first loads are issued, so that the fill and its barrier
// overlap those HBM loads instead of preceding them.
template <int N, int E, int KIND, bool CPLX, bool HALF, int S = 0, class Pro = NoPrologue>
__device__ __forceinline__ void mean_body(const KArgs& a, int64_t row, int bir, cf* lds_all,
                                          float* red, const cf* twl, float* ws, int tid,
                                          Pro&& prologue = Pro{}) {
  using PL = Plan<N, E, S>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  constexpr int NC = KIND == K_PSD ? 1 : (KIND == K_CSD ? 2 : 4);
  constexpr bool SPLIT = mean_split<N, KIND, HALF, CPLX>();
  static_assert(!SPLIT || PL::LDS % 2 == 0, "split exchange needs an even slot count");
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * (SPLIT ? PL::LDS / 2 : PL::LDS);
  const int unit0 = (bir * GPB + g) * a.run;

  using TH = TwiddleHome<N, E, S, mean_twl<N, KIND, HALF>()>;
  cf twr[TH::REGS];
  const cf* tw;
  if constexpr (TH::LDS) {
    tw = twl;
  } else {
    load_twiddles<N, E, 1, const MSIG_GLOBAL cf*, S>(t, as_global(a.twt), twr);
    tw = twr;
  }
  float w[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPF * e;
    w[e] = n < a.nperseg ? as_global(a.win)[n] : 0.f;
  }
  const size_t esz = CPLX ? sizeof(cf) : sizeof(float);
  UnitStream<N, E, CPLX, HALF> sx, sy;
  sx.row = static_cast<const char*>(a.x) + (size_t)row * a.ld_x * esz;
  sx.t = t;
  if constexpr (KIND != K_PSD) {
    sy.row = static_cast<const char*>(a.y) + (size_t)row * a.ld_y * esz;
    sy.t = t;
  }

  float acc[NC][E];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int e = 0; e < E; ++e) acc[c][e] = 0.f;

  // FFT + accumulate of a unit whose x points are detrended and windowed;
  // prep_y() fills zy after zx's FFT (two-signal kinds), so only one
  // signal's windowed points are live at a time.
  auto transform_acc = [&](cf (&zx)[E], cf (&zy)[E], bool active, auto&& prep_y) {
    if constexpr (KIND == K_PSD) {
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zx, tw, lds);
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) acc[0][e] = fmaf(zx[e].im, zx[e].im, fmaf(zx[e].re, zx[e].re, acc[0][e]));
      }
    } else {
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zx, tw, lds);
      group_sync<TPF>();  // last-pass reads of zx done before zy's stores
      prep_y();
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zy, tw, lds);
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          // conj(zx) zy accumulated with four FMAs per bin
          float qr = fmaf(zx[e].re, zy[e].re, acc[K_CSD == KIND ? 0 : 2][e]);
          float qi = fmaf(zx[e].re, zy[e].im, acc[K_CSD == KIND ? 1 : 3][e]);
          qr = fmaf(zx[e].im, zy[e].im, qr);
          qi = fmaf(-zx[e].im, zy[e].re, qi);
          if constexpr (KIND == K_CSD) {
            acc[0][e] = qr;
            acc[1][e] = qi;
          } else {
            acc[0][e] = fmaf(zx[e].im, zx[e].im, fmaf(zx[e].re, zx[e].re, acc[0][e]));
            acc[1][e] = fmaf(zy[e].im, zy[e].im, fmaf(zy[e].re, zy[e].re, acc[1][e]));
            acc[2][e] = qr;
            acc[3][e] = qi;
          }
        }
      }
    }
    group_sync<TPF>();  // reads of this unit's last pass done before next stores
  };

  // one unit: z already taken, detrend + window + FFT + accumulate
  auto unit_body = [&](cf (&zx)[E], cf (&zy)[E], bool active) {
    prep_unit<N, E>(zx, w, a.detrend, a.nperseg, red);
    transform_acc(zx, zy, active, [&] { prep_unit<N, E>(zy, w, a.detrend, a.nperseg, red); });
  };

  if constexpr (HALF) {
    // Units [unit0, u1) of this group; those below nseg / 2 have both
    // segments and run with unconditional loads (the prefetch past the last
    // of them re-reads that unit, in bounds and unused).  Unit u is
    // half-blocks h_2u, h_2u+1, h_2u+2, so consecutive units share a half:
    // its detrend group sum is carried to the next unit instead of being
    // recomputed (each unit sums its two new halves).  The partner-less unit
    // of an odd segment count runs once after them.
    const int u1 = unit0 + a.run < a.units ? unit0 + a.run : a.units;
    const int full = (int)(a.nseg / 2);
    const int uf = u1 < full ? u1 : full;
    const bool has_full = unit0 < uf;
    const bool has_tail = full >= unit0 && full < u1;
    const float inv = 1.0f / (float)a.nperseg;
    const float dinv = a.detrend ? inv : 0.f;  // zero means without detrend
    float sxc = 0.f, syc = 0.f;  // group sums of the carried half (detrend)
    if (has_full) {
      sx.fetch_full(unit0, true);
      if constexpr (KIND != K_PSD) sy.fetch_full(unit0, true);
    }
    prologue();
    if (a.detrend) {
      if (has_full) {
        sxc = sx.first_half_sum();
        if constexpr (KIND != K_PSD) syc = sy.first_half_sum();
      }
      group_sum2<TPF>(sxc, syc, red);  // every group: the same barriers (see below)
    }
    // segment means of the fetched unit of one signal (detrend)
    // Single-signal groups within one wave (TPF <= 64) run the sums whether
    // or not they detrend (means of zero without), which keeps the loop one
    // basic block (welch 64 x 2^20, nperseg 1024: 88.6 -> 87.6 us per
    // call).  The two-signal kinds keep the test: branch-free they ran 6-8%
    // slower at N = 1024.  Groups of several waves synchronise in
    // group_sum2 and must keep the test, so that their barrier sequence
    // matches prep_unit's.
    constexpr bool DALL = TPF <= 64 && KIND == K_PSD;
    auto means = [&](auto rev, const UnitStream<N, E, CPLX, HALF>& st, float& carry, float& mr,
                     float& mi) {
      constexpr bool REV = decltype(rev)::value;
      mr = 0.f;
      mi = 0.f;
      if (DALL || a.detrend) {
        float sb, sc;
        st.template new_half_sums<REV>(sb, sc);
        group_sum2<TPF>(sb, sc, red);
        st.half_means(carry, sb, sc, dinv, mr, mi);
      }
    };
    // welch from N = 1024: each unit's prefetch goes into the half-blocks the
    // unit no longer needs (UnitStream::fetch_next), so the loop runs two
    // units per trip with the roles swapped in the second and copies no
    // half-block (617 -> 602 VALU per pair in the standalone kernel's loop,
    // 127 -> 123 VGPRs; 64 x 2^20: 86.7 -> 85.7 us per call).  The loads then
    // issue after the unit is taken instead of ahead of it, which the short
    // transforms of small N do not cover (N = 256: 67.3 -> 69.5 us), so
    // those keep the copy.  Every welch body of one N (standalone, batch,
    // single-launch batch) runs the same loop, so a batched call keeps the
    // standalone call's bits.  The two-signal kinds keep one unit per trip
    // and the copied carry (fetch_full): unrolled, csd and coherence spilled
    // 54 and 82 VGPRs.
    constexpr bool ROTATE = KIND == K_PSD && N >= 1024;
    if constexpr (ROTATE) {
      auto full_unit = [&](auto rev, int nxt) {
        constexpr bool REV = decltype(rev)::value;
        cf zx[E], zy[E];
        float mrx, mix;
        means(rev, sx, sxc, mrx, mix);
        sx.template take_full_prepped<REV>(zx, w, DALL || a.detrend, mrx, mix);
        sx.template fetch_next<REV>(nxt);
        transform_acc(zx, zy, true, [] {});
      };
      if (has_full) {
        int unit = unit0;
        for (; unit + 1 < uf; unit += 2) {
          full_unit(RevTag<false>{}, unit + 1);
          full_unit(RevTag<true>{}, unit + 2 < uf ? unit + 2 : uf - 1);
        }
        if (unit < uf) full_unit(RevTag<false>{}, uf - 1);
      }
    } else if (has_full) {
      for (int unit = unit0; unit < uf; ++unit) {
        const int nxt = unit + 1 < uf ? unit + 1 : uf - 1;
        cf zx[E], zy[E];
        float mrx, mix;
        means(RevTag<false>{}, sx, sxc, mrx, mix);
        sx.take_full_prepped(zx, w, DALL || a.detrend, mrx, mix);
        sx.fetch_full(nxt, false);
        if constexpr (KIND == K_PSD) {
          transform_acc(zx, zy, true, [] {});
        } else {
          // y's points are built after x's FFT, and y's next loads are issued
          // then, so x's FFT holds one set of y half-blocks, not two (what
          // lets hop == N/2 coherence run two waves per SIMD)
          transform_acc(zx, zy, true, [&] {
            float mry, miy;
            means(RevTag<false>{}, sy, syc, mry, miy);
            sy.take_full_prepped(zy, w, DALL || a.detrend, mry, miy);
            sy.fetch_full(nxt, false);
          });
        }
      }
    }
    if (has_tail) {
      cf zx[E], zy[E];
      sx.fetch(full, true, a);
      sx.take(zx);
      if constexpr (KIND != K_PSD) {
        sy.fetch(full, true, a);
        sy.take(zy);
      }
      unit_body(zx, zy, true);
    }
    if constexpr (TPF > 64 && GPB > 1) {
      // Several multi-wave groups per block (N = 2048), which synchronise with
      // block barriers: every group of a block must run the same sequence of
      // barriers (different sequences in the two groups corrupted results on
      // MI355X).  Each unit above runs one detrend reduction per signal and
      // the FFTs in the same order, so a group with fewer units than the
      // block's first group (the last block of a row) runs inactive units on
      // zeros up to that count.
      const int ran = u1 > unit0 ? u1 - unit0 : 0;
      int most = a.units - bir * GPB * a.run;
      most = most < a.run ? most : a.run;
      for (int k = ran; k < most; ++k) {
        cf zx[E], zy[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          zx[e] = cf{0.f, 0.f};
          zy[e] = cf{0.f, 0.f};
        }
        unit_body(zx, zy, false);
      }
    }
  } else {
  prologue();  // before the loads here: the general stream's first unit is a full segment set
  sx.fetch(unit0, true, a);
  if constexpr (KIND != K_PSD) sy.fetch(unit0, true, a);
  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;
    const bool more = it + 1 < a.run;
    cf zx[E];
    sx.take(zx);
    if constexpr (KIND == K_PSD) {
      if (more) sx.fetch(unit + 1, false, a);
      prep_unit<N, E>(zx, w, a.detrend, a.nperseg, red);
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zx, tw, lds);
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) acc[0][e] = fmaf(zx[e].im, zx[e].im, fmaf(zx[e].re, zx[e].re, acc[0][e]));
      }
    } else {
      cf zy[E];
      sy.take(zy);
      if (more) {
        sx.fetch(unit + 1, false, a);
        sy.fetch(unit + 1, false, a);
      }
      prep_unit<N, E>(zx, w, a.detrend, a.nperseg, red);
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zx, tw, lds);
      group_sync<TPF>();  // last-pass reads of zx done before zy's stores
      prep_unit<N, E>(zy, w, a.detrend, a.nperseg, red);
      fft_passes<N, E, TH::LDS, 1, S, SPLIT>(t, zy, tw, lds);
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          // conj(zx) zy accumulated with four FMAs per bin
          float qr = fmaf(zx[e].re, zy[e].re, acc[K_CSD == KIND ? 0 : 2][e]);
          float qi = fmaf(zx[e].re, zy[e].im, acc[K_CSD == KIND ? 1 : 3][e]);
          qr = fmaf(zx[e].im, zy[e].im, qr);
          qi = fmaf(-zx[e].im, zy[e].re, qi);
          if constexpr (KIND == K_CSD) {
            acc[0][e] = qr;
            acc[1][e] = qi;
          } else {
            acc[0][e] = fmaf(zx[e].im, zx[e].im, fmaf(zx[e].re, zx[e].re, acc[0][e]));
            acc[1][e] = fmaf(zy[e].im, zy[e].im, fmaf(zy[e].re, zy[e].re, acc[1][e]));
            acc[2][e] = qr;
            acc[3][e] = qi;
          }
        }
      }
    }
    group_sync<TPF>();  // reads of this unit's last pass done before next stores
  }
  }

  // block reduction: groups summed in a fixed order
  float* redf = reinterpret_cast<float*>(lds_all);
  MSIG_GLOBAL float* dst = as_global_mut(ws) + ((size_t)row * a.bpr + bir) * NC * N;
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

// Waves per SIMD each kernel is compiled for.  Single-signal kinds fit three
// waves without spills up to N = 1024 (and the hop == N/2 form at 2048,
// whose deduplicated 16 KB twiddle table leaves LDS for three blocks) and
// two above; two-signal kinds hold both spectra and their prefetch buffers
// and run with two (CSD; at N = 4096 only the hop == N/2 form, whose
// twiddles move to LDS) or one (coherence, except the hop == N/2 form: two
// waves with 14-34 spilled VGPRs measured 281 -> 183 us on 64 x 2^20 at
// N = 1024, 364 -> 231 us at 2048 and 362 -> 251 us at 4096 (twiddles in
// LDS there), where one wave left the latencies exposed).
// CPLX: complex input.  The real-input csd on the general stream spilled
// 22 VGPRs at two waves from N = 512 (1024: 389 -> 339 us at one wave over
// 64 x 2^20 with noverlap 3N/4; 512: 335 -> 329).
template <int N, int KIND, bool HALF, bool CPLX = false>
constexpr int mean_waves() {
  if constexpr (N >= 2048) {
    if constexpr (KIND == K_PSD) return (N == 2048 && HALF) ? 3 : 2;
    if constexpr (KIND == K_CSD) return (N < 4096 || HALF) ? 2 : 1;
    return HALF ? 2 : 1;
  }
  if constexpr (KIND == K_CSD && !HALF && !CPLX && N >= 512) return 1;
  if constexpr (KIND == K_PSD && mean_split<N, KIND, HALF, CPLX>()) return 4;
  return KIND == K_PSD ? 3 : (KIND == K_CSD ? 2 : (HALF ? 2 : 1));
}
// The extension stream at N = 1024 spilled at three waves per SIMD; two run
// the stft 16 x 2^20 kernel in 48 us against 53 (at 512, three stay faster).
// The coherence trio on the general stream spilled 105-186 VGPRs at two
// waves from N = 512; one runs it 11% faster (csd frames stay at two).
template <int N, int OK, bool EXT = false, bool HALF = true>
constexpr int frames_waves() {
  if (EXT && N == 1024) return 2;
  if (!HALF && N >= 512 && OK == F_COH) return 1;
  return N >= 2048 ? ((OK == F_COMPLEX || OK == F_POWER) ? 2 : 1)
                   : ((OK == F_COMPLEX || OK == F_POWER) ? 3 : 2);
}

// HALF selects the hop == N/2 register-reuse stream (real input, nperseg == N).
template <int N, int E, int KIND, bool CPLX, bool HALF>
__global__ __launch_bounds__(BT, (mean_waves<N, KIND, HALF, CPLX>())) void k_mean(const KArgs a) {
  using PL = Plan<N, E>;
  using TH = TwiddleHome<N, E, 0, mean_twl<N, KIND, HALF>()>;
  __shared__ cf lds_all[mean_lds_cf<N, KIND, HALF, CPLX>()];
  __shared__ float red[2 * (BT / 64)];
  __shared__ cf twl[TH::LDS_ELEMS];
  mean_body<N, E, KIND, CPLX, HALF>(a, blockIdx.x / a.bpr, blockIdx.x % a.bpr, lds_all, red, twl,
                                    a.ws, threadIdx.x, [&] {
                                      if constexpr (TH::LDS) {
                                        fill_twiddle_lds<N, E>(twl, a.twt);
                                        __syncthreads();
                                      }
                                    });
}

// Per-frame spectra.  Real input: the two packed segments are separated with
// X_s = (Z[k] + conj Z[N-k]) / 2, X_{s+1} = -i (Z[k] - conj Z[N-k]) / 2;
// separate() returns 2 X_s and 2 X_{s+1}, and the callers fold the halving
// into their output scale (a power-of-two factor, so every output keeps its
// bits and the four multiplies per bin go).
// Bins go through LDS unpadded: the stores (bin(t, e), consecutive in t),
// the reads of k and the mirrored reads of N - k all fall on distinct banks
// (8-byte writes serve 16 lanes over 32 four-byte banks, 8-byte reads 32
// lanes over 64), where the pos + pos/16 padding of the exchanges put two
// lanes of every 32 on one bank pair.
template <int N, int E>
__device__ __forceinline__ void separate(int t, const cf* z, cf* lds, cf* x0, cf* x1) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  group_sync<TPF>();
#pragma unroll
  for (int e = 0; e < E; ++e) lds[PL::bin(t, e)] = z[e];
  group_sync<TPF>();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TPF * e;
    const cf p = lds[k];
    const cf q = lds[(N - k) & (N - 1)];
    x0[e] = cf{p.re + q.re, p.im - q.im};
    x1[e] = cf{p.im + q.im, q.re - p.re};
  }
  group_sync<TPF>();
}

template <int N, int E, int OK, bool CPLX, bool HALF, bool EXT = false>
__device__ __forceinline__ void frames_body(const KArgs& a, int64_t row, int bir, cf* lds_all,
                                            float* red, const cf* twl) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  constexpr bool PAIR = (OK == F_CSD || OK == F_COH);
  // real-input frames at 512-2048 exchange one component at a time
  // (fft_passes SPLIT, conflict-free linear slots): no per-element address
  // registers, which ends their spills (2-77 VGPRs at 1024) at the same
  // occupancy (at 4096 the power frames ran 3% slower and stay complex)
  constexpr bool SPLIT = !CPLX && N >= 512 && N <= 2048;
  const int tid = threadIdx.x;
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * PL::LDS;
  const int unit0 = (bir * GPB + g) * a.run;

  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  cf twr[TH::REGS];
  const cf* tw;
  if constexpr (TH::LDS) {
    tw = twl;
  } else {
    load_twiddles<N, E>(t, as_global(a.twt), twr);
    tw = twr;
  }
  float w[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPF * e;
    w[e] = n < a.nperseg ? as_global(a.win)[n] : 0.f;
  }
  const size_t esz = CPLX ? sizeof(cf) : sizeof(float);
  UnitStream<N, E, CPLX, HALF, EXT> sx, sy;
  sx.row = static_cast<const char*>(a.x) + (size_t)row * a.ld_x * esz;
  sx.t = t;
  if constexpr (PAIR) {
    sy.row = static_cast<const char*>(a.y) + (size_t)row * a.ld_y * esz;
    sy.t = t;
  }
  const int nb = a.nbins;
  const float os = a.oscale;

  // Power frames of real input at hop == N/2 (spectrogram psd, the median
  // average), groups of one wave: the mean kernels' full-unit stream.  The
  // units whose two segments both exist take their points without bounds
  // selects and detrend with the carried half-block sums (UnitStream
  // fetch_full / take_full_prepped / half_means; means of zero without
  // detrend, so the loop is one basic block); the partner-less unit of an
  // odd segment count runs once through the general fetch after them.
  constexpr bool FULL = HALF && !EXT && !CPLX && OK == F_POWER && TPF <= 64 && N >= 32;
  if constexpr (FULL) {
   if (nb == N / 2 + 1) {  // onesided
    const int u1 = unit0 + a.run < a.units ? unit0 + a.run : a.units;
    const int full = (int)(a.nseg / 2);
    const int uf = u1 < full ? u1 : full;
    const float osx = 0.25f * os;  // separate() returns 2 X
    const float fa = a.apply_factor ? 2.f : 1.f;
    // |X|^2 of segment s (bins of x = x0 or x1) to frame row * nseg + s:
    // registers e < E/2 hold bins below N/2 on every lane, bin N/2 is
    // register E/2 of lane 0
    auto store = [&](int64_t s, const cf (&x)[E]) {
      const int64_t fr = row * a.nseg + s;
      auto put = [&](int e, int k, float f) {
        static_cast<float*>(a.out0)[(size_t)fr * nb + k] =
            (x[e].re * x[e].re + x[e].im * x[e].im) * (osx * f);
      };
      put(0, t, t == 0 ? 1.f : fa);
#pragma unroll
      for (int e = 1; e < E / 2; ++e) put(e, t + TPF * e, fa);
      if (t == 0) put(E / 2, N / 2, 1.f);
    };
    if (unit0 < uf) {
      const float dinv = a.detrend ? 1.0f / (float)a.nperseg : 0.f;  // zero means without detrend
      sx.fetch_full(unit0, true);
      float carry = sx.first_half_sum(), unused = 0.f;
      group_sum2<TPF>(carry, unused, red);
      for (int unit = unit0; unit < uf; ++unit) {
        const int nxt = unit + 1 < uf ? unit + 1 : uf - 1;
        float sb, sc, mr, mi;
        sx.new_half_sums(sb, sc);
        group_sum2<TPF>(sb, sc, red);
        sx.half_means(carry, sb, sc, dinv, mr, mi);
        cf zx[E];
        sx.take_full_prepped(zx, w, true, mr, mi);
        sx.fetch_full(nxt, false);
        fft_passes<N, E, TH::LDS, 1, 0, SPLIT>(t, zx, tw, lds);
        cf x0[E], x1[E];
        separate<N, E>(t, zx, lds, x0, x1);
        store(2 * (int64_t)unit, x0);
        store(2 * (int64_t)unit + 1, x1);
      }
    }
    if (full >= unit0 && full < u1) {
      cf zx[E];
      sx.fetch(full, true, a);
      sx.take(zx);
      prep_unit<N, E>(zx, w, a.detrend, a.nperseg, red);
      fft_passes<N, E, TH::LDS, 1, 0, SPLIT>(t, zx, tw, lds);
      cf x0[E], x1[E];
      separate<N, E>(t, zx, lds, x0, x1);
      store(2 * (int64_t)full, x0);
      if (2 * (int64_t)full + 1 < a.nseg) store(2 * (int64_t)full + 1, x1);
    }
    return;
   }
  }

  sx.fetch(unit0, true, a);
  if constexpr (PAIR) sy.fetch(unit0, true, a);
  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;
    const bool more = it + 1 < a.run;
    cf zx[E], zy[E];
    sx.take(zx);
    if constexpr (PAIR) sy.take(zy);
    if (more) {
      sx.fetch(unit + 1, false, a);
      if constexpr (PAIR) sy.fetch(unit + 1, false, a);
    }
    prep_unit<N, E>(zx, w, a.detrend, a.nperseg, red);
    fft_passes<N, E, TH::LDS, 1, 0, SPLIT>(t, zx, tw, lds);
    if constexpr (CPLX) {
      // one segment per unit; bins come out in registers, no separation
      if constexpr (PAIR) {
        group_sync<TPF>();
        prep_unit<N, E>(zy, w, a.detrend, a.nperseg, red);
        fft_passes<N, E, TH::LDS, 1, 0, SPLIT>(t, zy, tw, lds);
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
      // 2 X per separated bin: complex outputs scale by os / 2, products
      // (powers, cross spectra) by os / 4
      const float osx = OK == F_COMPLEX ? 0.5f * os : 0.25f * os;
      cf x0[E], x1[E];
      separate<N, E>(t, zx, lds, x0, x1);
      cf y0[E], y1[E];
      if constexpr (PAIR) {
        prep_unit<N, E>(zy, w, a.detrend, a.nperseg, red);
        fft_passes<N, E, TH::LDS, 1, 0, SPLIT>(t, zy, tw, lds);
        separate<N, E>(t, zy, lds, y0, y1);
      }
      // One frame's bins to the outputs: bin k of register e, onesided
      // factor f (kinds other than complex frames).
      auto put = [&](int64_t fr, int h, int e, int k, float f) {
        const size_t o = (size_t)fr * nb + k;
        const cf X = h ? x1[e] : x0[e];
        if constexpr (OK == F_COMPLEX) {
          static_cast<cf*>(a.out0)[o] = cf{X.re * osx, X.im * osx};
        } else if constexpr (OK == F_POWER) {
          static_cast<float*>(a.out0)[o] = (X.re * X.re + X.im * X.im) * (osx * f);
        } else {
          const cf Y = h ? y1[e] : y0[e];
          const cf q = cmulc(X, Y);
          if constexpr (OK == F_CSD) {
            static_cast<cf*>(a.out0)[o] = cf{q.re * osx * f, q.im * osx * f};
          } else {
            static_cast<float*>(a.out0)[o] = (X.re * X.re + X.im * X.im) * (osx * f);
            static_cast<float*>(a.out1)[o] = (Y.re * Y.re + Y.im * Y.im) * (osx * f);
            static_cast<cf*>(a.out2)[o] = cf{q.re * osx * f, q.im * osx * f};
          }
        }
      };
      if (N >= 32 && active && nb == N / 2 + 1) {
        // onesided frames: registers e < E/2 hold bins below N/2 on every
        // lane, and bin N/2 is register E/2 of lane 0, so the bins are
        // stored without per-bin tests and the other registers' separation
        // is never computed (N = 16, one thread per FFT, keeps the generic
        // stores: both forms in one kernel cost it a wave of occupancy)
        const int64_t s0 = 2 * (int64_t)unit;
        const float fa = a.apply_factor ? 2.f : 1.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t s = s0 + h;
          if (s >= a.nseg) break;
          const int64_t fr = row * a.nseg + s;
          put(fr, h, 0, t, t == 0 ? 1.f : fa);
#pragma unroll
          for (int e = 1; e < E / 2; ++e) put(fr, h, e, t + TPF * e, fa);
          if (t == 0) put(fr, h, E / 2, N / 2, 1.f);
        }
      } else if (active) {
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
            float f = 1.f;
            if (a.apply_factor) f = (k == 0 || 2 * k == N) ? 1.f : 2.f;
            put(fr, h, e, k, f);
          }
        }
      }
    }
    group_sync<TPF>();
  }
}

// EXT: rows read through the virtual zero extension (stft's boundary zeros)
template <int N, int E, int OK, bool CPLX, bool HALF, bool EXT = false>
__global__ __launch_bounds__(BT, (frames_waves<N, OK, EXT, HALF>())) void k_frames(const KArgs a) {
  using PL = Plan<N, E>;
  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  __shared__ cf lds_all[(BT / PL::TPF) * PL::LDS];
  __shared__ float red[2 * (BT / 64)];
  __shared__ cf twl[TH::LDS_ELEMS];
  side_copy(a.side_src, a.side_dst, a.side_len);
  if constexpr (TH::LDS) {
    fill_twiddle_lds<N, E>(twl, a.twt);
    __syncthreads();
  }
  frames_body<N, E, OK, CPLX, HALF, EXT>(a, blockIdx.x / a.bpr, blockIdx.x % a.bpr, lds_all, red,
                                         twl);
}


// Inverse transform of one-sided spectra, two frames per complex FFT.
// Frames s and s+1 of a row (X_s, X_{s+1}: nbins_in = N/2 + 1 bins each,
// Hermitian extension X[k] = conj X[N - k] above N/2, the imaginary parts of
// the DC and Nyquist bins ignored as irfft does) are packed as
// Z = X_s + i X_{s+1}; ifft(Z) = conj(fft(conj Z)) / N has real part x_s and
// imaginary part x_{s+1}.  After the transform register e of thread t holds
// sample bin(t, e) of both frames, before the weff (= window * scale / N)
// factor.  Counterpart of the inverse kernel of istft_frames
// (_stft_metal.py:753-791).
template <int N, int E, bool TWL>
__device__ __forceinline__ void inverse_unit(const MSIG_GLOBAL cf* zrow, int nb, int64_t s0,
                                             bool active, bool two, int t, const cf* tw,
                                             cf* lds, cf (&z)[E]) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int M = N / 2;
  // X[k] of one frame at k = t + TPF e: registers e < E/2 hold k < M (read
  // directly), the others k >= M (the conjugate of bin N - k); the
  // imaginary parts of DC (t = 0, e = 0) and Nyquist (t = 0, e = E/2) drop
  auto bin_of = [&](const MSIG_GLOBAL cf* f, int e) -> cf {
    const int k = t + TPF * e;
    if (e < E / 2) {
      const cf c = cf{f[k].re, f[k].im};
      return cf{c.re, (e == 0 && t == 0) ? 0.f : c.im};
    }
    const cf c = cf{f[N - k].re, f[N - k].im};
    return cf{c.re, (e == E / 2 && t == 0) ? 0.f : -c.im};
  };
  const MSIG_GLOBAL cf* f0 = zrow + s0 * nb;
  const MSIG_GLOBAL cf* f1 = f0 + nb;
  if (active) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cf xs = bin_of(f0, e);
      const cf xt = two ? bin_of(f1, e) : cf{0.f, 0.f};
      // conj(X_s + i X_t)
      z[e] = cf{xs.re - xt.im, -(xs.im + xt.re)};
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = cf{0.f, 0.f};
  }
  fft_passes<N, E, TWL>(t, z, tw, lds);
}

// Frames of istft: x_s = Re F * weff, x_{s+1} = -Im F * weff written as
// [rows][nseg][N] float32 for the overlap-add.
template <int N, int E>
__device__ __forceinline__ void istft_body(const KArgs& a, int64_t row, int bir, cf* lds_all,
                                           const cf* twl) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  const int tid = threadIdx.x;
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * PL::LDS;
  const int unit0 = (bir * GPB + g) * a.run;

  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  cf twr[TH::REGS];
  const cf* tw;
  if constexpr (TH::LDS) {
    tw = twl;
  } else {
    load_twiddles<N, E>(t, as_global(a.twt), twr);
    tw = twr;
  }
  float w[E];  // weff at the output samples this thread holds
#pragma unroll
  for (int e = 0; e < E; ++e) w[e] = as_global(a.win)[PL::bin(t, e)];
  const int nb = a.nbins;  // bins per input frame (N/2 + 1)
  const MSIG_GLOBAL cf* zrow = as_global(static_cast<const cf*>(a.x)) + row * a.nseg * nb;
  MSIG_GLOBAL float* orow = as_global_mut(static_cast<float*>(a.out0)) + row * a.nseg * N;

  // Every group runs a.run units (those past the row's end transform zeros
  // and write nothing), so the groups of a block run the same barriers.
  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;  // group-uniform
    const int64_t s0 = 2 * (int64_t)unit;
    const bool two = active && s0 + 1 < a.nseg;
    cf z[E];  // FFT input position t + TPF e
    inverse_unit<N, E, TH::LDS>(zrow, nb, s0, active, two, t, tw, lds, z);
    // sample n = bin(t, e): x_s = Re F / N, x_t = -Im F / N (1/N folded in w)
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = PL::bin(t, e);
        orow[s0 * N + n] = z[e].re * w[e];
        if (two) orow[(s0 + 1) * N + n] = -z[e].im * w[e];
      }
    }
    group_sync<TPF>();
  }
}

// istft with the overlap-add fused, for hop = N/2.  Frame s covers
// half-blocks s and s+1 of the output (M = N/2 samples each), so half-block
// j is frame j-1's second half plus frame j's first half.  A group walks
// units u (frames 2u, 2u+1) in order: it writes half-blocks 2u (with the
// second half of frame 2u-1 carried in registers from the previous unit)
// and 2u+1, and carries frame 2u+1's second half; the row's last frame also
// writes the final half-block.  A group first transforms the unit before
// its own for that carry (every group does, on zeros at a row's start, so
// the groups of a block run the same barriers).  Register e of thread t
// holds sample t + TPF k with k = (e / R) + (E / R) (e % R) (R the last
// pass's radix), so a thread's first-half samples are k < E/2 and sample
// r + M is k + E/2 of the same thread.
//
// The overlap-add's divide by the window power (k_overlap_add: norm(i) =
// sum over the covering frames of win^2, taken where it exceeds 1e-10) is
// folded into the per-sample weights once per thread: a half-block covered
// by two frames has norm fma(win[r], win[r], fma(win[M + r], win[M + r], 0))
// at every j, so sample n of a frame weighs weff[n] / norm(n mod M), and
// the output is one multiply and one FMA of the two transforms' outputs.
// The row's two end half-blocks (one frame each) divide by their own norm.
// Written are samples first <= i < first + count, as out[row][i - first].
// KArgs: x = spectra, win = weff, y = window, out0 = output,
// ola_first / ola_count.
template <int N, int E>
__device__ __forceinline__ void istft_ola_body(const KArgs& a, int64_t row, int bir, cf* lds_all,
                                               const cf* twl) {
  using PL = Plan<N, E>;
  constexpr int TPF = PL::TPF;
  constexpr int GPB = BT / TPF;
  constexpr int M = N / 2;
  constexpr int H = E / 2;
  constexpr int R = PL::radix(PL::NP - 1);
  constexpr int ER = E / R;
  const int tid = threadIdx.x;
  const int g = tid / TPF, t = tid % TPF;
  cf* lds = lds_all + g * PL::LDS;
  const int unit0 = (bir * GPB + g) * a.run;

  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  cf twr[TH::REGS];
  const cf* tw;
  if constexpr (TH::LDS) {
    tw = twl;
  } else {
    load_twiddles<N, E>(t, as_global(a.twt), twr);
    tw = twr;
  }
  // register of sample t + TPF k
  auto reg = [](int k) { return (k % ER) * R + k / ER; };
  const MSIG_GLOBAL float* wf = as_global(a.win) + t;
  const MSIG_GLOBAL float* wn = as_global(static_cast<const float*>(a.y)) + t;
  float wi[E];  // weight of sample t + TPF k in a half-block covered by two frames
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float lo = wn[TPF * k], hi = wn[M + TPF * k];
    const float nrm = fmaf(lo, lo, fmaf(hi, hi, 0.f));
    const float d = nrm > 1e-10f ? nrm : 1.f;
    wi[k] = wf[TPF * k] / d;
    wi[k + H] = wf[M + TPF * k] / d;
  }
  // weight of sample t + TPF k in an end half-block (one frame)
  auto w_end = [&](int k) {
    const float v = wn[TPF * k];
    const float nrm = fmaf(v, v, 0.f);
    return wf[TPF * k] / (nrm > 1e-10f ? nrm : 1.f);
  };
  const int nb = a.nbins;
  const MSIG_GLOBAL cf* zrow = as_global(static_cast<const cf*>(a.x)) + row * a.nseg * nb;
  MSIG_GLOBAL float* orow = as_global_mut(static_cast<float*>(a.out0)) + row * a.ola_count;
  const int64_t first = a.ola_first, last = a.ola_first + a.ola_count;
  auto put = [&](int64_t j, int k, float v) {
    const int64_t i = j * M + t + TPF * k;
    if (i >= first && i < last) orow[i - first] = v;
  };

  float carry[H];  // frame 2u-1's second half (transform output) at r = t + TPF k
  {
    // lead-in: the unit before this group's first one, for the carry
    const int unit = unit0 - 1;
    const bool active = unit >= 0 && unit < a.units;
    const int64_t s0 = 2 * (int64_t)unit;
    const bool two = active && s0 + 1 < a.nseg;
    cf z[E];
    inverse_unit<N, E, TH::LDS>(zrow, nb, s0, active, two, t, tw, lds, z);
#pragma unroll
    for (int k = 0; k < H; ++k) carry[k] = -z[reg(k + H)].im;
    group_sync<TPF>();
  }
  for (int it = 0; it < a.run; ++it) {
    const int unit = unit0 + it;
    const bool active = unit < a.units;  // group-uniform
    const int64_t s0 = 2 * (int64_t)unit;
    const bool two = active && s0 + 1 < a.nseg;
    cf z[E];
    inverse_unit<N, E, TH::LDS>(zrow, nb, s0, active, two, t, tw, lds, z);
    if (active) {
      // x_s = Re F and x_{s+1} = -Im F (1/N and the window in the weights)
      if (s0 >= 1) {
#pragma unroll
        for (int k = 0; k < H; ++k)
          put(s0, k, fmaf(carry[k], wi[k + H], z[reg(k)].re * wi[k]));
      } else {
#pragma unroll
        for (int k = 0; k < H; ++k) put(s0, k, z[reg(k)].re * w_end(k));
      }
      if (two) {
#pragma unroll
        for (int k = 0; k < H; ++k)
          put(s0 + 1, k, fmaf(z[reg(k + H)].re, wi[k + H], -z[reg(k)].im * wi[k]));
#pragma unroll
        for (int k = 0; k < H; ++k) carry[k] = -z[reg(k + H)].im;
        if (unit == a.units - 1) {
#pragma unroll
          for (int k = 0; k < H; ++k) put(s0 + 2, k, carry[k] * w_end(k + H));
        }
      } else {
#pragma unroll
        for (int k = 0; k < H; ++k) put(s0 + 1, k, z[reg(k + H)].re * w_end(k + H));
      }
    }
    group_sync<TPF>();
  }
}

template <int N, int E>
__global__ __launch_bounds__(BT) void k_istft(const KArgs a) {
  using PL = Plan<N, E>;
  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  __shared__ cf lds_all[(BT / PL::TPF) * PL::LDS];
  __shared__ cf twl[TH::LDS_ELEMS];
  if constexpr (TH::LDS) {
    fill_twiddle_lds<N, E>(twl, a.twt);
    __syncthreads();
  }
  istft_body<N, E>(a, blockIdx.x / a.bpr, blockIdx.x % a.bpr, lds_all, twl);
}

template <int N, int E>
__global__ __launch_bounds__(BT, (N >= 4096 ? 1 : 2)) void k_istft_ola(const KArgs a) {
  using PL = Plan<N, E>;
  using TH = TwiddleHome<N, E, 0, frames_twl<N>()>;
  __shared__ cf lds_all[(BT / PL::TPF) * PL::LDS];
  __shared__ cf twl[TH::LDS_ELEMS];
  side_copy(a.side_src, a.side_dst, a.side_len);
  if constexpr (TH::LDS) {
    fill_twiddle_lds<N, E>(twl, a.twt);
    __syncthreads();
  }
  istft_ola_body<N, E>(a, blockIdx.x / a.bpr, blockIdx.x % a.bpr, lds_all, twl);
}

// One entry of a ragged batch's work queue: block `blk` (row * bpr + bir) of
// job `job`.  Entries of one FFT-size class are walked by a persistent grid.
struct WorkItem {
  int32_t job;
  int32_t blk;
};

// Ragged batch of mean jobs that share (N, KIND, CPLX, HALF): a persistent
// grid strides over the class's work queue; every entry runs exactly the
// block that a standalone launch of its job would run, so each job's result
// is the same alone and in the batch.
// The batch kernels run at the standalone kernels' occupancy (measured
// faster on ragged batches than one wave fewer), except where the job
// lookup's extra registers make the standalone bounds spill heavily: at
// N = 2048 (9-142 VGPRs) they keep two waves for welch and one for the
// two-signal kinds, and coherence keeps one wave everywhere (47-88 VGPRs at
// two).
template <int N, int KIND, bool HALF>
constexpr int batch_waves() {
  if constexpr (N == 2048) return KIND == K_PSD ? 2 : 1;
  if constexpr (KIND == K_COH) return 1;
  return mean_waves<N, KIND, HALF>();
}

template <int N, int E, int KIND, bool CPLX, bool HALF, int S>
__global__ __launch_bounds__(BT, (batch_waves<N, KIND, HALF>())) void k_mean_batch(
    const KArgs* __restrict__ jobs, const WorkItem* __restrict__ queue, int32_t nwork,
    float* ws_base) {
  using TH = TwiddleHome<N, E, S, mean_twl<N, KIND, HALF>()>;
  __shared__ cf lds_all[mean_lds_cf<N, KIND, HALF, CPLX>()];
  __shared__ float red[2 * (BT / 64)];
  __shared__ cf twl[TH::LDS_ELEMS];
  bool filled = false;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem item = queue[w];
    const KArgs& a = jobs[item.job];
    float* ws = ws_base + a.ws_off;
    if constexpr (TH::LDS) {
      if (!filled) {  // every job of a class has the same N: fill once
        fill_twiddle_lds<N, E, S>(twl, a.twt);
        filled = true;
      }
    }
    __syncthreads();
    if constexpr (HALF) {
      mean_body<N, E, KIND, false, true, S>(a, item.blk / a.bpr, item.blk % a.bpr, lds_all, red, twl,
                                         ws, threadIdx.x);
    } else {
      mean_body<N, E, KIND, CPLX, false, S>(a, item.blk / a.bpr, item.blk % a.bpr, lds_all, red, twl,
                                         ws, threadIdx.x);
    }
    __syncthreads();
  }
}

// ---- one persistent launch for a whole ragged batch ----
// The work items of every welch and csd job on real input with hop = N/2
// (any N) share one kernel: each item names its job, the job its class
// (N, kind), and a switch runs that class's mean body, the same block a
// standalone launch would run.  Items are taken in queue order from a 64-bit
// ticket counter: a run makes exactly nwork + gridDim.x fetches (every block
// stops at its first ticket past the work), so ticket mod (nwork + gridDim.x)
// is the run's own index and the counter never needs resetting.  LDS is one
// pool sized for the largest class; a class change refills its twiddle table.
// Classes are 2 log2(N / 16) + (kind == csd).
constexpr int RAGGED_CLASSES = 18;
constexpr int ragged_class(int N, int kind) {
  int l = 0;
  while ((16 << l) < N) ++l;
  return 2 * l + (kind == K_CSD ? 1 : 0);
}
template <int N, int KIND>
constexpr int ragged_pool() {
  using PL = Plan<N, (N < 16 ? N : 16)>;
  using TH = TwiddleHome<N, (N < 16 ? N : 16), 0, mean_twl<N, KIND, true>()>;
  return (BT / PL::TPF) * PL::LDS + (TH::LDS ? TH::LDS_ELEMS : 0);
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int RAGGED_POOL = cmax(cmax(ragged_pool<4096, K_CSD>(), ragged_pool<4096, K_PSD>()),
                                 cmax(ragged_pool<2048, K_CSD>(), ragged_pool<1024, K_CSD>()));

template <int N, int KIND>
__device__ __forceinline__ void ragged_item(const KArgs& a, int blk, bool refill, cf* pool,
                                            float* red, float* ws_base, int tid) {
  constexpr int E = N < 16 ? N : 16;
  using PL = Plan<N, E>;
  using TH = TwiddleHome<N, E, 0, mean_twl<N, KIND, true>()>;
  static_assert(ragged_pool<N, KIND>() <= RAGGED_POOL, "LDS pool too small");
  cf* twl = pool + (BT / PL::TPF) * PL::LDS;
  if constexpr (TH::LDS) {
    if (refill) {
      fill_twiddle_lds<N, E>(twl, a.twt);
      __syncthreads();
    }
  }
  mean_body<N, E, KIND, false, true>(a, blk / a.bpr, blk % a.bpr, pool, red, twl,
                                     ws_base + a.ws_off, tid);
}

}  // namespace msig
