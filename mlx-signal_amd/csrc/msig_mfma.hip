// Split-precision MFMA form of the fused welch mean for nfft = nperseg = 256.
//
// The 256-point DFT is factored 16 x 16 (n = 16 n1 + n2, k = k1 + 16 k2):
//   Y[n2][k1]  = sum_n1 z[16 n1 + n2] W16^(n1 k1)          (stage 1, matrix cores)
//   Y'[n2][k1] = Y[n2][k1] W256^(n2 k1)                   (twiddle, FP32 VALU)
//   Z[k1][k2]  = sum_n2 Y'[n2][k1] W16^(n2 k2)            (stage 2, matrix cores)
// Each stage is one 16 x 32 by 32 x 16 product per real output part
// (v_mfma_f32_16x16x32_{f16,bf16}): the K index interleaves the sum index with
// the real/imaginary part of the complex operand, k = 8g + 2r + part for sum
// index 4g + r, so the complex product folds into one real product of twice
// the depth.  That interleave also makes stage 1's accumulator layout (lane
// 16g + c holds rows 4g..4g+3 of column c) exactly the A-operand layout stage
// 2 reads (lane 16g + c holds row c, depth entries 8g..8g+7), so no lane
// exchange sits between the stages.
//
// Every operand is carried as a hi + lo pair of 16-bit values (hi = v rounded,
// lo = v - hi rounded) and each real product is hi*hi + hi*lo + lo*hi with
// FP32 accumulation; the dropped lo*lo term is below 2^-22 (FP16) or 2^-16
// (BF16) of the product.  FP16 has five exponent bits, so each segment pair is
// scaled by a power of two that puts its largest windowed sample in
// [2^10, 2^11) (stage-1 sums stay below 2^15) and the power is unscaled
// exactly in FP32.
//
// The segment pair is packed two-for-one (z = x_s w + i x_{s+1} w) exactly as
// in the FP32 kernels, and the block partials have the same [rows][bpr][256]
// layout, so the FP64 finalize (mirror fold, scale, 1/nseg, onesided factor)
// is shared.  Counterpart of rfft_power_mean (_stft_metal.py:530-557) for the
// precision study of the 256-point configuration.
#include "msig_launch.hpp"

namespace msig {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef __bf16 bh2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// hi = v rounded to 16 bits, lo = (v - hi) rounded; eight values as four
// packed pairs, one VGPR per pair.
template <class V, class P>
__device__ __forceinline__ void split8(const float (&v)[8], V& hi, V& lo) {
  P h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f2 x = {v[2 * q], v[2 * q + 1]};
    h[q] = __builtin_convertvector(x, P);
    const f2 back = __builtin_convertvector(h[q], f2);
    const f2 d = {x[0] - back[0], x[1] - back[1]};
    l[q] = __builtin_convertvector(d, P);
  }
  hi = __builtin_shufflevector(__builtin_shufflevector(h[0], h[1], 0, 1, 2, 3),
                               __builtin_shufflevector(h[2], h[3], 0, 1, 2, 3), 0, 1, 2, 3, 4,
                               5, 6, 7);
  lo = __builtin_shufflevector(__builtin_shufflevector(l[0], l[1], 0, 1, 2, 3),
                               __builtin_shufflevector(l[2], l[3], 0, 1, 2, 3), 0, 1, 2, 3, 4,
                               5, 6, 7);
}

template <int PREC>
struct SplitOps;

template <>
struct SplitOps<PREC_F16> {
  using V = h8;
  static constexpr bool kScale = true;  // five exponent bits: scale each pair
  static __device__ __forceinline__ void split(const float (&v)[8], V& hi, V& lo) {
    split8<V, h2>(v, hi, lo);
  }
  static __device__ __forceinline__ f4 mma(V a, V b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <>
struct SplitOps<PREC_BF16> {
  using V = b8;
  static constexpr bool kScale = false;  // FP32's exponent range
  static __device__ __forceinline__ void split(const float (&v)[8], V& hi, V& lo) {
    split8<V, bh2>(v, hi, lo);
  }
  static __device__ __forceinline__ f4 mma(V a, V b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

// hi*hi + hi*lo + lo*hi, smallest terms first
template <class S>
__device__ __forceinline__ f4 split_mma(const typename S::V& ah, const typename S::V& al,
                                        const typename S::V& bh, const typename S::V& bl) {
  f4 d = {0.f, 0.f, 0.f, 0.f};
  d = S::mma(al, bh, d);
  d = S::mma(ah, bl, d);
  return S::mma(ah, bh, d);
}

__device__ __forceinline__ float wave_max(float a) {
  a = fmaxf(a, dpp_f<0xB1>(a));
  a = fmaxf(a, dpp_f<0x4E>(a));
  a = fmaxf(a, dpp_f<0x141>(a));
  a = fmaxf(a, dpp_f<0x140>(a));
  a = fmaxf(a, swz16_f(a));
  return fmaxf(lane_f(a, 0), lane_f(a, 32));
}

constexpr int MFMA_WAVES = BT / 64;

template <int PREC>
struct MfmaLane {
  using S = SplitOps<PREC>;
  typename S::V bre_h, bre_l, bim_h, bim_l;  // DFT16 operands (constant)
  cf t2[4];                                  // W256^(n2 k1), n2 = 4g + r, k1 = c
  float win[4];                              // window at n = 64g + 16r + c
  float acc[4];                              // |Z|^2 at bins 4g + r + 16c
  float wbound;                              // 2 max|w| over the window
  int detrend;

  // One segment pair: x[2r] = x_s(n_r), x[2r+1] = x_{s+1}(n_r).  PARTNER is
  // false for the last segment of an odd count, whose partner is zero.
  template <bool PARTNER>
  __device__ __forceinline__ void pair(const float (&x)[8]) {
    float w0[4], w1[4];  // per-sample multipliers of the two segments
    float m0 = 0.f, m1 = 0.f;
    if (detrend) {
      m0 = (x[0] + x[2]) + (x[4] + x[6]);
      m1 = (x[1] + x[3]) + (x[5] + x[7]);
      m0 = lanes_sum<64>(m0);
      m1 = lanes_sum<64>(m1);
      m0 *= 1.f / MFMA_NFFT;
      m1 *= 1.f / MFMA_NFFT;
    }
    float unscale = 1.f;
    if constexpr (S::kScale) {
      // |(x - m) w| <= 2 max|x| max|w| (|m| <= max|x|): a power of two that
      // puts this bound for the pair in [2^10, 2^11)
      float b = fmaxf(fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3]))),
                      fmaxf(fmaxf(fabsf(x[4]), fabsf(x[5])), fmaxf(fabsf(x[6]), fabsf(x[7]))));
      b = wave_max(b) * wbound;
      int e = 11 - __builtin_amdgcn_frexp_expf(b);
      e = e < -40 ? -40 : (e > 40 ? 40 : e);
      const float sc = __builtin_ldexpf(1.f, e);
      unscale = __builtin_ldexpf(1.f, -2 * e);
#pragma unroll
      for (int r = 0; r < 4; ++r) w0[r] = win[r] * sc;
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) w0[r] = win[r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) w1[r] = PARTNER ? w0[r] : 0.f;
    float v[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[2 * r] = (x[2 * r] - m0) * w0[r];
      v[2 * r + 1] = (x[2 * r + 1] - m1) * w1[r];
    }
    // stage 1
    typename S::V ah, al;
    S::split(v, ah, al);
    const f4 yre = split_mma<S>(ah, al, bre_h, bre_l);
    const f4 yim = split_mma<S>(ah, al, bim_h, bim_l);
    // twiddle, laid out as stage 2's A operand (depth entry 2r + part)
    float y[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const cf z = cmul(cf{yre[r], yim[r]}, t2[r]);
      y[2 * r] = z.re;
      y[2 * r + 1] = z.im;
    }
    // stage 2
    S::split(y, ah, al);
    const f4 zre = split_mma<S>(ah, al, bre_h, bre_l);
    const f4 zim = split_mma<S>(ah, al, bim_h, bim_l);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if constexpr (S::kScale)
        acc[r] = fmaf(fmaf(zre[r], zre[r], zim[r] * zim[r]), unscale, acc[r]);
      else
        acc[r] = fmaf(zim[r], zim[r], fmaf(zre[r], zre[r], acc[r]));
    }
  }
};

template <int PREC>
__global__ __launch_bounds__(BT) void k_mean_mfma256(KArgs a) {
  using S = SplitOps<PREC>;
  constexpr int N = MFMA_NFFT;
  __shared__ float red[MFMA_WAVES][N];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int64_t row = blockIdx.x / a.bpr;
  const int bir = blockIdx.x - (int)(row * a.bpr);
  const MSIG_GLOBAL cf* tw = as_global(a.twt);

  MfmaLane<PREC> L;
  L.detrend = a.detrend;
  // B operands: depth entry j = 2r + part <-> sum index 4g + r, column c.
  // W16^m = tw[16 m] = (cos, -sin) of 2 pi m / 16, and
  // sum (p + i q)(cos - i sin) = (p cos + q sin) + i (q cos - p sin).
  {
    float bre[8], bim[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int m = ((4 * g + (j >> 1)) * c) & 15;
      const float cs = tw[16 * m].re, sn = -tw[16 * m].im;
      bre[j] = (j & 1) ? sn : cs;
      bim[j] = (j & 1) ? cs : -sn;
    }
    S::split(bre, L.bre_h, L.bre_l);
    S::split(bim, L.bim_h, L.bim_l);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = ((4 * g + r) * c) & (N - 1);
    L.t2[r] = cf{tw[e].re, tw[e].im};
    L.win[r] = as_global(a.win)[64 * g + 16 * r + c];
    L.acc[r] = 0.f;
  }
  L.wbound = 2.f * wave_max(fmaxf(fmaxf(fabsf(L.win[0]), fabsf(L.win[1])),
                                  fmaxf(fabsf(L.win[2]), fabsf(L.win[3]))));

  const MSIG_GLOBAL float* xrow =
      as_global(reinterpret_cast<const float*>(a.x)) + row * a.ld_x + 64 * g + c;
  const int64_t hop = a.hop;
  // Units [u0, u1) of this wave; units below `full` have both segments.  u is
  // wave-uniform, and the loads of a full unit are unconditional, so the
  // loads of the next unit stay in flight while this one is transformed.
  const int full = (int)(a.nseg / 2);
  const int u0 = (bir * MFMA_WAVES + wave) * a.run;
  const int u1 = u0 + a.run < a.units ? u0 + a.run : a.units;
  const int uf = u1 < full ? u1 : full;
  auto fetch = [&](int u, float (&v)[8]) {
    const MSIG_GLOBAL float* p0 = xrow + 2 * (int64_t)u * hop;
    const MSIG_GLOBAL float* p1 = p0 + hop;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[2 * r] = p0[16 * r];
      v[2 * r + 1] = p1[16 * r];
    }
  };
  float xa[8], xb[8];
  if (u0 < uf) {
    fetch(u0, xa);
    int u = u0;
    for (; u + 1 < uf; u += 2) {
      fetch(u + 1, xb);
      L.template pair<true>(xa);
      fetch(u + 2 < uf ? u + 2 : uf - 1, xa);  // past the end: a valid unit, unused
      L.template pair<true>(xb);
    }
    if (u < uf) L.template pair<true>(xa);
  }
  if (full >= u0 && full < u1) {  // unit `full` of an odd segment count: segment 2 full alone
    const MSIG_GLOBAL float* p0 = xrow + 2 * (int64_t)uf * hop;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      xa[2 * r] = p0[16 * r];
      xa[2 * r + 1] = 0.f;
    }
    L.template pair<false>(xa);
  }

  // lane 16g + c holds bins k1 + 16 k2 with k1 = 4g + r, k2 = c
#pragma unroll
  for (int r = 0; r < 4; ++r) red[wave][4 * g + r + 16 * c] = L.acc[r];
  __syncthreads();
  const int t = threadIdx.x;  // BT == N: one bin per thread, waves summed in order
  float s = red[0][t];
#pragma unroll
  for (int w = 1; w < MFMA_WAVES; ++w) s += red[w][t];
  as_global_mut(a.ws)[((size_t)row * a.bpr + bir) * N + t] = s;
}

static_assert(BT == MFMA_NFFT, "one bin per thread in the block reduction");

hipError_t launch_mean_mfma256(int prec, const KArgs& a, int64_t rows, hipStream_t s) {
  const dim3 grid((unsigned)(rows * a.bpr)), block(BT);
  hipError_t e;
  if (prec == PREC_F16)
    e = launch_timed(&k_mean_mfma256<PREC_F16>, grid, block, s, a);
  else if (prec == PREC_BF16)
    e = launch_timed(&k_mean_mfma256<PREC_BF16>, grid, block, s, a);
  else
    return hipErrorInvalidValue;
  return e != hipSuccess ? e : hipGetLastError();
}

}  // namespace msig
