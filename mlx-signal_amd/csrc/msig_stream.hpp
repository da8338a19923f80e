// Kernel arguments and the prefetching unit stream, host-compilable so the
// CPU emulator (tests/native/fft_emu.cpp) drives the same code the kernels do.
#pragma once

#include <stdint.h>

#include "msig_fft.hpp"

namespace msig {

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
  int32_t nbins;     // bins written per output row / frame
  int32_t apply_factor;
  float oscale;      // frames: power scale or sqrt(scale) for complex output
  int64_t ws_off;    // batch plans: this job's partials start at ws_base + ws_off
  int64_t lead;      // zero samples virtually in front of each row (frames kinds)
  int64_t nvalid;    // samples stored per row; the extended row reads zero past them
  int64_t ola_first, ola_count;  // fused istft: output samples [first, first + count)
  const float* side_src;  // side copy, grid-strided over all blocks: side_len floats to side_dst
  float* side_dst;
  int64_t side_len;
  int32_t cls;       // single-launch batch plans: the job's (N, kind) class
};

// Raw samples of consecutive units of one row, fetched one unit ahead so the
// HBM loads of unit u+1 are in flight while unit u is transformed.
//
// General mode: each unit loads its own segment(s) (overlap re-reads hit
// L1/L2).  HALF mode (real input, nperseg == N, hop == N/2): segments s and
// s+1 of a pair are half-blocks h_{2u}, h_{2u+1} and h_{2u+1}, h_{2u+2}, and
// h_{2u+2} is also the first half of the next pair, so each sample is loaded
// into registers exactly once per group.  Units past the row's end, and the
// missing partner of an odd segment count, load zeros.
//
// EXT: virtual zero extension (the complex per-frame kernel, KArgs::lead /
// nvalid): sample p of the extended row is stored sample p - lead, and reads
// as zero outside [0, nvalid).  Fetches whose span lies inside the stored
// samples take the unmasked loads; only the spans that cross an edge test
// each sample.  EXT = false compiles the plain fetches, so the other kernels
// carry none of this.
template <int N, int E, bool CPLX, bool HALF, bool EXT = false>
struct UnitStream {
  static_assert(!(CPLX && HALF), "half-hop reuse is for real input");
  static constexpr int TPF = N / E;
  static constexpr int H = E / 2;
  const char* row;
  int t;
  bool pair_ok;  // HALF: segment s0+1 of the fetched unit exists
  cf raw[HALF ? 1 : E];
  float h0[HALF ? H : 1], h1[HALF ? H : 1], h2[HALF ? H : 1];

  MSIG_HD void half_load(float* dst, int64_t h, bool valid) {
    const MSIG_GLOBAL float* p = as_global(reinterpret_cast<const float*>(row)) + h * (N / 2);
    if (valid) {
#pragma unroll
      for (int e = 0; e < H; ++e) dst[e] = p[t + TPF * e];
    } else {
#pragma unroll
      for (int e = 0; e < H; ++e) dst[e] = 0.f;
    }
  }

  // Offsets [lo, hi) of a span of len samples starting at stored index start
  // (uniform across the group) that are stored samples.
  static MSIG_HD void stored_window(int64_t start, int len, int64_t nvalid, int& lo, int& hi) {
    const int64_t a = -start, b = nvalid - start;
    lo = (int)(a < 0 ? 0 : (a > len ? len : a));
    hi = (int)(b < 0 ? 0 : (b > len ? len : b));
  }

  // half_load on the virtually extended row (stored samples [0, nvalid) at
  // extended positions [lead, lead + nvalid))
  MSIG_HD void half_load_ext(float* dst, int64_t h, bool valid, int64_t lead, int64_t nvalid) {
    const int64_t start = h * (N / 2) - lead;  // stored index of the half's first sample
    const MSIG_GLOBAL float* p = as_global(reinterpret_cast<const float*>(row)) + start;
    int lo, hi;
    stored_window(start, N / 2, nvalid, lo, hi);
    if (valid && lo == 0 && hi == N / 2) {
#pragma unroll
      for (int e = 0; e < H; ++e) dst[e] = p[t + TPF * e];
      return;
    }
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int o = t + TPF * e;
      dst[e] = (valid && o >= lo && o < hi) ? p[o] : 0.f;
    }
  }

  // first == true: nothing carried from a previous fetch
  MSIG_HD void fetch(int unit, bool first, const KArgs& a) {
    const bool active = unit < a.units;
    if constexpr (HALF) {
      const int64_t s0 = 2 * (int64_t)unit;
      const bool v1 = active && (s0 + 1 < a.nseg);
      pair_ok = v1;
      if (first) {
        if constexpr (EXT) half_load_ext(h0, s0, active, a.lead, a.nvalid);
        else half_load(h0, s0, active);
      } else {
#pragma unroll
        for (int e = 0; e < H; ++e) h0[e] = active ? h2[e] : 0.f;
      }
      if constexpr (EXT) {
        half_load_ext(h1, s0 + 1, active, a.lead, a.nvalid);
        half_load_ext(h2, s0 + 2, v1, a.lead, a.nvalid);
      } else {
        half_load(h1, s0 + 1, active);
        half_load(h2, s0 + 2, v1);
      }
    } else if constexpr (!CPLX) {
      const MSIG_GLOBAL float* xrow = as_global(reinterpret_cast<const float*>(row));
      const int64_t s0 = 2 * (int64_t)unit;
      const bool v1 = active && (s0 + 1 < a.nseg);
      const int64_t lead = EXT ? a.lead : 0;
      const int64_t i0 = s0 * a.hop - lead;  // stored index of segment s0's first sample
      const MSIG_GLOBAL float* p0 = xrow + i0;
      const MSIG_GLOBAL float* p1 = p0 + a.hop;
      const bool inside = !EXT || (i0 >= 0 && i0 + a.hop + a.nperseg <= a.nvalid);
      if (a.nperseg == N && v1 && inside) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          raw[e].re = p0[t + TPF * e];
          raw[e].im = p1[t + TPF * e];
        }
      } else if (inside) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int n = t + TPF * e;
          const bool in = n < a.nperseg;
          raw[e].re = (active && in) ? p0[n] : 0.f;
          raw[e].im = (v1 && in) ? p1[n] : 0.f;
        }
      } else {
        int loa, hia, lob, hib;
        stored_window(i0, a.nperseg, a.nvalid, loa, hia);
        stored_window(i0 + a.hop, a.nperseg, a.nvalid, lob, hib);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int n = t + TPF * e;
          raw[e].re = (active && n >= loa && n < hia) ? p0[n] : 0.f;
          raw[e].im = (v1 && n >= lob && n < hib) ? p1[n] : 0.f;
        }
      }
    } else {
      const int64_t i0 = (int64_t)unit * a.hop - (EXT ? a.lead : 0);
      const MSIG_GLOBAL cf* p0 = as_global(reinterpret_cast<const cf*>(row)) + i0;
      int lo = 0, hi = a.nperseg;
      if constexpr (EXT) stored_window(i0, a.nperseg, a.nvalid, lo, hi);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = t + TPF * e;
        if (active && n < a.nperseg && (!EXT || (n >= lo && n < hi))) {
          raw[e].re = p0[n].re;
          raw[e].im = p0[n].im;
        } else {
          raw[e] = cf{0.f, 0.f};
        }
      }
    }
  }

  // HALF: a unit whose two segments both exist (unit < nseg / 2).  No bounds
  // selects; `first` loads the shared half-block, else it is carried.
  MSIG_HD void fetch_full(int unit, bool first) {
    static_assert(HALF, "full-unit fetch is the half-hop stream's");
    const int64_t s0 = 2 * (int64_t)unit;
    if (first) {
      half_load(h0, s0, true);
    } else {
#pragma unroll
      for (int e = 0; e < H; ++e) h0[e] = h2[e];
    }
    half_load(h1, s0 + 1, true);
    half_load(h2, s0 + 2, true);
    pair_ok = true;
  }

  // HALF, the rotating form of the full-unit loop.  A unit's three
  // half-blocks are (first, h1, last) with first = h0 and last = h2 when REV
  // is false and the two swapped when it is true.  fetch_next<REV> loads the
  // next unit into h1 and into `first`, whose half is dead once the unit is
  // taken, so the next unit is (last, h1, first): REV alternates from unit
  // to unit and the shared half-block stays in its registers instead of
  // being copied (a loop unrolled by two names every array statically).
  template <bool REV>
  MSIG_HD float* first_half() {
    return REV ? h2 : h0;
  }
  template <bool REV>
  MSIG_HD const float* first_half() const {
    return REV ? h2 : h0;
  }
  template <bool REV>
  MSIG_HD const float* last_half() const {
    return REV ? h0 : h2;
  }
  template <bool REV>
  MSIG_HD void fetch_next(int unit) {
    static_assert(HALF, "full-unit fetch is the half-hop stream's");
    const int64_t s0 = 2 * (int64_t)unit;
    half_load(h1, s0 + 1, true);
    half_load(first_half<REV>(), s0 + 2, true);
    pair_ok = true;
  }

  // HALF: this thread's partial sum of the fetched unit's first half-block
  // (the one it shares with the previous unit).
  MSIG_HD float first_half_sum() const {
    static_assert(HALF, "half-block sums are the half-hop stream's");
    float s = h0[0];
#pragma unroll
    for (int e = 1; e < H; ++e) s += h0[e];
    return s;
  }

  // HALF: this thread's partial sums of the fetched unit's two new
  // half-blocks (h1 and the last one).
  template <bool REV = false>
  MSIG_HD void new_half_sums(float& pb, float& pc) const {
    static_assert(HALF, "half-block sums are the half-hop stream's");
    const float* hl = last_half<REV>();
    pb = h1[0];
    pc = hl[0];
#pragma unroll
    for (int e = 1; e < H; ++e) pb += h1[e];
#pragma unroll
    for (int e = 1; e < H; ++e) pc += hl[e];
  }

  // HALF: segment means of the fetched unit from the group sums of its
  // three half-blocks, the first carried from the previous unit.  Segment s
  // is (h0, h1) and s+1 is (h1, h2); carry becomes the sum of h2, which is
  // the next unit's first half.
  static MSIG_HD void half_means(float& carry, float sb, float sc, float inv, float& mr,
                                 float& mi) {
    mr = (carry + sb) * inv;
    mi = (sb + sc) * inv;
    carry = sc;
  }

  // HALF: the points of a unit fetched with fetch_full, detrended by the
  // segment means (mr, mi) when `detrend` and windowed by w[e] (sample
  // t + TPF e).
  template <bool REV = false>
  MSIG_HD void take_full_prepped(cf* z, const float* w, bool detrend, float mr, float mi) const {
    static_assert(HALF, "full-unit take is the half-hop stream's");
    const float* hf = first_half<REV>();
    const float* hl = last_half<REV>();
    if (detrend) {
#pragma unroll
      for (int e = 0; e < H; ++e) {
        z[e] = cf{(hf[e] - mr) * w[e], (h1[e] - mi) * w[e]};
        z[e + H] = cf{(h1[e] - mr) * w[e + H], (hl[e] - mi) * w[e + H]};
      }
    } else {
#pragma unroll
      for (int e = 0; e < H; ++e) {
        z[e] = cf{hf[e] * w[e], h1[e] * w[e]};
        z[e + H] = cf{h1[e] * w[e + H], hl[e] * w[e + H]};
      }
    }
  }

  // the fetched unit as complex points z[e] <-> sample t + TPF*e
  MSIG_HD void take(cf* z) const {
    if constexpr (HALF) {
      // a missing partner segment (odd segment count) is all zeros, including
      // the half it would share with segment s0
#pragma unroll
      for (int e = 0; e < H; ++e) {
        z[e] = cf{h0[e], pair_ok ? h1[e] : 0.f};
        z[e + H] = cf{h1[e], h2[e]};
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) z[e] = raw[e];
    }
  }
};

}  // namespace msig
