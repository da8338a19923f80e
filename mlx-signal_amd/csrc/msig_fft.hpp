// Register-resident Stockham FFT building blocks for the fused spectral kernels.
//
// One FFT of size N is carried by a "group" of TPF = N / E threads, each
// holding E complex points in registers.  Pass p of the radix schedule reads
// R_p points at stride N / R_p, applies the Stockham twiddles, runs an
// in-register DFT_R and writes its outputs back through LDS at the
// self-sorting positions, so the final pass leaves thread t holding bins
// t + TPF * e (natural order, no bit reversal).  The same templates compile
// for the host, which lets the index math be exercised lane by lane on a CPU
// (msig_fft_emu.cpp) before any GPU run.
//
// The real-input forward transform replaces the reference's per-segment
// even/odd packing + radix-2 Stockham in threadgroup memory
// (reference: src/mlx_signal_processing/_stft_metal.py:80-147); here two real
// segments ride in one complex FFT (z = a + i b) and are separated later with
// the conjugate-mirror identity, which needs no untangle twiddles at all.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MSIG_HD __host__ __device__ __forceinline__
#else
#define MSIG_HD inline
#endif

// Global-address-space qualifier for device code.  Pointers that reach a
// kernel through memory (batch job descriptors) are generic to the compiler,
// which then emits FLAT accesses; casting them to this space restores
// global_load / global_store.
#if defined(__HIP_DEVICE_COMPILE__)
#define MSIG_GLOBAL __attribute__((address_space(1)))
#else
#define MSIG_GLOBAL
#endif

namespace msig {

template <class T>
MSIG_HD const MSIG_GLOBAL T* as_global(const T* p) {
  return (const MSIG_GLOBAL T*)p;
}
template <class T>
MSIG_HD MSIG_GLOBAL T* as_global_mut(T* p) {
  return (MSIG_GLOBAL T*)p;
}

struct cf {
  float re, im;
};

MSIG_HD cf cadd(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
MSIG_HD cf csub(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
// a * b
MSIG_HD cf cmul(cf a, cf b) {
  return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// conj(a) * b
MSIG_HD cf cmulc(cf a, cf b) {
  return cf{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

// cos(2*pi*m/64), float32-rounded from FP64, m = 0..63.
MSIG_HD float cos64(int m) {
  constexpr float t[64] = {
      1.0f,          0.99518472f,   0.980785251f,  0.956940353f,  0.923879504f,
      0.881921291f,  0.831469595f,  0.773010433f,  0.707106769f,  0.634393275f,
      0.555570245f,  0.471396744f,  0.382683426f,  0.290284663f,  0.195090324f,
      0.0980171412f, 0.0f,          -0.0980171412f, -0.195090324f, -0.290284663f,
      -0.382683426f, -0.471396744f, -0.555570245f, -0.634393275f, -0.707106769f,
      -0.773010433f, -0.831469595f, -0.881921291f, -0.923879504f, -0.956940353f,
      -0.980785251f, -0.99518472f,  -1.0f,         -0.99518472f,  -0.980785251f,
      -0.956940353f, -0.923879504f, -0.881921291f, -0.831469595f, -0.773010433f,
      -0.707106769f, -0.634393275f, -0.555570245f, -0.471396744f, -0.382683426f,
      -0.290284663f, -0.195090324f, -0.0980171412f, 0.0f,          0.0980171412f,
      0.195090324f,  0.290284663f,  0.382683426f,  0.471396744f,  0.555570245f,
      0.634393275f,  0.707106769f,  0.773010433f,  0.831469595f,  0.881921291f,
      0.923879504f,  0.956940353f,  0.980785251f,  0.99518472f};
  return t[m & 63];
}

// v * exp(-2*pi*i*m/R) for a compile-time-foldable (m, R), R <= 64.
// Multiples of a quarter turn become swaps/negations; the rest cost one
// complex multiply against constants.
MSIG_HD cf mul_wconst(cf v, int m, int R) {
  m = ((m % R) + R) % R;
  if (m == 0) return v;
  if (4 * m == R) return cf{v.im, -v.re};       // * (-i)
  if (2 * m == R) return cf{-v.re, -v.im};      // * (-1)
  if (4 * m == 3 * R) return cf{-v.im, v.re};   // * (+i)
  const int q = m * (64 / R);
  const float c = cos64(q);
  const float s = cos64(q - 16);                // sin(2*pi*q/64)
  // (re + i im)(c - i s)
  return cf{v.re * c + v.im * s, v.im * c - v.re * s};
}

// In-register forward DFT of size R (natural order in and out):
// V[k] = sum_j v[j] exp(-2*pi*i*j*k/R).
template <int R>
struct Dft;

template <>
struct Dft<1> {
  static MSIG_HD void run(cf*) {}
};

template <>
struct Dft<2> {
  static MSIG_HD void run(cf* v) {
    cf a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <>
struct Dft<4> {
  static MSIG_HD void run(cf* v) {
    cf a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    cf c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cf{b.re + d.im, b.im - d.re};  // b - i d
    v[3] = cf{b.re - d.im, b.im + d.re};  // b + i d
  }
};

// Composite sizes: Cooley-Tukey R = R1 * R2 entirely in registers,
// n = R2*n1 + n2, k = k1 + R1*k2.
template <int R>
struct Dft {
  static constexpr int R1 = (R >= 16) ? 4 : 2;
  static constexpr int R2 = R / R1;
  static MSIG_HD void run(cf* v) {
    cf a[R2][R1];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) a[n2][n1] = v[R2 * n1 + n2];
      Dft<R1>::run(a[n2]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) a[n2][k1] = mul_wconst(a[n2][k1], n2 * k1, R);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      cf b[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
      Dft<R2>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
  }
};

// Radix schedule for an N-point FFT carried E points per thread.  N = E^m r
// with r < E; schedule S = 0 runs E, E, ..., r (remainder last), S = 1 runs
// r, E, ..., E (remainder first).  When N is a power of E both are E, ..., E.
// N and E are powers of two, E <= N.
//
// Pass 0 reads the thread's registers in place: register e holds position
// t + TPF u + (N / R0) j (u = e / R0, j = e % R0, R0 = radix(0)), which is
// sample t + TPF (u + (E / R0) j); load_order() maps register e to the
// sample index the caller loaded into register u + (E / R0) j.
template <int N, int E, int S = 0>
struct Plan {
  static_assert((N & (N - 1)) == 0 && (E & (E - 1)) == 0 && E <= N, "pow2 sizes");
  static constexpr int TPF = N / E;  // threads per FFT
  static constexpr int rem() {
    int r = N;
    while (r >= E) r /= E;
    return r;  // 1 when N is a power of E
  }
  static constexpr bool REM_FIRST = S == 1 && rem() > 1;
  static constexpr int radix(int p) {
    if (REM_FIRST) return p == 0 ? rem() : E;
    int r = N;
    for (int i = 0; i < p; ++i) r /= (r < E ? r : E);
    return r < E ? r : E;
  }
  static constexpr int npasses() {
    int rem = N, c = 0;
    while (rem > 1) {
      rem /= (rem < E ? rem : E);
      ++c;
    }
    return c;
  }
  static constexpr int ns(int p) {
    int s = 1;
    for (int i = 0; i < p; ++i) s *= radix(i);
    return s;
  }
  static constexpr int NP = npasses();
  // twiddle registers held per thread: passes p >= 1, (E/R)*(R-1) each
  static constexpr int tw_off(int p) {
    int o = 0;
    for (int i = 1; i < p; ++i) o += (E / radix(i)) * (radix(i) - 1);
    return o;
  }
  static constexpr int TWN = tw_off(NP) > 0 ? tw_off(NP) : 1;
  // LDS slot of logical position pos in the complex (8-byte) exchange: one
  // pad slot every 16 keeps the stride-16 stores conflict-free (8-byte
  // writes serve 16 lanes over 32 banks); the stride-1 reads of 32 lanes
  // meet on one bank pair once per 32 lanes (8-byte reads serve 32 lanes
  // over 64 banks), which no pad of this additive form avoids while keeping
  // the stores conflict-free.  The split exchange has its own pads (XPad).
  static MSIG_HD int pad(int pos) { return pos + (pos >> 4); }
  static constexpr int LDS = N + N / 16;

  // Split exchange (one float per slot, Plan::LDS slots per FFT group): the
  // slot of position pos at the boundary into pass P (pass P - 1's stores,
  // pass P's loads) is pos + k (pos >> sh).  (sh, k) is chosen per boundary
  // at compile time: among the pairs that keep every slot below LDS and keep
  // each slot a thread part plus a register part (slot(a + b) = slot(a) +
  // slot(b), so the register parts fold into the instructions' offset
  // fields), the one whose stores and loads repeat the fewest LDS banks
  // (XBANKS banks of 4 B, XLANES lanes served together); ties keep pad().
  // 32 banks and 32 lanes per cycle is the model that the bank-conflict
  // counter (SQ_LDS_BANK_CONFLICT) confirms for these 4-byte accesses on
  // gfx950: welch N = 1024 reads 2^23 conflict cycles per call with pad(),
  // 2^22 with the pads chosen for 64 banks x 64 lanes, and 0 with these.
  static constexpr int XBANKS = 32, XLANES = 32;
  struct XPad {
    int sh, k;
  };
  static constexpr int xpad_of(XPad x, int pos) { return pos + x.k * (pos >> x.sh); }
  // out_pos / in_pos as a thread part a(t) plus a register part b(e); false
  // when pass p's stores have no such form (NS > TPF, final passes only)
  static constexpr bool xsplit(int p, bool st) { return !st || TPF % ns(p) == 0; }
  static constexpr int xa(int p, bool st, int t) {
    return st ? (t / ns(p)) * ns(p) * radix(p) + (t % ns(p)) : t;
  }
  static constexpr int xb(int p, bool st, int e) {
    const int r = radix(p), u = e / r, j = e % r;
    return st ? u * TPF * r + j * ns(p) : TPF * u + (N / r) * j;
  }
  // extra bank cycles of one access set (stores of pass p or loads of p)
  static constexpr int xcost(XPad x, int p, bool st) {
    const int r = radix(p), nsp = ns(p), waves = TPF > 64 ? TPF / 64 : 1;
    int cost = 0;
    for (int e = 0; e < E; ++e)
      for (int w = 0; w < waves; ++w)
        for (int c = 0; c < 64; c += XLANES) {
          int cnt[XBANKS] = {};
          int worst = 0;
          for (int l = c; l < c + XLANES; ++l) {
            const int tid = w * 64 + l, g = tid / TPF, t = tid % TPF;
            const int u = e / r, j = e % r, vt = t + TPF * u;
            const int pos = st ? (vt / nsp) * nsp * r + (vt % nsp) + j * nsp : t + TPF * u + (N / r) * j;
            const int b = (g * LDS + xpad_of(x, pos)) % XBANKS;
            worst = ++cnt[b] > worst ? cnt[b] : worst;
          }
          cost += worst - 1;
        }
    return cost;
  }
  // slot(a + b) = slot(a) + slot(b) for every thread part a and register
  // part b when the largest low parts (mod 2^sh) of a and of b sum below 2^sh
  // (no carry into bit sh, so (a + b) >> sh = (a >> sh) + (b >> sh))
  static constexpr bool xlinear(XPad x, int p, bool st) {
    if (!xsplit(p, st)) return false;
    const int m = (1 << x.sh) - 1;
    int ma = 0, mb = 0;
    for (int t = 0; t < TPF; ++t) ma = (xa(p, st, t) & m) > ma ? (xa(p, st, t) & m) : ma;
    for (int e = 0; e < E; ++e) mb = (xb(p, st, e) & m) > mb ? (xb(p, st, e) & m) : mb;
    return ma + mb <= m;
  }
  static constexpr bool xvalid(XPad x, int p) {
    return x.k * ((N - 1) >> x.sh) + N <= LDS && (!xsplit(p - 1, true) || xlinear(x, p - 1, true)) &&
           xlinear(x, p, false);
  }
  static constexpr XPad xpad_choice(int p) {
    XPad best{4, 1};
    int best_cost = xcost(best, p - 1, true) + xcost(best, p, false);
    constexpr int ks[] = {1, 2, 3, 4, 5, 8};
    for (int sh = 4; sh <= 7; ++sh)
      for (int k : ks) {
        const XPad x{sh, k};
        if ((sh == 4 && k == 1) || !xvalid(x, p)) continue;
        const int c = xcost(x, p - 1, true) + xcost(x, p, false);
        if (c < best_cost) {
          best = x;
          best_cost = c;
        }
      }
    return best;
  }
  // final bin held by thread t in register e
  static MSIG_HD int bin(int t, int e) {
    constexpr int R = radix(NP - 1);
    constexpr int NS = ns(NP - 1);
    const int u = e / R, j = e % R;
    return t + TPF * u + NS * j;
  }
  // logical position written by thread t, register e, after pass p
  template <int P>
  static MSIG_HD int out_pos(int t, int e) {
    constexpr int R = radix(P);
    constexpr int NS = ns(P);
    const int u = e / R, j = e % R;
    const int vt = t + TPF * u;
    return (vt / NS) * NS * R + (vt % NS) + j * NS;
  }
  // logical position read by thread t into register e for pass P
  template <int P>
  static MSIG_HD int in_pos(int t, int e) {
    constexpr int R = radix(P);
    const int u = e / R, j = e % R;
    return t + TPF * u + (N / R) * j;
  }
  // twiddle exponent (units of exp(-2*pi*i/N)) for pass P, register e
  // caller register feeding pass-0 register e (identity for S = 0)
  static MSIG_HD constexpr int load_order(int e) {
    return (e / radix(0)) + (E / radix(0)) * (e % radix(0));
  }
  template <int P>
  static MSIG_HD int tw_exp(int t, int e) {
    constexpr int R = radix(P);
    constexpr int NS = ns(P);
    const int u = e / R, j = e % R;
    const int vt = t + TPF * u;
    return (j * (vt % NS) * (N / (NS * R))) % N;
  }

  // Shared LDS twiddle table (one per block).  Pass p's twiddle for
  // register u R + j depends on vt % NS, vt = t + TPF u.  When NS <= TPF
  // (NS divides TPF) that is t % NS for every u, so the pass can store R - 1
  // slots of NS columns ("narrow"); otherwise it stores (E / R)(R - 1) slots
  // of TPF columns.  Slot-major either way: the lanes of a wave read
  // consecutive (or broadcast) entries of one slot.  Narrow passes are used
  // only when the all-wide table would exceed 16 KB (N >= 2048 here): below
  // that the table costs no occupancy, and with one base address for a whole
  // narrow pass the compiler serialises its reads (one LDS round trip per
  // pair), measured 4-6% slower at N = 1024.
  static constexpr bool TWL_DEDUP = TWN * TPF * 8 > 16384;
  static constexpr bool twl_narrow(int p) { return TWL_DEDUP && ns(p) <= TPF; }
  static constexpr int twl_cols(int p) { return twl_narrow(p) ? ns(p) : TPF; }
  static constexpr int twl_slots(int p) {
    return twl_narrow(p) ? radix(p) - 1 : (E / radix(p)) * (radix(p) - 1);
  }
  static constexpr int twl_off(int p) {
    int o = 0;
    for (int i = 1; i < p; ++i) o += twl_slots(i) * twl_cols(i);
    return o;
  }
  static constexpr int TWL = twl_off(NP) > 0 ? twl_off(NP) : 1;
  // column of thread t in pass P's table
  template <int P>
  static MSIG_HD int twl_col(int t) {
    return twl_narrow(P) ? (t & (ns(P) - 1)) : t;
  }
  // slot of register u R + j (j >= 1) in pass P's table
  static constexpr int twl_slot(int p, int u, int j) {
    return twl_narrow(p) ? j - 1 : u * (radix(p) - 1) + (j - 1);
  }
  // twiddle exponent stored at table entry i (0 <= i < TWL)
  template <int P = 1>
  static MSIG_HD int twl_exp(int i) {
    if constexpr (P >= NP) {
      return 0;
    } else {
      constexpr int R = radix(P), NS = ns(P), C = twl_cols(P), SZ = twl_slots(P) * C;
      if (i >= SZ) return twl_exp<P + 1>(i - SZ);
      const int slot = i / C, col = i % C;
      const int u = twl_narrow(P) ? 0 : slot / (R - 1);
      const int j = twl_narrow(P) ? slot + 1 : slot % (R - 1) + 1;
      const int vt = col + TPF * u;
      return (j * (vt % NS) * (N / (NS * R))) % N;
    }
  }
};

// Load the per-thread twiddle registers from an N-entry table of
// exp(-2*pi*i*e/N) (computed in FP64 by the host, stored as float pairs).
template <int N, int E, int P = 1, class TablePtr, int S = 0>
MSIG_HD void load_twiddles(int t, TablePtr table, cf* tw) {
  using PL = Plan<N, E, S>;
  if constexpr (P < PL::NP) {
    constexpr int R = PL::radix(P);
    constexpr int off = PL::tw_off(P);
#pragma unroll
    for (int u = 0; u < E / R; ++u)
#pragma unroll
      for (int j = 1; j < R; ++j)
      {
        const int ex = PL::template tw_exp<P>(t, u * R + j);
        tw[off + u * (R - 1) + (j - 1)] = cf{table[ex].re, table[ex].im};
      }
    load_twiddles<N, E, P + 1, TablePtr, S>(t, table, tw);
  }
}

// Pass P compute step on registers already read for that pass.  TWL = false:
// tw is this thread's register array from load_twiddles.  TWL = true: tw is
// the block's shared table (Plan::twl_*), read at this thread's column, so
// every read is one ds_read_b64 at a constant offset from a per-pass base.
template <int N, int E, int P, bool TWL = false, int S = 0>
MSIG_HD void pass_compute(cf* v, const cf* tw, int t = 0) {
  using PL = Plan<N, E, S>;
  constexpr int R = PL::radix(P);
  const cf* twp = tw;
  if constexpr (TWL && P > 0) twp = tw + PL::twl_off(P) + PL::template twl_col<P>(t);
#pragma unroll
  for (int u = 0; u < E / R; ++u) {
    if constexpr (P > 0) {
      constexpr int off = PL::tw_off(P);
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const cf w = TWL ? twp[PL::twl_slot(P, u, j) * PL::twl_cols(P)] : twp[off + u * (R - 1) + (j - 1)];
        v[u * R + j] = cmul(v[u * R + j], w);
      }
    }
    Dft<R>::run(v + u * R);
  }
}

// LDS slots of pass P's outputs / inputs.  LIN (the split exchange): the
// boundary's own pad (Plan::xpad_choice), written as (thread part) +
// (register part) wherever that pad allows it.  !LIN (the complex exchange)
// keeps pad(pos), whose per-element addresses the compiler hoists out of the
// loop (measured: the two-signal means ran 2-4% slower with the linear form
// at N = 1024).
template <int N, int E, int P, int S = 0, bool LIN = true>
MSIG_HD int out_slot(int t, int e) {
  using PL = Plan<N, E, S>;
  if constexpr (!LIN) {
    return PL::pad(PL::template out_pos<P>(t, e));
  } else {
    constexpr typename PL::XPad x = PL::xpad_choice(P + 1);
    if constexpr (PL::xsplit(P, true) && PL::xvalid(x, P + 1))
      return PL::xpad_of(x, PL::xa(P, true, t)) + PL::xpad_of(x, PL::xb(P, true, e));
    else
      return PL::xpad_of(x, PL::template out_pos<P>(t, e));
  }
}
template <int N, int E, int P, int S = 0, bool LIN = true>
MSIG_HD int in_slot(int t, int e) {
  using PL = Plan<N, E, S>;
  if constexpr (!LIN) {
    return PL::pad(PL::template in_pos<P>(t, e));
  } else {
    constexpr typename PL::XPad x = PL::xpad_choice(P);
    if constexpr (PL::xvalid(x, P))
      return PL::xpad_of(x, t) + PL::xpad_of(x, PL::xb(P, false, e));
    else
      return PL::xpad_of(x, PL::template in_pos<P>(t, e));
  }
}

template <int N, int E, int P, int S = 0>
MSIG_HD void pass_store(int t, const cf* v, cf* lds) {
#pragma unroll
  for (int e = 0; e < E; ++e) lds[out_slot<N, E, P, S, false>(t, e)] = v[e];
}

template <int N, int E, int P, int S = 0>
MSIG_HD void pass_load(int t, cf* v, const cf* lds) {
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = lds[in_slot<N, E, P, S, false>(t, e)];
}

// One component (C = 0: real, 1: imaginary) of a pass exchange through a
// float buffer of Plan::LDS slots: the split exchange moves the two halves
// in turn through half the LDS.
template <int N, int E, int P, int S, int C>
MSIG_HD void pass_store_part(int t, const cf* v, float* lds) {
#pragma unroll
  for (int e = 0; e < E; ++e) lds[out_slot<N, E, P, S>(t, e)] = C ? v[e].im : v[e].re;
}

template <int N, int E, int P, int S, int C>
MSIG_HD void pass_load_part(int t, cf* v, const float* lds) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float x = lds[in_slot<N, E, P, S>(t, e)];
    if (C) {
      v[e].im = x;
    } else {
      v[e].re = x;
    }
  }
}

}  // namespace msig
