// Host emulation of the device FFT: runs the same msig_fft.hpp templates the
// kernels use, thread by thread with an emulated LDS array, in the same
// store / barrier / load order, and checks
//   1. the FFT against an O(N^2) FP64 DFT,
//   2. that every LDS slot of a pass is written exactly once and every read
//      slot was written (no stale or clobbered data),
//   3. the two-for-one separation and mirror fold used by the kernels:
//      |X_s|^2 + |X_{s+1}|^2 == (|Z[k]|^2 + |Z[N-k]|^2) / 2, and
//      conj(X_s)Y_s + conj(X_{s+1})Y_{s+1} == (Q[k] + conj Q[N-k]) / 2
//      with Q = conj(Zx) Zy.
//   4. the kernels' UnitStream (general and hop == N/2 register-reuse modes)
//      fed through the emulated FFT, detrend, window, accumulate and mirror
//      fold, against an FP64 Welch mean on the same float32 row, for odd and
//      even segment counts and nperseg < nfft.  The row is followed by NaN
//      sentinels so any read past its end poisons the result.
// Prints one line per case and exits non-zero on any failure.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "msig_fft.hpp"
#include "msig_stream.hpp"

using msig::cf;
using cd = std::complex<double>;

static std::vector<cd> dft(const std::vector<cd>& z) {
  const int N = (int)z.size();
  std::vector<cd> out(N);
  for (int k = 0; k < N; ++k) {
    cd s = 0;
    for (int n = 0; n < N; ++n) s += z[n] * std::polar(1.0, -2.0 * M_PI * double((int64_t)n * k % N) / N);
    out[k] = s;
  }
  return out;
}

template <int N, int E, int P, int S, int C>
static bool split_exchange(const std::vector<std::vector<cf>>& before, const std::vector<std::vector<cf>>& after) {
  using PL = msig::Plan<N, E, S>;
  constexpr int TPF = PL::TPF;
  std::vector<float> lf(PL::LDS, std::nanf(""));
  std::vector<int> count(PL::LDS, 0);
  std::vector<std::vector<cf>> v = before;
  for (int t = 0; t < TPF; ++t) {
    msig::pass_store_part<N, E, P - 1, S, C>(t, v[t].data(), lf.data());
    for (int e = 0; e < E; ++e) {
      const int slot = msig::out_slot<N, E, P - 1, S>(t, e);
      if (slot < 0 || slot >= PL::LDS || ++count[slot] > 1) {
        std::printf("N=%d pass %d: split store slot %d out of range or written twice\n", N, P - 1, slot);
        return false;
      }
    }
  }
  for (int t = 0; t < TPF; ++t) {
    for (int e = 0; e < E; ++e) {
      const int slot = msig::in_slot<N, E, P, S>(t, e);
      if (slot < 0 || slot >= PL::LDS || count[slot] != 1) {
        std::printf("N=%d pass %d: split load of unwritten slot %d\n", N, P, slot);
        return false;
      }
    }
    msig::pass_load_part<N, E, P, S, C>(t, v[t].data(), lf.data());
    for (int e = 0; e < E; ++e) {
      const float got = C ? v[t][e].im : v[t][e].re, want = C ? after[t][e].im : after[t][e].re;
      if (std::memcmp(&got, &want, sizeof(float)) != 0) {
        std::printf("N=%d pass %d: split exchange moved t=%d e=%d elsewhere\n", N, P, t, e);
        return false;
      }
    }
  }
  return true;
}

template <int N, int E, int P, int S = 0, bool TWL = false>
static bool run_pass(std::vector<std::vector<cf>>& v, const std::vector<std::vector<cf>>& tw,
                     const std::vector<cf>& twl, std::vector<cf>& lds, std::vector<int>& wcount) {
  using PL = msig::Plan<N, E, S>;
  constexpr int TPF = PL::TPF;
  if constexpr (P < PL::NP) {
    const std::vector<std::vector<cf>> before = v;
    std::fill(wcount.begin(), wcount.end(), 0);
    for (int t = 0; t < TPF; ++t) {
      msig::pass_store<N, E, P - 1, S>(t, v[t].data(), lds.data());
      for (int e = 0; e < E; ++e) wcount[PL::pad(PL::template out_pos<P - 1>(t, e))]++;
    }
    for (int i = 0; i < (int)wcount.size(); ++i)
      if (wcount[i] > 1) {
        std::printf("N=%d pass %d: LDS slot %d written %d times\n", N, P - 1, i, wcount[i]);
        return false;
      }
    for (int t = 0; t < TPF; ++t) {
      for (int e = 0; e < E; ++e) {
        const int slot = PL::pad(PL::template in_pos<P>(t, e));
        if (wcount[slot] != 1) {
          std::printf("N=%d pass %d: read of unwritten slot %d\n", N, P, slot);
          return false;
        }
      }
      msig::pass_load<N, E, P, S>(t, v[t].data(), lds.data());
    }
    // The split exchange (real parts, then imaginary parts, through LDS
    // floats with the boundary's own pad) must move every point where the
    // complex exchange does, writing each slot below LDS once.
    if (!split_exchange<N, E, P, S, 0>(before, v) || !split_exchange<N, E, P, S, 1>(before, v)) return false;
    for (int t = 0; t < TPF; ++t)
      msig::pass_compute<N, E, P, TWL, S>(v[t].data(), TWL ? twl.data() : tw[t].data(), t);
    return run_pass<N, E, P + 1, S, TWL>(v, tw, twl, lds, wcount);
  }
  return true;
}

// FFT of z (N complex points) exactly as a group of TPF device threads does it
// (radix schedule S; register e of a thread starts with sample t + TPF e and
// is renamed to pass 0's order as fft_passes does).  TWL = true reads the
// twiddles from the block-shared table filled as fill_twiddle_lds does
// (entry i = table[Plan::twl_exp(i)]) instead of per-thread registers.
template <int N, int E, int S = 0, bool TWL = false>
static bool emu_fft(const std::vector<cf>& table, const std::vector<cf>& z, std::vector<cf>& Z) {
  using PL = msig::Plan<N, E, S>;
  constexpr int TPF = PL::TPF;
  std::vector<std::vector<cf>> v(TPF, std::vector<cf>(E)), tw(TPF, std::vector<cf>(PL::TWN));
  std::vector<cf> twl(PL::TWL);
  for (int i = 0; i < PL::TWL; ++i) twl[i] = table[PL::twl_exp(i)];
  for (int t = 0; t < TPF; ++t) {
    msig::load_twiddles<N, E, 1, const cf*, S>(t, table.data(), tw[t].data());
    for (int e = 0; e < E; ++e) {
      const int src = PL::REM_FIRST ? PL::load_order(e) : e;
      if (PL::template in_pos<0>(t, e) != t + TPF * src) {
        std::printf("N=%d S=%d: pass-0 register %d does not hold sample t + TPF * %d\n", N, S, e, src);
        return false;
      }
      v[t][e] = z[t + TPF * src];
    }
    msig::pass_compute<N, E, 0, TWL, S>(v[t].data(), TWL ? twl.data() : tw[t].data(), t);
  }
  std::vector<cf> lds(PL::LDS);
  std::vector<int> wcount(PL::LDS);
  if (!run_pass<N, E, 1, S, TWL>(v, tw, twl, lds, wcount)) return false;
  Z.assign(N, cf{NAN, NAN});
  std::vector<int> seen(N, 0);
  for (int t = 0; t < TPF; ++t)
    for (int e = 0; e < E; ++e) {
      const int k = PL::bin(t, e);
      if (k < 0 || k >= N || seen[k]) {
        std::printf("N=%d: bad final bin map t=%d e=%d k=%d\n", N, t, e, k);
        return false;
      }
      seen[k] = 1;
      Z[k] = v[t][e];
    }
  return true;
}

template <int N>
static int check_size(std::mt19937_64& rng) {
  constexpr int E = N < 16 ? N : 16;
  std::vector<cf> table(N);
  for (int i = 0; i < N; ++i) {
    const double a = -2.0 * M_PI * i / N;
    table[i] = cf{(float)std::cos(a), (float)std::sin(a)};
  }
  std::normal_distribution<double> nd;
  // 1. FFT vs DFT on random complex input
  std::vector<cf> z(N), Z;
  std::vector<cd> zd(N);
  for (int i = 0; i < N; ++i) {
    z[i] = cf{(float)nd(rng), (float)nd(rng)};
    zd[i] = cd(z[i].re, z[i].im);
  }
  if (!emu_fft<N, E>(table, z, Z)) return 1;
  auto ref = dft(zd);
  double peak = 0, err = 0;
  for (int k = 0; k < N; ++k) {
    peak = std::max(peak, std::abs(ref[k]));
    err = std::max(err, std::abs(cd(Z[k].re, Z[k].im) - ref[k]));
  }
  const double rel = err / peak;
  // 2. two-for-one identities on real segment pairs / signal pairs
  std::vector<cf> za(N), zb(N), Za, Zb;
  std::vector<cd> xa(N), xb(N), ya(N), yb(N);
  for (int i = 0; i < N; ++i) {
    xa[i] = nd(rng);
    xb[i] = nd(rng);
    ya[i] = 0.5 * xa[i].real() + nd(rng);
    yb[i] = 0.5 * xb[i].real() + nd(rng);
    za[i] = cf{(float)xa[i].real(), (float)xb[i].real()};
    zb[i] = cf{(float)ya[i].real(), (float)yb[i].real()};
  }
  if (!emu_fft<N, E>(table, za, Za) || !emu_fft<N, E>(table, zb, Zb)) return 1;
  auto Xa = dft(xa), Xb = dft(xb), Ya = dft(ya), Yb = dft(yb);
  double ppk = 0, perr = 0, cpk = 0, cerr = 0;
  for (int k = 0; k < N; ++k) {
    const int mk = (N - k) & (N - 1);
    const double s = 0.5 * (std::norm(cd(Za[k].re, Za[k].im)) + std::norm(cd(Za[mk].re, Za[mk].im)));
    const double sref = std::norm(Xa[k]) + std::norm(Xb[k]);
    ppk = std::max(ppk, sref);
    perr = std::max(perr, std::abs(s - sref));
    const cd q = std::conj(cd(Za[k].re, Za[k].im)) * cd(Zb[k].re, Zb[k].im);
    const cd qm = std::conj(cd(Za[mk].re, Za[mk].im)) * cd(Zb[mk].re, Zb[mk].im);
    const cd c = 0.5 * (q + std::conj(qm));
    const cd cref = std::conj(Xa[k]) * Ya[k] + std::conj(Xb[k]) * Yb[k];
    cpk = std::max(cpk, std::abs(cref));
    cerr = std::max(cerr, std::abs(c - cref));
    // per-frame separation used by the frames kernel
    const cd a(Za[k].re, Za[k].im), b(Za[mk].re, Za[mk].im);
    const cd x0 = 0.5 * (a + std::conj(b));
    const cd x1 = cd(0, -0.5) * (a - std::conj(b));
    if (std::abs(x0 - Xa[k]) > 1e-4 * std::sqrt(ppk) || std::abs(x1 - Xb[k]) > 1e-4 * std::sqrt(ppk)) {
      std::printf("N=%d: separation mismatch at k=%d\n", N, k);
      return 1;
    }
  }
  const bool ok = rel < 2e-6 && perr / ppk < 2e-6 && cerr / cpk < 2e-6;
  std::printf("N=%5d E=%2d passes=%d fft_rel=%.3e psd_fold_rel=%.3e csd_fold_rel=%.3e %s\n", N, E,
              msig::Plan<N, E>::NP, rel, perr / ppk, cerr / cpk, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}


// Welch mean of one row exactly as k_mean does it for a group of TPF threads.
// lead / tail > 0: the row is read through the virtual zero extension
// (lead zeros in front of the n stored samples, tail zeros after them), as
// the per-frame kernels read stft rows; NaN guards on both sides of the
// stored samples poison any read outside them.
template <int N, bool HALF, bool EXT = false>
static int check_stream(std::mt19937_64& rng, int nperseg, int hop, int n, int detrend,
                        int lead = 0, int tail = 0) {
  constexpr int E = N < 16 ? N : 16;
  constexpr int TPF = N / E;
  std::normal_distribution<double> nd;
  const int guard = 2 * N;
  std::vector<float> buf(guard + n + 2 * N, NAN);
  float* x = buf.data() + guard;  // the stored row
  for (int i = 0; i < n; ++i) x[i] = (float)(nd(rng) + 0.5 * std::sin(0.3 * i) + 2.0);
  const int n_ext = lead + n + tail;
  std::vector<float> xe(n_ext, 0.f);  // the extended row, for the reference
  for (int i = 0; i < n; ++i) xe[lead + i] = x[i];
  std::vector<cf> table(N);
  for (int i = 0; i < N; ++i) {
    const double ang = -2.0 * M_PI * i / N;
    table[i] = cf{(float)std::cos(ang), (float)std::sin(ang)};
  }
  std::vector<double> wd(nperseg);
  for (int i = 0; i < nperseg; ++i) wd[i] = 0.5 - 0.5 * std::cos(2.0 * M_PI * i / nperseg);
  msig::KArgs a{};
  a.nseg = (n_ext - nperseg) / hop + 1;
  a.nperseg = nperseg;
  a.hop = hop;
  a.units = (int32_t)((a.nseg + 1) / 2);
  a.detrend = detrend;
  a.lead = lead;
  a.nvalid = n;
  std::vector<msig::UnitStream<N, E, false, HALF, EXT>> st(TPF);
  for (int t = 0; t < TPF; ++t) {
    st[t].row = reinterpret_cast<const char*>(x);
    st[t].t = t;
    st[t].fetch(0, true, a);
  }
  std::vector<double> S(N, 0.0);
  for (int u = 0; u <= a.units; ++u) {  // one unit past the end: must load zeros
    std::vector<cf> z(N), Z;
    for (int t = 0; t < TPF; ++t) {
      cf zt[E];
      st[t].take(zt);
      for (int e = 0; e < E; ++e) z[t + TPF * e] = zt[e];
      if (u < a.units) st[t].fetch(u + 1, false, a);
    }
    if (u == a.units) {
      for (int i = 0; i < N; ++i)
        if (z[i].re != 0.f || z[i].im != 0.f) {
          std::printf("stream N=%d: unit past the end is not zero\n", N);
          return 1;
        }
      break;
    }
    float sr = 0.f, si = 0.f;
    for (int i = 0; i < N; ++i) {
      sr += z[i].re;
      si += z[i].im;
    }
    const float mr = detrend ? sr / nperseg : 0.f, mi = detrend ? si / nperseg : 0.f;
    for (int i = 0; i < N; ++i) {
      const float w = i < nperseg ? (float)wd[i] : 0.f;
      z[i] = cf{(z[i].re - mr) * w, (z[i].im - mi) * w};
    }
    if (!emu_fft<N, E>(table, z, Z)) return 1;
    for (int k = 0; k < N; ++k) S[k] += (double)Z[k].re * Z[k].re + (double)Z[k].im * Z[k].im;
  }
  // FP64 reference: per-segment detrend, window, zero-pad to N, DFT
  std::vector<double> P(N, 0.0);
  for (int s = 0; s < a.nseg; ++s) {
    std::vector<cd> seg(N, 0.0);
    double m = 0;
    for (int i = 0; i < nperseg; ++i) m += xe[(size_t)s * hop + i];
    m = detrend ? m / nperseg : 0.0;
    for (int i = 0; i < nperseg; ++i) seg[i] = (xe[(size_t)s * hop + i] - m) * wd[i];
    auto X = dft(seg);
    for (int k = 0; k < N; ++k) P[k] += std::norm(X[k]);
  }
  double peak = 0, err = 0;
  for (int k = 0; k < N; ++k) {
    const double got = 0.5 * (S[k] + S[(N - k) & (N - 1)]);
    peak = std::max(peak, P[k]);
    err = std::isfinite(got) ? std::max(err, std::abs(got - P[k])) : INFINITY;  // NaN poisons
  }
  const double rel = std::isfinite(err) ? err / peak : 1.0;
  const bool ok = rel < 1e-5;
  std::printf("stream N=%5d %s nperseg=%d hop=%d nseg=%lld detrend=%d lead=%d tail=%d rel=%.3e %s\n",
              N, HALF ? "half" : "gen ", nperseg, hop, (long long)a.nseg, detrend, lead, tail, rel,
              ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// The hop == N/2 group loop of k_mean: each group walks units
// [unit0, min(unit0 + run, units)); units below nseg / 2 use fetch_full with
// the next fetch clamped to the group's last full unit, and take their points
// with take_full_prepped, detrended by means built from half-block group
// sums (first_half_sum once per group, new_half_sums per unit, half_means
// carrying the shared half's sum), as the kernel does.  rotate: the
// standalone welch kernel's loop instead, two units per trip with the
// half-block roles swapped in the second (fetch_next, REV alternating) and
// an odd last unit alone.  The partner-less unit of an odd segment count
// runs once through the general fetch.  Checks coverage (each unit once),
// that every fetch stays inside the row, and the mean against the FP64 DFT.
template <int N>
static int check_half_groups(std::mt19937_64& rng, int nseg, int run, int detrend,
                             bool rotate = false) {
  constexpr int E = N < 16 ? N : 16;
  constexpr int TPF = N / E;
  const int hop = N / 2;
  const int n = N + hop * (nseg - 1) + hop / 3;  // samples past the last segment are unused
  std::normal_distribution<double> nd;
  std::vector<float> x(n + 2 * N, NAN);
  for (int i = 0; i < n; ++i) x[i] = (float)(nd(rng) + 0.5 * std::sin(0.3 * i) - 1.0);
  std::vector<cf> table(N);
  for (int i = 0; i < N; ++i) {
    const double ang = -2.0 * M_PI * i / N;
    table[i] = cf{(float)std::cos(ang), (float)std::sin(ang)};
  }
  std::vector<double> wd(N);
  for (int i = 0; i < N; ++i) wd[i] = 0.5 - 0.5 * std::cos(2.0 * M_PI * i / N);
  msig::KArgs a{};
  a.nseg = nseg;
  a.nperseg = N;
  a.hop = hop;
  a.units = (int32_t)((nseg + 1) / 2);
  a.run = run;
  a.detrend = detrend;
  a.nvalid = n;
  const int groups = (a.units + run - 1) / run + 1;  // one group past the end
  std::vector<int> seen(a.units, 0);
  std::vector<double> S(N, 0.0);
  bool in_bounds = true;
  auto accumulate = [&](std::vector<cf>& z) -> bool {
    std::vector<cf> Z;
    if (!emu_fft<N, E>(table, z, Z)) return false;
    for (int k = 0; k < N; ++k) S[k] += (double)Z[k].re * Z[k].re + (double)Z[k].im * Z[k].im;
    return true;
  };
  auto process = [&](std::vector<cf>& z) -> bool {  // general prep (the tail unit)
    float sr = 0.f, si = 0.f;
    for (int i = 0; i < N; ++i) {
      sr += z[i].re;
      si += z[i].im;
    }
    const float mr = detrend ? sr / N : 0.f, mi = detrend ? si / N : 0.f;
    for (int i = 0; i < N; ++i) z[i] = cf{(z[i].re - mr) * (float)wd[i], (z[i].im - mi) * (float)wd[i]};
    return accumulate(z);
  };
  using US = msig::UnitStream<N, E, false, true>;
  std::vector<std::vector<float>> wt(TPF, std::vector<float>(E));  // window per thread register
  for (int t = 0; t < TPF; ++t)
    for (int e = 0; e < E; ++e) wt[t][e] = (float)wd[t + TPF * e];
  for (int g = 0; g < groups; ++g) {
    const int unit0 = g * run;
    const int u1 = unit0 + run < a.units ? unit0 + run : a.units;
    const int full = (int)(a.nseg / 2);
    const int uf = u1 < full ? u1 : full;
    std::vector<msig::UnitStream<N, E, false, true>> st(TPF);
    for (int t = 0; t < TPF; ++t) {
      st[t].row = reinterpret_cast<const char*>(x.data());
      st[t].t = t;
    }
    auto full_fetch = [&](int u, bool first) {
      in_bounds = in_bounds && (int64_t)(2 * u + 3) * hop <= n;
      for (int t = 0; t < TPF; ++t) st[t].fetch_full(u, first);
    };
    if (unit0 < uf) {
      full_fetch(unit0, true);
      float carry = 0.f;  // group sum of the shared half-block
      if (detrend)
        for (int t = 0; t < TPF; ++t) carry += st[t].first_half_sum();
      auto one_unit = [&](auto rev, int unit, int nxt) -> bool {
        constexpr bool REV = decltype(rev)::value;
        float mr = 0.f, mi = 0.f;
        if (detrend) {
          float sb = 0.f, sc = 0.f;
          for (int t = 0; t < TPF; ++t) {
            float b, c;
            st[t].template new_half_sums<REV>(b, c);
            sb += b;
            sc += c;
          }
          US::half_means(carry, sb, sc, 1.0f / N, mr, mi);
        }
        std::vector<cf> z(N);
        for (int t = 0; t < TPF; ++t) {
          cf zt[E];
          st[t].template take_full_prepped<REV>(zt, wt[t].data(), detrend != 0, mr, mi);
          for (int e = 0; e < E; ++e) z[t + TPF * e] = zt[e];
        }
        if (rotate) {
          in_bounds = in_bounds && (int64_t)(2 * nxt + 3) * hop <= n;
          for (int t = 0; t < TPF; ++t) st[t].template fetch_next<REV>(nxt);
        } else {
          full_fetch(nxt, false);
        }
        if (!accumulate(z)) return false;
        ++seen[unit];
        return true;
      };
      using F = std::integral_constant<bool, false>;
      using T = std::integral_constant<bool, true>;
      int unit = unit0;
      if (rotate) {
        for (; unit + 1 < uf; unit += 2) {
          if (!one_unit(F{}, unit, unit + 1)) return 1;
          if (!one_unit(T{}, unit + 1, unit + 2 < uf ? unit + 2 : uf - 1)) return 1;
        }
        if (unit < uf && !one_unit(F{}, unit, uf - 1)) return 1;
      } else {
        for (; unit < uf; ++unit)
          if (!one_unit(F{}, unit, unit + 1 < uf ? unit + 1 : uf - 1)) return 1;
      }
    }
    if (full >= unit0 && full < u1) {
      std::vector<cf> z(N);
      for (int t = 0; t < TPF; ++t) {
        cf zt[E];
        st[t].fetch(full, true, a);
        st[t].take(zt);
        for (int e = 0; e < E; ++e) z[t + TPF * e] = zt[e];
      }
      if (!process(z)) return 1;
      ++seen[full];
    }
  }
  bool once = true;
  for (int u = 0; u < a.units; ++u) once = once && seen[u] == 1;
  std::vector<double> P(N, 0.0);
  for (int s = 0; s < nseg; ++s) {
    std::vector<cd> seg(N, 0.0);
    double m = 0;
    for (int i = 0; i < N; ++i) m += x[(size_t)s * hop + i];
    m = detrend ? m / N : 0.0;
    for (int i = 0; i < N; ++i) seg[i] = (x[(size_t)s * hop + i] - m) * wd[i];
    auto X = dft(seg);
    for (int k = 0; k < N; ++k) P[k] += std::norm(X[k]);
  }
  double peak = 0, err = 0;
  for (int k = 0; k < N; ++k) {
    const double got = 0.5 * (S[k] + S[(N - k) & (N - 1)]);
    peak = std::max(peak, P[k]);
    err = std::isfinite(got) ? std::max(err, std::abs(got - P[k])) : INFINITY;  // NaN poisons
  }
  const double rel = std::isfinite(err) ? err / peak : 1.0;
  const bool ok = once && in_bounds && rel < 1e-5;
  std::printf("half groups N=%5d nseg=%d run=%d detrend=%d once=%d in_bounds=%d rel=%.3e %s\n",
              N, nseg, run, detrend, (int)once, (int)in_bounds, rel, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// Radix schedule 1 (remainder radix first) against the FP64 DFT.
template <int N>
static int check_schedule(std::mt19937_64& rng) {
  constexpr int E = N < 16 ? N : 16;
  static_assert(msig::Plan<N, E, 1>::REM_FIRST, "a size where the schedules differ");
  std::vector<cf> table(N);
  for (int i = 0; i < N; ++i) {
    const double a = -2.0 * M_PI * i / N;
    table[i] = cf{(float)std::cos(a), (float)std::sin(a)};
  }
  std::normal_distribution<double> nd;
  std::vector<cf> z(N), Z;
  std::vector<cd> zd(N);
  for (int i = 0; i < N; ++i) {
    z[i] = cf{(float)nd(rng), (float)nd(rng)};
    zd[i] = cd(z[i].re, z[i].im);
  }
  if (!emu_fft<N, E, 1>(table, z, Z)) return 1;
  auto ref = dft(zd);
  double peak = 0, err = 0;
  for (int k = 0; k < N; ++k) {
    peak = std::max(peak, std::abs(ref[k]));
    err = std::max(err, std::abs(cd(Z[k].re, Z[k].im) - ref[k]));
  }
  const double rel = err / peak;
  const bool ok = rel < 2e-6;
  std::printf("schedule 1 N=%5d radices %d first, passes=%d fft_rel=%.3e %s\n", N,
              msig::Plan<N, E, 1>::radix(0), msig::Plan<N, E, 1>::NP, rel, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// The block-shared LDS twiddle table: every (pass, thread, register) read
// lands on an entry holding that register's twiddle exponent, and the FFT
// through the table is bit-identical to the per-thread register form.
template <int N, int E, int S, int P = 1>
static bool twl_index_ok() {
  using PL = msig::Plan<N, E, S>;
  if constexpr (P >= PL::NP) {
    return true;
  } else {
    constexpr int R = PL::radix(P);
    for (int t = 0; t < PL::TPF; ++t)
      for (int u = 0; u < E / R; ++u)
        for (int j = 1; j < R; ++j) {
          const int i = PL::twl_off(P) + PL::twl_slot(P, u, j) * PL::twl_cols(P) + PL::template twl_col<P>(t);
          if (i < 0 || i >= PL::TWL || PL::twl_exp(i) != PL::template tw_exp<P>(t, u * R + j)) {
            std::printf("N=%d S=%d pass %d t=%d u=%d j=%d: table entry %d wrong\n", N, S, P, t, u, j, i);
            return false;
          }
        }
    return twl_index_ok<N, E, S, P + 1>();
  }
}

template <int N, int S = 0>
static int check_shared_table(std::mt19937_64& rng) {
  constexpr int E = N < 16 ? N : 16;
  using PL = msig::Plan<N, E, S>;
  std::vector<cf> table(N);
  for (int i = 0; i < N; ++i) {
    const double a = -2.0 * M_PI * i / N;
    table[i] = cf{(float)std::cos(a), (float)std::sin(a)};
  }
  std::normal_distribution<double> nd;
  std::vector<cf> z(N), Zr, Zl;
  for (int i = 0; i < N; ++i) z[i] = cf{(float)nd(rng), (float)nd(rng)};
  bool ok = twl_index_ok<N, E, S>();
  if (!PL::TWL_DEDUP) {
    // an all-wide table is column t = thread t's load_twiddles registers
    std::vector<cf> regs(PL::TWN);
    for (int t = 0; ok && t < PL::TPF; ++t) {
      msig::load_twiddles<N, E, 1, const cf*, S>(t, table.data(), regs.data());
      for (int s = 0; ok && s < PL::TWN; ++s)
        ok = std::memcmp(&regs[s], &table[PL::twl_exp(s * PL::TPF + t)], sizeof(cf)) == 0;
    }
  }
  ok = ok && emu_fft<N, E, S, false>(table, z, Zr) && emu_fft<N, E, S, true>(table, z, Zl);
  for (int k = 0; ok && k < N; ++k)
    ok = std::memcmp(&Zr[k], &Zl[k], sizeof(cf)) == 0;
  std::printf("shared twiddle table N=%5d S=%d entries=%d (%d bytes, per-thread form %d) %s\n", N, S,
              PL::TWL, PL::TWL * 8, PL::TWN * PL::TPF * 8, ok ? "ok" : "FAIL");
  return ok ? 0 : 1;
}

// The split exchange's per-boundary pads (Plan::xpad_choice): every slot
// below Plan::LDS, a thread part plus a register part, and no repeated bank
// among any 32 lanes of a store or load under the 32-bank, 32-lane service
// of 4-byte LDS accesses (the model the bank-conflict counter confirms).
template <int N, int S = 0>
static int check_split_pads() {
  using PL = msig::Plan<N, 16, S>;
  int bad = 0;
  for (int p = 1; p < PL::NP; ++p) {
    const auto x = PL::xpad_choice(p);
    const int cost = PL::xcost(x, p - 1, true) + PL::xcost(x, p, false);
    if (!PL::xvalid(x, p) || cost != 0) {
      std::printf("N=%d S=%d boundary %d: pad (%d, %d) valid %d, %d conflict cycles\n", N, S, p, x.sh,
                  x.k, (int)PL::xvalid(x, p), cost);
      bad = 1;
    }
  }
  if (!bad) std::printf("N=%d S=%d split pads conflict-free ok\n", N, S);
  return bad;
}

int main() {
  std::mt19937_64 rng(1234);
  int bad = 0;
  bad += check_size<16>(rng);
  bad += check_size<32>(rng);
  bad += check_size<64>(rng);
  bad += check_size<128>(rng);
  bad += check_size<256>(rng);
  bad += check_size<512>(rng);
  bad += check_size<1024>(rng);
  bad += check_size<2048>(rng);
  bad += check_size<4096>(rng);
  bad += check_shared_table<128>(rng);
  bad += check_shared_table<256>(rng);
  bad += check_shared_table<512>(rng);
  bad += check_shared_table<1024>(rng);
  bad += check_shared_table<2048>(rng);
  bad += check_shared_table<4096>(rng);
  bad += check_shared_table<512, 1>(rng);
  bad += check_shared_table<2048, 1>(rng);
  // hop == N/2 register reuse: odd and even segment counts
  bad += check_stream<64, true>(rng, 64, 32, 64 + 32 * 14, 1);   // nseg 15
  bad += check_stream<64, true>(rng, 64, 32, 64 + 32 * 15, 0);   // nseg 16
  bad += check_stream<256, true>(rng, 256, 128, 256 + 128 * 8 + 50, 1);
  bad += check_stream<1024, true>(rng, 1024, 512, 1024 + 512 * 6, 1);
  // general stream: arbitrary hop, nperseg < nfft
  bad += check_stream<64, false>(rng, 64, 7, 300, 1);
  bad += check_stream<64, false>(rng, 50, 20, 333, 1);
  bad += check_stream<256, false>(rng, 256, 100, 5000, 0);
  bad += check_stream<256, false>(rng, 200, 200, 4000, 1);
  // virtual zero extension (stft boundary='zeros' + padded): lead = nperseg / 2
  bad += check_stream<64, true, true>(rng, 64, 32, 64 * 7 + 5, 0, 32, 32 + 27);
  bad += check_stream<256, true, true>(rng, 256, 128, 3000, 1, 128, 128 + 72);
  bad += check_stream<64, false, true>(rng, 64, 16, 1001, 0, 32, 32 + 7);
  bad += check_stream<256, false, true>(rng, 200, 50, 2222, 1, 100, 100 + 28);
  // the extended stream with no zeros (lead 0, every sample stored) is the plain one
  bad += check_stream<256, true, true>(rng, 256, 128, 256 + 128 * 8 + 50, 1);
  bad += check_stream<256, false, true>(rng, 256, 100, 5000, 0);
  // the peeled hop == N/2 group loop: odd / even nseg, runs that do and do
  // not divide the unit count, a run of one
  bad += check_half_groups<64>(rng, 15, 3, 1);
  bad += check_half_groups<64>(rng, 16, 3, 0);
  bad += check_half_groups<256>(rng, 21, 4, 1);
  bad += check_half_groups<256>(rng, 9, 1, 1);
  bad += check_half_groups<1024>(rng, 13, 2, 1);
  bad += check_half_groups<1024>(rng, 1, 4, 1);
  // the same with the standalone welch kernel's rotating two-unit loop
  bad += check_half_groups<64>(rng, 15, 3, 1, true);
  bad += check_half_groups<64>(rng, 16, 3, 0, true);
  bad += check_half_groups<256>(rng, 21, 4, 1, true);
  bad += check_half_groups<256>(rng, 9, 1, 1, true);
  bad += check_half_groups<1024>(rng, 13, 2, 1, true);
  bad += check_half_groups<1024>(rng, 1, 4, 1, true);
  bad += check_half_groups<1024>(rng, 33, 16, 1, true);
  bad += check_half_groups<256>(rng, 40, 7, 0, true);
  // radix schedule 1 wherever it differs from schedule 0
  bad += check_schedule<32>(rng);
  bad += check_schedule<64>(rng);
  bad += check_schedule<128>(rng);
  bad += check_schedule<512>(rng);
  bad += check_schedule<1024>(rng);
  bad += check_schedule<2048>(rng);
  // split-exchange pads at every size and schedule the split kernels use
  bad += check_split_pads<32>();
  bad += check_split_pads<64>();
  bad += check_split_pads<128>();
  bad += check_split_pads<256>();
  bad += check_split_pads<512>();
  bad += check_split_pads<1024>();
  bad += check_split_pads<2048>();
  bad += check_split_pads<4096>();
  bad += check_split_pads<512, 1>();
  bad += check_split_pads<1024, 1>();
  bad += check_split_pads<2048, 1>();
  return bad ? 1 : 0;
}
