// C ABI of libmsig.so: validation, the resident twiddle tables, the FP64
// segment-reduction/finalize kernel, and dispatch to the per-size kernels.
//
// Reference interfaces replaced (src/mlx_signal_processing/_stft_metal.py):
// eligible :150-160, rfft_frames :192-220, csd_frames / coherence_frames
// :378-419, rfft_power_mean :530-557, csd_mean / coherence_mean :703-746,
// _group_params / _reduce_groups :429-455.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <climits>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "msig.h"
#include "msig_launch.hpp"

namespace msig {

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

static int hip_fail(hipError_t e, const char* what) {
  return fail(MSIG_EIO, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

// ---------------------------------------------------------------- per-device tables
// What the library keeps per device, looked up without locks on the hot
// path: the twiddle table of each fused size, the resident-block count of
// each standalone mean kernel, and the CU count.  Entries are written once
// (under g_dev_mu) and read atomically.
constexpr int MAX_DEVICES = 64;
constexpr int NSIZES = 13;  // log2(nfft) <= 12
struct DeviceTables {
  std::atomic<const cf*> tw[NSIZES];
  std::atomic<int64_t> resident[NSIZES][3][3];  // [log2 N][kind][complex, general, half-hop]
  std::atomic<int> cus;
};
static DeviceTables g_dev[MAX_DEVICES];
static std::mutex g_dev_mu;

static int log2i(int v);

static int current_device(int* dev) {
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (*dev < 0 || *dev >= MAX_DEVICES)
    return fail(MSIG_EINVAL, "device %d beyond the %d devices the library tracks", *dev, MAX_DEVICES);
  return 0;
}

// exp(-2 pi i e / N), e < N, computed in FP64 and rounded once to float;
// built on first use per (device, N) and kept until the process ends.
static int twiddles_on(int dev, int N, const cf** out) {
  std::atomic<const cf*>& slot = g_dev[dev].tw[log2i(N)];
  const cf* t = slot.load(std::memory_order_acquire);
  if (t) {
    *out = t;
    return 0;
  }
  std::lock_guard<std::mutex> lk(g_dev_mu);
  t = slot.load(std::memory_order_acquire);
  if (!t) {
    std::vector<cf> h(N);
    for (int i = 0; i < N; ++i) {
      const double ang = -2.0 * M_PI * (double)i / (double)N;
      h[i] = cf{(float)std::cos(ang), (float)std::sin(ang)};
    }
    cf* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(cf) * N);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(twiddles)");
    e = hipMemcpy(d, h.data(), sizeof(cf) * N, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return hip_fail(e, "hipMemcpy(twiddles)");
    }
    slot.store(d, std::memory_order_release);
    t = d;
  }
  *out = t;
  return 0;
}

static int twiddles(int N, const cf** out) {
  int dev = 0;
  int rc = current_device(&dev);
  return rc ? rc : twiddles_on(dev, N, out);
}

// Table of (device, N) if it exists; never allocates.
static const cf* resident_twiddles(int dev, int N) {
  if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
  return g_dev[dev].tw[log2i(N)].load(std::memory_order_acquire);
}

// ---------------------------------------------------------------- no-allocation scratch
// Block partials of msig_spectral_mean_noalloc: one statically allocated
// array in each device's memory, held until the process ends.  In host code
// `scratch` is only a stand-in symbol; its address on a device comes from
// hipGetSymbolAddress with that device current (msig_init), and only that
// address is ever passed to kernels.  Every call on a device shares its
// array: calls that overlap in time on one device race.  The mean kernel
// writes every partial it later reads, so nothing needs clearing.
static __device__ __align__(256) float scratch[MSIG_NOALLOC_SCRATCH_BYTES / sizeof(float)];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scr_mu;
static float* g_scr[SCRATCH_MAX_DEVICES] = {};

// with `dev` current
static int scratch_setup(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(MSIG_EINVAL, "device %d beyond the %d devices the scratch table covers", dev,
                SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (g_scr[dev]) return 0;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(MSIG_EIO, "hipGetSymbolAddress(scratch) returned NULL");
  g_scr[dev] = static_cast<float*>(p);
  return 0;
}

static float* scratch_of(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scr_mu);
  return g_scr[dev];
}

// ---------------------------------------------------------------- kernel timing
// Event pairs handed out to the next calls on the armed device, one pair per
// fused transform kernel.  The vector is sized at arm time and never grows
// while armed, so a handed-out pair stays valid until read.
struct TimingEvents {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
  int device = -1;
};
static std::mutex g_kt_mu;
static TimingEvents g_kt;
static std::atomic<bool> g_kt_armed{false};  // skip the lock when nothing is armed

static void timing_release() {
  g_kt_armed.store(false, std::memory_order_release);
  for (auto& p : g_kt.ev) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  g_kt.ev.clear();
  g_kt.used = 0;
  g_kt.device = -1;
}

static std::pair<hipEvent_t, hipEvent_t>* timing_slot() {
  if (!g_kt_armed.load(std::memory_order_acquire)) return nullptr;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (g_kt.used >= g_kt.ev.size()) return nullptr;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != g_kt.device) return nullptr;
  return &g_kt.ev[g_kt.used++];
}

thread_local LaunchTiming g_launch_timing;

// Run `body` (which makes one fused transform launch) with the events of a
// timing slot, if one is armed, handed to that launch (launch_timed).
template <class F>
static hipError_t timed_launch(hipStream_t, F&& body) {
  auto* slot = timing_slot();
  if (!slot) return body();
  g_launch_timing = LaunchTiming{slot->first, slot->second};
  hipError_t e = body();
  g_launch_timing = LaunchTiming{};
  return e;
}

// ---------------------------------------------------------------- finalize
struct FinArgs {
  const float* ws;
  void* out0;
  void* out1;
  void* out2;
  int64_t rows;
  int32_t bpr, N, nbins, nc, kind, fold, apply_factor;
  double scale;    // scale / nseg
  int64_t ws_off;  // batch plans: partials start at ws_base + ws_off
  const float* side_src;  // side copy (grid-strided): side_len floats to side_dst
  float* side_dst;
  int64_t side_len;
};

constexpr int FIN_BINS = 64;   // output bins per block
constexpr int FIN_LANES = 4;   // threads sharing one bin's block partials

// Sum the block partials of each row in FP64, fold the conjugate mirror of
// the two-for-one packing, apply scale, 1/nseg and the onesided factor.
// Thread (bin j, lane q) sums partials q, q+4, q+8, ... with all its loads in
// flight together; the four lane sums are then added in lane order, so the
// reduction order is fixed by the job's geometry.  Counterpart of
// _reduce_groups (_stft_metal.py:443-455) and the deferred onesided factor of
// csd (spectral.py:482-488).
__device__ __forceinline__ void finalize_write(const FinArgs& f, int64_t row, int k, int j,
                                               const double (&part)[FIN_LANES][4][FIN_BINS][2]);

__device__ __forceinline__ void finalize_chunk(const FinArgs& f, int64_t row, int chunk) {
  __shared__ double part[FIN_LANES][4][FIN_BINS][2];
  const int j = threadIdx.x % FIN_BINS, q = threadIdx.x / FIN_BINS;
  const int k = chunk * FIN_BINS + j;
  const int N = f.N;
  const int kk = k < f.nbins ? k : 0;
  const int mk = (N - kk) & (N - 1);
  const float* base = f.ws + (size_t)row * f.bpr * f.nc * N;
  const size_t bstride = (size_t)f.nc * N;
  for (int c = 0; c < f.nc; ++c) {
    double s = 0.0, m = 0.0;
    int b = q;
    for (; b + 3 * FIN_LANES < f.bpr; b += 4 * FIN_LANES) {
      float v[4], w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float* pb = base + (size_t)(b + i * FIN_LANES) * bstride + (size_t)c * N;
        v[i] = pb[kk];
        w[i] = f.fold ? pb[mk] : 0.f;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        s += (double)v[i];
        m += (double)w[i];
      }
    }
    for (; b < f.bpr; b += FIN_LANES) {
      const float* pb = base + (size_t)b * bstride + (size_t)c * N;
      s += (double)pb[kk];
      if (f.fold) m += (double)pb[mk];
    }
    part[q][c][j][0] = s;
    part[q][c][j][1] = m;
  }
  __syncthreads();
  if (q == 0 && k < f.nbins) finalize_write(f, row, k, j, part);
  __syncthreads();
}

__device__ __forceinline__ void finalize_write(const FinArgs& f, int64_t row, int k, int j,
                                               const double (&part)[FIN_LANES][4][FIN_BINS][2]) {
  const int N = f.N;
  double p[4] = {0, 0, 0, 0};
  for (int c = 0; c < f.nc; ++c) {
    double s = 0.0, m = 0.0;
    for (int l = 0; l < FIN_LANES; ++l) {
      s += part[l][c][j][0];
      m += part[l][c][j][1];
    }
    // powers and real parts fold symmetric, imaginary parts antisymmetric
    const bool imag = (f.nc == 2 && c == 1) || (f.nc == 4 && c == 3);
    p[c] = f.fold ? 0.5 * (imag ? s - m : s + m) : s;
  }
  double sc = f.scale;
  if (f.apply_factor && !(k == 0 || 2 * k == N)) sc *= 2.0;
  const size_t o = (size_t)row * f.nbins + k;
  if (f.kind == MSIG_PSD_MEAN) {
    static_cast<float*>(f.out0)[o] = (float)(p[0] * sc);
  } else if (f.kind == MSIG_CSD_MEAN) {
    static_cast<cf*>(f.out0)[o] = cf{(float)(p[0] * sc), (float)(p[1] * sc)};
  } else if (f.kind == MSIG_COH_TRIO_MEAN) {
    static_cast<float*>(f.out0)[o] = (float)(p[0] * sc);
    static_cast<float*>(f.out1)[o] = (float)(p[1] * sc);
    static_cast<cf*>(f.out2)[o] = cf{(float)(p[2] * sc), (float)(p[3] * sc)};
  } else {  // MSIG_COHERENCE: scales cancel in the ratio
    const double num = p[2] * p[2] + p[3] * p[3];
    static_cast<float*>(f.out0)[o] = (float)(num / p[0] / p[1]);
  }
}

__global__ __launch_bounds__(FIN_BINS * FIN_LANES) void k_finalize(const FinArgs f) {
  side_copy(f.side_src, f.side_dst, f.side_len);
  const int chunks = (f.nbins + FIN_BINS - 1) / FIN_BINS;
  finalize_chunk(f, blockIdx.x / chunks, (int)(blockIdx.x % chunks));
}

// Finalize of a batch plan: a persistent grid over (job, row * chunks + chunk).
__global__ __launch_bounds__(FIN_BINS * FIN_LANES) void k_finalize_batch(
    const FinArgs* __restrict__ fins, const WorkItem* __restrict__ queue, int32_t nwork,
    const float* ws_base) {
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem item = queue[w];
    FinArgs f = fins[item.job];
    f.ws = ws_base + f.ws_off;
    const int chunks = (f.nbins + FIN_BINS - 1) / FIN_BINS;
    finalize_chunk(f, item.blk / chunks, item.blk % chunks);
  }
}

// ---------------------------------------------------------------- overlap-add
// out[row][j][c] = sum_s frames[row][s][i - s hop][c] / norm(i), i = first + j,
// over the segments s that cover sample i, in increasing s; norm(i) =
// sum_s win[i - s hop]^2, applied only where it exceeds 1e-10 (scipy's rule).
// A gather: each output sample is written once, so the sum order is fixed.
// Counterpart of the overlap-add kernel of istft_frames (_stft_metal.py:793-815).
struct OlaArgs {
  const float* frames;
  const float* win;
  float* out;
  int64_t first, count, rows;
  int32_t nseg, nperseg, hop, channels;
};

// One thread per output value, grid-strided over all rows (64-bit indices:
// any row count).
__global__ __launch_bounds__(256) void k_overlap_add(OlaArgs o) {
  const int64_t per_row = o.count * o.channels;
  const int64_t total = o.rows * per_row;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  const MSIG_GLOBAL float* w = as_global(o.win);
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
    const int64_t row = g / per_row;
    const int64_t idx = g - row * per_row;  // j * channels + c
    const int c = (int)(idx % o.channels);
    const int64_t i = o.first + idx / o.channels;
    int64_t s_hi = i / o.hop;
    if (s_hi > o.nseg - 1) s_hi = o.nseg - 1;
    const int64_t s_lo = i >= o.nperseg ? (i - o.nperseg) / o.hop + 1 : 0;
    const MSIG_GLOBAL float* fr =
        as_global(o.frames) + row * (int64_t)o.nseg * o.nperseg * o.channels;
    float acc = 0.f, nrm = 0.f;
    for (int64_t s = s_lo; s <= s_hi; ++s) {
      const int64_t off = i - s * o.hop;
      acc += fr[(s * o.nperseg + off) * o.channels + c];
      const float wv = w[off];
      nrm = fmaf(wv, wv, nrm);
    }
    as_global_mut(o.out)[g] = acc / (nrm > 1e-10f ? nrm : 1.f);
  }
}

// ---------------------------------------------------------------- median
// Per-column median over segments: frames [rows][nseg][cols] float32 ->
// out[row][col], numpy's median (the mean of the two middle values for an
// even count; NaN if the column holds a NaN) times scale and the onesided
// factor of the column's bin.
//
// A block stages CPB columns of one row in LDS as order-preserving uint32
// keys (one read of the frames, eight loads in flight per thread, the
// column stride chosen so a wave's staging writes hit 64 banks) and takes
// each column's key range [kmin, kmax] on the way.  The rank (nseg - 1) / 2
// key is then selected by radix passes over the range-relative keys
// k - kmin, 8 bits per pass, counting into LDS-atomic histograms: digits cut
// from the column's own range spread its keys over the buckets (a column of
// power values spans a few binary exponents only), a constant column needs
// no pass, and only the first pass touches every key.  Each pass locates the
// bucket holding the rank with a log-step scan of the column's per-lane
// bucket sums.  For an even count one more pass finds the next order
// statistic.  Counterpart of the median average of csd (spectral.py:469-477,
// mx.median over the segment axis).
constexpr int MED_BT = 256;
constexpr int MED_MAX_CPB = 16;
constexpr int MED_BITS = 8;
constexpr int MED_BUCKETS = 1 << MED_BITS;
// histogram slot of bucket b: one pad word per 32 buckets, so per-lane sums
// over contiguous bucket ranges hit distinct banks
constexpr int MED_HSTRIDE = MED_BUCKETS + MED_BUCKETS / 32;
__device__ __forceinline__ int med_slot(int b) { return b + (b >> 5); }
constexpr int MED_LDS_BUDGET = 40 << 10;   // bytes per block: four blocks per CU

struct MedArgs {
  const float* in;
  float* out;
  int64_t rows;
  int32_t nseg, cols, ncb, channels, apply_factor, nfft, stride;
  float scale;
};

// Column stride >= nseg with stride = 64 / cpb (mod 64): a staging write
// instruction covers 64 / cpb consecutive segments of each of the cpb
// columns, which then fall in 64 distinct banks.
__host__ __device__ __forceinline__ int med_stride(int nseg, int cpb) {
  const int target = cpb > 1 ? 64 / cpb : 0;
  return nseg + ((target - nseg % 64) % 64 + 64) % 64;
}

__host__ __forceinline__ size_t med_lds_bytes(int cpb, int stride) {
  return ((size_t)cpb * stride + (size_t)cpb * MED_HSTRIDE + MED_BT + 7 * MED_MAX_CPB) * 4;
}

__device__ __forceinline__ uint32_t med_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return b ^ ((b & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u);
}

__device__ __forceinline__ float med_val(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

template <int CPB>
__global__ __launch_bounds__(MED_BT) void k_median(MedArgs a) {
  extern __shared__ uint32_t med_smem[];
  constexpr int TPC = MED_BT / CPB;          // threads per column
  constexpr int PER = MED_BUCKETS / TPC;     // buckets each thread sums (>= 1)
  const int nseg = a.nseg, stride = a.stride;
  uint32_t* tile = med_smem;                                   // [CPB][stride]
  uint32_t* hist = tile + CPB * stride;                        // [CPB][MED_HSTRIDE]
  uint32_t* part = hist + CPB * MED_HSTRIDE;                   // [MED_BT]
  uint32_t* sel_bucket = part + MED_BT;                        // [MED_MAX_CPB] x 7
  uint32_t* sel_rank = sel_bucket + MED_MAX_CPB;
  uint32_t* cnt_le = sel_rank + MED_MAX_CPB;
  uint32_t* min_gt = cnt_le + MED_MAX_CPB;
  uint32_t* has_nan = min_gt + MED_MAX_CPB;
  uint32_t* kmin_s = has_nan + MED_MAX_CPB;
  uint32_t* kmax_s = kmin_s + MED_MAX_CPB;
  const int tid = threadIdx.x;
  const int col = tid / TPC, lt = tid % TPC;
  // Blocks are dispatched round-robin over the 8 XCDs (block b to XCD b % 8,
  // as that XCD's (b / 8)-th block).  Each XCD is given a contiguous range
  // of (row, column-block) tiles, so neighbouring column blocks of a row,
  // which read the same cache lines of every segment, share one L2 instead
  // of each XCD fetching the lines for a few of their bytes.
  const uint32_t nb = gridDim.x, bx = blockIdx.x;
  const uint32_t per = nb / 8u, rem = nb % 8u, xcd = bx % 8u;
  const uint32_t tile_id = xcd * per + (xcd < rem ? xcd : rem) + bx / 8u;
  const int64_t row = tile_id / a.ncb;
  const int col0 = (int)(tile_id - row * a.ncb) * CPB;
  const int ncol = a.cols - col0 < CPB ? a.cols - col0 : CPB;  // columns present
  const MSIG_GLOBAL float* src = as_global(a.in) + row * (int64_t)nseg * a.cols + col0;

  if (tid < MED_MAX_CPB) {
    has_nan[tid] = 0u;
    cnt_le[tid] = 0u;
    min_gt[tid] = 0xFFFFFFFFu;
    kmin_s[tid] = 0xFFFFFFFFu;
    kmax_s[tid] = 0u;
  }
  __syncthreads();
  // staging: element i = (segment i / CPB, column i % CPB); every element of
  // a thread lies in column tid % CPB (MED_BT is a multiple of CPB)
  constexpr int UNR = 8;
  const int total = nseg * CPB;
  const int my_c = tid % CPB;
  uint32_t lmin = 0xFFFFFFFFu, lmax = 0u;
  bool nan_seen = false;
  for (int base = 0; base < total; base += MED_BT * UNR) {
    float v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int i = base + u * MED_BT + tid;
      v[u] = (i < total && my_c < ncol) ? src[(int64_t)(i / CPB) * a.cols + my_c] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int i = base + u * MED_BT + tid;
      if (i < total) {
        const uint32_t k = med_key(v[u]);
        nan_seen |= v[u] != v[u];
        tile[my_c * stride + i / CPB] = k;
        lmin = k < lmin ? k : lmin;
        lmax = k > lmax ? k : lmax;
      }
    }
  }
  if (lmin != 0xFFFFFFFFu) {
    atomicMin(&kmin_s[my_c], lmin);
    atomicMax(&kmax_s[my_c], lmax);
  }
  if (nan_seen) has_nan[my_c] = 1u;
  __syncthreads();

  const uint32_t r0 = (uint32_t)(nseg - 1) / 2;
  const uint32_t kmin = kmin_s[col];
  const uint32_t range = kmax_s[col] - kmin;
  const int nbits = range ? 32 - __builtin_clz(range) : 0;  // bits of k - kmin
  const uint32_t* keys = tile + col * stride;
  uint32_t* h = hist + col * MED_HSTRIDE;
  uint32_t prefix = 0u, rank = r0;  // prefix: bits >= hi of the selected k - kmin
  // every column of the block runs the block's largest pass count, so the
  // barriers are uniform; a column past its own bits idles
  int hi_max = 0;
#pragma unroll
  for (int c = 0; c < CPB; ++c) {
    const uint32_t rc = kmax_s[c] - kmin_s[c];
    const int nb = rc ? 32 - __builtin_clz(rc) : 0;
    hi_max = nb > hi_max ? nb : hi_max;
  }
  int hi = nbits;
  for (int hb = hi_max; hb > 0; hb -= MED_BITS) {
    const int lo = hi - MED_BITS > 0 ? hi - MED_BITS : 0;
    const int w = hi - lo;
    const bool active = w > 0;
    for (int b = lt; b < MED_HSTRIDE; b += TPC) h[b] = 0u;
    __syncthreads();
    if (active) {
      const uint32_t dm = (1u << w) - 1u;
      for (int i = lt; i < nseg; i += TPC) {
        const uint32_t r = keys[i] - kmin;
        if ((uint32_t)((uint64_t)r >> hi) == prefix)
          atomicAdd(&h[med_slot((int)((r >> lo) & dm))], 1u);
      }
    }
    __syncthreads();
    uint32_t c_own[PER];
    uint32_t sum = 0u;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      c_own[q] = h[med_slot(lt * PER + q)];
      sum += c_own[q];
    }
    // exclusive prefix of the column's per-thread sums: log-step scan over
    // the column's lanes of each wave, then the totals of earlier waves
    constexpr int CW = TPC < 64 ? TPC : 64;  // lanes of the column in one wave
    const int wl = tid & 63;
    uint32_t incl = sum;
#pragma unroll
    for (int d = 1; d < CW; d <<= 1) {
      const uint32_t v = __shfl_up(incl, d, CW);
      if ((wl % CW) >= d) incl += v;
    }
    if (TPC > 64 && wl == 63) part[tid >> 6] = incl;  // a column spans whole waves
    __syncthreads();
    uint32_t before = incl - sum;
    if constexpr (TPC > 64)
      for (int wv = (col * TPC) >> 6; wv < (tid >> 6); ++wv) before += part[wv];
    if (active && rank >= before && rank < before + sum) {
      uint32_t cum = before;
      bool found = false;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (!found && rank < cum + c_own[q]) {
          sel_bucket[col] = (uint32_t)(lt * PER + q);
          sel_rank[col] = rank - cum;
          found = true;
        }
        cum += c_own[q];
      }
    }
    __syncthreads();
    if (active) {
      prefix = (prefix << w) | sel_bucket[col];
      rank = sel_rank[col];
      hi = lo;
    }
    __syncthreads();  // sel_* and part read by every thread before the next pass writes them
  }
  const uint32_t k0 = kmin + prefix;
  const bool even = (nseg & 1) == 0;
  uint32_t k1 = k0;
  if (even) {
    uint32_t c = 0u, mn = 0xFFFFFFFFu;
    for (int i = lt; i < nseg; i += TPC) {
      const uint32_t k = keys[i];
      if (k <= k0) ++c;
      else mn = k < mn ? k : mn;
    }
    atomicAdd(&cnt_le[col], c);
    atomicMin(&min_gt[col], mn);
    __syncthreads();
    k1 = cnt_le[col] > r0 + 1 ? k0 : min_gt[col];
  }
  const int gc = col0 + col;
  if (lt == 0 && col < ncol) {
    float m = even ? 0.5f * (med_val(k0) + med_val(k1)) : med_val(k0);
    if (has_nan[col]) m = __uint_as_float(0x7FC00000u);
    const int bin = gc / a.channels;
    const float f = (a.apply_factor && !(bin == 0 || 2 * bin == a.nfft)) ? 2.f : 1.f;
    as_global_mut(a.out)[row * a.cols + gc] = m * (a.scale * f);
  }
}

// Register-resident median for columns of up to 8192 segments (the staged
// kernel above takes longer ones).  A 512-thread block takes CPB columns of
// one row (16 up to 1024 segments, 8 up to 2048, else 4); thread (column c,
// lane l) holds segments l, l + TPC, l + 2 TPC, ... of column c as keys in
// registers, TPC = 512 / CPB.  The lanes of a wave hold different columns,
// so one load instruction reads CPB contiguous floats of each of 64 / CPB
// segment rows (against 16 bytes per row in the staged kernel), issued as raw
// buffer loads whose per-load step is a scalar offset (reads past the row
// return zero), and a wave's histogram atomics go to different columns'
// bins.  Slots past the column's end hold copies of its minimum, which only
// shift the rank.  Selection: one radix pass on the top 9-10 bits of the
// range-relative keys picks the rank's bucket (a dozen keys of a spectral
// column, against some fifty at 8 bits); those keys are gathered (one
// reservation per thread), and a column with at most MEDR_RANKED of them
// takes its answer by counting, for each, the gathered keys below and equal
// to it.  Columns with more run further radix passes over the gathered keys
// (at most MEDR_CAP) or over their registers.  An even count adds one pass
// for the next order statistic.  Measured on 64 x 2047 x 513: 114 us against
// 182 us for the staged kernel (loads 58 us, first pass 14, gather 32, rank
// 12).
constexpr int MEDR_BT = 512;
constexpr int MEDR_CAP = 256;  // candidate keys kept per column after the first pass
constexpr int MEDR_RANKED = 64;  // candidates ranked by counting (else more radix passes)

template <int CPB, int KPT, int BTM = MEDR_BT>
__global__ __launch_bounds__(BTM) void k_median_reg(MedArgs a) {
  constexpr int TPC = BTM / CPB;             // threads per column
  // digits of BITS bits: the first pass then leaves about a dozen keys of a
  // spectral column in the rank's bucket (some fifty at 8 bits), few enough
  // to rank by counting
  constexpr int BITS = CPB >= 16 ? 9 : 10;
  constexpr int BUCKETS = 1 << BITS;
  constexpr int HSTRIDE = BUCKETS + BUCKETS / 32;
  constexpr int PER = BUCKETS / TPC;         // buckets each scan thread sums
  static_assert(PER >= 1 && BUCKETS % TPC == 0, "scan shape");
  __shared__ uint32_t hist[CPB * HSTRIDE];
  __shared__ uint32_t cand[CPB * MEDR_CAP];
  __shared__ uint32_t ccnt[CPB];
  __shared__ uint32_t part[BTM / 64];
  __shared__ uint32_t sel_bucket[CPB], sel_rank[CPB], cnt_le[CPB], min_gt[CPB], has_nan[CPB],
      kmin_s[CPB], kmax_s[CPB], done_s[CPB], sel_key[CPB];
  const int tid = threadIdx.x;
  const int c = tid % CPB, l = tid / CPB;  // keys: column c, segments l + TPC k
  // XCD-aware tile order, as in k_median
  const uint32_t nb = gridDim.x, bx = blockIdx.x;
  const uint32_t per = nb / 8u, rem = nb % 8u, xcd = bx % 8u;
  const uint32_t tile_id = xcd * per + (xcd < rem ? xcd : rem) + bx / 8u;
  const int64_t row = tile_id / a.ncb;
  const int col0 = (int)(tile_id - row * a.ncb) * CPB;
  const int ncol = a.cols - col0 < CPB ? a.cols - col0 : CPB;
  const int nseg = a.nseg;
  const bool colok = c < ncol;
  // slot k of this thread is segment l + TPC k: every thread's slots below
  // kfull are segments, none above it, and slot kfull is for l < tail (the
  // tests stay scalar except at k == kfull)
  const int kfull = nseg / TPC;
  const bool tail_ok = l < nseg - kfull * TPC;
  auto valid = [&](int k) { return k < kfull || (k == kfull && tail_ok); };
  if (tid < CPB) {
    done_s[tid] = 0u;
    has_nan[tid] = 0u;
    cnt_le[tid] = 0u;
    min_gt[tid] = 0xFFFFFFFFu;
    kmin_s[tid] = 0xFFFFFFFFu;
    kmax_s[tid] = 0u;
  }
  // Raw buffer loads over this row of frames (nseg * cols floats): the
  // per-load step (TPC segment rows) is a scalar offset, so the KPT loads
  // share one address register, and reads past the row return zero.
  const float* rowp = a.in + row * (int64_t)nseg * a.cols;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(rowp), (short)0, nseg * a.cols * (int)sizeof(float), 0x00020000);
  const int voff = (l * a.cols + col0 + c) * (int)sizeof(float);
  const int sstep = TPC * a.cols * (int)sizeof(float);
  uint32_t key[KPT];
#pragma unroll
  for (int k = 0; k < KPT; ++k)
    key[k] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, k * sstep, 0);
  uint32_t lmin = 0xFFFFFFFFu, lmax = 0u;
  bool nan_seen = false;
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const float x = __uint_as_float(key[k]);
    key[k] = med_key(x);
    if (valid(k)) {
      nan_seen |= x != x;
      lmin = key[k] < lmin ? key[k] : lmin;
      lmax = key[k] > lmax ? key[k] : lmax;
    }
  }
  __syncthreads();  // shared state initialised
  if (colok && valid(0)) {
    atomicMin(&kmin_s[c], lmin);
    atomicMax(&kmax_s[c], lmax);
  }
  if (colok && nan_seen) has_nan[c] = 1u;
  __syncthreads();

  // From here on keys are relative to the column minimum, and the slots past
  // the column's segments hold copies of the minimum (relative 0): PAD of
  // them sort first, so the rank-r element of the column is the rank
  // r + PAD element of the padded slots, and no pass tests validity.
  constexpr uint32_t SLOTS = (uint32_t)TPC * KPT;
  const uint32_t pad = SLOTS - (uint32_t)nseg;
  const uint32_t r0 = (uint32_t)(nseg - 1) / 2 + pad;
  const uint32_t kmin = kmin_s[c];
  const uint32_t range = kmax_s[c] - kmin;
  const int nbits = (colok && range) ? 32 - __builtin_clz(range) : 0;
#pragma unroll
  for (int k = 0; k < KPT; ++k) key[k] = valid(k) ? key[k] - kmin : 0u;
  int hi_max = 0;
#pragma unroll
  for (int cc = 0; cc < CPB; ++cc) {
    const uint32_t rc = cc < ncol ? kmax_s[cc] - kmin_s[cc] : 0u;
    const int b = rc ? 32 - __builtin_clz(rc) : 0;
    hi_max = b > hi_max ? b : hi_max;
  }
  // scan mapping: column sc's TPC threads are contiguous
  const int sc = tid / TPC, sl = tid % TPC;
  const uint32_t sc_kmin = kmin_s[sc < CPB ? sc : 0];
  const uint32_t sc_range = kmax_s[sc < CPB ? sc : 0] - sc_kmin;
  const int sc_nbits = (sc < ncol && sc_range) ? 32 - __builtin_clz(sc_range) : 0;
  uint32_t prefix = 0u, rank = r0;  // key-mapping view (column c)
  uint32_t s_rank = r0;             // scan-mapping view (column sc)
  int hi = nbits, s_hi = sc_nbits;
  int pass = 0;
  for (int hb = hi_max; hb > 0; hb -= BITS, ++pass) {
    const int lo = hi - BITS > 0 ? hi - BITS : 0;
    const int w = hi - lo;
    const int s_lo = s_hi - BITS > 0 ? s_hi - BITS : 0;
    const int s_w = s_hi - s_lo;
    for (int i = tid; i < CPB * HSTRIDE; i += BTM) hist[i] = 0u;
    __syncthreads();
    if (w > 0) {
      const uint32_t dm = (1u << w) - 1u;
      uint32_t* h = hist + c * HSTRIDE;
      const uint32_t nc = pass >= 1 ? ccnt[c] : MEDR_CAP + 1u;
      if (nc <= MEDR_CAP) {
        for (uint32_t i = (uint32_t)l; i < nc; i += TPC) {
          const uint32_t r = cand[c * MEDR_CAP + i];
          if ((uint32_t)((uint64_t)r >> hi) == prefix)
            atomicAdd(&h[med_slot((int)((r >> lo) & dm))], 1u);
        }
      } else if (pass == 0) {
        // every slot is under the empty prefix (r <= range < 2^hi): no test
#pragma unroll
        for (int k = 0; k < KPT; ++k) atomicAdd(&h[med_slot((int)((key[k] >> lo) & dm))], 1u);
      } else {
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
          const uint32_t r = key[k];
          if ((uint32_t)((uint64_t)r >> hi) == prefix)
            atomicAdd(&h[med_slot((int)((r >> lo) & dm))], 1u);
        }
      }
    }
    __syncthreads();
    const uint32_t* hs = hist + (sc < CPB ? sc : 0) * HSTRIDE;
    uint32_t sum = 0u;
#pragma unroll
    for (int q = 0; q < PER; ++q) sum += hs[med_slot(sl * PER + q)];
    constexpr int CW = TPC < 64 ? TPC : 64;
    const int wl = tid & 63;
    uint32_t incl = sum;
#pragma unroll
    for (int d = 1; d < CW; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, CW);
      if ((wl % CW) >= d) incl += t;
    }
    if (TPC > 64 && wl == 63) part[tid >> 6] = incl;
    __syncthreads();
    uint32_t before = incl - sum;
    if constexpr (TPC > 64)
      for (int wv = (sc * TPC) >> 6; wv < (tid >> 6); ++wv) before += part[wv];
    if (s_w > 0 && sc < ncol && s_rank >= before && s_rank < before + sum) {
      // the one lane whose buckets hold the rank reads them again
      uint32_t cum = before;
      bool found = false;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const uint32_t cq = hs[med_slot(sl * PER + q)];
        if (!found && s_rank < cum + cq) {
          sel_bucket[sc] = (uint32_t)(sl * PER + q);
          sel_rank[sc] = s_rank - cum;
          found = true;
        }
        cum += cq;
      }
    }
    __syncthreads();
    if (w > 0) {
      prefix = (prefix << w) | sel_bucket[c];
      rank = sel_rank[c];
      hi = lo;
    }
    if (s_w > 0) {
      s_rank = sel_rank[sc];
      s_hi = s_lo;
    }
    __syncthreads();  // sel_* and part read before the next pass writes them
    if (pass == 0 && hb > BITS) {
      // Collect each column's keys under the prefix the first pass chose
      // (a few dozen for spectra): a column with at most MEDR_RANKED takes
      // its answer by counting, for each candidate, the candidates below and
      // equal to it, and leaves the passes; one with more than MEDR_CAP keeps
      // reading its registers in them.
      if (tid < CPB) ccnt[tid] = 0u;
      __syncthreads();
      if (hi > 0) {
        // the thread's matches as a bit mask, one reservation for all of them
        static_assert(KPT <= 64, "match mask");
        using Mask = typename std::conditional<(KPT > 32), uint64_t, uint32_t>::type;
        Mask hit = 0u;
#pragma unroll
        for (int k = 0; k < KPT; ++k)
          hit |= (Mask)((uint32_t)((uint64_t)key[k] >> hi) == prefix) << k;
        if (hit) {
          uint32_t pos = atomicAdd(&ccnt[c], (uint32_t)__builtin_popcountll((uint64_t)hit));
#pragma unroll
          for (int k = 0; k < KPT; ++k) {
            if ((hit >> k) & 1u) {
              if (pos < MEDR_CAP) cand[c * MEDR_CAP + pos] = key[k];
              ++pos;
            }
          }
        }
      }
      __syncthreads();
      const uint32_t m = ccnt[sc];
      if (s_hi > 0 && sc < ncol && m <= (uint32_t)MEDR_RANKED) {
        const uint32_t* cs = cand + sc * MEDR_CAP;
        for (uint32_t j = (uint32_t)sl; j < m; j += TPC) {
          const uint32_t x = cs[j];
          uint32_t less = 0u, eq = 0u;
          // four candidates per step, their reads issued together
          for (uint32_t i0 = 0; i0 < m; i0 += 4) {
            uint32_t y[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) y[u] = cs[i0 + u < m ? i0 + u : m - 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const bool in = i0 + u < m;
              less += (in && y[u] < x) ? 1u : 0u;
              eq += (in && y[u] == x) ? 1u : 0u;
            }
          }
          if (less <= s_rank && s_rank < less + eq) sel_key[sc] = x;  // equal keys: one value
        }
        if (sl == 0) done_s[sc] = 1u;
      }
      __syncthreads();
      if (done_s[c]) hi = 0;
      if (done_s[sc]) s_hi = 0;
      bool more = false;  // any column still to refine (block-uniform)
#pragma unroll
      for (int cc = 0; cc < CPB; ++cc) {
        const uint32_t rc = cc < ncol ? kmax_s[cc] - kmin_s[cc] : 0u;
        const int b = rc ? 32 - __builtin_clz(rc) : 0;
        more |= done_s[cc] == 0u && b > BITS;
      }
      if (!more) break;
    }
  }
  (void)rank;
  if (done_s[c]) prefix = sel_key[c];  // relative key of the rank-r0 slot
  const uint32_t k0 = kmin + prefix;
  const bool even = (nseg & 1) == 0;
  uint32_t k1 = k0;
  if (even) {
    uint32_t cle = 0u, mn = 0xFFFFFFFFu;  // slots <= k0 (padding included); least key above
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      if (key[k] <= prefix) ++cle;
      else mn = key[k] + kmin < mn ? key[k] + kmin : mn;
    }
    if (colok) {
      atomicAdd(&cnt_le[c], cle);
      atomicMin(&min_gt[c], mn);
    }
    __syncthreads();
    k1 = cnt_le[c] > r0 + 1 ? k0 : min_gt[c];
  }
  if (l == 0 && colok) {
    const int gc = col0 + c;
    float m = even ? 0.5f * (med_val(k0) + med_val(k1)) : med_val(k0);
    if (has_nan[c]) m = __uint_as_float(0x7FC00000u);
    const int bin = gc / a.channels;
    const float f = (a.apply_factor && !(bin == 0 || 2 * bin == a.nfft)) ? 2.f : 1.f;
    as_global_mut(a.out)[row * a.cols + gc] = m * (a.scale * f);
  }
}

// (columns per block, keys per thread) of the register kernel for a column
// length, or {0, 0} when it is longer than the register kernel holds.
struct MedRegShape {
  int cpb, kpt, bt;
};
static MedRegShape medr_shape(int nseg, int64_t columns) {
  if (nseg <= 16 * (MEDR_BT / 16)) return {16, 16, MEDR_BT};
  // 513-1024 segments: 256-thread blocks of eight columns (four blocks per
  // CU) once the call has a few rounds of them: 128 x 1023 x 513 ran 187 ->
  // 132 us against 512-thread blocks of sixteen; small calls (16 x 1023 x
  // 129: 17.5 against 18.7 us) and 512 or fewer segments (256 x 511 x 513:
  // 151 against 186 us) keep the wider blocks.
  if (nseg <= 32 * (MEDR_BT / 16)) return columns >= 8192 ? MedRegShape{8, 32, 256} : MedRegShape{16, 32, MEDR_BT};
  // 256-thread blocks of four columns (four blocks per CU, 21 KB of LDS each)
  // against 512-thread blocks of eight (two per CU): 121.4 -> 117.5 us on
  // 64 x 2047 x 513.  Halving the blocks of the shorter shapes was slower.
  if (nseg <= 32 * 64) return {4, 32, 256};
  if (nseg <= 32 * (MEDR_BT / 4)) return {4, 32, MEDR_BT};
  if (nseg <= 64 * (MEDR_BT / 4)) return {4, 64, MEDR_BT};
  return {0, 0, 0};
}

// ---------------------------------------------------------------- dispatch
static bool pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

// hop == N/2 with full-length segments: each sample is loaded once per group
static bool half_hop(const KArgs& a, int N, bool cplx) {
  return !cplx && a.nperseg == N && 2 * a.hop == N;
}

static hipError_t dispatch_mean(int N, int kk, bool cplx, const KArgs& a, int64_t rows,
                                hipStream_t s) {
  const bool half = half_hop(a, N, cplx);
  switch (N) {
#define C(NV) \
  case NV:    \
    return launch_mean_##NV(kk, cplx, half, a, rows, s);
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  return hipErrorInvalidValue;
}

static hipError_t dispatch_frames(int N, int ok, bool cplx, bool ext, const KArgs& a,
                                  int64_t rows, hipStream_t s) {
  const bool half = half_hop(a, N, cplx);
  switch (N) {
#define C(NV) \
  case NV:    \
    return launch_frames_##NV(ok, cplx, half, ext, a, rows, s);
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  return hipErrorInvalidValue;
}

static hipError_t dispatch_istft_ola(int N, const KArgs& a, int64_t rows, hipStream_t s) {
  switch (N) {
#define C(NV) \
  case NV:    \
    return launch_istft_ola_##NV(a, rows, s);
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  return hipErrorInvalidValue;
}

static hipError_t dispatch_istft(int N, const KArgs& a, int64_t rows, hipStream_t s) {
  switch (N) {
#define C(NV) \
  case NV:    \
    return launch_istft_##NV(a, rows, s);
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  return hipErrorInvalidValue;
}

static hipError_t dispatch_mean_batch(int N, int kk, bool cplx, bool half, int sched,
                                      const KArgs* jobs,
                                      const WorkItem* queue, int32_t nwork, float* ws, int grid,
                                      hipStream_t s) {
  switch (N) {
#define C(NV) \
  case NV:    \
    return launch_mean_batch_##NV(kk, cplx, half, sched, jobs, queue, nwork, ws, grid, s);
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  return hipErrorInvalidValue;
}

static int ncomp(int kind) {
  switch (kind) {
    case MSIG_PSD_MEAN: return 1;
    case MSIG_CSD_MEAN: return 2;
    default: return 4;
  }
}

static bool is_mean(int kind) {
  return kind == MSIG_PSD_MEAN || kind == MSIG_CSD_MEAN || kind == MSIG_COH_TRIO_MEAN ||
         kind == MSIG_COHERENCE;
}

static bool is_pair(int kind) {
  return kind == MSIG_CSD_MEAN || kind == MSIG_COH_TRIO_MEAN || kind == MSIG_COHERENCE ||
         kind == MSIG_CSD_FRAMES || kind == MSIG_COH_FRAMES;
}

static int check_precision(const msig_job* j);

// samples stored per row (the extended row's length without the virtual zeros)
static int64_t stored_len(const msig_job* j) { return j->nvalid ? j->nvalid : j->n - j->lead; }

static int check_job(const msig_job* j) {
  if (!j) return fail(MSIG_EINVAL, "job is NULL");
  if (j->kind < MSIG_PSD_MEAN || j->kind > MSIG_COH_FRAMES)
    return fail(MSIG_EINVAL, "unknown kind %d", j->kind);
  if (j->rows < 1) return fail(MSIG_EINVAL, "rows must be >= 1 (got %lld)", (long long)j->rows);
  if (j->nperseg < 1) return fail(MSIG_EINVAL, "nperseg must be >= 1 (got %d)", j->nperseg);
  if (j->nstep < 1) return fail(MSIG_EINVAL, "nstep must be >= 1 (got %d)", j->nstep);
  if (j->nfft < j->nperseg)
    return fail(MSIG_EINVAL, "nfft (%d) must be >= nperseg (%d)", j->nfft, j->nperseg);
  if (j->n < j->nperseg)
    return fail(MSIG_EINVAL, "n (%lld) must be >= nperseg (%d)", (long long)j->n, j->nperseg);
  if (j->lead < 0 || j->nvalid < 0) return fail(MSIG_EINVAL, "lead and nvalid must be >= 0");
  if (j->side_len < 0 || (j->side_len > 0 && (!j->side_src || !j->side_dst)))
    return fail(MSIG_EINVAL, "side copy: side_len must be >= 0 with both buffers given");
  const bool ext = j->lead != 0 || j->nvalid != 0;
  if (ext && j->kind != MSIG_FRAMES_COMPLEX)
    return fail(MSIG_ENOTSUP,
                "the virtual zero extension (lead / nvalid) is for complex per-frame spectra");
  const int64_t stored = stored_len(j);
  if (stored < 1 || j->lead + stored > j->n)
    return fail(MSIG_EINVAL, "lead (%lld) + nvalid (%lld) must lie in [1, n = %lld]",
                (long long)j->lead, (long long)stored, (long long)j->n);
  if (j->ld_x < stored) return fail(MSIG_EINVAL, "ld_x < stored samples per row");
  if (is_pair(j->kind) && j->ld_y < stored) return fail(MSIG_EINVAL, "ld_y < stored samples per row");
  const int64_t nseg = (j->n - j->nperseg) / j->nstep + 1;
  if (nseg != j->nseg)
    return fail(MSIG_EINVAL, "nseg %d != (n - nperseg) / nstep + 1 = %lld", j->nseg,
                (long long)nseg);
  if (nseg > (int64_t)1 << 30) return fail(MSIG_EINVAL, "nseg too large");
  if (!pow2(j->nfft) || j->nfft < MSIG_MIN_NFFT || j->nfft > MSIG_MAX_NFFT)
    return fail(MSIG_ENOTSUP, "nfft %d outside the fused power-of-two range [%d, %d]", j->nfft,
                MSIG_MIN_NFFT, MSIG_MAX_NFFT);
  return check_precision(j);
}

static int nbins_of(const msig_job* j) {
  const bool cplx = j->flags & MSIG_F_COMPLEX_INPUT;
  return (!cplx && (j->flags & MSIG_F_ONESIDED)) ? j->nfft / 2 + 1 : j->nfft;
}

// Precision selected by the MSIG_F_MFMA_* flags.
static int precision_of(const msig_job* j) {
  if (j->flags & MSIG_F_MFMA_F16) return PREC_F16;
  if (j->flags & MSIG_F_MFMA_BF16) return PREC_BF16;
  return PREC_FP32;
}

// The split-precision form covers the welch mean at nperseg = nfft = 256.
static int check_precision(const msig_job* j) {
  const int prec = precision_of(j);
  if (prec == PREC_FP32) return 0;
  if ((j->flags & MSIG_F_MFMA_F16) && (j->flags & MSIG_F_MFMA_BF16))
    return fail(MSIG_EINVAL, "MSIG_F_MFMA_F16 and MSIG_F_MFMA_BF16 are exclusive");
  if (j->kind != MSIG_PSD_MEAN || (j->flags & MSIG_F_COMPLEX_INPUT) ||
      j->nfft != MFMA_NFFT || j->nperseg != MFMA_NFFT)
    return fail(MSIG_ENOTSUP,
                "the split-precision matrix-core form covers real-input PSD means with "
                "nperseg = nfft = %d only",
                MFMA_NFFT);
  return 0;
}

static int mean_kk(int kind) {
  return kind == MSIG_PSD_MEAN ? K_PSD : (kind == MSIG_CSD_MEAN ? K_CSD : K_COH);
}

static bool job_half(const msig_job* j) {
  return !(j->flags & MSIG_F_COMPLEX_INPUT) && j->nperseg == j->nfft && 2 * j->nstep == j->nfft;
}

// Blocks the current device holds at once of the standalone mean kernel for
// (N, kind, complex, half-hop): its occupancy per CU times the CUs, cached
// per device.  Without a device (host-only planning, nothing launched) the
// kernel's launch bound on 256 CUs stands in.
static int64_t resident_blocks(int N, int kk, bool cplx, bool half) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    (void)hipGetLastError();
    dev = -1;
  }
  const int var = cplx ? 0 : (half ? 2 : 1);
  std::atomic<int64_t>* slot = dev >= 0 ? &g_dev[dev].resident[log2i(N)][kk][var] : nullptr;
  if (slot) {
    const int64_t r = slot->load(std::memory_order_acquire);
    if (r > 0) return r;
  }
  int cus = 256;
  if (dev >= 0) {
    cus = g_dev[dev].cus.load(std::memory_order_acquire);
    if (cus <= 0) {
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          cus < 1) {
        (void)hipGetLastError();
        cus = 256;
      } else {
        g_dev[dev].cus.store(cus, std::memory_order_release);
      }
    }
  }
  int per_cu = 1;
  switch (N) {
#define C(NV) \
  case NV:    \
    per_cu = mean_blocks_per_cu_##NV(kk, cplx, half); \
    break;
    C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096)
#undef C
  }
  const int64_t r = (int64_t)(per_cu > 0 ? per_cu : 1) * cus;
  if (slot) slot->store(r, std::memory_order_release);
  return r;
}

// Plan of an FP32 mean job: sized to the resident blocks by default, or to
// a tuned block target (batch plans under a tuning table).
static Geo mean_geometry_fp32(const msig_job* j, int maxrun, int64_t target_blocks) {
  const bool cplx = j->flags & MSIG_F_COMPLEX_INPUT;
  if (target_blocks > 0) return geometry(j->rows, j->nseg, j->nfft, cplx, maxrun, target_blocks);
  const int64_t res = resident_blocks(j->nfft, mean_kk(j->kind), cplx, job_half(j));
  return geometry_resident(j->rows, j->nseg, j->nfft, cplx, maxrun, res);
}

static Geo mean_geometry(const msig_job* j) {
  if (precision_of(j) != PREC_FP32) return geometry_mfma(j->rows, j->nseg, MEAN_MAX_RUN);
  return mean_geometry_fp32(j, MEAN_MAX_RUN, 0);
}

// Mean-kind workspace: the block partials [rows][bpr][ncomp][nfft] float32.
static size_t mean_ws_bytes(const msig_job* j) {
  const Geo g = mean_geometry(j);
  return (size_t)j->rows * g.bpr * ncomp(j->kind) * j->nfft * sizeof(float);
}

static KArgs base_args(const msig_job* j, const void* x, const void* y, const float* win,
                       const cf* tw, const Geo& g) {
  KArgs a;
  std::memset(&a, 0, sizeof a);
  a.x = x;
  a.y = y;
  a.win = win;
  a.twt = tw;
  a.ld_x = j->ld_x;
  a.ld_y = j->ld_y;
  a.nseg = j->nseg;
  a.nperseg = j->nperseg;
  a.hop = j->nstep;
  a.units = g.units;
  a.run = g.run;
  a.bpr = g.bpr;
  a.detrend = (j->flags & MSIG_F_DETREND_CONST) ? 1 : 0;
  a.lead = j->lead;
  a.nvalid = stored_len(j);
  a.side_src = j->side_src;
  a.side_dst = j->side_dst;
  a.side_len = j->side_len;
  return a;
}

static int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

}  // namespace msig

// Opaque batch plan: the per-job descriptors and work queues on the host,
// and the launch list of one run.  The device copy lives in caller memory.
struct msig_batch {
  struct Class {
    int N, kk, sched;
    bool cplx, half;
    int32_t qoff, nwork;
    int grid;
  };
  int device = 0;
  std::vector<Class> classes;
  // the single launch's items occupy queue[0, single_nwork)
  int32_t single_nwork = 0;
  int single_grid = 0;
  std::vector<msig::KArgs> jobs;
  std::vector<msig::WorkItem> queue;
  std::vector<msig::FinArgs> fins;
  std::vector<msig::WorkItem> finq;
  int fin_grid = 1;
  size_t ws_bytes = 0;
  // byte offsets of the arrays and the single launch's 64-bit work counter
  // inside the descriptor block
  size_t off_jobs = 0, off_queue = 0, off_fins = 0, off_finq = 0, off_ticket = 0, desc_bytes = 0;
};

using namespace msig;

extern "C" {

int msig_abi_version(void) { return MSIG_ABI_VERSION; }

const char* msig_last_error(void) { return g_err.c_str(); }

int msig_device_count(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice) return 0;
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  return n;
}

int msig_init(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  for (int N = MSIG_MIN_NFFT; N <= MSIG_MAX_NFFT; N *= 2) {
    const cf* t = nullptr;
    int rc = twiddles(N, &t);
    if (rc) return rc;
    // every mean kernel's resident-block count, so later plans only look up
    for (int kk = K_PSD; kk <= K_COH; ++kk) {
      (void)resident_blocks(N, kk, true, false);
      (void)resident_blocks(N, kk, false, false);
      (void)resident_blocks(N, kk, false, true);
    }
  }
  return scratch_setup(device);
}

int msig_eligible(int32_t nperseg, int32_t nfft, int32_t onesided, int32_t detrend_const,
                  int32_t complex_input) {
  (void)onesided;
  (void)detrend_const;
  (void)complex_input;
  if (nperseg < 1 || nfft < nperseg) return 0;
  if (!pow2(nfft) || nfft < MSIG_MIN_NFFT || nfft > MSIG_MAX_NFFT) return 0;
  return 1;
}

int msig_workspace_bytes(const msig_job* job, size_t* bytes) {
  int rc = check_job(job);
  if (rc) return rc;
  if (!bytes) return fail(MSIG_EINVAL, "bytes is NULL");
  *bytes = is_mean(job->kind) ? mean_ws_bytes(job) : 0;
  return 0;
}

int msig_mean_plan(const msig_job* job, int32_t* plan4) {
  int rc = check_job(job);
  if (rc) return rc;
  if (!plan4) return fail(MSIG_EINVAL, "plan4 is NULL");
  if (!is_mean(job->kind)) return fail(MSIG_EINVAL, "kind %d is not a mean kind", job->kind);
  const Geo g = mean_geometry(job);
  plan4[0] = g.units;
  plan4[1] = g.gpb;
  plan4[2] = g.run;
  plan4[3] = g.bpr;
  return 0;
}

// Checks shared by the mean entry points.
static int check_mean_call(const msig_job* job, const void* x, const void* y, const float* win,
                           void* out0, void* out1, void* out2) {
  int rc = check_job(job);
  if (rc) return rc;
  if (!is_mean(job->kind)) return fail(MSIG_EINVAL, "kind %d is not a mean kind", job->kind);
  const bool pair = is_pair(job->kind);
  if (!x || !win || !out0 || (pair && !y)) return fail(MSIG_EINVAL, "NULL input/output pointer");
  if (job->kind == MSIG_COH_TRIO_MEAN && (!out1 || !out2))
    return fail(MSIG_EINVAL, "coherence trio needs three outputs");
  return 0;
}

// The two kernels of a checked mean job, block partials in `ws` (which holds
// at least mean_ws_bytes(job)); `tw` is the size's resident twiddle table.
static int mean_run(const msig_job* job, const void* x, const void* y, const float* win,
                    void* out0, void* out1, void* out2, float* ws, const cf* tw,
                    hipStream_t s) {
  const bool cplx = job->flags & MSIG_F_COMPLEX_INPUT;
  const int N = job->nfft;
  const int prec = precision_of(job);
  const Geo g = mean_geometry(job);
  KArgs a = base_args(job, x, y, win, tw, g);
  a.ws = ws;
  const int kk = mean_kk(job->kind);
  hipError_t e = timed_launch(s, [&] {
    return prec != PREC_FP32 ? launch_mean_mfma256(prec, a, job->rows, s)
                             : dispatch_mean(N, kk, cplx, a, job->rows, s);
  });
  if (e != hipSuccess) return hip_fail(e, "fused mean kernel launch");
  FinArgs f;
  std::memset(&f, 0, sizeof f);
  f.ws = a.ws;
  f.out0 = out0;
  f.out1 = out1;
  f.out2 = out2;
  f.rows = job->rows;
  f.bpr = g.bpr;
  f.N = N;
  f.nbins = nbins_of(job);
  f.nc = ncomp(job->kind);
  f.kind = job->kind;
  f.fold = cplx ? 0 : 1;
  f.apply_factor =
      (!cplx && (job->flags & MSIG_F_APPLY_FACTOR) && (job->flags & MSIG_F_ONESIDED)) ? 1 : 0;
  f.scale = job->scale / (double)job->nseg;
  f.side_src = job->side_src;
  f.side_dst = job->side_dst;
  f.side_len = job->side_len;
  const int64_t chunks = (f.nbins + FIN_BINS - 1) / FIN_BINS;
  const dim3 grid((unsigned)(job->rows * chunks)), block(FIN_BINS * FIN_LANES);
  hipLaunchKernelGGL(k_finalize, grid, block, 0, s, f);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "finalize kernel launch");
  return 0;
}

int msig_spectral_mean(const msig_job* job, const void* x, const void* y, const float* win,
                       void* out0, void* out1, void* out2, void* workspace, size_t ws_bytes,
                       void* stream) {
  int rc = check_mean_call(job, x, y, win, out0, out1, out2);
  if (rc) return rc;
  const size_t need = mean_ws_bytes(job);
  if (!workspace || ws_bytes < need)
    return fail(MSIG_EINVAL, "workspace of %zu bytes needed (got %zu)", need, ws_bytes);
  const cf* tw = nullptr;
  rc = twiddles(job->nfft, &tw);
  if (rc) return rc;
  return mean_run(job, x, y, win, out0, out1, out2, static_cast<float*>(workspace), tw,
                  static_cast<hipStream_t>(stream));
}

int msig_spectral_mean_noalloc(const msig_job* job, const void* x, const void* y,
                               const float* win, void* out0, void* out1, void* out2,
                               void* stream) {
  int rc = check_mean_call(job, x, y, win, out0, out1, out2);
  if (rc) return rc;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  float* ws = scratch_of(dev);
  if (!ws)
    return fail(MSIG_EINVAL,
                "msig_init(%d) must run once on this device before msig_spectral_mean_noalloc "
                "(it resolves the device's scratch array and builds the resident tables)",
                dev);
  const size_t need = mean_ws_bytes(job);
  if (need > MSIG_NOALLOC_SCRATCH_BYTES)
    return fail(MSIG_EINVAL,
                "block partials of %zu bytes exceed the %zu-byte scratch of "
                "msig_spectral_mean_noalloc (use msig_spectral_mean with a workspace)",
                need, (size_t)MSIG_NOALLOC_SCRATCH_BYTES);
  const cf* tw = resident_twiddles(dev, job->nfft);
  if (!tw) return fail(MSIG_EINVAL, "no resident twiddle table for nfft %d on device %d", job->nfft, dev);
  return mean_run(job, x, y, win, out0, out1, out2, ws, tw, static_cast<hipStream_t>(stream));
}

int msig_spectral_frames(const msig_job* job, const void* x, const void* y, const float* win,
                         void* out0, void* out1, void* out2, void* stream) {
  int rc = check_job(job);
  if (rc) return rc;
  if (is_mean(job->kind)) return fail(MSIG_EINVAL, "kind %d is not a per-frame kind", job->kind);
  const bool pair = is_pair(job->kind);
  if (!x || !win || !out0 || (pair && !y)) return fail(MSIG_EINVAL, "NULL input/output pointer");
  if (job->kind == MSIG_COH_FRAMES && (!out1 || !out2))
    return fail(MSIG_EINVAL, "coherence frames need three outputs");
  const bool cplx = job->flags & MSIG_F_COMPLEX_INPUT;
  const int N = job->nfft;
  const Geo g = geometry(job->rows, job->nseg, N, cplx, FRAMES_MAX_RUN, FRAMES_TARGET_BLOCKS);
  const cf* tw = nullptr;
  rc = twiddles(N, &tw);
  if (rc) return rc;
  KArgs a = base_args(job, x, y, win, tw, g);
  a.out0 = out0;
  a.out1 = out1;
  a.out2 = out2;
  a.nbins = nbins_of(job);
  a.apply_factor =
      (!cplx && (job->flags & MSIG_F_APPLY_FACTOR) && (job->flags & MSIG_F_ONESIDED)) ? 1 : 0;
  int ok = F_COMPLEX;
  switch (job->kind) {
    case MSIG_FRAMES_COMPLEX: ok = F_COMPLEX; break;
    case MSIG_FRAMES_POWER: ok = F_POWER; break;
    case MSIG_CSD_FRAMES: ok = F_CSD; break;
    default: ok = F_COH; break;
  }
  a.oscale = ok == F_COMPLEX ? (float)std::sqrt(job->scale) : (float)job->scale;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool ext = job->lead != 0 || job->nvalid != 0;
  hipError_t e =
      timed_launch(s, [&] { return dispatch_frames(N, ok, cplx, ext, a, job->rows, s); });
  if (e != hipSuccess) return hip_fail(e, "fused frames kernel launch");
  return 0;
}

// ---- operator-level entry points (reference _stft_metal.py functions) ----

static msig_job op_job(int kind, int64_t rows, int64_t n, int32_t nperseg, int32_t hop,
                       int32_t detrend) {
  msig_job j;
  std::memset(&j, 0, sizeof j);
  j.kind = kind;
  j.flags = MSIG_F_ONESIDED | (detrend ? MSIG_F_DETREND_CONST : 0);
  j.rows = rows;
  j.n = n;
  j.ld_x = n;
  j.ld_y = n;
  j.nperseg = nperseg;
  j.nstep = hop;
  j.nfft = nperseg;
  j.nseg = (hop >= 1 && n >= nperseg) ? (int32_t)((n - nperseg) / hop + 1) : 0;
  j.scale = 1.0;
  return j;
}

int msig_rfft_frames(const float* x, int64_t rows, int64_t n, const float* win, int32_t nperseg,
                     int32_t hop, int32_t detrend, int32_t power, void* out, void* stream) {
  msig_job j = op_job(power ? MSIG_FRAMES_POWER : MSIG_FRAMES_COMPLEX, rows, n, nperseg, hop, detrend);
  return msig_spectral_frames(&j, x, nullptr, win, out, nullptr, nullptr, stream);
}

int msig_rfft_power_mean(const float* x, int64_t rows, int64_t n, const float* win,
                         int32_t nperseg, int32_t hop, int32_t detrend, float* out,
                         void* workspace, size_t ws_bytes, void* stream) {
  msig_job j = op_job(MSIG_PSD_MEAN, rows, n, nperseg, hop, detrend);
  return msig_spectral_mean(&j, x, nullptr, win, out, nullptr, nullptr, workspace, ws_bytes, stream);
}

int msig_csd_frames(const float* x, const float* y, int64_t rows, int64_t n, const float* win,
                    int32_t nperseg, int32_t hop, int32_t detrend, void* out, void* stream) {
  msig_job j = op_job(MSIG_CSD_FRAMES, rows, n, nperseg, hop, detrend);
  return msig_spectral_frames(&j, x, y, win, out, nullptr, nullptr, stream);
}

int msig_coherence_frames(const float* x, const float* y, int64_t rows, int64_t n,
                          const float* win, int32_t nperseg, int32_t hop, int32_t detrend,
                          float* pxx, float* pyy, void* pxy, void* stream) {
  msig_job j = op_job(MSIG_COH_FRAMES, rows, n, nperseg, hop, detrend);
  return msig_spectral_frames(&j, x, y, win, pxx, pyy, pxy, stream);
}

int msig_csd_mean(const float* x, const float* y, int64_t rows, int64_t n, const float* win,
                  int32_t nperseg, int32_t hop, int32_t detrend, void* out, void* workspace,
                  size_t ws_bytes, void* stream) {
  msig_job j = op_job(MSIG_CSD_MEAN, rows, n, nperseg, hop, detrend);
  return msig_spectral_mean(&j, x, y, win, out, nullptr, nullptr, workspace, ws_bytes, stream);
}

int msig_coherence_mean(const float* x, const float* y, int64_t rows, int64_t n,
                        const float* win, int32_t nperseg, int32_t hop, int32_t detrend,
                        float* pxx, float* pyy, void* pxy, void* workspace, size_t ws_bytes,
                        void* stream) {
  msig_job j = op_job(MSIG_COH_TRIO_MEAN, rows, n, nperseg, hop, detrend);
  return msig_spectral_mean(&j, x, y, win, pxx, pyy, pxy, workspace, ws_bytes, stream);
}

}  // extern "C"

static void free_plan(msig_batch* p) { delete p; }

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

extern "C" {

int msig_batch_create(const msig_batch_job* jobs, int32_t njobs, const msig_tuning* tuning,
                      msig_batch** plan) {
  if (!plan) return fail(MSIG_EINVAL, "plan is NULL");
  *plan = nullptr;
  if (njobs < 1 || !jobs) return fail(MSIG_EINVAL, "a batch needs at least one job");
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
  std::vector<KArgs> hk(njobs);
  std::vector<FinArgs> hf(njobs);
  std::map<std::tuple<int, int, bool, bool>, std::vector<WorkItem>> qs;
  std::vector<WorkItem> finq;
  // jobs of the single launch, with the estimated cost of one of their blocks
  std::vector<std::pair<int64_t, int32_t>> single;
  const bool per_class = tuning && tuning->per_class_launches;
  size_t off = 0;  // floats
  for (int32_t i = 0; i < njobs; ++i) {
    const msig_batch_job& bj = jobs[i];
    const msig_job* j = &bj.job;
    int rc = check_job(j);
    if (rc) return fail(rc, "job %d: %s", i, msig_last_error());
    if (!is_mean(j->kind)) return fail(MSIG_EINVAL, "job %d: kind %d is not a mean kind", i, j->kind);
    if (precision_of(j) != PREC_FP32)
      return fail(MSIG_ENOTSUP, "job %d: batch plans run the FP32 form only", i);
    if (j->side_len) return fail(MSIG_EINVAL, "job %d: batch plans take no side copy", i);
    const bool pair = is_pair(j->kind);
    if (!bj.x || !bj.win || !bj.out0 || (pair && !bj.y))
      return fail(MSIG_EINVAL, "job %d: NULL input/output pointer", i);
    if (j->kind == MSIG_COH_TRIO_MEAN && (!bj.out1 || !bj.out2))
      return fail(MSIG_EINVAL, "job %d: coherence trio needs three outputs", i);
    const bool cplx = j->flags & MSIG_F_COMPLEX_INPUT;
    const int N = j->nfft;
    const int l2 = log2i(N);
    int maxrun = MEAN_MAX_RUN;
    int64_t target = 0;  // 0: the standalone call's resident-block plan
    if (tuning) {
      if (tuning->max_run[l2] > 0) maxrun = tuning->max_run[l2] < MEAN_MAX_RUN ? tuning->max_run[l2] : MEAN_MAX_RUN;
      if (tuning->target_blocks[l2] > 0) target = tuning->target_blocks[l2];
    }
    const Geo g = mean_geometry_fp32(j, maxrun, target);
    const cf* tw = nullptr;
    rc = twiddles(N, &tw);
    if (rc) return rc;
    KArgs a = base_args(j, bj.x, bj.y, bj.win, tw, g);
    a.ws = nullptr;
    a.ws_off = (int64_t)off;
    hk[i] = a;
    FinArgs f;
    std::memset(&f, 0, sizeof f);
    f.out0 = bj.out0;
    f.out1 = bj.out1;
    f.out2 = bj.out2;
    f.rows = j->rows;
    f.bpr = g.bpr;
    f.N = N;
    f.nbins = nbins_of(j);
    f.nc = ncomp(j->kind);
    f.kind = j->kind;
    f.fold = cplx ? 0 : 1;
    f.apply_factor =
        (!cplx && (j->flags & MSIG_F_APPLY_FACTOR) && (j->flags & MSIG_F_ONESIDED)) ? 1 : 0;
    f.scale = j->scale / (double)j->nseg;
    f.ws_off = (int64_t)off;
    hf[i] = f;
    const int kk = j->kind == MSIG_PSD_MEAN ? K_PSD : (j->kind == MSIG_CSD_MEAN ? K_CSD : K_COH);
    const bool half = !cplx && j->nperseg == N && 2 * j->nstep == N;
    const int64_t nblk = j->rows * g.bpr;
    if (nblk > INT32_MAX) return fail(MSIG_EINVAL, "job %d: too many blocks", i);
    const int sched = (tuning && tuning->radix_schedule[l2] == 1) ? 1 : 0;
    if (!per_class && half && kk != K_COH && sched == 0) {
      hk[i].cls = ragged_class(N, kk);
      // segment pairs per block x N log2 N, doubled for two signals
      const int64_t cost = (int64_t)g.gpb * g.run * N * l2 * (kk == K_CSD ? 2 : 1);
      single.emplace_back(cost, i);
    } else {
      auto& q = qs[std::make_tuple(N, kk, cplx, half)];
      for (int64_t b = 0; b < nblk; ++b) q.push_back(WorkItem{i, (int32_t)b});
    }
    const int64_t chunks = (f.nbins + FIN_BINS - 1) / FIN_BINS;
    for (int64_t b = 0; b < j->rows * chunks; ++b) finq.push_back(WorkItem{i, (int32_t)b});
    off += ((size_t)j->rows * g.bpr * f.nc * N + 63) / 64 * 64;
  }
  msig_batch* p = new msig_batch();
  p->device = dev;
  // Single launch: blocks are taken in queue order.  Each class's blocks
  // are contiguous (a CU then runs one class's code and twiddle table at a
  // time), classes in decreasing order of their costliest block, and inside
  // a class the costliest blocks first, jobs of one cost in job order.
  {
    std::map<int, int64_t> top;  // class -> its costliest block
    for (const auto& cj : single) {
      int64_t& t = top[hk[cj.second].cls];
      if (cj.first > t) t = cj.first;
    }
    std::stable_sort(single.begin(), single.end(),
                     [&](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& b) {
                       const int ca = hk[a.second].cls, cb = hk[b.second].cls;
                       if (ca != cb) {
                         const int64_t ta = top[ca], tb = top[cb];
                         return ta != tb ? ta > tb : ca > cb;
                       }
                       return a.first > b.first;
                     });
  }
  for (const auto& cj : single) {
    const int64_t nblk = hk[cj.second].bpr * jobs[cj.second].job.rows;
    for (int64_t b = 0; b < nblk; ++b) p->queue.push_back(WorkItem{cj.second, (int32_t)b});
  }
  if (p->queue.size() > (size_t)INT32_MAX) {
    delete p;
    return fail(MSIG_EINVAL, "too many blocks in one batch");
  }
  p->single_nwork = (int32_t)p->queue.size();
  if (p->single_nwork) {
    const int64_t cap = (int64_t)ragged_blocks_per_cu() * cus;
    p->single_grid = (int)(p->single_nwork < cap ? p->single_nwork : cap);
  }
  p->jobs = std::move(hk);
  p->fins = std::move(hf);
  for (auto& kv : qs) {
    msig_batch::Class c;
    c.N = std::get<0>(kv.first);
    c.kk = std::get<1>(kv.first);
    c.cplx = std::get<2>(kv.first);
    c.half = std::get<3>(kv.first);
    c.qoff = (int32_t)p->queue.size();
    c.nwork = (int32_t)kv.second.size();
    int bpc = 4;
    if (tuning && tuning->blocks_per_cu[log2i(c.N)] > 0) bpc = tuning->blocks_per_cu[log2i(c.N)];
    c.sched = (tuning && tuning->radix_schedule[log2i(c.N)] == 1) ? 1 : 0;
    const int64_t cap = (int64_t)bpc * cus;
    c.grid = (int)(c.nwork < cap ? c.nwork : cap);
    p->queue.insert(p->queue.end(), kv.second.begin(), kv.second.end());
    p->classes.push_back(c);
  }
  p->finq = std::move(finq);
  const int64_t nfin = (int64_t)p->finq.size();
  const int64_t fcap = (int64_t)8 * cus;
  p->fin_grid = (int)(nfin < fcap ? nfin : fcap);
  p->ws_bytes = off * sizeof(float);
  p->off_jobs = 0;
  p->off_queue = align256(p->off_jobs + sizeof(KArgs) * p->jobs.size());
  p->off_fins = align256(p->off_queue + sizeof(WorkItem) * p->queue.size());
  p->off_finq = align256(p->off_fins + sizeof(FinArgs) * p->fins.size());
  p->off_ticket = align256(p->off_finq + sizeof(WorkItem) * p->finq.size());
  p->desc_bytes = align256(p->off_ticket + sizeof(unsigned long long));
  *plan = p;
  return 0;
}

int msig_batch_descriptor_bytes(const msig_batch* plan, size_t* bytes) {
  if (!plan || !bytes) return fail(MSIG_EINVAL, "NULL plan or bytes");
  *bytes = plan->desc_bytes;
  return 0;
}

int msig_batch_pack(const msig_batch* plan, void* host, size_t bytes, void* device) {
  if (!plan || !host || !device) return fail(MSIG_EINVAL, "NULL plan or buffer");
  if (bytes < plan->desc_bytes)
    return fail(MSIG_EINVAL, "descriptor block of %zu bytes needed (got %zu)", plan->desc_bytes, bytes);
  char* h = static_cast<char*>(host);
  std::memcpy(h + plan->off_jobs, plan->jobs.data(), sizeof(KArgs) * plan->jobs.size());
  std::memcpy(h + plan->off_queue, plan->queue.data(), sizeof(WorkItem) * plan->queue.size());
  std::memcpy(h + plan->off_fins, plan->fins.data(), sizeof(FinArgs) * plan->fins.size());
  std::memcpy(h + plan->off_finq, plan->finq.data(), sizeof(WorkItem) * plan->finq.size());
  const unsigned long long ticket0 = 0;  // the work counter starts at zero
  std::memcpy(h + plan->off_ticket, &ticket0, sizeof ticket0);
  (void)device;  // the block holds no absolute pointers into itself
  return 0;
}

int msig_batch_workspace_bytes(const msig_batch* plan, size_t* bytes) {
  if (!plan || !bytes) return fail(MSIG_EINVAL, "NULL plan or bytes");
  *bytes = plan->ws_bytes;
  return 0;
}

int msig_batch_launches(const msig_batch* plan) {
  if (!plan) return fail(MSIG_EINVAL, "NULL plan");
  return (int)plan->classes.size() + (plan->single_nwork ? 1 : 0) + 1;
}

int msig_batch_run(const msig_batch* plan, const void* device, void* workspace, size_t ws_bytes,
                   void* stream) {
  if (!plan || !device) return fail(MSIG_EINVAL, "NULL plan or descriptor block");
  if (!workspace || ws_bytes < plan->ws_bytes)
    return fail(MSIG_EINVAL, "workspace of %zu bytes needed (got %zu)", plan->ws_bytes, ws_bytes);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev != plan->device)
    return fail(MSIG_EINVAL, "plan was built on device %d, current device is %d", plan->device, dev);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const char* d = static_cast<const char*>(device);
  const KArgs* jobs = reinterpret_cast<const KArgs*>(d + plan->off_jobs);
  const WorkItem* queue = reinterpret_cast<const WorkItem*>(d + plan->off_queue);
  const FinArgs* fins = reinterpret_cast<const FinArgs*>(d + plan->off_fins);
  const WorkItem* finq = reinterpret_cast<const WorkItem*>(d + plan->off_finq);
  float* ws = static_cast<float*>(workspace);
  if (plan->single_nwork) {
    unsigned long long* ticket = reinterpret_cast<unsigned long long*>(
        const_cast<char*>(d) + plan->off_ticket);
    e = launch_mean_ragged(jobs, queue, plan->single_nwork, ticket, ws, plan->single_grid, s);
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  }
  for (const auto& c : plan->classes) {
    e = dispatch_mean_batch(c.N, c.kk, c.cplx, c.half, c.sched, jobs, queue + c.qoff, c.nwork, ws,
                            c.grid, s);
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  }
  hipLaunchKernelGGL(k_finalize_batch, dim3(plan->fin_grid), dim3(FIN_BINS * FIN_LANES), 0, s, fins,
                     finq, (int32_t)plan->finq.size(), ws);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "batch finalize launch");
  return 0;
}

int msig_batch_destroy(msig_batch* plan) {
  free_plan(plan);
  return 0;
}

int msig_batch_autotune(const msig_batch_job* probe, int32_t njobs, const msig_tuning* base,
                        const msig_tile* cands, int32_t ncand, int32_t reps, float* ms_out,
                        int32_t* best, void* stream) {
  if (!probe || njobs < 1) return fail(MSIG_EINVAL, "autotuning needs at least one probe job");
  if (!cands || ncand < 1) return fail(MSIG_EINVAL, "autotuning needs at least one candidate");
  if (reps < 1) return fail(MSIG_EINVAL, "reps must be >= 1 (got %d)", reps);
  if (!ms_out || !best) return fail(MSIG_EINVAL, "NULL ms_out or best");
  const int N = probe[0].job.nfft;
  for (int32_t i = 0; i < njobs; ++i)
    if (probe[i].job.nfft != N)
      return fail(MSIG_EINVAL, "probe job %d has nfft %d; all probe jobs must share nfft %d", i,
                  probe[i].job.nfft, N);
  if (!pow2(N) || N < MSIG_MIN_NFFT || N > MSIG_MAX_NFFT)
    return fail(MSIG_ENOTSUP, "nfft %d outside the fused power-of-two range", N);
  const int l2 = log2i(N);
  for (int32_t c = 0; c < ncand; ++c) {
    const msig_tile& k = cands[c];
    if (k.max_run < 0 || k.target_blocks < 0 || k.blocks_per_cu < 0 || k.radix_schedule < 0 ||
        k.radix_schedule > 1)
      return fail(MSIG_EINVAL, "candidate %d: negative knob or radix_schedule not 0/1", c);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Owned resources, released on every exit.
  struct Owned {
    msig_batch* plan = nullptr;
    void* desc = nullptr;
    void* ws = nullptr;
    size_t desc_cap = 0, ws_cap = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Owned() {
      free_plan(plan);
      if (desc) (void)hipFree(desc);
      if (ws) (void)hipFree(ws);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } o;
  hipError_t e = hipEventCreate(&o.e0);
  if (e == hipSuccess) e = hipEventCreate(&o.e1);
  if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
  std::vector<char> host;
  // plan candidate c into o.plan with its descriptor block on the device
  auto prepare = [&](int32_t c) -> int {
    free_plan(o.plan);
    o.plan = nullptr;
    msig_tuning t;
    if (base) t = *base;
    else std::memset(&t, 0, sizeof t);
    t.max_run[l2] = cands[c].max_run;
    t.target_blocks[l2] = cands[c].target_blocks;
    t.blocks_per_cu[l2] = cands[c].blocks_per_cu;
    t.radix_schedule[l2] = cands[c].radix_schedule;
    int rc = msig_batch_create(probe, njobs, &t, &o.plan);
    if (rc) return fail(rc, "candidate %d: %s", c, msig_last_error());
    hipError_t err;
    if (o.plan->desc_bytes > o.desc_cap) {
      if (o.desc) (void)hipFree(o.desc);
      o.desc = nullptr;
      err = hipMalloc(&o.desc, o.plan->desc_bytes);
      if (err != hipSuccess) return hip_fail(err, "hipMalloc(descriptor block)");
      o.desc_cap = o.plan->desc_bytes;
    }
    const size_t wsb = o.plan->ws_bytes > 0 ? o.plan->ws_bytes : 256;
    if (wsb > o.ws_cap) {
      if (o.ws) (void)hipFree(o.ws);
      o.ws = nullptr;
      err = hipMalloc(&o.ws, wsb);
      if (err != hipSuccess) return hip_fail(err, "hipMalloc(workspace)");
      o.ws_cap = wsb;
    }
    host.assign(o.plan->desc_bytes, 0);
    rc = msig_batch_pack(o.plan, host.data(), host.size(), o.desc);
    if (rc) return rc;
    // the previous candidate's runs may still read the block: copy in stream order
    err = hipMemcpyAsync(o.desc, host.data(), host.size(), hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);  // pageable source: keep it alive
    if (err != hipSuccess) return hip_fail(err, "descriptor copy");
    return 0;
  };
  auto run = [&]() -> int { return msig_batch_run(o.plan, o.desc, o.ws, o.ws_cap, stream); };
  auto elapsed = [&](float* ms) -> int {
    hipError_t err = hipEventSynchronize(o.e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms, o.e0, o.e1);
    return err == hipSuccess ? 0 : hip_fail(err, "hipEventElapsedTime");
  };
  int rc = prepare(0);
  if (rc) return rc;
  // bring the clocks to their sustained level: ~60 ms of back-to-back runs
  float warm = 0.f;
  for (int it = 0; it < 64 && warm < 60.f; ++it) {
    e = hipEventRecord(o.e0, s);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    for (int k = 0; k < 8; ++k)
      if ((rc = run())) return rc;
    e = hipEventRecord(o.e1, s);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    float ms = 0.f;
    if ((rc = elapsed(&ms))) return rc;
    warm += ms;
  }
  int32_t arg = 0;
  for (int32_t c = 0; c < ncand; ++c) {
    if (c > 0 && (rc = prepare(c))) return rc;
    if ((rc = run())) return rc;  // untimed
    std::vector<float> ts((size_t)reps);
    for (int32_t r = 0; r < reps; ++r) {
      e = hipEventRecord(o.e0, s);
      if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
      if ((rc = run())) return rc;
      e = hipEventRecord(o.e1, s);
      if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
      if ((rc = elapsed(&ts[(size_t)r]))) return rc;
    }
    std::sort(ts.begin(), ts.end());
    ms_out[c] = ts[(size_t)reps / 2];
    if (ms_out[c] < ms_out[arg]) arg = c;
  }
  *best = arg;
  return 0;
}

int msig_mean_workspace_bytes(int32_t kind, int64_t rows, int64_t n, int32_t nperseg, int32_t hop,
                              size_t* bytes) {
  if (kind != MSIG_PSD_MEAN && kind != MSIG_CSD_MEAN && kind != MSIG_COH_TRIO_MEAN)
    return fail(MSIG_EINVAL, "kind %d has no mean workspace", kind);
  msig_job j = op_job(kind, rows, n, nperseg, hop, 0);
  return msig_workspace_bytes(&j, bytes);
}

int msig_istft_frames(const void* z, int64_t rows, int32_t nseg, int32_t nperseg,
                      const float* weff, float* frames, void* stream) {
  if (!z || !weff || !frames) return fail(MSIG_EINVAL, "NULL pointer");
  if (rows < 1 || nseg < 1) return fail(MSIG_EINVAL, "rows and nseg must be >= 1");
  if (!pow2(nperseg) || nperseg < MSIG_MIN_NFFT || nperseg > MSIG_MAX_NFFT)
    return fail(MSIG_ENOTSUP, "nperseg %d outside the fused power-of-two range [%d, %d]",
                nperseg, MSIG_MIN_NFFT, MSIG_MAX_NFFT);
  const int N = nperseg;
  const cf* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  const Geo g = geometry(rows, nseg, N, false, FRAMES_MAX_RUN);
  KArgs a;
  std::memset(&a, 0, sizeof a);
  a.x = z;
  a.win = weff;
  a.twt = tw;
  a.out0 = frames;
  a.nseg = nseg;
  a.nperseg = N;
  a.hop = N;
  a.units = g.units;
  a.run = g.run;
  a.bpr = g.bpr;
  a.nbins = N / 2 + 1;
  hipError_t e = dispatch_istft(N, a, rows, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "inverse frames kernel launch");
  return 0;
}

int msig_istft_ola(const void* z, int64_t rows, int32_t nseg, int32_t nperseg,
                   const float* weff, const float* win, int64_t first, int64_t count, float* out,
                   const float* side_src, float* side_dst, int64_t side_len, void* stream) {
  if (!z || !weff || !win || !out) return fail(MSIG_EINVAL, "NULL pointer");
  if (side_len < 0 || (side_len > 0 && (!side_src || !side_dst)))
    return fail(MSIG_EINVAL, "side copy: side_len must be >= 0 with both buffers given");
  if (rows < 1 || nseg < 1) return fail(MSIG_EINVAL, "rows and nseg must be >= 1");
  if (!pow2(nperseg) || nperseg < MSIG_MIN_NFFT || nperseg > MSIG_MAX_NFFT)
    return fail(MSIG_ENOTSUP, "nperseg %d outside the fused power-of-two range [%d, %d]",
                nperseg, MSIG_MIN_NFFT, MSIG_MAX_NFFT);
  const int64_t outlen = (int64_t)(nseg + 1) * (nperseg / 2);
  if (first < 0 || count < 0 || first + count > outlen)
    return fail(MSIG_EINVAL, "samples [%lld, %lld) outside the %lld-sample output",
                (long long)first, (long long)(first + count), (long long)outlen);
  if (count == 0) return 0;
  const int N = nperseg;
  const cf* tw = nullptr;
  int rc = twiddles(N, &tw);
  if (rc) return rc;
  const Geo g = geometry(rows, nseg, N, false, ISTFT_OLA_MAX_RUN, ISTFT_OLA_TARGET_BLOCKS);
  if (rows * (int64_t)g.bpr > 0x7fffffff) return fail(MSIG_EINVAL, "too many rows");
  KArgs a;
  std::memset(&a, 0, sizeof a);
  a.x = z;
  a.y = win;
  a.win = weff;
  a.twt = tw;
  a.out0 = out;
  a.nseg = nseg;
  a.nperseg = N;
  a.hop = N / 2;
  a.units = g.units;
  a.run = g.run;
  a.bpr = g.bpr;
  a.nbins = N / 2 + 1;
  a.ola_first = first;
  a.ola_count = count;
  a.side_src = side_src;
  a.side_dst = side_dst;
  a.side_len = side_len;
  hipError_t e = dispatch_istft_ola(N, a, rows, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "fused istft kernel launch");
  return 0;
}

int msig_overlap_add(const float* frames, int64_t rows, int32_t nseg, int32_t nperseg,
                     int32_t hop, int32_t channels, const float* win, int64_t first,
                     int64_t count, float* out, void* stream) {
  if (!frames || !win || !out) return fail(MSIG_EINVAL, "NULL pointer");
  if (rows < 1 || nseg < 1 || nperseg < 1 || hop < 1)
    return fail(MSIG_EINVAL, "rows, nseg, nperseg and hop must be >= 1");
  if (channels != 1 && channels != 2) return fail(MSIG_EINVAL, "channels must be 1 or 2");
  const int64_t outlen = nperseg + (int64_t)(nseg - 1) * hop;
  if (first < 0 || count < 0 || first + count > outlen)
    return fail(MSIG_EINVAL, "samples [%lld, %lld) outside the %lld-sample output",
                (long long)first, (long long)(first + count), (long long)outlen);
  if (count == 0) return 0;
  OlaArgs o{frames, win, out, first, count, rows, nseg, nperseg, hop, channels};
  const int64_t total = rows * count * channels;
  int64_t blocks = (total + 255) / 256;
  if (blocks > ((int64_t)1 << 20)) blocks = (int64_t)1 << 20;  // the rest grid-strided
  const dim3 grid((unsigned)blocks), block(256);
  hipLaunchKernelGGL(k_overlap_add, grid, block, 0, static_cast<hipStream_t>(stream), o);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "overlap-add kernel launch");
  return 0;
}

int msig_median_frames(const float* frames, int64_t rows, int32_t nseg, int32_t cols,
                       int32_t channels, double scale, int32_t apply_factor, int32_t nfft,
                       float* out, void* stream) {
  if (!frames || !out) return fail(MSIG_EINVAL, "NULL pointer");
  if (rows < 1 || nseg < 1 || cols < 1) return fail(MSIG_EINVAL, "rows, nseg and cols must be >= 1");
  if (channels != 1 && channels != 2) return fail(MSIG_EINVAL, "channels must be 1 or 2");
  if (nseg > MSIG_MEDIAN_MAX_NSEG)
    return fail(MSIG_ENOTSUP, "nseg %d above the staged median's %d", nseg, MSIG_MEDIAN_MAX_NSEG);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const MedRegShape rs = medr_shape(nseg, rows * (int64_t)cols);
  if (rs.cpb) {  // keys in registers
    const int ncb = (cols + rs.cpb - 1) / rs.cpb;
    if (rows * (int64_t)ncb > INT32_MAX) return fail(MSIG_EINVAL, "too many columns");
    MedArgs m{frames, out, rows, nseg, cols, ncb, channels, apply_factor ? 1 : 0, nfft, 0,
              (float)scale};
    const dim3 grid((unsigned)(rows * ncb)), block(rs.bt);
#define MEDR(CP, KP, BB)                                                          \
  else if (rs.cpb == CP && rs.kpt == KP && rs.bt == BB)                           \
    hipLaunchKernelGGL((k_median_reg<CP, KP, BB>), grid, block, 0, s, m);
    if (false) {
    }
    MEDR(16, 16, 512) MEDR(16, 32, 512) MEDR(8, 32, 512) MEDR(4, 32, 512) MEDR(4, 64, 512)
    MEDR(8, 16, 256) MEDR(8, 32, 256) MEDR(4, 32, 256)
    else return fail(MSIG_EINVAL, "median shape (%d, %d, %d) not built", rs.cpb, rs.kpt, rs.bt);
#undef MEDR
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "median kernel launch");
    return 0;
  }
  int cpb = MED_MAX_CPB;  // columns per block: the most that fit the LDS budget
  while (cpb > 1 && med_lds_bytes(cpb, med_stride(nseg, cpb)) > (size_t)MED_LDS_BUDGET) cpb /= 2;
  const int stride = med_stride(nseg, cpb);
  const size_t lds = med_lds_bytes(cpb, stride);
  const int ncb = (cols + cpb - 1) / cpb;
  if (rows * (int64_t)ncb > INT32_MAX) return fail(MSIG_EINVAL, "too many columns");
  MedArgs m{frames, out, rows, nseg, cols, ncb, channels, apply_factor ? 1 : 0, nfft, stride,
            (float)scale};
  const dim3 grid((unsigned)(rows * ncb)), block(MED_BT);
  switch (cpb) {
    case 16: hipLaunchKernelGGL(k_median<16>, grid, block, lds, s, m); break;
    case 8: hipLaunchKernelGGL(k_median<8>, grid, block, lds, s, m); break;
    case 4: hipLaunchKernelGGL(k_median<4>, grid, block, lds, s, m); break;
    case 2: hipLaunchKernelGGL(k_median<2>, grid, block, lds, s, m); break;
    default: hipLaunchKernelGGL(k_median<1>, grid, block, lds, s, m); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "median kernel launch");
  return 0;
}

int msig_kernel_timing_arm(int32_t max_launches) {
  if (max_launches < 0 || max_launches > (1 << 20))
    return fail(MSIG_EINVAL, "max_launches %d outside [0, 2^20]", max_launches);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_kt_mu);
  timing_release();
  for (int32_t i = 0; i < max_launches; ++i) {
    hipEvent_t a = nullptr, b = nullptr;
    e = hipEventCreate(&a);
    if (e == hipSuccess) {
      e = hipEventCreate(&b);
      if (e != hipSuccess) (void)hipEventDestroy(a);
    }
    if (e != hipSuccess) {
      timing_release();
      return hip_fail(e, "hipEventCreate");
    }
    g_kt.ev.emplace_back(a, b);
  }
  g_kt.device = dev;
  g_kt_armed.store(true, std::memory_order_release);
  return 0;
}

int msig_kernel_timing_read(double* total_ms, int32_t* launches) {
  if (!total_ms || !launches) return fail(MSIG_EINVAL, "NULL output pointer");
  std::lock_guard<std::mutex> lk(g_kt_mu);
  double tot = 0.0;
  for (size_t i = 0; i < g_kt.used; ++i) {
    hipError_t e = hipEventSynchronize(g_kt.ev[i].second);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, g_kt.ev[i].first, g_kt.ev[i].second);
    if (e != hipSuccess) {
      timing_release();
      return hip_fail(e, "kernel timing events");
    }
    tot += ms;
  }
  *total_ms = tot;
  *launches = (int32_t)g_kt.used;
  timing_release();
  return 0;
}

}  // extern "C"
