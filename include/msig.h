/*
 * msig.h - C ABI of the MI355X fused STFT/Welch library (libmsig.so).
 *
 * Every entry point takes plain device pointers, element counts and a
 * hipStream_t passed as void*; nothing throws across the ABI and no torch
 * type appears.  Calls are asynchronous on the given stream.  The caller owns
 * x, y, the window, all outputs and the workspace; the library owns only its
 * device-resident twiddle tables (one per device and FFT size, built on first
 * use and kept until the process ends).
 *
 * Return value: 0 on success, or a negative errno-style status:
 *   MSIG_EINVAL  (-22) bad descriptor / pointer / size
 *   MSIG_ENOTSUP (-95) the fused path does not cover this shape (the Python
 *                      planner then takes its composed path, as the
 *                      reference does when `eligible` is False)
 *   MSIG_EIO     (-5)  a HIP runtime error; text via msig_last_error()
 *
 * Layouts (all row-major, rows contiguous along the analysed axis):
 *   real input x, y      float32 [rows][ld]   (the first n samples are used)
 *   complex input        float32 pairs [rows][ld][2]
 *   mean outputs         [rows][bins]   bins = nfft/2+1 (onesided) or nfft
 *   per-frame outputs    [rows][nseg][bins]
 *   complex outputs      interleaved (re, im) float32 pairs (= complex64)
 *
 * Mean kinds run two kernels on the stream: the fused segment kernel writes
 * per-block partial sums into the caller's workspace and a finalize kernel
 * reduces them.  The workspace needs no initialisation; calls that overlap
 * in time must use different workspaces.
 */
#ifndef MSIG_H
#define MSIG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSIG_ABI_VERSION 8

#define MSIG_EINVAL (-22)
#define MSIG_ENOTSUP (-95)
#define MSIG_EIO (-5)

/* Smallest and largest FFT sizes (powers of two) of the fused kernels. */
#define MSIG_MIN_NFFT 16
#define MSIG_MAX_NFFT 4096

/* job kinds */
enum {
  MSIG_PSD_MEAN = 0,      /* mean |X|^2 over segments            -> out0 f32     */
  MSIG_CSD_MEAN = 1,      /* mean conj(X) Y                      -> out0 c64     */
  MSIG_COH_TRIO_MEAN = 2, /* mean |X|^2, |Y|^2, conj(X) Y        -> out0,1 f32, out2 c64 */
  MSIG_COHERENCE = 3,     /* |Pxy|^2 / (Pxx Pyy)                 -> out0 f32     */
  MSIG_FRAMES_COMPLEX = 4,/* per-frame X                         -> out0 c64     */
  MSIG_FRAMES_POWER = 5,  /* per-frame |X|^2                     -> out0 f32     */
  MSIG_CSD_FRAMES = 6,    /* per-frame conj(X) Y                 -> out0 c64     */
  MSIG_COH_FRAMES = 7     /* per-frame |X|^2, |Y|^2, conj(X) Y   -> out0,1 f32, out2 c64 */
};

/* job flags */
#define MSIG_F_COMPLEX_INPUT 0x1 /* x (and y) are complex64; output is two-sided  */
#define MSIG_F_ONESIDED 0x2      /* real input: emit bins 0..nfft/2 only           */
#define MSIG_F_APPLY_FACTOR 0x4  /* multiply onesided bins 1..nfft/2-1 by 2        */
#define MSIG_F_DETREND_CONST 0x8 /* subtract each segment's mean before windowing  */
/* Precision form (MSIG_PSD_MEAN, real input, nperseg = nfft = 256 only, else
 * MSIG_ENOTSUP): the two 16-point DFT stages run on the matrix cores with
 * hi + lo split FP16 or BF16 operands and FP32 accumulation; the twiddle
 * stage, detrend, window and the FP64 segment sum are as in the FP32 path.
 * No counterpart in the reference (precision study). */
#define MSIG_F_MFMA_F16 0x10
#define MSIG_F_MFMA_BF16 0x20

typedef struct msig_job {
  int32_t kind;    /* MSIG_* kind above                                         */
  int32_t flags;   /* MSIG_F_* bits                                             */
  int64_t rows;    /* independent rows (channels)                               */
  int64_t n;       /* samples per row the segments are cut from                 */
  int64_t ld_x;    /* row stride of x in elements (>= n)                        */
  int64_t ld_y;    /* row stride of y in elements (>= n; ignored without y)     */
  int32_t nperseg; /* segment length (window length), 1 <= nperseg <= nfft      */
  int32_t nstep;   /* hop between segment starts = nperseg - noverlap, >= 1     */
  int32_t nfft;    /* FFT length, power of two in [MSIG_MIN_NFFT, MSIG_MAX_NFFT] */
  int32_t nseg;    /* (n - nperseg) / nstep + 1, checked                        */
  double scale;    /* multiplies every mean/power output bin; complex per-frame
                      outputs are multiplied by sqrt(scale)                     */
  int64_t lead;    /* per-frame kinds only (0 for means and batches): zero
                      samples virtually in front of each row, so that `n` is
                      the length of the extended row the segments are cut
                      from (stft's boundary='zeros' extension plus padding)    */
  int64_t nvalid;  /* samples stored per row (0: n - lead).  Extended sample p
                      is stored sample p - lead, and reads as zero outside
                      [0, nvalid); ld_x / ld_y >= nvalid.                       */
  const float* side_src; /* optional side copy (side_len 0: none): the call's   */
  float* side_dst;       /* kernels also copy side_len floats from side_src to */
  int64_t side_len;      /* side_dst (device buffers, not overlapping), so a
                            caller returning cached constants alongside the
                            result (scipy's frequency and time grids,
                            spectral.py:301, 355-365) needs no copy launch of
                            its own.  Single calls only (batch plans reject it). */
} msig_job;

/* ABI version the library was built with (MSIG_ABI_VERSION). */
int msig_abi_version(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char* msig_last_error(void);

/* Number of visible HIP devices (0 when no GPU / no driver), or < 0 on error. */
int msig_device_count(void);

/* Make `device` current for the calling thread and build its resident
 * tables for every fused FFT size.  Optional: the entry points build what
 * they need on first use. */
int msig_init(int device);

/* 1 if the fused kernels cover this shape, else 0.
 * Counterpart of `eligible` (reference: src/mlx_signal_processing/_stft_metal.py:150-160),
 * widened to nfft >= nperseg (zero padding), two-sided and complex input. */
int msig_eligible(int32_t nperseg, int32_t nfft, int32_t onesided, int32_t detrend_const,
                  int32_t complex_input);

/* Workspace bytes the mean kinds need (0 for per-frame kinds). */
int msig_workspace_bytes(const msig_job* job, size_t* bytes);

/* Launch plan of a mean job on the current device: plan4 = {units per row
 * (segment pairs for real input), FFT groups per block, units per group,
 * blocks per row}.  Diagnostic, no counterpart in the reference (whose
 * group parameters are internal, _stft_metal.py:429-441): tests use it to
 * pick shapes whose groups walk several units and whose last block is
 * uneven. */
int msig_mean_plan(const msig_job* job, int32_t* plan4);

/* Generic entry points (all kinds).  y may be NULL for single-signal kinds;
 * win is float32[nperseg] on the device.  The mean kinds write
 * scale * mean_over_segments(...) (times the onesided factor when
 * MSIG_F_APPLY_FACTOR is set); MSIG_COHERENCE ignores scale. */
int msig_spectral_mean(const msig_job* job, const void* x, const void* y, const float* win,
                       void* out0, void* out1, void* out2, void* workspace, size_t ws_bytes,
                       void* stream);
int msig_spectral_frames(const msig_job* job, const void* x, const void* y, const float* win,
                         void* out0, void* out1, void* out2, void* stream);

/* Largest block-partial workspace (bytes) msig_spectral_mean_noalloc takes. */
#define MSIG_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)

/* msig_spectral_mean without a workspace argument and without any device
 * allocation: the block partials go to a statically allocated 16 MiB array
 * in the current device's memory (held until the process ends), so that
 * callers which cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation) can run the fused means.
 * No counterpart in the reference.  Set-up: msig_init(device) must have run
 * once on the device (it resolves the array's device address and builds the
 * resident tables); otherwise MSIG_EINVAL.  Jobs whose partials
 * (msig_workspace_bytes) exceed MSIG_NOALLOC_SCRATCH_BYTES get MSIG_EINVAL.
 * Every call on one device shares the array, so calls that overlap in time
 * on one device (several streams, several threads) race; calls in one
 * stream's order are safe.  Same kernels and launch plan as
 * msig_spectral_mean, so the outputs are bit-identical. */
int msig_spectral_mean_noalloc(const msig_job* job, const void* x, const void* y,
                               const float* win, void* out0, void* out1, void* out2,
                               void* stream);

/* ---- operator-level entry points: one per function of the reference's fused
 * module.  Real float32 input, rows contiguous (ld = n), onesided, nfft ==
 * nperseg, `win` carries any scale (as in the reference), detrend is 0 (none)
 * or 1 ('constant'); nseg = (n - nperseg) / hop + 1. ---- */

/* rfft_frames (reference: src/mlx_signal_processing/_stft_metal.py:192-220):
 * power == 0 -> out complex64 [rows][nseg][nperseg/2+1];
 * power != 0 -> out float32  [rows][nseg][nperseg/2+1]. */
int msig_rfft_frames(const float* x, int64_t rows, int64_t n, const float* win, int32_t nperseg,
                     int32_t hop, int32_t detrend, int32_t power, void* out, void* stream);

/* rfft_power_mean (reference: _stft_metal.py:530-557): out float32 [rows][nperseg/2+1]. */
int msig_rfft_power_mean(const float* x, int64_t rows, int64_t n, const float* win,
                         int32_t nperseg, int32_t hop, int32_t detrend, float* out,
                         void* workspace, size_t ws_bytes, void* stream);

/* csd_frames (reference: _stft_metal.py:411-413): out complex64 [rows][nseg][bins]. */
int msig_csd_frames(const float* x, const float* y, int64_t rows, int64_t n, const float* win,
                    int32_t nperseg, int32_t hop, int32_t detrend, void* out, void* stream);

/* coherence_frames (reference: _stft_metal.py:416-418): pxx, pyy f32, pxy c64 per frame. */
int msig_coherence_frames(const float* x, const float* y, int64_t rows, int64_t n,
                          const float* win, int32_t nperseg, int32_t hop, int32_t detrend,
                          float* pxx, float* pyy, void* pxy, void* stream);

/* csd_mean (reference: _stft_metal.py:739-741): out complex64 [rows][bins]. */
int msig_csd_mean(const float* x, const float* y, int64_t rows, int64_t n, const float* win,
                  int32_t nperseg, int32_t hop, int32_t detrend, void* out, void* workspace,
                  size_t ws_bytes, void* stream);

/* coherence_mean (reference: _stft_metal.py:744-746): mean pxx, pyy f32 and pxy c64. */
int msig_coherence_mean(const float* x, const float* y, int64_t rows, int64_t n,
                        const float* win, int32_t nperseg, int32_t hop, int32_t detrend,
                        float* pxx, float* pyy, void* pxy, void* workspace, size_t ws_bytes,
                        void* stream);

/* ---- ragged batches (many mean jobs of any shape in one submission) ----
 * No counterpart in the reference (it launches one kernel per call).  A plan
 * uploads per-job descriptors and work queues to the device once, and each
 * run is kernels only, so a run can be captured in a hipGraph and replayed:
 * one persistent launch walks the work of every welch / csd job on real
 * input with hop = nfft/2 (all FFT sizes in one kernel), one persistent
 * launch per (nfft, kind, input) class covers any other job (coherence,
 * other hops, complex input, radix schedule 1), and one launch finalizes
 * every job.  Every job is computed with the same block geometry as a
 * standalone msig_spectral_mean call with the same tuning, so its result is
 * bitwise the same alone and in any batch.  Runs of one plan on one device
 * must not overlap (they share the workspace and the work counter). */

typedef struct msig_batch_job {
  msig_job job;     /* a mean kind: MSIG_PSD_MEAN, MSIG_CSD_MEAN, MSIG_COH_TRIO_MEAN, MSIG_COHERENCE */
  const void* x;
  const void* y;    /* NULL for MSIG_PSD_MEAN */
  const float* win; /* nperseg float32 on the device */
  void* out0;
  void* out1;
  void* out2;
} msig_batch_job;

/* Per FFT size (index log2(nfft)): 0 selects the default.
 *   max_run:        most segment pairs one thread group sums sequentially (<= 32)
 *   target_blocks:  blocks a job's geometry aims for (default: the standalone
 *                   call's plan, whole waves of resident blocks)
 *   blocks_per_cu:  resident blocks per CU of a per-class persistent launch
 *                   (default 4; the single launch runs at its occupancy)
 *   radix_schedule: 0 = radix-16 passes first, remainder radix last (default);
 *                   1 = remainder radix first (nfft = 16^m r, r in {2, 4, 8});
 *                   identical for powers of 16 */
typedef struct msig_tuning {
  int32_t max_run[13];
  int32_t target_blocks[13];
  int32_t blocks_per_cu[13];
  int32_t radix_schedule[13];
  int32_t per_class_launches; /* 1: one persistent launch per class for every job
                                 (blocks_per_cu applies); 0 (default): the
                                 single launch for welch / csd at hop = nfft/2 */
} msig_tuning;

typedef struct msig_batch msig_batch; /* opaque plan */

/* Host-only planning: no device allocation and no copy. */
int msig_batch_create(const msig_batch_job* jobs, int32_t njobs, const msig_tuning* tuning,
                      msig_batch** plan);
/* Bytes of the plan's device descriptor block (job descriptors + work queues). */
int msig_batch_descriptor_bytes(const msig_batch* plan, size_t* bytes);
/* Write the descriptor block into host memory `host` (>= descriptor bytes),
 * laid out for the device address `device` it will be copied to.  The caller
 * copies it (e.g. hipMemcpyAsync from pinned memory on the run's stream)
 * before the first run that uses `device`. */
int msig_batch_pack(const msig_batch* plan, void* host, size_t bytes, void* device);
int msig_batch_workspace_bytes(const msig_batch* plan, size_t* bytes);
/* Launches on `stream` using the descriptor block at `device` (packed for that
 * address); allocates nothing and copies nothing from the host, so it can be
 * captured in a hipGraph. */
int msig_batch_run(const msig_batch* plan, const void* device, void* workspace, size_t ws_bytes,
                   void* stream);
/* Number of kernel launches one run makes (2 for a batch of welch / csd jobs at
 * hop = nfft/2; one more per other class). */
int msig_batch_launches(const msig_batch* plan);
int msig_batch_destroy(msig_batch* plan);

/* One candidate tile shape for one FFT size: the msig_tuning fields at that
 * size's index (0 selects the default). */
typedef struct msig_tile {
  int32_t max_run;
  int32_t target_blocks;
  int32_t blocks_per_cu;
  int32_t radix_schedule;
} msig_tile;

/* Autotuner (the counterpart of the proposed msig_autotune; the reference has
 * no tuner).  Times every candidate tile for the FFT size of the probe jobs,
 * which must all share one nfft: each candidate is `base` (NULL: defaults)
 * with that size's entries replaced by the candidate, planned over the probe
 * jobs, its descriptor block copied to a device buffer the call allocates,
 * run once untimed and then `reps` times between HIP events on `stream`.
 * Before the first candidate the first plan runs back to back for about
 * 60 ms so every candidate is timed at sustained clocks.  ms_out[i] receives
 * candidate i's median ms per run and *best the index of the fastest.  The
 * call allocates and frees device memory, writes the probe jobs' outputs,
 * and synchronises `stream`; it must not run inside a graph capture. */
int msig_batch_autotune(const msig_batch_job* probe, int32_t njobs, const msig_tuning* base,
                        const msig_tile* cands, int32_t ncand, int32_t reps, float* ms_out,
                        int32_t* best, void* stream);

/* Workspace bytes for the operator-level mean entry points
 * (kind = MSIG_PSD_MEAN, MSIG_CSD_MEAN or MSIG_COH_TRIO_MEAN). */
int msig_mean_workspace_bytes(int32_t kind, int64_t rows, int64_t n, int32_t nperseg, int32_t hop,
                              size_t* bytes);

/* istft, inverse half (reference: src/mlx_signal_processing/_stft_metal.py:753-791, the
 * inverse kernel of istft_frames :817-870): z complex64 [rows][nseg][nperseg/2+1]
 * one-sided spectra; weff float32[nperseg] = window * scale / nperseg;
 * frames float32 [rows][nseg][nperseg] = irfft(z) * weff per frame (the
 * imaginary parts of the DC and Nyquist bins are ignored, as irfft does).
 * nperseg a power of two in [MSIG_MIN_NFFT, MSIG_MAX_NFFT], else MSIG_ENOTSUP. */
int msig_istft_frames(const void* z, int64_t rows, int32_t nseg, int32_t nperseg,
                      const float* weff, float* frames, void* stream);

/* istft with the overlap-add fused, for hop = nperseg/2 (reference: the two kernels of
 * istft_frames, src/mlx_signal_processing/_stft_metal.py:753-815, called from istft,
 * spectral.py:824-992): msig_istft_frames followed by msig_overlap_add(channels = 1,
 * hop = nperseg/2) on its frames, without the [rows][nseg][nperseg] frames in memory.  The
 * window-power divide is folded into per-sample weights (weff / norm, rounded once), so
 * results differ from the two-call form by a few ulps.  z complex64 [rows][nseg][nperseg/2+1];
 * weff = window * scale / nperseg and win (the window, for the window-power
 * divide) float32[nperseg]; out float32 [rows][count] = samples first .. first+count-1,
 * first + count <= (nseg + 1) * nperseg/2.  side_src / side_dst / side_len: an optional
 * side copy as in msig_job (istft's time grid).  nperseg a power of two in
 * [MSIG_MIN_NFFT, MSIG_MAX_NFFT], else MSIG_ENOTSUP. */
int msig_istft_ola(const void* z, int64_t rows, int32_t nseg, int32_t nperseg,
                   const float* weff, const float* win, int64_t first, int64_t count, float* out,
                   const float* side_src, float* side_dst, int64_t side_len, void* stream);

/* Overlap-add with the window-power divide (reference: _stft_metal.py:793-815,
 * the gather kernel of istft_frames): frames float32 [rows][nseg][nperseg][channels]
 * (channels 1 = real, 2 = interleaved complex); for i = first + j, j < count:
 *   out[row][j][c] = sum_s frames[row][s][i - s*hop][c] / norm(i),
 *   norm(i) = sum_s win[i - s*hop]^2 (the divide only where norm > 1e-10),
 * summed over the segments covering i in increasing s (any row count);
 * first + count <= nperseg + (nseg - 1) * hop. */
int msig_overlap_add(const float* frames, int64_t rows, int32_t nseg, int32_t nperseg,
                     int32_t hop, int32_t channels, const float* win, int64_t first,
                     int64_t count, float* out, void* stream);

/* Per-column median over segments (reference: the median average of csd,
 * src/mlx_signal_processing/spectral.py:469-477): frames float32 [rows][nseg][cols];
 *   out[row][c] = median_s frames[row][s][c] * scale * f(c),
 * numpy's median (mean of the two middle values for an even nseg; NaN when
 * the column holds a NaN), f(c) = 2 when apply_factor and the column's bin
 * b = c / channels is neither 0 nor nfft/2, else 1 (channels 2: interleaved
 * complex, real and imaginary parts take their own medians).
 * nseg <= MSIG_MEDIAN_MAX_NSEG, else MSIG_ENOTSUP. */
#define MSIG_MEDIAN_MAX_NSEG 16000
int msig_median_frames(const float* frames, int64_t rows, int32_t nseg, int32_t cols,
                       int32_t channels, double scale, int32_t apply_factor, int32_t nfft,
                       float* out, void* stream);

/* Kernel timing (instrumentation; no counterpart in the reference, which
 * times whole calls).  After msig_kernel_timing_arm(k) on a device, the next
 * k calls of msig_spectral_mean / msig_spectral_frames (and the operator-level
 * entry points built on them) on that device hand a HIP event pair to the
 * launch of their fused transform kernel (hipExtLaunchKernel), which stamps
 * the pair with that dispatch's start and end on the call's stream; the
 * finalize is excluded.  msig_kernel_timing_read waits for the recorded pairs, returns
 * their summed elapsed milliseconds and count, and disarms.  k = 0 disarms.
 * Not for use under graph capture. */
int msig_kernel_timing_arm(int32_t max_launches);
int msig_kernel_timing_read(double* total_ms, int32_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* MSIG_H */
