"""Lane-level CPU model of the split-precision matrix-core kernel
(mlx-signal_amd/csrc/msig_mfma.hip).

The model restates the kernel's index maps exactly: which sample each lane
loads, the depth interleave k = 8g + 2r + part of the 16 x 32 by 32 x 16
products, the constant DFT16 operands, the stage-2 twiddles, the lane -> bin
map of the result, and the hi + lo operand split with the dropped lo*lo
term.  The matrix-core product itself (lane (g, c) holds A[c][8g+j],
B[8g+j][c] and D[4g+r][c]) is the layout the GPU probe of
v_mfma_f32_16x16x32_{f16,bf16} confirmed.  A mistake in any of the maps makes
the model disagree with numpy's FFT of the same packed segment pair.
"""

import numpy as np
import pytest

N = 256
LANES = np.arange(64)
G = LANES >> 4
C = LANES & 15


def _bf16(v):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    b = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _round(v, prec):
    v = np.asarray(v, np.float32)
    if prec == "f16":
        return v.astype(np.float16).astype(np.float32)
    if prec == "bf16":
        return _bf16(v)
    return v  # exact model


def _split(v, prec):
    hi = _round(v, prec)
    lo = _round(np.asarray(v, np.float32) - hi, prec)
    return hi, lo


def _operand_matrix(lane_vals, role):
    """lane_vals[l][j] -> the matrix the hardware sees.
    A (16 x 32): A[c][8g + j];  B (32 x 16): B[8g + j][c]."""
    m = np.zeros((16, 32) if role == "A" else (32, 16), np.float64)
    for l in range(64):
        for j in range(8):
            if role == "A":
                m[C[l], 8 * G[l] + j] = lane_vals[l][j]
            else:
                m[8 * G[l] + j, C[l]] = lane_vals[l][j]
    return m


def _accumulator_lanes(d):
    """D (16 x 16) -> lane (g, c) holds D[4g + r][c], r = 0..3."""
    return np.array([[d[4 * G[l] + r, C[l]] for r in range(4)] for l in range(64)])


def _mma(a_lanes, b_lanes, prec):
    if prec == "exact":
        return _operand_matrix(a_lanes, "A") @ _operand_matrix(b_lanes, "B")
    ah, al = _split(a_lanes, prec)
    bh, bl = _split(b_lanes, prec)
    A = lambda v: _operand_matrix(v, "A")
    B = lambda v: _operand_matrix(v, "B")
    return A(al) @ B(bh) + A(ah) @ B(bl) + A(ah) @ B(bh)


def _constants():
    tw = np.exp(-2j * np.pi * np.arange(N) / N).astype(np.complex64)
    bre = np.zeros((64, 8), np.float32)
    bim = np.zeros((64, 8), np.float32)
    for l in range(64):
        for j in range(8):
            m = ((4 * G[l] + (j >> 1)) * C[l]) & 15
            cs, sn = tw[16 * m].real, -tw[16 * m].imag
            bre[l, j] = sn if j & 1 else cs
            bim[l, j] = cs if j & 1 else -sn
    t2 = np.array([[tw[((4 * G[l] + r) * C[l]) & (N - 1)] for r in range(4)] for l in range(64)])
    return bre, bim, t2


def model_pair_spectrum(xs, xt, win, prec, detrend):
    """Z[k] of the packed pair z = (xs - ms) w + i (xt - mt) w, as the kernel
    forms it, scattered to natural bin order."""
    bre, bim, t2 = _constants()
    ms = xs.mean() if detrend else 0.0
    mt = xt.mean() if detrend else 0.0
    v = np.zeros((64, 8), np.float32)
    for l in range(64):
        for r in range(4):
            n = 64 * G[l] + 16 * r + C[l]
            v[l, 2 * r] = (xs[n] - ms) * win[n]
            v[l, 2 * r + 1] = (xt[n] - mt) * win[n]
    yre = _accumulator_lanes(_mma(v, bre, prec))
    yim = _accumulator_lanes(_mma(v, bim, prec))
    y = (yre + 1j * yim) * t2
    a2 = np.zeros((64, 8), np.float32)
    a2[:, 0::2] = y.real
    a2[:, 1::2] = y.imag
    zre = _accumulator_lanes(_mma(a2, bre, prec))
    zim = _accumulator_lanes(_mma(a2, bim, prec))
    out = np.zeros(N, np.complex128)
    for l in range(64):
        for r in range(4):
            out[4 * G[l] + r + 16 * C[l]] = zre[l, r] + 1j * zim[l, r]
    return out


def _pair(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    t = np.arange(N)
    xs = (rng.standard_normal(N) + 0.5 + np.sin(2 * np.pi * 17.5 * t / N)) * scale
    xt = (rng.standard_normal(N) - 0.25) * scale
    win = (0.5 - 0.5 * np.cos(2 * np.pi * t / N)).astype(np.float32)
    return xs.astype(np.float32), xt.astype(np.float32), win


@pytest.mark.parametrize("detrend", [False, True])
def test_index_maps_reproduce_the_dft(detrend):
    xs, xt, win = _pair(3)
    ms = xs.mean() if detrend else 0.0
    mt = xt.mean() if detrend else 0.0
    ref = np.fft.fft((xs - ms) * win + 1j * (xt - mt) * win)
    got = model_pair_spectrum(xs, xt, win, "exact", detrend)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-4 * np.abs(ref).max())


@pytest.mark.parametrize("prec,tol", [("f16", 2e-6), ("bf16", 1e-4)])
def test_split_precision_error_bound(prec, tol):
    xs, xt, win = _pair(5)
    ref = np.fft.fft(xs.astype(np.float64) * win + 1j * xt.astype(np.float64) * win)
    got = model_pair_spectrum(xs, xt, win, prec, False)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < tol, err


def test_split_beats_unsplit_rounding():
    """hi + lo must be far more accurate than rounding the operands once."""
    xs, xt, win = _pair(7)
    ref = np.fft.fft(xs.astype(np.float64) * win + 1j * xt.astype(np.float64) * win)
    split = np.abs(model_pair_spectrum(xs, xt, win, "f16", False) - ref).max()
    bre, bim, t2 = _constants()
    # one-term product: drop both lo parts
    v = np.zeros((64, 8), np.float32)
    for l in range(64):
        for r in range(4):
            n = 64 * G[l] + 16 * r + C[l]
            v[l, 2 * r], v[l, 2 * r + 1] = xs[n] * win[n], xt[n] * win[n]
    one = _operand_matrix(_round(v, "f16"), "A") @ _operand_matrix(_round(bre, "f16"), "B")
    exact = _operand_matrix(v, "A") @ _operand_matrix(bre, "B")
    assert split < 1e-5 * np.abs(ref).max()
    assert np.abs(one - exact).max() > 50 * split / 16  # one-term stage-1 error alone is larger


def test_planted_index_mistake_fails():
    """Swapping the stage-2 twiddle exponent's factors for a wrong map breaks
    the DFT: the model check is sensitive to the index maps."""
    xs, xt, win = _pair(9)
    ref = np.fft.fft(xs * win + 1j * xt * win)
    bre, bim, t2 = _constants()
    bad_t2 = np.array([[np.exp(-2j * np.pi * (((4 * G[l] + r) * C[l] + 1) % N) / N)
                        for r in range(4)] for l in range(64)])
    v = np.zeros((64, 8), np.float32)
    for l in range(64):
        for r in range(4):
            n = 64 * G[l] + 16 * r + C[l]
            v[l, 2 * r], v[l, 2 * r + 1] = xs[n] * win[n], xt[n] * win[n]
    y = (_accumulator_lanes(_mma(v, bre, "exact")) + 1j * _accumulator_lanes(_mma(v, bim, "exact"))) * bad_t2
    a2 = np.zeros((64, 8), np.float32)
    a2[:, 0::2], a2[:, 1::2] = y.real, y.imag
    z = _accumulator_lanes(_mma(a2, bre, "exact")) + 1j * _accumulator_lanes(_mma(a2, bim, "exact"))
    out = np.zeros(N, np.complex128)
    for l in range(64):
        for r in range(4):
            out[4 * G[l] + r + 16 * C[l]] = z[l, r]
    assert np.abs(out - ref).max() > 1e-2 * np.abs(ref).max()


def test_fp16_pair_scale_keeps_range():
    """Large-amplitude input: the kernel's power-of-two pair scale keeps the
    FP16 operands finite; the model applies the same scale and unscale."""
    xs, xt, win = _pair(11, scale=3e4)
    b = 2 * max(np.abs(xs).max(), np.abs(xt).max()) * np.abs(win).max()
    e = 11 - int(np.frexp(np.float32(b))[1])
    sc = np.float32(2.0 ** e)
    ref = np.fft.fft(xs.astype(np.float64) * win + 1j * xt.astype(np.float64) * win)
    got = model_pair_spectrum(xs * sc, xt * sc, win, "f16", False) / sc
    assert np.isfinite(got).all()
    assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()
    with np.errstate(over="ignore", invalid="ignore"):
        unscaled = model_pair_spectrum(xs, xt, win, "f16", False)
    assert not np.isfinite(unscaled).all()  # without the scale FP16 overflows


def _geometry_mfma(rows, nseg, maxrun=32, target_blocks=2048, gpb=4):
    units = (nseg + 1) // 2
    bprt = max(1, (target_blocks + rows - 1) // rows)
    run = min(max(1, (units + gpb * bprt - 1) // (gpb * bprt)), maxrun)
    bpr = (units + gpb * run - 1) // (gpb * run)
    return units, run, bpr


def _units_processed(nseg, run, bpr, units, gpb=4):
    """The kernel's per-wave control flow: the unrolled full-pair loop, then
    the partner-less tail unit."""
    full = nseg // 2
    seen = []
    for bir in range(bpr):
        for wave in range(gpb):
            u0 = (bir * gpb + wave) * run
            u1 = min(u0 + run, units)
            uf = min(u1, full)
            if u0 < uf:
                u = u0
                while u + 1 < uf:
                    seen += [(u, True), (u + 1, True)]
                    u += 2
                if u < uf:
                    seen.append((u, True))
            if u0 <= full < u1:
                seen.append((full, False))
    return seen


@pytest.mark.parametrize("rows", [1, 3, 64, 256, 1000])
def test_wave_partition_processes_each_unit_once(rows):
    for nseg in list(range(1, 80)) + [127, 511, 2047, 4097, 65535]:
        units, run, bpr = _geometry_mfma(rows, nseg)
        seen = _units_processed(nseg, run, bpr, units)
        assert sorted(u for u, _ in seen) == list(range(units)), (rows, nseg)
        for u, partner in seen:
            assert partner == (2 * u + 1 < nseg), (rows, nseg, u)
