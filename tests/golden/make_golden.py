"""Write the golden fixtures: float32 inputs + scipy.signal FP64 outputs.

Run from the repository root:  python tests/golden/make_golden.py

scipy.signal (the reference's own CPU route, spectral.py:433-450, and the
library its tests compare against) is called on float64 copies of the float32
inputs, so the expected values are FP64.  Inputs are committed (NumPy
Generator streams are not guaranteed stable across versions).  Cases follow
the reference tests' case lists (tests/test_spectral.py) plus the headline
single-channel configuration (1 x 2^16 N(0,1), seed 0, fs=48000, nperseg=256,
noverlap=128 -> 129 bins).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import scipy.signal as sps

HERE = Path(__file__).resolve().parent


def make_sig(rng, shape, complex_=False):
    x = rng.standard_normal(shape).astype(np.float32)
    n = shape[-1]
    t = np.arange(n, dtype=np.float32)
    x = x + 0.5 * np.sin(2 * np.pi * 0.12 * t).astype(np.float32)
    if complex_:
        x = (x + 1j * rng.standard_normal(shape)).astype(np.complex64)
    return x


def main():
    arrays = {}
    cases = []

    def add(name, func, inputs, kwargs, outputs):
        for k, v in inputs.items():
            arrays[f"{name}/in/{k}"] = v
        for i, v in enumerate(outputs):
            arrays[f"{name}/out/{i}"] = np.asarray(v)
        cases.append({"name": name, "func": func, "inputs": sorted(inputs), "kwargs": kwargs,
                      "noutputs": len(outputs)})

    # headline single-channel config
    rng0 = np.random.default_rng(0)
    x = rng0.standard_normal((1, 1 << 16)).astype(np.float32)
    kw = dict(fs=48000.0, window="hann", nperseg=256, noverlap=128, detrend="constant",
              scaling="density", average="mean")
    add("welch_1x2p16", "welch", {"x": x}, kw, sps.welch(x.astype(np.float64), **kw))

    rng = np.random.default_rng(1234)
    welch_cases = [
        {}, {"nperseg": 128}, {"nperseg": 128, "noverlap": 96},
        {"nperseg": 128, "window": "hamming"}, {"nperseg": 100, "window": ["tukey", 0.25]},
        {"nperseg": 128, "detrend": "linear"}, {"nperseg": 128, "detrend": False},
        {"nperseg": 128, "scaling": "spectrum"}, {"nperseg": 128, "return_onesided": False},
        {"nperseg": 128, "average": "median"}, {"nperseg": 128, "nfft": 256},
        {"nperseg": 129, "noverlap": 37, "fs": 250.0},
    ]
    for i, c in enumerate(welch_cases):
        x = make_sig(rng, (4093,))
        ck = dict(c)
        if "window" in ck and isinstance(ck["window"], list):
            ck["window"] = tuple(ck["window"])
        add(f"welch_case{i}", "welch", {"x": x}, c, sps.welch(x.astype(np.float64), **ck))

    x = make_sig(rng, (3, 4096))
    y = make_sig(rng, (3, 4096))
    for av in ("mean", "median"):
        kw = dict(nperseg=256, average=av)
        add(f"csd_{av}", "csd", {"x": x, "y": y}, kw,
            sps.csd(x.astype(np.float64), y.astype(np.float64), **kw))
    yc = (0.5 * x + 0.5 * make_sig(rng, (3, 4096))).astype(np.float32)
    kw = dict(nperseg=512, detrend=False)
    add("coherence", "coherence", {"x": x, "y": yc}, kw,
        sps.coherence(x.astype(np.float64), yc.astype(np.float64), **kw))

    x = make_sig(rng, (1, 4000))
    for i, c in enumerate([{}, {"nperseg": 128, "noverlap": 100}, {"boundary": "even"},
                           {"scaling": "psd"}, {"return_onesided": False}]):
        kw = dict(c, fs=8000.0)
        add(f"stft_case{i}", "stft", {"x": x}, kw, sps.stft(x.astype(np.float64), **kw))
    for mode in ("psd", "complex", "magnitude"):
        kw = dict(fs=100.0, mode=mode)
        add(f"spectrogram_{mode}", "spectrogram", {"x": x}, kw,
            sps.spectrogram(x.astype(np.float64), **kw))

    np.savez_compressed(HERE / "spectral_golden.npz", **arrays)
    (HERE / "spectral_golden.json").write_text(json.dumps(cases, indent=1) + "\n")
    total = sum(a.nbytes for a in arrays.values())
    print(f"wrote {len(cases)} cases, {total / 1e6:.2f} MB uncompressed")


if __name__ == "__main__":
    main()
