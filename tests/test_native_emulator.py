"""CPU checks of the device FFT: the host emulator (tests/native/fft_emu.cpp)
runs the kernels' own msig_fft.hpp templates lane by lane with an emulated
LDS and compares against an O(N^2) FP64 DFT, checks LDS slot coverage, that
the split exchange's pads repeat no LDS bank among 32 lanes, and the
two-for-one fold identities the kernels rely on."""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_fft_emulator(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "fft_emu"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{ROOT / 'mlx-signal_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "fft_emu.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 62
