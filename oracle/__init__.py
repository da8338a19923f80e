"""Parity oracle for the fused STFT/Welch path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to compare results against.  The product
package never imports it.  See spectral_fp64.py for what each function restates and how it
is pinned.
"""

from .spectral_fp64 import *  # noqa: F401,F403
