// Fused spectral kernels instantiated for FFT size 32.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(32)
