// Fused spectral kernels instantiated for FFT size 64.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(64)
