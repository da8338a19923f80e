// Fused spectral kernels instantiated for FFT size 16.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(16)
