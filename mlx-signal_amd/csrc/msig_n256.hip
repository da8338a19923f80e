// Fused spectral kernels instantiated for FFT size 256.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(256)
