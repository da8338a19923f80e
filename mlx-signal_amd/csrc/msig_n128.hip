// Fused spectral kernels instantiated for FFT size 128.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(128)
