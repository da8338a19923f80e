// Fused spectral kernels instantiated for FFT size 2048.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(2048)
