// Fused spectral kernels instantiated for FFT size 4096.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(4096)
