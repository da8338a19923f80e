// Fused spectral kernels instantiated for FFT size 1024.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(1024)
