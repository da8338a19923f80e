// Fused spectral kernels instantiated for FFT size 512.
#include "msig_launch.hpp"

MSIG_DEFINE_LAUNCHERS(512)
