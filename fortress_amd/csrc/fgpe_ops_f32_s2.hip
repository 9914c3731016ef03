// fgpe_ops_f32_s2.hip -- instantiation of the FP32, spin-2 kernels and launches.
#include "fgpe_ops.hpp"

FGPE_INSTANTIATE_SPIN(float, 2)
