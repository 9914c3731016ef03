// fgpe_ops_f32_s1.hip -- instantiation of the FP32, spin-1 kernels and launches.
#include "fgpe_ops.hpp"

FGPE_INSTANTIATE_SPIN(float, 1)
FGPE_INSTANTIATE_PREC(float)
