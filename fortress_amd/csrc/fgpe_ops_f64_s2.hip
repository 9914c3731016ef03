// fgpe_ops_f64_s2.hip -- instantiation of the FP64, spin-2 kernels and launches.
#include "fgpe_ops.hpp"

FGPE_INSTANTIATE_SPIN(double, 2)
