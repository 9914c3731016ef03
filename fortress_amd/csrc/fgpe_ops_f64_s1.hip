// fgpe_ops_f64_s1.hip -- instantiation of the FP64, spin-1 kernels and launches.
#include "fgpe_ops.hpp"

FGPE_INSTANTIATE_SPIN(double, 1)
FGPE_INSTANTIATE_PREC(double)
