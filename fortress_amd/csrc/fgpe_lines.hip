// fgpe_lines.hip -- the pass launches of one (precision, spin, line length),
// chosen by the build: FGPE_T (float or double), FGPE_SPIN (1 or 2), FGPE_N,
// and FGPE_LINES (2D/3D passes) and/or FGPE_ONED (the q1D kernel).
#include "fgpe_launch.hpp"

#if FGPE_LINES
FGPE_INSTANTIATE_LINES(FGPE_T, FGPE_SPIN, FGPE_N)
#if FGPE_SPIN == 1
FGPE_INSTANTIATE_YZ(FGPE_T, FGPE_N)
#endif
#endif
#if FGPE_ONED
FGPE_INSTANTIATE_1D(FGPE_T, FGPE_SPIN, FGPE_N)
#endif
