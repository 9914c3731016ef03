"""Run-to-run reproducibility of the propagator on the GPU: the same input
through fresh handles gives bit-identical states and norms.  Every pass reads
and writes disjoint tiles and reduces in a fixed order, so any difference is
a race.  The grids include 96-point y lines with large NX*NZ, where a pass
shape that spread one line over two waves once gave results that changed
from run to run (real time amplifies a single flipped bit to O(1) within ten
steps)."""
import numpy as np
import pytest

from cases import case, noisy
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64

pytestmark = pytest.mark.gpu

GRIDS_3D = [(32, 96, 192), (96, 96, 96), (64, 96, 256), (32, 192, 96), (64, 64, 192), (32, 384, 32), (32, 32, 512), (32, 512, 32), (32, 256, 64)]


def twice(inp, prec, nsteps, repeats=3):
    p, _ = fgpe.derive(inp, prec)
    outs = []
    for _ in range(repeats):
        with fgpe.Solver(p) as s:
            s.initial_guess()
            s.normalize()
            s.upload(noisy(s.download(), 3))
            s.run(nsteps)
            outs.append((s.download(), s.last_norms()))
    return outs


@pytest.mark.parametrize("prec", [FGPE_FP64, FGPE_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("n", GRIDS_3D, ids=lambda n: "x".join(map(str, n)))
def test_3d_bit_reproducible(n, spin, prec):
    inp = case(3, spin, n, (0.4, 0.3, 0.2), 5e-4, (0.5, 0.5, 0.5), imag=0, natoms=100000, nu=(20, 20, 20))
    print("3d", n, "spin", spin, "prec", prec, flush=True)
    outs = twice(inp, prec, 10)
    for phi, norms in outs[1:]:
        assert np.array_equal(phi, outs[0][0])
        assert np.array_equal(norms, outs[0][1])


@pytest.mark.parametrize("prec", [FGPE_FP64, FGPE_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [(96, 192), (256, 96), (384, 384)], ids=lambda n: "%dx%d" % n)
def test_2d_bit_reproducible(n, prec):
    inp = case(2, 2, n, (0.1, 0.1), 1e-3, (1.0, 0.5), imag=0, natoms=30000, nu=(20, 20, 400))
    print("2d", n, "prec", prec, flush=True)
    outs = twice(inp, prec, 20)
    for phi, norms in outs[1:]:
        assert np.array_equal(phi, outs[0][0])
        assert np.array_equal(norms, outs[0][1])
