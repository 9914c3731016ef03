"""Host-side argument checks of the batch API (no device needed: every check
runs before the library looks for a GPU)."""
import pytest

from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input


def params(n=64, spin=1, precision=FGPE_FP32, imag=1, ny=None, dim=2):
    a = (101.8, 100.4, 0.0) if spin == 1 else (87.93, 91.28, 99.18)
    grid = (n, ny or n) if dim == 2 else (n, n, n)
    inp = make_input(dim, spin, grid, (0.1,) * dim, 1e-3, 87.0, a, 10000, (20.0, 20.0, 400.0),
                     (0.5,) * dim, switch_im=imag)
    return fgpe.derive(inp, precision)[0]


def test_grid_sizes_outside_the_batched_set_are_refused():
    with pytest.raises(fgpe.FgpeError, match="64, 128, 192 or 256 points per axis"):
        fgpe.Batch([params(64), params(96)])
    with pytest.raises(fgpe.FgpeError, match="64, 128, 192 or 256 points per axis"):
        fgpe.Batch([params(64, ny=32)])


def test_mixed_spin_and_other_dimensions_are_refused():
    with pytest.raises(fgpe.FgpeError, match="mixed spin"):
        fgpe.Batch([params(spin=1), params(spin=2)])
    with pytest.raises(fgpe.FgpeError, match="only q2D"):
        fgpe.Batch([params(32, dim=3)])


def test_mixed_precision_and_time_mode_are_refused():
    with pytest.raises(fgpe.FgpeError, match="mixed precision"):
        fgpe.Batch([params(precision=FGPE_FP32), params(precision=FGPE_FP64)])
    with pytest.raises(fgpe.FgpeError, match="mixed real and imaginary"):
        fgpe.Batch([params(imag=1), params(imag=0)])
