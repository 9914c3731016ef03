"""Host-side argument checks of the batch API (no device needed: every check
runs before the library looks for a GPU)."""
import pytest

from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input


def params(n=64, spin=1, precision=FGPE_FP32, imag=1):
    inp = make_input(2, spin, (n, n), (0.1, 0.1), 1e-3, 87.0, (101.8, 100.4, 0.0), 10000, (20.0, 20.0, 400.0),
                     (0.5, 0.5), switch_im=imag)
    return fgpe.derive(inp, precision)[0]


def test_grid_sizes_outside_the_batched_set_are_refused():
    with pytest.raises(fgpe.FgpeError, match="square 64, 128, 192 or 256"):
        fgpe.Batch([params(64), params(96)])


def test_spin2_instances_are_refused():
    with pytest.raises(fgpe.FgpeError, match="spin-1"):
        fgpe.Batch([params(spin=2)])


def test_mixed_precision_and_time_mode_are_refused():
    with pytest.raises(fgpe.FgpeError, match="mixed precision"):
        fgpe.Batch([params(precision=FGPE_FP32), params(precision=FGPE_FP64)])
    with pytest.raises(fgpe.FgpeError, match="mixed real and imaginary"):
        fgpe.Batch([params(imag=1), params(imag=0)])
