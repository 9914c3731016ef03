"""fgpe_run_noalloc (the step loop with its NORMC partials in the library's
fixed device array) against fgpe_run, through the C ABI.

Both entry points run the same kernels, and the partials are summed in a
fixed order, so the results must be bit-identical: the state, the last NORMC
and the observables.  Calls in a row on one GPU share the array, and the
second and third calls here use smaller grids than the first, so leftovers
from a larger call must not leak into a smaller one.
"""
import numpy as np
import pytest

from cases import case, noisy
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64

pytestmark = pytest.mark.gpu


def start(inp, prec, seed):
    p, _ = fgpe.derive(inp, prec)
    with fgpe.Solver(p) as s:
        s.initial_guess()
        s.normalize()
        phi = s.download()
    return p, (phi if inp.switch_im else noisy(phi, seed))


def compare(inp, prec, nsteps, seed):
    p, phi0 = start(inp, prec, seed)
    out = []
    for noalloc in (False, True):
        with fgpe.Solver(p) as s:
            s.upload(phi0)
            (s.run_noalloc if noalloc else s.run)(nsteps)
            out.append((s.download(), s.last_norms(), s.observables()))
    (a, na, oa), (b, nb, ob) = out
    assert np.array_equal(a, b)
    assert np.array_equal(na, nb)
    assert oa[0] == ob[0] and oa[2] == ob[2]
    assert np.array_equal(oa[1], ob[1]) and np.array_equal(oa[3], ob[3])


CALLS = [
    # 3D spin-2 64^3 FP32 real time (16384 x-pass blocks of partials)
    ("3d_s2_64_fp32_real", case(3, 2, (64, 64, 64), (0.4, 0.4, 0.4), 5e-4, (0.5, 0.5, 0.5), imag=0, natoms=100000,
                                nu=(20, 20, 20)), FGPE_FP32, 12),
    # then a smaller input: 2D spin-1 64x32 FP64 imaginary time with the SOC = 0
    # projection (NORMC every step)
    ("2d_s1_64x32_fp64_imag", case(2, 1, (64, 32), (0.1, 0.2), 1e-3, (0.0, 0.0), mag=0.2, nu=(20, 20, 400)),
     FGPE_FP64, 20),
    # and the q1D kernel, which keeps its norms inside its one block
    ("1d_s2_256_fp64_imag", case(1, 2, (256,), (0.05,), 1.25e-4, (0.5,)), FGPE_FP64, 40),
]


def test_run_noalloc_matches_run_on_calls_in_a_row():
    for name, inp, prec, nsteps in CALLS:
        print("no-alloc run vs run:", name, flush=True)
        compare(inp, prec, nsteps, seed=3)
