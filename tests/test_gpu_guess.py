"""The device initial guess (every OPTION_FPC branch) against the oracle's.

INITIALIZE's guess branch (SRC/cgpe3d.f90:754-833, SRC/cgpe2d.f90:697-771,
SRC/cgpe1d.f90:707-786) is evaluated pointwise on the GPU by
fgpe_initial_guess; the oracle's guess is pinned by closed forms in
test_oracle_pins.py.  FP64 handles agree to rounding (libm and the device
math library may differ in the last bits of pow/exp); FP32 handles hold the
guess rounded to float.  The 2D spin-1 Thomas-Fermi branch keeps the
reference's NaN outside the TF radius (SRC/cgpe2d.f90:697-771): the device
must put its NaNs at the same points.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from cases import GUESS_CASES, case, rel_l2
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,spin,fpc,mag,n,d,natoms,nu,prof", GUESS_CASES)
def test_device_guess_matches_oracle(dim, spin, fpc, mag, n, d, natoms, nu, prof):
    inp = case(dim, spin, (n,) * dim, (d,) * dim, 1e-3, (0.0,) * dim, natoms=natoms, nu=nu, option_fpc=fpc,
               mag=mag)
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    ref = o.get_phi()
    o.close()
    for prec, tol in ((FGPE_FP64, 1e-13), (FGPE_FP32, 1e-6)):
        print("guess dim", dim, "spin", spin, "fpc", fpc, prof, "precision", prec, flush=True)
        pg, _ = fgpe.derive(inp, prec)
        with fgpe.Solver(pg) as s:
            s.initial_guess()
            got = s.download()
        want = ref
        if prof == "tf_nan":
            # the reference's sqrt of a negative number outside the TF radius:
            # NaN at the same points on both sides, agreement elsewhere
            bad = np.isnan(ref)
            assert bad.any() and np.array_equal(np.isnan(got), bad)
            got, want = np.where(bad, 0.0, got), np.where(bad, 0.0, ref)
        err = rel_l2(got, want)
        print("  rel L2", err, flush=True)
        assert np.isfinite(got).all() and err < tol
