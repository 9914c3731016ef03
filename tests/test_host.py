"""Host-side logic of the drop-in boundary, checked without a GPU.

* INPUT/input1D.dat is read with the reference's record semantics
  (SRC/cgpe1d.f90:233-260) and written back identically;
* fgpe_derive (the driver's setup arithmetic inside libfgpe) equals the
  oracle's restatement, which the golden header pins;
* libfgpe.so exports every entry point include/fgpe.h declares;
* without a usable device, fgpe_create fails loudly instead of falling back.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_ctypes as O
from cases import case
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64
from fortress_amd.inputs import read_input, write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_read_golden_input():
    inp = read_input(os.path.join(GOLD, "input1D.dat"), 1)
    assert (inp.spin, inp.niter, inp.stp, inp.nstp, inp.threads) == (2, 9900000, 400, 10000, 8)
    assert inp.n[0] == 2048 and inp.d[0] == 0.05 and inp.dt == 0.000125
    assert inp.mass == 23.0 and list(inp.a) == [34.9, 45.8, 64.5]
    assert inp.natoms == 10000 and list(inp.nu) == [20.0, 400.0, 400.0]
    assert inp.tol == 1e-6 and inp.switch_im == 1 and inp.gamma[0] == 0.5
    # SO coupling on: OPTION_FPC and MAG records are skipped
    assert inp.option_fpc == 4 and inp.mag == 0.0


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("soc", [0, 1])
def test_write_read_roundtrip(tmp_path, dim, soc):
    g = (0.5, 0.25, 0.125)[:dim] if soc else (0.0,) * dim
    inp = case(dim, 2, (64, 32, 16)[:dim], (0.1, 0.2, 0.3)[:dim], 1e-3, g, mag=0.3 if not soc else 0.0,
               option_fpc=2 if not soc else 4)
    f = tmp_path / ("input%dD.dat" % dim)
    write_input(str(f), inp)
    back = read_input(str(f), dim)
    for name, _ in inp._fields_:
        a, b = getattr(inp, name), getattr(back, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("gam", [0.0, 0.5])
def test_library_derive_matches_oracle(dim, spin, gam):
    inp = case(dim, spin, (64, 64, 64)[:dim], (0.1,) * dim, 1e-3, (gam,) * dim, natoms=123457,
               nu=(20.0, 33.0, 400.0), option_fpc=2, mag=0.25)
    po, ao = O.derive(inp)
    for prec in (FGPE_FP64, FGPE_FP32):
        pg, ag = fgpe.derive(inp, prec)
        assert pg.precision == prec
        assert ag == ao
        for f in ("dim", "spin", "imag", "soc", "option_fpc"):
            assert getattr(pg, f) == getattr(po, f), f
        for f in ("alpha", "gamma", "d", "n"):
            assert list(getattr(pg, f)) == list(getattr(po, f)), f
        # same formula; host compilers may contract a*b+c differently (1 ulp)
        assert np.allclose(list(pg.tau), list(po.tau), rtol=1e-15, atol=0), "tau"
        assert pg.dt == po.dt and pg.mag == po.mag


def test_derive_rejects_bad_input():
    inp = case(3, 2, (64, 64, 64), (0.1,) * 3, 1e-3, (0.5,) * 3)
    inp.spin = 3
    with pytest.raises(fgpe.FgpeError, match="spin"):
        fgpe.derive(inp)


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "fgpe.h")).read()
    declared = set(re.findall(r"\b(fgpe_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(fgpe.EXPORTS)
    lib = fgpe.lib()
    for name in declared:
        assert hasattr(lib, name), name


def test_create_fails_loudly_without_device():
    """No silent CPU path: with no HIP device fgpe_create reports an error."""
    hip = C.CDLL("libamdhip64.so")
    n = C.c_int(0)
    rc = hip.hipGetDeviceCount(C.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a HIP device is present")
    inp = case(3, 2, (32, 32, 32), (0.1,) * 3, 1e-3, (0.5,) * 3)
    p, _ = fgpe.derive(inp)
    with pytest.raises(fgpe.FgpeError, match="device"):
        fgpe.Solver(p)


def test_unsupported_grid_is_rejected_before_allocation():
    inp = case(3, 2, (48, 64, 64), (0.1,) * 3, 1e-3, (0.5,) * 3)
    p, _ = fgpe.derive(inp)
    with pytest.raises(fgpe.FgpeError):
        fgpe.Solver(p)
