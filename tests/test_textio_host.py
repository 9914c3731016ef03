"""cgpe.x's text files through the library (fgpe_write_dump,
fgpe_read_initial_sol, fgpe_format_f) against the Fortran runtime, on the CPU.

tests/fortran/dumps.f90 holds the reference's dump loops and FORMATs and its
list-directed initial_sol READ; it is compiled with amdflang here.  From the
same PHI and grids, every dump kind written by the library must be byte for
byte the Fortran program's file, and the state the library reads from an
initial_sol file must be bit for bit the one the Fortran READ forms.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from fortress_amd import fgpe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "fortran", "dumps.f90")
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

needs_flang = pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not available")

KINDS = {"1d_tmp": 1, "1d_sol": 2, "2d": 3, "3d": 4, "3d_xy": 5, "3d_xz": 6}


@pytest.fixture(scope="module")
def fortran_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("ftn")
    exe = str(d / "dumps")
    subprocess.run([FLANG, "-O2", SRC, "-o", exe], check=True, cwd=str(d), capture_output=True)
    return exe


def lib():
    L = fgpe.lib()
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.fgpe_write_dump.argtypes = [C.c_char_p, C.c_int32, ip, C.c_int32, dp, dp, dp, dp, C.c_int32]
    L.fgpe_read_initial_sol.argtypes = [C.c_char_p, C.c_int32, ip, C.c_int32, dp, dp, dp, C.c_int32]
    L.fgpe_format_f.argtypes = [C.c_double, C.c_int32, C.c_int32, C.c_char_p, C.c_size_t]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def state(dim, ms, n, seed):
    """A wave function with the values that stress the formats: ordinary
    ones, exact zeros (phase 0), negative zeros, a negative imaginary zero
    (phase -pi), tiny and large magnitudes."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    phi = (rng.standard_normal((ms, nz, ny, nx)) + 1j * rng.standard_normal((ms, nz, ny, nx))) * 0.3
    flat = phi.reshape(-1)
    flat[::97] = 0.0
    flat[5::89] = complex(-0.0, -0.0)
    flat[7::83] = complex(-0.4, -0.0)
    flat[11::79] *= 1e-6
    flat[13::71] *= 10.0  # densities up to ~100: F10.5 fields stay separated for the READ
    grids = [np.array([(-1.0 + 2.0 * i / m) * m * 0.1 / 2.0 for i in range(m)]) for m in n]
    return phi, grids


def write_stream(path, dim, ms, n, phi, grids):
    with open(path, "wb") as f:
        f.write(np.array([dim, ms, *n], dtype=np.int32).tobytes())
        # PHI(NX,NY,NZ,MS) in Fortran order is [m][z][y][x] in C order
        f.write(np.ascontiguousarray(phi).view(np.float64).tobytes())
        for g in grids:
            f.write(g.astype(np.float64).tobytes())


CASES = [(1, 5, (256, 1, 1)), (1, 3, (96, 1, 1)), (2, 5, (48, 40, 1)), (2, 3, (32, 64, 1)), (3, 5, (16, 12, 10)),
         (3, 3, (8, 16, 12))]


@needs_flang
@pytest.mark.parametrize("dim,ms,n", CASES)
def test_dumps_and_reader_match_fortran(tmp_path, fortran_exe, dim, ms, n):
    L = lib()
    phi, grids = state(dim, ms, n, seed=dim * 10 + ms)
    write_stream(tmp_path / "in.bin", dim, ms, n, phi, grids)
    # the library's dumps; its final-solution dump then serves as the
    # initial_sol that both readers take (the fields cgpe.x reads)
    sol = {1: "f_1d_sol.dat", 2: "f_2d.dat", 3: "f_3d.dat"}[dim]
    nn = np.array(n, dtype=np.int32)
    a = np.ascontiguousarray(phi).view(np.float64)
    gx, gy, gz = (np.ascontiguousarray(g) for g in grids)
    kinds = {1: ["1d_tmp", "1d_sol"], 2: ["2d"], 3: ["3d", "3d_xy", "3d_xz"]}[dim]
    for k in kinds:
        rc = L.fgpe_write_dump(str(tmp_path / ("c_%s.dat" % k)).encode(), KINDS[k],
                               nn.ctypes.data_as(C.POINTER(C.c_int32)), ms, _dp(a), _dp(gx), _dp(gy), _dp(gz), 3)
        assert rc == 0, fgpe.last_error()
    shutil.copy(tmp_path / ("c_%s.dat" % {1: "1d_sol", 2: "2d", 3: "3d"}[dim]), tmp_path / "sol_in.dat")
    subprocess.run([fortran_exe, str(tmp_path / "in.bin"), str(tmp_path / "sol_in.dat"), str(tmp_path / "out.bin")],
                   check=True, cwd=str(tmp_path), capture_output=True)
    for k in kinds:
        fname = {"1d_tmp": "f_1d_tmp.dat", "1d_sol": "f_1d_sol.dat", "2d": "f_2d.dat", "3d": "f_3d.dat",
                 "3d_xy": "f_3d_xy.dat", "3d_xz": "f_3d_xz.dat"}[k]
        want = open(tmp_path / fname, "rb").read()
        got = open(tmp_path / ("c_%s.dat" % k), "rb").read()
        assert got == want, k
    assert open(tmp_path / "sol_in.dat", "rb").read() == open(tmp_path / sol, "rb").read()
    # the reader against the Fortran READ of the same file
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    npts = n[0] * n[1] * n[2]
    f_phi = raw[:2 * npts * ms]
    f_x = raw[2 * npts * ms:2 * npts * ms + n[0]]
    f_y = raw[2 * npts * ms + n[0]:2 * npts * ms + n[0] + n[1]]
    got = np.zeros(2 * npts * ms)
    x = np.zeros(n[0])
    y = np.zeros(n[1])
    rc = L.fgpe_read_initial_sol(str(tmp_path / "sol_in.dat").encode(), dim, nn.ctypes.data_as(C.POINTER(C.c_int32)),
                                 ms, _dp(got), _dp(x), _dp(y), 4)
    assert rc == 0, fgpe.last_error()
    assert np.array_equal(got.view(np.uint64), f_phi.view(np.uint64))
    if dim <= 2:
        assert np.array_equal(x, f_x)
    if dim == 2:
        assert np.array_equal(y, f_y)


def fmt(v, w, d):
    buf = C.create_string_buffer(64)
    assert lib().fgpe_format_f(v, w, d, buf, 64) == 0
    return buf.value.decode()


# (value, w, d, what amdflang's runtime writes): ties to even on the exact
# binary value, signed zeros, the optional leading zero, overflow, Inf/NaN
FORTRAN_FIELDS = [
    (0.5, 16, 10, "    0.5000000000"), (-1e-12, 16, 10, "   -0.0000000000"), (-0.0, 16, 10, "   -0.0000000000"),
    (0.00048828125, 16, 10, "    0.0004882812"), (0.00146484375, 16, 10, "    0.0014648438"),
    (123456.123456789, 16, 10, "****************"), (-1.25e-10, 16, 10, "   -0.0000000001"),
    (1.5e-10, 16, 10, "    0.0000000001"), (2.5e-10, 16, 10, "    0.0000000003"),
    (3.5e-10, 16, 10, "    0.0000000003"), (0.125, 16, 10, "    0.1250000000"), (1.0e6, 14, 6, "1000000.000000"),
    (-99999.9999999999, 18, 10, " -99999.9999999999"), (-0.5, 4, 2, "-.50"), (0.5, 4, 2, "0.50"), (0.5, 3, 2, ".50"),
    (-9.999, 5, 2, "*****"), (-999.99999999999, 16, 10, "-1000.0000000000"), (99999.99999999999, 16, 10, "*" * 16),
    (float("inf"), 10, 5, "       Inf"), (float("-inf"), 10, 5, "      -Inf"), (float("nan"), 10, 5, "       NaN"),
]


@pytest.mark.parametrize("v,w,d,want", FORTRAN_FIELDS)
def test_format_f_fields(v, w, d, want):
    assert fmt(v, w, d) == want


def test_format_f_matches_exact_decimal_rounding():
    """Against Python's exact decimal expansion of the binary value, rounded
    half-even, on many values (what the Fortran runtime's F editing does)."""
    from decimal import ROUND_HALF_EVEN, Decimal
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-12, 5, 3000),
                           np.arange(-2000, 2000) / 2048.0, np.arange(0, 3000) * 1e-10])
    for v in vals:
        for w, d in ((16, 10), (14, 6), (10, 5), (8, 2), (18, 10), (14, 8)):
            q = Decimal(float(v)).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
            s = format(q, "f")
            if s.startswith("-") or (v < 0 or (v == 0 and np.signbit(v))):
                s = "-" + s.lstrip("-")
            if len(s) > w and s.lstrip("-").startswith("0."):
                s = s.replace("0.", ".", 1)
            want = "*" * w if len(s) > w else s.rjust(w)
            assert fmt(float(v), w, d) == want, (v, w, d)
