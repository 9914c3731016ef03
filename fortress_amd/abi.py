"""ctypes mirrors of the structs declared in include/fgpe.h."""
import ctypes as C


class FgpeInput(C.Structure):
    """One field per record of INPUT/input{1,2,3}D.dat (SRC/cgpe3d.f90:239-266)."""

    _fields_ = [
        ("dim", C.c_int32),
        ("spin", C.c_int32),
        ("niter", C.c_int32),
        ("stp", C.c_int32),
        ("nstp", C.c_int32),
        ("threads", C.c_int32),
        ("n", C.c_int32 * 3),
        ("d", C.c_double * 3),
        ("dt", C.c_double),
        ("mass", C.c_double),
        ("a", C.c_double * 3),
        ("natoms", C.c_int32),
        ("nu", C.c_double * 3),
        ("tol", C.c_double),
        ("switch_im", C.c_int32),
        ("gamma", C.c_double * 3),
        ("option_fpc", C.c_int32),
        ("mag", C.c_double),
    ]


class FgpeParams(C.Structure):
    """Dimensionless solver parameters consumed by fgpe_create."""

    _fields_ = [
        ("dim", C.c_int32),
        ("spin", C.c_int32),
        ("n", C.c_int32 * 3),
        ("imag", C.c_int32),
        ("precision", C.c_int32),
        ("option_fpc", C.c_int32),
        ("soc", C.c_int32),
        ("d", C.c_double * 3),
        ("dt", C.c_double),
        ("tau", C.c_double * 3),
        ("alpha", C.c_double * 3),
        ("gamma", C.c_double * 3),
        ("mag", C.c_double),
    ]


FGPE_FP64 = 0
FGPE_FP32 = 1


def make_input(dim, spin, n, d, dt, mass, a, natoms, nu, gamma, switch_im=1, niter=1000, stp=100,
               nstp=10**9, tol=1e-9, option_fpc=4, mag=0.0, threads=1):
    inp = FgpeInput()
    inp.dim, inp.spin, inp.niter, inp.stp, inp.nstp, inp.threads = dim, spin, niter, stp, nstp, threads
    n = list(n) + [1] * (3 - len(n))
    d = list(d) + [1.0] * (3 - len(d))
    a = list(a) + [0.0] * (3 - len(a))
    g = list(gamma) + [0.0] * (3 - len(gamma))
    for i in range(3):
        inp.n[i], inp.d[i], inp.a[i], inp.gamma[i], inp.nu[i] = n[i], d[i], a[i], g[i], nu[i]
    inp.dt, inp.mass, inp.natoms, inp.tol = dt, mass, int(natoms), tol
    inp.switch_im, inp.option_fpc, inp.mag = switch_im, option_fpc, mag
    return inp
