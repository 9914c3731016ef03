"""Reader for INPUT/input{1,2,3}D.dat with the reference's record semantics.

Each READ(10,*) consumes one record and takes as many list-directed items as
it needs; the rest of the record (the '#...' annotations) is ignored.  The
OPTION_FPC record is read only without SO coupling and the MAG record only in
imaginary time without SO coupling (SRC/cgpe3d.f90:239-266; 2D
SRC/cgpe2d.f90:234-261; 1D SRC/cgpe1d.f90:233-260).
"""
import re

from .abi import FgpeInput

EPS = 1.0e-40


def _items(line):
    toks = [t for t in re.split(r"[,\s]+", line.strip()) if t]
    return toks


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def read_input(path, dim):
    with open(path) as f:
        recs = f.read().splitlines()
    it = iter(recs)

    def take(k):
        toks = _items(next(it))
        if len(toks) < k:
            raise ValueError("record has %d items, %d needed" % (len(toks), k))
        return toks[:k]

    inp = FgpeInput()
    inp.dim = dim
    inp.spin = int(take(1)[0])
    inp.niter = int(take(1)[0])
    stp, nstp = take(2)
    inp.stp, inp.nstp = int(stp), int(nstp)
    inp.threads = int(take(1)[0])
    ns = take(dim)
    for i in range(3):
        inp.n[i] = int(ns[i]) if i < dim else 1
    ds = [_num(t) for t in take(dim + 1)]
    for i in range(3):
        inp.d[i] = ds[i] if i < dim else 1.0
    inp.dt = ds[dim]
    inp.mass = _num(take(1)[0])
    a = [_num(t) for t in take(2 if inp.spin == 1 else 3)]
    for i in range(3):
        inp.a[i] = a[i] if i < len(a) else 0.0
    inp.natoms = int(take(1)[0])
    nu = [_num(t) for t in take(3)]
    for i in range(3):
        inp.nu[i] = nu[i]
    inp.tol = _num(take(1)[0])
    inp.switch_im = int(take(1)[0])
    g = [_num(t) for t in take(dim)]
    for i in range(3):
        inp.gamma[i] = g[i] if i < dim else 0.0
    soc = any(abs(x) > EPS for x in g)
    inp.option_fpc = 4
    if not soc:
        inp.option_fpc = int(take(1)[0])
    else:
        next(it)
    inp.mag = 0.0
    if inp.switch_im == 1 and not soc:
        inp.mag = _num(take(1)[0])
    return inp


def write_input(path, inp):
    """Write an input file the reference's reader accepts (used by tests and
    tools that generate sweeps)."""
    dim, spin = inp.dim, inp.spin
    fm = lambda x: ("%.17g" % x).replace("e", "D") if "e" in ("%.17g" % x) else ("%.17gD0" % x)
    lines = [
        "%d  #SPIN" % spin,
        "%d  #NITER" % inp.niter,
        "%d, %d  #STP, NSTP" % (inp.stp, inp.nstp),
        "%d  #NUMBER OF THREADS" % inp.threads,
        ", ".join(str(inp.n[i]) for i in range(dim)) + "  #N",
        ", ".join(fm(inp.d[i]) for i in range(dim)) + ", " + fm(inp.dt) + "  #D, DT",
        fm(inp.mass) + "  #MASS",
        ", ".join(fm(inp.a[i]) for i in range(2 if spin == 1 else 3)) + "  #A",
        "%d  #NATOMS" % inp.natoms,
        ", ".join(fm(inp.nu[i]) for i in range(3)) + "  #NU",
        fm(inp.tol) + "  #TOL",
        "%d  #SWITCH_IM" % inp.switch_im,
        ", ".join(fm(inp.gamma[i]) for i in range(dim)) + "  #GAMMA",
        "%d  #OPTION_FPC" % inp.option_fpc,
        fm(inp.mag) + "  #MAG",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
