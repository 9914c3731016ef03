"""The err_msg.dat channel of cgpe.x, on the CPU.

Both cases stop before the device is touched, so they run without a GPU:

* no ./INPUT/inputND.dat: err_msg.dat and stdout say "'inputND.dat' does not
  exist." and the program stops (SRC/cgpe3d.f90:221-232, SRC/cgpe2d.f90 and
  SRC/cgpe1d.f90 name their own file);
* real time without ./INPUT/initial_sol_spinS.dat: the two lines of
  SRC/cgpe3d.f90:383-390, after the file1 header (written at 346-379) and
  before any row.

The third path, Newton-Raphson failure (SRC/cgpe3d.f90:1556-1559), needs the
time loop: tests/test_gpu_dropin_more.py runs it on the GPU, and the oracle's
own failure on the same input is checked here.
"""
import os
import subprocess

import pytest

import oracle_ctypes as O
from cases import NA_SPIN2, case
from fortress_amd.abi import make_input
from fortress_amd.inputs import write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortress_amd", "host", "cgpe.x")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="cgpe.x is not built (run __graft_entry__.build())")


def _run(tmp, dim):
    os.makedirs(tmp / "INPUT", exist_ok=True)
    os.makedirs(tmp / "OUTPUT", exist_ok=True)
    return subprocess.run([EXE, "%dD" % dim], cwd=str(tmp), capture_output=True, text=True, timeout=60)


def _lines(path):
    return [l.strip() for l in open(path).read().splitlines() if l.strip()]


@needs_exe
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_missing_input_file(tmp_path, dim):
    r = _run(tmp_path, dim)
    assert r.returncode == 0, r.stderr
    msg = "'input%dD.dat' does not exist." % dim
    assert _lines(tmp_path / "err_msg.dat") == [msg]
    assert msg in r.stdout
    assert not os.listdir(tmp_path / "OUTPUT")


@needs_exe
@pytest.mark.parametrize("dim,spin", [(1, 2), (2, 1), (3, 2)])
def test_missing_initial_solution_in_real_time(tmp_path, dim, spin):
    n = {1: (256,), 2: (64, 64), 3: (32, 32, 32)}[dim]
    inp = make_input(dim, spin, n, (0.1,) * dim, 1e-3, 87.0, (87.93, 91.28, 99.18) if spin == 2 else (101.8, 100.4),
                     10000, (20.0, 20.0, 20.0), (0.5,) * dim, switch_im=0, niter=10, stp=5)
    os.makedirs(tmp_path / "INPUT")
    write_input(str(tmp_path / "INPUT" / ("input%dD.dat" % dim)), inp)
    r = _run(tmp_path, dim)
    assert r.returncode == 0, r.stderr
    want = ['Check if the "initial_sol_spin%d.dat" is available' % spin,
            'if not, user can read the "solution_file_im_spin%d.dat" file as an initial solution.' % spin]
    assert _lines(tmp_path / "err_msg.dat") == want
    for w in want:
        assert w in r.stdout
    # the header of file1 is written, no row of file2/file3
    f1 = open(tmp_path / "OUTPUT" / ("file1_re_spin%d.dat" % spin)).read()
    assert "No. of space steps" in f1 and "Initial" not in f1
    assert _lines(tmp_path / "OUTPUT" / ("file2_re_spin%d.dat" % spin)) == []
    assert _lines(tmp_path / "OUTPUT" / ("file3_re_spin%d.dat" % spin)) == []


def test_oracle_newton_fails_for_unreachable_magnetisation():
    """MAG = 2.5 > F = 2 has no projection: Newton-Raphson (NTRIAL = 1000,
    TOL_NR = 1e-6, SRC/cgpe3d.f90:1516-1560) ends with a finite step above the
    tolerance, which the reference reports through err_msg.dat; MAG = 1.9
    converges."""
    for mag, fails in ((2.5, True), (1.9, False)):
        inp = case(1, 2, (256,), (0.05,), 1.25e-4, (0.0,), mag=mag, species=NA_SPIN2)
        p, _ = O.derive(inp)
        o = O.Oracle(p)
        o.initial_guess()
        o.normt()
        if fails:
            with pytest.raises(RuntimeError):
                o.run(1)
        else:
            o.run(3)
        o.close()
