"""cgpe.x in 2D and 3D against the FP64 oracle.

fortress_amd/host/cgpe.x (FP64 propagation) is run on generated
INPUT/input2D.dat / input3D.dat files (imaginary time, TOL below reach so the
run goes to NITER), and every row of OUTPUT/file2 (time, EN/2, RMS per
component) and OUTPUT/file3 (time, NORM per component, SUM(NORM), MZ) is
compared with the oracle advanced through the same events: the initial row
after NORMT, then one row every STP steps with the NORMC of the event's last
step (rows written at SRC/cgpe2d.f90:437-438,513-514 and
SRC/cgpe3d.f90:448-449,526-527).
The files print 10 decimals, so the bar is 2e-10 absolute.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from fortress_amd.abi import make_input
from fortress_amd.inputs import write_input

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortress_amd", "host", "cgpe.x")

CASES = {
    # name: (dim, spin, n, d, dt, mass, a, natoms, nu, gamma, niter, stp)
    "2D": (2, 1, (64, 64), (0.1, 0.1), 1e-3, 87.0, (101.8, 100.4), 10000, (20.0, 20.0, 400.0), (0.5, 0.5), 300, 100),
    "3D": (3, 2, (32, 32, 32), (0.4, 0.4, 0.4), 5e-4, 87.0, (87.93, 91.28, 99.18), 100000, (20.0, 20.0, 20.0),
           (0.5, 0.5, 0.5), 200, 100),
}


def rows(path):
    return np.array([[float(x) for x in line.split()] for line in open(path) if line.strip()])


@pytest.mark.parametrize("name", ["2D", "3D"])
def test_cgpe_rows_match_oracle(tmp_path, name):
    dim, spin, n, d, dt, mass, a, natoms, nu, gamma, niter, stp = CASES[name]
    ms = 2 * spin + 1
    inp = make_input(dim, spin, n, d, dt, mass, a, natoms, nu, gamma, switch_im=1, niter=niter, stp=stp,
                     nstp=10 ** 8, tol=1e-30)
    os.makedirs(tmp_path / "INPUT")
    os.makedirs(tmp_path / "OUTPUT")
    write_input(str(tmp_path / "INPUT" / ("input%dD.dat" % dim)), inp)
    print("running cgpe.x", name, flush=True)
    r = subprocess.run([EXE, name], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Program execution complete" in r.stdout, r.stdout + r.stderr
    f2 = rows(tmp_path / "OUTPUT" / ("file2_im_spin%d.dat" % spin))
    f3 = rows(tmp_path / "OUTPUT" / ("file3_im_spin%d.dat" % spin))
    assert f2.shape == (niter // stp + 1, 2 + ms) and f3.shape == (niter // stp + 1, 3 + ms)

    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    ref2, ref3 = [], []
    for ev in range(niter // stp + 1):
        if ev:
            o.run(stp)
            norms = o.last_norms()
        else:
            norms = o.normc()
        en, mu, mz = o.energy()
        t = 2.0 * ev * stp * dt
        ref2.append([t, en / 2.0] + list(o.rad()))
        ref3.append([t] + list(norms) + [norms.sum(), mz])
    o.close()
    err2 = np.abs(f2 - np.array(ref2)).max()
    err3 = np.abs(f3 - np.array(ref3)).max()
    print(name, "max |file2 - oracle|", err2, "max |file3 - oracle|", err3, flush=True)
    assert err2 <= 2e-10 and err3 <= 2e-10
