"""End-to-end drop-in check: cgpe.x on the reference's own run.

fortress_amd/host/cgpe.x (Fortran host over libfgpe, FP64 propagation) is run
in a scratch directory holding INPUT/input1D.dat of the reference's committed
run (1D spin-2 SOC, imaginary time, NX=2048, 399,200 steps to TOL=1e-6), and
every file of its OUTPUT/ is compared with the reference's OUTPUT/ (copied to
tests/golden/): the header and rows of file1, all rows of file2/file3 and
convergence.dat (list-directed there, so compared numerically), the exit
iteration, and the densities/phases of solution_file and the last
tmp_solution_file dump.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "fortress_amd", "host", "cgpe.x")


def _run(tmp, dim, inputs, extra=()):
    os.makedirs(os.path.join(tmp, "INPUT"), exist_ok=True)
    os.makedirs(os.path.join(tmp, "OUTPUT"), exist_ok=True)
    for name, src in inputs.items():
        shutil.copy(src, os.path.join(tmp, "INPUT", name))
    print("running cgpe.x %s in %s" % (dim, tmp), flush=True)
    r = subprocess.run([EXE, dim] + list(extra), cwd=tmp, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Program execution complete" in r.stdout, r.stdout + r.stderr
    return os.path.join(tmp, "OUTPUT")


@pytest.fixture(scope="module")
def golden_run(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("cgpe1d"))
    return _run(tmp, "1D", {"input1D.dat": os.path.join(GOLD, "input1D.dat")})


def test_file1_header_and_rows(golden_run):
    got = open(os.path.join(golden_run, "file1_im_spin2.dat")).read().splitlines()
    ref = open(os.path.join(GOLD, "file1_im_spin2.dat")).read().splitlines()
    # identical layout: lines 2-18 (header, rulers) byte for byte
    assert got[1:18] == ref[1:18]
    assert len(got) == len(ref)
    # Initial, 40 NSTP rows and the final row: same labels, numbers to print precision
    for g, r in zip(got[18:61], ref[18:61]):
        if set(r.strip()) <= {"-"}:
            assert g == r
            continue
        head = r.split(":")[0]
        assert g.split(":")[0] == head
        gv = np.array(g.split(":")[1].split(), float)
        rv = np.array(r.split(":")[1].split(), float)
        assert np.allclose(gv, rv, atol=2e-5), (g, r)
    # the run stops at the same STP bin: 'After   399199 iter.:'
    assert got[59].startswith("After   399199 iter.:")


def test_file2_file3_rows(golden_run):
    for name in ("file2_im_spin2.dat", "file3_im_spin2.dat"):
        g = np.loadtxt(os.path.join(golden_run, name))
        r = np.loadtxt(os.path.join(GOLD, name))
        assert g.shape == r.shape, name
        assert np.allclose(g[:, 0], r[:, 0], atol=1e-9)
        assert np.allclose(g[:, 1:], r[:, 1:], atol=2e-9, rtol=0), (name, np.abs(g - r).max())


def test_convergence_rows(golden_run):
    g = np.loadtxt(os.path.join(golden_run, "convergence.dat"))
    r = np.loadtxt(os.path.join(GOLD, "convergence.dat"))
    assert g.shape == r.shape
    assert np.allclose(g[:, 0], r[:, 0], atol=1e-9)
    assert np.allclose(g[:, 1], r[:, 1], rtol=1e-6)
    # CONV_EN is a difference of energies 400 steps apart: relative bar on
    # the energies (1e-13) over 2*STP*DT
    assert np.allclose(g[:, 2], r[:, 2], rtol=1e-5, atol=1e-9)


@pytest.mark.parametrize("name,dens_tol", [("solution_file_im_spin2.dat", 2e-9), ("tmp_solution_file_spin2.dat", 2e-6)])
def test_solution_dumps(golden_run, name, dens_tol):
    g = np.loadtxt(os.path.join(golden_run, name))
    r = np.loadtxt(os.path.join(GOLD, name))
    assert g.shape == r.shape == (2048, 11)
    assert np.allclose(g[:, 0], r[:, 0], atol=1e-5)
    dens_g, dens_r = g[:, 1::2], r[:, 1::2]
    assert np.allclose(dens_g, dens_r, atol=dens_tol)
    # phases are defined where the density is resolved by the format
    ph_g, ph_r = g[:, 2::2], r[:, 2::2]
    ok = dens_r > 1e-4
    d = np.angle(np.exp(1j * (ph_g - ph_r)))
    assert np.abs(d[ok]).max() < 1e-4
