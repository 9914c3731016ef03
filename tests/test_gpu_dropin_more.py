"""cgpe.x drop-in paths beyond the imaginary-time rows.

* The reference's own INPUT/input2D.dat (256^2 spin-2, 23Na, SOC 0.5) and
  INPUT/input3D.dat (128^3 spin-2, SOC 0.5 on every axis), copied to
  tests/golden/, run through cgpe.x with only the run length cut (2D: NITER
  200; 3D: NITER 20 with STP 10, which the FP64 oracle covers in under a
  minute); every file2/file3 row and the convergence.dat rows against the
  oracle advanced through the same events.
* Real time from ./INPUT/initial_sol_spinS.dat (SRC/cgpe3d.f90:383-410,
  cgpe2d.f90:378-400, cgpe1d.f90:372-383): the test writes the file in the
  reference's row order (3D: i, j, k with k fastest; 2D: x, y rows with a
  blank line after each x; 1D: x rows), cgpe.x reads it, and the *_re_* rows,
  the NSTP dumps (tmp_solution_file, and in 3D its _xy/_xz slices) and the
  final solution_file are compared with the oracle started from the same
  parsed values.
* Newton-Raphson failure (SRC/cgpe3d.f90:1556-1559): spin-2 without SO
  coupling at MAG = 2.5 stops with the reference's message in err_msg.dat
  and no row after t = 0.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from cases import noisy
from fortress_amd.abi import make_input
from fortress_amd.inputs import read_input, write_input

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "fortress_amd", "host", "cgpe.x")


def rows(path):
    return np.array([[float(x) for x in line.split()] for line in open(path) if line.strip()])


def run_cgpe(tmp, dim, timeout=300):
    os.makedirs(os.path.join(tmp, "OUTPUT"), exist_ok=True)
    print("running cgpe.x %dD in %s" % (dim, tmp), flush=True)
    r = subprocess.run([EXE, "%dD" % dim], cwd=str(tmp), capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def oracle_rows(inp, phi0, imag):
    """file2/file3 rows (and convergence rows in imaginary time) of the event
    loop, from phi0 (already normalised in imaginary time)."""
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.set_phi(phi0)
    stp, niter, dt = inp.stp, inp.niter, inp.dt
    r2, r3, conv = [], [], []
    en_old = None
    for ev in range(niter // stp + 1):
        if ev:
            o.run(stp - 1)
            prev = o.get_phi()
            o.run(1)
            norms = o.last_norms()
        else:
            norms = o.normc()
        en, mu, mz = o.energy()
        t = 2.0 * ev * stp * dt
        if ev and imag:
            cv = np.abs(o.get_phi() - prev).max() / (stp * 2.0 * dt)
            conv.append([t, cv, abs(en - en_old) / (stp * 2.0 * dt)])
        r2.append([t, en / 2.0] + list(o.rad()))
        r3.append([t] + list(norms) + [norms.sum(), mz])
        en_old = en
    phi = o.get_phi()
    o.close()
    return np.array(r2), np.array(r3), np.array(conv), phi


@pytest.mark.timeout(600)
@pytest.mark.parametrize("dim,niter,stp", [(2, 200, 100), (3, 20, 10)])
def test_reference_inputs_run_capped(tmp_path, dim, niter, stp):
    name = "input%dD.dat" % dim
    os.makedirs(tmp_path / "INPUT")
    lines = open(os.path.join(GOLD, name)).read().splitlines()
    lines[1] = "%d    #NITER (cut for the test)" % niter
    lines[2] = lines[2].replace("100,", "%d," % stp, 1)
    open(tmp_path / "INPUT" / name, "w").write("\n".join(lines) + "\n")
    inp = read_input(str(tmp_path / "INPUT" / name), dim)
    assert inp.niter == niter and inp.stp == stp and inp.spin == 2 and inp.switch_im == 1
    r = run_cgpe(tmp_path, dim)
    assert "Program execution complete" in r.stdout
    f2 = rows(tmp_path / "OUTPUT" / "file2_im_spin2.dat")
    f3 = rows(tmp_path / "OUTPUT" / "file3_im_spin2.dat")
    cv = rows(tmp_path / "OUTPUT" / "convergence.dat")
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    phi0 = o.get_phi()
    o.close()
    print("oracle: %d steps" % niter, flush=True)
    r2, r3, rc, _ = oracle_rows(inp, phi0, 1)
    assert f2.shape == r2.shape and f3.shape == r3.shape and cv.shape == rc.shape
    e2, e3 = np.abs(f2 - r2).max(), np.abs(f3 - r3).max()
    print("%s: max |file2 - oracle| %.2e, |file3 - oracle| %.2e" % (name, e2, e3), flush=True)
    assert e2 <= 2e-10 and e3 <= 2e-10
    assert np.allclose(cv[:, 0], rc[:, 0], atol=1e-12)
    assert np.allclose(cv[:, 1:], rc[:, 1:], rtol=1e-7, atol=1e-12)


def write_initial_sol(path, dim, phi, x, y):
    """initial_sol in the reference's row order, read list-directed by cgpe.x;
    phi is [x][y][z][m].  17 significant digits, and the state returned is
    the one cgpe.x forms from the printed values, sqrt(rho) exp(i theta)."""
    ms = phi.shape[-1]
    rho = repr_round(np.abs(phi) ** 2)
    th = repr_round(np.angle(phi))
    fmt = lambda v: " ".join("%.17g" % t for t in v)
    nx, ny, nz = phi.shape[:3]
    with open(path, "w") as f:
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    vals = [v for m in range(ms) for v in (rho[i, j, k, m], th[i, j, k, m])]
                    lead = [x[i]] if dim == 1 else [x[i], y[j]] if dim == 2 else []
                    f.write(fmt(lead + vals) + "\n")
            if dim == 2:
                f.write("\n")
    return np.sqrt(rho) * np.exp(1j * th)


def repr_round(a):
    return np.array([float("%.17g" % v) for v in a.ravel()]).reshape(a.shape)


def grid(n, d):
    return np.array([(-1.0 + 2.0 * i / n) * n * d / 2.0 for i in range(n)])


def to_xyzm(phi):
    """oracle layout (ms, nz, ny, nx) -> [x][y][z][m], the row order of the dumps"""
    return np.transpose(phi, (3, 2, 1, 0))


def from_xyzm(a):
    return np.transpose(a, (3, 2, 1, 0))


def dens_phase_close(got, ref_phi, dens_atol, phase_atol):
    ms = ref_phi.shape[-1]
    rho, th = got[:, 0::2][:, :ms], got[:, 1::2][:, :ms]
    rr = np.abs(ref_phi.reshape(-1, ms)) ** 2
    assert np.abs(rho - rr).max() <= dens_atol, np.abs(rho - rr).max()
    ok = rr > 1e-3
    d = np.angle(np.exp(1j * (th - np.angle(ref_phi.reshape(-1, ms)))))
    assert np.abs(d[ok]).max() <= phase_atol, np.abs(d[ok]).max()


RT_CASES = {
    # dim: (spin, n, d, dt, gamma, niter, stp, nstp)
    1: (1, (512,), (0.05,), 5e-4, (0.5,), 300, 100, 150),
    2: (2, (64, 96), (0.2, 0.15), 1e-3, (1.0, 0.5), 40, 10, 20),
    3: (2, (32, 32, 64), (0.4, 0.4, 0.2), 5e-4, (0.5, 0.3, 0.7), 20, 10, 20),
}


@pytest.mark.timeout(300)
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_real_time_from_initial_sol(tmp_path, dim):
    spin, n, d, dt, gamma, niter, stp, nstp = RT_CASES[dim]
    ms = 2 * spin + 1
    a = (87.93, 91.28, 99.18) if spin == 2 else (101.8, 100.4)
    nu = (20.0, 400.0, 400.0) if dim == 1 else (20.0, 20.0, 400.0) if dim == 2 else (20.0, 20.0, 20.0)
    # 2D: N = 3e4 keeps the noisy start well conditioned (at 1e5 a 1e-7
    # perturbation grows 60x in 20 steps)
    natoms = {1: 10000, 2: 30000, 3: 100000}[dim]
    inp = make_input(dim, spin, n, d, dt, 87.0, a, natoms, nu, gamma, switch_im=0, niter=niter, stp=stp, nstp=nstp)
    os.makedirs(tmp_path / "INPUT")
    write_input(str(tmp_path / "INPUT" / ("input%dD.dat" % dim)), inp)
    # a start state: the Gaussian guess, normalised, plus seeded 1% noise
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    phi = noisy(o.get_phi(), seed=50 + dim)
    o.close()
    x = grid(n[0], d[0])
    y = grid(n[1], d[1]) if dim >= 2 else None
    phi_x = write_initial_sol(str(tmp_path / "INPUT" / ("initial_sol_spin%d.dat" % spin)), dim, to_xyzm(phi), x, y)
    phi0 = from_xyzm(phi_x)
    r = run_cgpe(tmp_path, dim)
    assert "Program execution complete" in r.stdout
    out = tmp_path / "OUTPUT"
    f2 = rows(out / ("file2_re_spin%d.dat" % spin))
    f3 = rows(out / ("file3_re_spin%d.dat" % spin))
    r2, r3, _, phi_end = oracle_rows(inp, phi0, 0)
    assert f2.shape == r2.shape and f3.shape == r3.shape
    # 10 printed decimals; energies above 1 compared relative
    e2 = (np.abs(f2 - r2) / np.maximum(1.0, np.abs(r2))).max()
    e3 = (np.abs(f3 - r3) / np.maximum(1.0, np.abs(r3))).max()
    print("%dD real time: max |file2 - oracle| %.2e, |file3 - oracle| %.2e" % (dim, e2, e3), flush=True)
    assert e2 <= 2e-10 and e3 <= 2e-10
    # the final solution and the last NSTP dump (taken at NITER) in their
    # reference formats: 1D (F11.5,10F18.10) / (F11.5,10F14.6), 2D
    # (2F8.2,10F10.5) with a blank line after each x, 3D (10F16.10)
    sol = rows(out / ("solution_file_re_spin%d.dat" % spin))
    tmp = rows(out / ("tmp_solution_file_spin%d.dat" % spin))
    ref = to_xyzm(phi_end)
    ncoord = {1: 1, 2: 2, 3: 0}[dim]
    assert sol.shape == (int(np.prod(n)), ncoord + 2 * ms) and tmp.shape == sol.shape
    tol_sol = {1: 1e-9, 2: 1e-5, 3: 1e-9}[dim]
    tol_tmp = {1: 1e-6, 2: 1e-5, 3: 1e-9}[dim]
    dens_phase_close(sol[:, ncoord:], ref, tol_sol, 1e-3 if dim == 2 else 1e-6)
    dens_phase_close(tmp[:, ncoord:], ref, tol_tmp, 1e-3 if dim != 3 else 1e-6)
    if dim == 2:
        text = open(out / ("solution_file_re_spin%d.dat" % spin)).read().split("\n")
        # n[1] rows then one blank line per x
        assert all(text[(k + 1) * (n[1] + 1) - 1].strip() == "" for k in range(n[0]))
        assert len(text[0]) == 2 * 8 + 2 * ms * 10
    if dim == 3:
        for sl, shape in (("_xy", (n[0], n[1])), ("_xz", (n[0], n[2]))):
            s = rows(out / ("tmp_solution_file_spin%d%s.dat" % (spin, sl)))
            assert s.shape == (shape[0] * shape[1], 2 + 2 * ms)
            cut = ref[:, :, n[2] // 2, :] if sl == "_xy" else ref[:, n[1] // 2, :, :]
            dens_phase_close(s[:, 2:], cut, 1e-8, 1e-5)


@pytest.mark.timeout(120)
def test_newton_failure_stops_with_message(tmp_path):
    inp = make_input(1, 2, (256,), (0.05,), 1.25e-4, 23.0, (34.9, 45.8, 64.5), 10000, (20.0, 400.0, 400.0), (0.0,),
                     switch_im=1, niter=100, stp=10, mag=2.5)
    os.makedirs(tmp_path / "INPUT")
    write_input(str(tmp_path / "INPUT" / "input1D.dat"), inp)
    r = run_cgpe(tmp_path, 1)
    assert "Program execution complete" not in r.stdout
    msg = open(tmp_path / "err_msg.dat").read()
    assert "Newton-Raphason subroutine could not calculate the roots within" in msg, msg
    assert "tolerance" in msg
    assert rows(tmp_path / "OUTPUT" / "file2_im_spin2.dat").shape[0] == 1
    assert rows(tmp_path / "OUTPUT" / "file3_im_spin2.dat").shape[0] == 1
