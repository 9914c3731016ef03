"""Binary checkpoint/resume on the GPU (fgpe_checkpoint_*; no counterpart in
the reference, which restarts only real time from text, SRC/cgpe3d.f90:383-410).

* Library: run K steps, save, run K more; a fresh handle loads the file and
  runs K steps: the states, last NORMC and observables are bit-identical --
  FP64 q2D spin-2 without SO coupling in imaginary time (the Newton state is
  carried), FP32 3D spin-2 real time, FP64 q1D spin-1 imaginary time.
* A handle with other parameters (DT, grid, precision) refuses the file.
* cgpe.x: a run stopped at II = 200 with FGPE_CHECKPOINT, resumed with
  FGPE_RESUME to NITER = 400, writes the same file2/file3/convergence rows
  after 200 and the same final solution, byte for byte, as a run never
  stopped.
"""
import os
import subprocess

import numpy as np
import pytest

from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input
from fortress_amd.inputs import write_input

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortress_amd", "host", "cgpe.x")

CASES = {
    "2d_s2_newton_fp64": (make_input(2, 2, (64, 64), (0.1, 0.1), 5e-4, 87.0, (87.93, 91.28, 99.18), 100000,
                                     (20.0, 20.0, 400.0), (0.0, 0.0), switch_im=1, mag=0.6), FGPE_FP64, 60),
    "3d_s2_real_fp32": (make_input(3, 2, (32, 64, 32), (0.4, 0.2, 0.4), 5e-4, 87.0, (87.93, 91.28, 99.18), 100000,
                                   (20.0, 20.0, 20.0), (0.5, 0.3, 0.7), switch_im=0), FGPE_FP32, 20),
    "1d_s1_imag_fp64": (make_input(1, 1, (512,), (0.05,), 5e-4, 87.0, (101.8, 100.4), 10000, (20.0, 400.0, 400.0),
                                   (0.5,), switch_im=1), FGPE_FP64, 300),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_resume_is_bit_identical(tmp_path, name):
    inp, prec, k = CASES[name]
    p, _ = fgpe.derive(inp, prec)
    path = str(tmp_path / (name + ".ckpt"))
    with fgpe.Solver(p) as a:
        a.initial_guess()
        a.normalize()
        a.run(k)
        a.checkpoint_save(path, ii=k, en_old=1.25, user=(7.0, 8.0))
        a.run(k)
        ref, ref_n, ref_obs = a.download(), a.last_norms(), a.observables()
    info = fgpe.checkpoint_info(path)
    assert info["ii"] == k and info["dim"] == inp.dim and info["precision"] == prec
    with fgpe.Solver(p) as b:
        ii, en_old, user = b.checkpoint_load(path)
        assert ii == k and en_old == 1.25 and user[:2] == (7.0, 8.0)
        b.run(k)
        got, got_n, got_obs = b.download(), b.last_norms(), b.observables()
    print(name, "max |resumed - uninterrupted|", np.abs(got - ref).max(), flush=True)
    assert np.array_equal(got, ref)
    assert np.array_equal(got_n, ref_n)
    assert got_obs[0] == ref_obs[0] and np.array_equal(got_obs[1], ref_obs[1])


def test_load_refuses_other_parameters(tmp_path):
    inp, prec, _ = CASES["2d_s2_newton_fp64"]
    p, _ = fgpe.derive(inp, prec)
    path = str(tmp_path / "x.ckpt")
    with fgpe.Solver(p) as a:
        a.initial_guess()
        a.checkpoint_save(path)
    others = [
        make_input(2, 2, (64, 64), (0.1, 0.1), 4e-4, 87.0, (87.93, 91.28, 99.18), 100000, (20.0, 20.0, 400.0),
                   (0.0, 0.0), switch_im=1, mag=0.6),
        make_input(2, 2, (64, 32), (0.1, 0.1), 5e-4, 87.0, (87.93, 91.28, 99.18), 100000, (20.0, 20.0, 400.0),
                   (0.0, 0.0), switch_im=1, mag=0.6),
    ]
    for other, pr in ((others[0], prec), (others[1], prec), (inp, FGPE_FP32)):
        q, _ = fgpe.derive(other, pr)
        with fgpe.Solver(q) as b:
            with pytest.raises(fgpe.FgpeError):
                b.checkpoint_load(path)


def run_cgpe(tmp, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    print("cgpe.x", env_extra, flush=True)
    r = subprocess.run([EXE, "2D"], cwd=str(tmp), capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "Program execution complete" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("imag", [1, 0])
def test_cgpe_resume_matches_uninterrupted_run(tmp_path, imag):
    inp = make_input(2, 2, (64, 64), (0.1, 0.1), 5e-4, 87.0, (87.93, 91.28, 99.18), 100000, (20.0, 20.0, 400.0),
                     (0.0, 0.0) if imag else (1.0, 0.5), switch_im=imag, niter=400, stp=100, nstp=10 ** 8, tol=1e-30,
                     mag=0.6 if imag else 0.0)
    tag = "im" if imag else "re"
    dirs = {}
    for name, niter, env in (("full", 400, {"FGPE_CHECKPOINT": "ck_full.bin", "FGPE_CHECKPOINT_EVERY": "200"}),
                             ("first", 200, {"FGPE_CHECKPOINT": "ck.bin", "FGPE_CHECKPOINT_EVERY": "200"}),
                             ("resumed", 400, {"FGPE_CHECKPOINT": "ck2.bin", "FGPE_CHECKPOINT_EVERY": "200"})):
        d = tmp_path / name
        os.makedirs(d / "INPUT")
        os.makedirs(d / "OUTPUT")
        inp.niter = niter
        write_input(str(d / "INPUT" / "input2D.dat"), inp)
        if not imag:
            # real time starts from an initial_sol: the Gaussian guess written as density and phase
            rows = []
            x = [(-1.0 + 2.0 * i / 64) * 6.4 / 2.0 for i in range(64)]
            for i in range(64):
                for j in range(64):
                    r2 = x[i] ** 2 + x[j] ** 2
                    rho = np.exp(-r2) / np.pi * 0.2
                    rows.append("%.17g %.17g " % (x[i], x[j]) + " ".join("%.17g 0.3" % rho for _ in range(5)))
                rows.append("")
            open(d / "INPUT" / "initial_sol_spin2.dat", "w").write("\n".join(rows) + "\n")
        if name == "resumed":
            os.rename(tmp_path / "first" / "ck.bin", d / "ck.bin")
            env = dict(env, FGPE_RESUME="ck.bin")
        run_cgpe(d, env)
        dirs[name] = d / "OUTPUT"
    files = ["file2_%s_spin2.dat" % tag, "file3_%s_spin2.dat" % tag] + (["convergence.dat"] if imag else [])
    for f in files:
        full = open(dirs["full"] / f).read().splitlines()
        res = open(dirs["resumed"] / f).read().splitlines()
        # the resumed run's rows are the full run's rows after II = 200
        assert res == full[len(full) - len(res):], f
        assert len(res) == 2
    for f in ("solution_file_%s_spin2.dat" % tag,):
        assert open(dirs["full"] / f).read() == open(dirs["resumed"] / f).read()
