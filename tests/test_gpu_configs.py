"""GPU runs at the workloads' own sizes.

* q2D spin-2 Rashba, 256^2, DX = DY = 0.1, DT = 0.001, 87Rb spin-2
  (A0/A2/A4 = 87.93/91.28/99.18 a_B), N = 1e5, trap 20/20/400 Hz,
  GAMMAX = GAMMAY = 1.0: FP64 against the FP64 oracle (1e-10) and the FP32
  densities against the oracle (1e-5), imaginary time from the Gaussian guess
  and real time from 200 imaginary-time steps plus seeded 1% noise.
* 3D spin-2, 256^3, DX = DY = DZ = 0.1, DT = 0.0005, GAMMA = 0.5 on every
  axis, N = 1e5, 20 Hz trap: FP64 against the oracle for 2 real-time steps
  (phi, norms and observables to 1e-10), then FP32 against FP64 over 10^4
  real-time steps from an imaginary-time relaxed state plus seeded 1% noise,
  the component densities within 1e-5 relative L2.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from cases import noisy, rel_l2
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input

pytestmark = pytest.mark.gpu


def q2d_spin2(imag):
    return make_input(2, 2, (256, 256), (0.1, 0.1), 0.001, 87.0, (87.93, 91.28, 99.18), 100000,
                      (20.0, 20.0, 400.0), (1.0, 1.0), switch_im=imag, tol=1e-9)


def d3_spin2(imag, dt=5e-4):
    return make_input(3, 2, (256, 256, 256), (0.1, 0.1, 0.1), dt, 87.0, (87.93, 91.28, 99.18), 100000,
                      (20.0, 20.0, 20.0), (0.5, 0.5, 0.5), switch_im=imag, stp=100)


def oracle_start(inp, seed):
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    phi = o.get_phi()
    o.close()
    return phi if inp.switch_im else noisy(phi, seed)


def relaxed_start(inp_im, steps, seed):
    """FP64 imaginary-time relaxation on the GPU from the Gaussian guess, plus
    seeded 1% noise: a real-time start whose trajectory is well conditioned
    (the bare Gaussian at N = 1e5 is far from equilibrium, and 20 real-time
    FP64 steps from it turn a 1e-14 perturbation into 1e-7)."""
    p, _ = fgpe.derive(inp_im, FGPE_FP64)
    with fgpe.Solver(p) as s:
        s.initial_guess()
        s.normalize()
        s.run(steps)
        phi = s.download()
    return noisy(phi, seed)


@pytest.mark.parametrize("imag", [1, 0])
def test_q2d_spin2_256(imag):
    inp = q2d_spin2(imag)
    phi0 = oracle_start(inp, seed=21) if imag else relaxed_start(q2d_spin2(1), 200, seed=21)
    nsteps = 20
    p_or, _ = O.derive(inp)
    o = O.Oracle(p_or)
    o.set_phi(phi0)
    o.run(nsteps)
    ref = o.get_phi()
    rn = o.last_norms()
    ren, rmu, rmz = o.energy()
    rrms = o.rad()
    o.close()
    for prec in (FGPE_FP64, FGPE_FP32):
        p, _ = fgpe.derive(inp, prec)
        with fgpe.Solver(p) as s:
            s.upload(phi0)
            s.run(nsteps)
            got = s.download()
            gn = s.last_norms()
            en, mu, mz, rms = s.observables()
        err = rel_l2(got, ref)
        dens = rel_l2(np.abs(got) ** 2, np.abs(ref) ** 2)
        print("q2D spin-2 256^2 imag=%d %s: phi %.3e densities %.3e EN %.12g (oracle %.12g)"
              % (imag, "fp64" if prec == FGPE_FP64 else "fp32", err, dens, en, ren), flush=True)
        bar = 1e-10 if prec == FGPE_FP64 else 1e-5
        if prec == FGPE_FP64:
            assert err < bar
        assert dens < bar
        assert np.allclose(gn, rn, rtol=bar, atol=1e-14)
        assert abs(en - ren) <= bar * abs(ren)
        assert np.allclose(mu, rmu, rtol=bar, atol=1e-12)
        assert abs(mz - rmz) <= bar * max(1.0, abs(rmz))
        assert np.allclose(rms, rrms, rtol=bar)


@pytest.mark.timeout(600)
def test_3d_spin2_256_fp64_against_oracle():
    inp = d3_spin2(0)
    phi0 = oracle_start(inp, seed=33)
    p_or, _ = O.derive(inp)
    o = O.Oracle(p_or)
    o.set_phi(phi0)
    print("3D spin-2 256^3: oracle, 2 real-time steps", flush=True)
    o.run(2)
    ref = o.get_phi()
    rn = o.last_norms()
    ren, rmu, rmz = o.energy()
    rrms = o.rad()
    o.close()
    p, _ = fgpe.derive(inp, FGPE_FP64)
    with fgpe.Solver(p) as s:
        s.upload(phi0)
        del phi0
        s.run(2)
        got = s.download()
        gn = s.last_norms()
        en, mu, mz, rms = s.observables()
    err = rel_l2(got, ref)
    print("3D spin-2 256^3 fp64, 2 steps: phi %.3e EN %.12g (oracle %.12g)" % (err, en, ren), flush=True)
    assert err < 1e-10
    assert np.allclose(gn, rn, rtol=1e-10, atol=1e-14)
    assert abs(en - ren) <= 1e-10 * abs(ren)
    assert np.allclose(mu, rmu, rtol=1e-10, atol=1e-12)
    assert abs(mz - rmz) <= 1e-10 * max(1.0, abs(rmz))
    assert np.allclose(rms, rrms, rtol=1e-10)


@pytest.mark.timeout(600)
def test_3d_spin2_256_fp32_densities_after_1e4_steps():
    p_im, _ = fgpe.derive(d3_spin2(1, dt=0.005), FGPE_FP64)
    with fgpe.Solver(p_im) as s:
        s.initial_guess()
        s.normalize()
        s.run(400)
        phi = s.download()
    phi0 = noisy(phi, seed=2025)
    del phi
    out = {}
    for prec in (FGPE_FP64, FGPE_FP32):
        p, _ = fgpe.derive(d3_spin2(0), prec)
        with fgpe.Solver(p) as s:
            s.upload(phi0)
            for k in range(10):
                s.run(1000)
                s.sync()
                print("  %s: %d steps" % ("fp64" if prec == FGPE_FP64 else "fp32", 1000 * (k + 1)), flush=True)
            out[prec] = s.download()
    dens = rel_l2(np.abs(out[FGPE_FP32]) ** 2, np.abs(out[FGPE_FP64]) ** 2)
    err = rel_l2(out[FGPE_FP32], out[FGPE_FP64])
    print("256^3 spin-2, 1e4 real-time steps: FP32 vs FP64 densities %.3e, phi %.3e" % (dens, err), flush=True)
    assert dens < 1e-5
