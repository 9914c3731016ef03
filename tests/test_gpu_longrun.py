"""FP32 propagation over a long real-time run against FP64 on the GPU.

3D spin-2, 128^3 at DX = 0.2 (the 25.6 box of the 256^3 DX = 0.1 workload),
87Rb spin-2 couplings, N = 1e5, 20 Hz trap, GAMMA = 0.5 on every axis,
DT = 0.0005; start: 400 imaginary-time steps from the Gaussian guess plus
seeded 1% complex noise.  After 10^4 steps the FP32 component densities must
match the FP64 ones within 1e-5 relative L2 (the FP64 path itself is held to
1e-10 against the CPU oracle in test_gpu_parity).  A rounding bias that
repeats every step would grow linearly and fail this; unbiased rounding
random-walks to a few 1e-6.
"""
import numpy as np
import pytest

from cases import rel_l2
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input

pytestmark = pytest.mark.gpu


def workload(imag, dt):
    return make_input(3, 2, (128, 128, 128), (0.2, 0.2, 0.2), dt, 87.0, (87.93, 91.28, 99.18), 100000,
                      (20.0, 20.0, 20.0), (0.5, 0.5, 0.5), switch_im=imag, stp=100)


def test_fp32_densities_after_1e4_steps():
    p_im, _ = fgpe.derive(workload(1, 0.005), FGPE_FP64)
    with fgpe.Solver(p_im) as s:
        s.initial_guess()
        s.normalize()
        s.run(400)
        phi = s.download()
    rng = np.random.default_rng(2024)
    amp = 0.01 * np.abs(phi).max() / np.sqrt(2.0)
    phi0 = phi + amp * (rng.standard_normal(phi.shape) + 1j * rng.standard_normal(phi.shape))
    out = {}
    for prec in (FGPE_FP64, FGPE_FP32):
        p, _ = fgpe.derive(workload(0, 5e-4), prec)
        with fgpe.Solver(p) as s:
            s.upload(phi0)
            s.run(10000)
            out[prec] = s.download()
    dens = rel_l2(np.abs(out[FGPE_FP32]) ** 2, np.abs(out[FGPE_FP64]) ** 2)
    err = rel_l2(out[FGPE_FP32], out[FGPE_FP64])
    print("128^3 spin-2, 1e4 real-time steps: FP32 vs FP64 densities %.3e, phi %.3e" % (dens, err), flush=True)
    assert dens < 1e-5
