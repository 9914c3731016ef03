"""Shared small configurations for the oracle/GPU parity tests.

Physical parameters follow the benchmark configurations (87Rb spin-2:
A0/A2/A4 = 87.93/91.28/99.18 a_B; spin-1: 101.8/100.4 a_B) and the
reference's committed 1D run (23Na, INPUT/input1D.dat), at grid sizes the
FP64 oracle finishes in seconds.
"""
import numpy as np

from fortress_amd.abi import make_input

RB_SPIN2 = dict(mass=87.0, a=(87.93, 91.28, 99.18))
RB_SPIN1 = dict(mass=87.0, a=(101.8, 100.4))
NA_SPIN2 = dict(mass=23.0, a=(34.9, 45.8, 64.5))


def case(dim, spin, n, d, dt, gamma, imag=1, natoms=10000, nu=(20.0, 400.0, 400.0), option_fpc=4, mag=0.0,
         species=None):
    sp = species or (RB_SPIN2 if spin == 2 else RB_SPIN1)
    return make_input(dim, spin, n, d, dt, sp["mass"], sp["a"], natoms, nu, gamma, switch_im=imag,
                      option_fpc=option_fpc, mag=mag)


def noisy(phi, seed, amp=0.01):
    """phi + amp*max|phi|*(xi_r + i xi_i)/sqrt(2), xi ~ N(0,1), fixed seed."""
    rng = np.random.default_rng(seed)
    scale = amp * np.abs(phi).max() / np.sqrt(2.0)
    return phi + scale * (rng.standard_normal(phi.shape) + 1j * rng.standard_normal(phi.shape))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# Initial-guess cases over OPTION_FPC 1-4, with the branch (Thomas-Fermi or
# Gaussian profile) that the reference takes for each.
GUESS_CASES = [
    # dim, spin, OPTION_FPC, MAG, n, d, natoms, nu, profile of the run's branch
    (3, 1, 1, 0.3, 128, 0.3, 3e7, (20.0, 25.0, 30.0), "tf"),
    (3, 1, 1, -0.6, 64, 0.3, 1e4, (20.0, 25.0, 30.0), "gauss"),
    (3, 2, 1, -0.8, 128, 0.25, 1e7, (20.0, 25.0, 30.0), "tf"),
    (3, 2, 1, 0.5, 64, 0.3, 1e5, (20.0, 25.0, 30.0), "gauss"),
    (2, 2, 1, 1.2, 256, 0.15, 1e6, (20.0, 25.0, 400.0), "tf"),
    (2, 1, 1, 0.2, 128, 0.15, 1e3, (20.0, 25.0, 400.0), "gauss"),
    (1, 2, 1, 0.7, 512, 0.05, 1e2, (20.0, 400.0, 400.0), "gauss"),
    (3, 1, 2, 0.4, 64, 0.3, 1e4, (20.0, 25.0, 30.0), "gauss"),
    (2, 2, 2, -0.6, 128, 0.15, 1e4, (20.0, 25.0, 400.0), "gauss"),
    (3, 2, 3, 0.25, 64, 0.3, 1e4, (20.0, 25.0, 30.0), "gauss"),
    (1, 2, 3, -0.5, 512, 0.05, 1e2, (20.0, 400.0, 400.0), "gauss"),
    (3, 1, 4, 0.0, 64, 0.3, 1e4, (20.0, 25.0, 30.0), "gauss"),
    (3, 2, 4, 0.0, 64, 0.3, 1e4, (20.0, 25.0, 30.0), "gauss"),
    (2, 2, 4, 0.0, 128, 0.15, 1e4, (20.0, 25.0, 400.0), "gauss"),
    (1, 1, 4, 0.0, 512, 0.05, 1e2, (20.0, 400.0, 400.0), "gauss"),
    # 1D Thomas-Fermi: the reference raises to 0.67 (not 2/3), see tf1d below
    (1, 1, 1, 0.3, 512, 0.05, 1e4, (20.0, 400.0, 400.0), "tf1d"),
    (1, 2, 1, -0.4, 512, 0.05, 1e3, (20.0, 400.0, 400.0), "tf1d"),
    # 2D spin-1 Thomas-Fermi: the support test compares V with the 8th power
    # of the TF scale, so the reference takes sqrt of a negative number
    # outside the TF radius (NaN there)
    (2, 1, 1, 0.2, 256, 0.15, 1e6, (20.0, 25.0, 400.0), "tf_nan"),
]
