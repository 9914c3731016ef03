"""Learned initial guess for q2D spin-1 ground states.

A network maps (GAMMAX, GAMMAY, A0, A2, NATOMS, trap aspect) to the three
complex components of the ground state on a coarse 64x64 grid covering the
same box as the production grid; the prediction is interpolated spectrally
onto the production grid and used in place of the Gaussian guess of
INITIALIZE.  Training targets are ground states the propagator itself
produces (imaginary time from the Gaussian guess, run to a tolerance on the
coarse grid).  The imaginary-time solve to TOL judges the guess: the same
event loop and convergence test as cgpe.x (CONV = max|PHI - PHI_OLD|/(STP 2
DT), CONV_EN = |EN - EN_OLD|/(STP 2 DT), exit when either <= TOL;
SRC/cgpe2d.f90:486-495).  The plain Gaussian guess stays the default.
"""
import math
import time

import numpy as np

from . import fgpe
from .abi import FGPE_FP64, make_input
from .mlp import MLP

MASS = 87.0
A0_CENTER = 101.8
BOX = 25.6            # box length of the production and coarse grids
COARSE = 64
NU_X, NU_Z = 20.0, 400.0


def draw(rng):
    """One parameter draw: GAMMAX, GAMMAY ~ U[0,2]; A0 ~ U[99.8,103.8];
    A2 - A0 ~ U[-5,5] a_B; NATOMS ~ logU[1e3,1e5]; trap aspect ~ U[0.8,1.25]."""
    a0 = rng.uniform(A0_CENTER - 2.0, A0_CENTER + 2.0)
    return dict(gx=rng.uniform(0, 2), gy=rng.uniform(0, 2), a0=a0, a2=a0 + rng.uniform(-5, 5),
                natoms=int(round(10 ** rng.uniform(3, 5))), aspect=rng.uniform(0.8, 1.25))


def features(d):
    return np.array([d["gx"] - 1.0, d["gy"] - 1.0, (d["a0"] - A0_CENTER) / 2.0, (d["a2"] - d["a0"]) / 5.0,
                     math.log10(d["natoms"]) - 4.0, (d["aspect"] - 1.025) / 0.225], dtype=np.float32)


def problem(d, n, dt=1e-3, precision=FGPE_FP64):
    """Solver parameters of draw d on an n x n grid over the fixed box."""
    inp = make_input(2, 1, (n, n), (BOX / n, BOX / n), dt, MASS, (d["a0"], d["a2"], 0.0), d["natoms"],
                     (NU_X, NU_X * d["aspect"], NU_Z), (d["gx"], d["gy"]), switch_im=1)
    return fgpe.derive(inp, precision)[0]


def solve(params, phi0=None, stp=100, tol=1e-9, max_steps=400000):
    """Imaginary time to TOL from phi0 (or the Gaussian guess).  Returns
    (phi, steps, energy, device+host seconds)."""
    t0 = time.perf_counter()
    with fgpe.Solver(params) as s:
        if phi0 is None:
            s.initial_guess()
        else:
            s.upload(phi0)
        s.normalize()
        en_old = s.observables()[0]
        steps = 0
        den = stp * 2.0 * params.dt
        while steps < max_steps:
            s.run(stp - 1)
            s.keep_old()
            s.run(1)
            steps += stp
            conv = s.conv_max() / den
            en = s.observables()[0]
            conv_en = abs(en - en_old) / den
            en_old = en
            if conv <= tol or conv_en <= tol:
                break
        phi = s.download()
        s.sync()
    return phi, steps, en, time.perf_counter() - t0


def to_vector(phi):
    """(3, 1, n, n) complex -> [re, im] of the three components, flattened."""
    p = phi.reshape(3, phi.shape[-2], phi.shape[-1])
    return np.concatenate([p.real.ravel(), p.imag.ravel()]).astype(np.float32)


def from_vector(v, n=COARSE):
    v = np.asarray(v, dtype=np.float64)
    h = v.size // 2
    return (v[:h] + 1j * v[h:]).reshape(3, 1, n, n)


def _pad_spectrum(c, n, axis):
    """Zero-pad a length-m spectrum along axis to length n (Nyquist split)."""
    m = c.shape[axis]
    h = m // 2
    shape = list(c.shape)
    shape[axis] = n
    out = np.zeros(shape, dtype=np.complex128)
    take = lambda a, sl: a[(slice(None),) * (axis % c.ndim) + (sl,)]
    put = lambda sl: (slice(None),) * (axis % c.ndim) + (sl,)
    out[put(slice(0, h))] = take(c, slice(0, h))
    out[put(slice(n - h + 1, n))] = take(c, slice(m - h + 1, m))
    out[put(slice(h, h + 1))] = 0.5 * take(c, slice(h, h + 1))
    out[put(slice(n - h, n - h + 1))] += 0.5 * take(c, slice(h, h + 1))
    return out


def upsample(phi, n):
    """Trigonometric interpolation of (3, 1, m, m) onto (3, 1, n, n), n >= m:
    the coarse points x_i = -L/2 + i L/m coincide with every (n/m)-th fine
    point, and zero-padding the DFT (Nyquist split evenly) interpolates."""
    m = phi.shape[-1]
    if n == m:
        return phi.copy()
    c = np.fft.fft2(phi[:, 0], axes=(-2, -1))
    c = _pad_spectrum(_pad_spectrum(c, n, -1), n, -2)
    return (np.fft.ifft2(c, axes=(-2, -1)) * (n / m) ** 2)[:, None]


class Surrogate:
    """Network + output scale, trained on solver-produced ground states."""

    def __init__(self, hidden=512, layers=4, max_batch=64, seed=0):
        dims = [6] + [hidden] * layers + [6 * COARSE * COARSE]
        self.net = MLP(dims, max_batch=max_batch, seed=seed)
        self.mean, self.sd = 0.0, 1.0

    def fit(self, feats, targets, epochs, batch, lr, seed=0, lr_decay=0.97, log=None):
        """Trains on standardised targets: per output, minus its mean over the
        training set, over its spread (floored at 1e-3 of the largest spread),
        so the loss weighs how the state changes with the parameters rather
        than the state all draws share."""
        t = np.asarray(targets, dtype=np.float64)
        self.mean = t.mean(axis=0)
        sd = t.std(axis=0)
        self.sd = np.maximum(sd, 1e-3 * sd.max()) if sd.max() > 0 else np.ones_like(sd)
        self.net.set_dataset(feats, ((t - self.mean) / self.sd).astype(np.float32))
        return self.net.train(epochs, batch, lr, seed=seed, lr_decay=lr_decay, log=log)

    def predict_vectors(self, feats):
        return self.net.forward(feats) * self.sd + self.mean

    def predict(self, d):
        return from_vector(self.predict_vectors(features(d)[None])[0])

    def guess(self, d, n):
        return upsample(self.predict(d), n)

    def save(self, path):
        """The trained state (FP32 master weights, the targets' mean and
        spread) as an .npz file that load() restores into a network of the
        same shape."""
        np.savez(path, params=self.net.get_params(), mean=np.asarray(self.mean, np.float64),
                 sd=np.asarray(self.sd, np.float64), dims=np.asarray(self.net.dims, np.int64))

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        if list(z["dims"]) != list(self.net.dims):
            raise ValueError("saved network %s does not match %s" % (list(z["dims"]), self.net.dims))
        self.net.set_params(z["params"])
        self.mean, self.sd = z["mean"], z["sd"]

    def close(self):
        self.net.close()


def solve_batch(b, stp, tol, max_steps, chunk=20000, log=None):
    """fgpe_batch_solve in step budgets of `chunk` (the solve resumes where it
    stopped), calling log(steps_so_far, converged) after each budget."""
    limit, nconv = 0, 0
    while limit < max_steps:
        limit = min(limit + chunk, max_steps)
        nconv = b.solve(stp, tol, limit)
        if log:
            log(limit, nconv)
        if nconv == b.count:
            break
    return nconv


def make_dataset(draws, tol=1e-7, stp=100, max_steps=400000, log=None):
    """Coarse ground states (64 x 64 over the same box) for each draw, solved
    together as one batch (fgpe_batch_*: the same step and convergence test
    as solve()).  Returns (features, targets, steps per draw)."""
    params = [problem(d, COARSE) for d in draws]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        solve_batch(b, stp, tol, max_steps, log=log)
        steps = [b.result(i).steps for i in range(len(draws))]
        targets = [to_vector(b.download(i)) for i in range(len(draws))]
    return np.array([features(d) for d in draws]), np.array(targets), steps
