"""Imaginary-time ground states to TOL, with cgpe.x's events and exit test,
in FP64 or with FP32 bulk propagation finished in FP64.

The event loop is that of cgpe.x (SRC/cgpe3d.f90:451-531; 2D
SRC/cgpe2d.f90:440-495; 1D SRC/cgpe1d.f90:433-507) without its files: every
STP steps the observables are taken, CONV = max|PHI(II) - PHI(II-1)|/(STP 2 DT)
(PHI_OLD is the state one step earlier, SRC/cgpe3d.f90:530) and
CONV_EN = |EN(II) - EN(II-STP)|/(STP 2 DT); the run stops when either is
<= TOL (the reported iteration is then II - 1, SRC/cgpe3d.f90:579) or at NITER.

FP32 propagation cannot reach a small TOL: its per-step change never drops
below the rounding of the state, so CONV stalls near 2 eps32 max|PHI| /
(STP 2 DT) (eps32 = 2^-24; measured 8e-8 on a q2D spin-1 64^2 case where
this estimate gives 9e-8), and CONV_EN is rounding noise once the energy
changes less than its FP32 rounding -- noise that now and then falls below
TOL by chance (q2D spin-2 256^2: CONV_EN 8.8e-10 at CONV 1e-6, far from the
ground state).  With finish=True the FP32 run therefore never exits: it
hands its state to an FP64 solver (upload converts it) once CONV is within
`switch_factor` (2) of that floor, or is within 30x of it and no longer
falling (above 0.9 of its value 20 events earlier; a healthy exponential
decay loses more in 20 events), or CONV_EN is <= TOL; the FP64 solver
continues from the same iteration, EN_OLD and event phase and applies the
exit test.  What FP32 leaves in the slowest modes (~1e-7 of the state) still
relaxes at their rate, so the FP64 part covers the decay from the FP32 floor
to TOL.  finish=False runs cgpe.x's test in FP32 as it stands.
"""
import time

import numpy as np

from . import fgpe
from .abi import FGPE_FP32, FGPE_FP64

EPS32 = 2.0 ** -24


class Result:
    """What a solve ended with: the printed exit iteration (II - 1 after an
    exit, NITER otherwise), the observables at the last event, the state, and
    how the steps and the wall clock divide between the precisions."""

    def __init__(self):
        self.converged = False
        self.ii = 0
        self.ii_print = 0
        self.en = self.mz = self.conv = self.conv_en = float("nan")
        self.mu = self.rms = self.norms = None
        self.phi = None
        self.steps_fp32 = self.steps_fp64 = 0
        self.switch_ii = None
        self.wall_s = 0.0
        self.wall_fp32_s = 0.0
        self.events = []  # (II, CONV, CONV_EN, precision)

    def as_dict(self):
        return dict(converged=self.converged, ii=self.ii, ii_print=self.ii_print, en=self.en,
                    mu=[float(v) for v in self.mu], mz=self.mz, rms=[float(v) for v in self.rms],
                    conv=self.conv, conv_en=self.conv_en, steps_fp32=self.steps_fp32, steps_fp64=self.steps_fp64,
                    switch_ii=self.switch_ii, wall_s=self.wall_s, wall_fp32_s=self.wall_fp32_s)


def fp32_conv_floor(center_abs_max, stp, dt):
    """CONV that FP32 rounding of the state alone produces."""
    return 2.0 * EPS32 * center_abs_max / (stp * 2.0 * dt)


def solve(inp, precision=FGPE_FP64, phi0=None, finish=True, switch_factor=2.0, keep_phi=True, log=None):
    """Run the imaginary-time event loop of `inp` (an FgpeInput with
    switch_im = 1) from phi0 (normalised first) or the Gaussian guess of
    OPTION_FPC, to TOL or NITER.  precision FGPE_FP32 with finish=True ends
    in FP64 (see the module docstring).  log(ii, conv, conv_en, prec) is
    called at every event."""
    if inp.switch_im != 1:
        raise ValueError("solve() runs imaginary time (SWITCH_IM = 1)")
    stp, niter, dt, tol = int(inp.stp), int(inp.niter), float(inp.dt), float(inp.tol)
    den = stp * 2.0 * dt
    res = Result()
    t0 = time.perf_counter()
    prec = precision
    p, _ = fgpe.derive(inp, prec)
    s = fgpe.Solver(p)
    try:
        if phi0 is None:
            s.initial_guess()
        else:
            s.upload(phi0)
        s.normalize()
        en = s.observables()[0]
        # EN_OLD is the initial EN at the first test unless STP = 1
        # (SRC/cgpe3d.f90:307,418,529)
        en_old = en if stp > 1 else 0.0
        ii = 0
        history = []  # CONV of the FP32 events
        floor = None
        while ii < niter:
            ii0 = ii
            nextev = min(niter, (ii // stp + 1) * stp)
            at_stp = nextev % stp == 0
            if at_stp:
                if nextev - 1 > ii:
                    s.run(nextev - 1 - ii)
                    ii = nextev - 1
                if ii > 0:
                    s.keep_old()
            s.run(nextev - ii)
            steps = nextev - ii0
            if prec == FGPE_FP32:
                res.steps_fp32 += steps
            else:
                res.steps_fp64 += steps
            ii = nextev
            if not at_stp:
                continue
            en, mu, mz, rms = s.observables()
            conv = s.conv_max() / den
            conv_en = abs(en - en_old) / den
            res.events.append((ii, conv, conv_en, "fp32" if prec == FGPE_FP32 else "fp64"))
            if log:
                log(ii, conv, conv_en, prec)
            res.en, res.mu, res.mz, res.rms, res.conv, res.conv_en = en, mu, mz, rms, conv, conv_en
            fp32_phase = prec == FGPE_FP32 and finish
            if not fp32_phase and (conv <= tol or conv_en <= tol):
                res.converged = True
                break
            en_old = en
            if fp32_phase:
                if floor is None:
                    floor = fp32_conv_floor(float(np.max(s.center_abs())), stp, dt)
                history.append(conv)
                stalled = len(history) > 20 and conv <= 30.0 * floor and conv > 0.9 * history[-21]
                if conv <= switch_factor * floor or stalled or conv <= tol or conv_en <= tol:
                    phi = s.download()
                    s.close()
                    res.wall_fp32_s = time.perf_counter() - t0
                    res.switch_ii = ii
                    prec = FGPE_FP64
                    p, _ = fgpe.derive(inp, prec)
                    s = fgpe.Solver(p)
                    s.upload(phi)
                    del phi
        res.ii = ii
        res.ii_print = ii - 1 if res.converged else niter
        res.norms = s.last_norms()
        if res.mu is None:
            res.en, res.mu, res.mz, res.rms = s.observables()
        if keep_phi:
            res.phi = s.download()
        s.sync()
    finally:
        s.close()
    res.wall_s = time.perf_counter() - t0
    if res.switch_ii is None and precision == FGPE_FP32:
        res.wall_fp32_s = res.wall_s
    return res
