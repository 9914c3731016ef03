"""The 3D pencil schedule (fgpe_set_schedule): z-chunked x-line/forward-y
phases and x-chunked z-line/inverse-y phases on one to three streams give
bit-identical results to the whole-grid schedule, in real and imaginary time,
with SO coupling and with the SOC = 0 projection; chunk sizes that do not
split an axis into whole blocks fall back to the whole grid.  The whole-grid
schedule itself is held to the oracle in test_gpu_parity."""
import numpy as np
import pytest

from cases import case, noisy
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64

pytestmark = pytest.mark.gpu

CASES = {
    "s2_soc_real_fp32": (case(3, 2, (64, 32, 64), (0.4, 0.8, 0.4), 5e-4, (0.5, 0.3, 0.7), imag=0, natoms=100000,
                              nu=(20, 20, 20)), FGPE_FP32, 7),
    "s2_soc_imag_fp64": (case(3, 2, (32, 32, 64), (0.4, 0.4, 0.2), 5e-4, (0.5, 0.5, 0.5), natoms=100000,
                              nu=(20, 20, 20)), FGPE_FP64, 6),
    "s1_sigma_imag_fp32": (case(3, 1, (64, 64, 32), (0.4, 0.4, 0.4), 5e-4, (0.0, 0.0, 0.0), mag=0.25,
                                natoms=100000, nu=(20, 20, 20)), FGPE_FP32, 6),
}
# (zchunk, xchunk, streams)
SCHEDULES = [(0, 0, 1), (8, 16, 1), (16, 32, 2), (8, 16, 3), (4, 0, 3), (12, 24, 2), (0, 16, 2)]


@pytest.mark.parametrize("name", sorted(CASES))
def test_schedules_are_bit_identical(name):
    inp, prec, nsteps = CASES[name]
    p, _ = fgpe.derive(inp, prec)
    with fgpe.Solver(p) as s:
        s.initial_guess()
        s.normalize()
        phi0 = s.download()
    if not inp.switch_im:
        phi0 = noisy(phi0, 5)
    ref = None
    for zc, xc, ns in SCHEDULES:
        with fgpe.Solver(p) as s:
            s.set_schedule(zc, xc, ns)
            eff = s.schedule()
            s.upload(phi0)
            s.run(nsteps)
            s.run(nsteps)
            got = (s.download(), s.last_norms())
        print(name, "schedule", (zc, xc, ns), "in effect", eff, flush=True)
        assert eff[0] == (zc if zc and p.n[2] % zc == 0 else p.n[2])
        # x-chunks hold whole pass-D block pairs: 32 columns
        q = 32
        assert eff[1] == (xc if xc and p.n[0] % xc == 0 and xc % q == 0 and xc < p.n[0] else p.n[0])
        assert eff[2] == ns
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])


def test_tuner_restores_state_and_keeps_results():
    """fgpe_tune_schedule times its candidates on the handle's own state and
    puts that state (and the norms) back: a run after tuning equals a run on
    an untuned handle bit for bit."""
    inp, prec, nsteps = CASES["s2_soc_real_fp32"]
    p, _ = fgpe.derive(inp, prec)
    with fgpe.Solver(p) as s:
        s.initial_guess()
        s.normalize()
        phi0 = noisy(s.download(), 9)
    with fgpe.Solver(p) as s:
        s.upload(phi0)
        s.run(nsteps)
        ref = (s.download(), s.last_norms())
    with fgpe.Solver(p) as s:
        s.upload(phi0)
        before = s.download()
        ms, chosen = s.tune_schedule(3)
        print("tuner ms/step", ms, "chosen", chosen, flush=True)
        assert np.array_equal(s.download(), before)
        assert (ms > 0).all() and chosen == s.schedule()
        s.run(nsteps)
        assert np.array_equal(s.download(), ref[0])
        assert np.array_equal(s.last_norms(), ref[1])
