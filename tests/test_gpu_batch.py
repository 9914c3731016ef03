"""Batched q2D instances (fgpe_batch_*) against the FP64 oracle and the
single-instance path.

Instances follow the parameter sweep of the batch workload (square grids of
64, 128, 192 and 256 points, DX = 0.1, DT = 0.001, spin-1 87Rb with A2 - A0
varied, GAMMAX/GAMMAY, NATOMS and the in-plane trap aspect varied, imaginary
time), plus rectangular grids and spin-2 87Rb batches (with and without SOC:
the SOC = 0 instances take the Newton projection).  Bars: FP64 1e-10
relative on phi, norms and observables; FP32 densities 1e-5.  Each
instance's result must not depend on what else is in the batch or on the
schedule the autotuner picks (FP64: 1e-12).
"""
import numpy as np
import pytest

import oracle_ctypes as O
from cases import rel_l2
from fortress_amd import fgpe
from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input

pytestmark = pytest.mark.gpu

SWEEP = [  # n, GAMMAX, GAMMAY, A2 - A0, NATOMS, trap aspect
    (64, 0.3, 1.7, -4.0, 2000, 0.9),
    (128, 1.2, 0.4, 3.0, 30000, 1.2),
    (192, 0.8, 0.8, 1.5, 10000, 1.0),
    (256, 1.9, 0.1, -2.5, 80000, 0.85),
]


def inst(n, gx, gy, da, natoms, aspect):
    return make_input(2, 1, (n, n), (0.1, 0.1), 1e-3, 87.0, (101.8, 101.8 + da, 0.0), natoms,
                      (20.0, 20.0 * aspect, 400.0), (gx, gy), switch_im=1)


def inst_shape(nx, ny, spin, gx, gy, natoms, aspect, mag=0.0):
    a = (87.93, 91.28, 99.18) if spin == 2 else (101.8, 103.3, 0.0)
    return make_input(2, spin, (nx, ny), (0.1, 0.1), 1e-3, 87.0, a, natoms, (20.0, 20.0 * aspect, 400.0), (gx, gy),
                      switch_im=1, mag=mag)


SPIN2_SHAPES = [  # nx, ny, spin, GAMMAX, GAMMAY, NATOMS, aspect, MAG
    (64, 64, 2, 1.0, 1.0, 50000, 1.0, 0.0),
    (128, 64, 2, 0.6, 1.3, 20000, 1.1, 0.0),
    (64, 192, 2, 1.4, 0.3, 80000, 0.9, 0.0),
    (128, 128, 2, 0.0, 0.0, 30000, 1.0, 0.3),
]
SPIN1_SHAPES = [
    (64, 128, 1, 0.3, 1.7, 2000, 0.9),
    (256, 64, 1, 1.2, 0.4, 30000, 1.2),
    (192, 128, 1, 0.8, 0.8, 10000, 1.0),
]


def oracle_start(inp):
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    return o


def check_fp64_against_oracle(inps, steps=20):
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        starts = [b.download(i) for i in range(len(inps))]
        b.run(steps)
        for i, inp in enumerate(inps):
            ms = 2 * params[i].spin + 1
            print("instance", i, "grid", tuple(params[i].n[:2]), "spin", params[i].spin, flush=True)
            o = oracle_start(inp)
            assert rel_l2(starts[i], o.get_phi()) < 1e-13  # guess + NORMT
            o.run(steps - 1)
            prev = o.get_phi()
            o.run(1)
            ref = o.get_phi()
            got = b.download(i)
            r = b.result(i)
            err = rel_l2(got, ref)
            print("  rel_l2 phi", err, flush=True)
            assert err < 1e-10
            assert r.steps == steps and r.err == 0
            assert np.allclose(np.array(r.norm[:ms]), o.last_norms(), rtol=1e-10, atol=1e-14)
            en, mu, mz = o.energy()
            assert abs(r.en - en) <= 1e-10 * abs(en)
            assert np.allclose(np.array(r.mu[:ms]), mu, rtol=1e-10, atol=1e-12)
            assert abs(r.mz - mz) <= 1e-10 * max(1.0, abs(mz))
            assert np.allclose(np.array(r.rms[:ms]), o.rad(), rtol=1e-10)
            maxdiff = np.abs(ref - prev).max()
            assert abs(r.maxdiff - maxdiff) <= 1e-9 * maxdiff
            o.close()


def test_batch_fp64_matches_oracle():
    check_fp64_against_oracle([inst(*s) for s in SWEEP])


def test_batch_rectangular_spin1_fp64_matches_oracle():
    check_fp64_against_oracle([inst_shape(*s) for s in SPIN1_SHAPES])


def test_batch_spin2_fp64_matches_oracle():
    """Spin-2 instances of several shapes, one of them with SOC = 0 and
    MAG = 0.3 (the Newton projection from its (20, 20) start)."""
    check_fp64_against_oracle([inst_shape(*s) for s in SPIN2_SHAPES])


def test_batch_spin2_fp32_densities():
    inps = [inst_shape(*s) for s in SPIN2_SHAPES]
    params = [fgpe.derive(i, FGPE_FP32)[0] for i in inps]
    steps = 20
    with fgpe.Batch(params) as b:
        b.initial_guess()
        b.run(steps)
        for i, inp in enumerate(inps):
            o = oracle_start(inp)
            o.run(steps)
            dens = rel_l2(np.abs(b.download(i)) ** 2, np.abs(o.get_phi()) ** 2)
            print("instance", i, "grid", tuple(params[i].n[:2]), "fp32 densities", dens, flush=True)
            assert dens < 1e-5
            o.close()


def test_rectangular_spin2_alone_in_batch_and_single_instance_agree():
    inps = [inst_shape(*s) for s in SPIN2_SHAPES[1:3]]
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        b.run(15)
        together = [b.download(i) for i in range(len(params))]
    for i in range(len(params)):
        with fgpe.Batch([params[i]]) as b1:
            b1.initial_guess()
            b1.run(15)
            alone = b1.download(0)
        assert rel_l2(alone, together[i]) < 1e-14
        with fgpe.Solver(params[i]) as s:
            s.initial_guess()
            s.normalize()
            s.run(15)
            single = s.download()
        assert rel_l2(single, together[i]) < 1e-12


def test_batch_fp32_densities():
    inps = [inst(*s) for s in SWEEP]
    params = [fgpe.derive(i, FGPE_FP32)[0] for i in inps]
    steps = 20
    with fgpe.Batch(params) as b:
        b.initial_guess()
        b.run(steps)
        for i, inp in enumerate(inps):
            o = oracle_start(inp)
            o.run(steps)
            ref = o.get_phi()
            got = b.download(i)
            dens = rel_l2(np.abs(got) ** 2, np.abs(ref) ** 2)
            print("instance", i, "n", SWEEP[i][0], "fp32 densities", dens, flush=True)
            assert dens < 1e-5
            en, _, _ = o.energy()
            assert abs(b.result(i).en - en) <= 1e-5 * abs(en)
            o.close()


def test_batch_matches_single_instance_path_and_is_independent():
    """An instance gives the same numbers alone, in a batch, and through
    fgpe_run (sizes the single-instance path supports)."""
    sweep = [SWEEP[0], SWEEP[1], SWEEP[3]]
    inps = [inst(*s) for s in sweep]
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        b.run(15)
        together = [b.download(i) for i in range(3)]
    for i in range(3):
        with fgpe.Batch([params[i]]) as b1:
            b1.initial_guess()
            b1.run(15)
            alone = b1.download(0)
        assert rel_l2(alone, together[i]) < 1e-14
        with fgpe.Solver(params[i]) as s:
            s.initial_guess()
            s.normalize()
            s.run(15)
            single = s.download()
        assert rel_l2(single, together[i]) < 1e-12


def test_batch_tune_restores_state_and_keeps_results():
    inps = [inst(*s) for s in SWEEP[:3]]
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        before = [b.download(i) for i in range(3)]
        ms, chosen = b.tune(3)
        print("tune ms", ms, "chosen", chosen, flush=True)
        after = [b.download(i) for i in range(3)]
        for x, y in zip(before, after):
            assert np.array_equal(x, y)
        assert chosen[0] in (0, 1) and chosen[1] in (2, 3) and chosen[2] in (4, 5) and chosen[3] == -1
        assert all(ms[s] > 0 for s in range(6))
        b.run(10)
        tuned = [b.download(i) for i in range(3)]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        b.run(10)
        for i in range(3):
            assert rel_l2(b.download(i), tuned[i]) < 1e-12


def test_batch_solve_convergence_semantics():
    """solve() stops an instance at the first STP event with CONV <= TOL or
    CONV_EN <= TOL (CONV = max|phi - phi_old|/(STP 2 DT), phi_old one step
    earlier; CONV_EN against the previous event's energy), and an instance
    that is still running at max_steps has passed no event."""
    inps = [inst(64, 0.5, 0.5, 1.0, 2000, 1.0), inst(64, 1.5, 0.2, -3.0, 5000, 1.1),
            inst(128, 0.2, 1.0, 2.0, 20000, 0.9)]
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    # CONV of these instances (oracle): the first passes 2e-4 after about 1000
    # steps, the other two stay above it through 3000
    stp, tol, max_steps = 50, 2e-4, 3000
    with fgpe.Batch(params) as b:
        b.initial_guess()
        en0 = [b.result(i).en for i in range(3)]
        nconv = b.solve(stp, tol, max_steps)
        res = [b.result(i) for i in range(3)]
    print("converged", nconv, "steps", [r.steps for r in res], flush=True)
    assert nconv == sum(r.done for r in res) and nconv >= 1
    for i, r in enumerate(res):
        assert r.steps % stp == 0 and r.steps > 0
        if r.done:
            assert r.conv <= tol or r.conv_en <= tol
        else:
            assert r.steps + stp > max_steps
        # replay with the oracle: the first event that passes must be the last one run
        o = oracle_start(inps[i])
        en_old = en0[i]
        first = None
        for ev in range(1, r.steps // stp + 1):
            o.run(stp - 1)
            prev = o.get_phi()
            o.run(1)
            cur = o.get_phi()
            conv = np.abs(cur - prev).max() / (stp * 2.0 * params[i].dt)
            en = o.energy()[0]
            conv_en = abs(en - en_old) / (stp * 2.0 * params[i].dt)
            en_old = en
            if conv <= tol or conv_en <= tol:
                first = ev
                break
        o.close()
        assert first == (r.steps // stp if r.done else None)
