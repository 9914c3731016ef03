#!/usr/bin/env python3
"""Throughput of the split-step propagator on one MI355X.

Workload (one "step" = one full time step of the reference loop,
SRC/cgpe3d.f90:452-495): 3D spin-2 real-time dynamics on a 256^3 grid,
DX=DY=DZ=0.1, DT=0.0005, 87Rb spin-2 (A0/A2/A4 = 87.93/91.28/99.18 a_B),
NATOMS=1e5, trap 20/20/20 Hz, GAMMAX=GAMMAY=GAMMAZ=0.5, FP32 propagation with
FP64 norm/energy reductions, observables (ENERGY, RAD, NORMC) once per STP=100
steps.  Initial state: an imaginary-time relaxation of the reference's
Gaussian guess (FP32, DT=0.005, --prep-steps steps) plus seeded 1% complex
Gaussian noise phi += 0.01 max|phi| (xi_r + i xi_i)/sqrt(2).  Synthetic
parameters; no data is read.

Prints one JSON line.  value = steps/s of the whole job (one GPU): the timed
region runs the steps with an observables call at every multiple of STP, and
the difference between steps/STP calls and the calls the window held is added
at the cost of one call timed on its own (timing.observables_ms), so value
carries the observables at one per STP steps for any --steps.  The 3D pencil
schedule is autotuned on the state before the timed region.  The roofline
object is the dominant kernel's algorithmic HBM bytes per launch (2*S,
S = N*MS*8 B: every pass reads and writes the whole spinor once) over its
average whole-grid launch time measured with HIP events on the library's
stream.  fp64: the same workload propagated in FP64, the reference's
arithmetic.
cpu_baseline: the FP64 CPU restatement (oracle/, OpenMP) timed on a bounded
sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D spin-2 256^3 steps/s (%HBM roofline); time-to-TOL, surrogate vs plain guess"
PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
# workload (256^3 spin-2 FP32), scaled by the counters' readings for a kernel
# that reads and writes the spinor once: profiles/passes_3d_spin2_256_hbm_traffic.md
TRAFFIC_256 = {"x": 1343233715, "y": 1342329931}


def workload(n, imag, dt):
    from fortress_amd.abi import make_input
    return make_input(3, 2, (n, n, n), (0.1, 0.1, 0.1), dt, 87.0, (87.93, 91.28, 99.18), 100000,
                      (20.0, 20.0, 20.0), (0.5, 0.5, 0.5), switch_im=imag, stp=100)


def noisy(phi, seed, amp=0.01):
    rng = np.random.default_rng(seed)
    scale = amp * np.abs(phi).max() / np.sqrt(2.0)
    out = np.empty_like(phi)
    for m in range(phi.shape[0]):  # per component to bound host memory
        out[m] = phi[m] + scale * (rng.standard_normal(phi.shape[1:]) + 1j * rng.standard_normal(phi.shape[1:]))
    return out


def cpu_baseline(phi0, n, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    O.lib().or_set_threads(threads)
    p, _ = O.derive(workload(n, 0, 5e-4))
    o = O.Oracle(p)
    o.set_phi(phi0)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": steps / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "%d FP64 steps of the same %d^3 spin-2 real-time workload and initial state "
                      "(oracle/fgpe_oracle.c, OpenMP, %d threads), %.1f s" % (steps, n, threads, dt)}


def reference_cost(n, gpu_steps_per_s):
    """The reference's own step cannot be timed (it does not build here), so
    oracle/refcost times the two library calls that dominate it -- the FFTW
    3D transforms of FFT/BFT and the two ZHEEV calls per point of the spin-2
    SE step -- on the same host threads; their sum bounds the reference's
    step time from below."""
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    exe = os.path.join(ROOT, "oracle", "refcost")
    try:
        out = subprocess.run([exe, str(n), str(threads), "200000"], capture_output=True, text=True, timeout=300)
        if out.returncode != 0:
            return {"error": (out.stderr or out.stdout).strip()[-300:]}
        r = json.loads(out.stdout.strip().splitlines()[-1])
    except (OSError, subprocess.SubprocessError, ValueError) as e:
        return {"error": str(e)}
    r["reference_steps_per_s_at_most"] = round(1.0 / r["step_lower_bound_s"], 5)
    r["gpu_vs_reference_at_least"] = round(gpu_steps_per_s * r["step_lower_bound_s"], 1)
    return r


def accuracy(args):
    """FP32 against FP64 propagation of the same workload from the same state:
    relative L2 of the component densities |phi_m|^2 at checkpoints."""
    from fortress_amd import fgpe
    from fortress_amd.abi import FGPE_FP32, FGPE_FP64
    p_im, _ = fgpe.derive(workload(args.n, 1, 0.005), FGPE_FP32)
    with fgpe.Solver(p_im) as s:
        s.initial_guess()
        s.normalize()
        s.run(args.prep_steps)
        phi_g = s.download()
    phi0 = noisy(phi_g, seed=1234)
    del phi_g
    marks = sorted(set(int(x) for x in args.accuracy.split(",")))
    p32, _ = fgpe.derive(workload(args.n, 0, 5e-4), FGPE_FP32)
    p64, _ = fgpe.derive(workload(args.n, 0, 5e-4), FGPE_FP64)
    s32, s64 = fgpe.Solver(p32), fgpe.Solver(p64)
    s32.upload(phi0)
    s64.upload(phi0)
    del phi0
    done, rows = 0, []
    for mark in marks:
        s32.run(mark - done)
        s64.run(mark - done)
        done = mark
        a = s64.download()
        num = den = 0.0
        b = s32.download()
        for m in range(a.shape[0]):
            da = np.abs(a[m]) ** 2
            db = np.abs(b[m]) ** 2
            num += float(np.sum((db - da) ** 2))
            den += float(np.sum(da ** 2))
        phi_err = float(np.sqrt(np.sum(np.abs(b - a) ** 2) / np.sum(np.abs(a) ** 2)))
        del a, b
        rows.append({"steps": mark, "density_rel_l2": float(np.sqrt(num / den)), "phi_rel_l2": phi_err})
        print(json.dumps(rows[-1]), flush=True)
    s32.close()
    s64.close()
    print(json.dumps({"metric": "FP32 vs FP64 component densities, rel L2", "n": args.n, "rows": rows}), flush=True)


def config1_input(imag, niter):
    """q1D spin-1: NX=512, DX=0.05, DT=0.0005, 87Rb A0/A2 = 101.8/100.4 a_B,
    NATOMS=1e4, trap 20/400/400 Hz, GAMMAX=0.5, TOL=1e-9, STP=100."""
    from fortress_amd.abi import make_input
    return make_input(1, 1, (512,), (0.05,), 0.0005, 87.0, (101.8, 100.4), 10000, (20.0, 400.0, 400.0), (0.5,),
                      switch_im=imag, niter=niter, stp=100, tol=1e-9)


def oracle_ground_state(O, inp):
    """The oracle through the same events as fortress_amd.groundstate.solve
    (the cgpe.x loop): (state, exit iteration, EN, CONV, CONV_EN)."""
    p, _ = O.derive(inp)
    o = O.Oracle(p)
    o.initial_guess()
    o.normt()
    en_old = o.energy()[0]
    stp, den = inp.stp, inp.stp * 2.0 * inp.dt
    ii = 0
    conv = conv_en = en = None
    while ii < inp.niter:
        o.run(stp - 1)
        prev = o.get_phi()
        o.run(1)
        ii += stp
        en = o.energy()[0]
        conv = float(np.abs(o.get_phi() - prev).max() / den)
        conv_en = abs(en - en_old) / den
        if conv <= inp.tol or conv_en <= inp.tol:
            break
        en_old = en
    phi = o.get_phi()
    o.close()
    return phi, ii, en, conv, conv_en


def config1_bench(args):
    """q1D spin-1 ground state in FP64 (cgpe.x's loop, NITER=20000, TOL=1e-9),
    then 5000 real-time steps from it plus seeded 1% complex noise; the same
    two runs by the FP64 oracle on the host are the CPU baseline and the
    parity check."""
    from fortress_amd import fgpe, groundstate
    from fortress_amd.abi import FGPE_FP64
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    O.lib().or_set_threads(1)
    inp = config1_input(1, args.config_niter)
    # warm the device path once (module load, first-launch costs)
    groundstate.solve(config1_input(1, 200), FGPE_FP64, keep_phi=False)
    t0 = time.perf_counter()
    g = groundstate.solve(inp, FGPE_FP64)
    phi0 = noisy(g.phi, seed=4321)
    p_re, _ = fgpe.derive(config1_input(0, 5000), FGPE_FP64)
    with fgpe.Solver(p_re) as s:
        s.upload(phi0)
        s.run(args.config_rt_steps)
        phi_rt = s.download()
        en_rt = s.observables()[0]
    gpu_s = time.perf_counter() - t0
    # the oracle: same events, same noisy start
    t0 = time.perf_counter()
    ophi, oii, oen, oconv, oconv_en = oracle_ground_state(O, inp)
    t_gs = time.perf_counter() - t0
    p, _ = O.derive(config1_input(0, 5000))
    o = O.Oracle(p)
    o.set_phi(phi0)
    t1 = time.perf_counter()
    o.run(args.config_rt_steps)
    t_rt = time.perf_counter() - t1
    oen_rt = o.energy()[0]
    ophi_rt = o.get_phi()
    o.close()
    cpu_s = time.perf_counter() - t0
    rel = lambda a, b: float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    line = {
        "metric": "q1D spin-1 512 ground state (cgpe.x loop to TOL 1e-9 or NITER) + 5000 real-time steps, wall s",
        "value": round(gpu_s, 4), "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "fp64",
        "data": "synthetic: Gaussian guess; real time from the ground state + seeded 1% complex noise",
        "config": {"workload": "q1D spin-1, NX=512, DX=0.05, DT=0.0005, 87Rb A0/A2=101.8/100.4, NATOMS=1e4, "
                               "trap 20/400/400 Hz, GAMMAX=0.5, TOL=1e-9, STP=100, NITER=%d; then %d real-time "
                               "steps" % (args.config_niter, args.config_rt_steps)},
        "ground_state": {"converged": g.converged, "ii_print": g.ii_print, "en": g.en / 2.0, "conv": g.conv,
                         "conv_en": g.conv_en, "wall_s": round(g.wall_s, 4),
                         "steps_per_s": round(g.ii / g.wall_s, 1)},
        "real_time": {"steps": args.config_rt_steps, "en": en_rt / 2.0},
        "vs_oracle": {"ground_state_phi_rel_l2": rel(g.phi, ophi), "exit_ii": oii, "en_rel": abs(g.en - oen) / abs(oen),
                      "conv_rel": abs(g.conv - oconv) / oconv, "real_time_density_rel_l2":
                          rel(np.abs(phi_rt) ** 2, np.abs(ophi_rt) ** 2),
                      "real_time_en_rel": abs(en_rt - oen_rt) / abs(oen_rt)},
        "cpu_baseline": {"value": round(cpu_s, 3), "unit": "s", "cores": 1, "kind": "port",
                         "sample": "the whole workload (ground state %.2f s + real time %.2f s) by the FP64 oracle "
                                   "(oracle/fgpe_oracle.c), one thread" % (t_gs, t_rt)},
    }
    print(json.dumps(line), flush=True)


def config2_bench(args):
    """q2D spin-2 Rashba ground state to TOL=1e-9 (cgpe.x's loop, STP=100):
    FP64 throughout, against FP32 propagation (FP64 reductions) finished in
    FP64 once CONV reaches the FP32 floor (fortress_amd.groundstate); wall
    clock, steps in each precision and the EN/MU/RMS error of the mixed run
    against the FP64 one."""
    from fortress_amd import groundstate
    from fortress_amd.abi import FGPE_FP32, FGPE_FP64, make_input
    from fortress_amd import fgpe
    inp = make_input(2, 2, (256, 256), (0.1, 0.1), 0.001, 87.0, (87.93, 91.28, 99.18), 100000,
                     (20.0, 20.0, 400.0), (1.0, 1.0), switch_im=1, niter=args.max_steps, stp=100, tol=1e-9)
    # start: the Gaussian guess plus seeded 1% noise, so the branch of the
    # spin texture is picked by the start and not by rounding
    with fgpe.Solver(fgpe.derive(inp, FGPE_FP64)[0]) as s:
        s.initial_guess()
        s.normalize()
        phi0 = noisy(s.download(), seed=777)
    groundstate.solve(make_input(2, 2, (256, 256), (0.1, 0.1), 0.001, 87.0, (87.93, 91.28, 99.18), 100000,
                                 (20.0, 20.0, 400.0), (1.0, 1.0), switch_im=1, niter=300, stp=100, tol=1e-9),
                      FGPE_FP32, keep_phi=False)
    logs = {}

    def logger(tag):
        def f(ii, conv, conv_en, prec):
            if ii % 50000 == 0:
                print("%s II=%d CONV=%.3e CONV_EN=%.3e %s" % (tag, ii, conv, conv_en,
                                                             "fp32" if prec == FGPE_FP32 else "fp64"), flush=True)
        return f
    r64 = groundstate.solve(inp, FGPE_FP64, phi0=phi0, log=logger("fp64"))
    rmx = groundstate.solve(inp, FGPE_FP32, phi0=phi0, finish=True, log=logger("fp32+fp64"))
    r32 = groundstate.solve(inp, FGPE_FP32, phi0=phi0, finish=False, keep_phi=False,
                            log=None) if args.config_fp32_only else None
    rel = lambda a, b: float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))
    line = {
        "metric": "q2D spin-2 256^2 ground state to TOL 1e-9: wall s, FP32 bulk + FP64 finish",
        "value": round(rmx.wall_s, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "fp32 propagation + fp64 finish, fp64 reductions",
        "data": "synthetic: Gaussian guess + seeded 1% complex noise",
        "config": {"workload": "q2D spin-2 Rashba, 256^2, DX=DY=0.1, DT=0.001, 87Rb A0/A2/A4=87.93/91.28/99.18, "
                               "NATOMS=1e5, trap 20/20/400 Hz, GAMMAX=GAMMAY=1.0, TOL=1e-9, STP=100"},
        "fp64": r64.as_dict(),
        "fp32_fp64_finish": rmx.as_dict(),
        "speedup_vs_fp64": round(r64.wall_s / rmx.wall_s, 3),
        "error_vs_fp64": {"en_rel": abs(rmx.en - r64.en) / abs(r64.en), "mu_rel_max": rel(rmx.mu, r64.mu),
                          "rms_rel_max": rel(rmx.rms, r64.rms), "mz_abs": abs(rmx.mz - r64.mz),
                          "phi_density_rel_l2": float(np.linalg.norm(np.abs(rmx.phi) ** 2 - np.abs(r64.phi) ** 2)
                                                      / np.linalg.norm(np.abs(r64.phi) ** 2))},
    }
    if r32 is not None:
        line["fp32_only"] = r32.as_dict()
    print(json.dumps(line), flush=True)


def surrogate_bench(args):
    """Learned initial guess against the Gaussian guess: ground-state
    time-to-TOL of q2D spin-1 instances on the production grid for held-out
    parameter draws, same solver and convergence test for both.  Training
    states and both held-out solves run as batches (fgpe_batch_*); the
    time-to-TOL of an instance is its number of steps (every step of a grid
    costs the same), and the wall-clock of the two batched solves is given
    beside it."""
    from fortress_amd import fgpe
    from fortress_amd import surrogate as S
    rng = np.random.default_rng(args.seed)
    train = [S.draw(rng) for _ in range(args.train)]
    held = [S.draw(rng) for _ in range(args.held_out)]
    # a slice of the held-out draws (the evaluation of a large draw set split
    # over several runs of the same trained network)
    held = held[args.held_offset:args.held_offset + (args.held_count or len(held))]
    def log(tag):
        return lambda k, n: print(json.dumps({tag: {"steps": k, "converged": n}}), flush=True)
    sur = S.Surrogate(hidden=args.hidden, layers=4, max_batch=64, seed=args.seed)
    hist, val, spread, gen_s, train_s = [float("nan")], None, None, None, None
    if args.surrogate_load:
        sur.load(args.surrogate_load)
    else:
        t0 = time.perf_counter()
        feats, targets, tsteps = S.make_dataset(train, tol=args.train_tol, max_steps=args.max_steps,
                                                log=log("training_states"))
        gen_s = time.perf_counter() - t0
        print(json.dumps({"train_states": len(train), "generation_s": round(gen_s, 1),
                          "train_steps_median": int(np.median(tsteps))}), flush=True)
        # the last tenth of the training draws is kept out of training as a
        # validation set (same distribution as the held-out draws)
        nv = max(len(train) // 10, 1)
        t1 = time.perf_counter()
        def epoch_log(e, loss):
            if e % 25 == 0 or e == args.epochs - 1:
                print(json.dumps({"epoch": e, "loss": loss, "s": round(time.perf_counter() - t1, 1)}), flush=True)
        hist = sur.fit(feats[:-nv], targets[:-nv], epochs=args.epochs, batch=min(32, len(train) - nv), lr=args.lr,
                       lr_decay=args.lr_decay, log=epoch_log)
        train_s = time.perf_counter() - t1
        pred = sur.predict_vectors(feats[-nv:])
        tv = targets[-nv:]
        # validation error relative to the targets' spread (1 = no better than
        # the mean state), and the spread of the predictions across the
        # validation draws against that of the targets (0: the inputs are ignored)
        val = float(np.mean((pred - tv) ** 2) / np.mean((tv - tv.mean(axis=0)) ** 2))
        spread = float(np.mean(pred.std(axis=0)) / np.mean(tv.std(axis=0)))
        print(json.dumps({"train_loss_first": hist[0], "train_loss_last": hist[-1], "validation_loss": val,
                          "prediction_spread": spread, "train_s": round(train_s, 1)}), flush=True)
        del feats, targets
        if args.surrogate_save:
            sur.save(args.surrogate_save)
        if args.surrogate_train_only:
            sur.close()
            return
    params = [S.problem(d, args.n) for d in held]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        t2 = time.perf_counter()
        S.solve_batch(b, 100, args.tol, args.max_steps, log=log("plain"))
        plain_s = time.perf_counter() - t2
        rp = [b.result(i) for i in range(len(held))]
    t3 = time.perf_counter()
    guesses = [sur.guess(d, args.n) for d in held]
    infer_s = time.perf_counter() - t3
    sur.close()
    with fgpe.Batch(params) as b:
        for i, g in enumerate(guesses):
            b.upload(i, g)
        b.normalize()
        t4 = time.perf_counter()
        S.solve_batch(b, 100, args.tol, args.max_steps, log=log("surrogate"))
        sur_s = time.perf_counter() - t4
        rs = [b.result(i) for i in range(len(held))]
    sp = np.array([r.steps for r in rp], float)
    ss = np.array([r.steps for r in rs], float)
    gap = np.array([abs(s.en / p.en - 1) for s, p in zip(rs, rp)])
    both = np.array([bool(p.done) and bool(s.done) for p, s in zip(rp, rs)], dtype=bool)
    same = both & (gap <= 1e-8)
    for i in range(len(held)):
        print(json.dumps({"draw": i, "plain_steps": int(sp[i]), "surrogate_steps": int(ss[i]),
                          "energy_rel_gap": float(gap[i]), "plain_done": int(rp[i].done),
                          "surrogate_done": int(rs[i].done)}), flush=True)
    out = {
        "metric": "ground-state time-to-TOL, surrogate guess vs plain Gaussian guess",
        "value": round(float(sp.sum() / ss.sum()), 3), "unit": "x (total steps to TOL, plain / surrogate)",
        "higher_is_better": True,
        "median_step_ratio": round(float(np.median(sp / ss)), 3),
        "same_state_step_ratio": round(float(sp[same].sum() / ss[same].sum()), 3) if same.any() else None,
        "same_state_fraction": round(float(same.mean()), 3),
        "wall_clock_ratio": round(plain_s / (sur_s + infer_s), 3),
        "solve_s": {"plain": round(plain_s, 1), "surrogate": round(sur_s, 1), "inference": round(infer_s, 2)},
        "converged": {"plain": int(sum(r.done for r in rp)), "surrogate": int(sum(r.done for r in rs))},
        "max_energy_rel_gap": float(gap.max()), "median_energy_rel_gap": float(np.median(gap)),
        "config": {"workload": "q2D spin-1 imaginary time, %d^2 over a 25.6 box, DT=0.001, TOL=%g, STP=100, FP64; "
                               "draws: GAMMAX,GAMMAY~U[0,2], A0~U[99.8,103.8], A2-A0~U[-5,5] a_B, "
                               "NATOMS~logU[1e3,1e5], aspect~U[0.8,1.25], trap 20 Hz in plane, 400 Hz out"
                               % (args.n, args.tol),
                   "network": "6-%dx4-24576 (3 complex components on 64^2), BF16 MFMA, FP32 accumulation"
                              % args.hidden,
                   "train_states": len(train), "train_tol": args.train_tol, "held_out": len(held),
                   "held_offset": args.held_offset, "network_loaded": bool(args.surrogate_load),
                   "epochs": args.epochs, "lr": args.lr, "lr_decay": args.lr_decay, "max_steps": args.max_steps},
        "train_loss_last": hist[-1], "validation_loss": val, "prediction_spread": spread,
        "data_generation_s": None if gen_s is None else round(gen_s, 1),
        "training_s": None if train_s is None else round(train_s, 1),
        "dtype": "fp64 solves, bf16/fp32 network",
    }
    print(json.dumps(out), flush=True)


def surrogate_ceiling(args):
    """What a coarse-grid guess can gain at best: for args.surrogate_ceiling
    held-out-style draws, the exact coarse ground state (64^2 and 128^2 over
    the same box, FP64, TOL args.tol) spectrally interpolated onto the
    production grid, against the Gaussian guess; steps to TOL of the
    production solve, same batched loop for all."""
    from fortress_amd import fgpe
    from fortress_amd import surrogate as S
    rng = np.random.default_rng(args.seed + 1)
    draws = [S.draw(rng) for _ in range(args.surrogate_ceiling)]

    def solve(params, guesses):
        with fgpe.Batch(params) as b:
            if guesses is None:
                b.initial_guess()
            else:
                for i, g in enumerate(guesses):
                    b.upload(i, g)
                b.normalize()
            S.solve_batch(b, 100, args.tol, args.max_steps)
            return [b.result(i) for i in range(len(params))], [b.download(i) for i in range(len(params))]

    fine = [S.problem(d, args.n) for d in draws]
    rp, _ = solve(fine, None)
    sp = np.array([r.steps for r in rp], float)
    out = {"metric": "steps to TOL from the exact coarse ground state, plain / coarse-guess", "draws": len(draws),
           "config": {"workload": "q2D spin-1 imaginary time, %d^2 over a 25.6 box, DT=0.001, TOL=%g, STP=100, FP64; "
                                  "coarse states solved to the same TOL" % (args.n, args.tol)}}
    for m in (64, 128):
        _, pc = solve([S.problem(d, m) for d in draws], None)
        rg, _ = solve(fine, [S.upsample(p, args.n) for p in pc])
        sg = np.array([r.steps for r in rg], float)
        out["from_%d" % m] = {"total_ratio": round(float(sp.sum() / sg.sum()), 2),
                              "median_ratio": round(float(np.median(sp / sg)), 1),
                              "max_energy_rel_gap": float(max(abs(a.en / b.en - 1) for a, b in zip(rg, rp)))}
        print(json.dumps({"coarse": m, **out["from_%d" % m]}), flush=True)
    print(json.dumps(out), flush=True)


def batch_bench(args):
    """Batched q2D sweep: args.instances imaginary-time instances with seeded
    grid sizes from {64,128,192,256}^2 (--batch-rect: each axis drawn on its
    own), DX = 0.1, DT = 0.001; 87Rb spin 1 with A0 = 101.8 a_B, A2 - A0 ~
    U[-5,5] (--batch-spin 2: spin 2 with A0, A2, A4 = 87.93, 91.28, 99.18 a_B
    each scaled by U[0.97,1.03]), GAMMAX, GAMMAY ~ U[0,2], NATOMS ~
    logU[1e3,1e5], trap 20 Hz in-plane with aspect ~ U[0.8,1.25], FP32.
    After the autotuner, args.steps batched steps are timed; the same steps
    of a sample of the instances one at a time (fgpe_run on a single-instance
    handle each) give the unbatched rate beside it."""
    import numpy as np
    from fortress_amd import fgpe
    from fortress_amd.abi import FGPE_FP32, make_input
    rng = np.random.default_rng(args.seed)
    spin = args.batch_spin
    inps = []
    for _ in range(args.instances):
        nx = int(rng.choice([64, 128, 192, 256]))
        ny = int(rng.choice([64, 128, 192, 256])) if args.batch_rect else nx
        if spin == 1:
            a0 = 101.8
            a = (a0, a0 + rng.uniform(-5, 5), 0.0)
        else:
            a = tuple(v * rng.uniform(0.97, 1.03) for v in (87.93, 91.28, 99.18))
        inps.append(make_input(2, spin, (nx, ny), (0.1, 0.1), 1e-3, 87.0, a,
                               int(round(10 ** rng.uniform(3, 5))), (20.0, 20.0 * rng.uniform(0.8, 1.25), 400.0),
                               (rng.uniform(0, 2), rng.uniform(0, 2)), switch_im=1))
    if args.batch_solve:
        if args.batch_rect or spin != 1:
            sys.exit("--batch-solve runs the square spin-1 sweep")
        return batch_solve(args, inps)
    params = [fgpe.derive(i, FGPE_FP32)[0] for i in inps]
    sizes = [p.n[0] for p in params]
    pts = sum(p.n[0] * p.n[1] for p in params)
    ms_comp = 2 * spin + 1
    with fgpe.Batch(params) as b:
        b.initial_guess()
        ms_tune, chosen = b.tune(3)
        b.run(args.warmup)
        b.run(args.steps)
        ms = b.last_ms()
    rate = args.instances * args.steps / (ms * 1e-3)
    # one at a time: a bounded sample of the instances
    sample = list(range(len(params)))[: args.sample]
    t_single = 0.0
    for i in sample:
        with fgpe.Solver(params[i]) as s:
            s.initial_guess()
            s.normalize()
            s.run(args.warmup)
            s.sync()
            t0 = time.perf_counter()
            s.run(args.steps)
            s.sync()
            t_single += time.perf_counter() - t0
    single_rate = len(sample) * args.steps / t_single
    # HBM bytes per step: two passes, each reading and writing the FP32 state
    bytes_step = 2 * 2 * pts * ms_comp * 8
    line = {
        "metric": "batched q2D spin-%d sweep: instance-steps/s" % spin,
        "value": round(rate, 1), "unit": "instance-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 propagation, fp64 reductions",
        "data": "synthetic seeded parameter sweep, Gaussian initial guesses",
        "config": {"workload": "%d q2D spin-%d imaginary-time instances, grids from {64,128,192,256}%s, DX=0.1, "
                               "DT=0.001, autotuned schedules"
                               % (args.instances, spin, " per axis" if args.batch_rect else "^2"),
                   "instances_per_x_size": {str(n): sizes.count(n) for n in (64, 128, 192, 256)},
                   "tune_ms": [round(x, 4) for x in ms_tune], "chosen": list(chosen)},
        "step_GBps": round(bytes_step * args.steps / (ms * 1e-3) / 1e9, 1),
        "unbatched": {"value": round(single_rate, 1), "unit": "instance-steps/s",
                      "sample": "%d instances, one handle and fgpe_run each" % len(sample)},
        "speedup_vs_unbatched": round(rate / single_rate, 2),
    }
    print(json.dumps(line), flush=True)


def batch_solve(args, inps):
    """The sweep solved to TOL (FP64, imaginary time, STP = 100, cgpe.x's
    CONV/CONV_EN test, at most args.max_steps steps per instance) as one
    batch, after the autotuner; the time the same instance-steps would take
    one instance at a time comes from the unbatched FP64 rate measured on a
    sample."""
    from fortress_amd import fgpe
    from fortress_amd import surrogate as S
    from fortress_amd.abi import FGPE_FP64
    for i in inps:
        i.tol = args.batch_tol
    params = [fgpe.derive(i, FGPE_FP64)[0] for i in inps]
    sizes = [p.n[0] for p in params]
    with fgpe.Batch(params) as b:
        b.initial_guess()
        ms_tune, chosen = b.tune(3)
        t0 = time.perf_counter()
        nconv = S.solve_batch(b, 100, args.batch_tol, args.max_steps,
                              log=lambda k, n: print(json.dumps({"sweep": {"steps": k, "converged": n}}), flush=True))
        sweep_s = time.perf_counter() - t0
        steps = np.array([b.result(i).steps for i in range(len(params))])
    by_size = {n: steps[[i for i, m in enumerate(sizes) if m == n]] for n in (64, 128, 192, 256)}
    # unbatched FP64 rate per size class, one instance of each
    rate = {}
    for n in (64, 128, 192, 256):
        i = sizes.index(n)
        with fgpe.Solver(params[i]) as s:
            s.initial_guess()
            s.normalize()
            s.run(100)
            s.sync()
            t = time.perf_counter()
            s.run(args.steps)
            s.sync()
            rate[n] = args.steps / (time.perf_counter() - t)
    unbatched_s = sum(by_size[n].sum() / rate[n] for n in by_size)
    line = {
        "metric": "batched q2D spin-1 sweep solved to TOL: wall clock",
        "value": round(sweep_s, 1), "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "converged": int(nconv), "instances": len(params), "total_instance_steps": int(steps.sum()),
        "median_steps": int(np.median(steps)), "capped": int((steps >= args.max_steps).sum()),
        "unbatched_estimate_s": round(unbatched_s, 1), "speedup_vs_unbatched": round(unbatched_s / sweep_s, 2),
        "unbatched_steps_per_s": {str(n): round(r, 1) for n, r in rate.items()},
        "dtype": "fp64", "data": "synthetic seeded parameter sweep, Gaussian initial guesses",
        "config": {"workload": "%d q2D spin-1 imaginary-time instances, grids from {64,128,192,256}^2, DX=0.1, "
                               "DT=0.001, TOL=%g, STP=100, max %d steps, autotuned schedules"
                               % (len(params), args.batch_tol, args.max_steps),
                   "tune_ms": [round(x, 4) for x in ms_tune], "chosen": list(chosen)},
    }
    print(json.dumps(line), flush=True)


def gemm_bench(args):
    """TFLOP/s of the surrogate network's matrix-core product (k_gemm) on the
    products of the production 6-512x4-24576 network: inference of a batch of
    1024 draws (forward of the output layer), its weight gradient, the input
    gradient (split-K over 24576), the training step's batch-32 forward, and a
    square 4096^3 product for scale."""
    from fortress_amd import mlp
    shapes = [("fwd_out_b1024", 1024, 24576, 512, 0, 0), ("dW_out_b1024", 512, 24576, 1024, 1, 0),
              ("dH_out_b1024", 1024, 512, 24576, 0, 1), ("fwd_hidden_b1024", 1024, 512, 512, 0, 0),
              ("fwd_out_b32", 32, 24576, 512, 0, 0), ("square_4096", 4096, 4096, 4096, 0, 0)]
    res = {}
    for name, m, n, k, ta, tb in shapes:
        ms = mlp.gemm_time(m, n, k, ta, tb, reps=args.steps if args.steps < 200 else 50)
        tf = 2.0 * m * n * k / (ms * 1e-3) / 1e12
        res[name] = {"m": m, "n": n, "k": k, "ta": ta, "tb": tb, "ms": round(ms, 5), "TFLOPs": round(tf, 1)}
        print(name, res[name], flush=True)
    best = max(v["TFLOPs"] for v in res.values())
    line = {"metric": "surrogate GEMM throughput (k_gemm, v_mfma_f32_32x32x16_bf16)", "value": best,
            "unit": "TFLOP/s", "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 operands, fp32 accumulation",
            "data": "synthetic operands filled on the device",
            "roofline": {"bound": "mfma", "achieved": best, "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(best / 2500.0, 4)},
            "shapes": res}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--prep-steps", type=int, default=1500)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-refcost", action="store_true",
                    help="skip timing the reference's FFTW and ZHEEV calls beside the CPU baseline")
    ap.add_argument("--schedule", default=None, metavar="ZCHUNK,XCHUNK,STREAMS",
                    help="fixed 3D schedule instead of the autotuner (0,0,1: whole-grid launches, as profiled)")
    ap.add_argument("--tune-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 steps/s figure")
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--accuracy", default=None, metavar="STEPS,...",
                    help="instead of timing, compare FP32 with FP64 propagation at these step counts")
    ap.add_argument("--surrogate", action="store_true",
                    help="instead of the propagator throughput, the learned initial guess's time-to-TOL gain")
    ap.add_argument("--train", type=int, default=4000)
    ap.add_argument("--held-out", type=int, default=128)
    ap.add_argument("--held-offset", type=int, default=0, help="with --surrogate: first held-out draw evaluated")
    ap.add_argument("--held-count", type=int, default=0, help="with --surrogate: held-out draws evaluated (0: all)")
    ap.add_argument("--surrogate-save", default=None, help="with --surrogate: save the trained network (.npz)")
    ap.add_argument("--surrogate-load", default=None,
                    help="with --surrogate: evaluate a saved network instead of generating data and training")
    ap.add_argument("--surrogate-train-only", action="store_true",
                    help="with --surrogate: stop after training (and saving)")
    ap.add_argument("--train-tol", type=float, default=1e-9)
    ap.add_argument("--surrogate-ceiling", type=int, default=0,
                    help="instead of the 3D propagator, the gain of exact coarse ground states as guesses (draws)")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--epochs", type=int, default=1200)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--lr-decay", type=float, default=0.997)
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--batch", action="store_true",
                    help="instead of the 3D propagator, the batched q2D parameter sweep")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--batch-spin", type=int, default=1, choices=(1, 2), help="with --batch: spin of the sweep")
    ap.add_argument("--batch-rect", action="store_true", help="with --batch: draw each grid axis on its own")
    ap.add_argument("--sample", type=int, default=16)
    ap.add_argument("--batch-solve", action="store_true",
                    help="with --batch: solve the sweep to --batch-tol in FP64 instead of timing steps")
    ap.add_argument("--batch-tol", type=float, default=1e-8)
    ap.add_argument("--config1", action="store_true",
                    help="instead of the 3D propagator, the q1D spin-1 ground state + real-time run against the oracle")
    ap.add_argument("--config2", action="store_true",
                    help="instead of the 3D propagator, the q2D spin-2 ground state to TOL: FP64 vs FP32 + FP64 finish")
    ap.add_argument("--config-niter", type=int, default=20000)
    ap.add_argument("--config-rt-steps", type=int, default=5000)
    ap.add_argument("--config-fp32-only", action="store_true", help="with --config2: also FP32 without the finish")
    ap.add_argument("--gemm", action="store_true",
                    help="instead of the 3D propagator, the surrogate network's GEMM TFLOP/s")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py runs on exactly one GPU: this build has no multi-GPU path "
                 "(the slab decomposition over RCCL is design-only); use --gpus 1")

    if args.accuracy:
        return accuracy(args)
    if args.batch:
        return batch_bench(args)
    if args.surrogate:
        return surrogate_bench(args)
    if args.surrogate_ceiling:
        return surrogate_ceiling(args)
    if args.config1:
        return config1_bench(args)
    if args.config2:
        return config2_bench(args)
    if args.gemm:
        return gemm_bench(args)

    from fortress_amd import fgpe
    from fortress_amd.abi import FGPE_FP32, FGPE_FP64

    stp = 100
    # --- initial state: imaginary-time relaxation of the Gaussian guess + 1% noise
    p_im, _ = fgpe.derive(workload(args.n, 1, 0.005), FGPE_FP32)
    with fgpe.Solver(p_im) as s:
        s.initial_guess()
        s.normalize()
        s.run(args.prep_steps)
        prep_en = s.observables()[0]
        phi_g = s.download()
    phi0 = noisy(phi_g, seed=1234)
    del phi_g

    p, _ = fgpe.derive(workload(args.n, 0, 5e-4), FGPE_FP32)
    s = fgpe.Solver(p)
    s.upload(phi0)
    s.observables()
    # pencil schedule: the autotuner on this state (restored afterwards), or a fixed one
    tuned = None
    if args.schedule:
        s.set_schedule(*[int(v) for v in args.schedule.split(",")])
    else:
        tune_ms, _ = s.tune_schedule(args.tune_steps)
        tuned = [round(float(v), 4) for v in tune_ms]
    schedule = s.schedule()
    # one ENERGY + RAD + NORMC evaluation, timed on its own
    s.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        s.observables()
    t_obs = (time.perf_counter() - t0) / 3

    total = [0, 0]  # steps done, observables calls in the timed window

    def advance(k, timed):
        done = 0
        while done < k:
            c = min(stp - total[0] % stp, k - done)
            s.run(c)
            done += c
            total[0] += c
            if total[0] % stp == 0:
                s.observables()
                total[1] += timed

    advance(args.warmup, 0)
    s.sync()
    t0 = time.perf_counter()
    advance(args.steps, 1)
    s.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # observables once per STP steps: the window holds total[1] calls; the
    # difference to steps/STP calls is added at the separately timed cost
    amortised = elapsed + t_obs * (args.steps / stp - total[1])
    steps_per_s = args.steps / amortised
    en, mu, mz, rms = s.observables()

    # --- per-kernel times (HIP events on the library stream, whole-grid launches)
    passes = s.pass_times(50)
    s.close()
    fp64 = None
    if not args.no_fp64:
        # the reference's arithmetic: the same workload propagated in FP64
        p64, _ = fgpe.derive(workload(args.n, 0, 5e-4), FGPE_FP64)
        with fgpe.Solver(p64) as s64:
            s64.upload(phi0)
            s64.tune_schedule(5)
            s64.run(10)
            s64.sync()
            t0 = time.perf_counter()
            s64.run(args.fp64_steps)
            s64.sync()
            dt64 = time.perf_counter() - t0
            fp64 = {"steps_per_s": round(args.fp64_steps / dt64, 2), "ms_per_step": round(dt64 / args.fp64_steps * 1e3, 4),
                    "steps": args.fp64_steps, "schedule": list(s64.schedule()),
                    "step_roofline_frac": round(2 * 4 * args.n ** 3 * 5 * 16 / (dt64 / args.fp64_steps) / 1e9 / PEAK_GBS,
                                                4)}
    S = args.n ** 3 * 5 * 8
    b_pass = 2 * S  # algorithmic bytes per launch of every line pass
    dom = max((k for k in passes if k in ("x", "y", "z", "k")), key=lambda k: passes[k][0])
    dom_ms = passes[dom][0]
    achieved = b_pass / (dom_ms * 1e-3) / 1e9
    kernel_names = {"x": "k_xpass (inverse-x + SE + SP + NORMC + forward-x, all components)",
                    "y": "k_ypass (forward-y + kx/ky KE, one component)",
                    "z": "k_zpass (forward-z + kz KE + inverse-z, one component)",
                    "k": "k_dpass (SOC + inverse-y + 1/N, all components)"}
    per_pass = {k: {"ms": round(v[0], 4), "launches": v[1],
                    "GBps": round(b_pass / (v[0] * 1e-3) / 1e9, 1) if k in ("x", "y", "z", "k") else None}
                for k, v in passes.items()}
    b_alg = 2 * 4 * S  # four read+write passes per step
    line = {
        "metric": METRIC,
        "value": round(steps_per_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / steps_per_s, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 propagation, fp64 reductions",
        "data": "synthetic: imaginary-time relaxed Gaussian guess + seeded 1% complex noise",
        "config": {"workload": "3D spin-2 real-time split-step, %d^3, DX=0.1, DT=0.0005, 87Rb spin-2, N=1e5, "
                               "trap 20/20/20 Hz, gamma=(0.5,0.5,0.5); observables (ENERGY, RAD, NORMC) once per "
                               "STP=100 steps, amortised into value" % args.n,
                   "grid": [args.n] * 3, "components": 5, "stp": stp, "prep": {"imag_steps": args.prep_steps,
                                                                              "dt": 0.005, "energy": prep_en}},
        "timing": {"timed_region_s": round(elapsed, 4), "observables_in_window": total[1],
                   "observables_ms": round(t_obs * 1e3, 3), "propagation_steps_per_s": round(args.steps / elapsed, 2),
                   "schedule": list(schedule), "tuner_ms_per_step": tuned},
        "roofline": {"bound": "hbm", "kernel": kernel_names[dom], "achieved": round(achieved, 1),
                     "peak": PEAK_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": TRAFFIC_256.get(dom) if args.n == 256 else None,
                     "bytes_per_launch": b_pass},
        "step_roofline": {"bytes_per_step": b_alg, "achieved_GBps": round(b_alg * steps_per_s / 1e9, 1),
                          "frac": round(b_alg * steps_per_s / 1e9 / PEAK_GBS, 4)},
        "passes": per_pass,
        "fp64": fp64,
        "final_energy": en,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(phi0, args.n, args.cpu_steps)
        if not args.no_refcost:
            line["cpu_baseline"]["reference_cost_model"] = reference_cost(args.n, steps_per_s)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
