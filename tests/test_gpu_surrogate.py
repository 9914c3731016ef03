"""What the learned initial guess gains over the plain Gaussian guess.

A small network (the production architecture with narrower hidden layers) is
trained on ground states the solver produces for seeded parameter draws.  On
held-out draws its guess (after NORMT) must lie closer to the converged
ground state than the Gaussian guess does, and the imaginary-time solve from
it must reach TOL in fewer steps in total.  Closeness is measured by
quantities that do not depend on the global phase of the spinor (the ground
state is fixed only up to it): the relative L2 distance of the component
densities |phi_m|^2, and the infidelity 1 - |<guess, phi_gs>| of the
normalised spinors, which is printed.  The benchmark (bench.py --surrogate)
reports the time-to-TOL ratio on the production grid.
"""
import numpy as np
import pytest

from fortress_amd import fgpe
from fortress_amd import surrogate as S

pytestmark = pytest.mark.gpu


def normalised(params, phi):
    with fgpe.Solver(params) as s:
        if phi is None:
            s.initial_guess()
        else:
            s.upload(phi)
        s.normalize()
        return s.download()


def test_surrogate_guess_is_closer_and_converges_sooner():
    rng = np.random.default_rng(77)
    train = [S.draw(rng) for _ in range(24)]
    held = [S.draw(rng) for _ in range(3)]
    feats, targets = [], []
    for d in train:
        phi, _, _, _ = S.solve(S.problem(d, S.COARSE), tol=1e-5, max_steps=200000)
        feats.append(S.features(d))
        targets.append(S.to_vector(phi))
    sur = S.Surrogate(hidden=256, layers=4, max_batch=32, seed=3)
    hist = sur.fit(np.array(feats), np.array(targets), epochs=200, batch=8, lr=1e-3)
    print("train loss", hist[0], "->", hist[-1], flush=True)
    assert hist[-1] < 0.2 * hist[0]
    d_plain, d_sur, steps_plain, steps_sur = [], [], [], []
    for d in held:
        p = S.problem(d, S.COARSE)
        gs, kp, enp, _ = S.solve(p, tol=1e-7, max_steps=200000)
        g0 = normalised(p, None)
        g1 = normalised(p, sur.guess(d, S.COARSE))
        dist = lambda g: float(np.linalg.norm(np.abs(g) ** 2 - np.abs(gs) ** 2) / np.linalg.norm(np.abs(gs) ** 2))
        infid = lambda g: 1.0 - abs(np.vdot(g, gs)) / (np.linalg.norm(g) * np.linalg.norm(gs))
        d_plain.append(dist(g0))
        d_sur.append(dist(g1))
        _, ks, ens, _ = S.solve(p, g1, tol=1e-7, max_steps=200000)
        steps_plain.append(kp)
        steps_sur.append(ks)
        print("held-out draw: density distance plain %.3f surrogate %.3f; infidelity plain %.3f surrogate %.3f; "
              "steps to TOL plain %d surrogate %d; energy gap %.1e"
              % (d_plain[-1], d_sur[-1], infid(g0), infid(g1), kp, ks, abs(ens / enp - 1)), flush=True)
    sur.close()
    assert np.mean(d_sur) < np.mean(d_plain)
    assert sum(steps_sur) < sum(steps_plain)


def test_saved_network_predicts_the_same(tmp_path):
    """save() and load() carry the trained state between processes: a fresh
    network loaded from the file predicts bit for bit what the saved one
    does, and a network of another shape is refused."""
    rng = np.random.default_rng(3)
    draws = [S.draw(rng) for _ in range(4)]
    feats = np.array([S.features(d) for d in draws])
    a = S.Surrogate(hidden=32, layers=4, max_batch=8, seed=5)
    a.mean = rng.standard_normal(6 * S.COARSE * S.COARSE)
    a.sd = rng.uniform(0.5, 2.0, 6 * S.COARSE * S.COARSE)
    path = str(tmp_path / "net.npz")
    a.save(path)
    want = a.predict_vectors(feats)
    a.close()
    b = S.Surrogate(hidden=32, layers=4, max_batch=8, seed=6)
    b.load(path)
    got = b.predict_vectors(feats)
    b.close()
    assert np.array_equal(got, want)
    c = S.Surrogate(hidden=16, layers=4, max_batch=8, seed=6)
    with pytest.raises(ValueError):
        c.load(path)
    c.close()
