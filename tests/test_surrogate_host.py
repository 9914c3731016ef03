"""Host-side pieces of the learned initial guess (fortress_amd/surrogate.py)."""
import numpy as np

from fortress_amd import surrogate as S


def test_upsample_interpolates_band_limited_data_exactly():
    m, n, L = 64, 256, S.BOX
    xc = -L / 2 + L * np.arange(m) / m
    xf = -L / 2 + L * np.arange(n) / n

    def f(x, y):
        k = 2 * np.pi / L
        return (np.cos(3 * k * x[None, :]) * np.sin(5 * k * y[:, None]) + 0.3j * np.cos(31 * k * x[None, :])
                + np.exp(1j * 7 * k * (x[None, :] + 2 * y[:, None])))

    coarse = np.stack([f(xc, xc), 2 * f(xc, xc), f(xc, xc).conj()])[:, None]
    fine = np.stack([f(xf, xf), 2 * f(xf, xf), f(xf, xf).conj()])[:, None]
    up = S.upsample(coarse, n)
    assert np.abs(up - fine).max() < 1e-10
    assert np.abs(up[..., ::4, ::4] - coarse).max() < 1e-12


def test_vector_roundtrip_and_features():
    rng = np.random.default_rng(0)
    phi = rng.standard_normal((3, 1, 64, 64)) + 1j * rng.standard_normal((3, 1, 64, 64))
    assert np.array_equal(S.from_vector(S.to_vector(phi).astype(np.float64)), phi.astype(np.complex64).astype(np.complex128))
    d = S.draw(rng)
    f = S.features(d)
    assert f.shape == (6,) and np.all(np.abs(f) <= 1.01)
    p = S.problem(d, 256)
    assert p.n[0] == 256 and abs(p.d[0] - 0.1) < 1e-15 and abs(p.alpha[1] - d["aspect"]) < 1e-12


def test_mlp_host_rejects_mismatched_rows():
    """The library copies rows * width values from the host buffers, so the
    wrapper checks the shapes before any call (no GPU needed here)."""
    import pytest

    from fortress_amd.fgpe import FgpeError
    from fortress_amd.mlp import MLP

    net = object.__new__(MLP)
    net.dims, net.max_batch = [6, 8, 10], 4
    x, t = np.zeros((3, 6)), np.zeros((3, 10))
    assert net._rows(x, t)[0].shape == (3, 6)
    for bad_x, bad_t in ((np.zeros((3, 5)), t), (x, np.zeros((2, 10))), (x, np.zeros((3, 9))), (np.zeros(18), t)):
        with pytest.raises(FgpeError):
            net._rows(bad_x, bad_t)
    with pytest.raises(FgpeError):
        net.gradients(np.zeros((5, 6)), np.zeros((5, 10)))
