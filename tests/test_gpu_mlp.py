"""The surrogate network's HIP kernels: the matrix-core product against numpy,
gradients against the CPU restatement (tests/mlp_oracle.py, itself pinned by
finite differences), and training that fits a smooth map."""
import numpy as np
import pytest

import mlp_oracle as M
from fortress_amd import mlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(70, 130, 50), (64, 512, 4096), (96, 1000, 6), (32, 512, 24576)])
def test_gemm_matches_numpy(ta, tb, m, n, k):
    rng = np.random.default_rng(m + n + k + 2 * ta + tb)
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32)
    ab, bb = mlp.to_bf16_bits(a), mlp.to_bf16_bits(b)
    a64, b64 = mlp.from_bf16_bits(ab).astype(np.float64), mlp.from_bf16_bits(bb).astype(np.float64)
    got = mlp.gemm(ab.T.copy() if ta else ab, bb.T.copy() if tb else bb, m, n, k, bool(ta), bool(tb))
    ref = a64 @ b64
    scale = np.abs(a64) @ np.abs(b64)
    err = np.abs(got - ref) / scale
    print("ta", ta, "tb", tb, "shape", (m, n, k), "max err/scale", err.max(), flush=True)
    assert err.max() < 2e-6


@pytest.mark.parametrize("dims", [[6, 64, 64, 96], [6, 64, 64, 24576], [6, 512, 512, 512, 512, 24576]])
def test_gradients_match_cpu_restatement(dims):
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (16, dims[0])).astype(np.float32)
    t = rng.standard_normal((16, dims[-1])).astype(np.float32) * 0.5
    with mlp.MLP(dims, max_batch=32, seed=11) as net:
        p = net.get_params()
        g, loss = net.gradients(x, t)
        y = net.forward(x)
    hs, hb = M.forward(p, dims, x)
    # FP32 accumulation (GPU) and FP64 (restatement) can round a hidden value
    # to different BF16 neighbours; each such flip moves an output by at most
    # one BF16 spacing (2^-8 relative) of |h| |W|.  So: relative L2 1e-5
    # overall and at most 2^-7 of sum_j |h_j| |W_jk| at any output.
    w_last = M.unpack(p, dims)[-1][0]
    scale = np.abs(hb[-2]) @ np.abs(w_last)
    dy = np.abs(y - hs[-1])
    yerr = np.linalg.norm(y - hs[-1]) / np.linalg.norm(hs[-1])
    print("forward rel l2", yerr, "max |dy|", dy.max(), "max |dy|/scale", (dy / scale).max(), flush=True)
    assert yerr < 1e-5
    assert np.all(dy <= 2.0**-7 * scale + 1e-7)
    ref_loss, ref_g = M.loss_and_grad(p, dims, x, t)
    print("loss", loss, ref_loss, "grad rel l2", np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g), flush=True)
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    assert np.linalg.norm(g - ref_g) <= 2e-3 * np.linalg.norm(ref_g)
    # per layer too, so a wrong transpose in one layer cannot hide
    o = 0
    for k, n in zip(dims[:-1], dims[1:]):
        sl = slice(o, o + k * n + n)
        assert np.linalg.norm(g[sl] - ref_g[sl]) <= 5e-3 * np.linalg.norm(ref_g[sl])
        o += k * n + n


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(1)
    dims = [6, 128, 128, 64]
    a = rng.standard_normal((6, 64))
    x = rng.uniform(-1, 1, (512, 6)).astype(np.float32)
    y = np.sin(x @ a).astype(np.float32)
    with mlp.MLP(dims, max_batch=64, seed=5) as net:
        net.set_dataset(x, y)
        hist = net.train(epochs=60, batch=64, lr=2e-3, seed=2)
    print("loss first/last epoch", hist[0], hist[-1], flush=True)
    assert hist[-1] < 0.1 * hist[0]
