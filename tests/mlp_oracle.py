"""CPU restatement of the surrogate network (fortress_amd/csrc/fgpe_mlp.hip),
for the gradient checks: y = act(x W + b) per layer (ReLU between layers,
none after the last), mean-squared error, and its backward pass.

With bf16=True the operands of every product are rounded to BF16 exactly
where the HIP kernels round them (network input, hidden activations, weights,
the loss gradient and the masked hidden gradients), and all arithmetic is
FP64; the kernels accumulate in FP32, so the two agree to FP32 rounding.
"""
import numpy as np


def bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    out, o = [], 0
    for k, n in zip(dims[:-1], dims[1:]):
        w = params[o:o + k * n].reshape(k, n).astype(np.float64)
        o += k * n
        b = params[o:o + n].astype(np.float64)
        o += n
        out.append((w, b))
    return out


def _round(bf):
    if bf:
        return lambda a: bf16(np.asarray(a, dtype=np.float32))
    return lambda a: np.asarray(a, dtype=np.float64)


def forward(params, dims, x, bf=True):
    r = _round(bf)
    layers = unpack(params, dims)
    hs = [np.asarray(x, dtype=np.float64)]
    hb = [r(x)]
    for i, (w, b) in enumerate(layers):
        z = hb[-1] @ r(w) + b
        last = i == len(layers) - 1
        h = z if last else np.maximum(z, 0.0)
        hs.append(h)
        hb.append(r(h))
    return hs, hb


def loss_and_grad(params, dims, x, t, bf=True):
    r = _round(bf)
    layers = unpack(params, dims)
    hs, hb = forward(params, dims, x, bf)
    y = hs[-1]
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = 2.0 * (y - t) / n
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        dzb = r(dz)
        gw = hb[i].T @ dzb
        gb = dz.sum(axis=0)
        grads[i] = (gw, gb)
        if i == 0:
            break
        dh = dzb @ r(w).T
        dz = dh * (hs[i] > 0.0)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat
