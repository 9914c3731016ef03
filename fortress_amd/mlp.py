"""ctypes binding of the surrogate network (fgpe_mlp_* in include/fgpe.h).

Parameters are packed layer by layer as W ([K][N], row major) then b ([N]).
There is no CPU fallback: every call runs the library's HIP kernels.
"""
import ctypes as C

import numpy as np

from .fgpe import FgpeError, _check, lib

_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, i32, i64, fp, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.fgpe_mlp_create.argtypes = [ip, i32, i32, C.c_uint64, C.POINTER(vp)]
    L.fgpe_mlp_destroy.argtypes = [vp]
    L.fgpe_mlp_destroy.restype = None
    L.fgpe_mlp_param_count.argtypes = [vp, C.POINTER(i64)]
    L.fgpe_mlp_get_params.argtypes = [vp, fp]
    L.fgpe_mlp_set_params.argtypes = [vp, fp]
    L.fgpe_mlp_forward.argtypes = [vp, fp, i32, fp]
    L.fgpe_mlp_gradients.argtypes = [vp, fp, fp, i32, fp, dp]
    L.fgpe_mlp_set_dataset.argtypes = [vp, fp, fp, i32]
    L.fgpe_mlp_train_step.argtypes = [vp, ip, i32, C.c_float, dp]
    L.fgpe_mlp_gemm_test.argtypes = [i32, i32, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), fp, i32, i32, i32]
    L.fgpe_mlp_gemm_time.argtypes = [i32, i32, i32, i32, i32, i32, C.POINTER(C.c_double)]
    for n in ("fgpe_mlp_create", "fgpe_mlp_param_count", "fgpe_mlp_get_params", "fgpe_mlp_set_params",
              "fgpe_mlp_forward", "fgpe_mlp_gradients", "fgpe_mlp_set_dataset", "fgpe_mlp_train_step",
              "fgpe_mlp_gemm_test", "fgpe_mlp_gemm_time"):
        getattr(L, n).restype = C.c_int
    _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def to_bf16_bits(x):
    """Round FP32 values to BF16 (nearest even) and return the 16-bit patterns."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def gemm(a_bits, b_bits, m, n, k, ta=False, tb=False):
    """C = op(A) op(B) on the matrix cores (A given [K][M] if ta, B [N][K] if tb)."""
    L = _bind()
    a = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b = np.ascontiguousarray(b_bits, dtype=np.uint16)
    c = np.empty((m, n), dtype=np.float32)
    _check(L.fgpe_mlp_gemm_test(int(ta), int(tb), a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                b.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(c), m, n, k), "fgpe_mlp_gemm_test")
    return c


def gemm_time(m, n, k, ta=False, tb=False, reps=20):
    """Milliseconds per product C[m][n] = op(A) op(B) on device-resident BF16
    operands (the mean of reps launches after one untimed launch)."""
    L = _bind()
    ms = C.c_double()
    _check(L.fgpe_mlp_gemm_time(int(ta), int(tb), int(m), int(n), int(k), int(reps), C.byref(ms)),
           "fgpe_mlp_gemm_time")
    return ms.value


class MLP:
    def __init__(self, dims, max_batch, seed=0):
        L = _bind()
        self.dims = [int(d) for d in dims]
        arr = (C.c_int32 * len(self.dims))(*self.dims)
        h = C.c_void_p()
        _check(L.fgpe_mlp_create(arr, len(self.dims) - 1, int(max_batch), C.c_uint64(seed), C.byref(h)),
               "fgpe_mlp_create")
        self.h = h
        self.max_batch = int(max_batch)
        n = C.c_int64()
        _check(L.fgpe_mlp_param_count(self.h, C.byref(n)), "fgpe_mlp_param_count")
        self.nparams = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().fgpe_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        w = np.empty(self.nparams, dtype=np.float32)
        _check(lib().fgpe_mlp_get_params(self.h, _fp(w)), "fgpe_mlp_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparams:
            raise FgpeError("expected %d parameters, got %d" % (self.nparams, w.size))
        _check(lib().fgpe_mlp_set_params(self.h, _fp(w)), "fgpe_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.dims[0])
        out = np.empty((x.shape[0], self.dims[-1]), dtype=np.float32)
        for s in range(0, x.shape[0], self.max_batch):
            xb = np.ascontiguousarray(x[s:s + self.max_batch])
            yb = np.empty((xb.shape[0], self.dims[-1]), dtype=np.float32)
            _check(lib().fgpe_mlp_forward(self.h, _fp(xb), xb.shape[0], _fp(yb)), "fgpe_mlp_forward")
            out[s:s + xb.shape[0]] = yb
        return out

    def _rows(self, x, t):
        """x as [rows][in] and t as [rows][out], or FgpeError: the library
        copies rows * width values from each host buffer."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.dims[0]:
            raise FgpeError("inputs must have shape (rows, %d), got %s" % (self.dims[0], x.shape))
        if t.shape != (x.shape[0], self.dims[-1]):
            raise FgpeError("targets must have shape (%d, %d), got %s" % (x.shape[0], self.dims[-1], t.shape))
        return x, t

    def gradients(self, x, t):
        x, t = self._rows(x, t)
        if x.shape[0] > self.max_batch:
            raise FgpeError("batch of %d rows exceeds max_batch %d" % (x.shape[0], self.max_batch))
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        _check(lib().fgpe_mlp_gradients(self.h, _fp(x), _fp(t), x.shape[0], _fp(g), C.byref(loss)),
               "fgpe_mlp_gradients")
        return g, loss.value

    def set_dataset(self, x, y):
        self._x, self._y = self._rows(x, y)
        _check(lib().fgpe_mlp_set_dataset(self.h, _fp(self._x), _fp(self._y), self._x.shape[0]),
               "fgpe_mlp_set_dataset")
        self.rows = self._x.shape[0]

    def train_step(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        _check(lib().fgpe_mlp_train_step(self.h, idx.ctypes.data_as(C.POINTER(C.c_int32)), idx.size, float(lr),
                                         C.byref(loss)), "fgpe_mlp_train_step")
        return loss.value

    def train(self, epochs, batch, lr, seed=0, lr_decay=1.0, log=None):
        """Shuffled minibatch epochs over the device dataset; returns the mean
        loss of each epoch (log(epoch, loss) is called after each)."""
        rng = np.random.default_rng(seed)
        hist = []
        for e in range(epochs):
            perm = rng.permutation(self.rows)
            tot, nb = 0.0, 0
            for s in range(0, self.rows - batch + 1, batch):
                tot += self.train_step(perm[s:s + batch], lr)
                nb += 1
            hist.append(tot / max(nb, 1))
            if log:
                log(e, hist[-1])
            lr *= lr_decay
        return hist
