"""ctypes binding of libfgpe (include/fgpe.h).

``Solver`` owns one fgpe handle.  Wave functions are numpy complex128 arrays
of shape (ms, nz, ny, nx), which is the reference's PHI(NX,NY,NZ,MS) Fortran
layout read as C order.  There is no CPU fallback: loading fails loudly when
the library is missing, and every compute call raises when the HIP runtime
reports an error.
"""
import ctypes as C
import os

import numpy as np

from .abi import FGPE_FP32, FGPE_FP64, FgpeInput, FgpeParams  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfgpe.so")

_lib = None

EXPORTS = (
    "fgpe_derive",
    "fgpe_create",
    "fgpe_destroy",
    "fgpe_initial_guess",
    "fgpe_upload",
    "fgpe_download",
    "fgpe_normalize",
    "fgpe_run",
    "fgpe_run_noalloc",
    "fgpe_set_schedule",
    "fgpe_get_schedule",
    "fgpe_tune_schedule",
    "fgpe_last_norms",
    "fgpe_normc",
    "fgpe_observables",
    "fgpe_keep_old",
    "fgpe_conv_max",
    "fgpe_center_abs",
    "fgpe_sync",
    "fgpe_checkpoint_save",
    "fgpe_checkpoint_load",
    "fgpe_checkpoint_info",
    "fgpe_write_dump",
    "fgpe_format_f",
    "fgpe_read_initial_sol",
    "fgpe_last_run_ms",
    "fgpe_pass_times",
    "fgpe_last_error",
    "fgpe_batch_create",
    "fgpe_batch_destroy",
    "fgpe_batch_initial_guess",
    "fgpe_batch_normalize",
    "fgpe_batch_upload",
    "fgpe_batch_download",
    "fgpe_batch_run",
    "fgpe_batch_solve",
    "fgpe_batch_result_get",
    "fgpe_batch_tune",
    "fgpe_batch_last_ms",
    "fgpe_mlp_create",
    "fgpe_mlp_destroy",
    "fgpe_mlp_param_count",
    "fgpe_mlp_get_params",
    "fgpe_mlp_set_params",
    "fgpe_mlp_forward",
    "fgpe_mlp_gradients",
    "fgpe_mlp_set_dataset",
    "fgpe_mlp_train_step",
    "fgpe_mlp_gemm_test",
    "fgpe_mlp_gemm_time",
)

BATCH_NSCHED = 8
BATCH_SCHEDULES = ((64, 8), (64, 4), (128, 8), (128, 16), (192, 12), (192, 6), (256, 8), (256, 16))


class BatchResult(C.Structure):
    _fields_ = [("steps", C.c_int64), ("done", C.c_int32), ("err", C.c_int32), ("en", C.c_double),
                ("mz", C.c_double), ("conv", C.c_double), ("conv_en", C.c_double), ("maxdiff", C.c_double),
                ("mu", C.c_double * 5), ("rms", C.c_double * 5), ("norm", C.c_double * 5)]


class LoopState(C.Structure):
    """fgpe_loop_state: the caller's loop state stored in a checkpoint."""
    _fields_ = [("ii", C.c_int64), ("en_old", C.c_double), ("user", C.c_double * 6)]


class CheckpointInfo(C.Structure):
    _fields_ = [("version", C.c_int32), ("dim", C.c_int32), ("spin", C.c_int32), ("precision", C.c_int32),
                ("imag", C.c_int32), ("n", C.c_int32 * 3), ("dt", C.c_double), ("ii", C.c_int64),
                ("en_old", C.c_double), ("payload_bytes", C.c_int64)]


class FgpeError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FgpeError("libfgpe.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, dp, i32, i64 = C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.c_int64
        L.fgpe_derive.argtypes = [C.POINTER(FgpeInput), i32, C.POINTER(FgpeParams), dp]
        L.fgpe_create.argtypes = [C.POINTER(FgpeParams), C.POINTER(vp)]
        L.fgpe_destroy.argtypes = [vp]
        L.fgpe_destroy.restype = None
        for name in ("fgpe_initial_guess", "fgpe_keep_old", "fgpe_sync"):
            getattr(L, name).argtypes = [vp]
        L.fgpe_upload.argtypes = [vp, dp]
        L.fgpe_download.argtypes = [vp, dp]
        L.fgpe_normalize.argtypes = [vp, dp]
        L.fgpe_run.argtypes = [vp, i64]
        L.fgpe_run_noalloc.argtypes = [vp, i64]
        L.fgpe_set_schedule.argtypes = [vp, i32, i32, i32]
        L.fgpe_get_schedule.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        L.fgpe_tune_schedule.argtypes = [vp, i32, dp, C.POINTER(i32)]
        L.fgpe_last_norms.argtypes = [vp, dp]
        L.fgpe_normc.argtypes = [vp, dp]
        L.fgpe_observables.argtypes = [vp, dp, dp, dp, dp]
        L.fgpe_conv_max.argtypes = [vp, dp]
        L.fgpe_center_abs.argtypes = [vp, dp]
        L.fgpe_last_run_ms.argtypes = [vp, dp]
        L.fgpe_checkpoint_save.argtypes = [vp, C.c_char_p, C.POINTER(LoopState)]
        L.fgpe_checkpoint_load.argtypes = [vp, C.c_char_p, C.POINTER(LoopState)]
        L.fgpe_checkpoint_info.argtypes = [C.c_char_p, C.POINTER(CheckpointInfo)]
        L.fgpe_pass_times.argtypes = [vp, i64, dp, C.POINTER(C.c_int32)]
        L.fgpe_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.fgpe_batch_create.argtypes = [C.POINTER(FgpeParams), i32, C.POINTER(vp)]
        L.fgpe_batch_destroy.argtypes = [vp]
        L.fgpe_batch_destroy.restype = None
        L.fgpe_batch_initial_guess.argtypes = [vp]
        L.fgpe_batch_normalize.argtypes = [vp]
        L.fgpe_batch_upload.argtypes = [vp, i32, dp]
        L.fgpe_batch_download.argtypes = [vp, i32, dp]
        L.fgpe_batch_run.argtypes = [vp, i64]
        L.fgpe_batch_solve.argtypes = [vp, i32, C.c_double, i64, C.POINTER(C.c_int32)]
        L.fgpe_batch_result_get.argtypes = [vp, i32, C.POINTER(BatchResult)]
        L.fgpe_batch_tune.argtypes = [vp, i32, dp, C.POINTER(C.c_int32)]
        L.fgpe_batch_last_ms.argtypes = [vp, dp]
        for name in EXPORTS:
            if getattr(L, name).restype is not None and name not in ("fgpe_destroy", "fgpe_batch_destroy"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def checkpoint_info(path):
    """Header of a checkpoint file, its checksum verified (no device needed)."""
    info = CheckpointInfo()
    _check(lib().fgpe_checkpoint_info(os.fsencode(path), C.byref(info)), "fgpe_checkpoint_info")
    return dict(version=info.version, dim=info.dim, spin=info.spin, precision=info.precision, imag=info.imag,
                n=tuple(info.n), dt=info.dt, ii=info.ii, en_old=info.en_old, payload_bytes=info.payload_bytes)


def last_error():
    buf = C.create_string_buffer(512)
    lib().fgpe_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise FgpeError("%s failed (%d): %s" % (what, rc, last_error()))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def derive(inp, precision=FGPE_FP64):
    p = FgpeParams()
    aosc = C.c_double()
    _check(lib().fgpe_derive(C.byref(inp), precision, C.byref(p), C.byref(aosc)), "fgpe_derive")
    return p, aosc.value


class Solver:
    """One propagator instance on the current HIP device."""

    def __init__(self, params):
        self.p = params
        self.ms = 2 * params.spin + 1
        self.shape = (self.ms, params.n[2] if params.dim == 3 else 1, params.n[1] if params.dim >= 2 else 1,
                      params.n[0])
        h = C.c_void_p()
        _check(lib().fgpe_create(C.byref(params), C.byref(h)), "fgpe_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fgpe_destroy(self.h)
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

    def initial_guess(self):
        _check(lib().fgpe_initial_guess(self.h), "fgpe_initial_guess")

    def upload(self, phi):
        a = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128).reshape(self.shape))
        _check(lib().fgpe_upload(self.h, _dp(a.view(np.float64))), "fgpe_upload")

    def download(self):
        a = np.empty(self.shape, dtype=np.complex128)
        _check(lib().fgpe_download(self.h, _dp(a.view(np.float64))), "fgpe_download")
        return a

    def normalize(self):
        t = C.c_double()
        _check(lib().fgpe_normalize(self.h, C.byref(t)), "fgpe_normalize")
        return t.value

    def run(self, nsteps):
        _check(lib().fgpe_run(self.h, int(nsteps)), "fgpe_run")

    def run_noalloc(self, nsteps):
        """fgpe_run with its scratch in the library's fixed device array (no
        device allocation); calls overlapping on one GPU race."""
        _check(lib().fgpe_run_noalloc(self.h, int(nsteps)), "fgpe_run_noalloc")

    def set_schedule(self, zchunk, xchunk=0, nstreams=1):
        """3D pencil schedule: z-planes per chunk of the x-line/forward-y
        phase, x-columns per chunk of the z-line/inverse-y phase, streams."""
        _check(lib().fgpe_set_schedule(self.h, int(zchunk), int(xchunk), int(nstreams)), "fgpe_set_schedule")

    def schedule(self):
        z, x, n = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().fgpe_get_schedule(self.h, C.byref(z), C.byref(x), C.byref(n)), "fgpe_get_schedule")
        return z.value, x.value, n.value

    NSCHED_CAND = 8

    def tune_schedule(self, trial_steps=20):
        """Time the candidate 3D schedules on this state (restored afterwards)
        and keep the fastest: (ms per step of each candidate, chosen schedule)."""
        ms = np.zeros(self.NSCHED_CAND)
        chosen = (C.c_int32 * 3)()
        _check(lib().fgpe_tune_schedule(self.h, int(trial_steps), _dp(ms), chosen), "fgpe_tune_schedule")
        return ms, tuple(chosen)

    def last_norms(self):
        a = np.empty(self.ms)
        _check(lib().fgpe_last_norms(self.h, _dp(a)), "fgpe_last_norms")
        return a

    def normc(self):
        a = np.empty(self.ms)
        _check(lib().fgpe_normc(self.h, _dp(a)), "fgpe_normc")
        return a

    def observables(self):
        """(EN, MU[ms], MZ, RMS[ms]) in the reference's scaled units."""
        en, mz = C.c_double(), C.c_double()
        mu, rms = np.empty(self.ms), np.empty(self.ms)
        _check(lib().fgpe_observables(self.h, C.byref(en), _dp(mu), C.byref(mz), _dp(rms)), "fgpe_observables")
        return en.value, mu, mz.value, rms

    def keep_old(self):
        _check(lib().fgpe_keep_old(self.h), "fgpe_keep_old")

    def conv_max(self):
        v = C.c_double()
        _check(lib().fgpe_conv_max(self.h, C.byref(v)), "fgpe_conv_max")
        return v.value

    def center_abs(self):
        a = np.empty(self.ms)
        _check(lib().fgpe_center_abs(self.h, _dp(a)), "fgpe_center_abs")
        return a

    def sync(self):
        _check(lib().fgpe_sync(self.h), "fgpe_sync")

    def checkpoint_save(self, path, ii=0, en_old=0.0, user=()):
        """Binary checkpoint: state as held on the device, device scalars and
        the loop state (ii, en_old, up to six user values)."""
        st = LoopState(ii=int(ii), en_old=float(en_old))
        for i, v in enumerate(user):
            st.user[i] = v
        _check(lib().fgpe_checkpoint_save(self.h, os.fsencode(path), C.byref(st)), "fgpe_checkpoint_save")

    def checkpoint_load(self, path):
        """Resume from a checkpoint of a run with the same parameters:
        (ii, en_old, user values)."""
        st = LoopState()
        _check(lib().fgpe_checkpoint_load(self.h, os.fsencode(path), C.byref(st)), "fgpe_checkpoint_load")
        return st.ii, st.en_old, tuple(st.user)

    PASS_NAMES = ("x", "x0", "y", "k", "finalize", "line1d", "z")

    def pass_times(self, nsteps):
        """Run nsteps with per-kernel HIP events: {pass: (avg ms, launches)}."""
        ms = np.zeros(len(self.PASS_NAMES))
        cnt = np.zeros(len(self.PASS_NAMES), dtype=np.int32)
        _check(lib().fgpe_pass_times(self.h, int(nsteps), _dp(ms), cnt.ctypes.data_as(C.POINTER(C.c_int32))),
               "fgpe_pass_times")
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(self.PASS_NAMES) if cnt[i]}

    def last_run_ms(self):
        v = C.c_double()
        _check(lib().fgpe_last_run_ms(self.h, C.byref(v)), "fgpe_last_run_ms")
        return v.value


class Batch:
    """Independent q2D instances advanced together (fgpe_batch_*): spin 1 or
    2 (one spin per batch), 64-256 points per axis."""

    def __init__(self, params_list):
        arr = (FgpeParams * len(params_list))(*params_list)
        self.params = list(params_list)
        self.count = len(params_list)
        b = C.c_void_p()
        _check(lib().fgpe_batch_create(arr, self.count, C.byref(b)), "fgpe_batch_create")
        self.h = b

    def close(self):
        if getattr(self, "h", None):
            lib().fgpe_batch_destroy(self.h)
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

    def shape(self, i):
        p = self.params[i]
        return (2 * p.spin + 1, 1, p.n[1], p.n[0])

    def initial_guess(self):
        _check(lib().fgpe_batch_initial_guess(self.h), "fgpe_batch_initial_guess")

    def normalize(self):
        _check(lib().fgpe_batch_normalize(self.h), "fgpe_batch_normalize")

    def upload(self, i, phi):
        a = np.ascontiguousarray(np.asarray(phi, dtype=np.complex128).reshape(self.shape(i)))
        _check(lib().fgpe_batch_upload(self.h, int(i), _dp(a.view(np.float64))), "fgpe_batch_upload")

    def download(self, i):
        a = np.empty(self.shape(i), dtype=np.complex128)
        _check(lib().fgpe_batch_download(self.h, int(i), _dp(a.view(np.float64))), "fgpe_batch_download")
        return a

    def run(self, nsteps):
        _check(lib().fgpe_batch_run(self.h, int(nsteps)), "fgpe_batch_run")

    def solve(self, stp, tol, max_steps):
        n = C.c_int32()
        _check(lib().fgpe_batch_solve(self.h, int(stp), float(tol), int(max_steps), C.byref(n)), "fgpe_batch_solve")
        return n.value

    def result(self, i):
        r = BatchResult()
        _check(lib().fgpe_batch_result_get(self.h, int(i), C.byref(r)), "fgpe_batch_result_get")
        return r

    def tune(self, trial_steps):
        ms = np.zeros(BATCH_NSCHED)
        chosen = (C.c_int32 * 4)()
        _check(lib().fgpe_batch_tune(self.h, int(trial_steps), _dp(ms), chosen), "fgpe_batch_tune")
        return ms, list(chosen)

    def last_ms(self):
        v = C.c_double()
        _check(lib().fgpe_batch_last_ms(self.h, C.byref(v)), "fgpe_batch_last_ms")
        return v.value
