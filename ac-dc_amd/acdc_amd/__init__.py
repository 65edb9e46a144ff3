"""Python view of the gfx950 PCDM library (``lib/libacdc_hip.so``).

A thin ctypes layer over the C ABI declared in ``include/acdc_hip.h`` and
``include/acdc_gen.h``.  Every call goes to the HIP library; there is no CPU
fallback, and importing this module fails if the library has not been built.

The objects mirror the reference's solver interface:
``Problem`` ~ ``ProblemData`` (cpp/src/solver/structures.h:30-97),
``Solver.run_updates`` ~ ``executeBulkOfIterations``
(cpp/src/solver/MulticoreEngineExecutor.h:83-89),
``Solver.recompute_residuals`` ~ ``recomputeResiduals``
(cpp/src/solver/lossFunctions/squareLoss/MulticoreSquareLoss.h:42-53),
``Solver.report`` ~ ``computeObjectiveValue`` + ``getResultLogRow``
(squareLoss/MulticoreL1RegSquareLoss.h:30-39, SVMDualLoss/SVMDualLoss.h:31-42).
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libacdc_hip.so")

LOSS_L1_SQUARE = 0
LOSS_SVM_DUAL_HINGE = 2
LOSS_NNLS = 10
BATCH_NONFINITE = 1
PREC_FP64 = 0
PREC_FP32_REFRESH = 1
BETA_OMEGA = 0
BETA_FORCED = 1
BETA_NONE = 2
BETA_SPECTRAL = 3
STEP_SYNC = 0
STEP_ASYNC = 1
LAUNCH_AUTO = 0
LAUNCH_PER_STEP = 1
LAUNCH_PERSISTENT = 2
LAUNCH_PER_STEP_STREAMED = 3
LAUNCH_ASYNC_DIRECT = 4
LAUNCH_ASYNC_STAGED = 5

NNZ_FIXED = 0
NNZ_ZIPF_TRUNC = 1
NNZ_ZIPF_CLIP = 2

_STATUS = {0: "ACDC_OK", 1: "ACDC_ERR_INVALID", 2: "ACDC_ERR_HIP",
           3: "ACDC_ERR_NOMEM", 4: "ACDC_ERR_STATE", 5: "ACDC_ERR_NO_DEVICE"}


class AcdcError(RuntimeError):
    """A non-zero acdc_status, with the library's error text."""

    def __init__(self, status: int, text: str):
        super().__init__(f"{_STATUS.get(status, status)}: {text}")
        self.status = status


class Settings(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("tau", ctypes.c_int64),
                ("seed", ctypes.c_uint64), ("beta_mode", ctypes.c_int32),
                ("forced_beta", ctypes.c_double), ("precision", ctypes.c_int32),
                ("step_mode", ctypes.c_int32), ("lanes_per_coord", ctypes.c_int32),
                ("refresh_period", ctypes.c_int32), ("launch_mode", ctypes.c_int32)]


class _Report(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("primal", "dual", "gap", "rel_gap", "grad_inf", "x_l1", "res_sq",
                 "kkt", "accuracy", "nnz_x")] + \
               [("updates", ctypes.c_uint64), ("epochs", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("eval_ms", ctypes.c_double),
                ("step_ms", ctypes.c_double), ("step_launches", ctypes.c_uint64)]


class GenLassoCfg(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("nnz_law", ctypes.c_int32), ("nnz_fixed", ctypes.c_int32),
                ("nnz_lo", ctypes.c_int32), ("nnz_hi", ctypes.c_int32),
                ("zipf_s", ctypes.c_double), ("xstar_nnz", ctypes.c_int64),
                ("xstar_normal", ctypes.c_int32), ("noise", ctypes.c_double),
                ("lambda_frac", ctypes.c_double), ("seed", ctypes.c_uint64)]


class GenSvmCfg(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int64), ("n_features", ctypes.c_int64),
                ("row_median", ctypes.c_double), ("row_sigma", ctypes.c_double),
                ("wstar_nnz", ctypes.c_int64), ("flip_frac", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class _DatasetInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("lambda_", ctypes.c_double),
                ("lambda_max", ctypes.c_double), ("xstar_nnz", ctypes.c_int64)]


class BatchSettings(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("precision", ctypes.c_int32),
                ("lanes_per_coord", ctypes.c_int32), ("refresh_period", ctypes.c_int32),
                ("max_epochs", ctypes.c_int32), ("gap_tol", ctypes.c_double),
                ("use_graph", ctypes.c_int32), ("step_mode", ctypes.c_int32),
                ("tau", ctypes.c_int32), ("queue_chunk", ctypes.c_int32),
                ("loss", ctypes.c_int32), ("beta_mode", ctypes.c_int32),
                ("screening", ctypes.c_int32), ("ws_newton_iters", ctypes.c_int32)]


# CG iterations of a working set's support Newton step when screening is on
# and the caller names none
WS_NEWTON_ITERS = 200


class _BatchResult(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("lambda_", "primal", "dual", "rel_gap", "nnz_x")] + \
               [("epochs", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("updates", ctypes.c_uint64), ("device_us", ctypes.c_double),
                ("ws_size", ctypes.c_double), ("status", ctypes.c_int32),
                ("newton_steps", ctypes.c_int32)]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"acdc_amd: {LIB_PATH} is missing; build it with `make -C ac-dc_amd` "
            "(or __graft_entry__.build()).  There is no CPU fallback.")
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "acdc_last_error": ([], ctypes.c_char_p),
    "acdc_device_count": ([_I32P], ctypes.c_int),
    "acdc_problem_create_csc": ([ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _I32P, _I32P,
                                 _F64P, _F64P, ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_problem_create_csr": ([ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _I32P, _I32P,
                                 _F64P, _F64P, ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_problem_destroy": ([_P], ctypes.c_int),
    "acdc_problem_omega": ([_P, _I32P], ctypes.c_int),
    "acdc_solver_create": ([_P, ctypes.c_int32, ctypes.POINTER(Settings), ctypes.POINTER(_P)],
                           ctypes.c_int),
    "acdc_solver_destroy": ([_P], ctypes.c_int),
    "acdc_run_updates": ([_P, ctypes.c_uint64], ctypes.c_int),
    "acdc_recompute_residuals": ([_P], ctypes.c_int),
    "acdc_report_get": ([_P, ctypes.POINTER(_Report)], ctypes.c_int),
    "acdc_get_x": ([_P, _F64P, ctypes.c_int64], ctypes.c_int),
    "acdc_set_x": ([_P, _F64P, ctypes.c_int64], ctypes.c_int),
    "acdc_get_residuals": ([_P, _F64P, ctypes.c_int64], ctypes.c_int),
    "acdc_get_lipschitz": ([_P, _F64P, ctypes.c_int64], ctypes.c_int),
    "acdc_set_lambda": ([_P, ctypes.c_double], ctypes.c_int),
    "acdc_get_beta": ([_P, _F64P], ctypes.c_int),
    "acdc_sample_coordinates": ([_P, ctypes.c_uint64, ctypes.c_uint64, _I32P], ctypes.c_int),
    "acdc_sample_coordinates_noalloc": ([_P, ctypes.c_uint64, ctypes.c_uint64, _I32P],
                                        ctypes.c_int),
    "acdc_synchronize": ([_P], ctypes.c_int),
    "acdc_profile_step_kernels": ([_P, ctypes.c_uint64, _F64P, _F64P], ctypes.c_int),
    "acdc_batch_create": ([ctypes.POINTER(_P), ctypes.c_int32, _F64P, ctypes.c_int32,
                           ctypes.POINTER(BatchSettings), ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_batch_run": ([_P, _F64P], ctypes.c_int),
    "acdc_batch_results": ([_P, ctypes.POINTER(_BatchResult), ctypes.c_int64], ctypes.c_int),
    "acdc_batch_get_x": ([_P, ctypes.c_int32, _F64P, ctypes.c_int64], ctypes.c_int),
    "acdc_batch_instance_info": ([_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), _F64P,
                                  _I32P], ctypes.c_int),
    "acdc_batch_destroy": ([_P], ctypes.c_int),
    "acdc_batch_tune": ([ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(BatchSettings),
                         ctypes.POINTER(BatchSettings), _F64P], ctypes.c_int),
    "acdc_solver_tune_lanes": ([_P, ctypes.c_int32, ctypes.POINTER(Settings), _I32P, _F64P],
                               ctypes.c_int),
    "acdc_gen_lasso": ([ctypes.POINTER(GenLassoCfg), ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_gen_svm": ([ctypes.POINTER(GenSvmCfg), ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_gen_ttd": ([ctypes.c_int32, ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_dataset_info_get": ([_P, ctypes.POINTER(_DatasetInfo)], ctypes.c_int),
    "acdc_dataset_copy": ([_P, _I32P, _I32P, _F64P, _F64P, _F64P], ctypes.c_int),
    "acdc_dataset_problem": ([_P, ctypes.POINTER(_P)], ctypes.c_int),
    "acdc_dataset_destroy": ([_P], ctypes.c_int),
}
for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res


# Every live handle, so that they are all released by an atexit hook -- which
# runs before the interpreter tears modules down and before the C runtime's
# exit handlers (a profiler's finalisation among them) -- instead of by
# __del__ during interpreter shutdown.  Solvers and batches go before the
# problems they were built from.
_live_solvers: "weakref.WeakSet" = weakref.WeakSet()
_live_problems: "weakref.WeakSet" = weakref.WeakSet()


def close_all() -> None:
    """Destroys every solver, batch and problem handle still open."""
    for obj in list(_live_solvers):
        obj.close()
    for obj in list(_live_problems):
        obj.close()


atexit.register(close_all)


def _check(status: int) -> None:
    if status != 0:
        raise AcdcError(status, lib.acdc_last_error().decode())


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def device_count() -> int:
    c = ctypes.c_int32(0)
    _check(lib.acdc_device_count(ctypes.byref(c)))
    return c.value


# --------------------------------------------------------------------- data
@dataclass
class Dataset:
    """Host arrays of one generated instance.  kind 0: CSC A (m x n) + b;
    kind 1: CSR samples (n samples x m features) + labels."""
    kind: int
    m: int
    n: int
    ptr: np.ndarray
    idx: np.ndarray
    vals: np.ndarray
    rhs: np.ndarray
    xstar: np.ndarray
    lam: float
    lambda_max: float

    @property
    def nnz(self) -> int:
        return int(self.idx.size)

    def problem(self) -> "Problem":
        if self.kind == 0:
            return Problem.from_csc(self.m, self.n, self.ptr, self.idx, self.vals, self.rhs)
        return Problem.from_csr(self.n, self.m, self.ptr, self.idx, self.vals, self.rhs)


def _take_dataset(handle: _P) -> Dataset:
    info = _DatasetInfo()
    try:
        _check(lib.acdc_dataset_info_get(handle, ctypes.byref(info)))
        nseg = info.n
        ptr = np.empty(nseg + 1, np.int32)
        idx = np.empty(info.nnz, np.int32)
        vals = np.empty(info.nnz, np.float64)
        rhs = np.empty(info.m if info.kind == 0 else info.n, np.float64)
        xstar = np.empty(info.n if info.kind == 0 else info.m, np.float64)
        _check(lib.acdc_dataset_copy(handle, _ptr(ptr, ctypes.c_int32), _ptr(idx, ctypes.c_int32),
                                     _ptr(vals, ctypes.c_double), _ptr(rhs, ctypes.c_double),
                                     _ptr(xstar, ctypes.c_double)))
    finally:
        lib.acdc_dataset_destroy(handle)
    return Dataset(info.kind, info.m, info.n, ptr, idx, vals, rhs, xstar,
                   info.lambda_, info.lambda_max)


def gen_lasso(m, n, *, nnz_law=NNZ_FIXED, nnz_fixed=20, nnz_lo=1, nnz_hi=1, zipf_s=1.1,
              xstar_nnz=0, xstar_normal=False, noise=0.01, lambda_frac=0.1, seed=1) -> Dataset:
    cfg = GenLassoCfg(m, n, nnz_law, nnz_fixed, nnz_lo, nnz_hi, zipf_s, xstar_nnz,
                      int(bool(xstar_normal)), noise, lambda_frac, seed)
    h = _P()
    _check(lib.acdc_gen_lasso(ctypes.byref(cfg), ctypes.byref(h)))
    return _take_dataset(h)


def gen_svm(n_samples, n_features, *, row_median=300.0, row_sigma=0.901, wstar_nnz=10000,
            flip_frac=0.1, seed=3) -> Dataset:
    cfg = GenSvmCfg(n_samples, n_features, row_median, row_sigma, wstar_nnz, flip_frac, seed)
    h = _P()
    _check(lib.acdc_gen_svm(ctypes.byref(cfg), ctypes.byref(h)))
    return _take_dataset(h)


def gen_ttd(cells: int) -> Dataset:
    h = _P()
    _check(lib.acdc_gen_ttd(int(cells), ctypes.byref(h)))
    return _take_dataset(h)


class Problem:
    """Host copy of one instance inside the library (ProblemData)."""

    def __init__(self, handle: _P, kind: int, m: int, n: int, nnz: int):
        self._h = handle
        self.kind, self.m, self.n, self.nnz = kind, m, n, nnz
        _live_problems.add(self)

    @classmethod
    def from_csc(cls, m, n, col_ptr, row_idx, vals, b) -> "Problem":
        cp, ri, v, bb = _i32(col_ptr), _i32(row_idx), _f64(vals), _f64(b)
        if cp.size != n + 1 or bb.size != m or ri.size != v.size:
            raise ValueError("from_csc: inconsistent array sizes")
        h = _P()
        _check(lib.acdc_problem_create_csc(m, n, ri.size, _ptr(cp, ctypes.c_int32),
                                           _ptr(ri, ctypes.c_int32), _ptr(v, ctypes.c_double),
                                           _ptr(bb, ctypes.c_double), ctypes.byref(h)))
        return cls(h, 0, m, n, ri.size)

    @classmethod
    def from_csr(cls, n_samples, n_features, row_ptr, col_idx, vals, labels) -> "Problem":
        rp, ci, v, y = _i32(row_ptr), _i32(col_idx), _f64(vals), _f64(labels)
        if rp.size != n_samples + 1 or y.size != n_samples or ci.size != v.size:
            raise ValueError("from_csr: inconsistent array sizes")
        h = _P()
        _check(lib.acdc_problem_create_csr(n_samples, n_features, ci.size,
                                           _ptr(rp, ctypes.c_int32), _ptr(ci, ctypes.c_int32),
                                           _ptr(v, ctypes.c_double), _ptr(y, ctypes.c_double),
                                           ctypes.byref(h)))
        # reference orientation for the SVM: n = samples, m = features
        return cls(h, 1, n_features, n_samples, ci.size)

    @property
    def omega(self) -> int:
        o = ctypes.c_int32()
        _check(lib.acdc_problem_omega(self._h, ctypes.byref(o)))
        return o.value

    def close(self) -> None:
        if self._h:
            lib.acdc_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        if not sys.is_finalizing():  # at shutdown the atexit hook has run
            self.close()


class Solver:
    """One device solver (the executor + loss of the reference)."""

    def __init__(self, problem: Problem, loss: int, *, lam: float, tau: int = 1, seed: int = 1,
                 beta_mode: int = BETA_OMEGA, forced_beta: float = 1.0,
                 precision: int = PREC_FP64, step_mode: int = STEP_SYNC,
                 lanes_per_coord: int = 0, refresh_period: int = 0,
                 launch_mode: int = LAUNCH_AUTO):
        self.problem = problem
        self.loss = loss
        self.n = problem.n
        self.res_len = problem.m
        st = Settings(lam, tau, seed, beta_mode, forced_beta, precision, step_mode,
                      lanes_per_coord, refresh_period, launch_mode)
        self.settings = st
        h = _P()
        _check(lib.acdc_solver_create(problem._h, loss, ctypes.byref(st), ctypes.byref(h)))
        self._h = h
        _live_solvers.add(self)

    def run_updates(self, n_updates: int) -> None:
        """Queues n_updates coordinate updates on the solver's stream and
        returns; report(), x and synchronize() wait for them."""
        _check(lib.acdc_run_updates(self._h, int(n_updates)))

    def recompute_residuals(self) -> None:
        _check(lib.acdc_recompute_residuals(self._h))

    def report(self) -> dict:
        r = _Report()
        _check(lib.acdc_report_get(self._h, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _Report._fields_}

    @property
    def x(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        _check(lib.acdc_get_x(self._h, _ptr(out, ctypes.c_double), self.n))
        return out

    @x.setter
    def x(self, value) -> None:
        v = _f64(value)
        _check(lib.acdc_set_x(self._h, _ptr(v, ctypes.c_double), v.size))

    @property
    def residuals(self) -> np.ndarray:
        out = np.empty(self.res_len, np.float64)
        _check(lib.acdc_get_residuals(self._h, _ptr(out, ctypes.c_double), self.res_len))
        return out

    @property
    def lipschitz(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        _check(lib.acdc_get_lipschitz(self._h, _ptr(out, ctypes.c_double), self.n))
        return out

    @property
    def beta(self) -> float:
        b = ctypes.c_double()
        _check(lib.acdc_get_beta(self._h, ctypes.byref(b)))
        return b.value

    def set_lambda(self, lam: float) -> None:
        _check(lib.acdc_set_lambda(self._h, float(lam)))

    def sample_coordinates(self, step0: int, nsteps: int) -> np.ndarray:
        out = np.empty(nsteps * self.settings.tau, np.int32)
        _check(lib.acdc_sample_coordinates(self._h, step0, nsteps, _ptr(out, ctypes.c_int32)))
        return out.reshape(nsteps, self.settings.tau)

    def sample_coordinates_noalloc(self, step0: int, nsteps: int) -> np.ndarray:
        """sample_coordinates through the allocation-free entry point."""
        out = np.empty(nsteps * self.settings.tau, np.int32)
        _check(lib.acdc_sample_coordinates_noalloc(self._h, step0, nsteps,
                                                   _ptr(out, ctypes.c_int32)))
        return out.reshape(nsteps, self.settings.tau)

    def profile_step_kernels(self, nsteps: int) -> tuple:
        """Mean device ms per launch of the (gather, scatter) kernels."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib.acdc_profile_step_kernels(self._h, nsteps, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def synchronize(self) -> None:
        _check(lib.acdc_synchronize(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.acdc_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        if not sys.is_finalizing():
            self.close()


def _newton_iters(screening, iters):
    if iters is None:
        return WS_NEWTON_ITERS if screening else 0
    return int(iters)


class Batch:
    """Many Lasso, NNLS (loss=LOSS_NNLS) or hinge-SVM dual
    (loss=LOSS_SVM_DUAL_HINGE, problems from Problem.csr) instances, each along
    a warm-started lambda path, in one persistent launch (acdc_batch_* of the
    C ABI).  For SVM instances x(i) is alpha."""

    def __init__(self, problems, lambdas, *, seed=1, precision=PREC_FP32_REFRESH,
                 lanes_per_coord=0, refresh_period=5, max_epochs=200, gap_tol=1e-6,
                 use_graph=True, step_mode=STEP_SYNC, tau=0, queue_chunk=0,
                 loss=LOSS_L1_SQUARE, beta_mode=BETA_OMEGA, screening=False,
                 ws_newton_iters=None):
        lam = _f64(lambdas)
        if lam.ndim != 2 or lam.shape[0] != len(problems):
            raise ValueError("lambdas must be (instances, path_len)")
        self.count, self.path_len = lam.shape
        self.problems = list(problems)
        self.n = [p.n for p in problems]
        arr = (_P * self.count)(*[p._h for p in problems])
        st = BatchSettings(seed, precision, lanes_per_coord, refresh_period, max_epochs, gap_tol,
                           int(bool(use_graph)), step_mode, tau, queue_chunk, loss, beta_mode,
                           int(bool(screening)), _newton_iters(screening, ws_newton_iters))
        h = _P()
        _check(lib.acdc_batch_create(arr, self.count, _ptr(lam, ctypes.c_double), self.path_len,
                                     ctypes.byref(st), ctypes.byref(h)))
        self._h = h
        _live_solvers.add(self)

    def run(self) -> float:
        """Solves the whole batch from x = 0; returns the device time in ms."""
        ms = ctypes.c_double()
        _check(lib.acdc_batch_run(self._h, ctypes.byref(ms)))
        return ms.value

    def results(self) -> np.ndarray:
        out = (_BatchResult * (self.count * self.path_len))()
        _check(lib.acdc_batch_results(self._h, out, self.count * self.path_len))
        dt = np.dtype([("lambda", "f8"), ("primal", "f8"), ("dual", "f8"), ("rel_gap", "f8"),
                       ("nnz_x", "f8"), ("epochs", "i4"), ("converged", "i4"), ("updates", "u8"),
                       ("device_us", "f8"), ("ws_size", "f8"), ("status", "i4"),
                       ("newton_steps", "i4")])
        return np.frombuffer(bytes(out), dtype=dt).reshape(self.count, self.path_len).copy()

    def x(self, i: int) -> np.ndarray:
        out = np.empty(self.n[i], np.float64)
        _check(lib.acdc_batch_get_x(self._h, i, _ptr(out, ctypes.c_double), out.size))
        return out

    def info(self, i: int) -> tuple:
        tau, beta, lds = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int32()
        _check(lib.acdc_batch_instance_info(self._h, i, ctypes.byref(tau), ctypes.byref(beta),
                                            ctypes.byref(lds)))
        return tau.value, beta.value, bool(lds.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.acdc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if not sys.is_finalizing():
            self.close()


def batch_tune(problems, *, seed=1, precision=PREC_FP32_REFRESH, step_mode=STEP_ASYNC,
               refresh_period=5, max_epochs=200, gap_tol=1e-6, use_graph=True,
               loss=LOSS_L1_SQUARE, beta_mode=BETA_OMEGA, screening=False,
               ws_newton_iters=None) -> dict:
    """acdc_batch_tune: lanes per coordinate, tau and queue chunking for a batch."""
    arr = (_P * len(problems))(*[p._h for p in problems])
    base = BatchSettings(seed, precision, 0, refresh_period, max_epochs, gap_tol,
                         int(bool(use_graph)), step_mode, 0, 0, loss, beta_mode,
                         int(bool(screening)), _newton_iters(screening, ws_newton_iters))
    out = BatchSettings()
    score = ctypes.c_double()
    _check(lib.acdc_batch_tune(arr, len(problems), ctypes.byref(base), ctypes.byref(out),
                               ctypes.byref(score)))
    return {"lanes_per_coord": out.lanes_per_coord, "tau": out.tau,
            "queue_chunk": out.queue_chunk, "score": score.value}


def solver_tune_lanes(problem, loss, *, lam, tau, seed=1, precision=PREC_FP32_REFRESH,
                      step_mode=STEP_ASYNC) -> dict:
    """acdc_solver_tune_lanes: the fastest lanes per coordinate for a solver."""
    st = Settings(lam, tau, seed, BETA_OMEGA, 1.0, precision, step_mode, 0, 0, LAUNCH_AUTO)
    lanes = ctypes.c_int32()
    rate = ctypes.c_double()
    _check(lib.acdc_solver_tune_lanes(problem._h, loss, ctypes.byref(st), ctypes.byref(lanes),
                                      ctypes.byref(rate)))
    return {"lanes_per_coord": lanes.value, "updates_per_s": rate.value}


def lambda_path(lambda_max: float, path_len: int = 32, ratio: float = 1e-3) -> np.ndarray:
    """Geometric path from lambda_max down to ratio * lambda_max."""
    return lambda_max * np.geomspace(1.0, ratio, path_len)


# ------------------------------------------------------------- configurations
def config_lasso_c1(seed: int = 1) -> Dataset:
    """2,000 x 10,000 Lasso, 20 nnz per column, 100 +-1 planted nonzeros,
    b = Ax* + 0.01 N(0,1), lambda = 0.1 ||A^T b||_inf."""
    return gen_lasso(2000, 10000, nnz_law=NNZ_FIXED, nnz_fixed=20, xstar_nnz=100,
                     xstar_normal=False, noise=0.01, lambda_frac=0.1, seed=seed)


def config_lasso_c2(seed: int = 2, m: int = 200_000, n: int = 1_000_000) -> Dataset:
    """200,000 x 1,000,000 Lasso, column nnz truncated Zipf(1.1) on [5,200],
    5,000 N(0,1) planted nonzeros, lambda = 0.05 ||A^T b||_inf."""
    return gen_lasso(m, n, nnz_law=NNZ_ZIPF_TRUNC, nnz_lo=5, nnz_hi=200, zipf_s=1.1,
                     xstar_nnz=5000 * n // 1_000_000, xstar_normal=True, noise=0.01,
                     lambda_frac=0.05, seed=seed)


def config_svm_c3(seed: int = 3, n_samples: int = 20_000, n_features: int = 1_350_000) -> Dataset:
    """20,000 samples x 1,350,000 features, lognormal row nnz (median 300),
    10,000-sparse w*, 10% label noise; lambda = 1/n."""
    return gen_svm(n_samples, n_features, row_median=300.0, row_sigma=0.901,
                   wstar_nnz=min(10000, n_features), flip_frac=0.1, seed=seed)


def config_batch_c5(count: int = 256, seed: int = 5, m_range=(1000, 50000),
                    n_range=(5000, 200000)):
    """Batched Lasso workload: instance sizes m ~ logU(m_range), n ~ logU(n_range)
    (numpy PCG64 keyed by seed), column nnz truncated Zipf(1.1) on [3,100],
    unit N(0,1) columns, x* with n/100 entries +-1, b = Ax* + 0.01 N(0,1).
    Instance i is generated with seed 1000*seed + i."""
    rng = np.random.default_rng(seed)
    ms = np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1]), count)).astype(np.int64)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count)).astype(np.int64)
    return [gen_lasso(int(m), int(n), nnz_law=NNZ_ZIPF_TRUNC, nnz_lo=3, nnz_hi=100, zipf_s=1.1,
                      xstar_nnz=max(1, int(n) // 100), xstar_normal=False, noise=0.01,
                      lambda_frac=1.0, seed=1000 * seed + i)
            for i, (m, n) in enumerate(zip(ms, ns))]
