"""ctypes bindings of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker the HIP library is compared
against, never part of the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None

_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_i64, _u64, _f64 = ctypes.c_int64, ctypes.c_uint64, ctypes.c_double


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        sig = {
            "orc_philox4x32_10": ([_U32P, _U32P, _U32P], None),
            "orc_step_keys": ([_u64, _u64, _U32P], None),
            "orc_perm": ([ctypes.c_uint32, ctypes.c_uint32, _U32P], ctypes.c_uint32),
            "orc_sample_steps": ([_u64, _u64, _u64, ctypes.c_uint32, ctypes.c_uint32, _I32P], None),
            "orc_soft_threshold": ([_f64, _f64], _f64),
            "orc_csc_residual": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _F64P, _F64P], None),
            "orc_csc_lipschitz": ([_i64, _I32P, _F64P, _f64, _F64P], None),
            "orc_omega": ([_i64, _i64, _I32P], ctypes.c_int32),
            "orc_eso_beta": ([ctypes.c_int32, _i64, _i64], _f64),
            "orc_pm_start": ([_u64, _i64], _f64),
            "orc_spectral_lmax": ([_i64, _i64, _I32P, _I32P, _F64P, _u64, ctypes.c_int], _f64),
            "orc_cd_run": ([ctypes.c_int, _i64, _i64, _I32P, _I32P, _F64P, _F64P, _f64, _u64, _i64,
                            _u64, _u64, _F64P, _F64P], None),
            "orc_lasso_eval": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _F64P, _f64, _F64P], None),
            "orc_nnls_eval": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _F64P, _F64P], None),
            "orc_svm_lipschitz": ([_i64, _I32P, _F64P, _f64, _F64P], None),
            "orc_svm_w": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _f64, _F64P, _F64P], None),
            "orc_svm_run": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _F64P, _f64, _u64, _i64, _u64,
                             _u64, _F64P, _F64P], None),
            "orc_svm_eval": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _f64, _F64P, _F64P], None),
            "orc_svm_list": ([_i64, _i64, _I32P, _I32P, _F64P, _F64P, _F64P, _f64, _I32P, _u64,
                              _F64P, _F64P], None),
            "orc_cd_async_omp": ([ctypes.c_int, _i64, _i64, _I32P, _I32P, _F64P, _F64P, _f64, _u64,
                                  _u64, _u64, ctypes.c_int, _F64P, _F64P], ctypes.c_int),
            "orc_async_coords": ([_u64, ctypes.c_uint32, _u64, _u64, ctypes.c_uint32, _I32P], None),
            "orc_cd_list": ([ctypes.c_int, _i64, _i64, _I32P, _I32P, _F64P, _F64P, _f64, _I32P, _u64,
                             _F64P, _F64P], None),
        }
        for name, (a, r) in sig.items():
            f = getattr(L, name)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def step_keys(seed, step):
    k = (ctypes.c_uint32 * 4)()
    lib().orc_step_keys(seed, step, k)
    return k


def perm(i, n, keys):
    return lib().orc_perm(i, n, keys)


def sample_steps(seed, step0, nsteps, n, tau) -> np.ndarray:
    out = np.empty(nsteps * tau, np.int32)
    lib().orc_sample_steps(seed, step0, nsteps, n, tau, _p(out, ctypes.c_int32))
    return out.reshape(nsteps, tau)


def soft_threshold(t, v):
    return lib().orc_soft_threshold(t, v)


def async_coords(seed, worker, k0, count, n) -> np.ndarray:
    out = np.empty(count, np.int32)
    lib().orc_async_coords(seed, worker, k0, count, n, _p(out, ctypes.c_int32))
    return out


def omega(nidx, idx) -> int:
    idx = i32(idx)
    return lib().orc_omega(nidx, idx.size, _p(idx, ctypes.c_int32))


def eso_beta(om, tau, n) -> float:
    return lib().orc_eso_beta(om, tau, n)


def pm_start(seed, c) -> float:
    return lib().orc_pm_start(seed, c)


def spectral_lmax(nseg, nidx, ptr, idx, vals, seed, iters=20) -> float:
    """Largest eigenvalue of D^-1/2 A^T A D^-1/2 by power steps (segments as
    columns)."""
    ptr, idx, vals = i32(ptr), i32(idx), f64(vals)
    return lib().orc_spectral_lmax(nseg, nidx, _p(ptr, ctypes.c_int32), _p(idx, ctypes.c_int32),
                                   _p(vals, ctypes.c_double), seed, iters)


def csc_residual(m, n, cp, ri, v, x, b) -> np.ndarray:
    cp, ri, v, x, b = i32(cp), i32(ri), f64(v), f64(x), f64(b)
    r = np.empty(m)
    lib().orc_csc_residual(m, n, _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32),
                           _p(v, ctypes.c_double), _p(x, ctypes.c_double),
                           _p(b, ctypes.c_double), _p(r, ctypes.c_double))
    return r


def csc_lipschitz(n, cp, v, beta) -> np.ndarray:
    cp, v = i32(cp), f64(v)
    L = np.empty(n)
    lib().orc_csc_lipschitz(n, _p(cp, ctypes.c_int32), _p(v, ctypes.c_double), beta,
                            _p(L, ctypes.c_double))
    return L


class CD:
    """Lasso (loss 0) / NNLS (loss 10) oracle state over one CSC instance."""

    def __init__(self, loss, m, n, cp, ri, v, b, lam, beta=1.0):
        self.loss, self.m, self.n = loss, m, n
        self.cp, self.ri, self.v, self.b = i32(cp), i32(ri), f64(v), f64(b)
        self.lam = float(lam)
        self.L = csc_lipschitz(n, self.cp, self.v, beta)
        self.x = np.zeros(n)
        self.r = csc_residual(m, n, self.cp, self.ri, self.v, self.x, self.b)

    def _args(self):
        return (_p(self.cp, ctypes.c_int32), _p(self.ri, ctypes.c_int32), _p(self.v, ctypes.c_double))

    def run(self, seed, tau, step0, nsteps):
        lib().orc_cd_run(self.loss, self.m, self.n, *self._args(), _p(self.L, ctypes.c_double),
                         self.lam, seed, tau, step0, nsteps, _p(self.x, ctypes.c_double),
                         _p(self.r, ctypes.c_double))

    def run_async(self, seed, k0, per_worker, threads):
        """threads workers, updates [k0, k0+per_worker) of each worker's stream"""
        return lib().orc_cd_async_omp(self.loss, self.m, self.n, *self._args(),
                                      _p(self.L, ctypes.c_double), self.lam, seed, k0,
                                      per_worker, threads, _p(self.x, ctypes.c_double),
                                      _p(self.r, ctypes.c_double))

    def run_list(self, coords):
        c = i32(coords)
        lib().orc_cd_list(self.loss, self.m, self.n, *self._args(), _p(self.L, ctypes.c_double),
                          self.lam, _p(c, ctypes.c_int32), c.size, _p(self.x, ctypes.c_double),
                          _p(self.r, ctypes.c_double))

    def eval(self, x=None) -> dict:
        x = self.x if x is None else f64(x)
        out = np.empty(8)
        if self.loss == 10:
            lib().orc_nnls_eval(self.m, self.n, *self._args(), _p(self.b, ctypes.c_double),
                                _p(x, ctypes.c_double), _p(out, ctypes.c_double))
            keys = ("primal", "kkt", "xg", "gmin", "grad_inf", "x_l1", "res_sq", "nnz_x")
        else:
            lib().orc_lasso_eval(self.m, self.n, *self._args(), _p(self.b, ctypes.c_double),
                                 _p(x, ctypes.c_double), self.lam, _p(out, ctypes.c_double))
            keys = ("primal", "dual", "gap", "rel_gap", "grad_inf", "x_l1", "res_sq", "nnz_x")
        return dict(zip(keys, out))


class SVM:
    """Hinge-SVM dual oracle state over one CSR (samples x features) instance."""

    def __init__(self, ns, nf, rp, ci, v, y, lam, beta=1.0):
        self.ns, self.nf = ns, nf
        self.rp, self.ci, self.v, self.y = i32(rp), i32(ci), f64(v), f64(y)
        self.lam = float(lam)
        self.L = np.empty(ns)
        lib().orc_svm_lipschitz(ns, _p(self.rp, ctypes.c_int32), _p(self.v, ctypes.c_double), beta,
                                _p(self.L, ctypes.c_double))
        self.alpha = np.zeros(ns)
        self.w = np.zeros(nf)

    def _args(self):
        return (_p(self.rp, ctypes.c_int32), _p(self.ci, ctypes.c_int32),
                _p(self.v, ctypes.c_double), _p(self.y, ctypes.c_double))

    def recompute_w(self):
        lib().orc_svm_w(self.ns, self.nf, *self._args(), self.lam, _p(self.alpha, ctypes.c_double),
                        _p(self.w, ctypes.c_double))

    def run(self, seed, tau, step0, nsteps):
        lib().orc_svm_run(self.ns, self.nf, *self._args(), _p(self.L, ctypes.c_double), self.lam,
                          seed, tau, step0, nsteps, _p(self.alpha, ctypes.c_double),
                          _p(self.w, ctypes.c_double))

    def run_list(self, coords):
        c = i32(coords)
        lib().orc_svm_list(self.ns, self.nf, *self._args(), _p(self.L, ctypes.c_double), self.lam,
                           _p(c, ctypes.c_int32), c.size, _p(self.alpha, ctypes.c_double),
                           _p(self.w, ctypes.c_double))

    def eval(self, alpha=None) -> dict:
        a = self.alpha if alpha is None else f64(alpha)
        out = np.empty(8)
        lib().orc_svm_eval(self.ns, self.nf, *self._args(), self.lam, _p(a, ctypes.c_double),
                           _p(out, ctypes.c_double))
        keys = ("gap", "primal", "dual", "accuracy", "rel_gap", "w_sq", "sum_alpha", "sum_hinge")
        return dict(zip(keys, out))
