"""Writes the golden fixtures under tests/golden/ (run in the build container).

* philox4x32_10_kat.json -- known-answer vectors of Philox4x32-10 as published
  with the Random123 library of Salmon et al. (file kat_vectors, entries
  "philox4x32 10"): counter, key, expected output.
* nnls_golden.npz -- small nonnegative least-squares instances solved by
  scipy.optimize.nnls (Lawson-Hanson active set), the independent solver the
  NNLS coordinate step is checked against.
* svm_golden.npz -- small hinge-SVM dual instances: the box-constrained dual
  maximised by scipy.optimize.minimize(L-BFGS-B) to 1e-13 and, for 6
  samples, by exhaustive enumeration of the active sets {0, 1, free}^n.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np
import scipy.optimize as so

HERE = os.path.dirname(os.path.abspath(__file__))

PHILOX_KAT = [
    {"ctr": [0x00000000, 0x00000000, 0x00000000, 0x00000000], "key": [0x00000000, 0x00000000],
     "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
    {"ctr": [0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], "key": [0xffffffff, 0xffffffff],
     "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
    {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], "key": [0xa4093822, 0x299f31d0],
     "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
]


def sparse_csc(rng, m, n, density):
    cols, rows, vals, ptr = [], [], [], [0]
    for j in range(n):
        k = max(1, int(rng.binomial(m, density)))
        r = np.sort(rng.choice(m, size=k, replace=False))
        v = rng.standard_normal(k)
        rows.extend(r.tolist())
        vals.extend(v.tolist())
        ptr.append(len(rows))
    return np.array(ptr, np.int32), np.array(rows, np.int32), np.array(vals, np.float64)


def dense_of(m, n, ptr, rows, vals):
    A = np.zeros((m, n))
    for j in range(n):
        A[rows[ptr[j]:ptr[j + 1]], j] = vals[ptr[j]:ptr[j + 1]]
    return A


def make_nnls():
    rng = np.random.default_rng(20240610)
    out = {}
    for t, (m, n, dens) in enumerate([(30, 20, 0.3), (15, 40, 0.25), (60, 60, 0.1)]):
        ptr, rows, vals = sparse_csc(rng, m, n, dens)
        A = dense_of(m, n, ptr, rows, vals)
        b = rng.standard_normal(m)
        x, _ = so.nnls(A, b, maxiter=50 * n)
        r = A @ x - b
        out[f"m{t}"] = np.array([m, n])
        out[f"ptr{t}"], out[f"rows{t}"], out[f"vals{t}"] = ptr, rows, vals
        out[f"b{t}"], out[f"x{t}"] = b, x
        out[f"obj{t}"] = np.array([0.5 * r @ r])
    out["count"] = np.array([3])
    np.savez(os.path.join(HERE, "nnls_golden.npz"), **out)


def svm_dual_obj(alpha, X, y, lam):
    n = X.shape[0]
    w = (alpha * y) @ X / (lam * n)
    return -lam / 2 * w @ w + alpha.sum() / n, w


def svm_lbfgs(X, y, lam):
    n = X.shape[0]

    def f(a):
        D, w = svm_dual_obj(a, X, y, lam)
        g = (y * (X @ w) - 1.0) / n
        return -D, g

    res = so.minimize(f, np.full(n, 0.5), jac=True, method="L-BFGS-B",
                      bounds=[(0.0, 1.0)] * n,
                      options={"ftol": 1e-16, "gtol": 1e-13, "maxiter": 20000, "maxcor": 50})
    return res.x


def svm_enumerate(X, y, lam):
    """Exact maximiser of the box-constrained dual by active-set enumeration."""
    n = X.shape[0]
    # D(a) = -(1/(2 lam n^2)) a^T K a + (1/n) 1^T a with K = (yX)(yX)^T
    K = (y[:, None] * X) @ (y[:, None] * X).T
    best, best_a = -np.inf, None
    for pat in itertools.product((0, 1, 2), repeat=n):
        a = np.array([0.0 if p == 0 else 1.0 for p in pat])
        free = [i for i, p in enumerate(pat) if p == 2]
        if free:
            F = np.array(free)
            fixed = np.array([i for i in range(n) if pat[i] != 2], dtype=int)
            # stationarity on the free set: (1/(lam n^2)) K_FF a_F = 1/n - (1/(lam n^2)) K_Fc a_c
            rhs = np.full(len(F), lam * n) - (K[np.ix_(F, fixed)] @ a[fixed] if fixed.size else 0.0)
            try:
                aF = np.linalg.solve(K[np.ix_(F, F)], rhs)
            except np.linalg.LinAlgError:
                continue
            if np.any(aF < -1e-12) or np.any(aF > 1 + 1e-12):
                continue
            a[F] = np.clip(aF, 0, 1)
        D, _ = svm_dual_obj(a, X, y, lam)
        if D > best:
            best, best_a = D, a
    return best_a


def make_svm():
    rng = np.random.default_rng(7)
    out = {}
    cases = [(30, 50, 0.2, 1.0 / 30), (6, 8, 0.5, 1.0 / 6), (24, 40, 0.25, 0.01)]
    for t, (ns, nf, dens, lam) in enumerate(cases):
        ptr, cols, vals, X = [0], [], [], np.zeros((ns, nf))
        for i in range(ns):
            k = max(1, int(rng.binomial(nf, dens)))
            c = np.sort(rng.choice(nf, size=k, replace=False))
            v = np.exp(rng.standard_normal(k))
            v /= np.linalg.norm(v)
            X[i, c] = v
            cols.extend(c.tolist())
            vals.extend(v.tolist())
            ptr.append(len(cols))
        y = np.where(rng.standard_normal(ns) > 0, 1.0, -1.0)
        a = svm_enumerate(X, y, lam) if ns <= 8 else svm_lbfgs(X, y, lam)
        D, w = svm_dual_obj(a, X, y, lam)
        P = lam / 2 * w @ w + np.maximum(0, 1 - y * (X @ w)).mean()
        out[f"shape{t}"] = np.array([ns, nf])
        out[f"ptr{t}"] = np.array(ptr, np.int32)
        out[f"cols{t}"] = np.array(cols, np.int32)
        out[f"vals{t}"] = np.array(vals)
        out[f"y{t}"] = y
        out[f"lam{t}"] = np.array([lam])
        out[f"alpha{t}"] = a
        out[f"dual{t}"] = np.array([D])
        out[f"primal{t}"] = np.array([P])
    out["count"] = np.array([len(cases)])
    np.savez(os.path.join(HERE, "svm_golden.npz"), **out)


def main():
    with open(os.path.join(HERE, "philox4x32_10_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors, philox4x32 10", "vectors": PHILOX_KAT},
                  f, indent=1)
    make_nnls()
    make_svm()


if __name__ == "__main__":
    main()
