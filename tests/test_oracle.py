"""Pins the CPU FP64 oracle before anything is compared against it.

Each check is chosen so that a plausible slip (a dropped term, a sign, an
index, a transposed operand) fails it: published known-answer vectors,
closed forms, brute-force enumeration, independent solvers (scipy fixtures)
and planted instances with a known optimum.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def rand_csc(rng, m, n, k):
    ptr = np.arange(n + 1, dtype=np.int32) * k
    rows = np.concatenate([np.sort(rng.choice(m, size=k, replace=False)) for _ in range(n)])
    vals = rng.standard_normal(n * k)
    return ptr, rows.astype(np.int32), vals


def dense(m, n, ptr, rows, vals):
    A = np.zeros((m, n))
    for j in range(n):
        A[rows[ptr[j]:ptr[j + 1]], j] += vals[ptr[j]:ptr[j + 1]]
    return A


# ----------------------------------------------------------------- sampling
def test_philox_known_answers(orc):
    with open(os.path.join(HERE, "golden", "philox4x32_10_kat.json")) as f:
        kat = json.load(f)["vectors"]
    assert len(kat) == 3
    for v in kat:
        assert orc.philox(v["ctr"], v["key"]) == v["out"]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 100, 1000, 4097, 65536, 100003])
def test_perm_is_bijection(orc, n):
    for step in (0, 1, 12345):
        keys = orc.step_keys(7, step)
        img = np.array([orc.perm(i, n, keys) for i in range(n)])
        assert img.min() >= 0 and img.max() < n
        assert np.unique(img).size == n


def test_perm_slot0_uniform(orc):
    # tau = 1 draws perm_s(0): must be uniform over [0, n) across steps
    n, steps = 97, 97 * 300
    s = orc.sample_steps(11, 0, steps, n, 1).ravel()
    cnt = np.bincount(s, minlength=n)
    exp = steps / n
    chi2 = ((cnt - exp) ** 2 / exp).sum()
    # 96 degrees of freedom: P(chi2 > 160) < 1e-4
    assert chi2 < 160


def test_sample_steps_distinct_and_seeded(orc):
    a = orc.sample_steps(5, 10, 20, 1000, 64)
    for row in a:
        assert np.unique(row).size == 64
    b = orc.sample_steps(5, 10, 20, 1000, 64)
    c = orc.sample_steps(6, 10, 20, 1000, 64)
    assert (a == b).all() and not (a == c).all()
    # step counter addressing: steps 10..29 == rows 10..29 of a longer replay
    d = orc.sample_steps(5, 0, 30, 1000, 64)
    assert (d[10:] == a).all()


# ---------------------------------------------------------------- scalar rule
@pytest.mark.parametrize("t,v,want", [
    (1.0, 3.0, 2.0), (1.0, -3.0, -2.0), (1.0, 1.0, 0.0), (1.0, -1.0, 0.0),
    (1.0, 0.5, 0.0), (0.0, 2.5, 2.5), (0.0, -0.0, 0.0), (0.25, -0.75, -0.5),
    (2.0, 1e300, 1e300)])
def test_soft_threshold_table(orc, t, v, want):
    assert orc.soft_threshold(t, v) == want


def test_omega_and_beta(orc):
    rng = np.random.default_rng(1)
    idx = rng.integers(0, 50, size=1000)
    assert orc.omega(50, idx) == np.bincount(idx, minlength=50).max()
    assert orc.eso_beta(1, 4096, 10**6) == 1.0          # separable: no penalty
    assert orc.eso_beta(300, 1, 10**6) == 1.0           # serial
    assert orc.eso_beta(300, 1000, 1000) == pytest.approx(300.0)  # fully parallel: beta = omega
    assert orc.eso_beta(304, 4096, 10**6) == pytest.approx(1 + 303 * 4095 / 999999)


def _csc(cols, m):
    """CSC arrays of a list of dense columns (zeros dropped)."""
    ptr, idx, val = [0], [], []
    for c in cols:
        nz = np.nonzero(c)[0]
        idx += nz.tolist()
        val += np.asarray(c)[nz].tolist()
        ptr.append(len(idx))
    return np.array(ptr), np.array(idx), np.array(val, dtype=float)


def test_spectral_lmax_closed_forms(orc):
    rng = np.random.default_rng(7)
    m = 40
    # orthogonal segments (disjoint supports): the normalised Gram matrix is I
    cols = []
    for k in range(10):
        c = np.zeros(m)
        c[4 * k:4 * k + 4] = rng.normal(size=4)
        cols.append(c)
    ptr, idx, val = _csc(cols, m)
    assert orc.spectral_lmax(10, m, ptr, idx, val, seed=3) == pytest.approx(1.0, rel=1e-12)
    # n equal segments (any scales): the normalised Gram matrix is all ones
    base = rng.normal(size=m)
    cols = [base * s for s in rng.uniform(0.5, 2.0, size=12)]
    ptr, idx, val = _csc(cols, m)
    assert orc.spectral_lmax(12, m, ptr, idx, val, seed=3) == pytest.approx(12.0, rel=1e-12)
    # two disjoint groups of 5 and 3 equal segments: eigenvalues 5 and 3 (and 0)
    a, b = np.zeros(m), np.zeros(m)
    a[:20] = rng.normal(size=20)
    b[20:] = rng.normal(size=20)
    cols = [a * 1.5] * 5 + [b * 0.7] * 3
    ptr, idx, val = _csc(cols, m)
    assert orc.spectral_lmax(8, m, ptr, idx, val, seed=3) == pytest.approx(5.0, rel=1e-8)
    # an empty segment is left out of the operator
    ptr2 = np.concatenate([ptr, [ptr[-1]]])
    assert orc.spectral_lmax(9, m, ptr2, idx, val, seed=3) == pytest.approx(5.0, rel=1e-8)


def test_spectral_lmax_matches_eigvalsh(orc):
    """A random sparse matrix plus a shared direction, so the top eigenvalue
    of D^-1/2 A^T A D^-1/2 stands well clear of the rest and 20 power steps
    reach it."""
    rng = np.random.default_rng(11)
    m, n = 60, 30
    shared = rng.normal(size=m)
    A = rng.normal(size=(m, n)) * (rng.uniform(size=(m, n)) < 0.2) + 0.8 * shared[:, None]
    ptr, idx, val = _csc([A[:, j] for j in range(n)], m)
    d = 1.0 / np.linalg.norm(A, axis=0)
    M = (A * d).T @ (A * d)
    ev = np.linalg.eigvalsh(M)
    assert ev[-1] > 3 * ev[-2]
    assert orc.spectral_lmax(n, m, ptr, idx, val, seed=5) == pytest.approx(ev[-1], rel=1e-9)
    # and it lies in [1, omega]
    om = np.count_nonzero(A, axis=1).max()
    assert 1.0 <= ev[-1] <= om


def test_pm_start_is_seeded_uniform(orc):
    u = np.array([orc.pm_start(9, c) for c in range(4000)])
    assert ((u >= 0) & (u < 1)).all()
    assert abs(u.mean() - 0.5) < 0.02
    assert orc.pm_start(9, 17) != orc.pm_start(10, 17)


# -------------------------------------------------------------- square loss
def test_one_column_closed_form(orc):
    rng = np.random.default_rng(2)
    m = 40
    rows = np.sort(rng.choice(m, 12, replace=False)).astype(np.int32)
    a = rng.standard_normal(12)
    b = rng.standard_normal(m)
    ptr = np.array([0, 12], np.int32)
    aTb = a @ b[rows]
    lam = 0.3 * abs(aTb)
    s = orc.CD(0, m, 1, ptr, rows, a, b, lam)
    s.run(seed=1, tau=1, step0=0, nsteps=1)
    want = np.sign(aTb) * max(abs(aTb) - lam, 0) / (a @ a)
    assert s.x[0] == pytest.approx(want, rel=1e-14)
    # residual stays r = A x - b
    assert np.allclose(s.r, orc.csc_residual(m, 1, ptr, rows, a, s.x, b), atol=1e-14)
    nn = orc.CD(10, m, 1, ptr, rows, a, -np.sign(aTb) * b, 0.0)
    nn.run(seed=1, tau=1, step0=0, nsteps=1)
    assert nn.x[0] == 0.0  # a^T b < 0 after the sign flip: clipped at zero
    nn2 = orc.CD(10, m, 1, ptr, rows, a, np.sign(aTb) * b, 0.0)
    nn2.run(seed=1, tau=1, step0=0, nsteps=1)
    assert nn2.x[0] == pytest.approx(abs(aTb) / (a @ a), rel=1e-14)


def test_serial_step_is_exact_minimiser(orc):
    rng = np.random.default_rng(3)
    m, n = 60, 80
    ptr, rows, vals = rand_csc(rng, m, n, 6)
    b = rng.standard_normal(m)
    s = orc.CD(0, m, n, ptr, rows, vals, b, 0.2)
    prev = s.eval()["primal"]
    for step in range(400):
        s.run(seed=9, tau=1, step0=step, nsteps=1)
        cur = s.eval()["primal"]
        assert cur <= prev + 1e-12 * abs(prev)
        prev = cur


def planted_lasso(rng, m, n, k, nsupp, lam):
    """Nesterov-style instance with a known minimiser x* (own construction
    after generator_nesterov.h:24-186): scale the columns of a random sparse A
    so that A^T r* = -lam sgn(x*) on the support and |A^T r*| < lam off it,
    then b = A x* - r*."""
    ptr, rows, vals = rand_csc(rng, m, n, k)
    r_star = rng.standard_normal(m)
    supp = rng.choice(n, nsupp, replace=False)
    x_star = np.zeros(n)
    x_star[supp] = rng.choice([-1.0, 1.0], nsupp) * rng.uniform(0.5, 2.0, nsupp)
    for j in range(n):
        sl = slice(ptr[j], ptr[j + 1])
        c = vals[sl] @ r_star[rows[sl]]
        if j in supp:
            vals[sl] *= -lam * np.sign(x_star[j]) / c
        else:
            vals[sl] *= lam * rng.uniform(0.1, 0.9) / abs(c)
    A = dense(m, n, ptr, rows, vals)
    b = A @ x_star - r_star
    F_star = 0.5 * r_star @ r_star + lam * np.abs(x_star).sum()
    return ptr, rows, vals, b, x_star, F_star


def test_lasso_planted_optimum(orc):
    rng = np.random.default_rng(4)
    m, n, lam = 120, 200, 0.5
    ptr, rows, vals, b, x_star, F_star = planted_lasso(rng, m, n, 8, 15, lam)
    s = orc.CD(0, m, n, ptr, rows, vals, b, lam)
    e_star = s.eval(x_star)
    assert e_star["primal"] == pytest.approx(F_star, rel=1e-13)
    assert abs(e_star["gap"]) <= 1e-11 * F_star          # dual point is exact at x*
    assert e_star["grad_inf"] == pytest.approx(lam, rel=1e-12)
    for epoch in range(400):
        s.run(seed=13, tau=1, step0=epoch * n, nsteps=n)
        e = s.eval()
        assert e["gap"] >= -1e-12 * F_star                 # weak duality
        assert e["gap"] >= (e["primal"] - F_star) - 1e-10 * F_star
    assert e["primal"] == pytest.approx(F_star, rel=1e-11)
    assert np.allclose(s.x, x_star, atol=1e-7)
    assert e["rel_gap"] < 1e-10


def test_lasso_gap_at_zero_closed_form(orc):
    rng = np.random.default_rng(5)
    m, n = 50, 70
    ptr, rows, vals = rand_csc(rng, m, n, 5)
    b = rng.standard_normal(m)
    A = dense(m, n, ptr, rows, vals)
    lmax = np.abs(A.T @ b).max()
    for frac in (0.1, 0.5, 0.99, 1.5):
        s = orc.CD(0, m, n, ptr, rows, vals, b, frac * lmax)
        e = s.eval()
        sc = min(1.0, frac)
        assert e["primal"] == pytest.approx(0.5 * b @ b, rel=1e-14)
        assert e["gap"] == pytest.approx(0.5 * (b @ b) * (1 - sc) ** 2, rel=1e-12, abs=1e-15)


def test_sync_step_is_jacobi(orc):
    """tau-step: every delta from the same state (dense Jacobi restatement)."""
    rng = np.random.default_rng(6)
    m, n, tau, lam = 30, 40, 8, 0.1
    ptr, rows, vals = rand_csc(rng, m, n, 5)
    b = rng.standard_normal(m)
    A = dense(m, n, ptr, rows, vals)
    s = orc.CD(0, m, n, ptr, rows, vals, b, lam, beta=2.0)
    s.run(seed=3, tau=tau, step0=0, nsteps=5)
    x0 = s.x.copy()
    S = orc.sample_steps(3, 5, 1, n, tau)[0]
    Lj = 1.0 / (2.0 * (A ** 2).sum(0))
    g = A.T @ (A @ x0 - b)
    want = x0.copy()
    for j in S:
        v = x0[j] - Lj[j] * g[j]
        want[j] = np.sign(v) * max(abs(v) - Lj[j] * lam, 0.0)
    s.run(seed=3, tau=tau, step0=5, nsteps=1)
    assert np.allclose(s.x, want, rtol=1e-13, atol=1e-15)
    assert np.allclose(s.r, A @ s.x - b, atol=1e-13)


def test_nnls_matches_scipy_golden(orc):
    g = np.load(os.path.join(HERE, "golden", "nnls_golden.npz"))
    for t in range(int(g["count"][0])):
        m, n = (int(v) for v in g[f"m{t}"])
        s = orc.CD(10, m, n, g[f"ptr{t}"], g[f"rows{t}"], g[f"vals{t}"], g[f"b{t}"], 0.0)
        for ep in range(3000):
            s.run(seed=21, tau=1, step0=ep * n, nsteps=n)
        e = s.eval()
        obj = float(g[f"obj{t}"][0])
        assert e["primal"] == pytest.approx(obj, rel=1e-8, abs=1e-12)
        assert (s.x >= 0).all()
        assert e["kkt"] < 1e-6


def test_svm_matches_golden(orc):
    g = np.load(os.path.join(HERE, "golden", "svm_golden.npz"))
    for t in range(int(g["count"][0])):
        ns, nf = (int(v) for v in g[f"shape{t}"])
        lam = float(g[f"lam{t}"][0])
        s = orc.SVM(ns, nf, g[f"ptr{t}"], g[f"cols{t}"], g[f"vals{t}"], g[f"y{t}"], lam)
        e0 = s.eval(g[f"alpha{t}"])
        assert e0["dual"] == pytest.approx(float(g[f"dual{t}"][0]), rel=1e-12)
        assert e0["primal"] == pytest.approx(float(g[f"primal{t}"][0]), rel=1e-12)
        for ep in range(4000):
            s.run(seed=17, tau=1, step0=ep * ns, nsteps=ns)
        e = s.eval()
        assert e["gap"] >= -1e-13
        assert e["dual"] == pytest.approx(float(g[f"dual{t}"][0]), rel=1e-8)
        assert e["rel_gap"] < 1e-7


def test_svm_w_identity(orc):
    rng = np.random.default_rng(8)
    ns, nf = 20, 30
    X = (rng.random((ns, nf)) < 0.3) * rng.random((ns, nf))
    rp = np.concatenate([[0], np.cumsum((X != 0).sum(1))]).astype(np.int32)
    ci = np.nonzero(X)[1].astype(np.int32)
    v = X[X != 0]
    y = rng.choice([-1.0, 1.0], ns)
    s = orc.SVM(ns, nf, rp, ci, v, y, 0.05)
    s.alpha[:] = rng.uniform(-0.5, 1.5, ns)
    s.recompute_w()
    a = np.clip(s.alpha, 0, 1)
    assert np.allclose(s.w, (a * y) @ X / (0.05 * ns), rtol=1e-14, atol=1e-16)


def test_svm_list_equals_serial_run(orc):
    rng = np.random.default_rng(14)
    ns, nf = 60, 40
    X = (rng.random((ns, nf)) < 0.25) * rng.standard_normal((ns, nf))
    X[3] = 0.0
    rp = np.concatenate([[0], np.cumsum((X != 0).sum(1))]).astype(np.int32)
    ci = np.nonzero(X)[1].astype(np.int32)
    v = X[X != 0]
    y = rng.choice([-1.0, 1.0], ns)
    a = orc.SVM(ns, nf, rp, ci, v, y, 0.02)
    s = orc.SVM(ns, nf, rp, ci, v, y, 0.02)
    a.run_list(orc.sample_steps(6, 0, 2000, ns, 1))
    s.run(seed=6, tau=1, step0=0, nsteps=2000)
    assert (a.alpha == s.alpha).all() and (a.w == s.w).all()
    assert a.alpha.min() >= 0 and a.alpha.max() <= 1 and a.alpha[3] == 0


def test_async_coords_multiply_shift_exact(orc):
    """The 64x32-bit multiply-shift map against exact integer arithmetic on
    the Philox words, and distinct workers get distinct streams."""
    n = 1_000_003
    seed = 0x1234_5678_9ABC
    for w in (0, 1, 77):
        c = orc.async_coords(seed, w, 5, 40, n)
        for k in range(5, 45):
            blk = k >> 1
            o = orc.philox([blk & 0xFFFFFFFF, blk >> 32, w, 0x50434441],
                           [seed & 0xFFFFFFFF, seed >> 32])
            lo, hi = (o[2], o[3]) if k & 1 else (o[0], o[1])
            assert c[k - 5] == ((hi << 32 | lo) * n) >> 64
    a, b = orc.async_coords(seed, 0, 0, 64, n), orc.async_coords(seed, 1, 0, 64, n)
    assert (a != b).any()


def test_async_coords_uniform(orc):
    from scipy.stats import chisquare
    n = 50
    c = np.concatenate([orc.async_coords(3, w, 0, 4000, n) for w in range(5)])
    assert c.min() >= 0 and c.max() < n
    assert chisquare(np.bincount(c, minlength=n)).pvalue > 1e-3


def test_cd_list_equals_serial_run(orc):
    rng = np.random.default_rng(12)
    m, n = 70, 90
    ptr, rows, vals = rand_csc(rng, m, n, 7)
    b = rng.standard_normal(m)
    a = orc.CD(0, m, n, ptr, rows, vals, b, 0.1)
    s = orc.CD(0, m, n, ptr, rows, vals, b, 0.1)
    a.run_list(orc.sample_steps(4, 0, 3000, n, 1))
    s.run(seed=4, tau=1, step0=0, nsteps=3000)
    assert (a.x == s.x).all() and (a.r == s.r).all()


def test_async_one_thread_equals_serial(orc):
    rng = np.random.default_rng(9)
    m, n = 70, 90
    ptr, rows, vals = rand_csc(rng, m, n, 7)
    b = rng.standard_normal(m)
    a = orc.CD(0, m, n, ptr, rows, vals, b, 0.1)
    s = orc.CD(0, m, n, ptr, rows, vals, b, 0.1)
    assert a.run_async(seed=4, k0=100, per_worker=5000, threads=1) == 1
    s.run_list(orc.async_coords(4, 0, 100, 5000, n))
    assert (a.x == s.x).all() and (a.r == s.r).all()


def test_async_threads_converge(orc):
    rng = np.random.default_rng(10)
    m, n, lam = 120, 200, 0.5
    ptr, rows, vals, b, x_star, F_star = planted_lasso(rng, m, n, 8, 15, lam)
    a = orc.CD(0, m, n, ptr, rows, vals, b, lam, beta=orc.eso_beta(
        orc.omega(m, rows), 4, n))
    a.run_async(seed=4, k0=0, per_worker=100 * n, threads=4)
    assert a.eval()["primal"] == pytest.approx(F_star, rel=1e-8)
