"""CPU-side checks of the C ABI library: it loads, exports every declared
entry point, validates its inputs, and its host-side generators produce the
configured instances.  No kernel runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("acdc_hip.h", "acdc_gen.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(acdc_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported(acdc):
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(acdc.lib, n), n


def test_problem_validation(acdc):
    ptr = np.array([0, 2, 3], np.int32)
    with pytest.raises(acdc.AcdcError, match="index out of range"):
        acdc.Problem.from_csc(3, 2, ptr, [0, 5, 1], [1.0, 2.0, 3.0], np.zeros(3))
    with pytest.raises(acdc.AcdcError, match="ptr"):
        acdc.Problem.from_csc(3, 2, np.array([0, 2, 1], np.int32), [0, 1, 1], [1.0, 2.0, 3.0],
                              np.zeros(3))
    p = acdc.Problem.from_csc(3, 2, ptr, [0, 2, 2], [1.0, 2.0, 3.0], np.zeros(3))
    assert p.omega == 2  # row 2 holds two entries
    p.close()


def test_solver_without_device_fails_loudly(acdc):
    if acdc.device_count() > 0:
        pytest.skip("a gfx950 device is present")
    p = acdc.Problem.from_csc(3, 2, [0, 2, 3], [0, 2, 2], [1.0, 2.0, 3.0], np.ones(3))
    with pytest.raises(acdc.AcdcError, match="ACDC_ERR_NO_DEVICE"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1)


def test_solver_argument_errors(acdc):
    p = acdc.Problem.from_csc(3, 2, [0, 2, 3], [0, 2, 2], [1.0, 2.0, 3.0], np.ones(3))
    with pytest.raises(acdc.AcdcError, match="tau"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, tau=3)
    with pytest.raises(acdc.AcdcError, match="loss"):
        acdc.Solver(p, 1, lam=0.1)
    with pytest.raises(acdc.AcdcError, match="SVM"):
        acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=0.1)
    with pytest.raises(acdc.AcdcError, match="lanes"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, lanes_per_coord=12)
    # a block per coordinate: synchronous steps only
    with pytest.raises(acdc.AcdcError, match="block per coordinate"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, lanes_per_coord=256,
                    step_mode=acdc.STEP_ASYNC)
    # the streamed scatter: synchronous steps, tau >= 2, n coefficients in LDS
    with pytest.raises(acdc.AcdcError, match="streamed scatter"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, tau=1,
                    launch_mode=acdc.LAUNCH_PER_STEP_STREAMED)
    with pytest.raises(acdc.AcdcError, match="streamed scatter"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, tau=2, step_mode=acdc.STEP_ASYNC,
                    launch_mode=acdc.LAUNCH_PER_STEP_STREAMED)
    n = 21_000  # 168 KB of FP64 coefficients
    wide = acdc.Problem.from_csc(1, n, np.arange(n + 1), np.zeros(n), np.ones(n), np.ones(1))
    with pytest.raises(acdc.AcdcError, match="streamed scatter"):
        acdc.Solver(wide, acdc.LOSS_L1_SQUARE, lam=0.1, tau=2,
                    launch_mode=acdc.LAUNCH_PER_STEP_STREAMED)
    with pytest.raises(acdc.AcdcError, match="launch mode"):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.1, tau=2, launch_mode=4)


def test_batch_argument_errors(acdc):
    """acdc_batch_create checks its arguments before it looks for a device."""
    csc = acdc.Problem.from_csc(3, 2, [0, 2, 3], [0, 2, 2], [1.0, 2.0, 3.0], np.ones(3))
    csr = acdc.Problem.from_csr(2, 3, [0, 2, 3], [0, 2, 2], [1.0, 2.0, 3.0], [1.0, -1.0])
    lam = np.full((1, 2), 0.1)
    with pytest.raises(acdc.AcdcError, match="loss must be"):
        acdc.Batch([csc], lam, loss=1)
    with pytest.raises(acdc.AcdcError, match="SVM problems must be sample rows"):
        acdc.Batch([csc], lam, loss=acdc.LOSS_SVM_DUAL_HINGE)
    with pytest.raises(acdc.AcdcError, match="must be column-major"):
        acdc.Batch([csr], lam, loss=acdc.LOSS_L1_SQUARE)
    with pytest.raises(acdc.AcdcError, match="SVM lambdas must be > 0"):
        acdc.Batch([csr], np.array([[0.1, 0.0]]), loss=acdc.LOSS_SVM_DUAL_HINGE)
    with pytest.raises(acdc.AcdcError, match="lambdas must be >= 0"):
        acdc.Batch([csc], np.array([[0.1, -1.0]]))
    with pytest.raises(acdc.AcdcError, match="lanes_per_coord"):
        acdc.Batch([csc], lam, lanes_per_coord=12)
    with pytest.raises(acdc.AcdcError, match="beta_mode"):
        acdc.Batch([csc], lam, beta_mode=acdc.BETA_FORCED)
    for it in (-1, 10001):
        with pytest.raises(acdc.AcdcError, match="ws_newton_iters"):
            acdc.Batch([csc], lam, screening=True, ws_newton_iters=it)
    with pytest.raises(acdc.AcdcError, match="SVM problems must be sample rows"):
        acdc.batch_tune([csc], loss=acdc.LOSS_SVM_DUAL_HINGE)


def test_config1_generator(acdc):
    d = acdc.config_lasso_c1()
    assert (d.m, d.n, d.nnz) == (2000, 10000, 200000)
    cnt = np.diff(d.ptr)
    assert (cnt == 20).all()
    for j in range(0, d.n, 997):
        r = d.idx[d.ptr[j]:d.ptr[j + 1]]
        assert (np.diff(r) > 0).all() and r.min() >= 0 and r.max() < d.m
    norms = np.sqrt(np.add.reduceat(d.vals ** 2, d.ptr[:-1]))
    assert np.allclose(norms, 1.0, atol=1e-14)
    assert np.count_nonzero(d.xstar) == 100 and set(np.abs(d.xstar[d.xstar != 0])) == {1.0}
    # b = A x* + 0.01 N(0,1); lambda = 0.1 ||A^T b||_inf
    ax = np.zeros(d.m)
    col = np.repeat(np.arange(d.n), cnt)
    np.add.at(ax, d.idx, d.vals * d.xstar[col])
    noise = d.rhs - ax
    assert 0.0085 < noise.std() < 0.0115 and abs(noise.mean()) < 0.002
    atb = np.zeros(d.n)
    np.add.at(atb, col, d.vals * d.rhs[d.idx])
    assert d.lambda_max == pytest.approx(np.abs(atb).max(), rel=1e-12)
    assert d.lam == pytest.approx(0.1 * d.lambda_max, rel=1e-15)
    # a pure function of the seed
    e = acdc.config_lasso_c1()
    assert (e.idx == d.idx).all() and (e.vals == d.vals).all()
    f = acdc.config_lasso_c1(seed=99)
    assert not (f.idx == d.idx).all()


def test_truncated_zipf_law(acdc):
    d = acdc.gen_lasso(50_000, 200_000, nnz_law=acdc.NNZ_ZIPF_TRUNC, nnz_lo=5, nnz_hi=200,
                       zipf_s=1.1, xstar_nnz=10, seed=2)
    cnt = np.diff(d.ptr)
    k = np.arange(5, 201)
    p = k ** -1.1
    p /= p.sum()
    assert cnt.min() >= 5 and cnt.max() <= 200
    assert cnt.mean() == pytest.approx((k * p).sum(), rel=0.01)       # E = 46.74
    assert (cnt == 5).mean() == pytest.approx(p[0], rel=0.03)


def test_clipped_zipf_law(acdc):
    d = acdc.gen_lasso(5_000, 100_000, nnz_law=acdc.NNZ_ZIPF_CLIP, nnz_lo=3, nnz_hi=100,
                       zipf_s=1.1, xstar_nnz=10, seed=5)
    cnt = np.diff(d.ptr)
    # literal clip of Zipf(1.1) to [3,100]: E = 66.4, P(k = 100) = P(Zipf >= 100)
    assert cnt.mean() == pytest.approx(66.4, rel=0.02)


def test_ttd_ground_structure(acdc):
    t = acdc.gen_ttd(40)
    assert (t.m, t.n, t.nnz) == (201_720, 844_840, 3_460_479)
    assert (t.rhs != 0).sum() == 41 * 41 and set(t.rhs[t.rhs != 0]) == {-1.0}
    # the top-face loads sit in the z block of the rows
    nfree = t.m // 3
    assert (np.nonzero(t.rhs)[0] >= 2 * nfree).all()
    mags = np.unique(np.round(np.abs(t.vals), 12))
    assert set(mags) <= {1.0, 0.5, round(1 / 3, 12)}
    cnt = np.diff(t.ptr)
    assert cnt.min() >= 1 and cnt.max() == 6
    # a bar with both ends free has column sum zero (equal and opposite end forces)
    sums = np.add.reduceat(t.vals, t.ptr[:-1])
    assert np.abs(sums[cnt == 6]).max() < 1e-15
    small = acdc.gen_ttd(2)
    # 27 nodes, 9 fixed: 54 rows; 158 bars, 20 inside the fixed face
    assert (small.m, small.n) == (54, 138)


def test_svm_generator(acdc):
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=3)
    assert d.kind == 1 and (d.m, d.n) == (50_000, 2000)
    cnt = np.diff(d.ptr)
    assert np.median(cnt) == pytest.approx(300, rel=0.06)
    assert cnt.mean() == pytest.approx(450, rel=0.06)
    norms = np.sqrt(np.add.reduceat(d.vals ** 2, d.ptr[:-1]))
    assert np.allclose(norms, 1.0, atol=1e-14) and (d.vals > 0).all()
    assert set(np.unique(d.rhs)) == {-1.0, 1.0}
    rows = np.repeat(np.arange(d.n), cnt)
    margin = np.zeros(d.n)
    np.add.at(margin, rows, d.vals * d.xstar[d.idx])
    clean = np.where(margin < 0, -1.0, 1.0)
    assert (clean != d.rhs).mean() == pytest.approx(0.1, abs=0.025)
    assert d.lam == pytest.approx(1 / 2000)
