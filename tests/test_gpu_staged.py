"""Staged asynchronous workers (ACDC_LAUNCH_ASYNC_STAGED: a loader wave per
eight 64-lane workers copies records and entries into LDS ahead of use)
against the CPU FP64 oracle, with the same statements as the direct kernel:

(a) one worker (tau = 1) is the serial method: the FP64 trajectory over the
    replayed coordinate stream matches the oracle within 1e-12 relative, for
    segments inside the LDS slot and for segments longer than it (the slot
    holds 255 FP32 / 128 FP64 entries; the rest are read from global memory),
    for the Lasso, the NNLS and the SVM dual;
(b) many workers (tau not a multiple of the block's eight workers too):
    the FP64 oracle's evaluation of the device iterate reaches the gap
    target, and the device's FP64 report equals the oracle's to 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_SERIAL = 1e-12


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


@pytest.fixture(scope="module")
def c1(acdc):
    return acdc.config_lasso_c1()


def staged(acdc, p, loss, **kw):
    return acdc.Solver(p, loss, step_mode=acdc.STEP_ASYNC, launch_mode=acdc.LAUNCH_ASYNC_STAGED,
                       **kw)


def long_problem(acdc, lo, hi, seed=5, m=700, n=300):
    """Columns of lo..hi entries (every 17th empty) at distinct rows."""
    rng = np.random.default_rng(seed)
    cnt = rng.integers(lo, hi, n)
    cnt[::17] = 0
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(m, c, replace=False)) for c in cnt]).astype(np.int32)
    vals = rng.standard_normal(idx.size)
    b = rng.standard_normal(m)
    return m, n, ptr, idx, vals, b


def test_staged_rejects_other_lanes_and_sync(acdc, c1):
    p = c1.problem()
    with pytest.raises(acdc.AcdcError):
        staged(acdc, p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=8, lanes_per_coord=32)
    with pytest.raises(acdc.AcdcError):
        acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=8, step_mode=acdc.STEP_SYNC,
                    launch_mode=acdc.LAUNCH_ASYNC_STAGED)


def test_staged_one_worker_is_serial_fp64(acdc, orc, c1):
    p = c1.problem()
    s = staged(acdc, p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=1, seed=21)
    o = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam)
    k = 0
    for chunk in (1, 7, 3000, 2 * c1.n):  # short launches exercise the pipeline's ends
        s.run_updates(chunk)
        o.run_list(orc.async_coords(21, 0, k, chunk, c1.n))
        k += chunk
        assert rel_inf(s.x, o.x) < TOL_SERIAL, k
        assert rel_inf(s.residuals, o.r) < TOL_SERIAL, k
    s.close()


@pytest.mark.parametrize("lo,hi", [(1, 120), (100, 400)])
def test_staged_one_worker_segments_past_the_slot(acdc, orc, lo, hi):
    """(1, 120): inside both slots; (100, 400): most columns longer than the
    FP64 slot (128) and many longer than the FP32 one (255)."""
    m, n, ptr, idx, vals, b = long_problem(acdc, lo, hi)
    p = acdc.Problem.from_csc(m, n, ptr, idx, vals, b)
    for loss in (acdc.LOSS_L1_SQUARE, acdc.LOSS_NNLS):
        s = staged(acdc, p, loss, lam=0.3, tau=1, seed=8)
        o = orc.CD(loss, m, n, ptr, idx, vals, b, 0.3 if loss == 0 else 0.0)
        s.run_updates(5 * n)
        o.run_list(orc.async_coords(8, 0, 0, 5 * n, n))
        assert rel_inf(s.x, o.x) < TOL_SERIAL, loss
        assert rel_inf(s.residuals, o.r) < TOL_SERIAL, loss
        s.close()


def test_staged_one_worker_svm_fp64(acdc, orc):
    d = acdc.gen_svm(400, 20_000, row_median=150.0, row_sigma=0.9, wstar_nnz=500,
                     flip_frac=0.1, seed=3)
    assert np.diff(d.ptr).max() > 128  # rows past the FP64 slot
    p = d.problem()
    s = staged(acdc, p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=1, seed=12)
    o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    s.run_updates(6 * d.n)
    o.run_list(orc.async_coords(12, 0, 0, 6 * d.n, d.n))
    assert rel_inf(s.x, o.alpha) < TOL_SERIAL
    assert rel_inf(s.residuals, o.w) < TOL_SERIAL
    s.close()


@pytest.mark.parametrize("prec,tau", [("fp64", 128), ("fp32", 128), ("fp32", 131)])
def test_staged_workers_reach_fp64_optimum(acdc, orc, c1, prec, tau):
    p = c1.problem()
    precision = acdc.PREC_FP64 if prec == "fp64" else acdc.PREC_FP32_REFRESH
    s = staged(acdc, p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=3, precision=precision,
               refresh_period=5)
    s.run_updates(150 * c1.n)
    rep = s.report()
    e = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam).eval(s.x)
    assert e["rel_gap"] <= 1e-6, e["rel_gap"]
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-3, abs=1e-12)
    s.close()


def test_staged_workers_long_segments_fp32(acdc, orc):
    """Columns of 100..400 entries (most past the FP32 slot) with 64 FP32
    workers: after the same epochs the staged kernel's iterate is as good as
    the direct kernel's (oracle gap within 10x), and the device's report of
    it equals the oracle's."""
    m, n, ptr, idx, vals, b = long_problem(acdc, 100, 400, seed=6, m=6000, n=1500)
    lam = 0.3 * float(np.max(np.abs(np.bincount(np.repeat(np.arange(n), np.diff(ptr)),
                                                weights=vals * b[idx], minlength=n))))
    p = acdc.Problem.from_csc(m, n, ptr, idx, vals, b)
    o = orc.CD(0, m, n, ptr, idx, vals, b, lam)
    gaps = {}
    for mode in (acdc.LAUNCH_ASYNC_DIRECT, acdc.LAUNCH_ASYNC_STAGED):
        s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=lam, tau=64, seed=9,
                        step_mode=acdc.STEP_ASYNC, launch_mode=mode, lanes_per_coord=64,
                        precision=acdc.PREC_FP32_REFRESH, refresh_period=5)
        s.run_updates(100 * n)
        e = o.eval(s.x)
        rep = s.report()
        assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
        gaps[mode] = e["rel_gap"]
        s.close()
    assert gaps[acdc.LAUNCH_ASYNC_STAGED] <= 10 * gaps[acdc.LAUNCH_ASYNC_DIRECT] + 1e-9, gaps


def test_staged_svm_and_nnls_workers(acdc, orc):
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=3)
    p = d.problem()
    s = staged(acdc, p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=128, seed=10,
               precision=acdc.PREC_FP32_REFRESH, refresh_period=5)
    s.run_updates(60 * d.n)
    o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    a = s.x
    assert a.min() >= 0.0 and a.max() <= 1.0
    e = o.eval(a)
    rep = s.report()
    assert e["gap"] <= 1e-5, e["gap"]
    assert rep["gap"] == pytest.approx(e["gap"], rel=1e-6, abs=1e-12)
    s.close()
    t = acdc.gen_ttd(10)
    pt = t.problem()
    s = staged(acdc, pt, acdc.LOSS_NNLS, lam=0.0, tau=256, seed=6,
               precision=acdc.PREC_FP32_REFRESH, refresh_period=5)
    s.run_updates(300 * t.n)
    o = orc.CD(10, t.m, t.n, t.ptr, t.idx, t.vals, t.rhs, 0.0, beta=s.beta)
    p0 = o.eval()["primal"]
    x = s.x
    e = o.eval(x)
    rep = s.report()
    assert x.min() >= 0.0
    assert e["primal"] <= 5e-3 * p0
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    s.close()
