"""Asynchronous workers (ACDC_STEP_ASYNC) against the CPU FP64 oracle.

(a) each worker's coordinate stream is bit-exact against the host replay;
(b) one worker (tau = 1) is the serial method: the FP64 trajectory over the
    replayed coordinates matches the oracle within 1e-12 relative, through
    both the cached-entry kernel and the long-segment (tail) kernel;
(c) with tau > 1 workers the interleaving is not reproducible, so the
    statement is about the limit: after a fixed number of epochs the FP64
    oracle's evaluation of the device iterate has relative duality gap
    <= 1e-6 (Lasso) or gap <= 1e-5 (SVM), the NNLS objective is within 1.5x
    of the oracle's synchronous run after the same epochs, and the device's
    own FP64 report agrees with the oracle's to 1e-9.
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


@pytest.mark.parametrize("n,tau,steps", [(1, 1, 5), (1000, 1, 300), (10007, 37, 20),
                                         (1_000_000, 4096, 3)])
def test_async_sampler_replay_bit_exact(acdc, orc, n, tau, steps):
    ptr = np.arange(n + 1, dtype=np.int32)
    p = acdc.Problem.from_csc(1, n, ptr, np.zeros(n, np.int32), np.ones(n), np.ones(1))
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.0, tau=tau, seed=999,
                    step_mode=acdc.STEP_ASYNC)
    got = s.sample_coordinates(5, steps)
    for w in sorted({0, tau // 2, tau - 1}):
        assert (got[:, w] == orc.async_coords(999, w, 5, steps, n)).all(), w
    s.close()


@pytest.mark.parametrize("lanes", [8, 64])
def test_async_one_worker_is_serial_fp64(acdc, orc, c1, lanes):
    # C1 columns hold 20 entries: 8 lanes x 4 cached chunks = 32 (no tail),
    # and the 64-lane kernel likewise; the tail kernel is covered below
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=1, seed=21,
                    step_mode=acdc.STEP_ASYNC, lanes_per_coord=lanes)
    o = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam)
    k = 0
    for chunk in (7, 3000, 2 * c1.n):  # odd counts exercise the pipeline's ends
        s.run_updates(chunk)
        o.run_list(orc.async_coords(21, 0, k, chunk, c1.n))
        k += chunk
        assert rel_inf(s.x, o.x) < TOL_SERIAL, k
        assert rel_inf(s.residuals, o.r) < TOL_SERIAL, k
    s.close()


def test_async_one_worker_long_segments(acdc, orc):
    """Columns longer than the cached entries take the tail kernel."""
    rng = np.random.default_rng(5)
    m, n = 400, 300
    cnt = rng.integers(1, 120, n)
    cnt[::17] = 0
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(m, c, replace=False)) for c in cnt]).astype(np.int32)
    vals = rng.standard_normal(idx.size)
    b = rng.standard_normal(m)
    p = acdc.Problem.from_csc(m, n, ptr, idx, vals, b)
    for loss in (acdc.LOSS_L1_SQUARE, acdc.LOSS_NNLS):
        s = acdc.Solver(p, loss, lam=0.3, tau=1, seed=8, step_mode=acdc.STEP_ASYNC,
                        lanes_per_coord=8)
        o = orc.CD(loss, m, n, ptr, idx, vals, b, 0.3 if loss == 0 else 0.0)
        s.run_updates(5 * n)
        o.run_list(orc.async_coords(8, 0, 0, 5 * n, n))
        assert rel_inf(s.x, o.x) < TOL_SERIAL, loss
        assert rel_inf(s.residuals, o.r) < TOL_SERIAL, loss
        s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_async_workers_reach_fp64_optimum(acdc, orc, c1, prec):
    p = c1.problem()
    precision = acdc.PREC_FP64 if prec == "fp64" else acdc.PREC_FP32_REFRESH
    # 128 workers on 10,000 coordinates (tau/n = 1.3%); the ESO beta of tau
    # = 128 is 2.8 here and 150 epochs reach the FP64 floor of the gap
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=128, seed=3,
                    step_mode=acdc.STEP_ASYNC, precision=precision, refresh_period=5)
    s.run_updates(150 * c1.n)
    rep = s.report()
    e = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam).eval(s.x)
    assert e["rel_gap"] <= 1e-6, e["rel_gap"]
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-3, abs=1e-12)
    # the serial oracle run to its optimum agrees on the objective
    o = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam)
    o.run(seed=3, tau=1, step0=0, nsteps=150 * c1.n)
    assert e["primal"] == pytest.approx(o.eval()["primal"], rel=1e-6)
    s.close()


def test_async_nnls_ttd_fp32(acdc, orc):
    """TTD ground structure (10^3 cells, 13,810 bars), 256 workers, FP32 +
    refresh: after 300 epochs the objective is within 1.5x of the oracle's
    synchronous tau = 256 run after the same epochs (both ~0.2% of the
    objective at x = 0), x stays nonnegative up to racing same-bar updates,
    and the device's FP64 report equals the oracle's evaluation of x."""
    t = acdc.gen_ttd(10)
    p = t.problem()
    s = acdc.Solver(p, acdc.LOSS_NNLS, lam=0.0, tau=256, seed=6, step_mode=acdc.STEP_ASYNC,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5)
    s.run_updates(300 * t.n)
    o = orc.CD(10, t.m, t.n, t.ptr, t.idx, t.vals, t.rhs, 0.0, beta=s.beta)
    p0 = o.eval()["primal"]
    o.run(seed=6, tau=256, step0=0, nsteps=300 * t.n // 256)
    e_sync = o.eval()
    x = s.x
    e = o.eval(x)
    rep = s.report()
    assert x.min() > -1e-9 * max(1.0, x.max())
    assert e["primal"] <= 1.5 * e_sync["primal"], (e["primal"], e_sync["primal"])
    assert e["primal"] <= 5e-3 * p0
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["kkt"] == pytest.approx(e["kkt"], rel=1e-6, abs=1e-12)
    s.close()


def test_async_svm(acdc, orc):
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=3)
    p = d.problem()
    s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=128, seed=10,
                    step_mode=acdc.STEP_ASYNC)
    s.run_updates(60 * d.n)
    o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    o.alpha[:] = s.x
    e = o.eval()
    rep = s.report()
    assert e["gap"] <= 1e-5, e["gap"]
    assert rep["gap"] == pytest.approx(e["gap"], rel=1e-6, abs=1e-12)
    s.close()


def test_async_warm_start_and_reports(acdc, orc, c1):
    """set_x / set_lambda between asynchronous bulks (the lambda path)."""
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lambda_max * 0.5, tau=64, seed=4,
                    step_mode=acdc.STEP_ASYNC)
    s.run_updates(20 * c1.n)
    x_prev = s.x
    s.set_lambda(c1.lam)
    s.x = x_prev
    r = orc.csc_residual(c1.m, c1.n, c1.ptr, c1.idx, c1.vals, x_prev, c1.rhs)
    assert rel_inf(s.residuals, r) < 1e-13
    s.run_updates(60 * c1.n)
    e = orc.CD(0, c1.m, c1.n, c1.ptr, c1.idx, c1.vals, c1.rhs, c1.lam).eval(s.x)
    assert e["rel_gap"] <= 1e-6
    s.close()


def test_async_fp32_svm_set_lambda_keeps_alpha(acdc, orc):
    """FP32 asynchronous SVM: run_updates with no refresh, set_lambda, more
    updates.  The live alpha is in the workers' records when set_lambda
    rescales w; the running w afterwards must still equal
    A^T (alpha y) / (lambda n) for the device's alpha (FP32 corrections:
    1e-4 of |w|), and the oracle's gap of that alpha matches the report."""
    d = acdc.gen_svm(1500, 30_000, row_median=100.0, row_sigma=0.9, wstar_nnz=500,
                     flip_frac=0.1, seed=4)
    p = d.problem()
    s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=32, seed=7,
                    step_mode=acdc.STEP_ASYNC, precision=acdc.PREC_FP32_REFRESH,
                    refresh_period=0)
    s.run_updates(3 * d.n + 5)
    s.set_lambda(0.5 * d.lam)
    s.run_updates(3 * d.n)
    w_dev = s.residuals
    a = s.x
    o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, 0.5 * d.lam)
    o.alpha[:] = a
    o.recompute_w()
    assert float(np.max(np.abs(w_dev - o.w))) <= 1e-4 * float(np.max(np.abs(o.w)))
    rep = s.report()
    assert rep["gap"] == pytest.approx(o.eval()["gap"], rel=1e-9, abs=1e-12)
    s.close()
