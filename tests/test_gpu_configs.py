"""The five benchmark configurations at their stated sizes, through the C ABI,
checked against the CPU FP64 oracle on the same data.

C1  2,000 x 10,000 Lasso, tau = 1: the FP64 trajectory over 100 epochs
    (100 n updates) within 1e-12 of the serial oracle, epoch by epoch, with
    the objective and gap logged per epoch equal to the oracle's.
C2  200,000 x 1,000,000 Lasso (4.67e7 nnz), tau = 4096: synchronous FP64
    steps within 1e-10 of the oracle's tau-step; FP32 asynchronous workers
    (direct and staged kernels) until the oracle's evaluation of the device
    x has relative gap <= 1e-6, the device's FP64 report within 1e-9 of it.
C3  20,000 x 1,350,000 hinge-SVM dual: FP64 b = 512 minibatches within 1e-10
    of the oracle's SDCA steps; FP32 b = 8192 (streamed scatter) to relative
    gap <= 1e-6 as the oracle evaluates the device alpha.
C4  40^3-cell truss ground structure (844,840 bars, 201,720 DOFs) NNLS:
    synchronous FP64 tau = 1024 and 8192 steps within 1e-10 of the oracle;
    FP32 asynchronous epochs whose objective and KKT residual the oracle
    reproduces from the device x.
C5  the 256-instance batched lambda-path workload in one graph launch, as
    measured (working sets, support Newton steps, 500 epochs per pair at
    most): the (instance, lambda) pairs reach relative gap 1e-6 (all of
    them in the measured runs; at most 8 of 8,192 may end short), and
    eight instances (nnz quantiles) are checked by the oracle at the last
    lambda of their path -- primal within 1e-9 and gap within 1e-6
    (relative) of the device's report, and the oracle's relative gap
    <= 1e-6.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


def log(msg):
    print(msg, flush=True)


# --------------------------------------------------------------------- C1
def test_c1_tau1_trajectory_100_epochs(acdc, orc):
    d = acdc.config_lasso_c1()
    p = d.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=1, seed=1)
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    worst = 0.0
    for epoch in range(100):
        s.run_updates(d.n)
        o.run(seed=1, tau=1, step0=epoch * d.n, nsteps=d.n)
        if epoch % 10 == 9 or epoch < 3:
            err = max(rel_inf(s.x, o.x), rel_inf(s.residuals, o.r))
            worst = max(worst, err)
            assert err < 1e-12, (epoch, err)
        rep, e = s.report(), o.eval()
        assert rep["primal"] == pytest.approx(e["primal"], rel=1e-12), epoch
        assert rep["gap"] == pytest.approx(e["gap"], rel=1e-9, abs=1e-12), epoch
    log(f"C1: 100 epochs, worst relative deviation {worst:.2e}, final rel gap {rep['rel_gap']:.3e}")
    s.close()


# --------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def c2(acdc):
    t = time.perf_counter()
    d = acdc.config_lasso_c2()
    log(f"C2 generated: {d.m} x {d.n}, nnz {d.nnz} in {time.perf_counter() - t:.1f}s")
    return d


@pytest.fixture(scope="module")
def c2_problem(c2):
    p = c2.problem()
    yield p
    p.close()


def test_c2_sync_fp64_steps_match_oracle(acdc, orc, c2, c2_problem):
    tau, steps = 4096, 6
    s = acdc.Solver(c2_problem, acdc.LOSS_L1_SQUARE, lam=c2.lam, tau=tau, seed=2,
                    launch_mode=acdc.LAUNCH_PER_STEP)
    o = orc.CD(0, c2.m, c2.n, c2.ptr, c2.idx, c2.vals, c2.rhs, c2.lam, beta=s.beta)
    s.run_updates(steps * tau)
    o.run(seed=2, tau=tau, step0=0, nsteps=steps)
    assert int(np.count_nonzero(o.x)) > 0
    assert rel_inf(s.x, o.x) < 1e-10
    assert rel_inf(s.residuals, o.r) < 1e-10
    rep, e = s.report(), o.eval()
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-10)
    assert rep["gap"] == pytest.approx(e["gap"], rel=1e-8)
    s.close()


@pytest.mark.parametrize("launch", ["direct", "staged"])
def test_c2_fp32_async_reaches_oracle_gap(acdc, orc, c2, c2_problem, launch):
    mode = acdc.LAUNCH_ASYNC_DIRECT if launch == "direct" else acdc.LAUNCH_ASYNC_STAGED
    s = acdc.Solver(c2_problem, acdc.LOSS_L1_SQUARE, lam=c2.lam, tau=4096, seed=2,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                    step_mode=acdc.STEP_ASYNC, lanes_per_coord=64, launch_mode=mode)
    epochs, rep = 0, s.report()
    while rep["rel_gap"] > 1e-6 and epochs < 400:
        for _ in range(5):
            s.run_updates(c2.n)
        epochs += 5
        rep = s.report()
    e = orc.CD(0, c2.m, c2.n, c2.ptr, c2.idx, c2.vals, c2.rhs, c2.lam).eval(s.x)
    log(f"C2 {launch}: {epochs} epochs, device rel gap {rep['rel_gap']:.3e}, "
        f"oracle rel gap {e['rel_gap']:.3e}")
    assert e["rel_gap"] <= 1e-6
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["gap"] == pytest.approx(e["gap"], rel=1e-6)
    s.close()


# --------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3(acdc):
    t = time.perf_counter()
    d = acdc.config_svm_c3()
    log(f"C3 generated: {d.n} samples x {d.m} features, nnz {d.nnz} in "
        f"{time.perf_counter() - t:.1f}s")
    return d


def test_c3_fp64_b512_matches_oracle(acdc, orc, c3):
    p = c3.problem()
    b, steps = 512, 12
    s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=c3.lam, tau=b, seed=3)
    o = orc.SVM(c3.n, c3.m, c3.ptr, c3.idx, c3.vals, c3.rhs, c3.lam, beta=s.beta)
    s.run_updates(steps * b)
    o.run(seed=3, tau=b, step0=0, nsteps=steps)
    assert rel_inf(s.x, o.alpha) < 1e-10
    assert rel_inf(s.residuals, o.w) < 1e-10
    rep, e = s.report(), o.eval()
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-10)
    assert rep["dual"] == pytest.approx(e["dual"], rel=1e-10)
    s.close()
    p.close()


def test_c3_fp32_b8192_streamed_reaches_oracle_gap(acdc, orc, c3):
    p = c3.problem()
    s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=c3.lam, tau=8192, seed=3,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                    beta_mode=acdc.BETA_SPECTRAL)
    epochs, rep = 0, s.report()
    while rep["rel_gap"] > 1e-6 and epochs < 200:
        for _ in range(5):
            s.run_updates(c3.n)
        epochs += 5
        rep = s.report()
    a = s.x
    e = orc.SVM(c3.n, c3.m, c3.ptr, c3.idx, c3.vals, c3.rhs, c3.lam).eval(a)
    log(f"C3 b=8192: {epochs} epochs, oracle rel gap {e['rel_gap']:.3e}")
    assert a.min() >= 0.0 and a.max() <= 1.0
    assert e["rel_gap"] <= 1e-6
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["gap"] == pytest.approx(e["gap"], rel=1e-6)
    s.close()
    p.close()


# --------------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def c4(acdc):
    d = acdc.gen_ttd(40)
    assert (d.n, d.m, d.nnz) == (844_840, 201_720, 3_460_479)
    return d


@pytest.mark.parametrize("tau", [1024, 8192])
def test_c4_sync_fp64_steps_match_oracle(acdc, orc, c4, tau):
    p = c4.problem()
    steps = 8
    s = acdc.Solver(p, acdc.LOSS_NNLS, lam=0.0, tau=tau, seed=4, launch_mode=acdc.LAUNCH_PER_STEP)
    o = orc.CD(10, c4.m, c4.n, c4.ptr, c4.idx, c4.vals, c4.rhs, 0.0, beta=s.beta)
    s.run_updates(steps * tau)
    o.run(seed=4, tau=tau, step0=0, nsteps=steps)
    assert rel_inf(s.x, o.x) < 1e-10
    assert rel_inf(s.residuals, o.r) < 1e-10
    rep, e = s.report(), o.eval()
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-10)
    assert rep["kkt"] == pytest.approx(e["kkt"], rel=1e-8, abs=1e-14)
    s.close()
    p.close()


def test_c4_fp32_async_kkt_matches_oracle(acdc, orc, c4):
    p = c4.problem()
    s = acdc.Solver(p, acdc.LOSS_NNLS, lam=0.0, tau=8192, seed=4,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                    step_mode=acdc.STEP_ASYNC)
    p0 = s.report()["primal"]
    s.run_updates(50 * c4.n)
    x = s.x
    rep = s.report()
    e = orc.CD(10, c4.m, c4.n, c4.ptr, c4.idx, c4.vals, c4.rhs, 0.0).eval(x)
    log(f"C4 fp32 async: objective {e['primal']:.6e} from {p0:.6e}, KKT {e['kkt']:.3e}")
    assert x.min() >= 0.0
    assert e["primal"] < 0.1 * p0
    assert rep["primal"] == pytest.approx(e["primal"], rel=1e-9)
    assert rep["kkt"] == pytest.approx(e["kkt"], rel=1e-6, abs=1e-14)
    s.close()
    p.close()


# --------------------------------------------------------------------- C5
def test_c5_batched_workload_checked_by_oracle(acdc, orc):
    t = time.perf_counter()
    ds = acdc.config_batch_c5()
    nnz = np.array([d.nnz for d in ds])
    log(f"C5 generated: {len(ds)} instances, nnz {nnz.sum()} in {time.perf_counter() - t:.1f}s")
    probs = [d.problem() for d in ds]
    lams = np.stack([acdc.lambda_path(d.lambda_max, 32, 1e-3) for d in ds])
    # the measured configuration: working sets with Anderson extrapolation
    # and support Newton steps, 500 epochs of n per pair at most (about a
    # minute of device time)
    b = acdc.Batch(probs, lams, seed=5, precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                   max_epochs=500, gap_tol=1e-6, use_graph=True, step_mode=acdc.STEP_ASYNC,
                   lanes_per_coord=8, tau=128, beta_mode=acdc.BETA_SPECTRAL, screening=True)
    ms = b.run()
    res = b.results()
    log(f"C5: {ms:.0f} ms device, {res['converged'].mean():.4f} of pairs converged, "
        f"{int(res['newton_steps'].sum())} Newton steps kept")
    # every pair converged in the measured runs; the asynchronous workers'
    # interleaving is not reproducible, so up to 0.1% of the pairs (8) may
    # end short without failing the run
    short = np.argwhere(res["converged"] == 0)
    assert len(short) <= 8, short[:10]
    order = np.argsort(nnz)
    sample = [int(order[int(q * (len(ds) - 1))]) for q in np.linspace(0.0, 1.0, 8)]
    for i in sample:
        d = ds[i]
        x = b.x(i)  # the x of the last lambda
        e = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i, -1]).eval(x)
        r = res[i, -1]
        assert r["primal"] == pytest.approx(e["primal"], rel=1e-9), i
        assert r["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-6, abs=1e-12), i
        assert e["rel_gap"] <= 1e-6, (i, e["rel_gap"])
    b.close()
    for p in probs:
        p.close()
