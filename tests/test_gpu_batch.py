"""Batched multi-instance Lasso, NNLS and SVM (acdc_batch_*) against the FP64 oracle.

Instance i of a batch samples with key seed + i and its own tau and ESO beta,
so its trajectory is the oracle's synchronous run with those parameters; the
lambda path is warm-started, so the oracle carries x from one lambda to the
next.  Covers both residual placements (LDS for short instances, global for
one taller than the LDS holds), FP64 (tight tolerance) and FP32 + anchored
refresh (converges to the FP64 optimum), and graph replays against an eager
run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(acdc):
    shapes = [(300, 1000, 6), (1200, 2500, 10), (800, 4000, 4), (2500, 1500, 20),
              (45000, 2000, 30)]  # the last one does not fit the LDS residual
    ds = []
    for i, (m, n, k) in enumerate(shapes):
        ds.append(acdc.gen_lasso(m, n, nnz_law=acdc.NNZ_FIXED, nnz_fixed=k,
                                 xstar_nnz=max(1, n // 100), noise=0.01, lambda_frac=1.0,
                                 seed=100 + i))
    return ds


def lambdas_of(acdc, ds, path_len, ratio):
    return np.stack([acdc.lambda_path(d.lambda_max, path_len, ratio) for d in ds])


def oracle_path(orc, d, lams, seed, tau, beta, epochs):
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[0], beta=beta)
    steps = epochs * (-(-d.n // tau))
    step0 = 0
    for lam in lams:
        o.lam = float(lam)
        o.run(seed=seed, tau=tau, step0=step0, nsteps=steps)
        step0 += steps
    return o


@pytest.mark.parametrize("lanes,chunk", [(8, 0), (16, 0), (32, 0), (16, 3)])
def test_batch_fp64_fixed_epochs_matches_oracle(acdc, orc, small_batch, lanes, chunk):
    """chunk 3: a block takes three instances per queue access"""
    ds = small_batch
    lams = lambdas_of(acdc, ds, 3, 0.1)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=40, precision=acdc.PREC_FP64, lanes_per_coord=lanes,
                   refresh_period=1, max_epochs=2, gap_tol=-1.0, use_graph=False,
                   queue_chunk=chunk)
    b.run()
    res = b.results()
    assert (res["epochs"] == 2).all()
    # the refreshes at the last lambda stream every entry in the scatter (x
    # denser than n/64 after the previous lambda) for some instances; the
    # earlier, sparser refreshes take the nonzero columns one by one
    dense = res[:, -2]["nnz_x"] * 64 > np.array([d.n for d in ds])
    assert dense.any(), res[:, -2]["nnz_x"]
    for i, d in enumerate(ds):
        tau, beta, lds = b.info(i)
        assert lds == (d.m <= 19968)
        o = oracle_path(orc, d, lams[i], 40 + i, tau, beta, 2)
        x = b.x(i)
        err = np.max(np.abs(x - o.x)) / max(1.0, np.max(np.abs(o.x)))
        assert err < 1e-10, (i, err)
        e = o.eval()
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-10)
        assert res[i, -1]["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-6, abs=1e-12)
    b.close()


def test_batch_fp32_converges_to_fp64_optimum(acdc, orc, small_batch):
    ds = small_batch
    lams = lambdas_of(acdc, ds, 4, 0.05)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=7, precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                   max_epochs=400, gap_tol=1e-7, use_graph=True)
    b.run()
    res = b.results()
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i, -1])
        e = o.eval(b.x(i))
        # the FP64 oracle agrees with the device's own FP64 evaluation
        assert e["rel_gap"] <= 1.01e-7
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-9)
    b.close()


def test_batch_graph_replays_match_eager(acdc, small_batch):
    ds = small_batch
    lams = lambdas_of(acdc, ds, 3, 0.1)
    probs = [d.problem() for d in ds]
    kw = dict(seed=3, precision=acdc.PREC_FP64, refresh_period=2, max_epochs=4, gap_tol=-1.0)
    eager = acdc.Batch(probs, lams, use_graph=False, **kw)
    eager.run()
    ref = [eager.x(i) for i in range(len(ds))]
    g = acdc.Batch(probs, lams, use_graph=True, **kw)
    for rep in range(3):
        g.run()
        for i in range(len(ds)):
            x = g.x(i)  # eager host work and a synchronisation between replays
            assert np.max(np.abs(x - ref[i])) <= 1e-10 * max(1.0, np.max(np.abs(ref[i]))), (rep, i)
    eager.close()
    g.close()


def test_batch_async_one_worker_is_serial(acdc, orc, small_batch):
    """ASYNC instances with tau = 1: the one worker's stream (async_coord
    keyed by seed + i, worker 0) replayed by the serial oracle over the
    warm-started path, at FP64 (LDS and global residuals both)."""
    ds = small_batch
    lams = lambdas_of(acdc, ds, 3, 0.1)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=70, precision=acdc.PREC_FP64, refresh_period=1,
                   max_epochs=2, gap_tol=-1.0, use_graph=False, step_mode=acdc.STEP_ASYNC, tau=1)
    b.run()
    res = b.results()
    assert (res["epochs"] == 2).all()
    # the refreshes at the last lambda stream every entry in the scatter (x
    # denser than n/64 after the previous lambda) for some instances; the
    # earlier, sparser refreshes take the nonzero columns one by one
    dense = res[:, -2]["nnz_x"] * 64 > np.array([d.n for d in ds])
    assert dense.any(), res[:, -2]["nnz_x"]
    for i, d in enumerate(ds):
        assert b.info(i)[0] == 1
        o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i][0])
        k = 0
        for lam in lams[i]:
            o.lam = float(lam)
            o.run_list(orc.async_coords(70 + i, 0, k, 2 * d.n, d.n))
            k += 2 * d.n
        x = b.x(i)
        err = np.max(np.abs(x - o.x)) / max(1.0, np.max(np.abs(o.x)))
        assert err < 1e-10, (i, err)
        assert res[i, -1]["primal"] == pytest.approx(o.eval()["primal"], rel=1e-10)
    b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_async_workers_reach_fp64_optimum(acdc, orc, small_batch, prec):
    ds = small_batch
    lams = lambdas_of(acdc, ds, 4, 0.05)
    probs = [d.problem() for d in ds]
    precision = acdc.PREC_FP32_REFRESH if prec == "fp32" else acdc.PREC_FP64
    b = acdc.Batch(probs, lams, seed=9, precision=precision, refresh_period=5, max_epochs=400,
                   gap_tol=1e-7, use_graph=True, step_mode=acdc.STEP_ASYNC)
    b.run()
    res = b.results()
    print("epochs", res["epochs"].tolist(), "ws", res["ws_size"].tolist(), flush=True)
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        e = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i, -1]).eval(b.x(i))
        assert e["rel_gap"] <= 1.01e-7
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-9)
    b.close()


def test_autotuner_picks_valid_settings(acdc, small_batch):
    probs = [d.problem() for d in small_batch]
    t = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC)
    assert t["lanes_per_coord"] in (8, 16, 32)
    assert 1 <= t["tau"] <= 1024 // t["lanes_per_coord"]
    assert t["queue_chunk"] == 1  # fewer instances than CUs
    assert t["score"] > 0
    lams = lambdas_of(acdc, small_batch, 2, 0.3)
    b = acdc.Batch(probs, lams, seed=3, precision=acdc.PREC_FP32_REFRESH, max_epochs=400,
                   gap_tol=1e-6, step_mode=acdc.STEP_ASYNC, lanes_per_coord=t["lanes_per_coord"],
                   tau=t["tau"])
    b.run()
    assert b.results()["converged"].all()
    b.close()
    d = small_batch[1]
    s = acdc.solver_tune_lanes(d.problem(), acdc.LOSS_L1_SQUARE, lam=d.lam, tau=256)
    assert s["lanes_per_coord"] in (16, 32, 64) and s["updates_per_s"] > 0


@pytest.fixture(scope="module")
def ttd_batch(acdc):
    # truss ground structures of uneven sizes (NNLS: x >= 0, no lambda)
    return [acdc.gen_ttd(k) for k in (2, 3, 4, 6)]


def nnls_scale(d):
    g = np.zeros(d.n)
    np.add.at(g, np.repeat(np.arange(d.n), np.diff(d.ptr)), d.vals * d.rhs[d.idx])
    return np.max(np.abs(g))


def test_batch_nnls_sync_matches_oracle(acdc, orc, ttd_batch):
    """NNLS instances, synchronous FP64 steps at fixed epochs: the oracle's
    projected coordinate steps (loss 10) with the same sampling, tau and
    beta; the refresh's objective and relative KKT residual."""
    ds = ttd_batch
    probs = [d.problem() for d in ds]
    lams = np.zeros((len(ds), 1))
    epochs = 3
    b = acdc.Batch(probs, lams, seed=50, precision=acdc.PREC_FP64, refresh_period=1,
                   max_epochs=epochs, gap_tol=-1.0, use_graph=False, loss=acdc.LOSS_NNLS)
    b.run()
    res = b.results()
    for i, d in enumerate(ds):
        tau, beta, _ = b.info(i)
        o = orc.CD(10, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, 0.0, beta=beta)
        o.run(seed=50 + i, tau=tau, step0=0, nsteps=epochs * (-(-d.n // tau)))
        x = b.x(i)
        assert (x >= 0).all()
        assert np.max(np.abs(x - o.x)) / max(1.0, np.max(np.abs(o.x))) < 1e-10
        e = o.eval()
        assert res[i, 0]["primal"] == pytest.approx(e["primal"], rel=1e-10)
        assert res[i, 0]["rel_gap"] == pytest.approx(e["kkt"] / nnls_scale(d), rel=1e-8, abs=1e-14)
    b.close()


def test_batch_nnls_async_one_worker_is_serial(acdc, orc, ttd_batch):
    ds = ttd_batch
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, np.zeros((len(ds), 1)), seed=60, precision=acdc.PREC_FP64,
                   refresh_period=1, max_epochs=2, gap_tol=-1.0, use_graph=False,
                   step_mode=acdc.STEP_ASYNC, tau=1, loss=acdc.LOSS_NNLS)
    b.run()
    res = b.results()
    for i, d in enumerate(ds):
        o = orc.CD(10, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, 0.0)
        o.run_list(orc.async_coords(60 + i, 0, 0, 2 * d.n, d.n))
        x = b.x(i)
        assert np.max(np.abs(x - o.x)) / max(1.0, np.max(np.abs(o.x))) < 1e-10
        assert res[i, 0]["primal"] == pytest.approx(o.eval()["primal"], rel=1e-10)
    b.close()


def test_batch_nnls_async_workers_converge(acdc, orc, ttd_batch):
    """Many FP32 workers with compare-and-swap steps: x stays >= 0 and every
    instance takes its relative KKT residual from 1 (at x = 0) below 1e-4
    (the truss systems are ill-conditioned: four orders within the epoch
    cap), which the FP64 oracle's evaluation of the device's x confirms."""
    ds = ttd_batch
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, np.zeros((len(ds), 1)), seed=61, precision=acdc.PREC_FP32_REFRESH,
                   refresh_period=5, max_epochs=3000, gap_tol=1e-4, use_graph=True,
                   step_mode=acdc.STEP_ASYNC, loss=acdc.LOSS_NNLS)
    b.run()
    res = b.results()
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        x = b.x(i)
        assert (x >= 0).all()
        o = orc.CD(10, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, 0.0)
        e = o.eval(x)
        assert e["kkt"] / nnls_scale(d) == pytest.approx(res[i, 0]["rel_gap"], rel=1e-6, abs=1e-12)
        assert e["kkt"] / nnls_scale(d) <= 1e-4
        assert res[i, 0]["primal"] == pytest.approx(e["primal"], rel=1e-9)
    b.close()


@pytest.fixture(scope="module")
def svm_batch(acdc):
    # hinge-SVM dual instances (coordinates = samples, residual = w over the
    # features); the last one's w does not fit the LDS residual
    shapes = [(400, 3000, 30.0), (1500, 20000, 60.0), (900, 50000, 80.0)]
    return [acdc.gen_svm(ns, nf, row_median=med, row_sigma=0.9, wstar_nnz=nf // 10,
                         flip_frac=0.05, seed=200 + i)
            for i, (ns, nf, med) in enumerate(shapes)]


def svm_lambdas(ds):
    # a warm-started path 10/n -> 1/n per instance
    return np.stack([np.array([10.0, 3.0, 1.0]) / d.n for d in ds])


def svm_oracle(orc, d, lam, beta=1.0):
    return orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, lam, beta=beta)


@pytest.mark.parametrize("lanes", [16, 32])
def test_batch_svm_sync_matches_oracle(acdc, orc, svm_batch, lanes):
    """SVM instances, synchronous FP64 steps at fixed epochs along the path:
    the oracle's box-clipped SDCA steps with the same sampling, tau and
    beta (w recomputed from alpha at each new lambda), and the refresh's
    primal, dual and relative gap."""
    ds = svm_batch
    lams = svm_lambdas(ds)
    probs = [d.problem() for d in ds]
    epochs = 2
    b = acdc.Batch(probs, lams, seed=80, precision=acdc.PREC_FP64, lanes_per_coord=lanes,
                   refresh_period=1, max_epochs=epochs, gap_tol=-1.0, use_graph=False,
                   loss=acdc.LOSS_SVM_DUAL_HINGE)
    b.run()
    res = b.results()
    assert (res["epochs"] == epochs).all()
    assert not b.info(2)[2]  # the widest instance's w lives in global memory
    for i, d in enumerate(ds):
        tau, beta, _ = b.info(i)
        o = svm_oracle(orc, d, lams[i, 0], beta)
        steps = epochs * (-(-d.n // tau))
        step0 = 0
        for lam in lams[i]:
            o.lam = float(lam)
            o.recompute_w()
            o.run(seed=80 + i, tau=tau, step0=step0, nsteps=steps)
            step0 += steps
        a = b.x(i)
        assert a.min() >= 0.0 and a.max() <= 1.0
        assert np.max(np.abs(a - o.alpha)) < 1e-10, i
        e = o.eval()
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-10)
        assert res[i, -1]["dual"] == pytest.approx(e["dual"], rel=1e-10)
        assert res[i, -1]["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-7, abs=1e-13)
        assert abs(res[i, -1]["nnz_x"] - np.count_nonzero(o.alpha)) <= 2
    b.close()


def test_batch_svm_async_one_worker_is_serial(acdc, orc, svm_batch):
    """ASYNC SVM instances with tau = 1: the worker's stream replayed by the
    serial oracle (orc_svm_list) along the warm-started path, at FP64."""
    ds = svm_batch
    lams = svm_lambdas(ds)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=90, precision=acdc.PREC_FP64, refresh_period=1,
                   max_epochs=2, gap_tol=-1.0, use_graph=False, step_mode=acdc.STEP_ASYNC,
                   tau=1, loss=acdc.LOSS_SVM_DUAL_HINGE)
    b.run()
    res = b.results()
    for i, d in enumerate(ds):
        assert b.info(i)[0] == 1
        o = svm_oracle(orc, d, lams[i, 0])
        k = 0
        for lam in lams[i]:
            o.lam = float(lam)
            o.recompute_w()
            o.run_list(orc.async_coords(90 + i, 0, k, 2 * d.n, d.n))
            k += 2 * d.n
        a = b.x(i)
        assert np.max(np.abs(a - o.alpha)) < 1e-10, i
        assert res[i, -1]["primal"] == pytest.approx(o.eval()["primal"], rel=1e-10)
    b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_svm_async_workers_converge(acdc, orc, svm_batch, prec):
    """Many workers with compare-and-swap steps on the box [0, 1]: every
    (instance, lambda) reaches relative gap 1e-5, which the FP64 oracle's
    evaluation of the device's alpha confirms."""
    ds = svm_batch
    lams = svm_lambdas(ds)
    probs = [d.problem() for d in ds]
    precision = acdc.PREC_FP32_REFRESH if prec == "fp32" else acdc.PREC_FP64
    b = acdc.Batch(probs, lams, seed=91, precision=precision, refresh_period=5,
                   max_epochs=2000, gap_tol=1e-5, use_graph=True, step_mode=acdc.STEP_ASYNC,
                   loss=acdc.LOSS_SVM_DUAL_HINGE)
    b.run()
    res = b.results()
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        a = b.x(i)
        assert a.min() >= 0.0 and a.max() <= 1.0
        e = svm_oracle(orc, d, lams[i, -1]).eval(a)
        assert e["rel_gap"] <= 1.0001e-5
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-9)
        assert res[i, -1]["dual"] == pytest.approx(e["dual"], rel=1e-9)
    b.close()


def test_batch_svm_autotuner(acdc, svm_batch):
    probs = [d.problem() for d in svm_batch]
    t = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC, loss=acdc.LOSS_SVM_DUAL_HINGE)
    assert t["lanes_per_coord"] in (8, 16, 32)
    assert 1 <= t["tau"] <= 1024 // t["lanes_per_coord"]
    assert t["score"] > 0


def test_batch_spectral_beta(acdc, orc, small_batch):
    """beta_mode = spectral: each instance's beta is the single solver's
    spectral beta at the instance's tau and seed, and the synchronous FP64
    steps with it follow the oracle's."""
    ds = small_batch[:3]
    lams = lambdas_of(acdc, ds, 2, 0.1)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=44, precision=acdc.PREC_FP64, lanes_per_coord=16,
                   refresh_period=1, max_epochs=2, gap_tol=-1.0, use_graph=False,
                   beta_mode=acdc.BETA_SPECTRAL)
    b.run()
    for i, d in enumerate(ds):
        tau, beta, _ = b.info(i)
        s = acdc.Solver(probs[i], acdc.LOSS_L1_SQUARE, lam=0.0, tau=tau, seed=44 + i,
                        beta_mode=acdc.BETA_SPECTRAL)
        assert beta == pytest.approx(s.beta, rel=1e-12)  # atomic sums in the power method
        s.close()
        omega_beta = orc.eso_beta(orc.omega(d.m, d.idx), tau, d.n)
        assert 1.0 <= beta <= omega_beta + 1e-12
        o = oracle_path(orc, d, lams[i], 44 + i, tau, beta, 2)
        x = b.x(i)
        assert np.max(np.abs(x - o.x)) <= 1e-10 * max(1.0, np.max(np.abs(o.x))), i
    b.close()


def test_batch_spectral_beta_svm(acdc, svm_batch):
    """Spectral beta for SVM instances: the single solver's, at the
    instance's tau and seed (its power method runs on the sample rows)."""
    ds = svm_batch[:2]
    lams = svm_lambdas(ds)
    probs = [d.problem() for d in ds]
    b = acdc.Batch(probs, lams, seed=46, precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                   max_epochs=100, gap_tol=1e-5, use_graph=False, step_mode=acdc.STEP_ASYNC,
                   loss=acdc.LOSS_SVM_DUAL_HINGE, beta_mode=acdc.BETA_SPECTRAL)
    b.run()
    res = b.results()
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        tau, beta, _ = b.info(i)
        s = acdc.Solver(probs[i], acdc.LOSS_SVM_DUAL_HINGE, lam=float(lams[i, 0]), tau=tau,
                        seed=46 + i, beta_mode=acdc.BETA_SPECTRAL)
        assert beta == pytest.approx(s.beta, rel=1e-12)
        s.close()
        assert 1.0 <= beta
    b.close()


@pytest.mark.parametrize("prec,beta,newton", [("fp32", "omega", None), ("fp64", "omega", None),
                                              ("fp32", "spectral", None), ("fp32", "spectral", 0)])
def test_batch_screening_reaches_oracle_gap(acdc, orc, small_batch, prec, beta, newton):
    """Working sets (screening on: strong rule, gap-safe screening, a cap by
    |g_j| rank, a cyclic sweep, Anderson extrapolation, support Newton steps
    unless newton = 0; under the spectral rule W's ESO constant from the
    power method) for asynchronous Lasso instances along a path down to
    1e-3 lambda_max: every pair converges, the oracle's evaluation of each
    instance's last x has relative gap <= 1e-6 and equals the device's
    report, and the workers drew from working sets smaller than n along the
    path."""
    ds = small_batch
    lams = lambdas_of(acdc, ds, 8, 1e-3)
    probs = [d.problem() for d in ds]
    precision = acdc.PREC_FP32_REFRESH if prec == "fp32" else acdc.PREC_FP64
    beta_mode = acdc.BETA_SPECTRAL if beta == "spectral" else acdc.BETA_OMEGA
    # the serial method needs ~2,600-6,700 epochs of n at the last lambda of
    # the three wide instances (measured with the oracle on like data)
    b = acdc.Batch(probs, lams, seed=11, precision=precision, refresh_period=5, max_epochs=40000,
                   gap_tol=1e-6, use_graph=True, step_mode=acdc.STEP_ASYNC, screening=True,
                   beta_mode=beta_mode, ws_newton_iters=newton)
    b.run()
    res = b.results()
    print("epochs", res["epochs"].tolist(), "ws", res["ws_size"].tolist(), flush=True)
    assert res["converged"].all(), res["rel_gap"]
    for i, d in enumerate(ds):
        e = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i, -1]).eval(b.x(i))
        assert e["rel_gap"] <= 1e-6, (i, e["rel_gap"])
        assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-9)
        assert res[i, -1]["rel_gap"] == pytest.approx(e["rel_gap"], rel=1e-6, abs=1e-12)
    ns = np.array([d.n for d in ds])[:, None]
    assert (res["ws_size"][:, 1:] < ns).mean() > 0.5, res["ws_size"]
    b.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_newton_steps_on_near_interpolating_pairs(acdc, orc, prec):
    """Wide instances (n = 8-20 m) down to 1e-3 lambda_max, where the support
    of the small-lambda solutions nears m: with the support Newton steps
    every pair reaches gap 1e-6 (the oracle's evaluation of each last x
    agrees) within a cap of 500 epochs per lambda, and the path takes at
    most a third of the epochs the working sets without them take (capped
    at 2,000 per lambda)."""
    ds = [acdc.gen_lasso(m, n, nnz_law=acdc.NNZ_ZIPF_TRUNC, nnz_lo=3, nnz_hi=100, zipf_s=1.1,
                         xstar_nnz=max(1, n // 100), xstar_normal=False, noise=0.01,
                         lambda_frac=0.1, seed=700 + i)
          for i, (m, n) in enumerate([(400, 8000), (900, 9000), (600, 12000), (1000, 20000)])]
    lams = lambdas_of(acdc, ds, 12, 1e-3)
    probs = [d.problem() for d in ds]
    precision = acdc.PREC_FP32_REFRESH if prec == "fp32" else acdc.PREC_FP64
    out = {}
    for it, cap in ((200, 500), (0, 2000)):
        b = acdc.Batch(probs, lams, seed=13, precision=precision, refresh_period=5,
                       max_epochs=cap, gap_tol=1e-6, use_graph=True, step_mode=acdc.STEP_ASYNC,
                       screening=True, beta_mode=acdc.BETA_SPECTRAL, ws_newton_iters=it)
        b.run()
        res = b.results()
        print("newton", it, "epochs", res["epochs"].tolist(), "nnz", res["nnz_x"][:, -1].tolist(),
              "steps kept", res["newton_steps"].tolist(), flush=True)
        if it:
            assert res["converged"].all(), res["rel_gap"]
            for i, d in enumerate(ds):
                e = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[i, -1]).eval(b.x(i))
                assert e["rel_gap"] <= 1e-6, (i, e["rel_gap"])
                assert res[i, -1]["primal"] == pytest.approx(e["primal"], rel=1e-9)
        out[it] = int(res["epochs"].sum())
        b.close()
    assert out[200] * 3 <= out[0], out


@pytest.mark.parametrize("mode", ["sync", "async", "async_screened"])
def test_batch_nonfinite_instance_is_abandoned(acdc, small_batch, mode):
    """A NaN in one instance's data marks that instance (every lambda of its
    path) ACDC_BATCH_NONFINITE; the other instances run to their gap target
    as usual."""
    ds = small_batch
    bad = 2
    probs = []
    for i, d in enumerate(ds):
        vals = d.vals.copy()
        if i == bad:
            vals[5] = np.nan
        probs.append(acdc.Problem.from_csc(d.m, d.n, d.ptr, d.idx, vals, d.rhs))
    lams = lambdas_of(acdc, ds, 3, 0.1)
    b = acdc.Batch(probs, lams, seed=3, precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
                   max_epochs=400, gap_tol=1e-6, use_graph=True,
                   step_mode=acdc.STEP_SYNC if mode == "sync" else acdc.STEP_ASYNC,
                   screening=mode == "async_screened")
    b.run()
    res = b.results()
    assert (res["status"][bad] == acdc.BATCH_NONFINITE).all()
    assert not res["converged"][bad].any()
    good = [i for i in range(len(ds)) if i != bad]
    assert (res["status"][good] == 0).all()
    assert res["converged"][good].all(), res["rel_gap"][good]
    b.close()
