"""Parity of the gfx950 HIP path (through the C ABI) with the CPU FP64 oracle.

(a) the counter-based coordinate stream is bit-exact against the host replay;
(b) at tau = 1 the FP64 trajectory matches the serial oracle within 1e-12
    relative (differences come only from the order of the FP64 dot products);
(c) for tau > 1 the synchronous FP64 step matches the oracle's synchronous
    step within 1e-10 (plus FP64 atomic-order differences in the residual),
    and the FP32 + FP64-refresh path reaches the FP64 oracle's objective
    within the stated FP32 tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_SERIAL = 1e-12    # tau = 1, FP64
TOL_SYNC = 1e-10      # tau > 1, FP64 (atomic summation order)
TOL_FP32_OBJ = 1e-6   # FP32 + FP64 refresh: objective vs FP64 oracle


def rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


@pytest.fixture(scope="module")
def c1(acdc):
    return acdc.config_lasso_c1()


def oracle_cd(orc, d, loss, lam, beta=1.0):
    return orc.CD(loss, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lam, beta=beta)


def test_device_present(acdc):
    assert acdc.device_count() >= 1, "no gfx950 device visible: the HIP path cannot run"


@pytest.mark.parametrize("n,tau,steps", [(1, 1, 5), (7, 3, 50), (1000, 1, 2000),
                                         (10007, 64, 100), (1_000_000, 4096, 8)])
def test_sampler_replay_bit_exact(acdc, orc, n, tau, steps):
    ptr = np.arange(n + 1, dtype=np.int32)
    idx = np.zeros(n, np.int32)
    p = acdc.Problem.from_csc(1, n, ptr, idx, np.ones(n), np.ones(1))
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.0, tau=tau, seed=12345)
    got = s.sample_coordinates(17, steps)
    want = orc.sample_steps(12345, 17, steps, n, tau)
    assert (got == want).all()
    s.close()


def test_setup_matches_oracle(acdc, orc, c1):
    p = c1.problem()
    tau = 256
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=1)
    om = orc.omega(c1.m, c1.idx)
    assert p.omega == om
    beta = orc.eso_beta(om, tau, c1.n)
    assert s.beta == pytest.approx(beta, rel=1e-15)
    L = orc.csc_lipschitz(c1.n, c1.ptr, c1.vals, beta)
    assert rel_inf(s.lipschitz, L) < 1e-14
    assert np.allclose(s.residuals, -c1.rhs, rtol=0, atol=0)
    rep = s.report()
    e = oracle_cd(orc, c1, 0, c1.lam).eval(np.zeros(c1.n))
    for k in ("primal", "dual", "gap", "grad_inf"):
        assert rep[k] == pytest.approx(e[k], rel=1e-12), k
    s.close()


def test_tau1_trajectory_fp64(acdc, orc, c1):
    """tau = 1 FP64 matches the serial oracle within 1e-12, epoch by epoch."""
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=1, seed=1)
    o = oracle_cd(orc, c1, 0, c1.lam)
    for epoch in range(5):
        s.run_updates(c1.n)
        o.run(seed=1, tau=1, step0=epoch * c1.n, nsteps=c1.n)
        x = s.x
        assert rel_inf(x, o.x) < TOL_SERIAL, epoch
        assert rel_inf(s.residuals, o.r) < TOL_SERIAL, epoch
        rep = s.report()
        e = o.eval()
        assert rep["primal"] == pytest.approx(e["primal"], rel=TOL_SERIAL)
        assert rep["gap"] == pytest.approx(e["gap"], rel=1e-9, abs=1e-12 * e["primal"])
        assert rep["updates"] == (epoch + 1) * c1.n
    s.close()


@pytest.mark.parametrize("lanes,launch", [(8, 1), (16, 1), (32, 1), (64, 1), (256, 1),
                                          (8, 2), (16, 2), (32, 2), (64, 2), (256, 2), (16, 3), (256, 3)])
def test_sync_tau_fp64_matches_oracle(acdc, orc, c1, lanes, launch):
    """launch 1: one gather + one scatter launch per step; 2: persistent grid;
    3: per step with the streamed scatter over the transposed copy; 256
    lanes: a whole block per coordinate."""
    tau = 512
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=3, lanes_per_coord=lanes,
                    launch_mode=launch)
    o = oracle_cd(orc, c1, 0, c1.lam, beta=s.beta)
    steps = 2 * c1.n // tau
    for k in range(3):
        s.run_updates(steps * tau)
        o.run(seed=3, tau=tau, step0=k * steps, nsteps=steps)
        assert rel_inf(s.x, o.x) < TOL_SYNC
        assert rel_inf(s.residuals, o.r) < TOL_SYNC
    rep, e = s.report(), o.eval()
    assert rep["primal"] == pytest.approx(e["primal"], rel=TOL_SYNC)
    s.close()


def test_fp32_refresh_reaches_fp64_objective(acdc, orc, c1):
    tau = 512
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=5,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5)
    o = oracle_cd(orc, c1, 0, c1.lam, beta=s.beta)
    steps = 10 * c1.n // tau
    for k in range(4):
        s.run_updates(steps * tau)
        o.run(seed=5, tau=tau, step0=k * steps, nsteps=steps)
    x = s.x
    e_gpu = o.eval(x)            # FP64 oracle objective at the FP32 path's x
    e_orc = o.eval()
    rep = s.report()
    assert e_gpu["primal"] == pytest.approx(e_orc["primal"], rel=TOL_FP32_OBJ)
    # the device's own FP64 evaluation (with FP32-rounded A) agrees with the oracle's
    assert rep["primal"] == pytest.approx(e_gpu["primal"], rel=1e-6)
    # same progress as the FP64 oracle on the same coordinate stream
    assert e_gpu["rel_gap"] == pytest.approx(e_orc["rel_gap"], rel=1e-3)
    # the refreshed residual is A x - b of the FP64 oracle up to FP32 rounding of A
    s.recompute_residuals()
    r = orc.csc_residual(c1.m, c1.n, c1.ptr, c1.idx, c1.vals, x, c1.rhs)
    assert rel_inf(s.residuals, r) < 1e-6
    s.close()


@pytest.mark.parametrize("launch", [1, 2])
def test_fp32_anchored_accuracy_below_fp32_floor(acdc, orc, c1, launch):
    """The FP32 path keeps r = r64 + dr with an FP64 anchor: it converges to
    FP64-level gaps (a plain FP32 residual stalls near 1e-7)."""
    tau = 64
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=5,
                    precision=acdc.PREC_FP32_REFRESH, refresh_period=5, launch_mode=launch)
    o = oracle_cd(orc, c1, 0, c1.lam, beta=s.beta)
    steps = 10 * c1.n // tau
    for k in range(12):
        s.run_updates(steps * tau)
        o.run(seed=5, tau=tau, step0=k * steps, nsteps=steps)
    e_gpu, e_orc = o.eval(s.x), o.eval()
    assert e_orc["rel_gap"] < 1e-13
    assert e_gpu["rel_gap"] < 1e-10
    assert e_gpu["primal"] == pytest.approx(e_orc["primal"], rel=1e-10)
    rep = s.report()
    assert rep["rel_gap"] < 1e-9
    s.close()


def test_persistent_launch_covers_long_runs(acdc, orc, c1):
    """More steps than one persistent launch holds, FP64, against the oracle."""
    tau = 8
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=11,
                    launch_mode=acdc.LAUNCH_PERSISTENT)
    o = oracle_cd(orc, c1, 0, c1.lam, beta=s.beta)
    steps = 10_000  # > 4096 steps per launch
    s.run_updates(steps * tau)
    o.run(seed=11, tau=tau, step0=0, nsteps=steps)
    assert rel_inf(s.x, o.x) < TOL_SYNC
    s.close()


def test_nnls_ttd_matches_oracle(acdc, orc):
    t = acdc.gen_ttd(4)
    p = t.problem()
    for tau in (1, 64):
        s = acdc.Solver(p, acdc.LOSS_NNLS, lam=0.0, tau=tau, seed=7)
        o = oracle_cd(orc, t, 10, 0.0, beta=s.beta)
        steps = 3 * t.n // tau
        s.run_updates(steps * tau)
        o.run(seed=7, tau=tau, step0=0, nsteps=steps)
        tol = TOL_SERIAL if tau == 1 else TOL_SYNC
        assert rel_inf(s.x, o.x) < tol
        assert (s.x >= 0).all()
        rep, e = s.report(), o.eval()
        assert rep["primal"] == pytest.approx(e["primal"], rel=tol)
        assert rep["kkt"] == pytest.approx(e["kkt"], rel=1e-8, abs=1e-14)
        s.close()


def test_spectral_beta_matches_oracle(acdc, orc, c1):
    """ACDC_BETA_SPECTRAL: the device's 20 power steps give the oracle's
    eigenvalue (same seeded start), for CSC columns (Lasso) and CSR sample
    rows (SVM); n equal segments give lmax = n, so beta = tau."""
    p = c1.problem()
    for tau in (1, 512):
        s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lam, tau=tau, seed=21,
                        beta_mode=acdc.BETA_SPECTRAL)
        lmax = orc.spectral_lmax(c1.n, c1.m, c1.ptr, c1.idx, c1.vals, seed=21)
        want = 1.0 + (tau - 1) * (lmax - 1.0) / (c1.n - 1.0)
        assert s.beta == pytest.approx(want, rel=1e-10)
        assert 1.0 <= s.beta <= orc.eso_beta(orc.omega(c1.m, c1.idx), tau, c1.n) + 1e-12
        s.close()
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=3)
    s = acdc.Solver(d.problem(), acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=256, seed=22,
                    beta_mode=acdc.BETA_SPECTRAL)
    lmax = orc.spectral_lmax(d.n, d.m, d.ptr, d.idx, d.vals, seed=22)
    assert s.beta == pytest.approx(1.0 + 255 * (lmax - 1.0) / (d.n - 1.0), rel=1e-10)
    s.close()
    # 64 columns, each a multiple of one vector: the normalised Gram matrix is all ones
    m, n = 50, 64
    rng = np.random.default_rng(4)
    base = rng.normal(size=m)
    A = base[:, None] * rng.uniform(0.5, 2.0, size=n)[None, :]
    cp = np.arange(n + 1) * m
    ri = np.tile(np.arange(m), n)
    q = acdc.Problem.from_csc(m, n, cp, ri, A.T.ravel(), rng.normal(size=m))
    s = acdc.Solver(q, acdc.LOSS_L1_SQUARE, lam=0.1, tau=16, seed=1, beta_mode=acdc.BETA_SPECTRAL)
    assert s.beta == pytest.approx(16.0, rel=1e-12)
    s.close()


def test_svm_matches_oracle(acdc, orc):
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=3)
    p = d.problem()
    # automatic lanes: a block per sample row (rows average ~400 entries); 64:
    # one wave per row
    for tau, lanes in ((1, 0), (256, 0), (256, 64)):
        s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=tau, seed=9,
                        lanes_per_coord=lanes)
        o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=s.beta)
        steps = 2 * d.n // tau
        s.run_updates(steps * tau)
        o.run(seed=9, tau=tau, step0=0, nsteps=steps)
        tol = TOL_SERIAL if tau == 1 else TOL_SYNC
        assert rel_inf(s.x, o.alpha) < tol
        assert rel_inf(s.residuals, o.w) < tol
        rep, e = s.report(), o.eval()
        assert rep["gap"] == pytest.approx(e["gap"], rel=1e-9)
        assert rep["primal"] == pytest.approx(e["primal"], rel=tol)
        assert rep["accuracy"] == pytest.approx(e["accuracy"], abs=1.0 / d.n)
        assert (s.x >= 0).all() and (s.x <= 1).all()
        s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_streamed_scatter_svm(acdc, orc, prec):
    """SVM minibatches with the streamed scatter (automatic at tau >= n/16,
    forced below) against the oracle's SDCA steps (FP64), and against the
    atomic per-step scatter on the same stream (FP32: the two differ only in
    the order of the FP32 additions into w)."""
    # rows of ~450 entries: automatic lanes give a block per row, which runs
    # per step (no persistent grid), so tau = 512 >= n/16 streams by itself
    d = acdc.gen_svm(2000, 50_000, row_median=300.0, row_sigma=0.901, wstar_nnz=1000,
                     flip_frac=0.1, seed=5)
    p = d.problem()
    precision = acdc.PREC_FP64 if prec == "fp64" else acdc.PREC_FP32_REFRESH
    for tau, launch in ((512, 0), (64, 3)):
        s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=tau, seed=12,
                        precision=precision, refresh_period=0, launch_mode=launch)
        steps = 3 * d.n // tau
        s.run_updates(steps * tau)
        if prec == "fp64":
            o = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=s.beta)
            o.run(seed=12, tau=tau, step0=0, nsteps=steps)
            assert rel_inf(s.x, o.alpha) < TOL_SYNC
            assert rel_inf(s.residuals, o.w) < TOL_SYNC
            assert s.report()["primal"] == pytest.approx(o.eval()["primal"], rel=TOL_SYNC)
        else:
            a = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=tau, seed=12,
                            precision=precision, refresh_period=0,
                            launch_mode=acdc.LAUNCH_PER_STEP)
            a.run_updates(steps * tau)
            assert rel_inf(s.x, a.x) < 1e-4
            assert s.report()["primal"] == pytest.approx(a.report()["primal"], rel=1e-5)
            a.close()
        assert (s.x >= 0).all() and (s.x <= 1).all()
        s.close()


def test_edge_cases(acdc, orc):
    rng = np.random.default_rng(0)
    m, n = 30, 12
    # empty columns, a duplicated row inside one column, tau = n
    cnt = np.array([0, 3, 5, 0, 1, 4, 4, 2, 0, 6, 3, 2])
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(m, c, replace=False)) for c in cnt]).astype(np.int32)
    idx[ptr[2] + 1] = idx[ptr[2]]
    vals = rng.standard_normal(idx.size)
    b = rng.standard_normal(m)
    p = acdc.Problem.from_csc(m, n, ptr, idx, vals, b)
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=0.05, tau=n, seed=2)
    o = orc.CD(0, m, n, ptr, idx, vals, b, 0.05, beta=s.beta)
    s.run_updates(20 * n)
    o.run(seed=2, tau=n, step0=0, nsteps=20)
    assert rel_inf(s.x, o.x) < TOL_SYNC
    assert (s.x[cnt == 0] == 0).all()
    s.close()
    # lambda above lambda_max keeps x = 0 and the gap at 0
    A_T_b = np.zeros(n)
    np.add.at(A_T_b, np.repeat(np.arange(n), cnt), vals * b[idx])
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=2 * np.abs(A_T_b).max(), tau=4, seed=2)
    s.run_updates(10 * n)
    assert (s.x == 0).all() and abs(s.report()["gap"]) < 1e-14
    s.close()
    # a single coordinate
    p1 = acdc.Problem.from_csc(3, 1, [0, 2], [0, 2], [3.0, 4.0], [1.0, 0.0, 2.0])
    s = acdc.Solver(p1, acdc.LOSS_L1_SQUARE, lam=1.0, tau=1, seed=0)
    s.run_updates(3)
    assert s.x[0] == pytest.approx((3 * 1 + 4 * 2 - 1.0) / 25.0, rel=1e-15)
    s.close()


@pytest.mark.parametrize("loss", [0, 10])
def test_streamed_scatter_edge_cases(acdc, orc, loss):
    """The streamed scatter's block ranges: a row in every column (more
    entries than a block's target: a block of its own), a run of empty rows,
    empty columns; FP64 against the oracle's synchronous steps (Lasso, NNLS)."""
    rng = np.random.default_rng(31)
    m, n, k = 3000, 1500, 5
    cols = []
    for j in range(n):
        if j % 97 == 5:
            cols.append(np.array([], dtype=np.int64))  # empty column
            continue
        rows = 100 + rng.choice(m - 100, k, replace=False)  # rows 1..99 stay empty
        cols.append(np.sort(np.concatenate([[0], rows])))
    ptr = np.concatenate([[0], np.cumsum([c.size for c in cols])]).astype(np.int32)
    idx = np.concatenate(cols).astype(np.int32)
    vals = rng.standard_normal(idx.size)
    b = rng.standard_normal(m)
    p = acdc.Problem.from_csc(m, n, ptr, idx, vals, b)
    lam = 0.05 if loss == 0 else 0.0
    tau = 256
    s = acdc.Solver(p, loss, lam=lam, tau=tau, seed=8, launch_mode=acdc.LAUNCH_PER_STEP_STREAMED)
    o = orc.CD(loss, m, n, ptr, idx, vals, b, lam, beta=s.beta)
    steps = 3 * (-(-n // tau))
    s.run_updates(steps * tau)
    o.run(seed=8, tau=tau, step0=0, nsteps=steps)
    assert rel_inf(s.x, o.x) < TOL_SYNC
    assert rel_inf(s.residuals, o.r) < TOL_SYNC
    assert s.report()["primal"] == pytest.approx(o.eval()["primal"], rel=TOL_SYNC)
    s.close()


def test_warm_start_and_lambda_path(acdc, orc, c1):
    p = c1.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=c1.lambda_max * 0.5, tau=256, seed=4)
    s.run_updates(5 * c1.n)
    x_prev = s.x
    s.set_lambda(c1.lam)
    s.x = x_prev  # warm start recomputes r = A x - b
    r = orc.csc_residual(c1.m, c1.n, c1.ptr, c1.idx, c1.vals, x_prev, c1.rhs)
    assert rel_inf(s.residuals, r) < 1e-13
    s.run_updates(5 * c1.n)
    e = oracle_cd(orc, c1, 0, c1.lam).eval(s.x)
    assert s.report()["primal"] == pytest.approx(e["primal"], rel=1e-12)
    s.close()


@pytest.mark.gpu
def test_queued_runs_match_one_run(acdc, c1):
    """acdc_run_updates returns once its work is queued and keeps each call's
    device times as pending event pairs: 300 one-step calls (600 marks, past
    the 512 after which a call folds them in) end on the same iterate as one
    call of 300 steps, and the report counts every call's time.  tau = 1:
    the serial path adds in one order (tau > 1 steps scatter with atomics,
    whose order, and so the last bits, vary between runs)."""
    p = c1.problem()
    kw = dict(lam=c1.lam, tau=1, seed=9, precision=acdc.PREC_FP64)
    many = acdc.Solver(p, acdc.LOSS_L1_SQUARE, **kw)
    one = acdc.Solver(p, acdc.LOSS_L1_SQUARE, **kw)
    for _ in range(300):
        many.run_updates(1)
    one.run_updates(300)
    ra, rb = many.report(), one.report()
    assert np.array_equal(many.x, one.x)
    assert ra["primal"] == rb["primal"]
    for r in (ra, rb):
        assert r["updates"] == 300
        assert 0 < r["step_ms"] <= r["kernel_ms"]
    many.close()
    one.close()
