#!/usr/bin/env python3
"""PCDM hot-path benchmark on one MI355X.

Workload: the 200,000 x 1,000,000 Lasso (column nnz truncated Zipf(1.1) on
[5,200], unit columns, 5,000-sparse planted x*, lambda = 0.05 ||A^T b||_inf)
with tau = 4096 and the ESO beta of tau.  One "step" of this benchmark is one
epoch: n = 1,000,000 coordinate updates over data resident in HBM.

value: coordinate updates/s of the FP32 + FP64-refresh path (4096
asynchronous workers, FP32 residual corrections anchored on an FP64 refresh
every 5 epochs), from x = 0, with the direct asynchronous kernel (each
64-lane worker loads its own records and entries).  Beside it: the same
workers in FP64 (the precision alone), the plain FP64 path (synchronous
4096-coordinate steps, one gather + one scatter launch per step), the staged
asynchronous kernel (a loader wave per eight workers copies records and
entries into LDS ahead of use), the synchronous FP32 persistent kernel, the wall time each path needs from x = 0 to an FP64 relative duality
gap <= 1e-6 (gap evaluated every 5 epochs, evaluations included), the
roofline of the dominant kernel (timed with HIP events on the solver's
stream), the other configurations (the 2,000 x 10,000 serial Lasso, the SVM,
the truss NNLS and the batched lambda-path workloads), and the CPU baseline
(the oracle's OpenMP asynchronous restatement of the reference solver) on the
host cores.  Every reported iterate is evaluated afterwards by the FP64
oracle in the CPU leg (the "oracle" fields), which is the only place the
oracle runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ac-dc_amd"))

METRIC = "coord updates/s and time-to-1e-6 duality gap, FP32-batched vs FP64 plain, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_value(name, kernel, key):
    """A per-launch figure of `kernel` from a committed rocprofv3 summary
    (lines `<key>[<kernel>]: <value>` in profiles/<name>), or None."""
    import re
    try:
        text = open(os.path.join(ROOT, "profiles", name)).read()
    except OSError:
        return None
    m = re.search(r"^" + re.escape(key) + r"\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", text, re.M)
    return float(m.group(1)) if m else None


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    FETCH_SIZE/WRITE_SIZE summary (scaled by a copy of known size), or None."""
    return profile_value("lasso_c2_async_traffic.md", kernel, "traffic_bytes_per_launch")


def profiled_per_launch(kernel, key):
    """A per-launch counter figure of `kernel` from the committed counter
    summary of this workload, or None."""
    return profile_value("lasso_c2_async_atomics.md", kernel, key)


def device_probes():
    """Two rates of this device, measured in a child process with plain
    PyTorch operations (PyTorch brings its own HIP runtime, which cannot
    start in a process where the library's runtime already runs): see
    probe_device.  Returns (None, None) when the child fails."""
    import subprocess
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--probe-device"],
                             capture_output=True, text=True, timeout=300)
        line = [l for l in out.stdout.splitlines() if l.startswith("{")][-1]
        j = json.loads(line)
        return j["copy_gbs"], j["atomics_per_s"]
    except (subprocess.SubprocessError, IndexError, ValueError, KeyError) as e:
        log(f"[bench] device probes failed: {e}")
        return None, None


def probe_device():
    """The achievable HBM streaming rate (five copies of a 2 GiB buffer, read
    + write bytes over their HIP-event time) and the random FP32 atomic-add
    rate (index_add_ of 2^26 ones at uniform random rows of a 200,000-row
    array, the size of the 200,000 x 1,000,000 Lasso's residual)."""
    import torch
    dev = torch.device("cuda:0")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a = torch.ones(1 << 28, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * a.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    r = torch.zeros(200_000, dtype=torch.float32, device=dev)
    idx = torch.randint(0, r.numel(), (1 << 26,), device=dev)
    v = torch.ones(1 << 26, dtype=torch.float32, device=dev)
    r.index_add_(0, idx, v)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        r.index_add_(0, idx, v)
    e1.record()
    torch.cuda.synchronize()
    atom = 5 * idx.numel() / (e0.elapsed_time(e1) * 1e-3)
    return copy_gbs, atom


class OracleChecks:
    """Iterates the GPU legs report, evaluated by the FP64 oracle in the CPU
    leg.  Each entry: the data set, the loss, lambda, the device iterate and
    the dict of the result row the oracle's figures go into."""

    def __init__(self):
        self.jobs = []

    def add(self, row, d, loss, lam, x, device=None):
        self.jobs.append((row, d, loss, lam, x, device or {}))

    def run(self, orc):
        import numpy as np
        for row, d, loss, lam, x, dev in self.jobs:
            if loss == 2:
                e = orc.SVM(d.n, d.m, d.ptr, d.idx, d.vals, d.rhs, lam).eval(x)
                out = {"primal": e["primal"], "dual": e["dual"], "rel_gap": e["rel_gap"]}
            else:
                e = orc.CD(loss, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lam).eval(x)
                out = {"primal": e["primal"]}
                if loss == 10:
                    out["kkt"] = e["kkt"]
                else:
                    out.update(dual=e["dual"], rel_gap=e["rel_gap"])
            # the device's own FP64 evaluation of the same iterate, against the oracle's
            for k, v in dev.items():
                if k in out and out[k] != 0:
                    out[f"{k}_device_rel_err"] = abs(v - out[k]) / abs(out[k])
            if loss == 10:
                out["x_min"] = float(np.min(x)) if len(x) else 0.0
            row["oracle"] = out


def time_to_gap(solver, n, target, check_every, max_epochs):
    """Wall time from x = 0 to rel gap <= target: one run_updates per epoch,
    the FP64 gap every check_every epochs (evaluations included)."""
    import numpy as np
    solver.x = np.zeros(n)
    solver.synchronize()
    t0 = time.perf_counter()
    epochs, rep = 0, solver.report()
    while epochs < max_epochs and rep["rel_gap"] > target:
        for _ in range(check_every):
            solver.run_updates(n)
        epochs += check_every
        rep = solver.report()
    return time.perf_counter() - t0, epochs, rep


def ttg_row(solver, n, args, checks, d, lam, loss=0):
    tt, ep, rep = time_to_gap(solver, n, args.gap_target, args.gap_every, args.max_epochs)
    row = {"seconds": tt, "epochs": ep, "rel_gap": rep["rel_gap"],
           "reached": rep["rel_gap"] <= args.gap_target}
    checks.add(row, d, loss, lam, solver.x, {"primal": rep["primal"], "rel_gap": rep["rel_gap"]})
    return row


def throughput(solver, n, epochs):
    """Coordinate updates/s over `epochs` epochs (wall clock around them), and
    the device time and launch count of the update kernels inside them."""
    solver.synchronize()
    r0 = solver.report()
    t0 = time.perf_counter()
    for _ in range(epochs):
        solver.run_updates(n)
    solver.synchronize()
    wall = time.perf_counter() - t0
    r1 = solver.report()
    upd = r1["updates"] - r0["updates"]
    return {"value": upd / wall, "wall": wall, "updates": upd,
            "step_ms": r1["step_ms"] - r0["step_ms"],
            "launches": r1["step_launches"] - r0["step_launches"]}


def lasso_c1(acdc, args, checks):
    """The 2,000 x 10,000 Lasso (20 entries per column, 100 planted +-1,
    lambda = 0.1 ||A^T b||_inf) at tau = 1 for 100 n updates: the serial
    method in FP64 on the GPU (one wave), logging the objective and duality
    gap per epoch.  Beside it the CPU rate of the same serial method (the
    oracle on one core, same coordinates) and, evaluated in the CPU leg, the
    largest relative deviation of the logged objectives from the oracle's."""
    import numpy as np
    d = acdc.config_lasso_c1()
    p = d.problem()
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=1, seed=1)
    s.synchronize()
    t0 = time.perf_counter()
    s.run_updates(100 * d.n)
    s.synchronize()
    wall = time.perf_counter() - t0
    rep = s.report()
    s.close()
    # the logged trajectory: objective and gap after every epoch
    s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=1, seed=1)
    obj, gap = [], []
    for _ in range(100):
        s.run_updates(d.n)
        r = s.report()
        obj.append(r["primal"])
        gap.append(r["gap"])
    x = s.x
    s.close()
    p.close()
    out = {"workload": f"lasso {d.m}x{d.n}, 20 nnz per column, tau 1, 100 epochs",
           "updates": 100 * d.n, "gpu_fp64_serial": {"value": 100 * d.n / wall,
                                                     "unit": "coord updates/s", "wall_s": wall},
           "final_rel_gap": rep["rel_gap"]}
    log(f"[bench] c1: gpu serial fp64 {100 * d.n / wall:.3e} upd/s, rel gap {rep['rel_gap']:.2e}")
    out["_trajectory"] = (d, np.array(obj), np.array(gap), x)
    return out


def lasso_c1_cpu(orc, c1):
    """The CPU half of the C1 leg: the serial oracle over the same coordinate
    stream (timed, one core) and the per-epoch deviation of the GPU's logged
    objective and gap from it."""
    import numpy as np
    d, obj, gap, x = c1.pop("_trajectory")
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    t0 = time.perf_counter()
    o.run(seed=1, tau=1, step0=0, nsteps=100 * d.n)
    cw = time.perf_counter() - t0
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam)
    dev_obj, dev_gap = 0.0, 0.0
    for ep in range(100):
        o.run(seed=1, tau=1, step0=ep * d.n, nsteps=d.n)
        e = o.eval()
        dev_obj = max(dev_obj, abs(obj[ep] - e["primal"]) / abs(e["primal"]))
        # the gap falls to rounding level (~1e-16 of the objective), so its
        # deviation is measured against the objective
        dev_gap = max(dev_gap, abs(gap[ep] - e["gap"]) / abs(e["primal"]))
    c1["cpu_serial_fp64"] = {"value": 100 * d.n / cw, "unit": "coord updates/s", "cores": 1,
                             "kind": "port"}
    c1["max_rel_deviation"] = {"x": float(np.max(np.abs(x - o.x)) / max(1.0, np.max(np.abs(o.x)))),
                               "objective": dev_obj, "gap_over_objective": dev_gap}
    log(f"[bench] c1: cpu serial {100 * d.n / cw:.3e} upd/s, max deviation objective "
        f"{dev_obj:.2e} x {c1['max_rel_deviation']['x']:.2e}")


def batched(acdc, args, checks):
    """The batched workload: 256 Lasso instances of uneven sizes (m log-uniform
    in [1k,50k], n in [5k,200k], column nnz truncated Zipf(1.1) on [3,100]),
    each along a 32-value warm-started path lambda_max -> 1e-3 lambda_max,
    every (instance, lambda) to FP64 relative gap <= 1e-6; one graph-captured
    persistent launch with asynchronous workers inside each instance, the
    spectral ESO beta per instance and working sets (screening: strong rule
    and gap-safe screening, capped by |g_j| rank, one cyclic order, Anderson
    extrapolation, support Newton-CG steps), in FP32 + FP64 refresh and, the same kernel and
    settings, in FP64: the measured whole-workload ratio, and the ratio of
    the two runs' device time summed over the pairs both brought to the gap.
    Each pair runs until it meets the gap or its cap of max_epochs epochs.
    Eight instances (nnz quantiles) of the FP32 run are evaluated by the
    oracle."""
    import numpy as np
    t = time.perf_counter()
    ds = acdc.config_batch_c5(count=args.batch_count)
    nnz = np.array([d.nnz for d in ds])
    log(f"[bench] batched: generated {len(ds)} instances, nnz {nnz.sum()} in "
        f"{time.perf_counter() - t:.1f}s")
    probs = [d.problem() for d in ds]
    lams = np.stack([acdc.lambda_path(d.lambda_max, args.batch_path, 1e-3) for d in ds])
    tuned = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC, beta_mode=acdc.BETA_SPECTRAL)
    log(f"[bench] batched: autotuned lanes {tuned['lanes_per_coord']}, tau {tuned['tau']}, "
        f"queue chunk {tuned['queue_chunk']}")
    newton = acdc.WS_NEWTON_ITERS if args.batch_newton_iters is None else args.batch_newton_iters
    out = {"workload": f"{len(ds)} lasso instances x {args.batch_path} lambdas",
           "nnz": int(nnz.sum()), "max_epochs_per_lambda": args.batch_max_epochs,
           "gap_tol": args.gap_target, "beta_mode": "spectral (working sets: omega rule)",
           "screening": "working sets: strong rule + gap-safe screening, capped by |g| rank, "
                        "cyclic order, Anderson extrapolation, support Newton-CG steps",
           "newton_cg_iters": newton, "autotuned": tuned}
    order = np.argsort(nnz)
    sample = [int(order[int(q * (len(ds) - 1))]) for q in np.linspace(0.0, 1.0, 8)]
    conv, pair_us = {}, {}
    for prec, name in ((acdc.PREC_FP32_REFRESH, "fp32"), (acdc.PREC_FP64, "fp64")):
        b = acdc.Batch(probs, lams, use_graph=True, seed=5, precision=prec, refresh_period=5,
                       max_epochs=args.batch_max_epochs, gap_tol=args.gap_target,
                       step_mode=acdc.STEP_ASYNC, lanes_per_coord=tuned["lanes_per_coord"],
                       tau=tuned["tau"], queue_chunk=tuned["queue_chunk"],
                       beta_mode=acdc.BETA_SPECTRAL, screening=True, ws_newton_iters=newton)
        t0 = time.perf_counter()
        ms = b.run()
        wall = time.perf_counter() - t0
        r = b.results()
        upd = float(r["updates"][:, -1].astype(np.float64).sum())
        row = {"device_ms": ms, "wall_s": wall, "updates": upd, "value": upd / wall,
               "unit": "coord updates/s", "pairs_converged": float(r["converged"].mean()),
               "worst_rel_gap": float(r["rel_gap"].max()),
               "median_working_set_share": float(np.median(
                   r["ws_size"] / np.array([d.n for d in ds])[:, None]))}
        if name == "fp32":
            row["oracle_sampled"] = []
            for i in sample:
                rr = {"instance": i, "m": ds[i].m, "n": ds[i].n, "nnz": ds[i].nnz,
                      "device": {"primal": float(r["primal"][i, -1]),
                                 "rel_gap": float(r["rel_gap"][i, -1])}}
                checks.add(rr, ds[i], 0, float(lams[i, -1]), b.x(i), rr["device"])
                row["oracle_sampled"].append(rr)
        out[name] = row
        conv[name], pair_us[name] = r["converged"].astype(bool), r["device_us"].astype(np.float64)
        log(f"[bench] batched {name}: {ms:.0f} ms device, {upd:.3e} updates, "
            f"{row['pairs_converged']:.4f} of pairs converged, worst gap {row['worst_rel_gap']:.2e}")
        b.close()
    out["value"] = out["fp32"]["value"]
    out["unit"] = "coord updates/s"
    out["wall_ratio_fp64_over_fp32"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    both = conv["fp32"] & conv["fp64"]
    out["pairs_converged_both"] = float(both.mean())
    out["time_to_gap_ratio_fp64_over_fp32_on_pairs_converged_in_both"] = float(
        pair_us["fp64"][both].sum() / max(pair_us["fp32"][both].sum(), 1e-9))
    if out["fp32"]["pairs_converged"] == 1.0 and out["fp64"]["pairs_converged"] == 1.0:
        out["time_to_gap_speedup_fp32_vs_fp64"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    for p in probs:
        p.close()
    return out


def nnls_scale(d):
    """||A^T b||_inf of a column-major dataset: the NNLS KKT residual's scale."""
    import numpy as np
    g = np.zeros(d.n)
    np.add.at(g, np.repeat(np.arange(d.n), np.diff(d.ptr)), d.vals * d.rhs[d.idx])
    return float(np.max(np.abs(g))) or 1.0


def batched_plain_sample(acdc, args, ds, probs, lams, loss, kw, sample, checks, fixed_epochs=0):
    """Plain FP64 path (synchronous per-step launches, one instance at a
    time, gap checked every 5 epochs as the batch refreshes) against the
    batched kernel alone on the same instances and lambdas.  fixed_epochs > 0:
    both run that many epochs instead of stopping at the gap target."""
    rows = []
    for i in sample:
        d, p = ds[i], probs[i]
        P = lams.shape[1]
        bi = acdc.Batch([p], lams[i:i + 1], use_graph=False, **kw)
        bi.run()
        tb = bi.run() * 1e-3
        ri = bi.results()
        tau = int(bi.info(0)[0])
        xb = bi.x(0)
        bi.close()
        s = acdc.Solver(p, loss, lam=float(lams[i, 0]), tau=tau, seed=5,
                        precision=acdc.PREC_FP64, launch_mode=acdc.LAUNCH_PER_STEP,
                        beta_mode=kw.get("beta_mode", acdc.BETA_OMEGA))
        s.synchronize()
        t0 = time.perf_counter()
        for l in range(P):
            s.set_lambda(float(lams[i, l]))
            ep, rep = 0, s.report()
            cap = fixed_epochs or kw["max_epochs"]
            while (fixed_epochs or rep["rel_gap"] > kw["gap_tol"]) and ep < cap:
                for _ in range(5):
                    s.run_updates(d.n)
                ep += 5
                rep = s.report()
        tp = time.perf_counter() - t0
        up = rep["updates"]
        s.close()
        row = {"instance": i, "m": d.m, "n": d.n, "nnz": d.nnz, "lambdas": P,
               "batched_s": tb, "batched_updates": int(ri["updates"][0, -1]),
               "batched_rel_gap": float(ri["rel_gap"][0, -1]),
               "fp64_plain_s": tp, "fp64_plain_updates": int(up),
               "fp64_plain_rel_gap": float(rep["rel_gap"] if loss != acdc.LOSS_NNLS
                                           else rep["kkt"] / nnls_scale(d)),
               "speedup": tp / tb}
        checks.add(row, d, loss, float(lams[i, -1]), xb, {"primal": float(ri["primal"][0, -1])})
        rows.append(row)
        log(f"[bench] batched sample {i} ({d.m}x{d.n}): batched {tb:.3f}s vs plain fp64 "
            f"{tp:.3f}s = {tp / tb:.1f}x")
    return rows


def plain_estimate(out, updates, wall):
    """The whole batch at the plain FP64 path's measured cost per update (the
    sampled instances run one after another on the whole device)."""
    import numpy as np
    rows = out["fp64_plain_sampled"]
    per = [r["fp64_plain_s"] / max(1, r["fp64_plain_updates"]) for r in rows]
    out["fp64_plain_estimate_s"] = float(np.mean(per) * updates)
    out["speedup_vs_fp64_plain_estimate"] = out["fp64_plain_estimate_s"] / wall


def batched_svm(acdc, args, checks):
    """Batched hinge-SVM duals: 256 instances of uneven sizes (samples
    log-uniform in [1k, 8k], features in [10k, 200k], row nnz lognormal with
    median log-uniform in [20, 200]), each along an 8-value warm-started path
    lambda = 10/n -> 1/n, stopping at FP64 relative duality gap <= the gap
    target or 100 epochs per lambda; one graph-captured launch (FP32 + FP64
    refresh of w every 5 epochs, asynchronous workers with compare-and-swap
    steps on [0, 1], the spectral ESO beta of mSDCA.cpp per instance), and the
    same launch in FP64; beside them the plain FP64 path (same beta rule) on
    three sampled instances over the same paths."""
    import numpy as np
    t = time.perf_counter()
    rng = np.random.default_rng(6)
    count = 256
    ns = np.exp(rng.uniform(np.log(1000), np.log(8000), count)).astype(np.int64)
    nf = np.exp(rng.uniform(np.log(10000), np.log(200000), count)).astype(np.int64)
    med = np.exp(rng.uniform(np.log(20), np.log(200), count))
    ds = [acdc.gen_svm(int(a), int(f), row_median=float(r), row_sigma=0.9,
                       wstar_nnz=max(1, int(f) // 100), flip_frac=0.05, seed=6000 + i)
          for i, (a, f, r) in enumerate(zip(ns, nf, med))]
    nnz = np.array([d.nnz for d in ds])
    log(f"[bench] batched svm: generated {count} instances, nnz {nnz.sum()} in "
        f"{time.perf_counter() - t:.1f}s")
    probs = [d.problem() for d in ds]
    lams = np.stack([np.geomspace(10.0, 1.0, 8) / d.n for d in ds])
    tuned = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC, loss=acdc.LOSS_SVM_DUAL_HINGE,
                            beta_mode=acdc.BETA_SPECTRAL)
    kw = dict(seed=5, precision=acdc.PREC_FP32_REFRESH, refresh_period=5, max_epochs=100,
              gap_tol=args.gap_target, step_mode=acdc.STEP_ASYNC,
              lanes_per_coord=tuned["lanes_per_coord"], tau=tuned["tau"],
              queue_chunk=tuned["queue_chunk"], loss=acdc.LOSS_SVM_DUAL_HINGE,
              beta_mode=acdc.BETA_SPECTRAL)
    out = {"workload": f"{count} svm dual instances x 8 lambdas (10/n -> 1/n)",
           "nnz": int(nnz.sum()), "max_epochs_per_lambda": 100, "gap_tol": args.gap_target,
           "autotuned": tuned}
    conv, pair_us = {}, {}
    for prec, name in ((acdc.PREC_FP32_REFRESH, "fp32"), (acdc.PREC_FP64, "fp64")):
        b = acdc.Batch(probs, lams, use_graph=True, **dict(kw, precision=prec))
        t0 = time.perf_counter()
        ms = b.run()
        wall = time.perf_counter() - t0
        r = b.results()
        upd = float(r["updates"][:, -1].astype(np.float64).sum())
        out[name] = {"device_ms": ms, "wall_s": wall, "updates": upd, "value": upd / wall,
                     "unit": "coord updates/s", "pairs_converged": float(r["converged"].mean()),
                     "worst_rel_gap": float(r["rel_gap"].max())}
        conv[name], pair_us[name] = r["converged"].astype(bool), r["device_us"].astype(np.float64)
        log(f"[bench] batched svm {name}: {ms:.0f} ms device, {upd:.3e} updates, "
            f"{out[name]['pairs_converged']:.3f} of pairs converged")
        b.close()
    out["value"] = out["fp32"]["value"]
    out["unit"] = "coord updates/s"
    out["wall_ratio_fp64_over_fp32"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    both = conv["fp32"] & conv["fp64"]
    out["pairs_converged_both"] = float(both.mean())
    out["time_to_gap_ratio_fp64_over_fp32_on_pairs_converged_in_both"] = float(
        pair_us["fp64"][both].sum() / max(pair_us["fp32"][both].sum(), 1e-9))
    if out["fp32"]["pairs_converged"] == 1.0 and out["fp64"]["pairs_converged"] == 1.0:
        out["time_to_gap_speedup_fp32_vs_fp64"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    order = np.argsort(nnz)
    sample = [int(order[int(q * (count - 1))]) for q in (0.25, 0.5, 0.75)]
    out["fp64_plain_sampled"] = batched_plain_sample(acdc, args, ds, probs, lams,
                                                     acdc.LOSS_SVM_DUAL_HINGE, kw, sample, checks)
    plain_estimate(out, out["fp32"]["updates"], out["fp32"]["wall_s"])
    for p in probs:
        p.close()
    return out


def batched_ttd(acdc, args, checks):
    """Batched truss-topology NNLS: 32 ground structures of 4^3 to 15^3 cells
    (1,000 to 55,000 bars), 200 epochs each in one graph-captured launch
    (FP32 + FP64 refresh every 5 epochs, asynchronous workers with
    compare-and-swap steps on x >= 0); beside it the plain FP64 path on three
    sampled instances for the same 200 epochs.  The relative KKT residual
    after the epochs is reported for both."""
    import numpy as np
    t = time.perf_counter()
    count, epochs = 32, 200
    ds = [acdc.gen_ttd(4 + (i * 7) % 12) for i in range(count)]
    nnz = np.array([d.nnz for d in ds])
    log(f"[bench] batched ttd: generated {count} instances, nnz {nnz.sum()} in "
        f"{time.perf_counter() - t:.1f}s")
    probs = [d.problem() for d in ds]
    lams = np.zeros((count, 1))
    tuned = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC, loss=acdc.LOSS_NNLS)
    kw = dict(seed=5, precision=acdc.PREC_FP32_REFRESH, refresh_period=5, max_epochs=epochs,
              gap_tol=-1.0, step_mode=acdc.STEP_ASYNC, lanes_per_coord=tuned["lanes_per_coord"],
              tau=tuned["tau"], queue_chunk=tuned["queue_chunk"], loss=acdc.LOSS_NNLS)
    b = acdc.Batch(probs, lams, use_graph=True, **kw)
    t0 = time.perf_counter()
    ms = b.run()
    wall = time.perf_counter() - t0
    r = b.results()
    upd = float(r["updates"][:, -1].astype(np.float64).sum())
    out = {"workload": f"{count} ttd nnls instances, {epochs} epochs each", "nnz": int(nnz.sum()),
           "device_ms": ms, "wall_s": wall, "updates": upd, "value": upd / wall,
           "unit": "coord updates/s", "median_rel_kkt": float(np.median(r["rel_gap"][:, 0])),
           "worst_rel_kkt": float(r["rel_gap"].max()), "autotuned": tuned}
    log(f"[bench] batched ttd: {ms:.0f} ms device, {upd:.3e} updates = {upd / wall:.3e} upd/s, "
        f"median relative KKT {out['median_rel_kkt']:.2e}")
    b.close()
    order = np.argsort(nnz)
    sample = [int(order[int(q * (count - 1))]) for q in (0.25, 0.5, 0.75)]
    out["fp64_plain_sampled"] = batched_plain_sample(acdc, args, ds, probs, lams, acdc.LOSS_NNLS,
                                                     kw, sample, checks, fixed_epochs=epochs)
    plain_estimate(out, upd, wall)
    for p in probs:
        p.close()
    return out


def svm_c3(acdc, args, checks):
    """Hinge-SVM dual minibatch SDCA on 20,000 samples x 1,350,000 features
    (lognormal row nnz, median 300): synchronous minibatches b in {512, 2048,
    8192} (tau = b, the spectral ESO beta of mSDCA.cpp:188-286, the rule of
    the reference's minibatch SDCA experiments), FP32 + FP64 refresh of w
    every 5 epochs and the plain FP64 path, and b asynchronous FP32 workers
    (the multicore executor's semantics and its omega beta,
    MultiCoreSolver.cpp:55-62, 108-110); updates/s over the first 20 epochs
    and the FP64 relative duality gap after them; then the time from alpha = 0
    to relative gap 1e-6 for FP32 and FP64 at each b, whose final alphas the
    oracle evaluates."""
    d = acdc.config_svm_c3()
    p = d.problem()
    out = {"workload": f"svm dual {d.n} samples x {d.m} features", "nnz": d.nnz,
           "epochs": 20, "rows": []}
    for b in (512, 2048, 8192):
        for prec, mode, bm, name in (
                (acdc.PREC_FP32_REFRESH, acdc.STEP_SYNC, acdc.BETA_SPECTRAL, "fp32"),
                (acdc.PREC_FP64, acdc.STEP_SYNC, acdc.BETA_SPECTRAL, "fp64"),
                (acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, acdc.BETA_OMEGA, "fp32 async")):
            s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=b, seed=3,
                            precision=prec, refresh_period=5, step_mode=mode, beta_mode=bm)
            s.synchronize()
            t0 = time.perf_counter()
            s.run_updates(20 * d.n)
            s.synchronize()
            wall = time.perf_counter() - t0
            r = s.report()
            out["rows"].append({"b": b, "precision": name, "beta": s.beta,
                                "beta_rule": "spectral" if bm == acdc.BETA_SPECTRAL else "omega",
                                "updates_per_s": 20 * d.n / wall, "rel_gap": r["rel_gap"],
                                "accuracy": r["accuracy"]})
            log(f"[bench] svm b={b} {name}: {20 * d.n / wall:.3e} upd/s, rel gap {r['rel_gap']:.2e}")
            s.close()
    out["time_to_gap"] = []
    for b in (512, 2048, 8192):
        row = {"b": b}
        for prec, name in ((acdc.PREC_FP32_REFRESH, "fp32"), (acdc.PREC_FP64, "fp64")):
            s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=b, seed=3,
                            precision=prec, refresh_period=5, beta_mode=acdc.BETA_SPECTRAL)
            tt, ep, rep = time_to_gap(s, d.n, args.gap_target, 5, 400)
            row[name] = {"seconds": tt, "epochs": ep, "rel_gap": rep["rel_gap"],
                         "reached": rep["rel_gap"] <= args.gap_target}
            checks.add(row[name], d, 2, d.lam, s.x,
                       {"primal": rep["primal"], "rel_gap": rep["rel_gap"]})
            s.close()
        if row["fp32"]["reached"] and row["fp64"]["reached"]:
            row["speedup_fp32_vs_fp64"] = row["fp64"]["seconds"] / row["fp32"]["seconds"]
        out["time_to_gap"].append(row)
        log(f"[bench] svm b={b} time to gap: fp32 {row['fp32']['seconds']:.3f}s "
            f"({row['fp32']['epochs']} epochs) vs fp64 {row['fp64']['seconds']:.3f}s")
    p.close()
    return out


def ttd_c4(acdc, args, checks):
    """Truss topology design NNLS on the 40^3-cell ground structure (844,840
    bars, 201,720 DOFs): tau in {1024, 8192}, FP32 + FP64 refresh and FP64
    asynchronous workers (the precision alone) and the plain FP64 path
    (synchronous per-step launches); updates/s over 50 epochs from x = 0 and
    the FP64 objective and KKT residual after them, which the oracle
    recomputes from the device x."""
    d = acdc.gen_ttd(40)
    p = d.problem()
    out = {"workload": f"ttd nnls 40^3 cells, {d.n} bars x {d.m} dofs", "nnz": d.nnz,
           "epochs": 50, "rows": []}
    for tau in (1024, 8192):
        for prec, mode, name in ((acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, "fp32 async"),
                                 (acdc.PREC_FP64, acdc.STEP_ASYNC, "fp64 async"),
                                 (acdc.PREC_FP64, acdc.STEP_SYNC, "fp64 plain")):
            s = acdc.Solver(p, acdc.LOSS_NNLS, lam=0.0, tau=tau, seed=4, precision=prec,
                            refresh_period=5, step_mode=mode,
                            launch_mode=acdc.LAUNCH_PER_STEP if mode == acdc.STEP_SYNC
                            else acdc.LAUNCH_AUTO)
            p0 = s.report()["primal"]
            s.synchronize()
            t0 = time.perf_counter()
            s.run_updates(50 * d.n)
            s.synchronize()
            wall = time.perf_counter() - t0
            r = s.report()
            row = {"tau": tau, "path": name, "beta": s.beta,
                   "updates_per_s": 50 * d.n / wall, "primal": r["primal"],
                   "primal_at_0": p0, "kkt": r["kkt"]}
            checks.add(row, d, 10, 0.0, s.x, {"primal": r["primal"], "kkt": r["kkt"]})
            out["rows"].append(row)
            log(f"[bench] ttd tau={tau} {name}: {50 * d.n / wall:.3e} upd/s, "
                f"objective {r['primal']:.4e} (from {p0:.4e}), kkt {r['kkt']:.2e}")
            s.close()
    p.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed epochs")
    ap.add_argument("--warmup", type=int, default=2, help="untimed epochs")
    ap.add_argument("--tau", type=int, default=4096)
    ap.add_argument("--kernel", choices=["staged", "direct"], default="direct",
                    help="asynchronous kernel of the headline path")
    ap.add_argument("--lanes", type=int, default=0,
                    help="lanes per coordinate of the direct kernel (0 = autotune)")
    ap.add_argument("--m", type=int, default=200_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--gap-target", type=float, default=1e-6)
    ap.add_argument("--gap-every", type=int, default=5)
    ap.add_argument("--max-epochs", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttg", action="store_true", help="skip the time-to-gap runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workloads")
    ap.add_argument("--no-svm-ttd", action="store_true", help="skip the SVM and TTD workloads")
    ap.add_argument("--no-c1", action="store_true", help="skip the 2,000 x 10,000 serial leg")
    ap.add_argument("--no-persistent", action="store_true",
                    help="skip the synchronous persistent (cooperative-launch) leg; rocprofv3 "
                         "ends with a segmentation fault at exit after cooperative launches")
    ap.add_argument("--no-probes", action="store_true",
                    help="skip the streaming-copy and atomic-rate probes (a child process "
                         "running PyTorch kernels)")
    ap.add_argument("--probe-device", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-path", type=int, default=32)
    ap.add_argument("--batch-max-epochs", type=int, default=500,
                    help="cap on the updates of one (instance, lambda) pair, in epochs of n")
    ap.add_argument("--batch-newton-iters", type=int, default=None,
                    help="CG iterations of the working sets' support Newton steps "
                         "(0: none; default: the library's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "run without --gpus or with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.probe_device:  # the child of device_probes: PyTorch only
        copy_gbs, atom = probe_device()
        print(json.dumps({"copy_gbs": copy_gbs, "atomics_per_s": atom}), flush=True)
        return

    import numpy as np
    import acdc_amd as acdc

    if acdc.device_count() < 1:
        raise SystemExit("bench.py: no gfx950 device visible")
    checks = OracleChecks()
    t = time.perf_counter()
    d = acdc.config_lasso_c2(m=args.m, n=args.n)
    log(f"[bench] generated {d.m} x {d.n}, nnz {d.nnz} in {time.perf_counter() - t:.1f}s")
    prob = d.problem()
    n, tau = d.n, args.tau
    mean_nnz = d.nnz / d.n
    maxlen = int(np.diff(d.ptr).max())

    def make(precision, mode, launch=acdc.LAUNCH_AUTO, refresh=5, lanes=0,
             beta_mode=acdc.BETA_OMEGA):
        return acdc.Solver(prob, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=tau, seed=2,
                           precision=precision, refresh_period=refresh, step_mode=mode,
                           launch_mode=launch, lanes_per_coord=lanes, beta_mode=beta_mode)

    # ---- the direct kernel's lanes per coordinate: the autotuner unless given
    if args.lanes == 0:
        tl = acdc.solver_tune_lanes(prob, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=tau, seed=2)
        args.lanes = tl["lanes_per_coord"]
        log(f"[bench] autotuned lanes per coordinate (direct kernel): {args.lanes}")
    staged = (acdc.LAUNCH_ASYNC_STAGED, 64)
    direct = (acdc.LAUNCH_ASYNC_DIRECT, args.lanes)
    head = staged if args.kernel == "staged" else direct

    def async_solver(precision, which, beta_mode=acdc.BETA_OMEGA):
        return make(precision, acdc.STEP_ASYNC, launch=which[0], lanes=which[1],
                    beta_mode=beta_mode)

    # ---- FP32 + FP64 refresh, asynchronous workers: the timed epochs
    s32 = async_solver(acdc.PREC_FP32_REFRESH, head)
    beta = s32.beta
    s32.run_updates(args.warmup * n)
    s32.x = np.zeros(n)
    tp = throughput(s32, n, args.steps)
    value, wall = tp["value"], tp["wall"]
    log(f"[bench] fp32 async ({args.kernel}): {tp['updates']} updates in {wall:.3f}s = "
        f"{value:.4e} upd/s")
    s32.close()

    # ---- dominant kernel over the same timed epochs (one launch per epoch),
    #      HIP events on the solver's stream around the launches
    launch_ms = tp["step_ms"] / max(1, tp["launches"])
    upd_per_launch = tp["updates"] / max(1, tp["launches"])
    bytes_per_update = 8.0 * mean_nnz + 32.0  # (row, value) per nnz; the 32-byte record
    bytes_per_launch = upd_per_launch * bytes_per_update
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    if args.kernel == "staged":
        kernel = f"k_async_staged<float,0,{'true' if maxlen > 255 else 'false'}>"
    else:  # segments longer than the R = 4 register chunks of a group take the TAIL variant
        kernel = f"k_async<float,{args.lanes},0,4,{'true' if maxlen > 4 * args.lanes else 'false'}>"
    log(f"[bench] {kernel}: {launch_ms:.3f} ms per launch ({upd_per_launch:.0f} updates), "
        f"{achieved:.1f} GB/s")
    copy_gbs, atomic_probe = (None, None) if args.no_probes else device_probes()
    if not args.no_probes:
        log(f"[bench] device: {copy_gbs:.0f} GB/s streaming copy, {atomic_probe:.3e} random FP32 "
            f"atomic adds/s")
    atomics_per_launch = profiled_per_launch(kernel, "atomics_per_launch")

    result = {
        "metric": METRIC, "value": value, "unit": "coord updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"lasso {d.m}x{d.n} zipf1.1[5,200] tau{tau}", "m": d.m, "n": d.n,
                   "nnz": d.nnz, "omega": prob.omega, "tau": tau, "beta": beta,
                   "lambda": d.lam, "mode": "asynchronous workers",
                   "kernel": kernel,
                   "precision": "fp32 dots/scatter on an fp64-anchored residual, fp64 x, "
                                "fp64 refresh every 5 epochs",
                   "step": "one epoch = n coordinate updates, from x = 0"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(kernel),
                     "kernel": kernel, "copy_gbs": copy_gbs,
                     "launch_ms": launch_ms,
                     "updates_per_launch": upd_per_launch,
                     "bytes_per_update": bytes_per_update, "bytes_per_launch": bytes_per_launch},
        # updates/s inside the update kernels alone (HIP events), beside `value`,
        # which also holds the refreshes every 5 epochs and the host loop
        "kernel_only": {"value": upd_per_launch / (launch_ms * 1e-3), "unit": "coord updates/s"},
        "atomics": {"per_s": (atomics_per_launch / (launch_ms * 1e-3)
                              if atomics_per_launch else None),
                    "probe_random_fp32_per_s": atomic_probe,
                    "frac": (atomics_per_launch / (launch_ms * 1e-3) / atomic_probe
                             if atomics_per_launch and atomic_probe else None)},
    }

    # ---- the other asynchronous kernel, and FP64 workers (the precision alone)
    other = direct if args.kernel == "staged" else staged
    so = async_solver(acdc.PREC_FP32_REFRESH, other)
    so.run_updates(args.warmup * n)
    so.x = np.zeros(n)
    tpo = throughput(so, n, max(1, args.steps // 5))
    result["fp32_async_other_kernel"] = {
        "kernel": "direct" if args.kernel == "staged" else "staged",
        "value": tpo["value"], "unit": "coord updates/s",
        "kernel_only": tpo["updates"] / (tpo["step_ms"] * 1e-3)}
    log(f"[bench] fp32 async (other kernel): {tpo['value']:.4e} upd/s")
    so.close()
    sa64 = async_solver(acdc.PREC_FP64, head)
    sa64.run_updates(args.warmup * n)
    sa64.x = np.zeros(n)
    tpa = throughput(sa64, n, max(1, args.steps // 5))
    result["fp64_async"] = {"value": tpa["value"], "unit": "coord updates/s",
                            "kernel_only": tpa["updates"] / (tpa["step_ms"] * 1e-3),
                            "mode": "the same asynchronous workers and kernel in FP64"}
    log(f"[bench] fp64 async: {tpa['value']:.4e} upd/s")

    # ---- plain FP64 path: synchronous steps, per-step launches
    s64 = make(acdc.PREC_FP64, acdc.STEP_SYNC, launch=acdc.LAUNCH_PER_STEP)
    s64.run_updates(args.warmup * n)
    s64.x = np.zeros(n)
    tp64 = throughput(s64, n, max(1, args.steps // 5))
    result["fp64_plain"] = {"value": tp64["value"], "unit": "coord updates/s",
                            "mode": "synchronous tau-steps, gather + scatter launch per step"}
    log(f"[bench] fp64 plain: {tp64['value']:.4e} upd/s")

    # ---- synchronous FP32 steps in the persistent cooperative kernel
    if not args.no_persistent:
        ss = make(acdc.PREC_FP32_REFRESH, acdc.STEP_SYNC)
        ss.run_updates(args.warmup * n)
        ss.x = np.zeros(n)
        tps = throughput(ss, n, max(1, args.steps // 5))
        result["fp32_sync_persistent"] = {"value": tps["value"], "unit": "coord updates/s",
                                          "launches_per_epoch": tps["launches"] / max(1, args.steps // 5)}
        log(f"[bench] fp32 sync: {tps['value']:.4e} upd/s")
        ss.close()

    if not args.no_ttg:
        s = async_solver(acdc.PREC_FP32_REFRESH, head)
        result["time_to_gap"] = dict(ttg_row(s, n, args, checks, d, d.lam),
                                     check_every_epochs=args.gap_every)
        s.close()
        log(f"[bench] fp32 async: rel gap {result['time_to_gap']['rel_gap']:.3e} after "
            f"{result['time_to_gap']['epochs']} epochs in {result['time_to_gap']['seconds']:.3f}s")
        result["fp64_async"]["time_to_gap"] = ttg_row(sa64, n, args, checks, d, d.lam)
        log(f"[bench] fp64 async: {result['fp64_async']['time_to_gap']['seconds']:.3f}s")
        result["fp64_plain"]["time_to_gap"] = ttg_row(s64, n, args, checks, d, d.lam)
        log(f"[bench] fp64 plain: {result['fp64_plain']['time_to_gap']['seconds']:.3f}s")
        t32 = result["time_to_gap"]
        for key, base in (("speedup_vs_fp64_async", result["fp64_async"]["time_to_gap"]),
                          ("speedup_vs_fp64_plain", result["fp64_plain"]["time_to_gap"])):
            if t32["reached"] and base["reached"]:
                t32[key] = base["seconds"] / t32["seconds"]
        # the same paths with the spectral ESO beta of mSDCA.cpp (20 power
        # steps on the normalised Gram matrix) instead of the configured
        # omega rule: a side result, all paths with the same beta
        spec = {}
        for name, mk in (
                ("fp32_async", lambda: async_solver(acdc.PREC_FP32_REFRESH, head, acdc.BETA_SPECTRAL)),
                ("fp64_async", lambda: async_solver(acdc.PREC_FP64, head, acdc.BETA_SPECTRAL)),
                ("fp64_plain", lambda: make(acdc.PREC_FP64, acdc.STEP_SYNC, launch=acdc.LAUNCH_PER_STEP,
                                            beta_mode=acdc.BETA_SPECTRAL))):
            sp = mk()
            spec[name] = ttg_row(sp, n, args, checks, d, d.lam)
            spec["beta"] = sp.beta
            log(f"[bench] spectral beta {sp.beta:.4f} {name}: {spec[name]['seconds']:.3f}s, "
                f"{spec[name]['epochs']} epochs")
            sp.close()
        for other_name in ("fp64_async", "fp64_plain"):
            if spec["fp32_async"]["reached"] and spec[other_name]["reached"]:
                spec[f"speedup_vs_{other_name}"] = (spec[other_name]["seconds"] /
                                                    spec["fp32_async"]["seconds"])
        result["time_to_gap_spectral_beta"] = spec
    s64.close()
    sa64.close()

    c1 = None
    if not args.no_c1:
        c1 = lasso_c1(acdc, args, checks)
        result["c1_serial"] = c1
    if not args.no_batch:
        result["batched"] = batched(acdc, args, checks)
        result["batched_svm"] = batched_svm(acdc, args, checks)
        result["batched_ttd"] = batched_ttd(acdc, args, checks)
    if not args.no_svm_ttd:
        result["svm"] = svm_c3(acdc, args, checks)
        result["ttd"] = ttd_c4(acdc, args, checks)

    if not args.no_cpu:
        # OpenMP threads spread over the host's cores (read when the OpenMP
        # runtime starts, with the oracle library below)
        os.environ.setdefault("OMP_PROC_BIND", "spread")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as orc
        threads = max(1, min(16, os.cpu_count() or 1))
        o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam,
                   beta=orc.eso_beta(prob.omega, threads, d.n))
        t0 = time.perf_counter()
        done, per = 0, max(1, n // 10 // threads)
        while time.perf_counter() - t0 < args.cpu_seconds and done < 30 * n:
            o.run_async(seed=2, k0=done // threads, per_worker=per, threads=threads)
            done += per * threads
        cw = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": done / cw, "unit": "coord updates/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"{done} asynchronous FP64 updates ({done / n:.1f} "
                                            f"epochs) of the same Lasso from x = 0 in {cw:.1f}s, "
                                            "CAS atomics"}
        log(f"[bench] cpu: {done / cw:.4e} upd/s on {threads} threads")
        # the same solver on one thread (the serial method, beta = 1), a
        # quarter of the time budget, and the host it ran on
        o1 = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=1.0)
        t0 = time.perf_counter()
        done1, per1 = 0, max(1, n // 20)
        while time.perf_counter() - t0 < args.cpu_seconds / 4 and done1 < 30 * n:
            o1.run_async(seed=2, k0=done1, per_worker=per1, threads=1)
            done1 += per1
        cw1 = time.perf_counter() - t0
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f
                              if ln.startswith("model name")), "")
        except OSError:
            pass
        result["cpu_baseline"].update({
            "one_thread": {"value": done1 / cw1, "unit": "coord updates/s",
                           "sample": f"{done1} serial FP64 updates in {cw1:.1f}s"},
            "host": {"cpu_model": model, "nproc": os.cpu_count(),
                     "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
                     "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}})
        log(f"[bench] cpu: {done1 / cw1:.4e} upd/s on 1 thread ({model})")
        if c1 is not None:
            lasso_c1_cpu(orc, c1)
        # the oracle's FP64 evaluation of every iterate the GPU legs reported
        t0 = time.perf_counter()
        checks.run(orc)
        log(f"[bench] oracle: {len(checks.jobs)} iterates evaluated in {time.perf_counter() - t0:.1f}s")
    elif c1 is not None:
        c1.pop("_trajectory", None)
    prob.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
