#!/usr/bin/env python3
"""PCDM hot-path benchmark on one MI355X.

Workload: the 200,000 x 1,000,000 Lasso (column nnz truncated Zipf(1.1) on
[5,200], unit columns, 5,000-sparse planted x*, lambda = 0.05 ||A^T b||_inf)
with tau = 4096 and the ESO beta of tau.  One "step" of this benchmark is one
epoch: n = 1,000,000 coordinate updates over data resident in HBM.

value: coordinate updates/s of the FP32 + FP64-refresh path (4096
asynchronous workers, FP32 residual corrections anchored on an FP64 refresh
every 5 epochs), from x = 0.  Beside it: the plain FP64 path (synchronous
4096-coordinate steps, one gather + one scatter launch per step), the
synchronous FP32 persistent kernel, the wall time each of the FP32 and plain
FP64 paths needs from x = 0 to an FP64 relative duality gap <= 1e-6 (gap
evaluated every 5 epochs, evaluations included), the roofline of the
dominant kernel (k_async, timed with HIP events on the solver's stream), and
the CPU baseline (the oracle's OpenMP asynchronous restatement of the
reference solver) on the host cores.
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


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by a copy of known size),
    or None when the summary has no line for that kernel."""
    import re
    path = os.path.join(ROOT, "profiles", "lasso_c2_async_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.]+)", text, re.M)
    return float(m.group(1)) if m else None


def profiled_per_launch(kernel, key):
    """A per-launch figure of `kernel` from the committed rocprofv3 counter
    summary of this workload (lines `<key>[<kernel>]: <value>`), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "lasso_c2_async_atomics.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^" + key + r"\[" + re.escape(kernel) + r"\]:\s*([0-9.]+)", text, re.M)
    return float(m.group(1)) if m else None


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


def batched(acdc, args):
    """The batched workload: 256 Lasso instances of uneven sizes (m log-uniform
    in [1k,50k], n in [5k,200k], column nnz truncated Zipf(1.1) on [3,100]),
    each along a 32-value warm-started path lambda_max -> 1e-3 lambda_max,
    stopping at FP64 relative gap <= 1e-6 or the epoch cap per lambda; one
    graph-captured persistent launch (FP32 + FP64 refresh every 5 epochs,
    asynchronous workers inside each instance, spectral ESO beta per
    instance; the plain path below takes the same beta).  Beside it, the plain FP64
    path (synchronous per-step launches, one instance at a time) on sampled
    instances over the first lambdas of their paths, against the batched
    kernel on the same instances and lambdas."""
    import numpy as np
    t = time.perf_counter()
    ds = acdc.config_batch_c5(count=args.batch_count)
    nnz = np.array([d.nnz for d in ds])
    log(f"[bench] batched: generated {len(ds)} instances, nnz {nnz.sum()} in "
        f"{time.perf_counter() - t:.1f}s")
    probs = [d.problem() for d in ds]
    lams = np.stack([acdc.lambda_path(d.lambda_max, args.batch_path, 1e-3) for d in ds])
    # spectral ESO beta per instance (the batch names no beta rule; 1.04
    # median against 1.69 from omega: more pairs reach the gap under the cap)
    tuned = acdc.batch_tune(probs, step_mode=acdc.STEP_ASYNC, beta_mode=acdc.BETA_SPECTRAL)
    log(f"[bench] batched: autotuned lanes {tuned['lanes_per_coord']}, tau {tuned['tau']}, "
        f"queue chunk {tuned['queue_chunk']}")
    kw = dict(seed=5, precision=acdc.PREC_FP32_REFRESH, refresh_period=5,
              max_epochs=args.batch_max_epochs, gap_tol=args.gap_target,
              step_mode=acdc.STEP_ASYNC, lanes_per_coord=tuned["lanes_per_coord"],
              tau=tuned["tau"], queue_chunk=tuned["queue_chunk"],
              beta_mode=acdc.BETA_SPECTRAL)
    b = acdc.Batch(probs, lams, use_graph=True, **kw)  # uploads, captures the graph
    t0 = time.perf_counter()
    ms = b.run()
    wall = time.perf_counter() - t0
    r = b.results()
    upd = float(r["updates"][:, -1].astype(np.float64).sum())
    out = {"workload": f"{len(ds)} lasso instances x {args.batch_path} lambdas",
           "nnz": int(nnz.sum()), "max_epochs_per_lambda": args.batch_max_epochs,
           "gap_tol": args.gap_target, "device_ms": ms, "wall_s": wall, "updates": upd,
           "value": upd / wall, "unit": "coord updates/s",
           "pairs_converged": float(r["converged"].mean()),
           "worst_rel_gap": float(r["rel_gap"].max()),
           "tau_per_instance": int(b.info(0)[0]), "beta_mode": "spectral",
           "beta_median": float(np.median([b.info(i)[1] for i in range(len(ds))])),
           "autotuned": tuned}
    log(f"[bench] batched: {ms:.0f} ms device, {upd:.3e} updates = {upd / wall:.3e} upd/s, "
        f"{out['pairs_converged']:.3f} of (instance, lambda) pairs converged")
    b.close()
    # plain FP64 vs the batched kernel on sampled instances (quartiles of nnz)
    order = np.argsort(nnz)
    sample = [int(order[int(q * (len(ds) - 1))]) for q in (0.25, 0.5, 0.75)]
    P = min(args.batch_sample_path, args.batch_path)
    rows = []
    for i in sample:
        d, p = ds[i], probs[i]
        bi = acdc.Batch([p], lams[i:i + 1, :P], use_graph=False, **kw)
        bi.run()
        tb = bi.run() * 1e-3
        ri = bi.results()
        tau = int(bi.info(0)[0])
        bi.close()
        s = acdc.Solver(p, acdc.LOSS_L1_SQUARE, lam=float(lams[i, 0]), tau=tau, seed=5,
                        precision=acdc.PREC_FP64, launch_mode=acdc.LAUNCH_PER_STEP,
                        beta_mode=acdc.BETA_SPECTRAL)
        s.synchronize()
        t0 = time.perf_counter()
        for l in range(P):
            s.set_lambda(float(lams[i, l]))
            ep, rep = 0, s.report()
            while rep["rel_gap"] > args.gap_target and ep < args.batch_max_epochs:
                for _ in range(5):
                    s.run_updates(d.n)
                ep += 5
                rep = s.report()
        tp = time.perf_counter() - t0
        up = s.report()["updates"]
        s.close()
        rows.append({"instance": i, "m": d.m, "n": d.n, "nnz": d.nnz, "lambdas": P,
                     "batched_s": tb, "batched_updates": int(ri["updates"][0, -1]),
                     "fp64_plain_s": tp, "fp64_plain_updates": int(up),
                     "speedup": tp / tb})
        log(f"[bench] batched sample {i} ({d.m}x{d.n}): batched {tb:.3f}s vs plain fp64 "
            f"{tp:.3f}s = {tp / tb:.1f}x")
    out["fp64_plain_sampled"] = rows
    # the whole workload at the plain path's measured cost per update
    cpu = [row["fp64_plain_s"] / max(1, row["fp64_plain_updates"]) for row in rows]
    out["fp64_plain_estimate_s"] = float(np.mean(cpu) * upd)
    out["speedup_vs_fp64_plain_estimate"] = out["fp64_plain_estimate_s"] / wall
    return out


def nnls_scale(d):
    """||A^T b||_inf of a column-major dataset: the NNLS KKT residual's scale."""
    import numpy as np
    g = np.zeros(d.n)
    np.add.at(g, np.repeat(np.arange(d.n), np.diff(d.ptr)), d.vals * d.rhs[d.idx])
    return float(np.max(np.abs(g))) or 1.0


def batched_plain_sample(acdc, args, ds, probs, lams, loss, kw, sample, fixed_epochs=0):
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
        rows.append({"instance": i, "m": d.m, "n": d.n, "nnz": d.nnz, "lambdas": P,
                     "batched_s": tb, "batched_updates": int(ri["updates"][0, -1]),
                     "batched_rel_gap": float(ri["rel_gap"][0, -1]),
                     "fp64_plain_s": tp, "fp64_plain_updates": int(up),
                     "fp64_plain_rel_gap": float(rep["rel_gap"] if loss != acdc.LOSS_NNLS
                                                 else rep["kkt"] / nnls_scale(d)),
                     "speedup": tp / tb})
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


def batched_svm(acdc, args):
    """Batched hinge-SVM duals: 256 instances of uneven sizes (samples
    log-uniform in [1k, 8k], features in [10k, 200k], row nnz lognormal with
    median log-uniform in [20, 200]), each along an 8-value warm-started path
    lambda = 10/n -> 1/n, stopping at FP64 relative duality gap <= the gap
    target or 100 epochs per lambda; one graph-captured launch (FP32 + FP64
    refresh of w every 5 epochs, asynchronous workers with compare-and-swap
    steps on [0, 1], the spectral ESO beta of mSDCA.cpp per instance); beside
    it the plain FP64 path (same beta rule) on three sampled instances over
    the same paths."""
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
    b = acdc.Batch(probs, lams, use_graph=True, **kw)
    t0 = time.perf_counter()
    ms = b.run()
    wall = time.perf_counter() - t0
    r = b.results()
    upd = float(r["updates"][:, -1].astype(np.float64).sum())
    out = {"workload": f"{count} svm dual instances x 8 lambdas (10/n -> 1/n)",
           "nnz": int(nnz.sum()), "max_epochs_per_lambda": 100, "gap_tol": args.gap_target,
           "device_ms": ms, "wall_s": wall, "updates": upd, "value": upd / wall,
           "unit": "coord updates/s", "pairs_converged": float(r["converged"].mean()),
           "worst_rel_gap": float(r["rel_gap"].max()), "autotuned": tuned}
    log(f"[bench] batched svm: {ms:.0f} ms device, {upd:.3e} updates = {upd / wall:.3e} upd/s, "
        f"{out['pairs_converged']:.3f} of pairs converged")
    b.close()
    order = np.argsort(nnz)
    sample = [int(order[int(q * (count - 1))]) for q in (0.25, 0.5, 0.75)]
    out["fp64_plain_sampled"] = batched_plain_sample(acdc, args, ds, probs, lams,
                                                     acdc.LOSS_SVM_DUAL_HINGE, kw, sample)
    plain_estimate(out, upd, wall)
    return out


def batched_ttd(acdc, args):
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
                                                     kw, sample, fixed_epochs=epochs)
    plain_estimate(out, upd, wall)
    return out


def svm_c3(acdc, args):
    """Hinge-SVM dual minibatch SDCA on 20,000 samples x 1,350,000 features
    (lognormal row nnz, median 300): synchronous minibatches b in {512, 2048,
    8192} (tau = b, the spectral ESO beta of mSDCA.cpp:188-286, the rule of
    the reference's minibatch SDCA experiments), FP32 + FP64 refresh of w
    every 5 epochs and the plain FP64 path, and b asynchronous FP32 workers
    (the multicore executor's semantics and its omega beta,
    MultiCoreSolver.cpp:55-62, 108-110); updates/s over the first 20 epochs
    and the FP64 relative duality gap after them."""
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
    # time to relative gap 1e-6 from alpha = 0, gap checked every 5 epochs
    # (checks included), FP32 + refresh against plain FP64, spectral beta
    out["time_to_gap"] = []
    for b in (512, 2048, 8192):
        row = {"b": b}
        for prec, name in ((acdc.PREC_FP32_REFRESH, "fp32"), (acdc.PREC_FP64, "fp64")):
            s = acdc.Solver(p, acdc.LOSS_SVM_DUAL_HINGE, lam=d.lam, tau=b, seed=3,
                            precision=prec, refresh_period=5, beta_mode=acdc.BETA_SPECTRAL)
            tt, ep, rep = time_to_gap(s, d.n, args.gap_target, 5, 400)
            row[name] = {"seconds": tt, "epochs": ep, "rel_gap": rep["rel_gap"],
                         "reached": rep["rel_gap"] <= args.gap_target}
            s.close()
        if row["fp32"]["reached"] and row["fp64"]["reached"]:
            row["speedup_fp32_vs_fp64"] = row["fp64"]["seconds"] / row["fp32"]["seconds"]
        out["time_to_gap"].append(row)
        log(f"[bench] svm b={b} time to gap: fp32 {row['fp32']['seconds']:.3f}s "
            f"({row['fp32']['epochs']} epochs) vs fp64 {row['fp64']['seconds']:.3f}s")
    return out


def ttd_c4(acdc, args):
    """Truss topology design NNLS on the 40^3-cell ground structure (844,840
    bars, 201,720 DOFs): tau in {1024, 8192}, FP32 + FP64 refresh
    (asynchronous workers) and the plain FP64 path (synchronous per-step
    launches); updates/s over 50 epochs from x = 0 and the FP64 objective and
    KKT residual after them."""
    d = acdc.gen_ttd(40)
    p = d.problem()
    out = {"workload": f"ttd nnls 40^3 cells, {d.n} bars x {d.m} dofs", "nnz": d.nnz,
           "epochs": 50, "rows": []}
    for tau in (1024, 8192):
        for prec, mode, name in ((acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, "fp32 async"),
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
            out["rows"].append({"tau": tau, "path": name, "beta": s.beta,
                                "updates_per_s": 50 * d.n / wall, "primal": r["primal"],
                                "primal_at_0": p0, "kkt": r["kkt"]})
            log(f"[bench] ttd tau={tau} {name}: {50 * d.n / wall:.3e} upd/s, "
                f"objective {r['primal']:.4e} (from {p0:.4e}), kkt {r['kkt']:.2e}")
            s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed epochs")
    ap.add_argument("--warmup", type=int, default=2, help="untimed epochs")
    ap.add_argument("--tau", type=int, default=4096)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per coordinate (0 = autotune)")
    ap.add_argument("--m", type=int, default=200_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--gap-target", type=float, default=1e-6)
    ap.add_argument("--gap-every", type=int, default=5)
    ap.add_argument("--max-epochs", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttg", action="store_true", help="skip the time-to-gap runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-svm-ttd", action="store_true", help="skip the SVM and TTD workloads")
    ap.add_argument("--no-probes", action="store_true",
                    help="skip the streaming-copy and atomic-rate probes (a child process "
                         "running PyTorch kernels)")
    ap.add_argument("--probe-device", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-path", type=int, default=32)
    ap.add_argument("--batch-max-epochs", type=int, default=100)
    ap.add_argument("--batch-sample-path", type=int, default=8)
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

    # ---- lanes per coordinate: the autotuner unless given
    if args.lanes == 0:
        tl = acdc.solver_tune_lanes(prob, acdc.LOSS_L1_SQUARE, lam=d.lam, tau=tau, seed=2)
        args.lanes = tl["lanes_per_coord"]
        log(f"[bench] autotuned lanes per coordinate: {args.lanes}")

    # ---- FP32 + FP64 refresh, asynchronous workers: the timed epochs
    s32 = make(acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, lanes=args.lanes)
    beta = s32.beta
    s32.run_updates(args.warmup * n)
    s32.x = np.zeros(n)
    tp = throughput(s32, n, args.steps)
    value, wall = tp["value"], tp["wall"]
    log(f"[bench] fp32 async: {tp['updates']} updates in {wall:.3f}s = {value:.4e} upd/s")
    s32.close()

    # ---- dominant kernel: k_async over the same timed epochs (one launch per
    #      epoch), HIP events on the solver's stream around the launches
    launch_ms = tp["step_ms"] / max(1, tp["launches"])
    upd_per_launch = tp["updates"] / max(1, tp["launches"])
    bytes_per_update = 8.0 * mean_nnz + 32.0  # (row, value) per nnz; ptr pair, L_j, gref_j, x_j
    bytes_per_launch = upd_per_launch * bytes_per_update
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    # segments longer than the R = 4 register chunks of a group take the TAIL variant
    kernel = f"k_async<float,{args.lanes},0,4,{'true' if maxlen > 4 * args.lanes else 'false'}>"
    log(f"[bench] k_async: {launch_ms:.3f} ms per launch ({upd_per_launch:.0f} updates), "
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
        # the residual scatter's L2 atomics: per launch from the committed
        # counter summary, over this run's launch time, against the random
        # FP32 atomic rate of the device measured just before
        "atomics": {"per_s": (atomics_per_launch / (launch_ms * 1e-3)
                              if atomics_per_launch else None),
                    "probe_random_fp32_per_s": atomic_probe,
                    "frac": (atomics_per_launch / (launch_ms * 1e-3) / atomic_probe
                             if atomics_per_launch and atomic_probe else None)},
    }

    # ---- plain FP64 path: synchronous steps, per-step launches
    s64 = make(acdc.PREC_FP64, acdc.STEP_SYNC, launch=acdc.LAUNCH_PER_STEP)
    s64.run_updates(args.warmup * n)
    s64.x = np.zeros(n)
    tp64 = throughput(s64, n, max(1, args.steps // 5))
    result["fp64_plain"] = {"value": tp64["value"], "unit": "coord updates/s",
                            "mode": "synchronous tau-steps, gather + scatter launch per step"}
    log(f"[bench] fp64 plain: {tp64['value']:.4e} upd/s")

    # ---- synchronous FP32 steps in the persistent cooperative kernel
    ss = make(acdc.PREC_FP32_REFRESH, acdc.STEP_SYNC)
    ss.run_updates(args.warmup * n)
    ss.x = np.zeros(n)
    tps = throughput(ss, n, max(1, args.steps // 5))
    result["fp32_sync_persistent"] = {"value": tps["value"], "unit": "coord updates/s",
                                      "launches_per_epoch": tps["launches"] / max(1, args.steps // 5)}
    log(f"[bench] fp32 sync: {tps['value']:.4e} upd/s")
    ss.close()

    x32 = None
    if not args.no_ttg:
        s = make(acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, lanes=args.lanes)
        tt, ep, rep = time_to_gap(s, n, args.gap_target, args.gap_every, args.max_epochs)
        result["time_to_gap"] = {"seconds": tt, "epochs": ep, "rel_gap": rep["rel_gap"],
                                 "reached": rep["rel_gap"] <= args.gap_target,
                                 "check_every_epochs": args.gap_every}
        log(f"[bench] fp32: rel gap {rep['rel_gap']:.3e} after {ep} epochs in {tt:.3f}s")
        x32 = s.x
        s.close()
        tt, ep, rep = time_to_gap(s64, n, args.gap_target, args.gap_every, args.max_epochs)
        result["fp64_plain"]["time_to_gap"] = {"seconds": tt, "epochs": ep,
                                               "rel_gap": rep["rel_gap"],
                                               "reached": rep["rel_gap"] <= args.gap_target}
        log(f"[bench] fp64 plain: rel gap {rep['rel_gap']:.3e} after {ep} epochs in {tt:.3f}s")
        if result["time_to_gap"]["reached"] and result["fp64_plain"]["time_to_gap"]["reached"]:
            result["time_to_gap"]["speedup_vs_fp64_plain"] = (
                result["fp64_plain"]["time_to_gap"]["seconds"] / result["time_to_gap"]["seconds"])
        # the same two paths with the spectral ESO beta of mSDCA.cpp (20 power
        # steps on the normalised Gram matrix) instead of the configured
        # omega rule: a side result, both paths with the same beta
        spec = {}
        for name, prec, mode, launch, lanes in (
                ("fp32_async", acdc.PREC_FP32_REFRESH, acdc.STEP_ASYNC, acdc.LAUNCH_AUTO, args.lanes),
                ("fp64_plain", acdc.PREC_FP64, acdc.STEP_SYNC, acdc.LAUNCH_PER_STEP, 0)):
            sp = make(prec, mode, launch=launch, lanes=lanes, beta_mode=acdc.BETA_SPECTRAL)
            tt, ep, rep = time_to_gap(sp, n, args.gap_target, args.gap_every, args.max_epochs)
            spec[name] = {"seconds": tt, "epochs": ep, "rel_gap": rep["rel_gap"],
                          "reached": rep["rel_gap"] <= args.gap_target}
            spec["beta"] = sp.beta
            log(f"[bench] spectral beta {sp.beta:.4f} {name}: rel gap {rep['rel_gap']:.3e} "
                f"after {ep} epochs in {tt:.3f}s")
            sp.close()
        if spec["fp32_async"]["reached"] and spec["fp64_plain"]["reached"]:
            spec["speedup_vs_fp64_plain"] = (spec["fp64_plain"]["seconds"] /
                                             spec["fp32_async"]["seconds"])
        result["time_to_gap_spectral_beta"] = spec
    s64.close()

    if not args.no_batch:
        result["batched"] = batched(acdc, args)
        result["batched_svm"] = batched_svm(acdc, args)
        result["batched_ttd"] = batched_ttd(acdc, args)
    if not args.no_svm_ttd:
        result["svm"] = svm_c3(acdc, args)
        result["ttd"] = ttd_c4(acdc, args)

    if not args.no_cpu:
        # OpenMP threads spread over the host's cores (read when the OpenMP
        # runtime starts, with the oracle library below)
        os.environ.setdefault("OMP_PROC_BIND", "spread")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as orc
        threads = max(1, min(16, os.cpu_count() or 1))
        o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam,
                   beta=orc.eso_beta(prob.omega, threads, d.n))
        if x32 is not None:
            result["time_to_gap"]["rel_gap_fp64_oracle"] = o.eval(x32)["rel_gap"]
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
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
