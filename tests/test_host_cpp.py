"""The C++ host mirror: loaders and log rows (CPU), and the command-line
front end driving the gfx950 executor (GPU) against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ac-dc_amd", "bin", "acdc_hip_solver")


def test_loaders_and_log_rows(tmp_path):
    exe = tmp_path / "test_loaders"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    os.path.join(ROOT, "tests", "cpp", "test_loaders.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "ok" in out.stdout


def test_cli_usage_and_errors(acdc, tmp_path):
    assert os.path.exists(CLI), "build the front end first (make -C ac-dc_amd)"
    assert subprocess.run([CLI], capture_output=True).returncode == 2
    r = subprocess.run([CLI, "-A", str(tmp_path / "nothing"), "-b", str(tmp_path / "b")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "file not found" in r.stderr


def write_csc(prefix, d):
    np.savetxt(f"{prefix}_values.txt", d.vals, fmt="%.17g")
    np.savetxt(f"{prefix}_colPtr.txt", d.ptr, fmt="%d")
    np.savetxt(f"{prefix}_rowIdx.txt", d.idx, fmt="%d")
    np.savetxt(f"{prefix}_b.txt", d.rhs, fmt="%.17g")


@pytest.mark.gpu
def test_cli_lasso_matches_oracle(acdc, orc, tmp_path):
    d = acdc.gen_lasso(400, 1500, nnz_fixed=12, xstar_nnz=15, noise=0.01, lambda_frac=0.1, seed=9)
    prefix = str(tmp_path / "lasso")
    write_csc(prefix, d)
    tau, bulks = 16, 6
    inner = 4 * d.n  # a multiple of tau
    # synchronous tau-steps with the omega rule: replayed by the oracle
    r = subprocess.run([CLI, "-A", prefix, "-b", prefix + "_b.txt", "-l", repr(d.lam),
                        "-t", str(tau), "-f", "0", "-I", str(inner), "-c", str(bulks),
                        "-i", "1", "-s", "21", "-M", "0", "-B", "0",
                        "-r", str(tmp_path / "log.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    last = float(re.search(r"Last objective value: (\S+)", r.stdout).group(1))
    beta = float(re.search(r"ESO beta: (\S+)", r.stdout).group(1))
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=beta)
    o.run(seed=21, tau=tau, step0=0, nsteps=bulks * inner // tau)
    assert last == pytest.approx(o.eval()["primal"], rel=1e-10)
    rows = open(tmp_path / "log.csv").read().strip().splitlines()
    assert len(rows) == bulks + 1 and rows[-1].startswith(f"{bulks},")


@pytest.mark.gpu
def test_svm_executor_test_accuracy(acdc, tmp_path):
    """The SVM log row's test-set 0-1 column (computed on the host from the
    device's w) equals the device's training 0-1 accuracy when the test set
    is the training set."""
    exe = tmp_path / "test_svm_test_accuracy"
    lib = os.path.join(ROOT, "ac-dc_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "ac-dc_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "test_svm_test_accuracy.cpp"),
                    "-L", lib, "-lacdc_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")


@pytest.mark.gpu
def test_cli_reference_rule_serial_matches_oracle(acdc, orc, tmp_path):
    """Default flags: the reference multicore executor's rule -- asynchronous
    workers, and the Lasso's L_j = 1/||a_j||^2 whatever sigma the front end
    computed (MulticoreSquareLoss.h:27-40).  With one worker the run is the
    serial method over worker 0's stream, which the oracle replays."""
    d = acdc.gen_lasso(300, 1200, nnz_fixed=10, xstar_nnz=12, noise=0.01, lambda_frac=0.1,
                       seed=19)
    prefix = str(tmp_path / "lasso")
    write_csc(prefix, d)
    inner, bulks = 3 * d.n, 4
    r = subprocess.run([CLI, "-A", prefix, "-b", prefix + "_b.txt", "-l", repr(d.lam),
                        "-t", "1", "-f", "0", "-I", str(inner), "-c", str(bulks), "-i", "1",
                        "-s", "23"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert float(re.search(r"ESO beta: (\S+)", r.stdout).group(1)) == 1.0
    last = float(re.search(r"Last objective value: (\S+)", r.stdout).group(1))
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=1.0)
    o.run_list(orc.async_coords(23, 0, 0, inner * bulks, d.n))
    assert last == pytest.approx(o.eval()["primal"], rel=1e-12)


@pytest.mark.gpu
def test_cli_reference_rule_async_reaches_oracle_optimum(acdc, orc, tmp_path):
    """tau = 16 asynchronous workers with beta = 1 (the reference multicore
    Lasso at 16 threads): the last logged objective and duality gap reach the
    optimum the serial oracle converges to."""
    d = acdc.gen_lasso(300, 1200, nnz_fixed=10, xstar_nnz=12, noise=0.01, lambda_frac=0.1,
                       seed=19)
    prefix = str(tmp_path / "lasso")
    write_csc(prefix, d)
    r = subprocess.run([CLI, "-A", prefix, "-b", prefix + "_b.txt", "-l", repr(d.lam),
                        "-t", "16", "-f", "0", "-I", str(50 * d.n), "-c", "6", "-i", "1",
                        "-s", "5", "-r", str(tmp_path / "log.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert float(re.search(r"ESO beta: (\S+)", r.stdout).group(1)) == 1.0
    last = float(re.search(r"Last objective value: (\S+)", r.stdout).group(1))
    o = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, d.lam, beta=1.0)
    o.run(seed=5, tau=1, step0=0, nsteps=400 * d.n)
    opt = o.eval()
    assert opt["rel_gap"] < 1e-12
    assert last == pytest.approx(opt["primal"], rel=1e-9)
    # the Lasso row's last column is the duality gap
    gap = float(open(tmp_path / "log.csv").read().strip().splitlines()[-1].split(",")[-1])
    assert 0 <= gap <= 1e-9 * opt["primal"]


@pytest.mark.gpu
def test_executor_beta_rules(acdc, tmp_path):
    """HipEngineExecutor's step scaling: the reference rule (Lasso/NNLS
    L_j = 1/||a_j||^2 at any thread count; SVM 1/(sigma ||x_i||^2) with the
    caller's sigma), an explicit omega rule, and forcedSigma."""
    exe = tmp_path / "test_executor_beta"
    lib = os.path.join(ROOT, "ac-dc_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "ac-dc_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "test_executor_beta.cpp"),
                    "-L", lib, "-lacdc_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")


PATH_CLI = os.path.join(ROOT, "ac-dc_amd", "bin", "acdc_hip_lasso_path")


def dense_csv(path_a, path_b, d):
    """The dataset as the dense CSV files the lambda-path experiment reads."""
    a = np.zeros((d.m, d.n))
    for j in range(d.n):
        a[d.idx[d.ptr[j]:d.ptr[j + 1]], j] += d.vals[d.ptr[j]:d.ptr[j + 1]]
    np.savetxt(path_a, a, fmt="%.17g", delimiter=",")
    np.savetxt(path_b, d.rhs[:, None], fmt="%.17g", delimiter=",")
    return a


def test_lasso_path_cli_reads_dense_csv(acdc, tmp_path):
    """The CSV reader (rows x columns from the file, column-major storage) and
    the argument checks; without a device the run stops after loading."""
    assert os.path.exists(PATH_CLI), "build the front end first (make -C ac-dc_amd)"
    assert subprocess.run([PATH_CLI], capture_output=True).returncode == 2
    d = acdc.gen_lasso(7, 5, nnz_fixed=3, seed=2)
    dense_csv(tmp_path / "a.csv", tmp_path / "b.csv", d)
    r = subprocess.run([PATH_CLI, "-A", str(tmp_path / "a.csv"), "-b", str(tmp_path / "b.csv"),
                        "-o", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert "Loaded data with m=7 and n=5" in r.stdout
    (tmp_path / "bad.csv").write_text("1,2\n3\n")
    r = subprocess.run([PATH_CLI, "-A", str(tmp_path / "bad.csv"), "-b", str(tmp_path / "b.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "different lengths" in r.stderr


@pytest.mark.gpu
def test_lasso_path_cli_matches_oracle(acdc, orc, tmp_path):
    """The lambda path of lasso_console_solver.cpp on the device: the log's
    rows and the per-lambda x files; the last lambda's x is the oracle's
    optimum (relative gap <= 1e-6 evaluated by the FP64 oracle at the
    printed lambda, which carries six significant digits as the reference
    prints it)."""
    d = acdc.gen_lasso(150, 400, nnz_fixed=8, xstar_nnz=8, noise=0.01, lambda_frac=0.1, seed=23)
    dense_csv(tmp_path / "a.csv", tmp_path / "b.csv", d)
    hi = float(d.lambda_max)
    r = subprocess.run([PATH_CLI, "-A", str(tmp_path / "a.csv"), "-b", str(tmp_path / "b.csv"),
                        "-H", repr(hi), "-L", repr(hi * 0.05), "-N", "4", "-R", "40", "-m", "50",
                        "-t", "4", "-p", "0", "-g", "1e-12", "-o", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lams = [float(v) for v in re.findall(r"^lambda (\S+)$", r.stdout, re.M)]
    assert len(lams) == 5
    rows = open(tmp_path / "lasso.log").read().strip().splitlines()
    assert rows[0].split(",")[:3] == ["4", str(d.n), str(d.m)]
    assert all(len(row.split(",")) == 8 for row in rows[1:])
    files = sorted(p.name for p in tmp_path.glob("lasso_x_*.log"))
    assert len(files) == 5
    name = re.findall(r"^lambda (\S+)$", r.stdout, re.M)[-1]
    x = np.loadtxt(tmp_path / f"lasso_x_{name}.log")
    e = orc.CD(0, d.m, d.n, d.ptr, d.idx, d.vals, d.rhs, lams[-1]).eval(x)
    assert e["rel_gap"] <= 1e-6, e


@pytest.mark.gpu
def test_cli_time_budget_stops_the_bulks(acdc, tmp_path):
    """-T: the bulk loop ends once its pure computation time reaches the
    budget, long before the (huge) bulk count."""
    import time
    d = acdc.gen_lasso(300, 1200, nnz_fixed=10, xstar_nnz=12, noise=0.01, lambda_frac=0.1,
                       seed=19)
    prefix = str(tmp_path / "lasso")
    write_csc(prefix, d)
    t0 = time.perf_counter()
    r = subprocess.run([CLI, "-A", prefix, "-b", prefix + "_b.txt", "-l", repr(d.lam),
                        "-t", "16", "-f", "0", "-I", str(d.n), "-c", "100000000", "-i", "1",
                        "-T", "0.2", "-r", str(tmp_path / "log.csv")],
                       capture_output=True, text=True, timeout=120)
    wall = time.perf_counter() - t0
    assert r.returncode == 0, r.stderr
    rows = open(tmp_path / "log.csv").read().strip().splitlines()
    assert 1 <= len(rows) < 100000000 and wall < 60
