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
    r = subprocess.run([CLI, "-A", prefix, "-b", prefix + "_b.txt", "-l", repr(d.lam),
                        "-t", str(tau), "-f", "0", "-I", str(inner), "-c", str(bulks),
                        "-i", "1", "-s", "21", "-r", str(tmp_path / "log.csv")],
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
