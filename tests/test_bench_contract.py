"""CPU-side checks of bench.py's contract: one GPU only (a request for more
exits with a message), and the committed profile summaries it reads parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_profiles_read_by_bench_parse():
    import bench
    t = bench.profiled_traffic("k_async<float,64,0,4,false>")
    assert t is not None and 5e8 < t < 2e9
    assert bench.profiled_traffic("no_such_kernel") is None
    a = bench.profiled_per_launch("k_async<float,64,0,4,false>", "atomics_per_launch")
    assert a is not None and a > 0
