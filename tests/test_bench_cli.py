"""bench.py on the CPU: it parses, refuses more than one GPU with a message,
and its roofline helpers are consistent."""
import os
import py_compile
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_bench_compiles():
    py_compile.compile(BENCH, doraise=True)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_mixed_peak_between_fp64_and_fp32():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.PEAK_FP64 < bench.mixed_peak() < bench.PEAK_FP32
    assert bench.read_traffic("mixed", "no-such-workload") is None
