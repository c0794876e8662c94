"""The benchmark's command-line contract, checked on the CPU: one GPU only
(--gpus above 1 exits with a message, before touching a device), and the
JSON line's roofline / CPU-baseline helpers."""
import importlib.util
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_profiled_traffic_reads_the_committed_summary():
    b = load_bench()
    t = b.profiled_traffic("axis_double_n4096_e16_tb4_col_ct3")
    # within 20 % of the algorithmic bytes of one pass over a 256 MiB array
    # (the three-step column tile, the roofline kernel of the 2D 4096^2 headline)
    assert t is not None and 0.8 * 2 * 2**28 < t < 1.2 * 2 * 2**28


def test_cpu_library_column():
    b = load_bench()
    x = np.random.default_rng(0).standard_normal(64 * 32) + 0j
    r = b.cpu_library_fft(x, (64, 32), 2)
    assert r["workers"] == 2 and r["ms_per_transform"] > 0
