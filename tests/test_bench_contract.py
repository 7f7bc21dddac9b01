"""bench.py's host-side contract (CPU, no GPU): one GPU only, and the roofline kernel choice."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_roofline_takes_the_largest_main_stream_kernel(bench):
    # (launches, total ms, algorithmic work): the second-stream early FP64 factors have the
    # largest total but overlap the main stream, so the FP32 factor is the dominant kernel
    stats = {"band_ldlt_batch_fp32": (222, 1875.0, 222 * 92.07e9), "band_ldlt_early_fp64": (155, 1900.0, 155 * 10e9),
             "band_ldlt_batch_fp64": (212, 910.0, 212 * 12e9), "band_gather": (222, 780.0, 222 * 4e9)}
    r = bench.roofline(stats)
    assert r["kernel"] == "band_ldlt_batch_fp32" and r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    assert r["peak"] == 157.3 and abs(r["achieved"] - 92.07e9 / (1875.0 / 222 / 1e3) / 1e12) < 1e-9
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-12
    assert r["traffic"] is not None and r["traffic"] > 0  # from the committed rocprofv3 summary


def test_roofline_of_an_hbm_kernel(bench):
    stats = {"band_gather": (10, 35.0, 10 * 4e9), "band_solve_batch_fp32": (5, 1.0, 5 * 1e9)}
    r = bench.roofline(stats)
    assert r["kernel"] == "band_gather" and r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["achieved"] - 4e9 / 3.5e-3 / 1e9) < 1e-9
