"""Host-side logic of bench.py (no GPU): the roofline object's arithmetic,
the algorithmic-bytes model of the Poisson transforms, the CPU baseline's
fields, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

import pllcart3d_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _prm(nx, ny, nz):
    return P.make_params(nx, ny, nz)


def test_transform_model_follows_plane_size(monkeypatch):
    monkeypatch.delenv("PLLC_DCT", raising=False)
    assert bench.dct_is_fast(_prm(256, 128, 128))
    assert bench.dct_is_fast(_prm(512, 256, 256))
    assert not bench.dct_is_fast(_prm(64, 32, 32))
    assert not bench.dct_is_fast(_prm(64, 128, 64))
    assert bench.touches_per_launch("dct", 12.0, _prm(512, 256, 256)) == 2.0
    assert bench.touches_per_launch("dct", 12.0, _prm(64, 32, 32)) is None
    monkeypatch.setenv("PLLC_DCT", "gemm")
    assert not bench.dct_is_fast(_prm(256, 128, 128))


def test_sweep_touches_are_the_substage_total_over_launches():
    prm = _prm(256, 128, 128)
    # x: 15 touches per substage, 3 substages, 6 launches per step
    assert bench.touches_per_launch("sweep_x", 6.0, prm) == pytest.approx(7.5)
    assert bench.touches_per_launch("sweep_y", 6.0, prm) == pytest.approx(5.5)
    assert bench.touches_per_launch("corr_pg", 3.0, prm) == 14


def test_roofline_fields_of_the_dominant_kernel():
    prm = _prm(256, 128, 128)
    cells = 256 * 128 * 128
    kernels = {
        "corr_pg": {"avg_ms": 0.2, "launches_per_step": 3.0, "ms_per_step": 0.6},
        "rhs_c": {"avg_ms": 0.1, "launches_per_step": 3.0, "ms_per_step": 0.3},
    }
    r = bench.roofline(kernels, cells, 8.0, prm)
    assert r["kernel"] == "corr_pg" and r["bound"] == "hbm" and r["unit"] == "GB/s"
    ab = 14 * cells * 8.0
    assert r["algorithmic_bytes_per_launch"] == ab
    assert r["achieved"] == pytest.approx(ab / 0.2e-3 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], rel=1e-3)
    assert r["peak"] == 8000.0
    # measured HBM bytes per launch of that kernel at this workload
    assert r["traffic"] == pytest.approx(bench.measured_traffic_mb(8.0)["corr_pg"] * 1e6)
    # the dense transform, when dominant, is priced against the matrix cores
    dense = {"dct": {"avg_ms": 1.0, "launches_per_step": 12.0, "ms_per_step": 12.0}}
    r2 = bench.roofline(dense, 64 * 32 * 32, 8.0, _prm(64, 32, 32))
    assert r2["bound"] == "mfma" and r2["peak"] == 78.6


def test_cpu_baseline_fields():
    prm = P.make_params(16, 8, 8)
    c0 = P.ic_fill(prm, P.IC_NOISE, 1, 0.01)
    cb = bench.cpu_baseline(prm, c0, 16 * 8 * 8, 2)
    assert cb["kind"] == "port" and cb["cores"] == 2 and cb["value"] > 0
    assert cb["one_thread"]["cores"] == 1 and cb["one_thread"]["value"] > 0
    assert cb["host"]["logical_cpus"] >= 1


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_traffic_comes_from_the_committed_profiles():
    """roofline.traffic is read from the rocprofv3 summaries in profiles/
    (read + write MB per dispatch of the kernel class)."""
    t8 = bench.measured_traffic_mb(8.0)
    t4 = bench.measured_traffic_mb(4.0)
    assert t8["sweep_x"] > 100 and t4["sweep_x"] > 50
    assert set(t8) >= {"sweep_x", "sweep_y", "sweep_z", "corr_pg", "rhs_c", "rhs_u", "dct"}
    assert t8["sweep_x"] == pytest.approx(2 * t4["sweep_x"], rel=0.1)
