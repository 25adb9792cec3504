"""The C++ host API (mosaic-art-rust_amd/host/mosaic.hpp) through its test
program, whose cases restate the reference's unit tests (src/similarity.rs,
src/adjacency.rs, src/optimizer.rs, src/color_adjustment.rs, src/lib.rs).
Host-only cases (bookkeeping, the similarity-database file format) run here;
the full program calls the GPU kernels."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mosaic-art-rust_amd")
BIN = os.path.join(PKG, "test_mosaic")


def _binary():
    if not os.path.exists(BIN):
        if not os.path.exists(os.path.join(PKG, "libmosaic_hip.so")):
            pytest.fail("libmosaic_hip.so is not built: run __graft_entry__.build() first")
        import importlib.util
        spec = importlib.util.spec_from_file_location("mh_build", os.path.join(PKG, "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build_host()
    return BIN


def _run(args, timeout):
    r = subprocess.run([_binary(), *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:])
    print(r.stderr[-2000:])
    return r


def test_host_api_host_cases():
    r = _run(["--host"], 120)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout


def test_host_api_lists_reference_cases():
    r = _run(["--list"], 60)
    names = r.stdout.split()
    for case in ["test_similarity_database_save_load", "test_adjacency_penalty_calculation",
                 "test_total_cost_calculation", "test_optimization_zero_iterations",
                 "test_greedy_optimization", "test_pixel_adjustment_hue_shift",
                 "test_calculate_average_lab_single_color", "test_usage_tracker"]:
        assert case in names


@pytest.mark.gpu
def test_host_api_all_cases_on_gpu():
    r = _run([], 300)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout
