"""CPU checks of bench.py's bookkeeping: the algorithmic bytes per stage,
the kernel each option set names, and the stored rocprofv3 traffic records
it reports beside its timings (a record is used only for the kernel and
workload it was counted on)."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test",
                                                  os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_algorithmic_bytes(bench):
    # 31 ids + 31 x 4 weights + 17 geometry values, then per member the own
    # state and the averaged stage epilogue (64 + 128 + 128 + 96) / 4
    assert bench.STATIC_B == 31 * 4 + 31 * 32 + 17 * 8 == 1252
    assert bench.STATE_B == 32 + (64 + 128 + 128 + 96) // 4 == 136
    assert bench.stage_bytes(40964, 1) == 40964 * 1388
    assert bench.stage_bytes(163844, 64) == 163844 * (1252 + 64 * 136)
    assert bench.STATIC_B32 == 31 * 4 + 31 * 16 + 17 * 4
    assert bench.stage_bytes(10, 1, fp32=True) == 10 * (688 + 68)


def test_kernel_names(bench):
    class A:
        fp32 = False
        members = 1
        kernel = "auto"
        nodes = 40962
        exact = False
        tiled = False
        member_block = 0
    a = A()
    assert bench.dominant_kernel(a) == "k_steps_resident"
    a.nodes = 655362
    assert bench.dominant_kernel(a) == "k_stage_pair"
    a.kernel = "tile"
    assert bench.dominant_kernel(a) == "k_stage_ptile"
    a.members, a.tiled = 64, True
    assert bench.dominant_kernel(a) == "k_stage_tspec"
    assert bench.leg_kernel({"members": 1, "kernel": "tile"}) == "k_stage_ptile"
    assert bench.leg_kernel({"members": 1, "kernel": "pair"}) == "k_stage_pair"
    assert bench.leg_kernel({"members": 64, "tiled": 1, "member_block": 32}) == "k_stage_tspec"
    assert bench.leg_kernel({"members": 64, "tiled": 0, "member_block": 64}) == "k_stage_mpair"


@pytest.mark.parametrize("kernel,n,b", [("k_steps_resident", 40962, 1),
                                        ("k_stage_ptile", 655362, 1),
                                        ("k_stage_tspec", 163842, 64),
                                        ("k_stage_pair", 40962, 1)])
def test_stored_traffic_records(bench, kernel, n, b):
    traffic, source = bench.stored_traffic(kernel, n, b)
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_n{n}_b{b}.json")
    with open(path) as fp:
        rec = json.load(fp)
    assert rec["kernel"] == kernel
    assert traffic == rec["hbm_bytes_per_launch"]
    assert traffic == pytest.approx(rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"],
                                    abs=2)
    assert "not measured in this run" in source and os.path.basename(path) in source


def test_stored_traffic_only_for_its_kernel(bench):
    assert bench.stored_traffic("k_stage_pair", 655362, 1) == (None, None)
    assert bench.stored_traffic("k_stage_tspec", 163842, 32) == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
