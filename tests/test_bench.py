"""The benchmark's host side: its workloads build valid cases, multi-GPU
requests are refused before any device work, and the CPU baseline leg times
the oracle on a bounded sample."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from higflow_amd import cases as hcases  # noqa: E402


@pytest.mark.parametrize("name", sorted(bench.WORKLOADS))
def test_workloads_build_cases(name):
    spec = bench.WORKLOADS[name]
    case = bench.make_case(spec, rtol=1e-10, atol=1e-10, maxit=100)
    cases = case if isinstance(case, list) else [case]
    assert cases and all(c["dim"] in (2, 3) and c["domains"] for c in cases)
    assert spec["steps"] > 0 and spec["dt"] > 0
    if spec.get("semi") or spec["kind"] in ("sweep", "contraction"):
        assert all(c["params"]["tempdiscrtype"] == hcases.SEMI_IMPLICIT_EULER for c in cases)
    if spec["kind"] == "sweep":
        assert len(cases) == spec["count"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_times_the_oracle():
    case = bench.make_case(dict(kind="cavity", n=8, Re=100.0, dt=1e-3, steps=3))
    rate, steps, wall = bench.cpu_baseline(case, 64, budget_s=0.2, max_steps=3)
    assert rate > 0 and 1 <= steps <= 3 and wall > 0


def test_all_core_leg_steps_sweep_members_in_parallel():
    from higflow_amd import cases
    # the seed-0 parameters on small sides keep the oracle steps short
    rows = [(8, re, wi, 5e-4, dict(rtol=1e-10, atol=1e-10, maxit=1000))
            for (_, re, wi) in cases.sweep_table(4, 0)]
    rate, done, wall = bench.cpu_baseline_all_cores(rows, 2, budget_s=60.0)
    assert rate > 0 and done == 4 and wall > 0
    assert bench.cpu_baseline_all_cores(None, 2) is None


def test_default_workload_is_the_largest_single_gpu_configuration():
    assert bench.DEFAULT == "sweep64_ve_cavity"
    rid, name, form = bench.WORKLOADS[bench.DEFAULT]["roof"]
    # the fine-level V-cycle launch of the velocity AMG-PCG dominates the
    # mixed step's kernel trace
    assert rid == 13 and form == "mixed" and "k_amg_prolong_post" in name
