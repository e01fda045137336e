"""bench.py's host side (CPU): the one-GPU guard, and the step's counter figures read from the
committed rocprofv3 summaries (the stage's kernel rows add up to the stage time the summary
states, and the measured bytes per cell-update follow from the rows)."""
import importlib.util
import os
import re
import subprocess
import sys

import pytest

import cfdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU per process" in r.stderr


def stated_stage(path):
    """The variants and the stage sum in the summary's own words."""
    text = open(path).read()
    m = re.search(r"Sum of the \w+ kernels of one RK stage \(variants ([^)]*)\): ([0-9.]+) us", text)
    return [v.strip() for v in m.group(1).split(", ")], float(m.group(2))


@pytest.mark.parametrize("workload", ["c2", "ns256"])
def test_counter_figures_match_the_summary(workload):
    bench = load_bench()
    path = os.path.join(ROOT, "profiles", f"block_kernels_{workload}_fp64.md")
    if not os.path.exists(path):
        pytest.skip(f"no profile summary for {workload}")
    variants, stage_us = stated_stage(path)
    # the tile indices of the stated variants
    face = [tuple(int(x) for x in v[v.index("<") + 1:-1].split(",")[2:]) for v in variants if v.startswith("k_face")]
    node = tuple(int(x) for x in variants[0][variants[0].index("<") + 1:-1].split(",")[1:])
    tiles = (cfdx.tile_variants(0).index(face[0]), cfdx.tile_variants(1).index(face[1]),
             cfdx.tile_variants(1).index(face[2]), cfdx.tile_variants(2).index(node))
    n = {"c2": 128 ** 3, "ns256": 256 ** 3}[workload]
    cf = bench.counter_figures(cfdx, workload, "fp64", tiles, True, n)
    assert cf["same_variants"] and sorted(cf["kernels"]) == sorted(variants)
    assert abs(cf["stage_profiled_us"] - stage_us) <= 0.1
    rows, _ = bench.profile_rows(workload, "fp64")
    assert cf["hbm_bytes_per_cell_update"] == pytest.approx(3 * sum(rows[v]["hbm_bytes"] for v in variants) / n)
    # every face row counts at least the 18 algorithmic values per face, the step at least 848 B
    for v in variants:
        if v.startswith("k_face"):
            assert rows[v]["hbm_bytes"] >= 0.95 * bench.face_bytes_per_node("fp64") * n
    assert cf["hbm_bytes_per_cell_update"] >= bench.ALG_BYTES_PER_CELL_UPDATE
