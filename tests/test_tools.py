"""Measurement tooling on the CPU: bench.py's one-GPU contract and the
rocprofv3 summary's calibration (no GPU needed)."""
import csv
import os
import subprocess
import sys

from fluid_amd import rocprof_summary as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


FIELDS = ["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]


def write_run(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_calibration_takes_the_largest_single_copy(tmp_path):
    """The 256 MiB calibration copy is one dispatch; later, smaller copies of
    the same kernel and grid must not pull the factor down (they did when
    the factor came from the mean over dispatches)."""
    copy = "__amd_rocclr_copyBuffer(...)"
    rows = [(1, copy, 131072, "FETCH_SIZE", 131072.0)]
    rows += [(2 + k, copy, 131072, "FETCH_SIZE", 24000.0) for k in range(20)]
    rows += [(30, "void pbf::(anonymous namespace)::k_lambda<true>(Dev)", 4096000,
              "FETCH_SIZE", 300000.0)]
    write_run(str(tmp_path / "fetch"), rows)
    f = rs.calibration(str(tmp_path / "fetch"), "FETCH_SIZE")
    assert f == (256 << 20) / 131072.0


def test_pass_mapping_covers_the_fused_kernels():
    assert rs.pass_of(rs.short("void pbf::(anonymous namespace)::k_nbr_lambda<false>(Dev)")) == "neighbors"
    assert rs.pass_of(rs.short("void pbf::(anonymous namespace)::k_vort_commit_predict<false>(Dev)")) == "vorticity"
    assert rs.pass_of(rs.short("pbf::(anonymous namespace)::k_rank(Dev, float4 const*)")) == "sort"
    assert rs.pass_of("__amd_rocclr_copyBuffer") is None
    # The prediction without movement belongs to the density queries' sort,
    # listed beside the step, not to the step's prediction.
    assert rs.pass_of(rs.short("void pbf::(anonymous namespace)::k_predict<false, true>(Dev)")) == "predict"
    assert rs.pass_of(rs.short("void pbf::(anonymous namespace)::k_predict<false, false>(Dev)")) is None
    assert rs.side_of("k_predict<false, false>") == "sort of the committed state"
    assert rs.side_of("k_density_field<true>") is not None
