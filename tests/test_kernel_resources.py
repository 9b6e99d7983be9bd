"""Register budgets of the per-lane list kernels that carry the 4,096,000-
particle step (CPU only: the kernels are compiled to gfx950 assembly here).

Their occupancy is set by their VGPR counts, and code for other state sizes
shares these kernels (the dense split's quad parts sit in the per-lane
launches).  A change that raised the per-lane build + first lambda kernel
from 72 to 107 VGPRs slowed the 4M step by ~1.3% through its build alone;
these budgets hold the counts of the measured build."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# kernel (mangled-name fragment) -> largest allowed .amdhsa_next_free_vgpr
BUDGET = {
    "k_nbr_lambdaILb0E": 72,          # per-lane build + first lambda
    "k_lambdaILb1E": 72,              # lambda (packed x*)
    "k_deltapILb0ELb0ELi4E": 72,      # delta-p
    "k_deltapILb0ELb1ELi4E": 88,      # last delta-p (writes the (x*, v) records)
    "k_xsphE": 104,                   # XSPH
    "k_vort_commit_predictILb0E": 72,  # vorticity + commit + next prediction
}


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="no hipcc")
def test_per_lane_list_kernels_keep_their_register_budgets(tmp_path):
    import sys
    sys.path.insert(0, ROOT)
    from fluid_amd import build as B

    flags = [f for f in B.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    out = str(tmp_path / "k.s")
    cmd = [B.hipcc(), *flags, "--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"),
           "-o", out, os.path.join(ROOT, "fluid_amd", "csrc", "pbf_kernels.hip")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    asm = open(out).read()
    found = {}
    for m in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S):
        for frag in BUDGET:
            if frag in m.group(1) and "GLOBAL__N_" in m.group(1):
                found[frag] = int(re.search(r"\.amdhsa_next_free_vgpr (\d+)", m.group(2)).group(1))
    assert set(found) == set(BUDGET), sorted(set(BUDGET) - set(found))
    over = {k: (v, BUDGET[k]) for k, v in found.items() if v > BUDGET[k]}
    assert not over, over
