"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m fluid_amd.rocprof_summary stats DIR --command "..." --out FILE.txt
    python -m fluid_amd.rocprof_summary counters --fetch DIR --write DIR --sq DIR
        --workload NAME --command "..." --out-json FILE.json --out-txt FILE.txt

`stats` turns run_kernel_stats.csv (from --kernel-trace --stats) into a
table.  `counters` reads three --pmc runs of the same program: FETCH_SIZE,
WRITE_SIZE and an SQ pass (SQ_INSTS_VALU, SQ_WAVES, ... with --kernel-trace).
FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by the factor that
makes the calibration copy (the 256 MiB device-to-device copy that
`bench.py --calibrate` runs first, a __amd_rocclr_copyBuffer dispatch) read
and write 256 MiB.  The JSON holds, per pass of the step, the calibrated HBM
bytes and the wave-level VALU instructions of one launch; bench.py reads it.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re

# Kernel name prefix -> pass name of the step (pbf.h PBF_PASS_* order).
PASS_OF = [
    ("k_predict", "predict"),
    ("k_scan_block", "scan"),
    ("k_scan_partials", "scan"),
    ("k_scan_add", "scan"),
    ("k_scatter", "sort"),
    ("k_rank", "sort"),
    ("k_nbr_build", "neighbors"),
    ("k_nbr_lambda", "neighbors"),      # list build fused with the first lambda
    ("k_lambda", "lambda"),
    ("k_deltap", "deltap"),
    ("k_xsph", "xsph"),
    ("k_vort_commit", "vorticity"),     # also k_vort_commit_predict
]
# Kernels that stand in for one another in a pass (one or the other runs in a
# given step): averaged by dispatches like template variants, not added up.
VARIANT_OF = {"k_vort_commit_predict": "k_vort_commit"}
COPY_BYTES = 256 << 20


def short(name: str) -> str:
    name = name.replace("pbf::(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\(.*", "", name)


# Kernels beside the step that the bench also runs (listed on their own).
SIDE_OF = [
    ("k_density_field", "density lattice / points"),
    ("k_densities", "load-time densities"),
    ("k_unpermute", "read back (load order)"),
]
VALU_PEAK_LANE_OPS = 256 * 128 * 2.4e9  # published FP32 rate, FMA counted once


def side_of(kernel: str):
    # The prediction without movement sorts the committed state for the
    # density queries; it is not the step's prediction.
    if kernel.startswith("k_predict") and kernel.endswith("false>"):
        return "sort of the committed state"
    for prefix, p in SIDE_OF:
        if kernel.startswith(prefix):
            return p
    return None


def pass_of(kernel: str):
    if side_of(kernel):
        return None
    for prefix, p in PASS_OF:
        if kernel.startswith(prefix):
            return p
    return None


def read_csv(path: str):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def stats(args) -> None:
    rows = read_csv(os.path.join(args.dir, "run_kernel_stats.csv"))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"# rocprofv3 --kernel-trace --stats: {args.title}", "",
             f"command: {args.command}", "",
             f"{'kernel':40s} {'calls':>6s} {'avg us':>9s} {'min us':>9s} {'max us':>9s} {'% time':>7s}"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"{short(r['Name'])[:40]:40s} {int(r['Calls']):6d} "
                     f"{float(r['AverageNs']) / 1e3:9.1f} {float(r['MinNs']) / 1e3:9.1f} "
                     f"{float(r['MaxNs']) / 1e3:9.1f} {100 * float(r['TotalDurationNs']) / total:7.2f}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def per_kernel(path: str):
    """Mean of every counter per (kernel, grid size), over dispatches, and
    the number of dispatches ("dispatches")."""
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in read_csv(os.path.join(path, "run_counter_collection.csv")):
        k = short(r["Kernel_Name"])
        vals[(k, r["Grid_Size"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for key, d in vals.items():
        out[key] = {c: sum(v) / len(v) for c, v in d.items()}
        out[key]["dispatches"] = max(len(v) for v in d.values())
    return out


def durations(path: str):
    f = os.path.join(path, "run_kernel_trace.csv")
    if not os.path.exists(f):
        return {}
    acc = collections.defaultdict(list)
    for r in read_csv(f):
        k = short(r["Kernel_Name"])
        acc[(k, r["Grid_Size_X"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return {key: sum(v) / len(v) for key, v in acc.items()}


def calibration(path: str, counter: str) -> float:
    """HBM bytes per counted KiB, from the largest single copyBuffer dispatch
    (the 256 MiB calibration copy; the bench's later copies are smaller)."""
    best = max((float(r["Counter_Value"])
                for r in read_csv(os.path.join(path, "run_counter_collection.csv"))
                if short(r["Kernel_Name"]).startswith("__amd_rocclr_copyBuffer")
                and r["Counter_Name"] == counter), default=0.0)
    if best <= 0:
        raise SystemExit(f"no calibration copy found for {counter}: run bench.py with --calibrate")
    return COPY_BYTES / best


def gpu_shape(path: str):
    """(CUs, XCDs) of the profiled GPU from rocprofv3's agent table."""
    for r in read_csv(os.path.join(path, "run_agent_info.csv")):
        if r.get("Name", "").startswith("gfx"):
            return int(r["Cu_Count"]), int(r["Num_Xcc"])
    raise SystemExit(f"no GPU agent in {path}/run_agent_info.csv")


def counters(args) -> None:
    fetch, write, sq = per_kernel(args.fetch), per_kernel(args.write), per_kernel(args.sq)
    # Optional pass: TA_TA_BUSY_sum (address unit busy cycles, summed over
    # the CUs) against GRBM_GUI_ACTIVE (GPU busy cycles, summed over the
    # XCDs): the fraction of the launch each CU's address unit was busy.
    ta = per_kernel(args.ta) if args.ta else {}
    ncu, nxcd = gpu_shape(args.ta) if args.ta else (0, 0)
    dur = durations(args.sq)
    ff, wf = calibration(args.fetch, "FETCH_SIZE"), calibration(args.write, "WRITE_SIZE")
    out = {"workload": args.workload, "fetch_factor": ff, "write_factor": wf, "passes": {},
           "kernels": {}}
    lines = [f"# rocprofv3 --pmc summary: {args.workload}", "",
             f"command (one run per counter set): {args.command}", "",
             f"HBM bytes = FETCH_SIZE KiB x 1024 x {ff / 1024:.4f} + WRITE_SIZE KiB x 1024 x {wf / 1024:.4f}: "
             f"the 256 MiB calibration copy counts {COPY_BYTES / ff:.0f} KiB FETCH_SIZE and "
             f"{COPY_BYTES / wf:.0f} KiB WRITE_SIZE",
             "", f"{'kernel':28s} {'grid':>8s} {'us':>7s} {'HBM rd MB':>10s} {'HBM wr MB':>10s} "
             f"{'VALU Minst':>10s} {'waves':>7s} {'VMEM rd M':>9s} {'LDS bank cf':>11s} {'TA busy':>7s}"]
    pass_recs = collections.defaultdict(list)
    for key in sorted(sq, key=lambda k: -sq[k].get("SQ_INSTS_VALU", 0)):
        k, grid = key
        p = pass_of(k)
        if p is None:
            continue
        rd = fetch.get(key, {}).get("FETCH_SIZE", 0.0) * ff
        wr = write.get(key, {}).get("WRITE_SIZE", 0.0) * wf
        s = sq[key]
        rec = {"pass": p, "grid": int(grid), "us": dur.get(key), "hbm_read_bytes": rd,
               "hbm_write_bytes": wr, "valu_insts": s.get("SQ_INSTS_VALU", 0.0),
               "waves": s.get("SQ_WAVES", 0.0), "vmem_rd_insts": s.get("SQ_INSTS_VMEM_RD", 0.0),
               "lds_bank_conflict": s.get("SQ_LDS_BANK_CONFLICT", 0.0),
               "lds_idx_active": s.get("SQ_LDS_IDX_ACTIVE", 0.0)}
        t = ta.get(key, {})
        if t.get("GRBM_GUI_ACTIVE"):
            rec["ta_busy_frac"] = t["TA_TA_BUSY_sum"] / ncu / (t["GRBM_GUI_ACTIVE"] / nxcd)
        out["kernels"][f"{k}[{grid}]"] = rec
        us = f"{rec['us']:7.1f}" if rec["us"] else "      -"
        tb = f"{rec['ta_busy_frac']:7.2f}" if "ta_busy_frac" in rec else "      -"
        lines.append(f"{k[:28]:28s} {grid:>8s} {us} {rd / 1e6:10.1f} {wr / 1e6:10.1f} "
                     f"{rec['valu_insts'] / 1e6:10.1f} {rec['waves']:7.0f} "
                     f"{rec['vmem_rd_insts'] / 1e6:9.2f} {rec['lds_bank_conflict']:11.0f} {tb}")
        pass_recs[p].append((k, s["dispatches"], rec))
    for p, recs in pass_recs.items():
        # Kernels of one pass that run one after the other (scan, sort) add
        # up; the template variants of one kernel (the last delta-p of a
        # step) and the kernels of VARIANT_OF are averaged, weighted by their
        # dispatches.
        fam = collections.defaultdict(list)
        for k, n, rec in recs:
            base = re.sub(r"<.*", "", k)
            fam[VARIANT_OF.get(base, base)].append((n, rec))
        tot = {}
        for f_ in ("hbm_read_bytes", "hbm_write_bytes", "valu_insts", "lds_bank_conflict",
                   "lds_idx_active"):
            tot[f_] = sum(sum(n * r[f_] for n, r in v) / sum(n for n, _ in v) for v in fam.values())
        tot["hbm_bytes"] = tot["hbm_read_bytes"] + tot["hbm_write_bytes"]
        tas = [(n, r["ta_busy_frac"]) for _, n, r in recs if "ta_busy_frac" in r]
        if tas:
            tot["ta_busy_frac"] = sum(n * f_ for n, f_ in tas) / sum(n for n, _ in tas)
        out["passes"][p] = tot
    side = []
    for key in sorted(sq, key=lambda k: -sq[k].get("SQ_INSTS_VALU", 0)):
        k, grid = key
        what = side_of(k)
        if what is None or not dur.get(key):
            continue
        rd = fetch.get(key, {}).get("FETCH_SIZE", 0.0) * ff
        wr = write.get(key, {}).get("WRITE_SIZE", 0.0) * wf
        v = sq[key].get("SQ_INSTS_VALU", 0.0)
        us = dur[key]
        rec = {"what": what, "grid": int(grid), "us": us, "hbm_bytes": rd + wr, "valu_insts": v,
               "hbm_GBs": (rd + wr) / (us * 1e-6) / 1e9,
               "valu_frac": v * 64 / (us * 1e-6) / VALU_PEAK_LANE_OPS}
        out.setdefault("beside", {})[f"{k}[{grid}]"] = rec
        side.append(f"{k[:28]:28s} {grid:>9s} {us:8.1f} {(rd + wr) / 1e6:9.1f} {rec['hbm_GBs']:8.0f} "
                    f"{v / 1e6:9.1f} {rec['valu_frac']:6.2f}  {what}")
    if side:
        lines += ["", "beside the step (HBM GB/s of the launch; VALU as a fraction of the published "
                  "FP32 rate, 157.3 TFLOP/s = 78.6e12 lane-ops/s):",
                  f"{'kernel':28s} {'grid':>9s} {'us':>8s} {'HBM MB':>9s} {'GB/s':>8s} "
                  f"{'VALU Minst':>9s} {'VALU':>6s}  what"] + side
    lines += ["", "per launch of each pass (HBM bytes = calibrated FETCH + WRITE):"]
    for p, t in out["passes"].items():
        ta_s = f"   TA busy {t['ta_busy_frac']:.2f}" if "ta_busy_frac" in t else ""
        lines.append(f"  {p:10s} HBM {t['hbm_bytes'] / 1e6:8.1f} MB   VALU {t['valu_insts'] / 1e6:7.1f} M wave-instructions{ta_s}")
    with open(args.out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    with open(args.out_txt, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--command", required=True)
    a.add_argument("--title", default="")
    a.add_argument("--out", required=True)
    b = sub.add_parser("counters")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--sq", required=True)
    b.add_argument("--ta", default=None, help="run dir with TA_TA_BUSY_sum and GRBM_GUI_ACTIVE")
    b.add_argument("--workload", required=True)
    b.add_argument("--command", required=True)
    b.add_argument("--out-json", required=True)
    b.add_argument("--out-txt", required=True)
    args = ap.parse_args()
    stats(args) if args.cmd == "stats" else counters(args)


if __name__ == "__main__":
    main()
