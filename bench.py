#!/usr/bin/env python3
"""Throughput of the MI355X PBF step (one GPU).

Default workload: the 4,096,000-particle two-sphere collision (h = 0.02031,
4 solver iterations, dt 0.005).  A step is one full Particles::timeStep
(src/particles.cpp:250-297) over all particles, with the state resident in
HBM.  Prints ONE JSON line:

* value      particle-steps/s = N * steps / wall time of the timed steps
             (one time_step(K) call and a synchronize).  The timed steps start
             from the scene's initial state (reloaded after the warm-up), so
             the default K = 100 is the whole run, through the impact of the
             two spheres near step 50;
* windows    the same steps again with HIP events around every launch, in
             two halves (steps 1-50 before the impact, 51-100 through it):
             ms per step of every pass and the mean / largest neighbour
             count at the end of each half;
* roofline   the dominant pass (most time per step), from HIP events recorded
             around every launch on the context stream during K more steps
             run right after the timed ones: algorithmic state bytes per
             launch (B/particle table below x N) / mean launch time, against
             8 TB/s.  `traffic` is the HBM bytes of one launch from the
             committed rocprofv3 counters (calibrated FETCH_SIZE +
             WRITE_SIZE); `valu` is the FP32 roof: wave-level VALU
             instructions of one launch (committed counters) x 64 lanes /
             mean launch time, against 157.3 TFLOP/s as 78.6e12 lane
             operations per second;
* passes     ms per step, GB/s (state-only and state+list) of every pass,
             and the LDS bank-conflict rate (SQ_LDS_BANK_CONFLICT /
             SQ_LDS_IDX_ACTIVE) from the committed counters;
* dam_break  the 262,144-particle dam break's whole 200-step run, in windows
             of 50 steps, with the same per-pass table;
* small_scenes  the reference's own scenes (particles/p.xml, 720 particles,
             and particles/spheres_p.xml, 2,106) at its native constants;
* reference_distance  per-step error and 100-step drift against the FP64
             all-pairs step (sequential and Jacobi sweeps), as measured by
             tests/test_gpu_reference_distance.py (profiles/reference_distance.json);
* readback, estimate_densities, density_field, snapshot, mesh, batched,
             adjoint  the paths beside the step (positions for the CPU
             consumers, load-time densities, the marching-cubes density
             lattice, checkpoint files, the triangle-mesh collider, many
             small scenes, gradients);
* cpu_baseline  the CPU restatement (oracle/, FP32, grid, Jacobi, OpenMP on
             every host core) timed on a few steps of the same scene, and
             (`workloads`) on the dam block's 100 steps and steps 1-50 of
             the dam break (all cores; one core on 100 and 2 steps), beside
             the GPU's rate on those workloads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X's published FP32 peak, 157.3 TFLOP/s, as lane operations (an FMA
# counts once): 256 CUs x 128 lanes x 2.4 GHz.  The figure counts packed FP32
# (two operations per lane and instruction); code built without packed FP32,
# as this library is, issues one wave64 VALU instruction per 4 cycles per
# SIMD, half of it.
VALU_PEAK_LANE_OPS = 256 * 128 * 2.4e9

# Algorithmic FP32 state bytes per particle of every pass (reads + writes of
# the particle state the pass needs; neighbour gathers assumed on-chip).
STATE_BYTES = {
    "predict": 64,    # x, v in; x*, v, cell out (+ histogram atomic)
    "scan": 0,        # 8 bytes per grid cell, added below
    "sort": 84,       # cell, rank, x, x*, v in; ids, x, x*, v out
    "neighbors": 16,  # x* in (+ the list it writes, state+list)
    "lambda": 16,     # x* in, lambda out
    "deltap": 28,     # x*, lambda in, x*' out
    "xsph": 48,       # x, x* in (v derived); v, omega out
    "vorticity": 60,  # x*, v, omega in; v, x out
}
LIST_PASSES = ("neighbors", "lambda", "deltap", "xsph", "vorticity")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def load_counters(workload: str):
    """Per-pass HBM bytes and VALU instructions of one launch, measured with
    rocprofv3 --pmc and committed as profiles/<workload>_counters.json
    (written by fluid_amd.rocprof_summary)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(sc, steps: int):
    from oracle import oracle

    threads = len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    log(f"cpu baseline: oracle grid/Jacobi FP32, {threads} threads, {steps} steps of {sc.name}")
    t0 = time.perf_counter()
    oracle.step(sc.params, sc.x, sc.v, steps)
    dt = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    # The same restatement on one core, one step.
    oracle.set_threads(1)
    t0 = time.perf_counter()
    oracle.step(sc.params, sc.x, sc.v, 1)
    dt1 = time.perf_counter() - t0
    oracle.set_threads(threads)
    log(f"cpu baseline: {sc.n * steps / dt:.3e} particle-steps/s on {threads} threads, "
        f"{sc.n / dt1:.3e} on 1")
    return {
        "value": sc.n * steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sc.name}: {steps} steps from the initial state, oracle/ restatement "
                  f"(FP32, uniform grid, Jacobi XSPH, OpenMP), {dt:.2f} s; host {model}, "
                  f"nproc {os.cpu_count()}, {threads} usable",
        "one_core": {"value": sc.n / dt1, "unit": "particle-steps/s", "cores": 1,
                     "sample": f"{sc.name}: 1 step, {dt1:.2f} s"},
    }


def cpu_baseline_workloads(dam_break_gpu):
    """The CPU restatement on the dam block (4,096 particles, its 100 steps,
    all host cores and one) and on a bounded sample of the dam break
    (262,144 particles: steps 1-50 of its 200 on all cores, ~10 s; steps 1-2
    on one), beside the GPU's rate (the dam block's 100 steps timed here,
    the dam break's whole run from its bench entry, whose later steps carry
    three times the neighbours of the sample's)."""
    from fluid_amd import scenes
    from fluid_amd.pbf import Particles
    from oracle import oracle

    threads = len(os.sched_getaffinity(0))
    out = {}

    def cpu(sc, steps, nthreads):
        oracle.set_threads(nthreads)
        t0 = time.perf_counter()
        oracle.step(sc.params, sc.x, sc.v, steps)
        dt = time.perf_counter() - t0
        return {"value": sc.n * steps / dt, "unit": "particle-steps/s", "cores": nthreads,
                "steps": steps, "seconds": dt}

    sc = scenes.dam_block()
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=sc.steps)
        c.load(sc.x, sc.v)
        c.sync()
        t0 = time.perf_counter()
        c.time_step(nsteps=sc.steps)
        c.sync()
        wall = time.perf_counter() - t0
    out[sc.name] = {"gpu": {"value": sc.n * sc.steps / wall, "unit": "particle-steps/s",
                            "steps": sc.steps, "ms_per_step": wall / sc.steps * 1e3},
                    "cpu_all_cores": cpu(sc, sc.steps, threads),
                    "cpu_one_core": cpu(sc, sc.steps, 1)}
    sc = scenes.dam_break()
    entry = {"cpu_all_cores": cpu(sc, 50, threads), "cpu_one_core": cpu(sc, 2, 1)}
    if dam_break_gpu:
        entry["gpu"] = {"value": dam_break_gpu["particle_steps_per_s"], "unit": "particle-steps/s",
                        "steps": dam_break_gpu["steps"], "ms_per_step": dam_break_gpu["ms_per_step"]}
    out[sc.name] = entry
    oracle.set_threads(threads)
    for name, e in out.items():
        log(f"cpu {name}: {e['cpu_all_cores']['value']:.3e} particle-steps/s on {threads} cores, "
            f"{e['cpu_one_core']['value']:.3e} on 1" +
            (f"; GPU {e['gpu']['value']:.3e}" if "gpu" in e else ""))
    return out


def batched_bench(steps: int, looped_steps: int):
    """512 scenes (counts log-uniform in [1000, 32768]): autotuned batched
    plan vs the plain path looped scene by scene; a sample of scenes is
    checked bit for bit between the two."""
    import numpy as np

    from fluid_amd import batch, scenes

    t0 = time.perf_counter()
    items, p = scenes.batched()
    counts = np.array([len(it[0]) for it in items])
    log(f"batched: {len(items)} scenes, {counts.sum()} particles "
        f"(min {counts.min()}, max {counts.max()}), generated in {time.perf_counter() - t0:.1f} s")
    plan, table = batch.autotune(p, items, steps=3, warmup=1, log=log)
    runner = batch.BatchRunner(p, items, plan)
    runner.step(1)
    t0 = time.perf_counter()
    runner.step(steps)
    wall_b = time.perf_counter() - t0
    rate_b = len(items) * steps / wall_b
    log(f"batched plan {plan.describe()}: {steps} steps in {wall_b:.3f} s, {rate_b:.1f} scene-steps/s")
    runner.close()
    # Parity sample: first scenes, batched plan vs looped plain path.
    sample = items[:4]
    r = batch.BatchRunner(p, sample, plan)
    r.step(looped_steps)
    got = r.read()
    r.close()
    wall_l, looped_states = batch.run_looped(p, items, looped_steps)
    rate_l = len(items) * looped_steps / wall_l
    exact = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                for g, s in zip(got, looped_states[:4]) for a, b in zip(g, s))
    log(f"looped plain path: {looped_steps} steps in {wall_l:.3f} s, {rate_l:.1f} scene-steps/s; "
        f"speedup {rate_b / rate_l:.2f}x, sample bit-exact {exact}")
    return {
        "metric": "batched vs looped scene-steps/s",
        "scenes": len(items),
        "particles": int(counts.sum()),
        "plan": plan.describe(),
        "batched_scene_steps_per_s": rate_b,
        "batched_steps": steps,
        "looped_scene_steps_per_s": rate_l,
        "looped_steps": looped_steps,
        "speedup": rate_b / rate_l,
        "sample_bit_exact": exact,
        "autotune": {pl.describe(): rt for pl, rt in table},
    }


def density_field_bench(ctx, sc):
    """Density on the marching-cubes corner lattice of the scene's box at
    step h/2 (the reference's H * FSTEPSIZE_RATIO, src/particles.cpp:396-397),
    after the timed steps: the first call sorts the committed state by cell,
    the second reuses it."""
    import numpy as np

    p = sc.params
    lo = np.array(p.box_lo, np.float32)
    hi = np.array(p.box_hi, np.float32)
    step = np.float32(p.h * 0.5)
    counts = [int((hi[a] - lo[a]) / step) + 2 for a in range(3)]
    m = int(np.prod(counts))
    t0 = time.perf_counter()
    ctx.density_lattice(lo, hi, float(step), counts)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat = ctx.density_lattice(lo, hi, float(step), counts)
    again = time.perf_counter() - t0
    log(f"density lattice {counts} = {m} points: {first * 1e3:.1f} ms (with sort), "
        f"{again * 1e3:.1f} ms cached")
    return {
        "metric": "density lattice points/s",
        "points": m,
        "counts": counts,
        "step": float(step),
        "ms_with_sort": first * 1e3,
        "ms": again * 1e3,
        "points_per_s": m / again,
        "surface_points": int((lat >= 0.95 * p.rest_density).sum()),
    }


def snapshot_bench(ctx, sc):
    """Binary snapshot of the committed state (pbf_save_snapshot) written to
    a temporary file and loaded back (pbf_load_snapshot, checksum verified);
    the state read after the load must equal the state before, bit for bit."""
    import tempfile

    import numpy as np

    before = ctx.read()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.pbfsnap")
        t0 = time.perf_counter()
        ctx.save_snapshot(path)
        save = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        ctx.load_snapshot(path)
        load = time.perf_counter() - t0
    after = ctx.read()
    exact = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(before, after))
    log(f"snapshot {size / 1e6:.1f} MB: save {save * 1e3:.1f} ms, load {load * 1e3:.1f} ms, "
        f"round trip bit-exact {exact}")
    return {
        "metric": "snapshot MB/s",
        "bytes": size,
        "save_ms": save * 1e3,
        "load_ms": load * 1e3,
        "save_MBs": size / save / 1e6,
        "load_MBs": size / load / 1e6,
        "round_trip_bit_exact": exact,
    }


def mesh_bench(sc, warmup: int, steps: int):
    """The same scene with the mesh collider: the box as 12 triangles plus an
    icosphere obstacle (5,120 triangles, radius 0.45) between the two
    spheres, which the fluid reaches within the first steps.  The plain
    path (analytic box walls) is timed over the same steps beside it."""
    import numpy as np

    from fluid_amd import mesh
    from fluid_amd.pbf import Particles

    p = sc.params
    with Particles(p) as ctx:
        ctx.load(sc.x, sc.v)
        ctx.time_step(nsteps=warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.time_step(nsteps=steps)
        ctx.sync()
        plain_ms = (time.perf_counter() - t0) / steps * 1e3
    m = mesh.merge(mesh.box(p.box_lo, p.box_hi), mesh.icosphere((2.5, 1.5, 1.5), 0.45, 4))
    skin = mesh.default_skin(m[0], p.box_lo, p.box_hi)
    ctx = Particles(p)
    ctx.set_mesh(*m, skin=skin)
    ctx.load(sc.x, sc.v)
    ctx.time_step(nsteps=warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.time_step(nsteps=steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    x, _, _ = ctx.read()
    ctx.close()
    # Particles on the obstacle's surface (within 2 h) and inside its
    # inscribed sphere (the icosphere's faces lie inside the circumsphere).
    r = np.linalg.norm(x.astype(np.float64) - np.array([2.5, 1.5, 1.5]), axis=1)
    on_surface = int((r < 0.45 + 2 * p.h).sum())
    inside = int((r < 0.45 * 0.98).sum())
    ms = wall / steps * 1e3
    log(f"mesh collider ({len(m[2])} triangles): {steps} steps, {ms:.3f} ms/step "
        f"(box walls {plain_ms:.3f}); {on_surface} particles within 2h of the obstacle, "
        f"{inside} inside")
    return {
        "metric": "particle-steps/s with the mesh collider",
        "triangles": int(len(m[2])),
        "obstacle": "icosphere r=0.45 at (2.5, 1.5, 1.5), 5120 triangles, plus the box walls",
        "steps": f"{warmup + 1}-{warmup + steps} from the initial state",
        "ms_per_step": ms,
        "box_walls_ms_per_step": plain_ms,
        "particle_steps_per_s": sc.n * steps / wall,
        "vs_box_walls": plain_ms / ms,
        "particles_near_obstacle": on_surface,
        "particles_inside_obstacle": inside,
    }


def pass_table(prof, sc, steps: int, nbar: float, counters) -> dict:
    """Per pass: ms per step, launches, mean launch time and state-only and
    state+list GB/s from HIP-event times `prof` ({pass: (ms, launches)} over
    `steps` steps); HBM GB/s, VALU fraction, LDS bank-conflict rate and TA
    busy from the committed rocprofv3 counters of the workload, if any."""
    ncell = 1
    for a in range(3):
        ncell *= int((sc.params.box_hi[a] - sc.params.box_lo[a]) / (sc.params.h * sc.params.cell_scale)) + 1
    passes = {}
    for name, (ms, launches) in prof.items():
        if launches == 0:
            continue
        per_launch_s = ms / launches * 1e-3
        state = STATE_BYTES[name] * sc.n + (8 * ncell if name == "scan" else 0)
        listb = (4.0 * nbar + 4.0) * sc.n if name in LIST_PASSES else 0.0
        pd = {
            "ms_per_step": ms / steps,
            "launches_per_step": launches / steps,
            "us_per_launch": per_launch_s * 1e6,
            "state_GBs": state / per_launch_s / 1e9,
            "state_list_GBs": (state + listb) / per_launch_s / 1e9,
        }
        c = counters.get(name, {})
        if c.get("hbm_bytes"):
            pd["hbm_GBs"] = c["hbm_bytes"] / per_launch_s / 1e9
        if c.get("valu_insts"):
            pd["valu_frac"] = c["valu_insts"] * 64 / per_launch_s / VALU_PEAK_LANE_OPS
        if c.get("lds_idx_active"):
            pd["lds_bank_conflict_rate"] = c.get("lds_bank_conflict", 0.0) / c["lds_idx_active"]
        elif "lds_idx_active" in c:
            pd["lds_bank_conflict_rate"] = None  # the pass uses no LDS
        if c.get("ta_busy_frac") is not None:
            # Share of the launch with gathers in flight in each CU's
            # address unit (TA_TA_BUSY_sum / CUs / GRBM_GUI_ACTIVE per XCD).
            pd["ta_busy_frac"] = c["ta_busy_frac"]
        passes[name] = pd
    return passes


def dam_break_bench(warmup: int):
    """The 262,144-particle dam break (64^3 lattice at spacing 0.025, h 0.05,
    4 iterations, dt 10 ms): its whole 200-step run from the initial state,
    then the same steps with HIP events around every launch, in windows of
    50 steps, with the per-pass table of the headline scene."""
    from fluid_amd import scenes
    from fluid_amd.pbf import Particles

    sc = scenes.dam_break()
    steps = sc.steps
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=warmup)
        c.load(sc.x, sc.v)
        c.sync()
        t0 = time.perf_counter()
        c.time_step(nsteps=steps)
        c.sync()
        wall = time.perf_counter() - t0
        c.load(sc.x, sc.v)
        prof, windows = {}, []
        for a in range(0, steps, 50):
            t0 = time.perf_counter()
            pw = c.time_step_profiled(nsteps=50)  # returns after the device finished
            ww = time.perf_counter() - t0
            wbar, wmax = c.neighbor_stats()
            windows.append({
                "steps": f"{a + 1}-{a + 50}",
                "ms_per_step": ww / 50 * 1e3,
                "pass_ms_per_step": {k: ms / 50 for k, (ms, n) in pw.items() if n},
                "mean_neighbors_last_step": wbar,
                "max_neighbors_last_step": wmax,
            })
            for k, (ms, n) in pw.items():
                m0, n0 = prof.get(k, (0.0, 0))
                prof[k] = (m0 + ms, n0 + n)
        nbar, nmax = c.neighbor_stats()
    counters = (load_counters(sc.name) or {}).get("passes", {})
    res = {
        "workload": sc.name,
        "particles": sc.n,
        "steps": steps,
        "ms_per_step": wall / steps * 1e3,
        "particle_steps_per_s": sc.n * steps / wall,
        "mean_neighbors": nbar,
        "max_neighbors": nmax,
        "windows": windows,
        "passes": pass_table(prof, sc, steps, nbar, counters),
    }
    log(f"dam break {sc.n}: {steps} steps at {res['ms_per_step']:.3f} ms/step, neighbours "
        f"mean {nbar:.1f} max {nmax} at the end")
    return res


def small_scenes_bench(warmup: int = 10, steps: int = 300):
    """The reference's own scenes, particles/p.xml (720 particles) and
    particles/spheres_p.xml (2,106), loaded as its -p path does, at its
    native constants (h 0.3, 12 iterations, dt 0.016, the Cornell box):
    ms per step of `steps` steps after `warmup`, per schedule."""
    from fluid_amd import params as P
    from fluid_amd import xmlio
    from fluid_amd.pbf import Particles

    out = {}
    for name in ("p.xml", "spheres_p.xml"):
        rho0, pos, vel = xmlio.load_particles(os.path.join(ROOT, "tests", "golden", name))
        p = P.reference_params().replace(rest_density=rho0)
        res = {"particles": len(pos), "steps": steps, "iterations": p.iterations}
        for schedule in ("flat", "graph"):
            with Particles(p) as c:
                c.set_schedule(schedule)
                c.load(pos, vel)
                c.time_step(nsteps=warmup)
                c.sync()
                t0 = time.perf_counter()
                c.time_step(nsteps=steps)
                c.sync()
                w = time.perf_counter() - t0
            res[schedule + "_ms_per_step"] = w / steps * 1e3
        # The reference's own use: one timeStep() per frame, each call
        # synchronous (the default schedule replays a captured whole step
        # once dt repeats).
        with Particles(p) as c:
            c.load(pos, vel)
            for _ in range(warmup):
                c.time_step()
            t0 = time.perf_counter()
            for _ in range(steps):
                c.time_step()
            res["one_step_calls_ms_per_step"] = (time.perf_counter() - t0) / steps * 1e3
        res["particle_steps_per_s"] = len(pos) * steps / (min(res["flat_ms_per_step"],
                                                               res["graph_ms_per_step"]) * 1e-3 * steps)
        log(f"{name}: {len(pos)} particles, flat {res['flat_ms_per_step']:.4f} ms/step, "
            f"graph {res['graph_ms_per_step']:.4f} ms/step, one-step calls "
            f"{res['one_step_calls_ms_per_step']:.4f} ms/step")
        out[name] = res
    return out


def adjoint_bench():
    """32^3 block, 60 steps, checkpoints every 10: wall time of the backward
    pass (recompute included) per step, and of the forward pass."""
    import numpy as np

    from fluid_amd import scenes
    from fluid_amd.adjoint import Rollout

    sc, target = scenes.diff_block()
    r = Rollout(sc.params, sc.n, sc.steps, checkpoint_every=10)
    r.forward(sc.x, sc.v, sc.params.k_tensile)  # warm-up
    t0 = time.perf_counter()
    xT, _ = r.forward(sc.x, sc.v, sc.params.k_tensile)
    fwd = time.perf_counter() - t0
    xbar = ((2.0 / sc.n) * (xT - target)).astype(np.float32)
    t0 = time.perf_counter()
    _, vb0, kb = r.backward(xbar)
    bwd = time.perf_counter() - t0
    ev_ms, contacts = r.stats()
    r.close()
    log(f"adjoint {sc.name}: forward {fwd * 1e3:.1f} ms, backward {bwd * 1e3:.1f} ms "
        f"({bwd / sc.steps * 1e3:.3f} ms/step), |dL/dv0| {np.linalg.norm(vb0):.3e}, dL/dk {kb:.4e}")
    return {
        "metric": "adjoint ms/step",
        "workload": sc.name,
        "particles": sc.n,
        "steps": sc.steps,
        "checkpoint_every": 10,
        "ms_per_step": bwd / sc.steps * 1e3,
        "ms_per_step_hip_events": ev_ms / sc.steps,
        "forward_ms_per_step": fwd / sc.steps * 1e3,
        "wall_contacts": contacts,
        "grad_norm_v0": float(np.linalg.norm(vb0)),
        "dL_dk": kb,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--no-dam-break", action="store_true")
    ap.add_argument("--only-dam-break", action="store_true",
                    help="run only the 262,144-particle dam break and print its result")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the 512-scene batch against the looped plain path")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run only the 32,768-particle, 60-step adjoint rollout")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--looped-steps", type=int, default=10)
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is not part of this build", file=sys.stderr)
        return 2
    if args.calibrate:
        import torch

        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 256 MiB read + 256 MiB write")

    if args.only_dam_break:
        print(json.dumps({"dam_break": dam_break_bench(args.warmup)}), flush=True)
        return 0
    if args.only_batched:
        print(json.dumps({"batched": batched_bench(args.batched_steps, args.looped_steps)}),
              flush=True)
        return 0
    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_bench()}), flush=True)
        return 0

    from fluid_amd import scenes
    from fluid_amd.pbf import Particles

    t0 = time.perf_counter()
    sc = scenes.two_spheres()
    log(f"scene {sc.name}: {sc.n} particles, h={sc.params.h}, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx = Particles(sc.params)
    ctx.load(sc.x, sc.v)
    ctx.time_step(nsteps=args.warmup)
    ctx.sync()
    log(f"warmup {args.warmup} steps done")
    # The timed steps run from the initial state: steps 1..K of the scene.
    ctx.load(sc.x, sc.v)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.time_step(nsteps=args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    value = sc.n * args.steps / wall
    log(f"timed steps 1-{args.steps}: {wall:.3f} s, {value:.3e} particle-steps/s")
    # Per-pass breakdown: the same steps again from the initial state with
    # HIP events recorded around every launch on the context's stream, in two
    # windows (before and through the impact).
    ctx.load(sc.x, sc.v)
    half = args.steps // 2 if args.steps > 1 else args.steps
    bounds = [(0, half), (half, args.steps)] if args.steps > 1 else [(0, args.steps)]
    prof = {}
    windows = []
    wall_profiled = 0.0
    for a, b in bounds:
        t0 = time.perf_counter()
        pw = ctx.time_step_profiled(nsteps=b - a)  # returns after the device finished
        ww = time.perf_counter() - t0
        wall_profiled += ww
        wbar, wmax = ctx.neighbor_stats()
        windows.append({
            "steps": f"{a + 1}-{b}",
            "ms_per_step": ww / (b - a) * 1e3,
            "pass_ms_per_step": {k: ms / (b - a) for k, (ms, n) in pw.items() if n},
            "mean_neighbors_last_step": wbar,
            "max_neighbors_last_step": wmax,
        })
        for k, (ms, n) in pw.items():
            m0, n0 = prof.get(k, (0.0, 0))
            prof[k] = (m0 + ms, n0 + n)
        log(f"profiled steps {a + 1}-{b}: {ww / (b - a) * 1e3:.3f} ms/step, neighbours "
            f"mean {wbar:.2f} max {wmax}")
    nbar, nmax = ctx.neighbor_stats()
    ctx.read()  # first call allocates the pinned staging slabs
    t0 = time.perf_counter()
    ctx.read()
    read_s = time.perf_counter() - t0
    readback = {"metric": "pbf_read of x, v, rho (load order) into new host arrays",
                "bytes": 28 * sc.n, "ms": read_s * 1e3, "GBs": 28 * sc.n / read_s / 1e9}
    log(f"read back x, v, rho: {read_s * 1e3:.1f} ms")
    ctx.estimate_densities()
    t0 = time.perf_counter()
    rho = ctx.estimate_densities()
    est_s = time.perf_counter() - t0
    densities = {"metric": "estimateDensities (load-time densities, self included) into a new "
                           "host array", "particles": sc.n, "ms": est_s * 1e3,
                 "mean_density": float(rho.mean())}
    log(f"estimate_densities: {est_s * 1e3:.1f} ms")
    field = density_field_bench(ctx, sc)
    snap = snapshot_bench(ctx, sc)
    ctx.close()
    counters = (load_counters(sc.name) or {}).get("passes", {})
    passes = pass_table(prof, sc, args.steps, nbar, counters)
    dom = max(passes, key=lambda k: passes[k]["ms_per_step"])
    d = passes[dom]
    per_launch_bytes = STATE_BYTES[dom] * sc.n
    cp = counters.get(dom, {})
    roof = {
        # The passes are bound by VALU issue (and their neighbour gathers),
        # not by HBM: achieved/peak/frac below are the state-only HBM figures
        # the metric is defined on, `valu` the roof that binds.
        "bound": "valu",
        "kernel": dom,
        "achieved": d["state_GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": d["state_GBs"] / HBM_PEAK_GBS,
        "traffic": cp.get("hbm_bytes"),
        "bytes_per_launch": per_launch_bytes,
        "basis": f"state-only {STATE_BYTES[dom]} B/particle x {sc.n} particles per launch",
        "state_list_frac": d["state_list_GBs"] / HBM_PEAK_GBS,
        "us_per_launch": d["us_per_launch"],
    }
    if cp.get("valu_insts"):
        lane_ops = cp["valu_insts"] * 64 / (d["us_per_launch"] * 1e-6)
        roof["valu"] = {
            "achieved": lane_ops,
            "peak": VALU_PEAK_LANE_OPS,
            "unit": "lane-ops/s",
            "frac": lane_ops / VALU_PEAK_LANE_OPS,
            "valu_insts_per_launch": cp["valu_insts"],
        }
    out = {
        "metric": "PBF particle-steps/s @4M",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": sc.name,
            "particles": sc.n,
            "h": sc.params.h,
            "iterations": sc.params.iterations,
            "dt": sc.params.dt,
            "constants": "scaled",
            "timed_steps": f"1-{args.steps} from the initial state",
            "mean_neighbors": nbar,
            "max_neighbors": nmax,
        },
        "roofline": roof,
        "ms_per_step_profiled": wall_profiled / args.steps * 1e3,
        "windows": windows,
        "passes": passes,
        "readback": readback,
        "estimate_densities": densities,
        "density_field": field,
        "snapshot": snap,
    }
    refd = os.path.join(ROOT, "profiles", "reference_distance.json")
    if os.path.exists(refd):
        with open(refd) as f:
            out["reference_distance"] = json.load(f)
    if not args.no_mesh:
        out["mesh"] = mesh_bench(sc, args.warmup, 10)
    if not args.no_dam_break:
        out["dam_break"] = dam_break_bench(args.warmup)
    if not args.no_small:
        out["small_scenes"] = small_scenes_bench()
    if not args.no_batched:
        out["batched"] = batched_bench(args.batched_steps, args.looped_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
        out["cpu_baseline"]["workloads"] = cpu_baseline_workloads(out.get("dam_break"))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
