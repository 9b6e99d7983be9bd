"""Distance of the GPU step (FP32, grid neighbour order, Jacobi XSPH) to
the reference's own step semantics: FP64, all-pairs neighbour search in
ascending j, and the in-place sequential XSPH/vorticity sweep
(src/particles.cpp:250-297, :285-288), as the oracle's double / brute /
sequential mode; and to the same FP64 all-pairs step with the Jacobi sweep.

Workloads: the 4,096-particle dam block (scene constants) and the
reference's own particles/p.xml (720 particles, its native constants).
Per step: from each state of the GPU's own 100-step trajectory, one GPU step
against one FP64 step from the same state.  Metrics (per step):
positions max_i |dx_i| / max(|x_i|, 1e-6) and |dx|_2 / |x|_2; velocities
|dv|_2 / |v|_2 and max |dv|.  Drift: the same after 100 free-running steps.

Where the 1e-5 per-step bar holds, and where it does not:
* dam block, particles with no coordinate within h of a box face: positions
  within 2.3e-6 (max) on every step, velocities 8.8e-6 norm-wise (median
  over the steps, 3.6e-5 at most): v = (x* - x)/dt divides the FP32
  rounding of x* (~1e-7 m at |x| ~ 0.5) by dt = 0.01.  Asserted: positions
  1e-5, velocities 1e-4.
* particles on the walls: the reference slides a predicted displacement
  along the wall unless it is exactly head-on (dot(d, n) > -1,
  src/particles.cpp:116), with the tangent normalised to unit length; a
  nearly vertical fall onto the floor that rounds to head-on in FP32 does
  not slide, while FP64 slides it up to half the remaining length (2.2 mm
  at step 45 of the dam block).  EPS_D = 1e-11 is also below FP32
  resolution at |x| ~ 1.  Position errors there reach 2e-3 (one particle);
  reported, not asserted.
* sequential against Jacobi XSPH: 2e-4 to 2.6e-3 norm-wise in velocity on
  interior particles of the dam block (1e-3 median): above 1e-5 on every
  step.  Reported.
* p.xml starts on a lattice of spacing h/2, so pairs sit at r = h exactly
  and whether r^2 <= H2 holds depends on the rounding; the list is frozen
  for the step's 12 iterations, so the first step differs by 1e-3 in
  position for every particle.  Once the lattice has broken up (steps 5 to
  15) the distance is 1e-7 (positions) and 2e-6 to 5e-6 (velocities);
  wall contacts (from step ~20) bring it back to 1e-3.  Reported.
* 100-step drift: O(1) (chaotic divergence from the first differing bits).

PBF_REPORT=<path>: the measured figures are also written there as JSON.
"""
import json
import os

import numpy as np
import pytest

from fluid_amd import params as P
from fluid_amd import scenes, xmlio
from fluid_amd.pbf import Particles

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEPS = 100
SAMPLE = 2  # per-step comparison from every 2nd state of the trajectory


def metrics(x, v, xr, vr, mask=None):
    x = x.astype(np.float64)
    v = v.astype(np.float64)
    if mask is not None:
        x, v, xr, vr = x[mask], v[mask], xr[mask], vr[mask]
    dx = np.linalg.norm(x - xr, axis=1)
    return {
        "x_max_rel": float((dx / np.maximum(np.linalg.norm(xr, axis=1), 1e-6)).max()),
        "x_norm": float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
        "v_norm": float(np.linalg.norm(v - vr) / max(np.linalg.norm(vr), 1e-30)),
        "v_max_abs": float(np.abs(v - vr).max()),
    }


def summarise(rows):
    out = {}
    for key in rows[0]:
        vals = np.array([r[key] for r in rows])
        out[key] = {"max": float(vals.max()), "median": float(np.median(vals))}
    return out


def workload(name):
    if name == "dam_block_4096":
        sc = scenes.dam_block()
        return sc.params, sc.x, sc.v
    rho0, pos, vel = xmlio.load_particles(os.path.join(ROOT, "tests", "golden", "p.xml"))
    return (P.reference_params().replace(rest_density=rho0), pos.astype(np.float32),
            vel.astype(np.float32))


@pytest.fixture(scope="module", params=["dam_block_4096", "p_xml_720"])
def measured(request, orc):
    name = request.param
    p, x0, v0 = workload(name)
    lo, hi = np.array(p.box_lo), np.array(p.box_hi)
    rows = {"jacobi_all": [], "jacobi_interior": [], "sequential_all": [],
            "sequential_interior": []}
    with Particles(p) as c:
        c.load(x0, v0)
        x, v = x0.astype(np.float32), v0.astype(np.float32)
        for k in range(STEPS):
            c.time_step(nsteps=1)
            xg, vg, _ = c.read()
            if k % SAMPLE == 0:
                xd, vd = x.astype(np.float64), v.astype(np.float64)
                interior = np.all((x > lo + p.h) & (x < hi - p.h), axis=1)
                for mode, xm in (("jacobi", orc.XSPH_JACOBI), ("sequential", orc.XSPH_SEQUENTIAL)):
                    xr, vr, _ = orc.step(p.as_dict(), xd, vd, 1, orc.NBR_BRUTE, xm,
                                         dtype=np.float64)
                    rows[mode + "_all"].append(metrics(xg, vg, xr, vr))
                    if interior.sum() > 0:
                        rows[mode + "_interior"].append(metrics(xg, vg, xr, vr, interior))
            x, v = xg, vg
    xr, vr, _ = orc.step(p.as_dict(), x0.astype(np.float64), v0.astype(np.float64), STEPS,
                         orc.NBR_BRUTE, orc.XSPH_SEQUENTIAL, dtype=np.float64)
    res = {"particles": int(len(x0)), "steps": STEPS, "sampled_states": STEPS // SAMPLE,
           "per_step": {k: summarise(r) for k, r in rows.items() if r},
           "drift_vs_sequential_fp64": metrics(x, v, xr, vr)}
    print(f"\n{name}: " + json.dumps(res))
    report = os.environ.get("PBF_REPORT")
    if report:
        data = json.load(open(report)) if os.path.exists(report) else {}
        data[name] = res
        with open(report, "w") as f:
            json.dump(data, f, indent=1)
    return name, res


def test_interior_step_within_bar_of_fp64_jacobi(measured):
    name, res = measured
    if name != "dam_block_4096":
        pytest.skip("p.xml: first step at r = h exactly; reported only")
    s = res["per_step"]["jacobi_interior"]
    assert s["x_max_rel"]["max"] <= 1e-5
    assert s["v_norm"]["max"] <= 1e-4
    assert s["v_norm"]["median"] <= 2e-5


def test_sequential_sweep_differs_more_than_rounding(measured):
    """The reference's sequential XSPH sweep against the Jacobi form: the
    difference is a property of the sweep, far above FP32 rounding."""
    name, res = measured
    s = res["per_step"]
    if name == "dam_block_4096":
        assert s["sequential_interior"]["v_norm"]["median"] > 10 * s["jacobi_interior"]["v_norm"]["median"]
    assert np.isfinite(res["drift_vs_sequential_fp64"]["x_norm"])
