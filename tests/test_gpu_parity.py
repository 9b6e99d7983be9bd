"""Parity of the HIP path (through the C ABI) with the CPU restatement.

Bar: neighbour sets and cell indices bit-exact; per-step float32 positions
and velocities within 1e-5 relative of the oracle.  The HIP path evaluates
the same expressions in the same order as the oracle's grid/Jacobi mode, so
those comparisons are expected to be bit-exact, and are asserted so.  Against
the reference's own neighbour order (brute force, ascending j) the results
differ by summation order only and are held to the 1e-5 bar.
"""
import os

import numpy as np
import pytest

from fluid_amd import params as P
from fluid_amd import scenes, xmlio
from fluid_amd.pbf import Particles, PbfError

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rel_pos(x, ref):
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-6))


def rel_vel(v, ref):
    return float(np.linalg.norm(v - ref) / max(np.linalg.norm(ref), 1e-12))


@pytest.fixture(scope="module")
def wall_state(orc):
    """Dam block after 30 oracle steps: particles rest on the floor and
    walls, so the collision branches are exercised."""
    sc = scenes.dam_block()
    x, v, _ = orc.step(sc.params, sc.x, sc.v, 30)
    return sc.params, x, v


def ctx_for(p, x, v):
    c = Particles(p)
    c.load(x, v)
    return c


def test_derived_constants_bit_exact(orc, wall_state):
    p, x, v = wall_state
    with ctx_for(p, x, v) as c:
        d, dims = orc.derive(p)
        assert bits_equal(c.derived(), d)
        assert c.grid_dims() == dims


def test_predict_stage_bit_exact(orc, wall_state):
    p, x, v = wall_state
    xs, vp = orc.predict(p, x, v)
    with ctx_for(p, x, v) as c:
        c.step_debug("predict")
        assert bits_equal(c.debug_read("pred")[:, :3], xs)
        assert bits_equal(c.debug_read("vel")[:, :3], vp)


def test_cells_and_neighbour_lists_bit_exact(orc, wall_state):
    p, x, v = wall_state
    xs, _ = orc.predict(p, x, v)
    og, lg = orc.neighbors(p, xs, orc.NBR_GRID)
    ob, lb = orc.neighbors(p, xs, orc.NBR_BRUTE)
    with ctx_for(p, x, v) as c:
        c.step_debug("neighbors")
        assert np.array_equal(c.cells(), orc.cells(p, xs))
        counts, ids = c.neighbors(stride=p.max_neighbors)
    assert np.array_equal(counts, np.diff(og))
    for i in range(len(x)):
        got = ids[i, : counts[i]]
        assert np.array_equal(got, lg[og[i]:og[i + 1]])
        assert set(got.tolist()) == set(lb[ob[i]:ob[i + 1]].tolist())


def test_solver_passes_bit_exact(orc, wall_state):
    p, x, v = wall_state
    xs, vp = orc.predict(p, x, v)
    off, lst = orc.neighbors(p, xs, orc.NBR_GRID)
    rho, lam = orc.lambda_pass(p, xs, off, lst)
    xs1 = orc.deltap_pass(p, xs, lam, off, lst)
    with ctx_for(p, x, v) as c:
        c.step_debug("lambda0")
        g = c.debug_read("pred")
        assert bits_equal(g[:, :3], xs) and bits_equal(g[:, 3], lam)
    with ctx_for(p, x, v) as c:
        c.step_debug("deltap0")
        assert bits_equal(c.debug_read("pred")[:, :3], xs1)
    cur = xs
    for _ in range(p.iterations):
        _, lam = orc.lambda_pass(p, cur, off, lst)
        cur = orc.deltap_pass(p, cur, lam, off, lst)
    vout, om, rho = orc.xsph_pass(p, x, cur, vp, off, lst)
    with ctx_for(p, x, v) as c:
        c.step_debug("solve")
        assert bits_equal(c.debug_read("pred")[:, :3], cur)
        assert bits_equal(c.debug_read("vel")[:, :3], (cur - x) * np.float32(1.0 / np.float32(p.dt)))
    with ctx_for(p, x, v) as c:
        c.step_debug("xsph")
        vx = c.debug_read("vxsph")
        assert bits_equal(vx[:, :3], vout) and bits_equal(vx[:, 3], rho)
        assert bits_equal(c.debug_read("omega")[:, :3], om)


def test_full_step_bit_exact_and_within_tolerance_of_reference_order(orc, wall_state):
    p, x, v = wall_state
    xg, vg, rg = orc.step(p, x, v, 1, orc.NBR_GRID)
    xb, vb, _ = orc.step(p, x, v, 1, orc.NBR_BRUTE)
    with ctx_for(p, x, v) as c:
        c.time_step()
        gx, gv, grho = c.read()
    assert bits_equal(gx, xg) and bits_equal(gv, vg) and bits_equal(grho, rg)
    assert rel_pos(gx, xb) <= TOL
    assert rel_vel(gv, vb) <= TOL


def test_hundred_steps_drift(orc):
    """Dam block, 100 free-running steps from identical input: drift of the
    HIP path against the oracle (grid/Jacobi) and against the reference order."""
    sc = scenes.dam_block()
    xg, vg, _ = orc.step(sc.params, sc.x, sc.v, 100, orc.NBR_GRID)
    xb, vb, _ = orc.step(sc.params, sc.x, sc.v, 100, orc.NBR_BRUTE)
    with ctx_for(sc.params, sc.x, sc.v) as c:
        c.time_step(nsteps=100)
        gx, gv, _ = c.read()
        assert c.simulate_time == pytest.approx(1.0, rel=1e-6)
    print(f"\n100-step drift vs oracle: pos {rel_pos(gx, xg):.3e} vel {rel_vel(gv, vg):.3e}; "
          f"vs reference order: pos {rel_pos(gx, xb):.3e} vel {rel_vel(gv, vb):.3e}")
    assert bits_equal(gx, xg) and bits_equal(gv, vg)


def test_each_step_within_tolerance_of_reference_order(orc):
    """Per-step error from identical input state, reference neighbour order."""
    sc = scenes.dam_block()
    x, v = sc.x, sc.v
    worst_x = worst_v = 0.0
    for s in range(20):
        xb, vb, _ = orc.step(sc.params, x, v, 1, orc.NBR_BRUTE)
        with ctx_for(sc.params, x, v) as c:
            c.time_step()
            gx, gv, _ = c.read()
        worst_x = max(worst_x, rel_pos(gx, xb))
        worst_v = max(worst_v, rel_vel(gv, vb))
        x, v = xb, vb
    print(f"\nworst per-step error vs reference order: pos {worst_x:.3e} vel {worst_v:.3e}")
    assert worst_x <= TOL and worst_v <= TOL


def test_reference_fixture_native_constants(orc):
    """particles/p.xml of the reference with its own constants and Cornell box
    (stop planes y = 1.49 and z = 1), one simulated second = 63 steps of
    dt 0.016 as the reference's drivers run it."""
    rho0, pos, vel = xmlio.load_particles(os.path.join(GOLDEN, "p.xml"))
    assert len(pos) == 720 and rho0 == 150.0
    p = P.reference_params().replace(rest_density=rho0)
    x, v = pos.astype(np.float32), vel.astype(np.float32)
    xg, vg, rg = orc.step(p, x, v, 63, orc.NBR_GRID)
    with ctx_for(p, x, v) as c:
        c.time_step(nsteps=63)
        gx, gv, grho = c.read()
    assert bits_equal(gx, xg) and bits_equal(gv, vg) and bits_equal(grho, rg)


def test_load_time_densities(orc):
    rho0, pos, vel = xmlio.load_particles(os.path.join(GOLDEN, "spheres_p.xml"))
    p = P.reference_params().replace(rest_density=rho0)
    x = pos.astype(np.float32)
    with ctx_for(p, x, vel.astype(np.float32)) as c:
        r = c.estimate_densities()
        gx, _, _ = c.read()
    assert bits_equal(gx, x)  # the density pass re-sorts but keeps the state
    assert bits_equal(r, orc.estimate_densities(p, x, orc.NBR_GRID))
    rb = orc.estimate_densities(p, x, orc.NBR_BRUTE)
    assert np.abs(r - rb).max() <= TOL * np.abs(rb).max()


def test_large_grid_multi_block_scan(orc):
    """A 120^3-cell grid takes the multi-tile scan path (over 1024 tiles)."""
    sc, _ = scenes.diff_block(n_side=10)
    p = sc.params
    xs, _ = orc.predict(p, sc.x, sc.v)
    og, lg = orc.neighbors(p, xs, orc.NBR_GRID)
    with ctx_for(p, sc.x, sc.v) as c:
        c.step_debug("neighbors")
        assert np.array_equal(c.cells(), orc.cells(p, xs))
        counts, ids = c.neighbors(stride=p.max_neighbors)
    assert np.array_equal(counts, np.diff(og))
    assert all(np.array_equal(ids[i, : counts[i]], lg[og[i]:og[i + 1]]) for i in range(len(xs)))


def test_small_state_single_workgroup_sort_over_several_scan_chunks(orc):
    """4,096 particles over a 27^3-cell grid: the scan, scatter and rank run
    in one workgroup (the scan in five 4,096-cell chunks); 5 steps bit-exact
    with the oracle, and with the multi-launch path forced in a context
    created under PBF_NO_SMALL_SORT, the same bits."""
    import os

    sc = scenes.dam_block()
    p = sc.params.replace(box_hi=(2.7, 2.7, 2.7))
    xg, vg, _ = orc.step(p, sc.x, sc.v, 5, orc.NBR_GRID)
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["PBF_NO_SMALL_SORT"] = env
        try:
            with ctx_for(p, sc.x, sc.v) as c:
                assert np.prod(c.grid_dims()) == 27 ** 3
                c.time_step(nsteps=5)
                outs.append(c.read())
        finally:
            os.environ.pop("PBF_NO_SMALL_SORT", None)
    for gx, gv, _ in outs:
        assert bits_equal(gx, xg) and bits_equal(gv, vg)


LANE_MODES = (
    {},  # the default rule: a wave per particle for small states
    {"PBF_WIDE_MAX": "0"},  # quads for lambda / delta-p / XSPH, lanes for the rest
    {"PBF_WIDE_MAX": "0", "PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": "0"},  # one lane throughout
)


def run_modes(p, x, v, nsteps, modes=LANE_MODES):
    """(read(), neighbor_stats()) after nsteps steps, per lane mode."""
    outs = []
    for env in modes:
        os.environ.update(env)
        try:
            with ctx_for(p, x, v) as c:
                c.time_step(nsteps=nsteps)
                outs.append((c.read(), c.neighbor_stats()))
        finally:
            for k in env:
                os.environ.pop(k, None)
    return outs


@pytest.mark.parametrize("spacing_over_h", [0.5, 0.3])
def test_wide_quad_and_lane_passes_equal(orc, spacing_over_h):
    """The list passes with one wave, one quad (4 lanes) or one lane per
    particle give the same bits, equal to the oracle's, for the dam block
    (about 31 neighbours) and for a block at spacing 0.3 h (about 150
    neighbours at first, 98 at most after three steps: lists of several
    64-entry blocks and of chunks past the end of a quad's last chunk)."""
    sc = scenes.dam_block()
    h = sc.params.h
    s = spacing_over_h * h
    x = scenes.lattice_block(16, 16, 16, s, (0.1, 0.1, 0.1)).astype(np.float32)
    x += np.random.default_rng(4).uniform(-1e-3 * s, 1e-3 * s, x.shape).astype(np.float32)
    v = np.random.default_rng(5).uniform(-0.2, 0.2, x.shape).astype(np.float32)
    p = sc.params.replace(max_neighbors=512)
    xg, vg, rg = orc.step(p, x, v, 3, orc.NBR_GRID)
    outs = run_modes(p, x, v, 3)
    if spacing_over_h < 0.4:
        assert outs[0][1][1] > 64  # several 64-entry blocks per list
    for (gx, gv, gr), _ in outs:
        assert bits_equal(gx, xg) and bits_equal(gv, vg) and bits_equal(gr, rg)


@pytest.mark.parametrize("dense_list,dense_cell", [("60", "24"), ("120", "48"), ("100000", "24"),
                                                    ("0", "0")])
def test_dense_split_bit_exact(orc, dense_list, dense_cell):
    """One lane per particle with the dense split: the build queues the
    particles with more than PBF_DENSE_LIST neighbours (a block at spacing
    0.3 h holds about 150 at first and 98 at most after three steps, so 60
    queues many particles at every step, 120 only in the first steps,
    100000 none) and the lambda, delta-p and XSPH launches give them to quad
    workgroups; the sort queues the particles of cells holding more than
    PBF_DENSE_CELL (about 37 per cell at this spacing: 24 queues most of
    them, 48 a few) and the build + first lambda launch builds their lists
    on quads; 0, 0 turns both off.  Three steps equal the oracle's."""
    sc = scenes.dam_block()
    h = sc.params.h
    s = 0.3 * h
    x = scenes.lattice_block(16, 16, 16, s, (0.1, 0.1, 0.1)).astype(np.float32)
    x += np.random.default_rng(4).uniform(-1e-3 * s, 1e-3 * s, x.shape).astype(np.float32)
    v = np.random.default_rng(5).uniform(-0.2, 0.2, x.shape).astype(np.float32)
    p = sc.params.replace(max_neighbors=512)
    xg, vg, rg = orc.step(p, x, v, 3, orc.NBR_GRID)
    env = {"PBF_WIDE_MAX": "0", "PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": dense_list,
           "PBF_DENSE_CELL": dense_cell}
    ((gx, gv, gr), (nbar, nmax)), = run_modes(p, x, v, 3, (env,))
    assert nmax > 60
    assert bits_equal(gx, xg) and bits_equal(gv, vg) and bits_equal(gr, rg)


def test_mid_size_state_quad_passes_bit_exact(orc):
    """49,152 particles, where the default rule runs lambda, delta-p and XSPH
    with a quad per particle and the build and vorticity with a lane: five
    steps equal the oracle's, and the one-lane passes give the same bits."""
    s = 0.05
    x = scenes.lattice_block(48, 32, 32, s, (0.1, 0.1, 0.1)).astype(np.float32)
    x += np.random.default_rng(1).uniform(-1e-3, 1e-3, x.shape).astype(np.float32)
    v = np.random.default_rng(2).uniform(-0.5, 0.5, x.shape).astype(np.float32)
    p = P.scene_params(0.1, s, (0, 0, 0), (4.0, 4.0, 3.0), dt=0.01)
    xg, vg, rg = orc.step(p, x, v, 5, orc.NBR_GRID)
    outs = run_modes(p, x, v, 5, (LANE_MODES[0], LANE_MODES[2]))
    for (gx, gv, gr), _ in outs:
        assert bits_equal(gx, xg) and bits_equal(gv, vg) and bits_equal(gr, rg)


def test_two_sphere_collision_bit_exact(orc):
    """A small two-sphere impact (v_x = +-2) with the c2 constants."""
    sc = scenes.two_spheres(per_sphere=3000, spacing=0.04, steps=5)
    sc.params = sc.params.replace(box_hi=(5.0, 3.0, 3.0))
    xg, vg, _ = orc.step(sc.params, sc.x, sc.v, 10, orc.NBR_GRID)
    with ctx_for(sc.params, sc.x, sc.v) as c:
        c.time_step(nsteps=10)
        gx, gv, _ = c.read()
    assert bits_equal(gx, xg) and bits_equal(gv, vg)


def test_short_math_sequences_verified_and_same_bits(orc):
    """The short division/reciprocal/square-root sequences pass the device's
    exhaustive check at create time, and a context that keeps the IEEE
    expansions (PBF_IEEE_MATH) gives the same bits as one that uses them."""
    sc = scenes.two_spheres(per_sphere=3000, spacing=0.04, steps=5)
    sc.params = sc.params.replace(box_hi=(5.0, 3.0, 3.0))
    outs = []
    for ieee in (False, True):
        if ieee:
            os.environ["PBF_IEEE_MATH"] = "1"
        try:
            with ctx_for(sc.params, sc.x, sc.v) as c:
                flags = c.math_flags()
                c.time_step(nsteps=10)
                outs.append(c.read())
        finally:
            os.environ.pop("PBF_IEEE_MATH", None)
        if ieee:
            assert flags == {"div": False, "rcp": False, "sqrt": False}
        else:
            assert flags == {"div": True, "rcp": True, "sqrt": True}, flags
    for a, b in zip(*outs):
        assert bits_equal(a, b)
    xg, vg, _ = orc.step(sc.params, sc.x, sc.v, 10, orc.NBR_GRID)
    assert bits_equal(outs[0][0], xg) and bits_equal(outs[0][1], vg)


def test_deterministic_across_runs(wall_state):
    p, x, v = wall_state
    outs = []
    for _ in range(2):
        with ctx_for(p, x, v) as c:
            c.time_step(nsteps=5)
            outs.append(c.read())
    for a, b in zip(*outs):
        assert bits_equal(a, b)


def test_batched_scenes_equal_plain_runs():
    items, p = scenes.batched(nscenes=6, nmin=500, nmax=3000)
    with Particles(p, scenes=[(len(x), lo, hi) for x, _, lo, hi in items]) as b:
        b.load(np.concatenate([it[0] for it in items]), np.concatenate([it[1] for it in items]))
        b.time_step(nsteps=8)
        bx, bv, brho = b.read()
    off = 0
    for x, v, lo, hi in items:
        with ctx_for(p.replace(box_lo=lo, box_hi=hi), x, v) as c:
            c.time_step(nsteps=8)
            gx, gv, grho = c.read()
        n = len(x)
        assert bits_equal(bx[off:off + n], gx) and bits_equal(bv[off:off + n], gv)
        assert bits_equal(brho[off:off + n], grho)
        off += n


def _batched_items(n=6, nmin=500, nmax=3000):
    items, p = scenes.batched(nscenes=n, nmin=nmin, nmax=nmax)
    spec = [(len(x), lo, hi) for x, _, lo, hi in items]
    x0 = np.concatenate([it[0] for it in items])
    v0 = np.concatenate([it[1] for it in items])
    return p, spec, x0, v0


@pytest.mark.parametrize("schedule,block", [("graph", 256), ("persistent", 256),
                                            ("persistent", 512), ("persistent", 1024)])
def test_schedules_bit_exact_with_flat(schedule, block):
    p, spec, x0, v0 = _batched_items()
    with Particles(p, scenes=spec) as f:
        f.load(x0, v0)
        f.time_step(nsteps=6)
        fx, fv, frho = f.read()
    with Particles(p, scenes=spec) as b:
        b.set_schedule(schedule, block)
        b.load(x0, v0)
        b.time_step(nsteps=6)
        bx, bv, brho = b.read()
        assert b.simulate_time == pytest.approx(6 * p.dt, rel=1e-6)
    assert bits_equal(bx, fx) and bits_equal(bv, fv) and bits_equal(brho, frho)


def test_graph_replays_each_match_eager():
    """Replay the captured steps over several calls (calls of three or more
    steps replay the captured inner step, calls of one step a captured
    whole step, calls of two steps run plain launches) with eager work
    (reads, synchronisations) in between; every call must equal the flat
    schedule's run."""
    p, spec, x0, v0 = _batched_items(4)
    with Particles(p, scenes=spec) as f, Particles(p, scenes=spec) as g:
        f.load(x0, v0)
        g.set_schedule("graph")
        g.load(x0, v0)
        for n in (4, 1, 5, 2, 3, 4):
            f.time_step(nsteps=n)
            g.time_step(nsteps=n)
            a, b = f.read(), g.read()
            for u, w in zip(a, b):
                assert bits_equal(u, w)


def test_graph_one_step_calls_match_flat(orc):
    """The reference's use: one timeStep per frame.  One-step calls replay a
    captured whole step with the graph schedule, and with the flat schedule
    once dt repeats; twelve calls, a read after each, a change of dt
    (recapture) after six and a three-step call between, all equal to plain
    launches (a context created with PBF_STEP_GRAPH=0), and five one-step
    calls equal the oracle's five steps."""
    sc = scenes.two_spheres(per_sphere=3000, spacing=0.04, steps=5)
    p = sc.params.replace(box_hi=(5.0, 3.0, 3.0))
    os.environ["PBF_STEP_GRAPH"] = "0"
    try:
        plain = ctx_for(p, sc.x, sc.v)
    finally:
        os.environ.pop("PBF_STEP_GRAPH", None)
    with plain as f, ctx_for(p, sc.x, sc.v) as g, ctx_for(p, sc.x, sc.v) as d:
        g.set_schedule("graph")
        for k in range(12):
            dt = None if k < 6 else 0.5 * p.dt
            n = 3 if k == 8 else 1
            for c in (f, g, d):
                c.time_step(dt=dt, nsteps=n)
            a = f.read()
            for c in (g, d):
                for u, w in zip(a, c.read()):
                    assert bits_equal(u, w)
    for sched in ("flat", "graph"):
        with ctx_for(p, sc.x, sc.v) as g:
            g.set_schedule(sched)
            for _ in range(5):
                g.time_step()
            gx, gv, _ = g.read()
        xg, vg, _ = orc.step(p, sc.x, sc.v, 5, orc.NBR_GRID)
        assert bits_equal(gx, xg) and bits_equal(gv, vg)


def test_async_steps_and_overflow_reported_by_sync():
    sc = scenes.dam_block()
    with ctx_for(sc.params, sc.x, sc.v) as c:
        c.step_async(3)
        c.step_async(2)
        c.sync()
        gx, gv, _ = c.read()
    with ctx_for(sc.params, sc.x, sc.v) as c:
        c.time_step(nsteps=5)
        rx, rv, _ = c.read()
    assert bits_equal(gx, rx) and bits_equal(gv, rv)
    with ctx_for(sc.params.replace(max_neighbors=4), sc.x, sc.v) as c:
        c.step_async(1)
        with pytest.raises(PbfError) as e:
            c.sync()
        assert e.value.code == -3


def test_error_paths():
    sc = scenes.dam_block()
    with Particles(sc.params) as c:
        with pytest.raises(PbfError) as e:
            c.time_step()
        assert e.value.code == -4  # step before load
        bad = sc.x.copy()
        bad[3, 1] = np.nan
        with pytest.raises(PbfError) as e:
            c.load(bad, sc.v)
        assert e.value.code == -1
        bad = sc.x.copy()
        bad[3, 1] = 1.5
        with pytest.raises(PbfError) as e:
            c.load(bad, sc.v)
        assert e.value.code == -5
        c.load(sc.x, sc.v)
        c.step_debug("lambda0")
        with pytest.raises(PbfError) as e:
            c.time_step()
        assert e.value.code == -4  # debug state must be reloaded
        c.load(sc.x, sc.v)
        with pytest.raises(PbfError):
            c.time_step(dt=0.0)
    with Particles(sc.params.replace(max_neighbors=4)) as c:
        c.load(sc.x, sc.v)
        with pytest.raises(PbfError) as e:
            c.time_step()
        assert e.value.code == -3
