"""Edge cases of the step on the GPU, each against a closed form or the
oracle: an isolated particle (semi-implicit Euler with the reference's
velocity update, src/particles.cpp:175-183 and :215-217), pairs at exactly
r = h (the neighbour test is inclusive, :258-265), particle counts around
the wave and tile sizes (1 .. 257), particles starting on the walls, and
the neighbour-list capacity at and one chunk below the largest list.  The
step cases run with each variant of the list passes: a wave, a quad (4
lanes) or a lane per particle, and a lane with the long lists split off to
quads."""
import os

import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd.pbf import Particles, PbfError

pytestmark = pytest.mark.gpu

f32 = np.float32


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# Context environments that select the list-pass variant for these small
# states: the default rule (a wave per particle), quads (lanes for the
# vorticity pass that carries the next prediction), one lane throughout,
# and one lane with the dense split.
MODES = {
    "wave": {},
    "quad": {"PBF_WIDE_MAX": "0"},
    "lane": {"PBF_WIDE_MAX": "0", "PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": "0"},
    # one lane per particle, with every list longer than 8 split off to quads
    "split": {"PBF_WIDE_MAX": "0", "PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": "8"},
}


def context(p, mode="wave"):
    env = MODES[mode]
    os.environ.update(env)
    try:
        return Particles(p)
    finally:
        for k in env:
            os.environ.pop(k, None)


def run(p, x, v, nsteps, mode="wave"):
    with context(p, mode) as c:
        c.load(x, v)
        c.time_step(nsteps=nsteps)
        return c.read()


@pytest.mark.parametrize("mode", list(MODES))
def test_isolated_particle_is_semi_implicit_euler(orc, mode):
    """No neighbours: rho = 0, delta-p = 0, XSPH and confinement add
    nothing, so each step is v.y -= g dt, x* = x + v dt, v = (x* - x) / dt,
    x = x*, in float32 in that order."""
    p = scenes.dam_block().params
    x = np.array([[0.5, 0.9, 0.5]], np.float32)
    v = np.array([[0.3, 0.0, -0.2]], np.float32)
    steps = 20
    gx, gv, grho = run(p, x, v, steps, mode)
    dt = f32(p.dt)
    gdt = f32(p.gravity) * dt
    inv_dt = f32(1.0) / dt
    cx, cv = x[0].copy(), v[0].copy()
    for _ in range(steps):
        cv[1] = cv[1] - gdt
        xs = cx + cv * dt
        cv = inv_dt * (xs - cx)
        cx = xs
    assert bits_equal(gx[0], cx) and bits_equal(gv[0], cv)
    assert grho[0] == 0.0
    ox, ov, _ = orc.step(p, x, v, steps, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov)


@pytest.mark.parametrize("mode", list(MODES))
def test_pairs_at_exactly_h_are_neighbours(orc, mode):
    """h = 1/8 so that h^2 is exact; with no gravity the predicted positions
    equal the loaded ones.  B and D sit exactly h from A (included), C one
    float above h (excluded)."""
    p = scenes.dam_block().params.replace(h=0.125, gravity=0.0)
    x = np.array([
        [0.5, 0.5, 0.5],                                      # A
        [0.625, 0.5, 0.5],                                    # B: dx = h
        [0.5, np.nextafter(f32(0.625), f32(1.0)), 0.5],       # C: dy = h + 1 ulp
        [0.5, 0.5, 0.375],                                    # D: dz = -h
    ], np.float32)
    v = np.zeros_like(x)
    with context(p, mode) as c:
        c.load(x, v)
        c.step_debug("neighbors")
        counts, ids = c.neighbors(stride=8)
    got = [sorted(ids[i, :counts[i]].tolist()) for i in range(4)]
    assert got == [[1, 3], [0], [], [0]]
    og, lg = orc.neighbors(p, x, orc.NBR_GRID)
    for i in range(4):
        assert np.array_equal(ids[i, :counts[i]], lg[og[i]:og[i + 1]])
    gx, gv, grho = run(p, x, v, 3, mode)
    ox, ov, orho = orc.step(p, x, v, 3, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(grho, orho)
    # poly6 and the spiky gradient vanish at r = h: A's density is 0.
    assert grho[0] == 0.0


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 129, 257])
def test_ragged_counts_bit_exact(orc, n, mode):
    """Counts around the 64-lane wave and the 256-particle workgroup (and,
    with quads, the 64-particle quad workgroup)."""
    sc = scenes.dam_block()
    rng = np.random.default_rng(n)
    x = sc.x[rng.choice(len(sc.x), size=n, replace=False)].copy()
    v = rng.uniform(-0.5, 0.5, size=x.shape).astype(np.float32)
    gx, gv, grho = run(sc.params, x, v, 3, mode)
    ox, ov, orho = orc.step(sc.params, x, v, 3, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(grho, orho)


@pytest.mark.parametrize("mode", list(MODES))
def test_particles_on_the_walls_bit_exact(orc, mode):
    """Particles loaded exactly on the faces, edges and corners of the box,
    moving outwards: ray clipping, the slide and the per-axis clamp."""
    sc = scenes.dam_block()
    lo, hi = np.array(sc.params.box_lo, np.float32), np.array(sc.params.box_hi, np.float32)
    rng = np.random.default_rng(7)
    pts, vel = [], []
    for k in range(96):
        q = rng.uniform(lo + 0.05, hi - 0.05).astype(np.float32)
        axes = rng.choice(3, size=1 + k % 3, replace=False)
        w = rng.uniform(-1.0, 1.0, size=3).astype(np.float32)
        for a in axes:
            side = rng.integers(2)
            q[a] = hi[a] if side else lo[a]
            w[a] = abs(w[a]) if side else -abs(w[a])
        pts.append(q)
        vel.append(w)
    x = np.array(pts, np.float32)
    v = np.array(vel, np.float32)
    gx, gv, grho = run(sc.params, x, v, 4, mode)
    ox, ov, orho = orc.step(sc.params, x, v, 4, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(grho, orho)
    assert np.all(gx >= lo) and np.all(gx <= hi)


@pytest.mark.parametrize("mode", list(MODES))
def test_list_capacity_at_and_below_the_largest_list(mode):
    """With each list build (per lane for debug stops and the lane mode, a
    wave or a quad with the first lambda in a step)."""
    sc = scenes.dam_block()
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.step_debug("neighbors")
        nmax = c.neighbor_stats()[1]
    assert nmax > 8
    ref = run(sc.params, sc.x, sc.v, 1)
    # Capacity = max_neighbors rounded up to whole 4-entry chunks.
    got = run(sc.params.replace(max_neighbors=nmax), sc.x, sc.v, 1, mode)
    for a, b in zip(ref, got):
        assert bits_equal(a, b)
    with context(sc.params.replace(max_neighbors=4 * ((nmax - 1) // 4)), mode) as c:
        c.load(sc.x, sc.v)
        with pytest.raises(PbfError) as e:
            c.time_step()
        assert e.value.code == -3


@pytest.mark.parametrize("nan_at,out_at,code,first", [
    (150_000, 170_000, -1, 150_000),
    (170_000, 150_000, -5, 150_000),
    (3, 199_999, -1, 3),
])
def test_load_reports_the_first_bad_particle(nan_at, out_at, code, first):
    """Loads of 2^16 particles or more are checked on several host threads;
    the error is still the one a sequential pass meets first (the lowest bad
    particle), and a failed load leaves the previous state in place."""
    sc = scenes.dam_block()
    p = sc.params
    rng = np.random.default_rng(9)
    lo, hi = np.array(p.box_lo, np.float32), np.array(p.box_hi, np.float32)
    x = rng.uniform(lo, hi, size=(200_000, 3)).astype(np.float32)
    x[nan_at, 1] = np.nan
    x[out_at, 2] = hi[2] + 1.0
    with Particles(p) as c:
        c.load(sc.x, sc.v)
        before = c.read()
        with pytest.raises(PbfError) as e:
            c.load(x, None)
        assert e.value.code == code
        assert f"particle {first} " in str(e.value)
        after = c.read()
    for a, b in zip(before, after):
        assert bits_equal(a, b)


def test_failed_allocation_leaves_the_context_unloaded():
    """A load whose device buffers cannot be allocated (20M particles with
    a 4096-entry list each: 328 GB of lists on a 288 GB device) fails with
    PBF_ERR_NOMEM and leaves the context unloaded: stepping and reading
    then fail with PBF_ERR_STATE instead of touching freed buffers, and a
    later load of a small scene works again."""
    from fluid_amd import params as P

    p = P.reference_params().replace(max_neighbors=4096)
    rng = np.random.default_rng(1)
    small = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.3, 0.9], (500, 3)).astype(np.float32)
    with Particles(p) as c:
        c.load(small)
        c.time_step(nsteps=1)
        n = 20_000_000
        big = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.3, 0.9], (n, 3)).astype(np.float32)
        with pytest.raises(PbfError) as e:
            c.load(big)
        assert e.value.code == -6  # PBF_ERR_NOMEM
        del big
        with pytest.raises(PbfError) as e:
            c.time_step(nsteps=1)
        assert e.value.code == -4  # PBF_ERR_STATE
        with pytest.raises(PbfError) as e:
            c.read()
        assert e.value.code == -4
        c.load(small)
        c.time_step(nsteps=1)
        x, _, _ = c.read()
        assert x.shape == (500, 3) and np.isfinite(x).all()


def test_contexts_on_concurrent_host_threads():
    """Contexts share no device state: two of them (a 4,096-particle block on
    the wave-per-particle passes and a 40,000-particle block on the
    lane-per-particle passes) stepped at the same time from two host
    threads, each 20 calls of 2 steps with a read back after each, give the
    bits of the same calls made one context after the other.  An error on
    one thread (a bad load) does not show up in the other's last error."""
    import threading

    from fluid_amd import params as P

    small = scenes.dam_block()
    p_big = P.scene_params(0.1, 0.05, (0.0, 0.0, 0.0), (3.0, 3.0, 3.0), dt=0.01)
    rng = np.random.default_rng(31)
    xb = scenes.lattice_block(40, 40, 25, 0.05, (0.1, 0.1, 0.1)).astype(np.float32)
    xb += rng.uniform(-1e-3, 1e-3, xb.shape).astype(np.float32)
    vb = rng.uniform(-0.5, 0.5, xb.shape).astype(np.float32)
    jobs = [(small.params, small.x, small.v), (p_big, xb, vb)]

    def work(p, x, v, out, idx, bad):
        with Particles(p) as c:
            c.load(x, v)
            states = []
            for k in range(20):
                c.time_step(nsteps=2)
                states.append(c.read())
                if bad and k == 10:
                    with pytest.raises(PbfError):
                        c.load(np.full((4, 3), np.nan, np.float32), None)
                    c.load(*states[-1][:2])  # back to the state just read
            out[idx] = states

    seq = [None, None]
    for i, (p, x, v) in enumerate(jobs):
        work(p, x, v, seq, i, False)
    par = [None, None]
    threads = [threading.Thread(target=work, args=(p, x, v, par, i, i == 0))
               for i, (p, x, v) in enumerate(jobs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i in range(2):
        assert par[i] is not None
        for k in range(10):  # before the reload on thread 0 (bits of every call)
            for a, b in zip(par[i][k], seq[i][k]):
                assert bits_equal(a, b), (i, k)
    for k in range(10, 20):  # thread 1 never saw thread 0's error
        for a, b in zip(par[1][k], seq[1][k]):
            assert bits_equal(a, b), (1, k)
