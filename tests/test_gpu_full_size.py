"""Parity at the benchmark's full sizes: the 4,096,000-particle two-sphere
collision, the 262,144-particle dam break and the 512-scene batch, against
the oracle (bit-exact, grid/Jacobi) and through size-independent properties
(determinism across runs and schedules, the batch equal to the plain path
scene by scene over 10 steps: the batch runs the per-lane passes, the
scenes alone the wave-per-particle passes of small states)."""
import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd.pbf import Particles

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def two_spheres():
    return scenes.two_spheres()


def test_two_spheres_4m_step_bit_exact(orc, two_spheres):
    sc = two_spheres
    assert sc.n == 4_096_000
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=1)
        gx, gv, grho = c.read()
        nbar, nmax = c.neighbor_stats()
    ox, ov, orho = orc.step(sc.params, sc.x, sc.v, 1, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(grho, orho)
    assert 25 < nbar < 40 and nmax <= sc.params.max_neighbors


def test_two_spheres_4m_through_the_impact_bit_exact(orc, two_spheres):
    """The compressed phase of the headline scene: the sphere surfaces start
    1.0 m apart and close at 4 m/s, so they meet near step 50 (dt 5 ms).
    55 steps on the GPU, then 6 more on the GPU and in the oracle from the
    state read back: the same bits, with the spheres in contact (particles
    within h of the midplane x = 2.5, more neighbours than at rest)."""
    sc = two_spheres
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=55)
        x55, v55, _ = c.read()
        nbar0, nmax0 = c.neighbor_stats()
        c.time_step(nsteps=6)
        gx, gv, grho = c.read()
        nbar, nmax = c.neighbor_stats()
    mid = np.abs(x55[:, 0] - 2.5) < sc.params.h
    left = np.arange(sc.n) < sc.n // 2
    assert (mid & left).any() and (mid & ~left).any()
    print(f"\nimpact: step 55 mean/max neighbours {nbar0:.2f}/{nmax0}, step 61 {nbar:.2f}/{nmax}; "
          f"{int(mid.sum())} particles within h of the midplane")
    ox, ov, orho = orc.step(sc.params, x55, v55, 6, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(grho, orho)


def test_two_spheres_4m_deterministic_across_runs_and_schedules(two_spheres):
    sc = two_spheres
    outs = []
    for schedule in ("flat", "flat", "graph"):
        with Particles(sc.params) as c:
            c.set_schedule(schedule)
            c.load(sc.x, sc.v)
            c.time_step(nsteps=3)
            outs.append(c.read())
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert bits_equal(a, b)


def test_dam_break_262k_steps_bit_exact(orc):
    """20 of the dam break's 200 steps, bit-exact, then its last 5 from the
    state after step 195 (read back from the GPU), when the column has
    spread along the floor and compressed (mean ~96 neighbours, lists of up
    to ~900)."""
    sc = scenes.dam_break()
    assert sc.n == 262_144
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=20)
        gx, gv, _ = c.read()
        c.time_step(nsteps=175)
        x195, v195, _ = c.read()
        c.time_step(nsteps=5)
        hx, hv, _ = c.read()
        nbar, nmax = c.neighbor_stats()
    assert nbar > 80 and nmax > 512
    ox, ov, _ = orc.step(sc.params, sc.x, sc.v, 20, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov)
    ox, ov, _ = orc.step(sc.params, x195, v195, 5, orc.NBR_GRID)
    assert bits_equal(hx, ox) and bits_equal(hv, ov)


def test_dam_break_schedules_with_dense_splits_agree():
    """The dam break's compressed steps run the dense splits (queues filled
    and drained inside the step's own kernels); all 200 steps in one flat
    call, with the graph schedule, and as 200 one-step calls (the captured
    whole step replayed) give the same bits."""
    sc = scenes.dam_break()
    outs = []
    for mode in ("flat", "graph", "calls"):
        with Particles(sc.params) as c:
            if mode == "graph":
                c.set_schedule("graph")
            c.load(sc.x, sc.v)
            if mode == "calls":
                for _ in range(200):
                    c.time_step()
            else:
                c.time_step(nsteps=200)
            outs.append(c.read())
            if mode == "flat":
                assert c.neighbor_stats()[1] > 512  # lists the split gives to quads
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert bits_equal(a, b)


@pytest.mark.parametrize("quad_max", [None, "0"])
def test_quad_and_lane_passes_100_steps_bit_exact(orc, quad_max):
    """100 steps of a 40,000-particle block with random velocities (walls
    reached on every side) in one call, with the default rule (a quad per
    particle in lambda, delta-p and XSPH, a lane in the build and the
    vorticity pass) and with one lane throughout (PBF_QUAD_MAX=0): the
    oracle's bits, so the 100-step drift of both paths is 0 too."""
    import os

    from fluid_amd import params as P

    p = P.scene_params(0.1, 0.05, (0.0, 0.0, 0.0), (3.0, 3.0, 3.0), dt=0.01)
    rng = np.random.default_rng(21)
    x = scenes.lattice_block(40, 40, 25, 0.05, (0.1, 0.1, 0.1)).astype(np.float32)
    x += rng.uniform(-1e-3, 1e-3, x.shape).astype(np.float32)
    v = rng.uniform(-1.0, 1.0, x.shape).astype(np.float32)
    if quad_max is not None:
        os.environ["PBF_QUAD_MAX"] = quad_max
    try:
        with Particles(p) as c:
            c.load(x, v)
            c.time_step(nsteps=100)
            gx, gv, gr = c.read()
    finally:
        os.environ.pop("PBF_QUAD_MAX", None)
    ox, ov, orho = orc.step(p, x, v, 100, orc.NBR_GRID)
    assert bits_equal(gx, ox) and bits_equal(gv, ov) and bits_equal(gr, orho)
    lo = p.box_lo[0] + 0.05
    assert (gx.min(axis=0) < lo).all()  # particles at the walls


def test_batch_512_scenes_equal_plain_path():
    items, p = scenes.batched()
    assert len(items) == 512
    spec = [(len(x), lo, hi) for x, _, lo, hi in items]
    with Particles(p, scenes=spec) as b:
        b.load(np.concatenate([it[0] for it in items]), np.concatenate([it[1] for it in items]))
        b.time_step(nsteps=10)
        bx, bv, _ = b.read()
    off = 0
    for x, v, lo, hi in items:
        with Particles(p.replace(box_lo=lo, box_hi=hi)) as c:
            c.load(x, v)
            c.time_step(nsteps=10)
            gx, gv, _ = c.read()
        n = len(x)
        assert bits_equal(bx[off:off + n], gx) and bits_equal(bv[off:off + n], gv)
        off += n
