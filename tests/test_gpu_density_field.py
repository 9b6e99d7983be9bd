"""GPU density field for the marching-cubes consumers
(Particles::estimateDensityAt, src/particles.cpp:446-453) through the C ABI,
against the oracle's grid restatement (bit-exact: the same 27 cells in the
same (cell, id) order) and the reference's all-particle order (within the
1e-5 bar)."""
import numpy as np
import pytest

from fluid_amd import params as P
from fluid_amd import scenes
from fluid_amd.pbf import Particles, PbfError

pytestmark = pytest.mark.gpu

TOL = 1e-5


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def stepped(orc):
    """Dam block after 20 steps, as the GPU holds it (the committed state
    is what the field is evaluated on)."""
    sc = scenes.dam_block()
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=20)
        x, v, _ = c.read()
    return sc.params, x, v


def queries(x, lo, hi, m, seed=5):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    pad = 0.1 * (hi - lo)
    q = rng.uniform(lo - pad, hi + pad, size=(m, 3))
    return np.concatenate([q, x[:: max(1, len(x) // 500)]]).astype(np.float32)


def test_density_at_bit_exact_with_oracle_grid(orc, stepped):
    p, x, v = stepped
    q = queries(x, p.box_lo, p.box_hi, 3000)
    with Particles(p) as c:
        c.load(x, v)
        got = c.density_at(q)
        gx, gv, _ = c.read()
    assert bits_equal(gx, x) and bits_equal(gv, v)  # the query re-sorts, never moves
    assert bits_equal(got, orc.density_at(p, x, q, orc.NBR_GRID))
    ref = orc.density_at(p, x, q, orc.NBR_BRUTE, np.float64)
    assert np.abs(got - ref).max() <= TOL * np.abs(ref).max()
    assert (ref == 0).any() and (ref > 0).sum() > 1000


def test_density_lattice_matches_points(orc, stepped):
    """The marching-cubes corner lattice: the same values as explicit points
    at (min(hi, lo + i*step), ...) computed in float."""
    p, x, v = stepped
    lo = np.array([0.0, 0.0, 0.0], np.float32)
    hi = np.array(p.box_hi, np.float32)
    step = np.float32(p.h * 0.5)
    counts = [int((hi[a] - lo[a]) / step) + 2 for a in range(3)]
    axes = [np.minimum(hi[a], lo[a] + np.arange(counts[a], dtype=np.float32) * step)
            for a in range(3)]
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    pts = np.stack([xx, yy, zz], -1).reshape(-1, 3)
    with Particles(p) as c:
        c.load(x, v)
        lat = c.density_lattice(lo, hi, float(step), counts)
        at = c.density_at(pts)
    assert lat.shape == (counts[2], counts[1], counts[0])
    assert bits_equal(lat.reshape(-1), at)
    assert bits_equal(lat.reshape(-1), orc.density_at(p, x, pts, orc.NBR_GRID))


def test_density_lattice_across_copy_slabs_and_buffer_reuse(stepped):
    """A lattice of 17.6M points returns through two pinned 64 MB slabs; a
    smaller lattice before and after it (the device buffer grows, then is
    reused) gives the same bits both times, and sampled corners of the large
    one equal explicit points."""
    p, x, v = stepped
    lo = np.array([0.0, 0.0, 0.0], np.float32)
    hi = np.array(p.box_hi, np.float32)
    counts = [260, 260, 260]
    step = np.float32(float(hi.min()) / 300.0)
    with Particles(p) as c:
        c.load(x, v)
        small = c.density_lattice(lo, hi, float(p.h * 0.5), [8, 9, 10])
        big = c.density_lattice(lo, hi, float(step), counts)
        again = c.density_lattice(lo, hi, float(p.h * 0.5), [8, 9, 10])
        rng = np.random.default_rng(3)
        idx = rng.integers(0, [counts[2], counts[1], counts[0]], size=(2000, 3))
        pts = np.stack([np.minimum(hi[a], lo[a] + idx[:, 2 - a].astype(np.float32) * step)
                        for a in range(3)], -1).astype(np.float32)
        at = c.density_at(pts)
    assert bits_equal(small, again)
    assert big.shape == (260, 260, 260)
    assert bits_equal(big[idx[:, 0], idx[:, 1], idx[:, 2]], at)
    assert np.count_nonzero(big) > 1000


def test_density_at_particles_is_load_time_density(stepped):
    p, x, v = stepped
    with Particles(p) as c:
        c.load(x, v)
        rho = c.estimate_densities()
        at = c.density_at(x)
    assert bits_equal(at, rho)


def test_stepping_after_density_query_is_unchanged(stepped):
    p, x, v = stepped
    outs = []
    for query in (False, True):
        with Particles(p) as c:
            c.load(x, v)
            if query:
                c.density_at(x[:100])
            c.time_step(nsteps=3)
            outs.append(c.read())
    for a, b in zip(*outs):
        assert bits_equal(a, b)


def test_density_at_one_scene_of_a_batch(orc):
    items, p = scenes.batched(nscenes=4, nmin=500, nmax=2000)
    spec = [(len(x), lo, hi) for x, _, lo, hi in items]
    with Particles(p, scenes=spec) as b:
        b.load(np.concatenate([it[0] for it in items]), np.concatenate([it[1] for it in items]))
        for s, (x, _, lo, hi) in enumerate(items):
            q = queries(x, lo, hi, 500, seed=s)
            got = b.density_at(q, scene=s)
            assert bits_equal(got, orc.density_at(p.replace(box_lo=lo, box_hi=hi), x, q,
                                                  orc.NBR_GRID))
        with pytest.raises(PbfError):
            b.density_at(q, scene=len(items))


def test_density_field_argument_errors(stepped):
    p, x, v = stepped
    with Particles(p) as c:
        c.load(x, v)
        with pytest.raises(PbfError):
            c.density_at(np.array([[np.nan, 0.0, 0.0]], np.float32))
        with pytest.raises(PbfError):
            c.density_lattice([0, 0, 0], [1, 1, 1], 0.0, [2, 2, 2])
        with pytest.raises(PbfError):
            c.density_lattice([0, 0, 0], [1, 1, 1], 0.1, [0, 2, 2])
        assert c.density_at(np.zeros((0, 3), np.float32)).shape == (0,)
