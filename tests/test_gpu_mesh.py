"""Mesh collider (pbf_set_mesh) through the C ABI against the oracle's
restatement: the BVH traversal on the GPU and the oracle's all-triangle scan
give the same nearest hit, so whole steps agree bit for bit."""
import numpy as np
import pytest

from fluid_amd import mesh, scenes
from fluid_amd import params as P
from fluid_amd.pbf import Particles, PbfError

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture
def orc_mesh(orc):
    yield orc
    orc.set_mesh(None)


def drop_scene():
    """An 8x8x8 block falling onto a sphere inside the unit box."""
    rng = np.random.default_rng(3)
    s = 0.05
    x = scenes.lattice_block(8, 8, 8, s, (0.3, 0.6, 0.3))
    x = (x + rng.uniform(-1e-3, 1e-3, x.shape)).astype(np.float32)
    p = P.scene_params(0.1, s, (0, 0, 0), (1, 1, 1), dt=0.01)
    m = mesh.merge(mesh.box((0, 0, 0), (1, 1, 1)), mesh.icosphere((0.5, 0.3, 0.5), 0.2, 2))
    skin = mesh.default_skin(m[0], (0, 0, 0), (1, 1, 1))
    return p, x, np.zeros_like(x), m + (skin,)


def test_prediction_with_mesh_bit_exact(orc_mesh):
    orc = orc_mesh
    p, _, _, m = drop_scene()
    rng = np.random.default_rng(9)
    d = rng.normal(size=(3000, 3))  # directions from the sphere centre
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = (np.array([0.5, 0.3, 0.5]) + d * rng.uniform(0.21, 0.3, (3000, 1))).astype(np.float32)
    x = np.clip(x, 0.01, 0.99).astype(np.float32)
    v = (-d * rng.uniform(1.0, 20.0, (3000, 1))).astype(np.float32)  # towards the sphere
    orc.set_mesh(*m)
    xs, vp = orc.predict(p, x, v)
    with Particles(p) as c:
        c.set_mesh(*m)
        c.load(x, v)
        c.step_debug("predict")
        assert bits_equal(c.debug_read("pred")[:, :3], xs)
        assert bits_equal(c.debug_read("vel")[:, :3], vp)
    orc.set_mesh(None)
    xs0, _ = orc.predict(p, x, v)
    assert (np.abs(xs0 - xs).max(axis=1) > 1e-4).sum() > 1000  # the sphere mattered


def test_falling_block_on_sphere_bit_exact(orc_mesh):
    orc = orc_mesh
    p, x, v, m = drop_scene()
    orc.set_mesh(*m)
    with Particles(p) as c:
        c.set_mesh(*m)
        c.load(x, v)
        ox, ov = x, v
        for _ in range(3):
            c.time_step(nsteps=10)
            ox, ov, _ = orc.step(p, ox, ov, 10, orc.NBR_GRID)
            gx, gv, _ = c.read()
            assert bits_equal(gx, ox) and bits_equal(gv, ov)
    # the block reached the sphere: no particle inside it, some resting on it
    r = np.linalg.norm(gx - np.array([0.5, 0.3, 0.5]), axis=1)
    assert r.min() > 0.18 and (r < 0.25).sum() > 50


def test_falling_slab_on_sphere_quad_and_lane_passes_bit_exact(orc_mesh):
    """A 48x20x48 slab (46,080 particles, above the wave-per-particle range)
    falling onto the sphere: the oracle's bits after 15 and 30 steps, with
    the delta-p pass and its collision calls run with a quad per particle
    (the default rule at this size), in a second context created with
    PBF_QUAD_MAX=0 and PBF_DENSE_LIST=0 with one lane per particle, and in a
    third with one lane and the dense splits at thresholds low enough that
    most particles take the split's quads (PBF_DENSE_LIST=8,
    PBF_DENSE_CELL=2)."""
    import os

    orc = orc_mesh
    rng = np.random.default_rng(5)
    s = 0.02
    x = scenes.lattice_block(48, 20, 48, s, (0.03, 0.54, 0.03))
    x = (x + rng.uniform(-5e-4, 5e-4, x.shape)).astype(np.float32)
    assert len(x) == 46_080 and x.max() < 1.0
    p = P.scene_params(2 * s, s, (0, 0, 0), (1, 1, 1), dt=0.01)
    m = mesh.merge(mesh.box((0, 0, 0), (1, 1, 1)), mesh.icosphere((0.5, 0.3, 0.5), 0.2, 2))
    m = m + (mesh.default_skin(m[0], (0, 0, 0), (1, 1, 1)),)
    orc.set_mesh(*m)
    v = np.zeros_like(x)
    def make(env):
        os.environ.update(env)
        try:
            return Particles(p)
        finally:
            for k in env:
                os.environ.pop(k, None)

    lane = make({"PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": "0"})
    split = make({"PBF_QUAD_MAX": "0", "PBF_DENSE_LIST": "8", "PBF_DENSE_CELL": "2"})
    with Particles(p) as c, lane, split:
        for ctx in (c, lane, split):
            ctx.set_mesh(*m)
            ctx.load(x, v)
        ox, ov = x, v
        for _ in range(2):
            for ctx in (c, lane, split):
                ctx.time_step(nsteps=15)
            ox, ov, _ = orc.step(p, ox, ov, 15, orc.NBR_GRID)
            for ctx in (c, lane, split):
                gx, gv, _ = ctx.read()
                assert bits_equal(gx, ox) and bits_equal(gv, ov)
    r = np.linalg.norm(gx - np.array([0.5, 0.3, 0.5]), axis=1)
    assert r.min() > 0.18 and (r < 0.25).sum() > 100  # on the sphere, none inside it


def test_box_mesh_matches_box_walls_and_removal_restores_them():
    sc = scenes.dam_block()
    p = sc.params
    with Particles(p) as a:
        a.load(sc.x, sc.v)
        a.time_step(nsteps=5)
        ax, av, _ = a.read()
    with Particles(p) as b:
        b.set_mesh(*mesh.box((0, 0, 0), (1, 1, 1)))
        b.load(sc.x, sc.v)
        b.time_step(nsteps=5)
        bx, bv, _ = b.read()
        assert np.abs(bx - ax).max() <= 1e-5 * np.abs(ax).max()
        b.set_mesh(None)
        b.load(sc.x, sc.v)
        b.time_step(nsteps=5)
        cx, cv, _ = b.read()
    assert bits_equal(cx, ax) and bits_equal(cv, av)


@pytest.mark.parametrize("center", [(0.5, 0.3, 0.5), (0.95, 0.3, 0.5)])
def test_mesh_cell_pruning_changes_no_bit(monkeypatch, center):
    """Segments whose box touches no grid cell near a triangle skip the walk;
    with the cells switched off (PBF_MESH_CELLS=0) every segment walks it.
    Random velocities up to 3 m/s (0.3 cells per step) and a few particles
    at 30 m/s (segments over two cells long) hit the sphere and the walls;
    the second sphere reaches out of the box (its cells clamp to the grid)."""
    p, x, _, m = drop_scene()
    if center != (0.5, 0.3, 0.5):
        m = mesh.merge(mesh.box((0, 0, 0), (1, 1, 1)), mesh.icosphere(center, 0.2, 2)) + (m[3],)
    rng = np.random.default_rng(11)
    v = rng.uniform(-3.0, 3.0, size=x.shape).astype(np.float32)
    v[::37] *= 10.0

    def run(cells):
        with monkeypatch.context() as mp:
            if not cells:
                mp.setenv("PBF_MESH_CELLS", "0")
            with Particles(p) as c:
                c.set_mesh(*m)
                c.load(x, v)
                out = []
                for _ in range(4):
                    c.time_step(nsteps=10)
                    out.append(c.read())
                return out

    walked, pruned = run(False), run(True)
    for a, b in zip(walked, pruned):
        assert all(bits_equal(u, w) for u, w in zip(a, b))
    r = np.linalg.norm(pruned[-1][0] - np.array(center), axis=1)
    assert r.min() > 0.18


def test_mesh_argument_errors():
    sc = scenes.dam_block()
    pos, nrm, tri = mesh.icosphere((0.5, 0.5, 0.5), 0.1, 1)
    with Particles(sc.params) as c:
        with pytest.raises(PbfError):
            c.set_mesh(pos, nrm, np.array([[0, 1, len(pos)]]))
        bad = pos.copy()
        bad[0, 0] = np.nan
        with pytest.raises(PbfError):
            c.set_mesh(bad, nrm, tri)
    items, p = scenes.batched(nscenes=2, nmin=500, nmax=600)
    with Particles(p, scenes=[(len(x), lo, hi) for x, _, lo, hi in items]) as b:
        with pytest.raises(PbfError, match="single-scene"):
            b.set_mesh(pos, nrm, tri)
