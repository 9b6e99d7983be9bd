"""Seeded synthetic scenes of the PBF workloads.

Every scene is a cubic lattice (the particle XML fixtures of the reference are
lattices too: particles/pgen.py:25-61) with uniform jitter drawn from
``numpy.random.default_rng(seed)`` (PCG64), lattice first, then jitter, then
a cast to float32.

* ``dam_block``      16^3 block, spacing 0.05, h 0.1, box [0,1]^3, 100 steps.
* ``dam_break``      64^3 block, spacing 0.025, h 0.05, box [0,4.8]x[0,3]x[0,1.6], 200 steps.
* ``two_spheres``    2 x 2,048,000 particles (4,096,000), v_x = +-2, box
  [0,5]x[0,3]x[0,3], dt 0.005, 100 steps.  Spacing 0.010155 and h = 2 x
  spacing, so that two R = 0.8 spheres hold 4.096M lattice points (spacing
  0.0128 would give about 2.05M).
* ``batched``        512 scenes, counts log-uniform in [1000, 32768], blocks or
  spheres at spacing 0.05, h 0.1, boxes U[1,3]^3 (grown when the lattice does
  not fit at half fill), one velocity U[-1,1]^3 per scene.
* ``diff_block``     32^3 block for the adjoint, 60 steps; the target cloud is
  the lattice shifted by (0.5, 0, 0.3) with its own jitter.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .params import Params, scene_params


@dataclasses.dataclass
class Scene:
    name: str
    x: np.ndarray  # (n, 3) float32
    v: np.ndarray  # (n, 3) float32
    params: Params
    steps: int

    @property
    def n(self) -> int:
        return len(self.x)


def _jitter(rng: np.random.Generator, pts: np.ndarray, amp: float) -> np.ndarray:
    return (pts + rng.uniform(-amp, amp, size=pts.shape)).astype(np.float32)


def lattice_block(nx: int, ny: int, nz: int, spacing: float, corner) -> np.ndarray:
    """Lattice points corner + spacing*(i, j, k), x fastest, then y, then z."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    return np.asarray(corner, dtype=np.float64) + spacing * pts


def lattice_sphere(count: int, spacing: float, centre) -> np.ndarray:
    """The `count` points of the lattice centre + spacing*(i,j,k) nearest to
    the centre; ties broken by (k, j, i) so the set is deterministic."""
    r = int(math.ceil((3.0 * count / (4.0 * math.pi)) ** (1.0 / 3.0))) + 2
    ax = np.arange(-r, r + 1)
    k, j, i = np.meshgrid(ax, ax, ax, indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    r2 = i * i + j * j + k * k
    order = np.lexsort((i, j, k, r2))[:count]
    pts = np.stack([i[order], j[order], k[order]], axis=1).astype(np.float64)
    return np.asarray(centre, dtype=np.float64) + spacing * pts


def dam_block(seed: int = 0, steps: int = 100, constants: str = "scaled") -> Scene:
    rng = np.random.default_rng(seed)
    s = 0.05
    x = _jitter(rng, lattice_block(16, 16, 16, s, (0.1, 0.1, 0.1)), 1e-3)
    p = scene_params(0.1, s, (0, 0, 0), (1, 1, 1), dt=0.01, constants=constants)
    return Scene("dam_block_4096", x, np.zeros_like(x), p, steps)


def dam_break(seed: int = 1, steps: int = 200, constants: str = "scaled") -> Scene:
    rng = np.random.default_rng(seed)
    s = 0.025
    x = _jitter(rng, lattice_block(64, 64, 64, s, (s / 2, s / 2, s / 2)), 5e-4)
    # With 4 iterations at dt 10 ms the collapsing column keeps compressing:
    # the mean neighbour count rises from 28 to 98 over the 200 steps and
    # the largest list from 45 to 898 (256 is passed near step 45): the list
    # holds 1,024.
    p = scene_params(0.05, s, (0, 0, 0), (4.8, 3.0, 1.6), dt=0.01, constants=constants,
                     max_neighbors=1024)
    return Scene("dam_break_262144", x, np.zeros_like(x), p, steps)


TWO_SPHERE_SPACING = 0.010155
TWO_SPHERE_PER = 2_048_000


def two_spheres(seed: int = 2, steps: int = 100, per_sphere: int = TWO_SPHERE_PER,
                spacing: float = TWO_SPHERE_SPACING, constants: str = "scaled") -> Scene:
    rng = np.random.default_rng(seed)
    a = lattice_sphere(per_sphere, spacing, (1.2, 1.5, 1.5))
    b = lattice_sphere(per_sphere, spacing, (3.8, 1.5, 1.5))
    x = _jitter(rng, np.concatenate([a, b]), 2e-4 * spacing / 0.0128)
    v = np.zeros_like(x)
    v[:per_sphere, 0] = 2.0
    v[per_sphere:, 0] = -2.0
    # The spheres meet near step 47 and reach the floor near step 75; with
    # 4 iterations at dt 5 ms the floor impact compresses the fluid to ~20x
    # rest density, and particles reach 367 neighbours by step 97 (30 at
    # rest): the list holds 512.
    p = scene_params(2 * spacing, spacing, (0, 0, 0), (5, 3, 3), dt=0.005, constants=constants,
                     max_neighbors=512)
    return Scene(f"two_spheres_{len(x)}", x, v, p, steps)


def diff_block(seed: int = 4, n_side: int = 32, steps: int = 60, constants: str = "scaled"):
    """Adjoint workload: block, target cloud, box large enough that no
    particle reaches a wall."""
    rng = np.random.default_rng(seed)
    s = 0.05
    corner = (-0.5 * s * (n_side - 1), 0.0, -0.5 * s * (n_side - 1))
    base = lattice_block(n_side, n_side, n_side, s, corner)
    x = _jitter(rng, base, 1e-3)
    trng = np.random.default_rng(seed + 1000)
    target = _jitter(trng, base + np.array([0.5, 0.0, 0.3]), 1e-3)
    p = scene_params(0.1, s, (-6, -6, -6), (6, 6, 6), dt=0.01, constants=constants)
    return Scene(f"diff_block_{len(x)}", x, np.zeros_like(x), p, steps), target


def batched(seed: int = 3, nscenes: int = 512, steps: int = 100, nmin: int = 1000,
            nmax: int = 32768, constants: str = "scaled"):
    """List of (x, v, box_lo, box_hi) and the shared parameters."""
    rng = np.random.default_rng(seed)
    s, h = 0.05, 0.1
    margin = 0.5 * s
    counts = np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nscenes)).astype(np.int64)
    out = []
    for c in counts:
        c = int(c)
        shape = "sphere" if rng.uniform() < 0.5 else "block"
        L = rng.uniform(1.0, 3.0, size=3)
        vel = rng.uniform(-1.0, 1.0, size=3)
        # Grow the box until the lattice holds 2c points.
        while True:
            cap = np.floor((L - 2 * margin) / s).astype(int) + 1
            if int(np.prod(cap)) >= 2 * c:
                break
            L = L * 1.1
        lo = np.zeros(3)
        if shape == "block":
            side = int(math.ceil(c ** (1.0 / 3.0)))
            nx, nz = min(cap[0], side), min(cap[2], side)
            ny = int(math.ceil(c / (nx * nz)))
            if ny > cap[1]:
                nx, nz = cap[0], cap[2]
                ny = int(math.ceil(c / (nx * nz)))
            pts = lattice_block(nx, ny, nz, s, lo + margin)
            order = np.lexsort((pts[:, 0], pts[:, 2], pts[:, 1]))[:c]  # y layers from the floor
            pts = pts[order]
        else:
            rad = s * (3.0 * c / (4.0 * math.pi)) ** (1.0 / 3.0) + s
            need = 2 * (rad + margin)
            if np.any(L < need):
                L = np.maximum(L, need)
            pts = lattice_sphere(c, s, L / 2.0)
        x = _jitter(rng, pts, 1e-3)
        x = np.clip(x, lo.astype(np.float32), L.astype(np.float32))
        v = np.broadcast_to(vel.astype(np.float32), x.shape).copy()
        out.append((x, v, tuple(lo), tuple(float(t) for t in L)))
    p = scene_params(h, s, (0, 0, 0), (1, 1, 1), dt=0.01, constants=constants)
    return out, p


CONFIGS = {0: dam_block, 1: dam_break, 2: two_spheres}
