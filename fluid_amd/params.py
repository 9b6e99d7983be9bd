"""Run-time parameters of the PBF step.

The reference fixes every constant with a ``#define`` (src/particles.cpp:7-48);
:func:`reference_params` returns those values, :class:`Params` carries any
other set (the synthetic scenes use h = 2 x lattice spacing, 4 iterations,
epsilon 600, k 0.1, dq 0.2, c 0.01, vorticity 5e-4, g 9.8).
"""
from __future__ import annotations

import dataclasses
import math

EPS_D = 0.00000000001  # CGL/include/CGL/misc.h:11
POLY6_LITERAL = 1.56668147106  # src/particles.cpp:140 (315/(64 pi), truncated)
SPIKY_LITERAL = 4.774648292756860  # src/particles.cpp:148 (15/pi)

# Wall bits: 2*axis + (0 lower, 1 upper).
WALL_Y_HI = 1 << 3
WALL_Z_HI = 1 << 5


@dataclasses.dataclass
class Params:
    h: float = 0.3
    rest_density: float = 1000.0
    iterations: int = 12
    epsilon: float = 2.0
    k_tensile: float = 0.0001
    n_tensile: int = 4
    dq: float = 0.1
    xsph_c: float = 0.001
    vort_eps: float = 0.001
    gravity: float = 10.0
    dt: float = 0.016
    box_lo: tuple = (-1.0, 0.0, -1.0)
    box_hi: tuple = (1.0, 1.49, 1.0)
    wall_stop_mask: int = WALL_Y_HI | WALL_Z_HI
    cell_scale: float = 1.0 + 1.0 / 1024.0
    max_neighbors: int = 256
    poly6_coef: float = POLY6_LITERAL
    spiky_coef: float = SPIKY_LITERAL
    eps_d: float = EPS_D

    def replace(self, **kw) -> "Params":
        return dataclasses.replace(self, **kw)

    def to_c(self):
        from ._lib import PbfParams

        p = PbfParams()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name in ("box_lo", "box_hi"):
                arr = getattr(p, f.name)
                for a in range(3):
                    arr[a] = float(v[a])
            else:
                setattr(p, f.name, v)
        return p

    def as_dict(self) -> dict:
        return dataclasses.asdict(self)


def reference_params() -> Params:
    """The reference's native constants and Cornell-box collision."""
    return Params()


def poly6_lattice_density(h: float, spacing: float, include_self: bool = False) -> float:
    """SPH density of an interior particle of a cubic lattice, in float64:
    sum over lattice offsets of the reference's poly6 (src/particles.cpp:134-141)."""
    h2 = h * h
    n = int(math.ceil(h / spacing)) + 1
    total = 0.0
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for k in range(-n, n + 1):
                if not include_self and i == 0 and j == 0 and k == 0:
                    continue
                r2 = (i * spacing) ** 2 + (j * spacing) ** 2 + (k * spacing) ** 2
                if r2 < h2:
                    total += POLY6_LITERAL * (h2 - r2) ** 3 / h ** 9
    return total


SCENE_K, SCENE_XSPH_C, SCENE_VORT_EPS = 0.1, 0.01, 5e-4


def scene_params(h: float, spacing: float, box_lo, box_hi, dt: float = 0.01,
                 constants: str = "scaled", **kw) -> Params:
    """Constants of the synthetic PBF scenes: h, 4 iterations, epsilon 600,
    k 0.1, n 4, dq 0.2 h, XSPH 0.01, vorticity 5e-4, g 9.8, all six walls
    treated as the reference's mesh walls, rest density = poly6 density of an
    interior lattice particle (self excluded, as in the step's density).

    ``constants="literal"`` passes k, XSPH c and vorticity epsilon as given.
    The reference's formulas sum raw kernel values (src/particles.cpp:195,
    :227-229, :239) without the SPH volume m/rho, so at rho0 ~ 6.5e3 these
    literal values give an XSPH gain c*sum(W) ~ 65 and the step diverges.
    ``constants="scaled"`` (default) reads the three values as
    density-normalised (m/rho0 per kernel sum, m = 1) and divides them by
    rho0, which keeps the reference's formulas and gives a stable flow.
    """
    rho0 = float(poly6_lattice_density(h, spacing))
    if constants == "literal":
        k, c, ev = SCENE_K, SCENE_XSPH_C, SCENE_VORT_EPS
    elif constants == "scaled":
        k, c, ev = SCENE_K / rho0, SCENE_XSPH_C / rho0, SCENE_VORT_EPS / rho0
    else:
        raise ValueError(f"constants must be 'scaled' or 'literal', not {constants!r}")
    base = dict(
        h=h,
        rest_density=rho0,
        iterations=4,
        epsilon=600.0,
        k_tensile=k,
        n_tensile=4,
        dq=0.2,
        xsph_c=c,
        vort_eps=ev,
        gravity=9.8,
        dt=dt,
        box_lo=tuple(float(v) for v in box_lo),
        box_hi=tuple(float(v) for v in box_hi),
        wall_stop_mask=0,
        # ~30 neighbours at rest; scenes compressed against walls reach 4x
        # that, and the list costs only memory beyond each particle's count.
        max_neighbors=256,
    )
    base.update(kw)
    return Params(**base)
