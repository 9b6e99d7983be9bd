"""Python handle on a PBF context (thin layer over the C ABI of include/pbf.h).

:class:`Particles` mirrors the reference's ``Particles`` load/step interface
(src/particles.h:105-140): construct with the parameters, load the particles
(the reference's repeated ``addParticle``), ``time_step`` (``timeStep``),
``estimate_densities`` (``estimateDensities``), ``simulate_time``.  State
read-back is in load order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import BUF, STAGE, check, lib
from .params import Params


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Particles:
    """One simulation context on the current HIP device.

    ``scenes``: optional list of (count, box_lo, box_hi) for a multi-scene
    (batched) context; all scenes share the physics of ``params``.
    """

    def __init__(self, params: Params, scenes=None):
        self.params = params
        self._ctx = C.c_void_p()
        cp = params.to_c()
        if scenes is None:
            check(lib().pbf_create(C.byref(cp), C.byref(self._ctx)))
            self.scene_counts = None
        else:
            counts = np.asarray([int(s[0]) for s in scenes], dtype=np.int64)
            lo = np.ascontiguousarray([s[1] for s in scenes], dtype=np.float32)
            hi = np.ascontiguousarray([s[2] for s in scenes], dtype=np.float32)
            check(
                lib().pbf_create_batch(
                    C.byref(cp),
                    len(scenes),
                    counts.ctypes.data_as(C.POINTER(C.c_int64)),
                    _fptr(lo),
                    _fptr(hi),
                    C.byref(self._ctx),
                )
            )
            self.scene_counts = counts
        self.n = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self._ctx:
            lib().pbf_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- reference interface ----------------------------------------------
    def load(self, x: np.ndarray, v: np.ndarray | None = None) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        check(lib().pbf_load(self._ctx, _fptr(x), None if vv is None else _fptr(vv), len(x)), self._ctx)
        self.n = len(x)

    def time_step(self, dt: float | None = None, nsteps: int = 1) -> None:
        if dt is None:
            check(lib().pbf_step(self._ctx, nsteps), self._ctx)
        else:
            check(lib().pbf_step_dt(self._ctx, float(dt), nsteps), self._ctx)

    def set_schedule(self, schedule: str, block: int = 256) -> None:
        """'flat', 'graph' (hipGraph-captured inner steps) or 'persistent' (one launch,
        workgroups run whole scenes); none changes the arithmetic."""
        check(lib().pbf_set_schedule(self._ctx, _lib.SCHEDULE[schedule], int(block)), self._ctx)

    def step_async(self, nsteps: int = 1, dt: float | None = None) -> None:
        dt = self.params.dt if dt is None else dt
        check(lib().pbf_step_async(self._ctx, float(dt), nsteps), self._ctx)

    def sync(self) -> None:
        check(lib().pbf_sync(self._ctx), self._ctx)

    def read(self):
        x = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        rho = np.empty(self.n, np.float32)
        check(lib().pbf_read(self._ctx, _fptr(x), _fptr(v), _fptr(rho)), self._ctx)
        return x, v, rho

    def estimate_densities(self) -> np.ndarray:
        rho = np.empty(self.n, np.float32)
        check(lib().pbf_estimate_densities(self._ctx, _fptr(rho)), self._ctx)
        return rho

    def set_mesh(self, pos=None, nrm=None, tri=None, skin: float = 0.0) -> None:
        """Triangle-mesh collider (vertex positions, vertex normals facing the
        fluid, vertex-index triples, skin clearance) in place of the box
        walls; None removes it."""
        if pos is None:
            check(lib().pbf_set_mesh(self._ctx, None, None, 0, None, 0, 0.0), self._ctx)
            return
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        check(lib().pbf_set_mesh(self._ctx, _fptr(pos), _fptr(nrm), len(pos),
                                 tri.ctypes.data_as(C.POINTER(C.c_int32)), len(tri),
                                 float(skin)), self._ctx)

    def save_snapshot(self, path: str) -> None:
        """Binary checkpoint of the committed state (x, v, rho, simulate_time)."""
        check(lib().pbf_save_snapshot(self._ctx, os.fsencode(path)), self._ctx)

    def load_snapshot(self, path: str) -> None:
        """Resume from a checkpoint written with the same parameters/scenes."""
        check(lib().pbf_load_snapshot(self._ctx, os.fsencode(path)), self._ctx)
        self.n = lib().pbf_count(self._ctx)

    def density_at(self, pts, scene: int = 0) -> np.ndarray:
        """Density of the particles at query points pts[m, 3]
        (estimateDensityAt, src/particles.cpp:446-453)."""
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.empty(len(pts), np.float32)
        check(lib().pbf_density_at(self._ctx, scene, _fptr(pts), len(pts), _fptr(out)), self._ctx)
        return out

    def density_lattice(self, lo, hi, step: float, counts, scene: int = 0) -> np.ndarray:
        """Density on the lattice (min(hi, lo + i*step)), shape
        (counts[2], counts[1], counts[0])."""
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        cnt = np.ascontiguousarray(counts, np.int32)
        out = np.empty(int(np.prod(cnt.astype(np.int64))), np.float32)
        check(lib().pbf_density_lattice(self._ctx, scene, _fptr(lo), _fptr(hi), float(step),
                                        cnt.ctypes.data_as(C.POINTER(C.c_int32)), _fptr(out)),
              self._ctx)
        return out.reshape(int(cnt[2]), int(cnt[1]), int(cnt[0]))

    def time_step_profiled(self, nsteps: int = 1, dt: float | None = None) -> dict:
        """Steps like time_step and returns {pass: (total ms, launches)} from
        HIP events recorded around every pass on the context stream."""
        dt = self.params.dt if dt is None else dt
        n = len(_lib.PASSES)
        ms = np.zeros(n, np.float32)
        cnt = np.zeros(n, np.int32)
        check(lib().pbf_step_profiled(self._ctx, float(dt), nsteps, _fptr(ms),
                                      cnt.ctypes.data_as(C.POINTER(C.c_int32)), n), self._ctx)
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_lib.PASSES)}

    @property
    def simulate_time(self) -> float:
        return lib().pbf_simulate_time(self._ctx)

    # -- test hooks ---------------------------------------------------------
    def neighbors(self, stride: int = 128):
        counts = np.empty(self.n, np.uint32)
        ids = np.empty((self.n, stride), np.uint32)
        check(lib().pbf_neighbors(self._ctx, _uptr(counts), _uptr(ids), stride), self._ctx)
        return counts, ids

    def neighbor_stats(self):
        """(mean, max) neighbour count of the last step."""
        tot, mx = C.c_uint64(), C.c_uint32()
        check(lib().pbf_neighbor_stats(self._ctx, C.byref(tot), C.byref(mx)), self._ctx)
        return tot.value / max(self.n, 1), mx.value

    def cells(self) -> np.ndarray:
        out = np.empty(self.n, np.uint32)
        check(lib().pbf_cells(self._ctx, _uptr(out)), self._ctx)
        return out

    def grid_dims(self, scene: int = 0):
        d = (C.c_int32 * 3)()
        check(lib().pbf_grid_dims(self._ctx, scene, d), self._ctx)
        return tuple(d)

    def step_debug(self, stage: str, dt: float | None = None) -> None:
        dt = self.params.dt if dt is None else dt
        check(lib().pbf_step_debug(self._ctx, float(dt), STAGE[stage]), self._ctx)

    def debug_read(self, which: str) -> np.ndarray:
        out = np.empty((self.n, 4), np.float32)
        check(lib().pbf_debug_read(self._ctx, BUF[which], _fptr(out)), self._ctx)
        return out

    def derived(self) -> np.ndarray:
        out = np.empty(32, np.float32)
        check(lib().pbf_debug_derived(self._ctx, _fptr(out), 25), self._ctx)
        return out[:25]

    def math_flags(self) -> dict:
        """Which short sequences (division by h^9, reciprocal, square root)
        passed the exhaustive device check and are in use."""
        out = np.empty(32, np.float32)
        check(lib().pbf_debug_derived(self._ctx, _fptr(out), 32), self._ctx)
        return {"div": bool(out[25]), "rcp": bool(out[26]), "sqrt": bool(out[27])}


PbfError = _lib.PbfError
