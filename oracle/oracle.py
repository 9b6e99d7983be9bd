"""ctypes wrapper of the CPU restatement (libpbf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker.  The product path never
imports it.

``p`` arguments are dicts (or objects with ``as_dict()``) holding the fields
of ``orc_params`` (pbf_oracle.h).  Arrays are numpy, positions (n, 3).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpbf_oracle.so")

NBR_GRID, NBR_BRUTE = 0, 1
XSPH_JACOBI, XSPH_SEQUENTIAL = 0, 1

DERIVED_NAMES = [
    "h", "h2", "h6", "h9", "c1", "m3c2", "tensile_S", "inv_rho0", "rho0", "eps", "negk",
    "xsph_c", "dt_vort", "inv_dt", "dt", "gdt", "cell", "inv_cell", "eps_d",
    "lo_e0", "lo_e1", "lo_e2", "hi_e0", "hi_e1", "hi_e2",
]


class OrcParams(C.Structure):
    _fields_ = [
        ("h", C.c_double),
        ("rest_density", C.c_double),
        ("iterations", C.c_int),
        ("epsilon", C.c_double),
        ("k_tensile", C.c_double),
        ("n_tensile", C.c_int),
        ("dq", C.c_double),
        ("xsph_c", C.c_double),
        ("vort_eps", C.c_double),
        ("gravity", C.c_double),
        ("dt", C.c_double),
        ("box_lo", C.c_double * 3),
        ("box_hi", C.c_double * 3),
        ("wall_stop_mask", C.c_uint),
        ("cell_scale", C.c_double),
        ("poly6_coef", C.c_double),
        ("spiky_coef", C.c_double),
        ("eps_d", C.c_double),
        ("mesh_pos", C.POINTER(C.c_double)),
        ("mesh_nrm", C.POINTER(C.c_double)),
        ("mesh_tri", C.POINTER(C.c_int32)),
        ("mesh_nt", C.c_int64),
        ("mesh_skin", C.c_double),
    ]


_mesh = None  # (pos, nrm, tri) of the mesh collider, or None


def set_mesh(pos=None, nrm=None, tri=None, skin=0.0):
    """Mesh collider used by every following call (None: the box walls).
    Positions, normals and the skin are rounded to float32 first, as the
    GPU path receives them."""
    global _mesh
    if pos is None:
        _mesh = None
        return
    pos = np.ascontiguousarray(np.asarray(pos, np.float32), np.float64).reshape(-1, 3)
    nrm = np.ascontiguousarray(np.asarray(nrm, np.float32), np.float64).reshape(-1, 3)
    tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
    _mesh = (pos, nrm, tri, float(np.float32(skin)))


_lib = None


def build() -> str:
    """Compiles the restatement with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True, stdout=subprocess.PIPE,
                   stderr=subprocess.STDOUT)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _as_dict(p) -> dict:
    return p.as_dict() if hasattr(p, "as_dict") else dict(p)


def params(p, single: bool = True) -> OrcParams:
    """orc_params from a parameter dict.  With single=True every value is
    first rounded to float32, so the double fields hold exactly the float
    parameters the GPU path receives."""
    d = _as_dict(p)
    o = OrcParams()
    conv = (lambda v: float(np.float32(v))) if single else float
    for name, ctype in OrcParams._fields_:
        if name not in d:
            continue
        v = d[name]
        if name in ("box_lo", "box_hi"):
            arr = getattr(o, name)
            for a in range(3):
                arr[a] = conv(v[a])
        elif ctype in (C.c_int, C.c_uint):
            setattr(o, name, int(v))
        else:
            setattr(o, name, conv(v))
    if _mesh is not None:
        pos, nrm, tri, skin = _mesh
        o.mesh_skin = skin
        o.mesh_pos = pos.ctypes.data_as(C.POINTER(C.c_double))
        o.mesh_nrm = nrm.ctypes.data_as(C.POINTER(C.c_double))
        o.mesh_tri = tri.ctypes.data_as(C.POINTER(C.c_int32))
        o.mesh_nt = len(tri)
        o._keep = _mesh
    return o


def _t(dtype):
    return "f" if dtype == np.float32 else "d"


def _ptr(a: np.ndarray):
    ct = C.c_float if a.dtype == np.float32 else (C.c_double if a.dtype == np.float64 else C.c_uint32)
    return a.ctypes.data_as(C.POINTER(ct))


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def derive(p, dtype=np.float32):
    out = np.zeros(len(DERIVED_NAMES), dtype=dtype)
    dims = np.zeros(3, dtype=np.int32)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_derive_" + _t(dtype))(C.byref(op), _ptr(out), dims.ctypes.data_as(C.POINTER(C.c_int)))
    return out, tuple(int(v) for v in dims)


def poly6(p, r2: float, dtype=np.float32) -> float:
    t = _t(dtype)
    fn = getattr(lib(), "orc_poly6_" + t)
    fn.restype = C.c_float if t == "f" else C.c_double
    fn.argtypes = [C.POINTER(OrcParams), C.c_float if t == "f" else C.c_double]
    op = params(p, t == "f")
    return fn(C.byref(op), r2)


def spiky(p, r, dtype=np.float32) -> np.ndarray:
    r = _arr(r, dtype)
    g = np.zeros(3, dtype)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_spiky_" + _t(dtype))(C.byref(op), _ptr(r), _ptr(g))
    return g


def clamp_response(p, pos, vel, dt, dtype=np.float32) -> np.ndarray:
    pos = _arr(pos, dtype).copy()
    vel = _arr(vel, dtype)
    t = _t(dtype)
    fn = getattr(lib(), "orc_clamp_response_" + t)
    fn.argtypes = [C.POINTER(OrcParams), C.c_void_p, C.c_void_p, C.c_float if t == "f" else C.c_double]
    op = params(p, t == "f")
    fn(C.byref(op), pos.ctypes.data, vel.ctypes.data, dt)
    return pos


def clamp(p, pos, dp, dtype=np.float32) -> np.ndarray:
    pos = _arr(pos, dtype).copy()
    dp = _arr(dp, dtype)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_clamp_" + _t(dtype))(C.byref(op), _ptr(pos), _ptr(dp))
    return pos


def cells(p, xs) -> np.ndarray:
    xs = _arr(xs, np.float32)
    out = np.zeros(len(xs), np.uint32)
    op = params(p)
    lib().orc_cells_f(C.byref(op), C.c_size_t(len(xs)), _ptr(xs), _ptr(out))
    return out


def neighbors(p, xs, mode=NBR_GRID, dtype=np.float32):
    """CSR neighbour lists (offsets[n+1], list)."""
    xs = _arr(xs, dtype)
    n = len(xs)
    cap = 64 * n + 64
    op = params(p, dtype == np.float32)
    fn = getattr(lib(), "orc_neighbors_" + _t(dtype))
    fn.restype = C.c_long
    while True:
        off = np.zeros(n + 1, np.uint32)
        lst = np.zeros(cap, np.uint32)
        tot = fn(C.byref(op), int(mode), C.c_size_t(n), _ptr(xs), _ptr(off), _ptr(lst), C.c_size_t(cap))
        if tot >= 0:
            return off, lst[:tot]
        if tot == -2:
            raise MemoryError("oracle neighbour build")
        cap *= 2


def predict(p, x, v, dt=None, dtype=np.float32):
    x = _arr(x, dtype)
    v = _arr(v, dtype).copy()
    xs = np.zeros_like(x)
    t = _t(dtype)
    dt = _as_dict(p)["dt"] if dt is None else dt
    fn = getattr(lib(), "orc_predict_" + t)
    fn.argtypes = [C.POINTER(OrcParams), C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_float if t == "f" else C.c_double]
    op = params(p, t == "f")
    fn(C.byref(op), len(x), x.ctypes.data, v.ctypes.data, xs.ctypes.data, dt)
    return xs, v


def lambda_pass(p, xs, off, lst, dtype=np.float32):
    xs = _arr(xs, dtype)
    rho = np.zeros(len(xs), dtype)
    lam = np.zeros(len(xs), dtype)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_lambda_" + _t(dtype))(C.byref(op), C.c_size_t(len(xs)), _ptr(xs), _ptr(off),
                                              _ptr(lst), _ptr(rho), _ptr(lam))
    return rho, lam


def deltap_pass(p, xs, lam, off, lst, dtype=np.float32):
    xs = _arr(xs, dtype)
    lam = _arr(lam, dtype)
    out = np.zeros_like(xs)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_deltap_" + _t(dtype))(C.byref(op), C.c_size_t(len(xs)), _ptr(xs), _ptr(lam),
                                              _ptr(off), _ptr(lst), _ptr(out))
    return out


def xsph_pass(p, x, xs, v_pred, off, lst, mode=XSPH_JACOBI, dt=None, dtype=np.float32):
    x, xs, v_pred = _arr(x, dtype), _arr(xs, dtype), _arr(v_pred, dtype)
    n = len(x)
    vout = np.zeros((n, 3), dtype)
    om = np.zeros((n, 3), dtype)
    rho = np.zeros(n, dtype)
    t = _t(dtype)
    dt = _as_dict(p)["dt"] if dt is None else dt
    fn = getattr(lib(), "orc_xsph_" + t)
    fn.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_size_t] + [C.c_void_p] * 8 + [
        C.c_float if t == "f" else C.c_double]
    op = params(p, t == "f")
    fn(C.byref(op), int(mode), n, x.ctypes.data, xs.ctypes.data, v_pred.ctypes.data, off.ctypes.data,
       lst.ctypes.data, vout.ctypes.data, om.ctypes.data, rho.ctypes.data, dt)
    return vout, om, rho


def vorticity_pass(p, xs, om, off, lst, v, dt=None, dtype=np.float32):
    xs, om = _arr(xs, dtype), _arr(om, dtype)
    v = _arr(v, dtype).copy()
    x = np.zeros_like(xs)
    t = _t(dtype)
    dt = _as_dict(p)["dt"] if dt is None else dt
    fn = getattr(lib(), "orc_vorticity_" + t)
    fn.argtypes = [C.POINTER(OrcParams), C.c_size_t] + [C.c_void_p] * 6 + [
        C.c_float if t == "f" else C.c_double]
    op = params(p, t == "f")
    fn(C.byref(op), len(xs), xs.ctypes.data, om.ctypes.data, off.ctypes.data, lst.ctypes.data,
       v.ctypes.data, x.ctypes.data, dt)
    return x, v


def step(p, x, v, nsteps=1, nbr_mode=NBR_GRID, xsph_mode=XSPH_JACOBI, dt=None, dtype=np.float32):
    """Whole steps; returns (x, v, rho) after nsteps."""
    x = _arr(x, dtype).copy()
    v = _arr(v, dtype).copy()
    rho = np.zeros(len(x), dtype)
    t = _t(dtype)
    dt = _as_dict(p)["dt"] if dt is None else dt
    fn = getattr(lib(), "orc_step_" + t)
    fn.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, C.c_size_t, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_float if t == "f" else C.c_double, C.c_int]
    op = params(p, t == "f")
    rc = fn(C.byref(op), int(nbr_mode), int(xsph_mode), len(x), x.ctypes.data, v.ctypes.data,
            rho.ctypes.data, dt, int(nsteps))
    if rc != 0:
        raise MemoryError("oracle step allocation failed")
    return x, v, rho


def estimate_densities(p, x, mode=NBR_GRID, dtype=np.float32):
    x = _arr(x, dtype)
    rho = np.zeros(len(x), dtype)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_estimate_densities_" + _t(dtype))(C.byref(op), int(mode), C.c_size_t(len(x)),
                                                          _ptr(x), _ptr(rho))
    return rho


def density_at(p, x, pts, mode=NBR_GRID, dtype=np.float32):
    """Density of particles x at query points pts (estimateDensityAt)."""
    x = _arr(x, dtype)
    pts = _arr(pts, dtype)
    out = np.zeros(len(pts), dtype)
    op = params(p, dtype == np.float32)
    getattr(lib(), "orc_density_at_" + _t(dtype))(C.byref(op), int(mode), C.c_size_t(len(x)), _ptr(x),
                                                  C.c_size_t(len(pts)), _ptr(pts), _ptr(out))
    return out


def csr_to_lists(off, lst):
    return [lst[off[i]:off[i + 1]] for i in range(len(off) - 1)]
