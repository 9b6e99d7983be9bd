"""C ABI of include/pbf.h: the library loads, exports every declared symbol,
and rejects bad arguments before touching a device (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from fluid_amd import _lib
from fluid_amd.params import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"PBF_API\s+[\w\s\*]+?\b(pbf_\w+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_symbols()
    for core in ["pbf_create", "pbf_load", "pbf_step", "pbf_step_dt", "pbf_read",
                 "pbf_estimate_densities", "pbf_destroy", "pbf_last_error", "pbf_neighbors",
                 "pbf_cells", "pbf_create_batch"]:
        assert core in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the header
    assert set(declared_symbols()) <= set(_lib.SIGNATURES) | set(getattr(_lib, "EXTRA_SIGNATURES", {}))


def test_defaults_are_the_reference_constants():
    p = _lib.PbfParams()
    _lib.lib().pbf_params_default(C.byref(p))
    ref = Params()
    assert p.h == pytest.approx(0.3) and p.iterations == 12 and p.n_tensile == 4
    assert p.epsilon == 2.0 and p.k_tensile == pytest.approx(1e-4)
    assert p.gravity == 10.0 and p.dt == pytest.approx(0.016)
    assert tuple(p.box_hi) == pytest.approx((1.0, 1.49, 1.0))
    assert p.wall_stop_mask == ref.wall_stop_mask


@pytest.mark.parametrize("field,value", [("h", 0.0), ("rest_density", -1.0), ("dt", 0.0),
                                         ("iterations", 0), ("cell_scale", 0.5),
                                         ("max_neighbors", 0), ("h", float("nan"))])
def test_create_rejects_bad_params(field, value):
    p = Params().replace(**{field: value}).to_c()
    ctx = C.c_void_p()
    rc = _lib.lib().pbf_create(C.byref(p), C.byref(ctx))
    assert rc == -1 and not ctx
    assert _lib.lib().pbf_last_error(None)


def test_create_rejects_empty_box_and_null():
    p = Params().replace(box_hi=(1.0, -1.0, 1.0)).to_c()
    ctx = C.c_void_p()
    assert _lib.lib().pbf_create(C.byref(p), C.byref(ctx)) == -1
    assert _lib.lib().pbf_create(None, C.byref(ctx)) == -1
    assert _lib.lib().pbf_destroy(None) == 0
    assert _lib.lib().pbf_step(None, 1) == -1
    assert _lib.lib().pbf_read(None, None, None, None) == -1
