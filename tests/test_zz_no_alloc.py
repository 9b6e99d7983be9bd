"""pbf_step_device_noalloc: steps caller-owned device arrays with every
buffer of the step in the library's fixed device workspace.  Checked bit for
bit against the allocating path (pbf_load + pbf_step_dt + pbf_read) on two
calls in a row, the second with fewer particles; sizes beyond the workspace
are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd._lib import PbfError, check, lib
from fluid_amd.pbf import Particles

pytestmark = pytest.mark.gpu


def allocating_path(p, x, v, steps):
    with Particles(p) as c:
        c.load(x, v)
        c.time_step(dt=p.dt, nsteps=steps)
        gx, gv, _ = c.read()
    return gx, gv


class DeviceArray:
    """float32 array in device memory, allocated by the test through the HIP
    runtime the library itself uses (the caller owns inputs and outputs)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            lib()  # the library's own HIP runtime is loaded first
            h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, a):
        self.host = np.ascontiguousarray(a, np.float32)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def noalloc_path(ctx, x, v, steps, dt):
    dx, dv = DeviceArray(x), DeviceArray(v)
    try:
        rc = lib().pbf_step_device_noalloc(ctx._ctx, dx.ptr, dv.ptr, len(x), float(dt), steps)
        check(rc, ctx._ctx)
        return dx.read(), dv.read()
    finally:
        dx.free()
        dv.free()


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_matches_allocating_path_on_two_calls():
    sc = scenes.dam_block()
    p = sc.params
    rng = np.random.default_rng(9)
    v = (sc.v + rng.uniform(-0.5, 0.5, sc.v.shape)).astype(np.float32)
    small = np.arange(0, sc.n, 3)  # 1,366 particles: the second, smaller call
    with Particles(p) as ctx:
        for x_in, v_in, steps in ((sc.x, v, 5), (sc.x[small], v[small], 3)):
            print(f"noalloc: {len(x_in)} particles, {steps} steps", flush=True)
            gx, gv = noalloc_path(ctx, x_in, v_in, steps, p.dt)
            rx, rv = allocating_path(p, x_in, v_in, steps)
            assert same_bits(gx, rx) and same_bits(gv, rv)
            assert not same_bits(gx, np.asarray(x_in, np.float32))


def test_noalloc_mid_size_states():
    """A 40,000-particle block (one lane per particle in the build and the
    vorticity pass, a quad in lambda, delta-p and XSPH; within the 64 MiB
    workspace at 256 neighbours), then every second particle (20,000: a
    wave in the build, lambda and XSPH, a quad in delta-p, a lane in the
    vorticity pass) in the same workspace: the allocating path's bits both
    times."""
    from fluid_amd import params as P

    p = P.scene_params(0.1, 0.05, (0.0, 0.0, 0.0), (3.0, 3.0, 3.0), dt=0.01)
    rng = np.random.default_rng(11)
    x = scenes.lattice_block(40, 40, 25, 0.05, (0.1, 0.1, 0.1)).astype(np.float32)
    x += rng.uniform(-1e-3, 1e-3, x.shape).astype(np.float32)
    v = rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)
    assert len(x) == 40_000
    with Particles(p) as ctx:
        for x_in, v_in, steps in ((x, v, 3), (x[::2], v[::2], 2)):
            print(f"noalloc: {len(x_in)} particles, {steps} steps", flush=True)
            gx, gv = noalloc_path(ctx, x_in, v_in, steps, p.dt)
            rx, rv = allocating_path(p, x_in, v_in, steps)
            assert same_bits(gx, rx) and same_bits(gv, rv)


def test_noalloc_refuses_sizes_beyond_the_workspace_and_bad_input():
    sc = scenes.dam_block()
    with Particles(sc.params) as ctx:
        n = 10_000_000  # far beyond the 64 MiB workspace; no device work is done
        rc = lib().pbf_step_device_noalloc(ctx._ctx, C.c_void_p(16), C.c_void_p(16), n,
                                           float(sc.params.dt), 1)
        assert rc == -1
        msg = lib().pbf_last_error(ctx._ctx).decode()
        assert "workspace" in msg and "at most" in msg
        x = sc.x.copy()
        x[17] = (5.0, 0.5, 0.5)  # outside the [0, 1]^3 box
        dx, dv = DeviceArray(x), DeviceArray(np.zeros_like(x))
        try:
            rc = lib().pbf_step_device_noalloc(ctx._ctx, dx.ptr, dv.ptr, len(x),
                                               float(sc.params.dt), 1)
        finally:
            dx.free()
            dv.free()
        assert rc == -5 and "particle 17" in lib().pbf_last_error(ctx._ctx).decode()
        with pytest.raises(PbfError):
            check(lib().pbf_step_device_noalloc(ctx._ctx, None, None, 0, 0.01, 1), ctx._ctx)
