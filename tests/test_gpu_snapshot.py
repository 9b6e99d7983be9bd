"""Checkpoint / resume through the C ABI's binary snapshots: a resumed run
gives the bits of the uninterrupted one, and damaged or mismatched files are
refused."""
import os

import numpy as np
import pytest

from fluid_amd import scenes
from fluid_amd.pbf import Particles, PbfError

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_resume_is_bit_exact(tmp_path):
    sc = scenes.dam_block()
    path = str(tmp_path / "run.pbfsnap")
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=6)
        c.save_snapshot(path)
        mid = c.read()
        t_mid = c.simulate_time
        c.time_step(nsteps=6)
        whole = c.read()
        t_end = c.simulate_time
    n = len(sc.x)
    assert os.path.getsize(path) == 64 + 92 + 8 + 24 + 4 * 7 * n  # header, params, count, box, x v rho
    with Particles(sc.params) as r:
        r.load_snapshot(path)
        assert r.n == n and r.simulate_time == t_mid
        for a, b in zip(r.read(), mid):
            assert bits_equal(a, b)  # x, v and the last densities
        r.time_step(nsteps=6)
        assert r.simulate_time == t_end
        for a, b in zip(r.read(), whole):
            assert bits_equal(a, b)


def test_resume_mid_size_state_is_bit_exact(tmp_path):
    """40,000 particles (a lane per particle in the vorticity pass, which
    also predicts the next step when the call goes on; a quad per particle
    in lambda, delta-p and XSPH): 12 steps in one call against 6, a
    snapshot, a fresh context and 6 more."""
    from fluid_amd import params as P

    p = P.scene_params(0.1, 0.05, (0.0, 0.0, 0.0), (3.0, 3.0, 3.0), dt=0.01)
    rng = np.random.default_rng(13)
    x = scenes.lattice_block(40, 40, 25, 0.05, (0.1, 0.1, 0.1)).astype(np.float32)
    x += rng.uniform(-1e-3, 1e-3, x.shape).astype(np.float32)
    v = rng.uniform(-0.5, 0.5, x.shape).astype(np.float32)
    path = str(tmp_path / "lane.pbfsnap")
    with Particles(p) as c:
        c.load(x, v)
        c.time_step(nsteps=12)
        whole = c.read()
    with Particles(p) as c:
        c.load(x, v)
        c.time_step(nsteps=6)
        c.save_snapshot(path)
    with Particles(p) as r:
        r.load_snapshot(path)
        r.time_step(nsteps=6)
        for a, b in zip(r.read(), whole):
            assert bits_equal(a, b)


def test_batched_resume_is_bit_exact(tmp_path):
    items, p = scenes.batched(nscenes=5, nmin=500, nmax=2000)
    spec = [(len(x), lo, hi) for x, _, lo, hi in items]
    x0 = np.concatenate([it[0] for it in items])
    v0 = np.concatenate([it[1] for it in items])
    path = str(tmp_path / "batch.pbfsnap")
    with Particles(p, scenes=spec) as b:
        b.load(x0, v0)
        b.time_step(nsteps=4)
        b.save_snapshot(path)
        b.time_step(nsteps=4)
        whole = b.read()
    with Particles(p, scenes=spec) as r:
        r.load_snapshot(path)
        r.time_step(nsteps=4)
        for a, b_ in zip(r.read(), whole):
            assert bits_equal(a, b_)
    # a context with other scenes refuses it
    with Particles(p, scenes=spec[:-1] + [(spec[-1][0], spec[-1][1], np.asarray(spec[-1][2]) + 1)]) as o:
        with pytest.raises(PbfError, match="boxes"):
            o.load_snapshot(path)


def test_damaged_or_foreign_snapshots_are_refused(tmp_path):
    sc = scenes.dam_block()
    path = str(tmp_path / "a.pbfsnap")
    with Particles(sc.params) as c:
        c.load(sc.x, sc.v)
        c.time_step(nsteps=1)
        c.save_snapshot(path)
    raw = open(path, "rb").read()
    cases = {
        "flip": raw[:200] + bytes([raw[200] ^ 1]) + raw[201:],   # payload bit: checksum
        "short": raw[:-4],                                        # truncated
        "long": raw + b"\0",                                      # trailing bytes
        "magic": b"XBFSNAP1" + raw[8:],                           # not a snapshot
    }
    for name, data in cases.items():
        bad = str(tmp_path / f"{name}.pbfsnap")
        open(bad, "wb").write(data)
        with Particles(sc.params) as r:
            with pytest.raises(PbfError):
                r.load_snapshot(bad)
    with Particles(sc.params.replace(k_tensile=sc.params.k_tensile * 2)) as r:
        with pytest.raises(PbfError, match="parameters"):
            r.load_snapshot(path)
    with Particles(sc.params) as r:
        with pytest.raises(PbfError):
            r.load_snapshot(str(tmp_path / "missing.pbfsnap"))


PRIME = 1099511628211
MASK = (1 << 64) - 1


def fnv1a(data: bytes, h: int) -> int:
    for b in data:
        h = ((h ^ b) * PRIME) & MASK
    return h


def fnv4(payload: bytes) -> int:
    """The version-2 checksum as include/pbf.h states it, in plain Python."""
    basis = 14695981039346656037
    lanes = [basis] * 4
    nblk = len(payload) // 32
    words = np.frombuffer(payload[:32 * nblk], dtype="<u8").reshape(-1, 4)
    for w in words:
        for l in range(4):
            lanes[l] = ((lanes[l] ^ int(w[l])) * PRIME) & MASK
    lanes[0] = fnv1a(payload[32 * nblk:], lanes[0])
    return fnv1a(b"".join(v.to_bytes(8, "little") for v in lanes), basis)


def test_snapshot_checksum_is_the_documented_one_and_version_1_loads(tmp_path):
    """A small scene (odd payload length, so the byte tail is used): the
    header's checksum equals the documented four-lane FNV-1a; the same
    payload rewritten as a version-1 file (byte-serial FNV-1a from basis
    1469598103934665603) loads to the same state."""
    sc = scenes.dam_block()
    x, v = sc.x[:333].copy(), sc.v[:333].copy()
    path = str(tmp_path / "v2.pbfsnap")
    with Particles(sc.params) as c:
        c.load(x, v)
        c.time_step(nsteps=3)
        c.save_snapshot(path)
        state = c.read()
    raw = open(path, "rb").read()
    assert raw[:8] == b"PBFSNAP2"
    payload = raw[64:]
    assert len(payload) % 32 != 0
    assert int.from_bytes(raw[40:48], "little") == fnv4(payload)
    v1 = bytearray(raw)
    v1[:8] = b"PBFSNAP1"
    v1[40:48] = fnv1a(payload, 1469598103934665603).to_bytes(8, "little")
    path1 = str(tmp_path / "v1.pbfsnap")
    open(path1, "wb").write(bytes(v1))
    with Particles(sc.params) as r:
        r.load_snapshot(path1)
        for a, b in zip(r.read(), state):
            assert bits_equal(a, b)
    v1[40] ^= 1  # a version-1 file with a wrong checksum is refused
    open(path1, "wb").write(bytes(v1))
    with Particles(sc.params) as r:
        with pytest.raises(PbfError, match="checksum"):
            r.load_snapshot(path1)
