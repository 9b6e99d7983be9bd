"""The C++ host path: pbf_driver (particle XML in, -d seconds of steps, XML
out) over the Particles facade of include/pbf_particles.hpp."""
import os
import subprocess

import numpy as np
import pytest

from fluid_amd import params as P
from fluid_amd import xmlio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fluid_amd", "host", "pbf_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(*args, timeout=120):
    return subprocess.run([DRIVER, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=timeout)


@pytest.fixture(scope="module", autouse=True)
def driver_built():
    if not os.path.exists(DRIVER):
        from fluid_amd import build

        build.build_all()
    assert os.path.exists(DRIVER)


def test_driver_rejects_missing_and_foreign_files(tmp_path):
    r = run("-p", str(tmp_path / "nope.xml"))
    assert r.returncode == 2 and "not found" in r.stderr
    bad = tmp_path / "bad.xml"
    bad.write_text("<scene><density>1</density></scene>")
    r = run("-p", str(bad))
    assert r.returncode == 2 and "not a particles file" in r.stderr
    r = run("-x")
    assert r.returncode == 2 and "usage" in r.stderr


def test_xml_roundtrip(tmp_path):
    rho0, pos, vel = xmlio.load_particles(os.path.join(GOLDEN, "spheres_p.xml"))
    assert len(pos) == 2106 and rho0 == 700.0
    out = tmp_path / "rt.xml"
    xmlio.save_particles(str(out), rho0, pos, vel)
    r2, p2, v2 = xmlio.load_particles(str(out))
    assert r2 == rho0 and np.array_equal(p2, pos) and np.array_equal(v2, vel)


@pytest.mark.gpu
def test_driver_one_second_matches_oracle(tmp_path, orc):
    """-p particles/p.xml -d 1: 63 steps of 0.016 s with the reference's
    constants, bit-exact with the CPU restatement."""
    out = tmp_path / "out.xml"
    r = run("-p", os.path.join(GOLDEN, "p.xml"), "-d", "1.0", "-q", "-o", str(out))
    assert r.returncode == 0, r.stderr
    assert "done: 63 steps" in r.stdout
    rho0, pos, vel = xmlio.load_particles(os.path.join(GOLDEN, "p.xml"))
    p = P.reference_params().replace(rest_density=rho0)
    xr, vr, _ = orc.step(p, pos.astype(np.float32), vel.astype(np.float32), 63)
    _, gx, gv = xmlio.load_particles(str(out))
    assert np.array_equal(gx.astype(np.float32), xr)
    assert np.array_equal(gv.astype(np.float32), vr)


@pytest.mark.gpu
def test_driver_checkpoint_resume_matches_one_run(tmp_path):
    """-d 0.5 -s snap, then -r snap -d 0.49: the 32 + 31 steps give the same
    file as one -d 1 run of 63 steps."""
    src = os.path.join(GOLDEN, "p.xml")
    one, two, snap = tmp_path / "one.xml", tmp_path / "two.xml", tmp_path / "mid.pbfsnap"
    r = run("-p", src, "-d", "1.0", "-q", "-o", str(one))
    assert r.returncode == 0, r.stderr
    r = run("-p", src, "-d", "0.5", "-q", "-s", str(snap))
    assert r.returncode == 0 and "done: 32 steps" in r.stdout, r.stderr
    r = run("-p", src, "-r", str(snap), "-d", "0.49", "-q", "-o", str(two))
    assert r.returncode == 0 and "done: 31 steps, simulate_time 1.008000" in r.stdout, r.stdout + r.stderr
    assert one.read_bytes() == two.read_bytes()
    r = run("-p", src, "-r", str(tmp_path / "missing.pbfsnap"))
    assert r.returncode == 1 and "cannot open" in r.stderr
