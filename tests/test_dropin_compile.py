"""The drop-in boundary at compile level: the reference's callers of
Particles (src/application.cpp:302-344, src/pathtracer.cpp:444-480) compile
against the replacement header over include/pbf_particles.hpp, with the
reference's own CGL::Vector3D, CGL::Color and tinyxml2 headers.

Reads the reference's headers where they lie, so it runs only where the
reference tree is present (the build container) and is skipped elsewhere.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/CGL/include"


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference tree absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_reference_callers_compile_against_the_facade(tmp_path):
    src = os.path.join(ROOT, "tests", "dropin", "reference_callers.cpp")
    obj = tmp_path / "callers.o"
    cmd = ["g++", "-std=gnu++11", "-Wall", "-Werror", "-c", src, "-o", str(obj),
           "-I", REF_INCLUDE, "-I", os.path.join(ROOT, "include")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    # The object refers to the C ABI the library exports, and to nothing of
    # the reference's step (CGL::Particles::timeStep and friends).
    syms = subprocess.run(["nm", "-C", "--undefined-only", str(obj)], stdout=subprocess.PIPE,
                          text=True, check=True).stdout
    for name in ("pbf_create", "pbf_load", "pbf_step_dt", "pbf_estimate_densities", "pbf_read",
                 "pbf_density_at"):
        assert name in syms, name
    assert "CGL::Particles::timeStep" not in syms
