"""Host code under AddressSanitizer and UBSan (CPU only, no device).

* fluid_amd/csrc/pbf_host.h, the device-free part of the C ABI: load
  validation and packing on host threads, the snapshot format with damaged
  files (truncated, extended, flipped bits, a header claiming 2^30
  particles, other parameters or scenes, version 1), and the BVH build of
  random and degenerate meshes (tests/sanitize/host_checks.cpp);
* the oracle in every mode, with and without a mesh collider
  (tests/sanitize/oracle_checks.c);
* the C++ driver's XML parsing on missing, foreign and malformed files
  (fluid_amd/host/pbf_driver.cpp; with no device the run stops at context
  creation, after the parse);
* the list passes' workgroup order (xcd_block of pbf_device.h, built by
  hipcc and run on the host): a permutation of the workgroups for every
  grid size, with the documented runs per XCD (tests/sanitize/xcd_order.hip).

Each is compiled here with -fsanitize=address,undefined and
-fno-sanitize-recover, so any finding fails the run; the host checks are
also built with ThreadSanitizer (the threaded load path).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="no host compiler")


def build(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def run(cmd, timeout=300):
    return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout, env=ENV)


def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_checks")
    build(["g++", "-std=c++17", "-Wall", *SAN, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "fluid_amd", "csrc"),
           os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), "-o", exe, "-pthread"])
    r = run([exe, str(tmp_path)])
    assert r.returncode == 0 and "all checks passed" in r.stdout, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_host_threads_under_thread_sanitizer(tmp_path):
    """The same host checks built with ThreadSanitizer: load validation and
    packing split over 1, 3 and 8 host threads (run_parts, the lowest bad
    particle kept across threads) show no data race.  (The oracle is left
    out: its OpenMP runtime is not built for ThreadSanitizer.)"""
    exe = str(tmp_path / "host_checks_tsan")
    build(["g++", "-std=c++17", "-Wall", "-g", "-O1", "-fsanitize=thread",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "fluid_amd", "csrc"),
           os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), "-o", exe, "-pthread"])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0 and "all checks passed" in r.stdout, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_checks")
    build(["gcc", "-std=c11", "-Wall", "-ffp-contract=off", "-fopenmp", *SAN,
           "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "pbf_oracle.c"),
           os.path.join(ROOT, "tests", "sanitize", "oracle_checks.c"), "-o", exe, "-lm"])
    r = run([exe])
    assert r.returncode == 0 and "all modes ran" in r.stdout, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_driver_parsing_under_sanitizers(tmp_path):
    lib = os.path.join(ROOT, "fluid_amd", "libpbf_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libpbf_hip.so not built")
    exe = str(tmp_path / "pbf_driver_asan")
    build(["g++", "-std=c++17", "-Wall", *SAN, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "fluid_amd", "host", "pbf_driver.cpp"), "-o", exe,
           "-L", os.path.join(ROOT, "fluid_amd"), "-lpbf_hip",
           "-Wl,-rpath," + os.path.join(ROOT, "fluid_amd")])
    cases = {
        "foreign.xml": ("<scene><density>1</density></scene>", "not a particles file"),
        "nodensity.xml": ("<particles><ps></ps></particles>", "no <density>"),
        "open.xml": ("<particles><density>1</density><ps><particle><pos>1 2 3</pos>",
                     "unterminated <particle>"),
        "badpos.xml": ("<particles><density>1</density><ps><particle><pos>1 x 3</pos>"
                       "<v>0 0 0</v></particle></ps></particles>", "malformed particle"),
        "nov.xml": ("<particles><density>1</density><ps><particle><pos>1 2 3</pos>"
                    "</particle></ps></particles>", "malformed particle"),
        "empty.xml": ("", "not a particles file"),
    }
    for name, (text, why) in cases.items():
        f = tmp_path / name
        f.write_text(text)
        r = run([exe, "-p", str(f), "-q"])
        assert r.returncode == 2 and why in r.stderr, (name, r.stderr)
    r = run([exe, "-p", str(tmp_path / "missing.xml")])
    assert r.returncode == 2 and "not found" in r.stderr
    r = run([exe, "-x"])
    assert r.returncode == 2 and "usage" in r.stderr
    # A valid file parses; without a device the run ends at context creation.
    r = run([exe, "-p", os.path.join(ROOT, "tests", "golden", "p.xml"), "-d", "0.02", "-q"])
    assert r.returncode in (0, 1), r.stderr
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="no hipcc")
def test_xcd_block_order_is_a_permutation(tmp_path):
    exe = str(tmp_path / "xcd_order")
    build(["hipcc", "-std=c++17", "--offload-arch=gfx950", "-g", "-O1",
           "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "fluid_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "sanitize", "xcd_order.hip"), "-o", exe])
    r = run([exe])
    assert r.returncode == 0 and "all grids checked" in r.stdout, r.stdout + r.stderr
    assert "runtime error" not in r.stderr
