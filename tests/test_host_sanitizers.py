"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
library's host-only sources (csrc/topology.cpp: the node-list builder, the
.vtklb reader and writer, the binary twin; csrc/vtu.cpp: the VTU writer) are
built with g++ and both sanitizers into the driver tests/host/topology_sanitize.cpp,
which runs them on small D2Q9 and D3Q19 geometries (one rank and several), a
single fluid cell, and their error paths (missing and truncated files, bad
arguments).  The run must end with status 0, leak-free, with no report."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_sources_clean_under_asan_ubsan():
    csrc = os.path.join(ROOT, "badchimp-cpp_amd", "csrc")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "topology_sanitize")
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
               "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "include"), "-I", csrc,
               os.path.join(ROOT, "tests", "host", "topology_sanitize.cpp"), os.path.join(csrc, "topology.cpp"),
               os.path.join(csrc, "vtu.cpp"), "-o", exe]
        subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
        # a preloaded library of the environment comes before the ASan runtime
        # in the link order: let ASan accept that instead of refusing to start
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
                   UBSAN_OPTIONS="print_stacktrace=1")
        res = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "host sanitizer run: ok" in res.stdout
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr
