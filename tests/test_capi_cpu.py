"""The C ABI library on the host: it loads, exports every function declared
in include/*.h, and its host-only entry points behave (no GPU needed)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import lbamd
from helpers import ROOT


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(lb_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_api():
    names = declared_functions()
    for required in ("lb_ctx_create", "lb_set_topology", "lb_upload_f", "lb_step_bgk", "lb_step_cg",
                     "lb_reduce_flux", "lb_download_scalar", "lb_topo_read_vtklb"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(lbamd.ffi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(lbamd.ffi.SIGNATURES)


def test_version_and_null_handling():
    L = lbamd.lib()
    assert b"gfx950" in L.lb_version()
    assert L.lb_ctx_destroy(None) == 0
    assert L.lb_last_error(None) == b"null context"
    assert L.lb_topo_num_nodes(None) == -1


def test_create_reports_missing_device_or_bad_args():
    L = lbamd.lib()
    h = ctypes.c_void_p()
    assert L.lb_ctx_create(7, 0, 0, ctypes.byref(h)) == -1
    assert L.lb_ctx_create(0, 9, 0, ctypes.byref(h)) == -1
    assert b"lattice" in L.lb_global_error() or b"precision" in L.lb_global_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="host without a GPU only")
def test_create_without_gpu_fails_loudly():
    with pytest.raises(lbamd.LBError) as e:
        lbamd.Context(lbamd.D3Q19)
    assert e.value.code == -5


def test_sphere_pack_argument_checks():
    with pytest.raises(lbamd.LBError):
        lbamd.sphere_pack((10, 10, 10), -1.0, 0.4, 1)
    with pytest.raises(lbamd.LBError):
        lbamd.sphere_pack((10, 10, 10), 2.0, 1.5, 1)


def test_topology_roundtrip_through_writer(tmp_path):
    geo, _ = lbamd.sphere_pack((18, 14, 12), 3.0, 0.5, 9, periodic="x", walls=True)
    t = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    rho = np.linspace(0.9, 1.1, t.n_nodes)
    p = str(tmp_path / "g.vtklb")
    t.write_vtklb(p, [("init_rho", rho)])
    r = lbamd.Topology.from_vtklb(p, lbamd.D3Q19)
    for m in ("neighbors", "positions", "node_types", "bulk"):
        assert np.array_equal(getattr(t, m)(), getattr(r, m)())
    got = r.attribute("init_rho")
    assert np.array_equal(got[1:], rho[1:])


def test_batch_footprint_without_gpu():
    """The footprint of a batched solve is host arithmetic: fluid count and
    the sum of 256-byte aligned regions (2 population buffers, coordinates,
    row windows, block partials)."""
    geo, _ = lbamd.sphere_pack((20, 18, 16), 3.0, 0.5, 4, periodic="x", walls=True)
    nbytes, nf = lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, geo, "x")
    assert nf == int((geo > 0).sum())
    NP = -(-nf // 256) * 256
    a = lambda v: -(-v // 256) * 256
    E = [20, 16, 14]  # periodic x spans the box; y, z span the fluid cells inside the walls
    nwin = (E[0] + 2) * (E[1] + 2) * (-(-E[2] // 32))
    assert nbytes == 2 * a(19 * NP * 4) + a(NP * 4) + a(nwin * 16) + a(-(-nf // 256) * 8)
    nb64, _ = lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP64, geo, "x")
    assert nb64 - nbytes == 2 * (a(19 * NP * 8) - a(19 * NP * 4))


def test_batch_create_without_gpu_fails_loudly():
    if lbamd.lib().lb_batch_create is None:
        pytest.skip("symbol missing")
    import ctypes as C
    h = C.c_void_p()
    rc = lbamd.lib().lb_batch_create(lbamd.D3Q19, lbamd.FP32S, 0, 1 << 20, 100, C.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        assert rc == 0
        lbamd.lib().lb_batch_destroy(h)
    else:
        assert rc != 0 and b"device" in lbamd.lib().lb_global_error()


def test_two_field_launches_leave_the_row_windows_above_4gib(tmp_path):
    """The row windows address one population buffer with 32-bit byte
    offsets: a launch over both fields of a two-phase FP64 context whose
    buffer exceeds 4 GiB (D3Q19, 14.1 M to 28.3 M nodes) must take the plain
    rows, while a one-field launch of the same context keeps the windows."""
    import subprocess
    src = tmp_path / "sel.cpp"
    src.write_text(r'''
#include <cstdio>
#include "layout_select.hpp"
int main() {
  using namespace lbamd;
  const long long nps[] = {256, 14000128, 14128384, 20000000, 28246784, 28246528 + 256 * 2};
  for (long long NP : nps)
    std::printf("%lld %d %d %d %d %d\n", NP, launch_layout(2, 19, NP, 1, 8), launch_layout(2, 19, NP, 2, 8),
                launch_layout(2, 19, NP, 2, 4), launch_layout(1, 19, NP, 2, 8), launch_layout(0, 19, NP, 2, 8));
  return 0;
}
''')
    exe = tmp_path / "sel"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "badchimp-cpp_amd", "csrc"), str(src),
                    "-o", str(exe)], check=True)
    rows = [list(map(int, l.split())) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                             check=True).stdout.split("\n") if l]
    for NP, one64, two64, two32, rows_layout, table in rows:
        assert one64 == (2 if 19 * NP * 8 < 2**32 else 1)
        assert two64 == (2 if 2 * 19 * NP * 8 < 2**32 else 1)
        assert two32 == (2 if 2 * 19 * NP * 4 < 2**32 else 1)
        assert rows_layout == 1 and table == 0
    # the case the selection exists for: one FP64 field fits, two do not
    assert any(r[1] == 2 and r[2] == 1 for r in rows)
