"""Build the native library for gfx950 (and nothing else).

    python badchimp-cpp_amd/build.py

Produces badchimp-cpp_amd/lib/libbadchimp_amd.so from csrc/.  The FP64
kernels are compiled with -ffp-contract=off (no multiply-add fusion) so that
their arithmetic matches the reference's double arithmetic operation by
operation.  The reduced-precision kernels fuse multiply-adds only within one
expression (-ffp-contract=on), so that every instantiation of a kernel (with
or without the optional stores and reductions) rounds the same way and a run
does not depend on how its steps are split into calls.  Every HIP file is
compiled without packed FP32 instructions.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
HIP = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
       "-munsafe-fp-atomics"]
SOURCES = {
    "kernels_f64.hip": HIP + ["-O3", "-ffp-contract=off"],
    "kernels_f32.hip": HIP + ["-O3", "-ffp-contract=on"],
    "capi.hip": HIP + ["-O2", "-ffp-contract=off"],
    "batch.hip": HIP + ["-O2", "-ffp-contract=off"],
    "calibrate.hip": HIP + ["-O3"],
    "topo_device.hip": HIP + ["-O3"],
    "halo.hip": HIP + ["-O3"],
    "topology.cpp": ["-O2", "-ffp-contract=off", "-x", "c++"],
    "vtu.cpp": ["-O2", "-x", "c++"],
}


def _compile(name, flags):
    src = os.path.join(CSRC, name)
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), res.stdout, res.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        futs = {ex.submit(_compile, n, f): n for n, f in SOURCES.items()}
        objs = []
        for fut in cf.as_completed(futs):
            objs.append(fut.result())
            if verbose:
                print("compiled", futs[fut])
    out = os.path.join(LIB, "libbadchimp_amd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + sorted(objs)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (res.stdout, res.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
