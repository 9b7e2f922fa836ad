"""Build libreina_rt.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python reina-vk_amd/build_lib.py [--jobs N]

Every translation unit is compiled with -ffp-contract=off (no fused
multiply-add: the kernels' float arithmetic is exactly the sequence written)
and with packed-FP32 instructions disabled.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libreina_rt.so")

# host-only translation units (no device code): g++ with the same flags
HOST_SOURCES = ["ingest_obj.cpp", "ingest_png.cpp", "ingest_jpeg.cpp", "ingest_gltf.cpp"]
SOURCES = ["api.cpp", "bvh_host.cpp", "bvh_device.hip", "integrator.hip", "trace_f32.hip", "trace_q8.hip", "trace_multi_f32.hip", "trace_multi_q8.hip", "server.hip", "server.cpp", "tail.hip", "post.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=gfx950", "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def compile_one(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src in HOST_SOURCES:
        cmd = [CXX] + COMMON + ["-c", os.path.join(CSRC, src), "-o", out]
    else:
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES + HOST_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # the headless C++ host (links the library; rpath next to it)
    host_src = os.path.join(HERE, "host", "reina_headless.cpp")
    host_bin = os.path.join(HERE, "reina_headless")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", host_src, "-o", host_bin + ".tmp", "-L" + HERE, "-lreina_rt",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(host_bin + ".tmp", host_bin)
    if verbose:
        print("built", LIB, "and", host_bin)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
