"""Build libcourant.so in-tree: host set-up with g++, device engine with hipcc for gfx950.

Host set-up code is compiled without floating-point contraction so the
exported plan matches the reference core bit for bit; device code is built for
gfx950 only, with packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcourant.so")

HOST_SOURCES = ["setup.cpp", "materials.cpp", "mode2d.cpp", "modeport.cpp", "board.cpp"]
DEVICE_SOURCES = ["engine.hip", "mode_gpu.hip"]

CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra",
              "-Wno-unused-parameter", "-fvisibility=hidden", "-I" + INCLUDE, "-I" + CSRC]
# -ffp-contract=off: every kernel rounds each operation as written (the FP32 CPU
# emulation of the oracle is built the same way), independent of the compiler's
# multiply-add fusion choices.
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "-ffp-contract=off", "-fvisibility=hidden", "-Wno-unused-result",
                "-I" + INCLUDE, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    """Compile every source and link courant_amd/libcourant.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(BUILD, s + ".o") for s in HOST_SOURCES + DEVICE_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-Wl,-Bsymbolic", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_napi()
    return LIB


NAPI_DIR = os.path.join(HERE, "napi")
NAPI_ADDON = os.path.join(HERE, "courant_napi.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_napi():
    """The Node addon (courant_napi.node, next to libcourant.so): wrappers generated from
    include/courant.h, compiled against the system's node_api.h; skipped where Node's
    headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    inc = os.path.join(BUILD, "courant_napi_gen.inc")
    _run([sys.executable, os.path.join(NAPI_DIR, "gen_addon.py"), os.path.join(INCLUDE, "courant.h"), inc])
    tmp = NAPI_ADDON + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I" + NODE_INCLUDE, "-I" + INCLUDE, "-I" + BUILD,
          os.path.join(NAPI_DIR, "courant_napi.cpp"), "-L" + HERE, "-lcourant", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, NAPI_ADDON)
    return NAPI_ADDON


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
