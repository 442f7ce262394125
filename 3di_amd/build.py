"""Build libtdi.so (gfx950) in-tree with hipcc, and the C++ host layer libtdi_host.so with g++.

    python 3di_amd/build.py            # compile if sources changed
    python 3di_amd/build.py --force    # always compile
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtdi.so")
SOURCES = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
HEADERS = sorted(glob.glob(os.path.join(HERE, "csrc", "*.h"))) + [os.path.join(ROOT, "include", "tdi_api.h")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST_LIB = os.path.join(HERE, "libtdi_host.so")
HOST_SOURCES = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
HOST_HEADERS = [os.path.join(ROOT, "include", "tdi_host.h"), os.path.join(ROOT, "include", "tdi_model_fitter.h")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # no packed FP32 instructions anywhere in the device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def stale() -> bool:
    if not os.path.exists(LIB) or not os.path.exists(HOST_LIB):
        return True
    t = min(os.path.getmtime(LIB), os.path.getmtime(HOST_LIB))
    return any(os.path.getmtime(p) > t for p in SOURCES + HEADERS + HOST_SOURCES + HOST_HEADERS + [__file__])


def build_host() -> str:
    """The C++ host layer (3di_amd/host) as libtdi_host.so, linked against libtdi.so."""
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include"] + HOST_SOURCES + [
           "-L" + HERE, "-ltdi", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib",
           "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build(force: bool = False, verbose: bool = False) -> str:
    if not force and not stale():
        return LIB
    objs = []
    procs = []
    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(odir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", src, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(" ".join(cmd) + "\n" + text)
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    build_host()
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
