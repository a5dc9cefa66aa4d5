"""Compile the native libraries in-tree.

* ``lib/libsimcoproj.so`` -- the HIP engine for gfx950 (hipcc).
* ``lib/libsimcogen.so``  -- the synthetic-configuration generator (g++).

Both are rebuilt on every call; the Python loaders only load what exists.
"""

import os
import subprocess
import sys

from ._paths import ENGINE_LIB, INCLUDE_DIR, LIB_DIR, PKG_DIR

CSRC = os.path.join(PKG_DIR, "csrc")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

# gfx950 only.  Packed FP32 instructions are disabled for every HIP object.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]

HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall"]


def engine_sources():
    d = os.path.join(CSRC, "engine")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".hip") or f.endswith(".cpp"))


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gen():
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libsimcogen.so")
    _run(["g++"] + HOST_FLAGS + ["-o", out + ".tmp", os.path.join(CSRC, "gen", "simco_gen.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_engine(jobs=8):
    os.makedirs(LIB_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(LIB_DIR, "libsimcoproj.so")
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = engine_sources()
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [hipcc] + [f for f in HIP_FLAGS if f != "-shared"] + [
            "-I", INCLUDE_DIR, "-I", os.path.join(CSRC, "engine"), "-c", s, "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("hipcc failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    """The CPU oracle; where the reference tree and amdflang are present, also
    the reference's own VTR writer with its test host (oracle/_ref)."""
    repo = os.path.dirname(PKG_DIR)
    _run(["make", "-s", "-C", os.path.join(repo, "oracle")])
    if os.path.isdir("/root/reference/src") and os.path.exists(FLANG):
        _run(["make", "-s", "-C", os.path.join(repo, "oracle"), "ref"])


def build_fortran():
    """The Fortran binding module (include/simco_proj.f90) and the Fortran
    test hosts under tests/fortran, linked against the engine library."""
    repo = os.path.dirname(PKG_DIR)
    if not os.path.exists(FLANG):
        print("flang not found: Fortran binding not built")
        return None
    src = os.path.join(repo, "tests", "fortran")
    out = os.path.join(src, "_build")
    os.makedirs(out, exist_ok=True)
    mod_obj = os.path.join(out, "simco_proj.o")
    _run([FLANG, "-O2", "-J", out, "-c", os.path.join(repo, "include", "simco_proj.f90"),
          "-o", mod_obj])
    lib_dir = os.path.dirname(ENGINE_LIB)
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(lib_dir, out)
    for prog in ("sp_layout", "sp_drive"):
        _run([FLANG, "-O2", "-J", out, os.path.join(src, prog + ".f90"), mod_obj,
              "-L" + lib_dir, "-lsimcoproj", rpath, "-o", os.path.join(out, prog)])
    return out


def build_all():
    build_gen()
    build_oracle()
    build_engine()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"gen": build_gen, "engine": build_engine, "oracle": build_oracle,
         "fortran": build_fortran, "all": build_all}[w]()
