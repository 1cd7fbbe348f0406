"""Build recipes for the native parts of reapers_amd.

- libreapers_hip.so  : the gfx950 HIP kernels behind the C ABI in
                       include/reapers_hip.h (hipcc, one object per .hip file)
- libreapers_host.so : the header-only C++ API instantiated for Python callers
                       (g++, links libreapers_hip.so)
- libreapers_cpu.so  : the same API built with REAPERS_NOGPU (the CPUImpl
                       backend) behind include/reapers_cpu.h (g++, no HIP)
- reapers_service    : the JSON request/response program (g++, links
                       libreapers_hip.so)
- examples/*         : example programs of the C++ API (g++)

Run ``python -m reapers_amd.build`` or call :func:`build_all`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["rp_runtime.hip", "rp_vec.hip", "rp_apply.hip", "rp_mlp.hip", "rp_batch.hip",
               "rp_dense.hip"]
HOST_SOURCES = ["host_api.cc"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_* FP32 instructions anywhere in the device code
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]
# per-source additions: the apply kernels branch on uniform values (the row
# flip of a sub, its class run) and are compiled without the structurizer's
# rewrite of uniform regions, which otherwise routes every accumulator of a
# multi-way uniform branch through copies at its join
HIP_FILE_FLAGS = {"rp_apply.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions"]}
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = [
    "-O2", "-fPIC", "-std=c++20", "-fopenmp", "-ffp-contract=off", "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(HERE, "include"),
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers(dirpath, exts=(".h", ".hpp")):
    out = []
    for base, _, files in os.walk(dirpath):
        for f in files:
            if f.endswith(exts):
                out.append(os.path.join(base, f))
    return out


def build_hip(force=False):
    os.makedirs(BUILD, exist_ok=True)
    deps = _headers(CSRC) + [os.path.join(ROOT, "include", "reapers_hip.h")]
    objs = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        if force or _stale(o, [s] + deps):
            _run([HIPCC] + HIP_FLAGS + HIP_FILE_FLAGS.get(src, []) + ["-c", s, "-o", o])
        objs.append(o)
    lib = os.path.join(HERE, "libreapers_hip.so")
    if force or _stale(lib, objs):
        tmp = lib + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
             + ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(tmp, lib)
    return lib


def build_host(force=False):
    deps = (_headers(os.path.join(HERE, "include")) + _headers(CSRC)
            + [os.path.join(ROOT, "include", "reapers_hip.h"),
               os.path.join(HERE, "libreapers_hip.so")])
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    lib = os.path.join(HERE, "libreapers_host.so")
    if force or _stale(lib, srcs + deps):
        tmp = lib + ".tmp"
        _run([CXX] + HOST_FLAGS + ["-shared", "-o", tmp] + srcs
             + ["-L" + HERE, "-lreapers_hip", "-Wl,-rpath,$ORIGIN",
                "-Wl,--no-as-needed"])
        os.replace(tmp, lib)
    return lib


def build_cpu(force=False):
    """libreapers_cpu.so: the C++ API with REAPERS_NOGPU (CPUImpl backend) behind
    include/reapers_cpu.h; no HIP library is linked."""
    deps = (_headers(os.path.join(HERE, "include"))
            + [os.path.join(ROOT, "include", "reapers_hip.h"),
               os.path.join(ROOT, "include", "reapers_cpu.h")])
    src = os.path.join(CSRC, "cpu_api.cc")
    lib = os.path.join(HERE, "libreapers_cpu.so")
    if force or _stale(lib, [src] + deps):
        tmp = lib + ".tmp"
        flags = [f for f in HOST_FLAGS if f not in ("-O2", "-ffp-contract=off")]
        _run([CXX, "-O3"] + flags + ["-shared", "-o", tmp, src])
        os.replace(tmp, lib)
    return lib


def build_service(force=False):
    """reapers_service: the JSON request/response program (csrc/service.cc)."""
    deps = (_headers(os.path.join(HERE, "include"))
            + [os.path.join(ROOT, "include", "reapers_hip.h"),
               os.path.join(HERE, "libreapers_hip.so")])
    src = os.path.join(CSRC, "service.cc")
    exe = os.path.join(HERE, "reapers_service")
    if force or _stale(exe, [src] + deps):
        tmp = exe + ".tmp"
        _run([CXX] + HOST_FLAGS + ["-o", tmp, src, "-L" + HERE, "-lreapers_hip",
                                   "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, exe)
    return exe


def build_examples(force=False):
    """The example programs under examples/ (g++, links libreapers_hip.so)."""
    deps = (_headers(os.path.join(HERE, "include"))
            + [os.path.join(ROOT, "include", "reapers_hip.h"),
               os.path.join(HERE, "libreapers_hip.so")])
    edir = os.path.join(HERE, "examples")
    out = []
    for src in sorted(os.listdir(edir)):
        if not src.endswith(".cc"):
            continue
        s = os.path.join(edir, src)
        exe = os.path.join(edir, src[:-3])
        if force or _stale(exe, [s] + deps):
            tmp = exe + ".tmp"
            _run([CXX] + HOST_FLAGS + ["-I" + os.path.join(HERE, "include", "reapers"),
                                       "-o", tmp, s, "-L" + HERE, "-lreapers_hip",
                                       "-Wl,-rpath,$ORIGIN/.."])
            os.replace(tmp, exe)
        out.append(exe)
    return out


def build_oracle(force=False):
    odir = os.path.join(ROOT, "oracle")
    if force:
        _run(["make", "-C", odir, "-B"])
    else:
        _run(["make", "-C", odir])
    return os.path.join(odir, "liboracle.so")


def build_all(force=False):
    build_hip(force)
    build_host(force)
    build_cpu(force)
    build_service(force)
    build_examples(force)
    build_oracle(force)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
