"""Build the HIP/C++ sources of langevindynamicsforpcp_amd in-tree for gfx950.

    python -m langevindynamicsforpcp_amd.build          # library + executables

Outputs:
  langevindynamicsforpcp_amd/lib/libldpcp.so      HIP kernels + C ABI (include/ldpcp.h)
  langevindynamicsforpcp_amd/bin/langevinDynamicsSingleBarrier
  langevindynamicsforpcp_amd/bin/langevinDynamicsDoubleBarrier
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIB_DIR, "libldpcp.so")
ARCH = os.environ.get("LDPCP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every multiply and add rounds on its own, as in the
# reference; kernels request fused multiply-adds explicitly.
# packed FP32 ops are disabled: the packed forms are not used by this library.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", os.path.join(REPO, "include")]

LIB_SOURCES = ["ldpcp_common.cpp", "ldpcp_run.hip", "ldpcp_engine.hip", "ldpcp_server.hip"]
HOST_COMMON = ["input_parser.cpp", "restart_file.cpp", "runner.cpp", "text_util.cpp"]
EXECUTABLES = {"langevinDynamicsSingleBarrier": ("main_single.cpp", 0),
               "langevinDynamicsDoubleBarrier": ("main_double.cpp", 1)}


def _run(cmd):
    print("+ " + " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.check_call(cmd)


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *HIP_FLAGS, "-c", *lang, src, "-o", obj]
    _run(cmd)
    return obj


def build_library():
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_executables():
    if not os.path.isdir(HOST):
        return []
    os.makedirs(BIN_DIR, exist_ok=True)
    common = [os.path.join(HOST, s) for s in HOST_COMMON if os.path.exists(os.path.join(HOST, s))]
    out = []
    for name, (main, _) in EXECUTABLES.items():
        main_src = os.path.join(HOST, main)
        if not os.path.exists(main_src):
            continue
        exe = os.path.join(BIN_DIR, name)
        _run(["g++", "-O2", "-std=c++17", "-pthread", "-Wall", "-I", os.path.join(REPO, "include"), "-I", HOST,
              main_src, *common, "-o", exe, f"-L{LIB_DIR}", "-lldpcp", "-Wl,-rpath,$ORIGIN/../lib"])
        out.append(exe)
    return out


def build():
    lib = build_library()
    exes = build_executables()
    return lib, exes


if __name__ == "__main__":
    build()
    sys.exit(0)
