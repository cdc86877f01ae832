"""Build libmgx.so (gfx950) in-tree with hipcc, and the CPU oracle with make.

Usage: python -m compressiontests_amd.build  (or __graft_entry__.build()).
The shared library lands next to this file (compressiontests_amd/libmgx.so) so that it
travels with the source tree; it is listed in .gitignore.
"""
import os
import shutil
import concurrent.futures
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmgx.so")

ARCH = "gfx950"
# -packed-fp32-ops: keep packed FP32 (v_pk_*_f32) instructions out of the generated code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I/opt/rocm/include"]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in ("mgx_kernels.hip", "mgx_block.hip", "mgx_nodeblock.hip", "mgx_duvi.hip", "mgx_api.hip"):
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
    # host-only code (setup, file I/O): no device code, no FMA contraction (bit-exact weights)
    for src in ("mgx_host.cpp", "mgx_io.cpp", "mgx_fem2d.cpp"):
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [hipcc, *HOST_FLAGS, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for o in outs:
            if o:
                print(o)
    objs = [j[0] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    """The CPU oracle is test infrastructure (see oracle/oracle.h); built here so that the
    tests and the bench's CPU baseline find it."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(verbose=False):
    lib = build_lib(verbose)
    orc = build_oracle()
    return lib, orc


if __name__ == "__main__":
    print(build_all(verbose=True))
