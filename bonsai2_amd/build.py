"""Builds bonsai2_amd/libbonsai2_amd.so in-tree for gfx950 with hipcc.

Every call compiles all sources (objects go to bonsai2_amd/_build/); the
shared library is written next to this file so it travels with the tree.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libbonsai2_amd.so")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["k_sort.hip", "k_tree.hip", "k_groups.hip", "k_walk.hip", "k_misc.hip",
               "k_learn.hip", "capi.hip"]
HOST_SOURCES = ["host_io.cpp"]

ARCH = os.environ.get("BONSAI_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 (v_pk_*_f32) is switched off for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(HERE, "..", "include")]
# BONSAI_HIPCC_EXTRA: further device-compiler options (e.g. -DBT_BUILD_WAVE_PRIO=0)
HIP_FLAGS += os.environ.get("BONSAI_HIPCC_EXTRA", "").split()
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(HERE, "..", "include")]
# what `hipconfig --cpp_config` prints: HIP's vector types (double4, ...) for
# host-compiled code that includes include/Treecode.h
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIP_HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__=", "-I" + os.path.join(ROCM, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return r.returncode, " ".join(cmd), r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    objs = [j[-1] for j in jobs]
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(_run, jobs))
    for rc, cmd, out in results:
        if verbose or rc != 0:
            sys.stderr.write(cmd + "\n" + out)
        if rc != 0:
            raise RuntimeError("compile failed: " + cmd)
    tmp = OUT + ".tmp"
    rc, cmd, out = _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
                        ["-lpthread"])
    if verbose or rc != 0:
        sys.stderr.write(cmd + "\n" + out)
    if rc != 0:
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    # the command-line driver, linked against the library next to it
    exe = os.path.join(HERE, "treecode")
    rc, cmd, out = _run(["g++"] + HOST_FLAGS + HIP_HOST_FLAGS +
                        [os.path.join(CSRC, "treecode_main.cpp"), "-o",
                                                 exe + ".tmp", "-L" + HERE, "-lbonsai2_amd",
                                                 "-Wl,-rpath,$ORIGIN"])
    if verbose or rc != 0:
        sys.stderr.write(cmd + "\n" + out)
    if rc != 0:
        raise RuntimeError("driver link failed")
    os.replace(exe + ".tmp", exe)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
