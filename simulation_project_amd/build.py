"""Builds libmpsr.so (HIP for gfx950 + C++ host code) in-tree.

    python -m simulation_project_amd.build          # compile what changed
    python -m simulation_project_amd.build --force  # compile everything

Every HIP translation unit is compiled with -ffp-contract=off (no fused multiply-adds: the
restated reference arithmetic rounds every product and sum separately) and without packed FP32
instructions.  Objects go to simulation_project_amd/_build/, the library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpsr.so")
RUNNER = os.path.join(HERE, "mpsr_run")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("MPSR_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["sort.hip", "particles.hip", "walls.hip", "inflow.hip", "spheres.hip", "mpnn.hip", "ppe.hip", "api.hip",
               "noalloc.hip"]
CXX_SOURCES = ["host_setup.cpp", "deck.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
             "-Wno-unknown-cuda-version"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + \
        [os.path.join(HERE, "..", "include", "mpsr.h")]


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + _headers())


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, r.returncode, r.stdout + r.stderr


def build(force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for f in HIP_SOURCES:
        src, obj = os.path.join(CSRC, f), os.path.join(BUILD, f + ".o")
        objs.append(obj)
        if force or _stale(src, obj):
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    for f in CXX_SOURCES:
        src, obj = os.path.join(CSRC, f), os.path.join(BUILD, f + ".o")
        objs.append(obj)
        if force or _stale(src, obj):
            jobs.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    nproc = min(8, max(1, len(jobs)))
    with ThreadPoolExecutor(nproc) as ex:
        results = list(ex.map(_compile, jobs))
    failed = [r for r in results if r[1] != 0]
    for cmd, rc, out in results:
        if verbose or rc != 0:
            noisy = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
            if noisy:
                sys.stderr.write("\n".join(noisy) + "\n")
    if failed:
        raise RuntimeError("libmpsr build failed: " + " ".join(failed[0][0]))
    if jobs or force or not os.path.exists(LIB):
        # link next to the library and rename over it, so a reader never sees a half-written file
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + \
            [f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("libmpsr link failed:\n" + r.stdout + r.stderr)
        os.replace(tmp, LIB)
    # the command-line driver: host code over the C ABI, linked against the library next to it
    src = os.path.join(CSRC, "mpsr_run.cpp")
    if force or not os.path.exists(RUNNER) or os.path.getmtime(RUNNER) < max(os.path.getmtime(src), os.path.getmtime(LIB)):
        cmd = ["g++"] + CXX_FLAGS + ["-pthread", src, "-o", RUNNER, f"-L{HERE}", "-lmpsr", "-Wl,-rpath,$ORIGIN",
                                     f"-Wl,-rpath-link,{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("mpsr_run build failed:\n" + r.stdout + r.stderr)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose="-v" in sys.argv)
    print(LIB)
