"""Builds libspruce_gpu.so (the HIP kernels + the C ABI of include/spruce_gpu.h) for gfx950, in place."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspruce_gpu.so")
SOURCES = ["domain.hip", "outputs.hip", "mhd.hip", "mhd2.hip", "solar.hip", "twofluid.hip", "tf_host.hip", "visc.hip", "e2.hip", "e2_host.hip", "pvisc.hip", "adjoint.hip", "sources.hip", "adj_modules.hip", "adj_pvisc.hip", "moc.hip"]

# -ffp-contract=off: no implicit multiply-add fusion (the reference rounds every product).
# -packed-fp32-ops off: packed FP32 instructions are not used on this hardware pool.
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
         "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    """Compiles every source to an object in parallel (objects under spruce_amd/_build), then links."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = "-I" + os.path.join(ROOT, "include")
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def comp(src: str) -> str:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + compile_flags + [inc, "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
        return obj

    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(comp, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_runner(verbose)
    return LIB


RUNNER = os.path.join(HERE, "spruce_run")


def build_runner(verbose: bool = False) -> str:
    """The command-line program (csrc/run_main.cpp): host C++ on the C ABI, linked to the library."""
    rocm_lib = "/opt/rocm/lib"
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "run_main.cpp"),
           "-L" + HERE, "-lspruce_gpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + rocm_lib, "-o", RUNNER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("runner build failed:\n" + r.stdout + r.stderr)
    os.replace(RUNNER + ".tmp", RUNNER)
    return RUNNER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
