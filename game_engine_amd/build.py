"""Builds the backend in-tree: libge_rt.so (HIP, gfx950) and libge_synth.so (host C++).

HIP flags: -ffp-contract=off (no contraction: every kernel's arithmetic is the IEEE operation
order written in the source, with a fused multiply-add only where one is written), -fno-fast-math, correctly rounded division and square root
(hipcc's HIP default), and -packed-fp32-ops disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("GE_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["rt_api.hip", "rt_bvh.hip", "rt_kernels.hip", "rt_wavefront.hip", "rt_shadow.hip", "rt_mlp.hip",
               "rt_clusters.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=CSRC)


def build(verbose=True):
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    rt = os.path.join(LIB, "libge_rt.so")
    synth = os.path.join(LIB, "libge_synth.so")
    # One object a source, compiled side by side (each carries its own device code), then linked.
    # GE_HIP_DEFINES: extra -D flags for the walk's tuning macros (GE_WALK_BATCH, GE_WALK_REFILL, ...).
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"] + os.environ.get("GE_HIP_DEFINES", "").split()
    objs = [os.path.join(OBJ, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
    jobs = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, ["hipcc", f"--offload-arch={ARCH}", *compile_flags, "-c", "-o", obj, src])
                  for src, obj in zip(HIP_SOURCES, objs)]:
            f.result()
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", rt + ".tmp", *objs])
    os.replace(rt + ".tmp", rt)
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-o", synth + ".tmp", "scene_synth.cpp"])
    os.replace(synth + ".tmp", synth)
    return rt, synth


if __name__ == "__main__":
    build()
    sys.exit(0)
