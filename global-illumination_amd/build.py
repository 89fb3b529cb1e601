"""Builds the HIP library libvgi.so for gfx950, in-tree (next to this file).

Float flags matter for parity: -ffp-contract=off keeps every a*b+c as two
rounded operations, and correctly rounded fp32 division and square root keep
`/` and sqrt IEEE, as the CPU restatement computes them.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvgi.so")
WORKLOADS_LIB = os.path.join(HERE, "libvgi_workloads.so")
BUILD = os.path.join(HERE, "_build")

INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_FLAGS = [
    "-I" + INCLUDE,
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-I" + INCLUDE, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math"]

HIP_SOURCES = ["vgi_kernels.hip", "vgi_wavefront.hip", "vgi_select.hip", "vgi_sun.hip", "vgi_mlp.hip", "vgi_view.hip", "vgi_host.cpp"]
HOST_SOURCES = ["vgi_host_scene.cpp"]
# The seeded workload generators of the tests and the benchmark: a library of
# their own, outside the product's boundary.
WORKLOAD_SOURCES = ["vgi_scenes.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False, jobs=None):
    """Compile every source (in parallel) and link libvgi.so. Always rebuilds."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    commands, objects = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        commands.append([hipcc] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        commands.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    jobs = jobs or max(1, min(len(commands), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for out in pool.map(_run, commands):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objects + ["-lpthread"])
    os.replace(tmp, LIB)
    tmp = WORKLOADS_LIB + ".tmp"
    _run(["g++"] + HOST_FLAGS + ["-shared", "-o", tmp] + [os.path.join(CSRC, src) for src in WORKLOAD_SOURCES] + ["-lpthread"])
    os.replace(tmp, WORKLOADS_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
