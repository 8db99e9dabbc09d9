"""Builds the GPU tracer library (librpx.so, in this package's directory) with hipcc for gfx950.

Every HIP source of csrc/ goes into one shared library. The flags keep single-precision
arithmetic IEEE-exact (no FMA contraction, correctly rounded division and square root, f32
denormals kept), which is what lets the GPU reproduce the reference renderer's bits.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librpx.so")
SOURCES = ["rpx_kernels.hip", "rpx_engine.hip", "rpx_scene.hip", "rpx_capi.hip", "rpx_sobol.hip", "rpx_generators.hip",
           "rpx_adaptive.hip", "rpx_denoise.hip", "rpx_caustics.hip"]

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    path = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return path if os.path.exists(path) else "hipcc"


OBJ_DIR = os.path.join(HERE, "_build")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    return res.returncode


def build(verbose=False, jobs=None):
    """Compiles librpx.so (always: the call is the build check): each source to an object in
    parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([hipcc()] + compile_flags + ["-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1), 8)
    with ThreadPoolExecutor(jobs) as pool:
        codes = list(pool.map(lambda c: _run(c, verbose), cmds))
    if any(codes):
        raise RuntimeError("hipcc failed compiling a source of librpx.so")
    tmp = LIB + ".tmp"
    if _run([hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared"] + objs + ["-o", tmp], verbose):
        raise RuntimeError("hipcc failed linking librpx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
