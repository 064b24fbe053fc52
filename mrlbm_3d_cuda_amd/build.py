"""Build recipe of the HIP library (libmrlbm.so, in-tree) for gfx950.

    python -m mrlbm_3d_cuda_amd.build

Compiles every source under csrc/ with hipcc into one shared library next to this file.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmrlbm.so")
ARCH = os.environ.get("MRLBM_OFFLOAD_ARCH", "gfx950")

# per-source extra flags: the FP64 kernels and the output arrays are compiled without FMA
# contraction; the FP32-arithmetic kernels divide and take square roots with the hardware
# reciprocal and square-root approximations (about 1 ulp, like the bulk path's reciprocal of rho)
# instead of the correctly rounded sequences, which dominated the boundary-node closed forms, and
# flush FP32 denormals (the moments are O(1) and their deviations far above 1e-38), which removes
# the operand scaling the approximations otherwise need to keep denormal results; they fuse
# multiply-adds only within one source expression, so a function rounds alike in every kernel it
# is inlined into (the batched and the single-case fused kernels give the same bits)
SOURCES = {
    "geometry.cpp": [],
    "kernels_f64.hip": ["-ffp-contract=off"],
    "kernels_f32.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
                        "-ffp-contract=on"],
    "capi.hip": [],
    "batch.hip": [],
    "output.hip": ["-ffp-contract=off"],
    # the allocation-free read-back decodes with the FP32-arithmetic kernels' flags
    "noalloc.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"],
}

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def build(verbose=False, extra_flags=()):
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    cmds = []
    for src, flags in SOURCES.items():
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([hipcc(), *HIPCC_FLAGS, *flags, *extra_flags, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the sources compile independently: run the compilers side by side (a few at a time)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "4")), os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        futures = []
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
            futures.append(pool.submit(subprocess.run, cmd, check=True))
        for f in futures:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
