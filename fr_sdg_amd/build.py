"""Build libfrsdg.so (HIP, gfx950) in-tree.

    python -m fr_sdg_amd.build        # or fr_sdg_amd.build.build()

The library is a plain C-ABI shared object (include/frsdg.h); it links the
HIP runtime only.  libfrsdg_peaks.so (copy and FP64 FMA-loop probes) is
built beside it for the benchmark.  Packed-FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfrsdg.so")
# measured machine peaks for the benchmark (copy bandwidth, FP64 FMA rate); not part of the residual ABI
PEAKS_LIB = os.path.join(HERE, "libfrsdg_peaks.so")
PEAKS_SRC = os.path.join(HERE, "csrc", "frsdg_peaks.hip")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("frsdg_capi.hip", "frsdg_newton.hip", "frsdg_batch.hip")]
HEADERS = [
    os.path.join(HERE, "csrc", f)
    for f in ("frsdg_scalar.hpp", "frsdg_physics.hpp", "frsdg_geometry.hpp", "frsdg_kernels.hpp",
              "frsdg_tables.hpp", "frsdg_adjoint.hpp", "frsdg_reverse.hpp", "frsdg_limiter.hpp", "frsdg_transform.hpp", "frsdg_color.hpp", "frsdg_ctx.hpp")
] + [os.path.join(ROOT, "include", "frsdg.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++23",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    """Compile each translation unit in parallel, then link libfrsdg.so."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [HIPCC, *cflags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    tmp = PEAKS_LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, PEAKS_SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PEAKS_LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
