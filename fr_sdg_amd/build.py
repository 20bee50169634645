"""Build libfrsdg.so (HIP, gfx950) in-tree.

    python -m fr_sdg_amd.build        # or fr_sdg_amd.build.build()

The library is a plain C-ABI shared object (include/frsdg.h); it links the
HIP runtime only.  Packed-FP32 instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfrsdg.so")
SOURCES = [os.path.join(HERE, "csrc", "frsdg_capi.hip")]
HEADERS = [
    os.path.join(HERE, "csrc", f)
    for f in ("frsdg_scalar.hpp", "frsdg_physics.hpp", "frsdg_geometry.hpp", "frsdg_kernels.hpp",
              "frsdg_tables.hpp", "frsdg_adjoint.hpp")
] + [os.path.join(ROOT, "include", "frsdg.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
