"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/frsdg.h declares, and the product path has no CPU fallback."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "frsdg.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(frsdg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("frsdg_create", "frsdg_residual", "frsdg_rhs", "frsdg_ifunction", "frsdg_jvp",
                 "frsdg_jvp_cstep", "frsdg_vjp", "frsdg_matmult", "frsdg_filter", "frsdg_ssprk_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import fr_sdg_amd as fr
    lib = fr.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", fr.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (frsdg_\w+)", out))
    assert set(declared()) <= exported
    assert set(fr.EXPORTS) <= exported


def test_strerror_and_invalid_descriptor_need_no_gpu():
    import fr_sdg_amd as fr
    lib = fr.load_library()
    assert lib.frsdg_strerror(0) == b"success"
    assert b"invalid" in lib.frsdg_strerror(1)
    # invalid descriptors are rejected before any HIP call
    d = fr.make_desc(dict(dims=2, nx=4, ny=4, p=9))
    h = ctypes.c_void_p()
    assert lib.frsdg_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d = fr.make_desc(dict(dims=2, nx=4, ny=4, p=3, bc_type=[7, 1, 1, 1]))
    assert lib.frsdg_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d = fr.make_desc(dict(dims=3, nx=4, ny=4, p=3))
    assert lib.frsdg_create(ctypes.byref(d), ctypes.byref(h)) == 1


def test_no_cpu_fallback():
    import torch

    import fr_sdg_amd as fr
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(fr.FrsdgError):
        fr.PhysicsEvaluator(dict(dims=2, nx=2, ny=2, p=2))


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "fr_sdg_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in txt and "fr_oracle" not in txt and "liboracle" not in txt, f


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N (N > 1) exits with a message and status 2 before touching a GPU."""
    import json
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
