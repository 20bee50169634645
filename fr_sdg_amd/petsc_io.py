"""PETSc binary Vec files, the format of the reference's
``-ts_monitor_solution binary:<file>`` output that its post-processing reads
(tools/plot.py:109-129, tools/check_iso_vortex.py:214-285).

Each vector is: big-endian int32 class id 1211214 (VEC_FILE_CLASSID),
big-endian int32 global length, then the values as big-endian float64.
Vectors are appended one after another (one per monitored time step).  The
values are in DMDA natural order, which is the layout of this library's state
arrays: 2D states reshape to (ny, nx, n, n, 4) (check_iso_vortex.py:244).
"""
import os

import numpy as np

VEC_FILE_CLASSID = 1211214


def _host(u):
    try:
        import torch
        if isinstance(u, torch.Tensor):
            return u.detach().to("cpu", torch.float64).numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(u, dtype=np.float64)


def write_vec(path, u, append=True):
    """Append (or write) one vector."""
    a = _host(u).reshape(-1)
    with open(path, "ab" if append else "wb") as f:
        np.array([VEC_FILE_CLASSID, a.size], dtype=">i4").tofile(f)
        a.astype(">f8").tofile(f)


def read_vecs(path):
    """All vectors of a file, as float64 arrays."""
    out = []
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        while f.tell() < size:
            hdr = np.fromfile(f, dtype=">i4", count=2)
            if hdr.size < 2:
                break
            if int(hdr[0]) != VEC_FILE_CLASSID:
                raise ValueError(f"not a PETSc Vec record (class id {int(hdr[0])})")
            n = int(hdr[1])
            out.append(np.fromfile(f, dtype=">f8", count=n).astype(np.float64))
    return out


def as_fields(u, nx, ny, p):
    """(ny, nx, n, n, 4) view of a 2D state (node (k,l) at [..., l, k, :])."""
    n = p + 1
    return _host(u).reshape(ny, nx, n, n, 4)
