"""PETSc binary Vec round trip, checked against the byte layout the
reference's reader expects (tools/plot.py:109-129): int32 class id 1211214,
int32 length, float64 data, all big-endian, vectors appended."""
import struct

import numpy as np

from fr_sdg_amd import petsc_io


def test_round_trip_and_layout(tmp_path):
    f = tmp_path / "sol.bin"
    a = np.arange(12, dtype=np.float64) * 0.5
    b = -np.linspace(0, 1, 7)
    petsc_io.write_vec(f, a, append=False)
    petsc_io.write_vec(f, b)
    raw = f.read_bytes()
    assert struct.unpack(">ii", raw[:8]) == (1211214, 12)
    assert struct.unpack(">d", raw[8 + 8 * 3:8 + 8 * 4])[0] == 1.5
    off = 8 + 8 * 12
    assert struct.unpack(">ii", raw[off:off + 8]) == (1211214, 7)
    got = petsc_io.read_vecs(f)
    assert len(got) == 2 and np.array_equal(got[0], a) and np.array_equal(got[1], b)


def test_field_view_matches_reference_reshape():
    nx, ny, p = 3, 2, 2
    n = p + 1
    u = np.arange(ny * nx * n * n * 4, dtype=np.float64)
    F = petsc_io.as_fields(u, nx, ny, p)
    # element (i=1, j=1), node (k=2, l=0), eq 3 at ((j*nx+i)*n*n + l*n + k)*4 + eq
    assert F[1, 1, 0, 2, 3] == u[((1 * nx + 1) * n * n + 0 * n + 2) * 4 + 3]
