"""Pins of the oracle's element stabilisation with the WENO limiter and shock
sensor (filter.hpp:486-755, physics_impl_{1D,2D}.hpp limit_solution).

The reference's own tests hold no vectors for this path, so the oracle is
pinned by what the mathematics fixes:
  * smoothness indicators of polynomials in closed form (D p = b for p = a + b xi,
    D^2 xi^2 = 2): beta(a + b xi) = 4 b^2, beta(c xi^2) = (16/3 + 64) c^2;
  * the limited line of a shocked element keeps its cell average;
  * equal smoothness gives the linear weights and no shock (identity);
  * with the spectral filter also on, the filter's output replaces WENO's;
  * 2D = 1D limiter along x averaged with the identity along y for data
    constant in y; and an array-level numpy restatement of the 2D sweep
    (neighbours, zero ghost layer, directions) on jumpy data.
"""
import numpy as np
import pytest

from pyoracle import Oracle

EPS = 1e-6


def cfg1d(nx=3, p=3, periodic=True, filt=0, gamma=0.998, theta=1.0, thr=0.8):
    bc = [5, 5, 5, 5] if periodic else [1, 1, 1, 1]
    return dict(dims=1, nx=nx, ny=1, p=p, lx=1.0, riemann=0, bc_type=bc, apply_filter=filt, filt_cutoff=0,
                filt_strength=16, filt_alpha=36.0, apply_weno=1, weno_gamma=gamma, shock_theta=theta,
                shock_threshold=thr)


def sensor(betas, gamma=0.998, theta=1.0):
    d = np.array([(1 - gamma) / 2, gamma, (1 - gamma) / 2])
    c = d / (EPS + np.asarray(betas)) ** 2
    om = c / c.sum()
    r = np.sum(np.abs(om / d - 1) ** theta) / (abs(1 / d.min() - 1) ** theta + 2)
    return r, om


def state1d(o, lines):
    """1D state [E][eq][node]: eq 0 from `lines` (E x n), eq 1 = 0, eq 2 = 2.5."""
    n = o.n
    E = len(lines)
    st = np.zeros((E, 3, n))
    st[:, 0, :] = lines
    st[:, 2, :] = 2.5
    return st.reshape(-1)


def test_closed_form_smoothness_and_weno_1d():
    o = Oracle(cfg1d())
    xi = o.tables()["csi"]
    w = o.tables()["w"]
    # prev slope 0.01, curr = 0.7 + 0.5 xi^2 (beta 69.33 * 0.25), next slope 0.3
    prev = 1.0 + 0.01 * xi
    curr = 0.7 + 0.5 * xi * xi
    nxt = 0.3 + 0.3 * xi
    betas = [4 * 0.01 ** 2, (16.0 / 3 + 64.0) * 0.25, 4 * 0.3 ** 2]
    r, om = sensor(betas)
    assert r >= 0.8  # the rough own element against a smooth neighbour trips the sensor
    out = o.filter(state1d(o, [prev, curr, nxt])).reshape(3, 3, -1)
    avg = lambda p: 0.5 * np.dot(w, p)
    want = om[0] * (prev - avg(prev) + avg(curr)) + om[1] * curr + om[2] * (nxt - avg(nxt) + avg(curr))
    np.testing.assert_allclose(out[1, 0], want, rtol=0, atol=1e-12)
    # equations 1 and 2 are constant: untouched, bitwise
    assert np.array_equal(out[:, 1], np.zeros((3, o.n)))
    assert np.array_equal(out[:, 2], np.full((3, o.n), 2.5))


def test_equal_smoothness_is_identity():
    o = Oracle(cfg1d(nx=4, p=4))
    xi = o.tables()["csi"]
    lines = [0.5 + e + 0.3 * xi for e in range(4)]  # same slope everywhere
    u = state1d(o, lines)
    assert np.array_equal(o.filter(u), u)


def test_shocked_lines_keep_cell_average():
    o = Oracle(cfg1d(nx=6, p=5))
    w = o.tables()["w"]
    rng = np.random.default_rng(3)
    lines = rng.uniform(0.1, 2.0, (6, o.n))
    lines[2, 3:] = 5.0  # a jump inside element 2
    u = state1d(o, lines)
    out = o.filter(u).reshape(6, 3, -1)
    a_in = lines @ w
    a_out = out[:, 0] @ w
    np.testing.assert_allclose(a_out, a_in, rtol=1e-14, atol=1e-14)
    assert not np.array_equal(out[:, 0], lines)  # something was limited


def test_zero_ghost_at_nonperiodic_end():
    o = Oracle(cfg1d(nx=3, p=3, periodic=False))
    xi = o.tables()["csi"]
    w = o.tables()["w"]
    curr = 1.0 + 0.2 * xi
    nxt = 1.4 + 0.2 * xi
    r, om = sensor([0.0, 4 * 0.04, 4 * 0.04])  # the ghost neighbour is the zero line
    assert r >= 0.8
    out = o.filter(state1d(o, [curr, nxt, nxt + 0.4])).reshape(3, 3, -1)
    ac = 0.5 * np.dot(w, curr)
    an = 0.5 * np.dot(w, nxt)
    want = om[0] * (0.0 - 0.0 + ac) + om[1] * curr + om[2] * (nxt - an + ac)
    np.testing.assert_allclose(out[0, 0], want, rtol=0, atol=1e-12)


def test_filter_replaces_weno():
    rng = np.random.default_rng(4)
    cw = cfg1d(nx=5, p=4, filt=1)
    cf = dict(cw, apply_weno=0)
    lines = rng.uniform(0.1, 2.0, (5, 5))
    lines[1, 2:] = 6.0
    u = state1d(Oracle(cw), lines)
    assert np.array_equal(Oracle(cw).filter(u), Oracle(cf).filter(u))


def cfg2d(nx=4, ny=3, p=3, per=(True, True)):
    bc = [5 if per[0] else 1, 5 if per[0] else 1, 5 if per[1] else 1, 5 if per[1] else 1]
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=1.0, ly=1.0, riemann=0, bc_type=bc, apply_filter=0,
                filt_cutoff=0, filt_strength=16, filt_alpha=36.0, apply_weno=1, weno_gamma=0.998,
                shock_theta=1.0, shock_threshold=0.8)


def test_2d_is_x_limiter_averaged_with_identity_for_y_constant_data():
    nx, p = 5, 3
    o2 = Oracle(cfg2d(nx=nx, ny=3, p=p))
    o1 = Oracle(cfg1d(nx=nx, p=p))
    n = p + 1
    rng = np.random.default_rng(8)
    lines = rng.uniform(0.2, 1.5, (nx, n))
    lines[2, 2:] = 4.0
    st = np.zeros((3, nx, n, n, 4))  # [j][i][l][k][eq]
    st[..., 0] = lines[None, :, None, :]
    st[..., 3] = 2.5
    out2 = o2.filter(st.reshape(-1)).reshape(st.shape)
    lim1 = o1.filter(state1d(o1, lines)).reshape(nx, 3, n)[:, 0]
    want = 0.5 * (lim1[None, :, None, :] + st[..., 0])
    np.testing.assert_allclose(out2[..., 0], np.broadcast_to(want, out2[..., 0].shape), rtol=0, atol=1e-13)


def np_limit2d(u, nx, ny, n, D, w, per, gamma=0.998, mistake=None):
    """Array-level restatement: u[j, i, l, k, eq] -> limited state.  Planted
    mistakes: "wrap" (periodic neighbours instead of the zero ghost layer),
    "dir" (the y-pass takes its neighbours along x)."""
    U = u.reshape(ny, nx, n, n, 4)
    if mistake == "wrap":
        per = (True, True)
    Dp = [np.linalg.matrix_power(D, s) for s in range(1, n)]

    def beta(lines):  # lines [..., n] -> [...]
        return sum(4.0 ** s * 0.5 * np.einsum("m,...m->...", w, (lines @ Dp[s - 1].T) ** 2) for s in range(1, n))

    def shift(A, di, dj):
        B = np.roll(A, (-dj, -di), axis=(0, 1))
        if not per[0] and di:
            B[:, -1 if di > 0 else 0] = 0.0
        if not per[1] and dj:
            B[-1 if dj > 0 else 0, :] = 0.0
        return B

    def limit(P, C, N):  # lines along the last axis
        b = np.stack([beta(P), beta(C), beta(N)])
        d = np.array([(1 - gamma) / 2, gamma, (1 - gamma) / 2]).reshape(3, *([1] * (b.ndim - 1)))
        c = d / (EPS + b) ** 2
        om = c / c.sum(axis=0)
        r = np.abs(om / d - 1).sum(axis=0) / (abs(1 / d.min() - 1) + 2)
        avg = lambda X: 0.5 * X @ w
        ac = avg(C)[..., None]
        lim = (om[0][..., None] * (P - avg(P)[..., None] + ac) + om[1][..., None] * C
               + om[2][..., None] * (N - avg(N)[..., None] + ac))
        return np.where((r >= 0.8)[..., None], lim, C), r

    X = np.moveaxis(U, 4, 2)  # [j, i, eq, l, k]: rows along k
    xo, rx = limit(shift(X, -1, 0), X, shift(X, 1, 0))
    Y = np.swapaxes(X, 3, 4)  # [j, i, eq, k, l]: columns along l
    if mistake == "dir":
        yo, ry = limit(shift(Y, -1, 0), Y, shift(Y, 1, 0))
    else:
        yo, ry = limit(shift(Y, 0, -1), Y, shift(Y, 0, 1))
    out = 0.5 * (xo + np.swapaxes(yo, 3, 4))
    return np.moveaxis(out, 2, 4).reshape(-1), rx, ry


@pytest.mark.parametrize("per", [(True, True), (False, False), (True, False)])
def test_2d_against_numpy_restatement(per):
    nx, ny, p = 5, 4, 2
    o = Oracle(cfg2d(nx, ny, p, per))
    n = p + 1
    t = o.tables()
    rng = np.random.default_rng(11)
    u = rng.uniform(0.3, 1.2, (ny, nx, n, n, 4))
    u[1:3, 2, :, 1:, :] += 3.0  # jumps inside a column of elements
    u[3, :, 2:, :, 0] -= 0.25
    want, rx, ry = np_limit2d(u, nx, ny, n, t["D"], t["w"], per)
    # keep clear of sensor ties so rounding-level differences cannot flip a line
    assert np.min(np.abs(rx - 0.8)) > 1e-6 and np.min(np.abs(ry - 0.8)) > 1e-6
    assert (rx >= 0.8).any() and (rx < 0.8).any()
    got = o.filter(u.reshape(-1))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mistake", ["wrap", "dir"])
def test_planted_mistake_is_caught(mistake):
    """Restatements with a planted mistake fail the same comparison."""
    nx, ny, p = 5, 4, 2
    o = Oracle(cfg2d(nx, ny, p, (False, False)))
    n = p + 1
    t = o.tables()
    rng = np.random.default_rng(11)
    u = rng.uniform(0.3, 1.2, (ny, nx, n, n, 4))
    u[1:3, 2, :, 1:, :] += 3.0
    u[3, :, 2:, :, 0] -= 0.25
    want, _, _ = np_limit2d(u, nx, ny, n, t["D"], t["w"], (False, False), mistake=mistake)
    got = o.filter(u.reshape(-1))
    assert np.max(np.abs(got - want)) > 1e-3
