"""Pins of the oracle's (and the product's) setup tables against closed forms.

The fixture tests/golden/gll_tables.json is written by
tests/golden/make_gll_fixture.py with mpmath at 50 digits.  Additional
identities fix what a plausible mistake would break: a transposed D fails
D.x^m = m x^(m-1); a sign slip in dL fails sum(w dL) = -1; a wrong filter
basis fails F P_j = sigma_j P_j.
"""
import json
import os

import numpy as np
import pytest

from pyoracle import Oracle

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gll_tables.json")))


def tables(p, **kw):
    cfg = dict(dims=1, nx=4, p=p, bc_type=[5, 5, 5, 5])
    cfg.update(kw)
    return Oracle(cfg).tables()


@pytest.mark.parametrize("p", range(1, 9))
def test_gll_nodes_weights_vs_mpmath(p):
    t = tables(p)
    ref = FIX[str(p)]
    assert np.max(np.abs(t["csi"] - ref["csi"])) <= 4e-16
    assert np.max(np.abs(t["w"] - ref["w"])) <= 1e-15
    assert abs(t["w"].sum() - 2.0) <= 1e-14  # reference_element.hpp:180-184 check_ws
    assert np.allclose(t["csi"], -t["csi"][::-1], atol=6e-16, rtol=0)


@pytest.mark.parametrize("p", range(1, 9))
def test_diff_matrix(p):
    t = tables(p)
    D, x, w = t["D"], t["csi"], t["w"]
    ref = np.array(FIX[str(p)]["D"])
    assert np.max(np.abs(D - ref)) <= 5e-14 * max(1, p * p)
    # exact on polynomials of degree <= p
    for m in range(p + 1):
        d = D @ x ** m
        exact = m * x ** (m - 1) if m > 0 else np.zeros_like(x)
        assert np.max(np.abs(d - exact)) <= 1e-12
    # SBP: M D + D^T M = diag(-1, 0, ..., 0, 1)  (tests/sbp.cpp:27-38)
    M = np.diag(w)
    B = np.zeros_like(D)
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.max(np.abs(M @ D + D.T @ M - B)) <= 1e-13
    assert abs(D[0, 0] + p * (p + 1) / 4.0) <= 1e-12


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("c,key", [(0.0, "c0"), (0.01, "c001")])
def test_vcjh_correction(p, c, key):
    t = tables(p, c_vcjh=c)
    ref = FIX[str(p)]
    assert np.max(np.abs(t["dL"] - ref["dL_" + key])) <= 1e-12
    assert np.max(np.abs(t["dR"] - ref["dR_" + key])) <= 1e-12
    w = t["w"]
    # g_R(1) - g_R(-1) = 1, g_L(1) - g_L(-1) = -1, integrated exactly by GLL (degree <= 2p-1)
    assert abs(w @ t["dR"] - 1.0) <= 1e-13
    assert abs(w @ t["dL"] + 1.0) <= 1e-13
    assert np.max(np.abs(t["dL"] + t["dR"][::-1])) <= 1e-12  # mirror symmetry


@pytest.mark.parametrize("p", range(1, 9))
def test_filter_matrix(p):
    cfg = dict(filt_cutoff=0, filt_strength=16, filt_alpha=36.0)
    t = tables(p, **cfg)
    assert np.max(np.abs(t["filt"] - np.array(FIX[str(p)]["filt_nc0_s16_a36"]))) <= 1e-12
    # F P_j(x) = sigma_j P_j(x) at the nodes
    x = t["csi"]
    for j in range(p + 1):
        Pj = np.polynomial.legendre.legval(x, [0] * j + [1])
        sig = np.exp(-36.0 * (j / p) ** 16)
        assert np.max(np.abs(t["filt"] @ Pj - sig * Pj)) <= 1e-12
    if p >= 2:
        t2 = tables(p, filt_cutoff=1, filt_strength=32, filt_alpha=36.0)
        assert np.max(np.abs(t2["filt"] - np.array(FIX[str(p)]["filt_nc1_s32_a36"]))) <= 1e-12
    t0 = tables(p, filt_cutoff=0, filt_strength=16, filt_alpha=0.0)
    assert np.max(np.abs(t0["filt"] - np.eye(p + 1))) <= 1e-12
