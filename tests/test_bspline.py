"""Greville collocation B-spline operators: oracle and product setup.

The oracle (scipy-based) and the product's host-side de Boor setup
(sz_bsplineop_alloc in libsuzerain_amd.so; pure host code, no GPU needed)
are both pinned against the reference's golden values:
* tests/test_bsmbsm.cpp:683-760: Mathematica k=6 M, D1, D2 bands on
  breakpoints {0, 1/4, 3/4, 1, 5/4, 2} (exact rationals);
* tests/test_bsplineop.cpp:610-741: exact cubic D_T0 band (kl = ku = 2),
  cubic apply goldens, almost-dense cubic D_T0/D_T1, septic (k=8) apply
  goldens with integer results;
plus, at the workload's stretched k=8 grid, the two implementations against
each other and the trimmed bandwidth kl = ku = k-2.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import oracle.bspline as obs
from conftest import golden

EPS = np.finfo(float).eps
F = lambda v: np.array([float(Fraction(s)) for s in v])


def product_op(k, bp, nderiv):
    from suzerain_amd._lib import load
    lib = load()
    bp = np.ascontiguousarray(bp, dtype=np.float64)
    info = C.c_int(0)
    w = lib.sz_bsplineop_alloc(k, len(bp), bp.ctypes.data_as(C.POINTER(C.c_double)),
                               nderiv, C.byref(info))
    assert w and info.value == 0, info.value
    n, ld, mku = lib.sz_bsplineop_n(w), lib.sz_bsplineop_ld(w), lib.sz_bsplineop_max_ku(w)
    out = {"n": n, "ld": ld, "max_ku": mku, "kl": [], "ku": [], "D": [],
           "points": np.array([lib.sz_bsplineop_collocation_point(w, i)
                               for i in range(n)])}
    for d in range(nderiv + 1):
        kl, ku = lib.sz_bsplineop_kl(w, d), lib.sz_bsplineop_ku(w, d)
        p = lib.sz_bsplineop_D_T(w, d)
        band = np.ctypeslib.as_array(p, shape=(n * ld,)).copy().reshape(n, ld)
        D = np.zeros((n, n))
        for i in range(n):              # D(i, j) = D_T(j, i) at i*ld+mku+j-i
            for j in range(max(0, i - ku), min(n, i + kl + 1)):
                D[i, j] = band[i, mku + j - i]
        out["kl"].append(kl)
        out["ku"].append(ku)
        out["D"].append(D)
    lib.sz_bsplineop_free(w)
    return out


def golden_k6():
    g = golden("bsmbsm_golden.json")["solve"]
    n = g["n"]
    dense = {}
    for name in ("M", "D1", "D2"):
        ab = F(g[name]).reshape(n, 9)          # column-major band, ld 9, ku 4
        dense[name] = obs.dense_of(ab, n, 4, 4)
    return [float(Fraction(b)) for b in g["breakpoints"]], dense


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_k6_mathematica_bands(impl):
    bp, dense = golden_k6()
    op = (obs.collocation(6, bp, 2) if impl == "oracle" else product_op(6, bp, 2))
    for d, name in enumerate(("M", "D1", "D2")):
        ref = dense[name]
        np.testing.assert_allclose(op["D"][d], ref, rtol=0,
                                   atol=1e3 * EPS * np.abs(ref).max())
    # k=6 trims to bandwidth 4 (tests/test_bsmbsm.cpp:801-811 uses kl=ku=4)
    assert max(op["kl"]) == 4 and max(op["ku"]) == 4


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_cubic_golden_band_and_apply(impl):
    g = golden("bsplineop_golden.json")["cubic"]
    bp = F(g["breakpoints"])
    op = obs.collocation(4, bp, 2) if impl == "oracle" else product_op(4, bp, 2)
    assert op["kl"][0] == g["D_T0_kl"] and op["ku"][0] == g["D_T0_ku"]
    n = op["n"]
    good = F(g["good_D_T0"]).reshape(n, g["D_T0_ld"])    # transposed band
    for i in range(n):
        for j in range(max(0, i - 2), min(n, i + 3)):
            assert abs(op["D"][0][i, j] - good[i, 2 + j - i]) < 1e3 * EPS
    x = F(g["apply0_in"]).reshape(g["apply_nrhs"], n)
    y = (op["D"][0] @ x.T).T
    np.testing.assert_allclose(y.reshape(-1), F(g["apply0_out"]), rtol=1e-12)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_almost_dense_cubic(impl):
    g = golden("bsplineop_golden.json")["cubic_dense"]
    bp = F(g["breakpoints"])
    op = obs.collocation(4, bp, 2) if impl == "oracle" else product_op(4, bp, 2)
    n = op["n"]
    for d, key in ((0, "good_D_T0"), (1, "good_D_T1")):
        good = F(g[key]).reshape(n, g["ld"])
        for i in range(n):
            for j in range(n):
                assert abs(op["D"][d][i, j] - good[i, g["ku"] + j - i]) < 1e3 * EPS


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_septic_apply_integers(impl):
    g = golden("bsplineop_golden.json")["septic"]
    bp = F(g["breakpoints"])
    op = obs.collocation(8, bp, 2) if impl == "oracle" else product_op(8, bp, 2)
    n = op["n"]
    for case in g["cases"]:
        x = F(case["in"]).reshape(case["nrhs"], n)
        y = case["alpha"] * (op["D"][case["nderiv"]] @ x.T).T
        np.testing.assert_allclose(y.reshape(-1), F(case["out"]), rtol=1e3 * EPS)


def test_breakpoints_product_vs_oracle():
    from suzerain_amd._lib import load
    lib = load()
    for ndof, k, h in ((96, 8, 3.0), (48, 6, 0.0), (50, 8, -2.0)):
        out = np.zeros(ndof - k + 2)
        nb = lib.sz_bspline_breakpoints(ndof, k, -1.0, 1.0, h,
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
        assert nb == ndof - k + 2
        np.testing.assert_array_equal(out, obs.breakpoints(ndof, k, -1.0, 1.0, h))


@pytest.mark.parametrize("ndof,k", [(96, 8), (48, 6), (192, 8)])
def test_stretched_grid_product_vs_oracle(ndof, k):
    bp = obs.breakpoints(ndof, k, -1.0, 1.0, 3.0)
    o = obs.collocation(k, bp, k - 2)
    p = product_op(k, bp, k - 2)
    assert o["n"] == p["n"] == ndof
    assert p["kl"] == o["kl"] and p["ku"] == o["ku"]
    assert max(p["kl"]) == k - 2 and max(p["ku"]) == k - 2
    np.testing.assert_allclose(p["points"], o["points"], rtol=0, atol=4 * EPS)
    for d in range(k - 1):
        scale = np.abs(o["D"][d]).max()
        np.testing.assert_allclose(p["D"][d], o["D"][d], rtol=0,
                                   atol=5e-13 * scale)
