"""Pins the C oracle to the reference.

* per-function vectors written by the reference's own Fortran routines
  (tests/golden/*.npz, made by tests/golden/make_golden.py from
  oracle/_ref/libhams_ref.so);
* end to end against the reference's shipped regression outputs
  (test/data/config1/expected_output, 5-6 printed digits);
* closed forms and invariants for what the vectors leave open.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

o = oracle.lib()
d = C.c_double


def p(a, t=C.c_double):
    return C.cast(a.ctypes.data, C.POINTER(t))


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def relmax(a, b, floor=1e-300):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


# ---------------------------------------------------------------------------
# special functions (FinGrnExtSubs.f90, InfGreen_Appr.f90)
# ---------------------------------------------------------------------------
def test_special_functions_match_reference():
    z = load("special.npz")
    xs = z["x"]
    for name in ("error", "e1xa", "eix", "struve_h0", "struve_h1", "bessel_j0", "bessel_j1"):
        got = np.array([getattr(o, "orc_" + name)(x) for x in xs])
        assert relmax(got, z[name]) <= 1e-15, name
    for n in range(1, 6):
        got = np.array([o.orc_enxa(n, x) for x in xs])
        assert relmax(got, z["enxa"][n - 1]) <= 1e-15
    for i, x in enumerate(xs):
        a, b = d(), d()
        o.orc_ik01a(x, C.byref(a), C.byref(b))
        assert relmax(np.array([a.value, b.value]), z["ik01a"][i]) <= 1e-15
        q = [d() for _ in range(4)]
        o.orc_jy01b(x, *[C.byref(t) for t in q])
        assert relmax(np.array([t.value for t in q]), z["jy01b"][i]) <= 1e-15


def test_limes_matches_reference():
    z = load("special.npz")
    for s, n, ref in zip(z["limes_seq"], z["limes_n"], z["limes"]):
        s = s[: n + 1].copy()
        ip, kp, err = C.c_int(), C.c_int(), d()
        got = o.orc_limes(int(n), p(s), C.byref(ip), C.byref(kp), C.byref(err))
        assert got == ref


def test_limes_sums_a_geometric_series_exactly():
    # Wynn's epsilon algorithm is exact for a geometric series: the [1/1] step
    # already gives a / (1 - q).
    a, q = 0.7, -0.45
    s = np.array([a * (1 - q ** (k + 1)) / (1 - q) for k in range(12)])
    ip, kp, err = C.c_int(), C.c_int(), d()
    got = o.orc_limes(11, p(s.copy()), C.byref(ip), C.byref(kp), C.byref(err))
    assert abs(got - a / (1 - q)) < 1e-14


def test_special_functions_against_scipy():
    sc = pytest.importorskip("scipy.special")
    xs = np.linspace(0.05, 30.0, 200)
    k0 = np.zeros(len(xs))
    k1 = np.zeros(len(xs))
    for i, x in enumerate(xs):
        a, b = d(), d()
        o.orc_ik01a(x, C.byref(a), C.byref(b))
        k0[i], k1[i] = a.value, b.value
    assert relmax(k0, sc.k0(xs)) < 1e-8 and relmax(k1, sc.k1(xs)) < 1e-8  # asymptotic branch x > 9
    assert relmax(np.array([o.orc_error(x) for x in xs[:50] / 5]), sc.erf(xs[:50] / 5)) < 1e-14
    assert relmax(np.array([o.orc_e1xa(x) for x in xs]), sc.exp1(xs)) < 2e-7  # Zhang-Jin rational fit
    j = np.array([[t.value for t in q] for q in [[d() for _ in range(4)] for _ in xs]])
    for i, x in enumerate(xs):
        q = [d() for _ in range(4)]
        o.orc_jy01b(x, *[C.byref(t) for t in q])
        j[i] = [t.value for t in q]
    for k, f in enumerate((sc.j0, sc.j1, sc.y0, sc.y1)):
        assert np.max(np.abs(j[:, k] - f(xs))) < 1e-7   # polynomial fits, ~1e-8 absolute
    h0 = np.array([o.orc_struve_h0(x) for x in xs])
    assert np.max(np.abs(h0 - sc.struve(0, xs))) < 1e-6


# ---------------------------------------------------------------------------
# dispersion relation
# ---------------------------------------------------------------------------
def test_dispersion_matches_reference_and_closed_form():
    z = load("dispersion.npz")
    for a, w in enumerate(z["w"]):
        for b, h in enumerate(z["h"]):
            wvn = np.zeros(200)
            o.orc_dispersion(w, h, 200, p(wvn))
            np.testing.assert_array_equal(wvn, z["roots"][a, b])
            # k tanh(kh) = w^2/g and u tan(u h) = -w^2/g with the routine's g = 9.807
            v = w * w / 9.807
            assert abs(wvn[0] * np.tanh(wvn[0] * h) - v) < 1e-9 * max(1.0, v)
            u = wvn[1:]
            assert np.max(np.abs(u * np.tan(u * h) + v)) < 1e-7 * max(1.0, v)
            assert np.all(np.diff(u) > 0)


# ---------------------------------------------------------------------------
# Green functions
# ---------------------------------------------------------------------------
def test_green_functions_match_reference():
    rows = load("green.npz")["rows"]
    worst = 0.0
    for r in rows:
        depth, w, flag = r[0], r[1], int(r[2])
        src, fld, ref = r[3:6], r[6:9], r[9:17]
        if flag >= 10:          # V = 0 / V = -1 limits: column 1 is V itself
            v, flag = w, flag - 10
            wvn = np.zeros(200)
        else:
            v = w * w / 9.80665
            wvn = np.zeros(200)
            if depth > 0:
                o.orc_dispersion(w, depth, 200, p(wvn))
        g = np.zeros(8)
        args = (src[0], fld[0], src[1], fld[1], src[2], fld[2], v)
        if depth < 0:
            o.orc_infgreen3d(*args, flag, p(g))
        else:
            o.orc_fingreen3d(*args, p(wvn), 200, depth, flag, p(g))
        e = np.max(np.abs(g - ref)) / max(np.max(np.abs(ref)), 1e-300)
        worst = max(worst, e)
    assert worst < 1e-13


def test_finite_depth_tends_to_deep_water():
    """FINGREEN3D approaches INFGREEN3D as h grows (kh >> 1): checks the
    series, image sums and Rankine bookkeeping against an independent formula."""
    w = 2.0
    v = w * w / 9.80665
    h = 400.0
    wvn = np.zeros(200)
    o.orc_dispersion(w, h, 200, p(wvn))
    src, fld = (1.0, -2.0, -3.0), (25.0, 14.0, -1.5)
    gi, gf = np.zeros(8), np.zeros(8)
    args = (src[0], fld[0], src[1], fld[1], src[2], fld[2], v)
    o.orc_infgreen3d(*args, 0, p(gi))
    o.orc_fingreen3d(*args, p(wvn), 200, h, 0, p(gf))
    assert np.max(np.abs(gi - gf)) / np.max(np.abs(gi)) < 2e-3


# ---------------------------------------------------------------------------
# Newman integrals and panel geometry on the reference's own mesh
# ---------------------------------------------------------------------------
def test_geometry_and_newman_match_reference():
    z = load("newman.npz")
    ne, nn = len(z["ncn"]), len(z["xyz"])
    xyz = np.asfortranarray(z["xyz"])
    ncn = z["ncn"].astype(np.int32)
    ncon = np.asfortranarray(z["ncon"].astype(np.int32))
    cen = np.zeros((ne, 3), order="F")
    area, pnsz = np.zeros(ne), np.zeros(ne)
    dxyz = np.zeros((ne, 6), order="F")
    o.orc_panel_geometry(nn, p(xyz), ne, p(ncn, C.c_int), p(ncon, C.c_int), p(z["xr"].copy()), p(cen), p(area),
                         p(pnsz), p(dxyz))
    for got, ref in ((cen, z["cen"]), (area, z["area"]), (pnsz, z["pnsz"]), (dxyz, z["dxyz"])):
        np.testing.assert_array_equal(got, ref)
    for (i, j), ref in zip(z["pairs"], z["sd"]):
        xv = np.zeros(12)
        for c in range(ncn[j]):
            xv[3 * c:3 * c + 3] = xyz[ncon[j, c] - 1]
        s, dv = d(), np.zeros(3)
        o.orc_sglint(int(ncn[j]), p(xv), p(cen[i].copy()), C.byref(s), p(dv))
        got = np.array([s.value, *dv])
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


def _gauss_panel(xv, xf, order=40):
    """Brute-force Gauss-Legendre integral of 1/r and grad(1/r) over a flat quad
    (bilinear map), for field points away from the panel."""
    g, wts = np.polynomial.legendre.leggauss(order)
    s = 0.0
    dsum = np.zeros(3)
    for a, wa in zip(g, wts):
        for b, wb in zip(g, wts):
            u, v = (a + 1) / 2, (b + 1) / 2
            x = (1 - u) * (1 - v) * xv[0] + u * (1 - v) * xv[1] + u * v * xv[2] + (1 - u) * v * xv[3]
            xu = (1 - v) * (xv[1] - xv[0]) + v * (xv[2] - xv[3])
            xw = (1 - u) * (xv[3] - xv[0]) + u * (xv[2] - xv[1])
            jac = np.linalg.norm(np.cross(xu, xw)) / 4
            r = xf - x
            rn = np.linalg.norm(r)
            s += wa * wb * jac / rn
            dsum += wa * wb * jac * r / rn ** 3
    return s, dsum


def test_newman_integral_against_quadrature():
    xv = np.array([[0.0, 0.0, -1.0], [1.2, 0.1, -1.0], [1.1, 0.9, -1.0], [-0.1, 1.0, -1.0]])
    for xf in ([0.4, 0.5, -2.5], [3.0, -1.0, -0.2], [0.5, 0.5, -1.3]):
        xf = np.array(xf)
        s, dv = d(), np.zeros(3)
        o.orc_sglint(4, p(xv.ravel().copy()), p(xf.copy()), C.byref(s), p(dv))
        sq, dq = _gauss_panel(xv, xf, 60)
        # SIJ = int 1/r dS and DIJ = int (x_f - x)/r^3 dS over the panel
        assert abs(s.value - sq) < 1e-12 * abs(sq)
        assert np.max(np.abs(dv - dq)) < 1e-12 * np.max(np.abs(dq))
    xt = xv[:3]
    for xf in ([0.4, 0.5, -2.5], [3.0, -1.0, -0.2], [0.3, 0.2, -1.1]):
        xf = np.array(xf)
        s, dv = d(), np.zeros(3)
        o.orc_sglint(3, p(xt.ravel().copy()), p(xf.copy()), C.byref(s), p(dv))
        sq, dq = _gauss_panel(np.array([xt[0], xt[1], xt[2], xt[2]]), xf, 80)
        assert abs(s.value - sq) < 1e-12 * abs(sq)
        assert np.max(np.abs(dv - dq)) < 1e-12 * np.max(np.abs(dq))


def test_incident_wave_matches_reference():
    rows = load("incident.npz")["rows"]
    for r in rows:
        depth, w, v, wk, x, y, z, beta = r[:8]
        dp, vi = np.zeros(6), np.zeros(2)
        o.orc_dinp(x, y, z, 1.5, -2.0, beta, v, w, wk, depth, 1.3, p(dp))
        o.orc_vinp(x, y, z, 1.5, -2.0, beta, v, w, wk, depth, 1.3, p(vi))
        assert np.max(np.abs(dp - r[8:14])) <= 1e-14 * max(1.0, np.max(np.abs(r[8:14])))
        assert np.max(np.abs(vi - r[14:16])) <= 1e-14 * max(1.0, np.max(np.abs(r[14:16])))


# ---------------------------------------------------------------------------
# LU (ZGETRF semantics)
# ---------------------------------------------------------------------------
def test_zgetrf_backward_error_and_pivots():
    rng = np.random.default_rng(3)
    n = 60
    A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(np.complex128)
    A = np.asfortranarray(A)
    LU = A.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert o.orc_zgetrf(n, p(LU), p(ipiv, C.c_int)) == 0
    PA = A.copy()
    for k in range(n):
        PA[[k, ipiv[k] - 1]] = PA[[ipiv[k] - 1, k]]
    L = np.tril(LU, -1) + np.eye(n)
    assert np.linalg.norm(PA - L @ np.triu(LU)) / np.linalg.norm(A) < 1e-14
    import scipy.linalg as sl
    lu, piv = sl.lu_factor(A)
    np.testing.assert_array_equal(ipiv - 1, piv)      # same pivot sequence as LAPACK
    b = np.asfortranarray(rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3)))
    x = b.copy(order="F")
    o.orc_zgetrs(n, p(LU), p(ipiv, C.c_int), 3, p(x))
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-13


def test_zgetrf_reports_zero_pivot():
    n = 5
    A = np.asfortranarray(np.eye(n, dtype=np.complex128))
    A[:, 2] = 0
    ipiv = np.zeros(n, dtype=np.int32)
    assert o.orc_zgetrf(n, p(A), p(ipiv, C.c_int)) == 3
