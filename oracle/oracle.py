"""ctypes wrapper of the C oracle (liboracle.so) and of the reference checker
(oracle/_ref/libhams_ref.so, built from the reference's Fortran).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
ORACLE_FMA_SO = os.path.join(HERE, "liboracle_fma.so")
REF_SO = os.path.join(HERE, "_ref", "libhams_ref.so")
REF_FMA_SO = os.path.join(HERE, "_ref", "fma", "libhams_ref.so")
NK = 200
G = 9.80665

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_d = C.c_double
_o = None
_r = None
_rf = None
_variant = {}


def _p(a, t=C.c_double):
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError("array must be contiguous")
    return C.cast(a.ctypes.data, C.POINTER(t))


def lib(fma: bool = False):
    """The oracle; fma=True gives the FMA-contracted build of the same source."""
    global _o
    if fma:
        if "fma" not in _variant:
            _variant["fma"] = _load(ORACLE_FMA_SO)
        return _variant["fma"]
    if _o is None:
        _o = _load(ORACLE_SO)
    return _o


def _load(path):
    if True:
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run `make -C oracle`")
        o = C.CDLL(path)
        for name in ("orc_error", "orc_e1xa", "orc_eix", "orc_struve_h0", "orc_struve_h1", "orc_bessel_j0",
                     "orc_bessel_j1"):
            getattr(o, name).restype = _d
            getattr(o, name).argtypes = [_d]
        o.orc_enxa.restype = _d
        o.orc_enxa.argtypes = [C.c_int, _d]
        o.orc_limes.restype = _d
        o.orc_limes.argtypes = [C.c_int, _dp, _ip, _ip, _dp]
        o.orc_ik01a.argtypes = [_d, _dp, _dp]
        o.orc_jy01b.argtypes = [_d, _dp, _dp, _dp, _dp]
        o.orc_dispersion.argtypes = [_d, _d, C.c_int, _dp]
        o.orc_infgreen3d.argtypes = [_d] * 7 + [C.c_int, _dp]
        o.orc_fingreen3d.argtypes = [_d] * 7 + [_dp, C.c_int, _d, C.c_int, _dp]
        o.orc_sglint.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        o.orc_dinp.argtypes = [_d] * 11 + [_dp]
        o.orc_vinp.argtypes = [_d] * 11 + [_dp]
        o.orc_panel_geometry.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
        o.orc_assemble.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _d, _d, _dp, C.c_int,
                                   _dp, _dp, C.c_int]
        o.orc_assemble_entry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp,
                                         _d, _d, _dp, C.c_int, _dp, _dp]
        o.orc_assemble_rows.argtypes = [C.c_int, _ip, C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, _d, _d,
                                        _dp, C.c_int, _dp, _dp, C.c_int]
        o.orc_rhs.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, C.c_int, _dp, _d, _d, _d, _d, _d, _d, _d, _dp]
        o.orc_zgetrf.restype = C.c_int
        o.orc_zgetrf.argtypes = [C.c_int, _dp, _ip]
        o.orc_zgetrs.argtypes = [C.c_int, _dp, _ip, C.c_int, _dp]
        o.orc_forces.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _ip, C.c_int, _dp, _d, _d, _d, _d, _d, _d,
                                 _d, _d, C.c_int, _dp, _dp, _dp]
        o.orc_solve_frequency_irr.restype = C.c_int
        o.orc_solve_frequency_irr.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip,
                                              C.c_int, _dp, _d, _d, C.c_int, _dp, _d, _d, _d, _d, C.c_int, _dp, _dp,
                                              _dp, C.c_int]
        o.orc_solve_potentials.restype = C.c_int
        o.orc_solve_potentials.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip, _d,
                                           _d, C.c_int, _dp, _d, _d, _d, _dp, C.c_int]
        o.orc_on_body.restype = C.c_int
        o.orc_on_body.argtypes = [_dp, C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp]
        o.orc_field_values.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip, _d, _d,
                                       C.c_int, _dp, _d, _d, _d, _d, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _ip,
                                       C.c_int]
        o.orc_field_values_sym.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip,
                                           C.c_int, _d, _d, C.c_int, _dp, _d, _d, _d, _d, C.c_int, _dp, C.c_int,
                                           _dp, C.c_int, _dp, _ip, C.c_int]
        o.orc_motion.restype = C.c_int
        o.orc_motion.argtypes = [C.c_int, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        o.orc_solve_frequency_sym.restype = C.c_int
        o.orc_solve_frequency_sym.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip,
                                              C.c_int, C.c_int, _dp, _d, _d, C.c_int, _dp, _d, _d, _d, _d, C.c_int,
                                              _dp, _dp, _dp, _dp, C.c_int]
        o.orc_solve_frequency.restype = C.c_int
        o.orc_solve_frequency.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, _dp, C.c_int, _ip, _d,
                                          _d, C.c_int, _dp, _d, _d, _d, _d, C.c_int, _dp, _dp, _dp, C.c_int]
    return o


def ref_available(fma: bool = False) -> bool:
    return os.path.exists(REF_FMA_SO if fma else REF_SO)


def ref(fma: bool = False):
    """The reference's own Fortran routines (built here from /root/reference);
    fma=True: the same sources built with FMA contraction."""
    global _r, _rf
    if fma:
        if _rf is None:
            _rf = _bind_ref(C.CDLL(REF_FMA_SO))
        return _rf
    if _r is None:
        _r = _bind_ref(C.CDLL(REF_SO))
    return _r


def _bind_ref(r):
    if True:
        for name in ("ref_error", "ref_e1xa", "ref_eix", "ref_struve_h0", "ref_struve_h1", "ref_bessel_j0",
                     "ref_bessel_j1"):
            getattr(r, name).restype = _d
            getattr(r, name).argtypes = [_d]
        r.ref_enxa.restype = _d
        r.ref_enxa.argtypes = [C.c_int, _d]
        r.ref_limes.restype = _d
        r.ref_limes.argtypes = [C.c_int, _dp]
        r.ref_ik01a.argtypes = [_d, _dp, _dp]
        r.ref_jy01b.argtypes = [_d, _dp, _dp, _dp, _dp]
        r.ref_dispersion.argtypes = [_d, _d, C.c_int, _dp]
        r.ref_wave_properties.argtypes = [C.c_int, _d, _d, _dp, _dp, _dp, _dp]
        r.ref_infgreen3d.argtypes = [_d] * 7 + [C.c_int, _dp]
        r.ref_fingreen3d.argtypes = [_d] * 7 + [_dp, C.c_int, _d, C.c_int, _dp]
        r.ref_set_mesh.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _dp]
        r.ref_calgreen.restype = _d
        r.ref_calgreen.argtypes = [C.c_int, C.c_int, _dp, _d, _dp, C.c_int, _d]
        r.ref_sglint.argtypes = [C.c_int, C.c_int, _dp, _dp]
        r.ref_sglint2.argtypes = [_dp, C.c_int, _dp]
        r.ref_dinp.argtypes = [_d] * 11 + [_dp]
        r.ref_vinp.argtypes = [_d] * 11 + [_dp]
        r.ref_geometry.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    return r


# ---------------------------------------------------------------------------
# Convenience wrappers over combined meshes (hamsx.mesh.Combined layout)
# ---------------------------------------------------------------------------
def wave(w1: float, h: float):
    """CalWaveProperts for a frequency in rad/s: v, wk, wvn (w1 = 0 / -1: the
    zero- / infinite-frequency limits, v = wk = w1)."""
    wvn = np.zeros(NK)
    if w1 in (0.0, -1.0):
        return w1, w1, wvn
    v = w1 * w1 / G
    if h > 0:
        lib().orc_dispersion(w1, h, NK, _p(wvn))
        wk = wvn[0]
    else:
        wk = v
    return v, wk, wvn


def geometry(mesh, xr):
    ne = mesh.nelem
    xyz = np.asfortranarray(mesh.xyz, dtype=np.float64)
    ncn = np.ascontiguousarray(mesh.ncn, dtype=np.int32)
    ncon = np.asfortranarray(mesh.ncon, dtype=np.int32)
    cen = np.zeros((ne, 3), order="F")
    area, pnsz = np.zeros(ne), np.zeros(ne)
    dxyz = np.zeros((ne, 6), order="F")
    xr = np.asarray(xr, dtype=np.float64)
    lib().orc_panel_geometry(mesh.nnode, _p(xyz), ne, _p(ncn, C.c_int), _p(ncon, C.c_int), _p(xr), _p(cen),
                             _p(area), _p(pnsz), _p(dxyz))
    return cen, area, pnsz, dxyz


def assemble(m, w1: float, h: float, nthreads: int = 8, fma: bool = False):
    v, wk, wvn = wave(w1, h)
    n = m.nelem
    A = np.zeros((n, n), dtype=np.complex128, order="F")
    S = np.zeros((n, n), dtype=np.complex128, order="F")
    lib(fma).orc_assemble(n, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int), _p(m.cen), _p(m.area),
                       _p(m.pnsz), _p(m.dxyz), v, h, _p(wvn), NK, _p(A), _p(S),
                       nthreads)
    return A, S


def assemble_rows(m, rows, w1: float, h: float, nthreads: int = 8):
    """Rows `rows` (0-based) of A and S, shape (len(rows), N)."""
    v, wk, wvn = wave(w1, h)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    nr, n = len(rows), m.nelem
    A = np.zeros((nr, n), dtype=np.complex128, order="F")
    S = np.zeros((nr, n), dtype=np.complex128, order="F")
    lib().orc_assemble_rows(nr, _p(rows, C.c_int), n, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int),
                            _p(m.ncon, C.c_int), _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), v, h, _p(wvn), NK,
                            _p(A), _p(S), nthreads)
    return A, S


def assemble_entry(m, i: int, j: int, w1: float, h: float):
    v, wk, wvn = wave(w1, h)
    a, s = np.zeros(2), np.zeros(2)
    lib().orc_assemble_entry(i, j, m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int),
                             _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), v, h, _p(wvn), NK, _p(a), _p(s))
    return complex(a[0], a[1]), complex(s[0], s[1])


def solve_frequency(m, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0, isol=1,
                    nthreads: int = 8, fma: bool = False):
    nd = 6 * m.nbody
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    amas = np.zeros((nd, nd), order="F")
    bdmp = np.zeros((nd, nd), order="F")
    exfc = np.zeros((nd, max(1, len(beta))), dtype=np.complex128, order="F")
    info = lib(fma).orc_solve_frequency(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int),
                                     _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody,
                                     _p(m.body_start, C.c_int), w1, h, len(beta), _p(beta), amp, xw[0], xw[1], rho,
                                     isol, _p(amas), _p(bdmp), _p(exfc), nthreads)
    return amas, bdmp, exfc[:, :len(beta)], info


def solve_frequency_irr(m, lid_cen, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0, isol=1,
                        nthreads: int = 8, fma: bool = False):
    """Irregular-frequency removal: lid_cen (L x 3) are the lid centroids."""
    nd = 6 * m.nbody
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    lc = np.asfortranarray(lid_cen, dtype=np.float64)
    amas = np.zeros((nd, nd), order="F")
    bdmp = np.zeros((nd, nd), order="F")
    exfc = np.zeros((nd, max(1, len(beta))), dtype=np.complex128, order="F")
    info = lib(fma).orc_solve_frequency_irr(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int),
                                         _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody,
                                         _p(m.body_start, C.c_int), lc.shape[0], _p(lc), w1, h, len(beta), _p(beta),
                                         amp, xw[0], xw[1], rho, isol, _p(amas), _p(bdmp), _p(exfc), nthreads)
    return amas, bdmp, exfc[:, :len(beta)], info


def solve_potentials(m, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), nthreads: int = 8,
                     fma: bool = False):
    """The solved potentials MXPOT (N x (6 nbody + nbeta)) of one frequency."""
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    nr = 6 * m.nbody + len(beta)
    phi = np.zeros((m.nelem, nr), dtype=np.complex128, order="F")
    info = lib(fma).orc_solve_potentials(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int),
                                      _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody,
                                      _p(m.body_start, C.c_int), w1, h, len(beta), _p(beta), amp, xw[0], xw[1],
                                      _p(phi), nthreads)
    return phi, info


def field_values(m, phi, xfp, w1: float, h: float, headings_rad, kind=0, amp=1.0, xw=(0.0, 0.0), rho=1000.0,
                 isol=1, nthreads: int = 8, fma: bool = False):
    """CalPotentialMulti / CalPressureMulti / CalElevationMulti (kind 0/1/2)
    at the field points xfp (P x 3) for every mode column of phi.  Returns
    (values P x (6 nbody + nbeta), onbody P)."""
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    xf = np.asfortranarray(xfp, dtype=np.float64).reshape(-1, 3)
    nfp, nr = xf.shape[0], 6 * m.nbody + len(beta)
    ph = np.asfortranarray(phi, dtype=np.complex128)
    assert ph.shape == (m.nelem, nr)
    out = np.zeros((nfp, nr), dtype=np.complex128, order="F")
    ob = np.zeros(nfp, dtype=np.int32)
    lib(fma).orc_field_values(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int), _p(m.cen),
                              _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody, _p(m.body_start, C.c_int), w1, h,
                              len(beta), _p(beta), amp, xw[0], xw[1], rho, isol, _p(ph), nfp, _p(xf), kind, _p(out),
                              _p(ob, C.c_int), nthreads)
    return out, ob.astype(bool)


def field_values_sym(m, isx: int, phi2, xfp, w1: float, h: float, headings_rad, kind=0, amp=1.0, xw=(0.0, 0.0),
                     rho=1000.0, isol=1, nthreads: int = 8, fma: bool = False):
    """field_values on the half mesh m with one symmetry plane (isx = 1: x = 0,
    else y = 0); phi2 (nelem x nr x 2) the symmetric solve's potentials."""
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    xf = np.asfortranarray(xfp, dtype=np.float64).reshape(-1, 3)
    nfp, nr = xf.shape[0], 6 * m.nbody + len(beta)
    ph = np.asfortranarray(phi2, dtype=np.complex128)
    assert ph.shape == (m.nelem, nr, 2)
    out = np.zeros((nfp, nr), dtype=np.complex128, order="F")
    ob = np.zeros(nfp, dtype=np.int32)
    lib(fma).orc_field_values_sym(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int), _p(m.cen),
                                  _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody, _p(m.body_start, C.c_int), int(isx),
                                  w1, h, len(beta), _p(beta), amp, xw[0], xw[1], rho, isol, _p(ph), nfp, _p(xf), kind,
                                  _p(out), _p(ob, C.c_int), nthreads)
    return out, ob.astype(bool)


def _blocks(mats):
    """nbody (6 x 6) matrices -> the nbody x 36 column-major block array."""
    return np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=np.float64).T.ravel() for m in mats]))


def motion(nbody, w1, amas, bdmp, exfc, mass, blnr, bqdr, crs, kstf, fma: bool = False):
    """SolveMotionMulti for every heading column of exfc (6 nbody x nbeta):
    each body matrix argument is a list of nbody 6 x 6 arrays.  Returns
    (dspl 6 nbody x nbeta, iterations per heading)."""
    nd = 6 * nbody
    am = np.asfortranarray(amas, dtype=np.float64)
    bd = np.asfortranarray(bdmp, dtype=np.float64)
    blk = [_blocks(m) for m in (mass, blnr, bqdr, crs, kstf)]
    ex = np.asarray(exfc, dtype=np.complex128).reshape(nd, -1)
    out = np.zeros_like(ex)
    its = []
    for b in range(ex.shape[1]):
        e = np.ascontiguousarray(ex[:, b])
        d = np.zeros(nd, dtype=np.complex128)
        its.append(lib(fma).orc_motion(nbody, w1, _p(am), _p(bd), _p(e), *[_p(x) for x in blk], _p(d)))
        out[:, b] = d
    return out, its


def forces(m, phi, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0, isol=1):
    """RFORCE_MULTI / EFORCE_MULTI from the potentials phi (N x (6 nbody + nbeta))."""
    v, wk, _ = wave(w1, h)
    nd = 6 * m.nbody
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    ph = np.asfortranarray(phi, dtype=np.complex128)
    amas = np.zeros((nd, nd), order="F")
    bdmp = np.zeros((nd, nd), order="F")
    exfc = np.zeros((nd, max(1, len(beta))), dtype=np.complex128, order="F")
    lib().orc_forces(m.nelem, _p(ph), _p(m.cen), _p(m.area), _p(m.dxyz), m.nbody, _p(m.body_start, C.c_int),
                     len(beta), _p(beta), v, w1, wk, h, amp, xw[0], xw[1], rho, isol, _p(amas), _p(bdmp), _p(exfc))
    return amas, bdmp, exfc[:, :len(beta)]


def solve_frequency_sym(m, isx: int, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0,
                        isol=1, nthreads: int = 8, potentials: bool = False, fma: bool = False, lid_cen=None):
    """NSYS = 2 on the half mesh m (isx = 1: mirrored in x, else in y);
    lid_cen (L x 3): the half lid's points (IRSP = 1), or None."""
    lc = np.asfortranarray(np.zeros((0, 3)) if lid_cen is None else lid_cen, dtype=np.float64)
    nd = 6 * m.nbody
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    amas = np.zeros((nd, nd), order="F")
    bdmp = np.zeros((nd, nd), order="F")
    exfc = np.zeros((nd, max(1, len(beta))), dtype=np.complex128, order="F")
    phi2 = np.zeros((m.nelem, nd + len(beta), 2), dtype=np.complex128, order="F")
    info = lib(fma).orc_solve_frequency_sym(m.nelem, m.nnode, _p(m.xyz), _p(m.ncn, C.c_int), _p(m.ncon, C.c_int),
                                         _p(m.cen), _p(m.area), _p(m.pnsz), _p(m.dxyz), m.nbody,
                                         _p(m.body_start, C.c_int), int(isx), lc.shape[0],
                                         _p(lc) if lc.shape[0] else None, w1, h, len(beta), _p(beta), amp, xw[0],
                                         xw[1], rho, isol, _p(amas), _p(bdmp), _p(exfc), _p(phi2), nthreads)
    out = (amas, bdmp, exfc[:, :len(beta)], info)
    return out + (phi2,) if potentials else out


# ---------------------------------------------------------------------------
# Rounding samples and solves from assembled matrices (parity-gate helpers)
# ---------------------------------------------------------------------------
def perturbed(m, kind: int):
    """A copy of the combined mesh m whose centroids are moved by one unit in
    the last place: kind 1 moves z towards -inf, kind 2 moves x and y towards
    +inf, kind 3 moves x and y towards -inf and z towards +inf.  The reference's answer on such a mesh differs from the unperturbed
    one only by the rounding sensitivity of its algorithm (the geometry moves
    by ~1e-16 relative), so each kind is one more sample of that spread."""
    import copy
    q = copy.copy(m)
    cen = np.array(m.cen, dtype=np.float64, order="F", copy=True)
    if kind == 1:
        cen[:, 2] = np.nextafter(cen[:, 2], -np.inf)
    elif kind == 2:
        cen[:, 0] = np.nextafter(cen[:, 0], np.inf)
        cen[:, 1] = np.nextafter(cen[:, 1], np.inf)
    elif kind == 3:
        cen[:, 0] = np.nextafter(cen[:, 0], -np.inf)
        cen[:, 1] = np.nextafter(cen[:, 1], -np.inf)
        cen[:, 2] = np.nextafter(cen[:, 2], np.inf)
    else:
        raise ValueError("kind must be 1, 2 or 3")
    q.cen = np.asfortranarray(cen)
    return q


def rhs(m, S, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0)):
    """ASSB_RBC_MULTI + ASSB_DBC_MULTI right-hand sides (N x (6 nbody + nbeta))
    from the single-layer matrix S."""
    v, wk, _ = wave(w1, h)
    beta = np.ascontiguousarray(headings_rad, dtype=np.float64)
    Sf = np.asfortranarray(S, dtype=np.complex128)
    out = np.zeros((m.nelem, 6 * m.nbody + len(beta)), dtype=np.complex128, order="F")
    lib().orc_rhs(m.nelem, _p(Sf), _p(m.cen), _p(m.dxyz), m.nbody, _p(m.body_start, C.c_int), len(beta), _p(beta),
                  v, w1, wk, h, amp, xw[0], xw[1], _p(out))
    return out


def solve_assembled(m, A, S, w1: float, h: float, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0, isol=1):
    """The frequency solve from oracle-assembled A and S: right-hand sides,
    LAPACK ZGETRF/ZGETRS (scipy's LAPACK; the reference calls MKL's at
    AssbMatxMulti.f90:112, :279, :346) and the force sums.  Returns
    (amas, bdmp, exfc, potentials)."""
    import scipy.linalg as sl
    B = rhs(m, S, w1, h, headings_rad, amp, xw)
    X = sl.lu_solve(sl.lu_factor(np.asfortranarray(A), check_finite=False), B, check_finite=False)
    am, bd, ex = forces(m, X, w1, h, headings_rad, amp, xw, rho, isol)
    return am, bd, ex, X


def ref_calgreen(m, w1: float, h: float, nthreads: int = 8) -> float:
    """Runs the reference's own pair loop CALGREEN_MULTI (its OpenMP loop over
    field panels, CalGreenFuncMulti.f90:112-183, built as shipped) over every
    pair of mesh m at (w1, h) on `nthreads` threads; returns its checksum.
    For timing the reference on the host (the CPU baseline)."""
    r = ref()
    v, wk, wvn = wave(w1, h)
    r.ref_set_mesh(m.nnode, _p(np.asfortranarray(m.xyz)), m.nelem, _p(np.ascontiguousarray(m.ncn), C.c_int),
                   _p(np.asfortranarray(m.ncon), C.c_int), _p(np.asfortranarray(m.cen)))
    return r.ref_calgreen(nthreads, m.nelem, _p(np.ascontiguousarray(m.pnsz)), v, _p(wvn), NK, h)
