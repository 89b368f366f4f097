"""The reference's output files, written from hamsx results (host side).

Layouts and Fortran edit descriptors follow the reference's WRITE statements:

* HAMS format ``OAMASS1..6`` / ``ODAMPING1..6`` (RFORCE_MULTI,
  src/PotentWavForceMulti.f90:214-219, FORMAT 1000 = F7.3,1X,F7.3,1X,6E14.5)
  and ``OEXFOR1..6`` (EFORCE_MULTI, :119, FORMAT 1010 = F7.3,1X,F7.3,1X,6E14.6);
* WAMIT format ``AmssDamp.1`` (RFORCE_MULTI :222-231 after ReorderAmssDamp
  :235-346: per frequency, row-major, ES14.6,2I6,2ES14.6), ``ExcForce.3``
  (EFORCE_MULTI :113-123, 2ES14.6,I6,4ES14.6), ``Motion.4``
  (SolveMotionMulti.f90:113-139), ``Hydrostat.hst`` (HydroStaticMulti.f90:100-115,
  2I6,2X,ES14.6) and the pressure/elevation ``.6p`` files
  (PressureElevationMulti.f90:407-590).

Every function writes text only; the numbers come from the device
(``FrequencySolver``) or any other source of the same arrays.
"""
from __future__ import annotations

import math

import numpy as np

G = 9.80665
PI = 3.141592653589793


# ---------------------------------------------------------------------------
# Fortran edit descriptors
# ---------------------------------------------------------------------------
def _special(x: float, w: int):
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    return None


def _exp_field(e: int) -> str:
    # Ew.d / ESw.d exponent: E+dd for |e| <= 99, +ddd without the E beyond
    return f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"


def fmt_e(x: float, w: int, d: int) -> str:
    """Ew.d: 0.ddddE+xx, right-justified."""
    sp = _special(x, w)
    if sp is not None:
        return sp
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        s = f"{abs(x):.{d - 1}E}"           # D.DDDDE+XX, correctly rounded to d digits
        mant, ex = s.split("E")
        body = "0." + mant.replace(".", "") + _exp_field(int(ex) + 1)
    return (sign + body).rjust(w)


def fmt_es(x: float, w: int, d: int) -> str:
    """ESw.d: D.DDDDDDE+XX, right-justified."""
    sp = _special(x, w)
    if sp is not None:
        return sp
    s = f"{abs(x):.{d}E}"
    mant, ex = s.split("E")
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    return (sign + mant + _exp_field(int(ex))).rjust(w)


def fmt_f(x: float, w: int, d: int) -> str:
    """Fw.d."""
    sp = _special(x, w)
    if sp is not None:
        return sp
    s = f"{x:.{d}f}"
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fmt_i(i: int, w: int) -> str:
    s = str(int(i))
    return s.rjust(w) if len(s) <= w else "*" * w


def atan2d(y: float, x: float) -> float:
    return math.atan2(y, x) * 180.0 / PI


def is_limit(tp: float) -> bool:
    """The zero- / infinite-frequency rows (TP = -1 / 0, ImplementSubs.f90:45-70)."""
    return abs(tp + 1.0) < np.float32(1.0e-6) or abs(tp) < np.float32(1.0e-6)


# ---------------------------------------------------------------------------
# HAMS format
# ---------------------------------------------------------------------------
def hams_force_lines(wk, w1, amas, bdmp, exfc, nbody: int):
    """The lines of OAMASS1..6, ODAMPING1..6 and OEXFOR1..6 for all
    frequencies: wk[k], w1[k]; amas[k], bdmp[k] (6 nbody square); exfc[k]
    (6 nbody x nbeta complex).  Returns {file name: [lines]}."""
    out = {f"{name}{md}.txt": [] for name in ("OAMASS", "ODAMPING", "OEXFOR") for md in range(1, 7)}
    for k in range(len(wk)):
        head = fmt_f(wk[k], 7, 3) + " " + fmt_f(w1[k], 7, 3) + " "
        for m in range(nbody):                 # RFORCE_MULTI per body pair (FILE_M, FILE_N)
            for n in range(nbody):
                for md in range(6):
                    row = 6 * m + md
                    out[f"OAMASS{md + 1}.txt"].append(head + "".join(fmt_e(amas[k][row, 6 * n + c], 14, 5)
                                                                      for c in range(6)))
                    out[f"ODAMPING{md + 1}.txt"].append(head + "".join(fmt_e(bdmp[k][row, 6 * n + c], 14, 5)
                                                                        for c in range(6)))
        ex = np.asarray(exfc[k])
        for b in range(ex.shape[1]):            # EFORCE_MULTI per heading, per body
            for m in range(nbody):
                for md in range(6):
                    v = ex[6 * m + md, b]
                    out[f"OEXFOR{md + 1}.txt"].append(head + fmt_e(v.real, 14, 6) + fmt_e(v.imag, 14, 6))
    return out


# ---------------------------------------------------------------------------
# WAMIT format
# ---------------------------------------------------------------------------
def amssdamp_lines(oufr, w1, tp, amas, bdmp, rho=1000.0):
    """AmssDamp.1 after ReorderAmssDamp: per frequency, rows i then columns
    j of A/rho and B/(rho w) (B column 0 at the frequency limits, where the
    reference writes A only)."""
    lines = []
    for k in range(len(oufr)):
        nd = amas[k].shape[0]
        lim = is_limit(tp[k])
        for i in range(nd):
            for j in range(nd):
                b = 0.0 if lim else bdmp[k][i, j] / (rho * w1[k])
                lines.append(fmt_es(oufr[k], 14, 6) + fmt_i(i + 1, 6) + fmt_i(j + 1, 6)
                             + fmt_es(amas[k][i, j] / rho, 14, 6) + fmt_es(b, 14, 6))
    return lines


def excforce_lines(oufr, tp, heads_deg, exfc, nbody: int, rho=1000.0, amp=1.0):
    """ExcForce.3: per frequency, heading and body, MOD, PHS, NREL, NIMG of
    EXFC / (rho g AMP) with NREL = -Im, NIMG = -Re; PHS is ATAN2D(NIMG, NREL)
    * 180 / pi as the reference writes it.  Limit frequencies are skipped."""
    lines = []
    for k in range(len(oufr)):
        if is_limit(tp[k]):
            continue
        ex = np.asarray(exfc[k])
        for b, beta in enumerate(heads_deg):
            for m in range(nbody):
                for md in range(6):
                    v = ex[6 * m + md, b]
                    rel, img = v.real / (rho * G * amp), v.imag / (rho * G * amp)
                    mod = abs(v) / (rho * G * amp)
                    nrel, nimg = -img, -rel
                    phs = atan2d(nimg, nrel) * 180.0 / PI
                    lines.append(fmt_es(oufr[k], 14, 6) + fmt_es(beta, 14, 6) + fmt_i(6 * m + md + 1, 6)
                                 + "".join(fmt_es(x, 14, 6) for x in (mod, phs, nrel, nimg)))
    return lines


def motion_lines(oufr, tp, heads_deg, dspl, amp=1.0, refl=1.0):
    """Motion.4: per frequency and heading, for every DOF, MAGN, PHS, NREL,
    NIMG of DSPL / (AMP / REFL^MEXP), MEXP = 2 (translations) or 3
    (rotations).  dspl[k]: 6 nbody x nbeta."""
    lines = []
    for k in range(len(oufr)):
        if is_limit(tp[k]):
            continue
        d = np.asarray(dspl[k])
        for b, beta in enumerate(heads_deg):
            for md in range(d.shape[0]):
                mexp = 2 if md % 6 < 3 else 3
                nfac = amp / refl ** mexp
                nrel, nimg = d[md, b].real / nfac, d[md, b].imag / nfac
                magn = math.sqrt(nrel ** 2 + nimg ** 2)
                lines.append(fmt_es(oufr[k], 14, 6) + fmt_es(beta, 14, 6) + fmt_i(md + 1, 6)
                             + "".join(fmt_es(x, 14, 6) for x in (magn, atan2d(nimg, nrel), nrel, nimg)))
    return lines


def hydrostat_lines(crs, rho=1000.0):
    """Hydrostat.hst: CRS / (rho g) on the block diagonal (crs: one 6 x 6 per
    body), every (I, J) of the 6 nbody square."""
    nb = len(crs)
    lines = []
    for i in range(6 * nb):
        for j in range(6 * nb):
            v = crs[i // 6][i % 6, j % 6] / (rho * G) if i // 6 == j // 6 else 0.0
            lines.append(fmt_i(i + 1, 6) + fmt_i(j + 1, 6) + "  " + fmt_es(v, 14, 6))
    return lines


def pressure_elevation_lines(oufr, heads_deg, rad, dif):
    """PressureElevation.6p (one file, radiation and diffraction): per
    frequency, per point the radiation values of every body (6 complex each,
    already scaled as hamsx.wamit.nondim_field does), then per heading the
    diffraction value of every point.  rad[k]: P x 6 nbody; dif[k]: P x nbeta."""
    lines = []
    for k in range(len(oufr)):
        r = np.asarray(rad[k])
        nb = r.shape[1] // 6
        for p in range(r.shape[0]):
            for b in range(nb):
                vals = "".join(fmt_es(x, 14, 6) for v in r[p, 6 * b:6 * b + 6] for x in (v.real, v.imag))
                lead = fmt_es(oufr[k], 14, 6) + fmt_i(p + 1, 10) if b == 0 else " " * 24
                lines.append(lead + vals)
        d = np.asarray(dif[k])
        for bi, beta in enumerate(heads_deg):
            for p in range(d.shape[0]):
                v = d[p, bi]
                lines.append(fmt_es(oufr[k], 14, 6) + fmt_es(beta, 14, 6) + fmt_i(p + 1, 10)
                             + fmt_es(v.real, 14, 6) + fmt_es(v.imag, 14, 6))
    return lines


def incidence_lines(oufr, heads_deg, inc):
    """PressureElevationIncidence.6p: per frequency, heading and point MOD,
    PHS (ATAN2D), REL, IMG of the scaled incident value; inc[k]: P x nbeta."""
    lines = []
    for k in range(len(oufr)):
        d = np.asarray(inc[k])
        for bi, beta in enumerate(heads_deg):
            for p in range(d.shape[0]):
                v = d[p, bi]
                lines.append(fmt_es(oufr[k], 14, 6) + fmt_es(beta, 14, 6) + fmt_i(p + 1, 10)
                             + "".join(fmt_es(x, 14, 6) for x in (abs(v), atan2d(v.imag, v.real), v.real, v.imag)))
    return lines


def write_lines(path: str, lines) -> None:
    with open(path, "w") as f:
        for ln in lines:
            f.write(ln + "\n")
