"""The output writers (hamsx.writers) against the text of the reference's
shipped output files (tests/golden/writers.npz): the numbers are parsed back
into the arrays the solver returns, written again, and the text must come
out the same -- character for character where the file prints the values
themselves, and to the last printed digit where it prints quantities derived
from them (MOD, PHS)."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from hamsx import writers as W

Z = np.load(os.path.join(GOLDEN, "writers.npz"))


def cols(line, widths):
    out, pos = [], 0
    for w in widths:
        out.append(line[pos:pos + w])
        pos += w
    return out


@pytest.mark.parametrize("x,w,d,expect", [
    (132840.0, 14, 5, "   0.13284E+06"),
    (-0.47814, 14, 5, "  -0.47814E+00"),
    (0.0, 14, 5, "   0.00000E+00"),
    (0.999996, 14, 5, "   0.10000E+01"),    # rounding carries into the exponent
    (float("nan"), 14, 6, "           NaN"),
])
def test_fortran_e(x, w, d, expect):
    assert W.fmt_e(x, w, d) == expect


@pytest.mark.parametrize("x,w,d,expect", [
    (92.60277, 14, 6, "  9.260277E+01"),
    (-7.782595e-05, 14, 6, " -7.782595E-05"),
    (0.0, 14, 6, "  0.000000E+00"),
    (9.9999996, 14, 6, "  1.000000E+01"),
])
def test_fortran_es(x, w, d, expect):
    assert W.fmt_es(x, w, d) == expect


def test_hams_tables_round_trip():
    """config1: 3 frequencies, 3 bodies, the OAMASS/ODAMPING/OEXFOR files
    rebuilt from their own numbers."""
    nb = 3
    am = {md: Z[f"config1_OAMASS{md}"] for md in range(1, 7)}
    nfreq = len(am[1]) // (nb * nb)
    nbeta = len(Z["config1_OEXFOR1"]) // (nfreq * nb)
    wk, w1 = [], []
    amas = np.zeros((nfreq, 6 * nb, 6 * nb))
    bdmp = np.zeros_like(amas)
    exfc = np.zeros((nfreq, 6 * nb, nbeta), dtype=np.complex128)
    for md in range(1, 7):
        for kind, arr in (("OAMASS", amas), ("ODAMPING", bdmp)):
            for q, line in enumerate(Z[f"config1_{kind}{md}"]):
                k, mn = divmod(q, nb * nb)
                m, n = divmod(mn, nb)
                f = cols(line, [7, 1, 7, 1] + [14] * 6)
                arr[k, 6 * m + md - 1, 6 * n:6 * n + 6] = [float(t) for t in f[4:]]
                if md == 1 and kind == "OAMASS" and mn == 0:
                    wk.append(float(f[0]))
                    w1.append(float(f[2]))
        for q, line in enumerate(Z[f"config1_OEXFOR{md}"]):
            k, rest = divmod(q, nbeta * nb)
            b, m = divmod(rest, nb)
            f = cols(line, [7, 1, 7, 1, 14, 14])
            exfc[k, 6 * m + md - 1, b] = complex(float(f[4]), float(f[5]))
    out = W.hams_force_lines(wk, w1, amas, bdmp, exfc, nb)
    for name, lines in out.items():
        assert lines == list(Z["config1_" + name.replace(".txt", "")]), name


def test_amssdamp_round_trip():
    """example2's AmssDamp.1 (60 DOF), first six rows of the first
    frequency: A / rho and B / (rho w) re-scaled and written back."""
    lines = list(Z["ex2_AmssDamp_1"])
    nd = 60
    oufr = float(lines[0][:14])
    w1 = 2 * math.pi / oufr                      # OUFT = 4: the period
    amas = np.zeros((nd, nd))
    bdmp = np.zeros((nd, nd))
    for line in lines:
        f = cols(line, [14, 6, 6, 14, 14])
        i, j = int(f[1]) - 1, int(f[2]) - 1
        amas[i, j] = float(f[3]) * 1000.0
        bdmp[i, j] = float(f[4]) * 1000.0 * w1
    got = W.amssdamp_lines([oufr], [w1], [oufr], [amas], [bdmp])[:len(lines)]
    assert got == lines


def test_hydrostat_round_trip():
    """Hydrostat.hst of example2 from its Hydrostatic_*.in restoring matrix
    (rows 1..6 read as 1, 2, ..., 6 on every row)."""
    crs = [np.tile(np.arange(1.0, 7.0), (6, 1)) for _ in range(10)]
    got = W.hydrostat_lines(crs)[:240]
    assert got == list(Z["ex2_Hydrostat_hst"])


def _close_lines(got, ref, exact_cols, widths):
    """Same layout; the columns in exact_cols identical, the others within
    one unit of the last printed digit."""
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert len(g) == len(r), (g, r)
        fg, fr = cols(g, widths), cols(r, widths)
        for c, (a, b) in enumerate(zip(fg, fr)):
            if c in exact_cols:
                assert a == b, (g, r)
            elif a != b:
                x, y = float(a), float(b)
                assert abs(x - y) <= 1.5e-6 * max(abs(y), 1e-30) * 10, (g, r)


def test_excforce_round_trip():
    """ExcForce.3: EXFC rebuilt from NREL / NIMG (= -Im, -Re of EXFC /
    (rho g A)); MOD and PHS recomputed."""
    ref = list(Z["ex2_ExcForce_3"])
    nb = 10
    oufr = float(ref[0][:14])
    ex = np.zeros((6 * nb, 1), dtype=np.complex128)
    for line in ref[:6 * nb]:
        f = cols(line, [14, 14, 6, 14, 14, 14, 14])
        md = int(f[2]) - 1
        nrel, nimg = float(f[5]), float(f[6])
        ex[md, 0] = complex(-nimg, -nrel) * 1000.0 * W.G
    got = W.excforce_lines([oufr], [oufr], [0.0], [ex], nb)
    _close_lines(got, ref[:6 * nb], {0, 1, 2, 5, 6}, [14, 14, 6, 14, 14, 14, 14])


def test_motion_round_trip():
    ref = list(Z["ex2_Motion_4"])
    oufr = float(ref[0][:14])
    d = np.zeros((60, 1), dtype=np.complex128)
    for line in ref[:60]:
        f = cols(line, [14, 14, 6, 14, 14, 14, 14])
        d[int(f[2]) - 1, 0] = complex(float(f[5]), float(f[6]))
    got = W.motion_lines([oufr], [oufr], [0.0], [d])
    _close_lines(got, ref[:60], {0, 1, 2, 5, 6}, [14, 14, 6, 14, 14, 14, 14])


def test_pressure_elevation_round_trip():
    """PressureElevation.6p: ten points' radiation lines (10 bodies x 6
    modes each) and heading 0's diffraction lines for all 100 points."""
    rad_lines, dif_lines = list(Z["ex2_6p_rad"]), list(Z["ex2_6p_dif"])
    oufr = float(rad_lines[0][:14])
    rad = np.zeros((10, 60), dtype=np.complex128)
    for q, line in enumerate(rad_lines):
        p, b = divmod(q, 10)
        v = [float(t) for t in cols(line[24:], [14] * 12)]
        rad[p, 6 * b:6 * b + 6] = np.array(v[0::2]) + 1j * np.array(v[1::2])
    dif = np.zeros((100, 1), dtype=np.complex128)
    for p, line in enumerate(dif_lines):
        f = cols(line, [14, 14, 10, 14, 14])
        dif[p, 0] = complex(float(f[3]), float(f[4]))
    got = W.pressure_elevation_lines([oufr], [0.0], [rad], [dif * 0])
    assert got[:100] == rad_lines
    got = W.pressure_elevation_lines([oufr], [0.0], [np.zeros((100, 0))], [dif])
    assert got == dif_lines


def test_incidence_round_trip():
    ref = list(Z["ex2_PressureElevationIncidence_6p"])[:100]
    oufr = float(ref[0][:14])
    inc = np.zeros((100, 1), dtype=np.complex128)
    for p, line in enumerate(ref):
        f = cols(line, [14, 14, 10, 14, 14, 14, 14])
        inc[p, 0] = complex(float(f[5]), float(f[6]))
    got = W.incidence_lines([oufr], [0.0], [inc])
    _close_lines(got, ref, {0, 1, 2, 5, 6}, [14, 14, 10, 14, 14, 14, 14])
