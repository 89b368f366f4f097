"""The deck host's file handling and frequency bookkeeping (no GPU): the
ControlFile.in / .pnl / Hydrostatic*.in writers read back through the
readers to the same values, and the frequency list and wave quantities follow
ReadControlFile and CalWaveProperts (src/InputFiles.f90:93-112,
src/ImplementSubs.f90:40-151)."""
import math
import os

import numpy as np

from hamsx import driver
from hamsx._lib import FREQ_LIMIT
from hamsx.deck import (ControlDeck, Hydrostatic, PanelMesh, read_control, read_hydrostatic, read_pnl,
                        write_control, write_hydrostatic, write_pnl)
from hamsx.mesh import cylinder


def deck(nbody, sybo=1, inft=4, isolv=2, nfp=3):
    rng = np.random.default_rng(3)
    return ControlDeck(depth=55.25, sybo=sybo, inft=inft, ouft=2, wvnb=list(rng.uniform(1, 9, 4)), nbody=nbody,
                       lcs=np.round(rng.uniform(-50, 50, (nbody, 4)), 3), headings_deg=[0.0, 22.5, 181.0 / 3],
                       xr=np.round(rng.normal(0, 30, (nbody, 3)), 4),
                       refl=1.0, isol=1, irsp=1, nthread=8, isolv=isolv, xfp=rng.normal(0, 40, (nfp, 3)))


def test_control_file_round_trip(tmp_path):
    for nb in (1, 3):
        c = deck(nb)
        p = os.path.join(tmp_path, f"ControlFile{nb}.in")
        write_control(p, c)
        r = read_control(p)
        for k in c.__dataclass_fields__:
            if k == "lcs" and nb == 1:
                continue  # a single body has no LCS lines (zeros)
            assert np.array_equal(np.asarray(getattr(c, k)), np.asarray(getattr(r, k))), k


def test_pnl_and_hydrostatic_round_trip(tmp_path):
    m = cylinder(4.3, 7.1, 20, 5, 3, 11.3, -2.7)
    pm = PanelMesh(m.xyz, m.ncn, m.ncon, 0, 1)
    write_pnl(os.path.join(tmp_path, "m.pnl"), pm)
    r = read_pnl(os.path.join(tmp_path, "m.pnl"))
    assert np.array_equal(r.xyz, pm.xyz) and np.array_equal(r.ncn, pm.ncn) and np.array_equal(r.ncon, pm.ncon)
    assert (r.isx, r.isy) == (0, 1)
    rng = np.random.default_rng(1)
    h = Hydrostatic(rng.normal(size=3), *[rng.normal(size=(6, 6)) for _ in range(5)])
    write_hydrostatic(os.path.join(tmp_path, "h.in"), h)
    g = read_hydrostatic(os.path.join(tmp_path, "h.in"))
    for k in ("xg", "mass", "blnr", "bqdr", "crs", "kstf"):
        assert np.array_equal(getattr(h, k), getattr(g, k)), k


def test_frequency_list_and_wave_quantities():
    c = deck(2, sybo=1, inft=4)
    fl = driver.frequency_list(c)
    assert fl[:2] == [(FREQ_LIMIT, 0.0), (FREQ_LIMIT, -1.0)] and [v for _, v in fl[2:]] == c.wvnb
    assert driver.wave_quantities(FREQ_LIMIT, 0.0, 0, 0, 0) == (-1.0, -1.0)
    assert driver.wave_quantities(FREQ_LIMIT, -1.0, -1, -1, -1) == (0.0, 0.0)
    tp, wl = driver.wave_quantities(4, 7.5, 0.1, 0.12, 2 * math.pi / 7.5)
    assert tp == 7.5 and wl == 2 * math.pi / 0.12            # INFT = 4: TP is the input
    tp, wl = driver.wave_quantities(5, 80.0, 0.1, 2 * math.pi / 80.0, 0.9)
    assert wl == 80.0 and tp == 2 * math.pi / 0.9             # INFT = 5: WL is the input
    assert [driver.output_frequency(k, 1, 2, 3, 4, 5) for k in range(1, 6)] == [1, 2, 3, 4, 5]
    assert driver.frequency_list(deck(1, sybo=0))[0][0] == 4
