"""The deck host end to end on the GPU: input files in the reference's
formats -> hamsx.driver.run_deck -> the HAMS and WAMIT result files.

1. The reference's own regression deck, test/data/config1 (3 bodies x 234
   panels, depth 70 m, 3 frequencies, 3 headings), rebuilt from
   tests/golden/config1.npz: the written OAMASS1-6, ODAMPING1-6 and OEXFOR1-6
   tables agree with the reference's shipped expected output to the last
   printed digit, or within the bound the oracle itself meets
   (tests/test_oracle_config1.py) widened to 4x the reference's own rounding
   spread on this deck (the oracle's rounding draws of tests/gates.py move
   the added mass by up to 7.8e-7 of its largest entry at 0.07 rad/s, more
   than that bound) plus the half unit of the last digit that printing our
   own values adds.
2. A deck with the frequency limits (SYBO = 1), periods as input (INFT = 4),
   field points and body matrices: every file the driver writes equals, line
   for line, the file written from the C-ABI's own step-by-step results.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from hamsx import writers as W
from hamsx.deck import ControlDeck, Hydrostatic, PanelMesh, read_hams_block_file, write_control, write_hydrostatic, \
    write_pnl
from hamsx.driver import AMP, RHO, frequency_list, run_deck
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver
from hamsx.wamit import nondim_field
import oracle
from test_oracle_config1 import combined_with_oracle_geometry, printed_ulps, tables_from

pytestmark = pytest.mark.gpu


def hydro(scale):
    eye = np.eye(6)
    return Hydrostatic(np.zeros(3), scale * 1e6 * eye, scale * 1e3 * eye, np.zeros((6, 6)), scale * 1e5 * eye,
                       np.zeros((6, 6)))


def write_deck(indir, ctl, bodies, hydros):
    os.makedirs(indir, exist_ok=True)
    write_control(os.path.join(indir, "ControlFile.in"), ctl)
    for b, (m, h) in enumerate(zip(bodies, hydros)):
        suffix = "" if ctl.nbody == 1 else f"_{b + 1}"
        write_pnl(os.path.join(indir, f"HullMesh{suffix}.pnl"), m)
        write_hydrostatic(os.path.join(indir, f"Hydrostatic{suffix}.in"), h)


def test_config1_deck_writes_the_reference_tables(tmp_path):
    z = np.load(os.path.join(GOLDEN, "config1.npz"))
    bodies = [PanelMesh(z[f"xyz{b}"], z[f"ncn{b}"], z[f"ncon{b}"]) for b in range(3)]
    ctl = ControlDeck(depth=float(z["depth"]), sybo=0, inft=int(z["inft"]), ouft=3, wvnb=[float(w) for w in z["omegas"]],
                      nbody=3, lcs=np.zeros((3, 4)), headings_deg=[float(b) for b in z["headings_deg"]],
                      xr=np.array(z["xr"]), refl=1.0, isol=int(z["isol"]), irsp=0, nthread=1)
    indir, outdir = os.path.join(tmp_path, "Input"), os.path.join(tmp_path, "Output")
    write_deck(indir, ctl, bodies, [hydro(1.0)] * 3)
    print(f"[gpu] deck config1: {indir}", flush=True)
    res = run_deck(indir, outdir)
    assert len(res.files) == 18 + 4
    # the reference's own rounding spread on this deck: the oracle's three
    # rounding draws (tests/gates.py), as tables like the files'
    m = combined_with_oracle_geometry(bodies, z["xr"])
    heads = np.deg2rad(z["headings_deg"])
    runs = {}
    for tag, mm, fma in (("oracle", m, False), ("fma", m, True), ("pz", oracle.perturbed(m, 1), False),
                         ("pxy", oracle.perturbed(m, 2), False)):
        runs[tag] = tables_from([oracle.solve_frequency(mm, float(w), float(z["depth"]), heads, nthreads=16,
                                                        fma=fma)[:3] for w in z["omegas"]], 3, len(heads))
    for q, (name, digits, tol) in enumerate((("OAMASS", 5, 5e-7), ("ODAMPING", 5, 1e-3), ("OEXFOR", 6, 5e-7))):
        for k in range(6):
            got = read_hams_block_file(os.path.join(outdir, "Hams_format", f"{name}{k + 1}.txt"))
            ref = z[f"{name}{k + 1}"]
            assert got.shape == ref.shape
            assert np.all(np.abs(got[:, :2] - ref[:, :2]) <= 5e-4)     # WK, W1 columns (F7.3)
            # both sides are printed values now: our own rounding to the printed
            # digits adds up to half a unit of the last digit to the oracle's bound
            g, r = got[:, 2:], ref[:, 2:]
            ulp = 10.0 ** (np.floor(np.log10(np.maximum(np.abs(r), 1e-300))) - (digits - 1))
            spread = max(float(np.max(np.abs(runs[t][q][k] - runs["oracle"][q][k]))) for t in ("fma", "pz", "pxy"))
            close = printed_ulps(g, r, digits) <= 1.5
            small = np.abs(g - r) <= max(tol * np.max(np.abs(r)), 4 * spread) + 0.5 * ulp
            frac = float(np.mean(printed_ulps(g, r, digits) <= 0.5))
            print(f"    {name}{k + 1}: {frac:.3f} of the entries print the same digits as the reference; "
                  f"largest difference {np.max(np.abs(g - r)) / np.max(np.abs(r)):.1e} of the table's largest "
                  f"entry, the reference's own rounding spread {spread / np.max(np.abs(r)):.1e}", flush=True)
            assert np.all(close | small), (name, k + 1)


def test_deck_files_equal_the_step_by_step_results(tmp_path):
    bodies = [cylinder(4.0, 8.0, 20, 5, 3, -15.0), cylinder(3.0, 6.0, 20, 5, 3, 15.0)]
    pms = [PanelMesh(b.xyz, b.ncn, b.ncon) for b in bodies]
    xr = np.array([[-15.0, 0.0, 0.0], [15.0, 0.0, 0.0]])
    xfp = np.array([[0.0, 20.0, 0.0], [30.0, -5.0, -3.0]])
    ctl = ControlDeck(depth=40.0, sybo=1, inft=4, ouft=4, wvnb=[9.0, 5.5], nbody=2, lcs=np.zeros((2, 4)),
                      headings_deg=[0.0, 45.0], xr=xr, refl=2.0, isol=1, irsp=0, nthread=1, xfp=xfp)
    hyd = [hydro(1.0), hydro(0.6)]
    indir, outdir = os.path.join(tmp_path, "Input"), os.path.join(tmp_path, "Output")
    write_deck(indir, ctl, pms, hyd)
    print(f"[gpu] deck limits + periods + field points: {indir}", flush=True)
    res = run_deck(indir, outdir)
    # the same loop through the C-ABI directly
    mesh = combine(pms, list(xr))
    wk, w1, tp, oufr, am, bd, ex, ds, rad, dif, inc = ([] for _ in range(11))
    with FrequencySolver(mesh) as s:
        s.set_waves(np.deg2rad(ctl.headings_deg), amp=AMP, rho=RHO, isol=1)
        s.set_hydrostatics(hyd)
        s.set_field_points(xfp)
        for inft, value in frequency_list(ctl):
            v, k, w, _ = s.calwaveproperts(value, ctl.depth, inft)
            s.assemble()
            assert s.factor() == 0
            s.radiation_solver()
            s.diffraction_solver()
            a, b, e, _ = s.forces()
            d, _ = s.motions()
            lim = inft == 0
            t = (-1.0 if value == 0.0 else 0.0) if lim else value
            wk.append(k), w1.append(w), tp.append(t), oufr.append(t)
            am.append(a), bd.append(b), ex.append(e), ds.append(d)
            vals, _, ic = s.field_values("auto", incident=True)
            nv = nondim_field(vals, xfp[:, 2], 2, w, AMP, RHO, ctl.refl, limit=lim)
            iv = nondim_field(np.concatenate([np.zeros((2, 12)), ic], axis=1), xfp[:, 2], 2, w, AMP, RHO, ctl.refl,
                              limit=lim)[:, 12:]
            rad.append(nv[:, :12]), dif.append(nv[:, 12:]), inc.append(iv)
    want = dict(W.hams_force_lines(wk, w1, am, bd, ex, 2))
    want.update({"AmssDamp.1": W.amssdamp_lines(oufr, w1, tp, am, bd, RHO),
                 "ExcForce.3": W.excforce_lines(oufr, tp, ctl.headings_deg, ex, 2, RHO, AMP),
                 "Motion.4": W.motion_lines(oufr, tp, ctl.headings_deg, ds, AMP, ctl.refl),
                 "Hydrostat.hst": W.hydrostat_lines([h.crs for h in hyd], RHO),
                 "PressureElevation.6p": W.pressure_elevation_lines(oufr, ctl.headings_deg, rad, dif),
                 "PressureElevationIncidence.6p": W.incidence_lines(oufr, ctl.headings_deg, inc)})
    assert len(res.files) == len(want)
    for path in res.files:
        with open(path) as f:
            lines = f.read().splitlines()
        assert lines == want[os.path.basename(path)], os.path.basename(path)
    assert res.tp[:2] == [-1.0, 0.0] and res.tp[2:] == ctl.wvnb
