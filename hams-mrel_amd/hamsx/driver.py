"""The reference driver's flow for one input deck, on the device.

``run_deck(indir, outdir)`` reads ``ControlFile.in``, ``HullMesh*.pnl``,
``Hydrostatic*.in`` and (IRSP = 1) ``WaterPlaneMesh*.pnl`` with the
reference's column layouts (:mod:`hamsx.deck`), solves every frequency on the
GPU through the C-ABI in the reference's call order -- CalWaveProperts,
CALGREEN + ASSB_LEFT, ZGETRF, the radiation and diffraction solves, RFORCE /
EFORCE, SolveMotionMulti, the field points -- and writes the HAMS and WAMIT
result files (:mod:`hamsx.writers`) under ``outdir/Hams_format`` and
``outdir/Wamit_format``.  It restates the multi-body frequency loop of
src/HAMS_Prog.f90:613-735 (single-body decks take the same path with one
body), the frequency list of ReadControlFile (src/InputFiles.f90:93-112:
with SYBO = 1 the zero- and infinite-frequency limits come first) and the
per-frequency wave quantities of CalWaveProperts (src/ImplementSubs.f90:40-151).

Command line, as the reference's ``hamsmrel [indir outdir]``
(src/HAMS_Prog.f90:125-132)::

    python -m hamsx.driver Input Output
"""
from __future__ import annotations

import os
import sys
from dataclasses import dataclass, field

import numpy as np

from . import writers as W
from ._lib import FREQ_LIMIT, HAMSX_ENOCONV, HamsxError
from .deck import read_deck, read_pnl
from .mesh import combine, panel_geometry
from .solver import FrequencySolver
from .wamit import nondim_field

PI = 3.141592653589793
G = 9.80665
RHO, AMP = 1000.0, 1.0  # Const_mod (WavDynMods.f90:146), InitialisationMulti (HydroStaticMulti.f90:45-46)


@dataclass
class DeckResult:
    """What the deck's frequency loop produced, per frequency (in loop order)."""
    wk: list = field(default_factory=list)
    w1: list = field(default_factory=list)
    tp: list = field(default_factory=list)
    oufr: list = field(default_factory=list)
    amas: list = field(default_factory=list)
    bdmp: list = field(default_factory=list)
    exfc: list = field(default_factory=list)
    dspl: list = field(default_factory=list)
    rad: list = field(default_factory=list)   # scaled field values (.6p), per frequency
    dif: list = field(default_factory=list)
    inc: list = field(default_factory=list)
    files: list = field(default_factory=list)


def frequency_list(ctl):
    """(inft, value) per loop index KK: the two limits first when SYBO = 1
    (ImplementSubs.f90:45-70), then the deck's frequencies of type INFT."""
    out = [(FREQ_LIMIT, 0.0), (FREQ_LIMIT, -1.0)] if ctl.sybo == 1 else []
    return out + [(ctl.inft, float(v)) for v in ctl.wvnb]


def wave_quantities(inft, value, v, wk, w1):
    """TP and WL as CalWaveProperts leaves them (ImplementSubs.f90:45-138)."""
    if inft == FREQ_LIMIT:
        return (-1.0, -1.0) if value == 0.0 else (0.0, 0.0)
    tp = value if inft == 4 else 2.0 * PI / w1
    wl = value if inft == 5 else 2.0 * PI / wk
    return tp, wl


def output_frequency(ouft, v, wk, w1, tp, wl):
    """OUFR by OUTPUT_FREQUENCY_TYPE (ImplementSubs.f90:140-150)."""
    return {1: v, 2: wk, 3: w1, 4: tp, 5: wl}[ouft]


def _lids(indir, ctl):
    """The waterplane (lid) meshes of IRSP = 1, placed by each body's LCS
    (ReadWTPLMeshMulti, ReadPanelMeshMulti.f90:76-118)."""
    lids = []
    for b in range(ctl.nbody):
        name = "WaterPlaneMesh.pnl" if ctl.nbody == 1 else f"WaterPlaneMesh_{b + 1}.pnl"
        lids.append(read_pnl(os.path.join(indir, name), ctl.lcs[b]))
    return lids


def run_deck(indir: str, outdir: str | None = None, precision: str = "fp64", device: int = 0,
             log=None) -> DeckResult:
    """Solve the deck in `indir` on GPU `device` and, when `outdir` is given,
    write the reference's result files there.  Returns the per-frequency
    results."""
    say = log or (lambda s: None)
    deck = read_deck(indir)
    ctl = deck.control
    mesh = combine(deck.bodies, [ctl.xr[b] for b in range(ctl.nbody)])
    heads_deg = [float(b) for b in ctl.headings_deg]
    isx = {int(getattr(b, "isx", 0)) for b in deck.bodies}
    isy = {int(getattr(b, "isy", 0)) for b in deck.bodies}
    if len(isx) > 1 or len(isy) > 1:
        raise ValueError("every body's mesh must declare the same symmetry planes (ISX, ISY)")
    isx, isy = isx.pop(), isy.pop()
    res = DeckResult()
    with FrequencySolver(mesh, device=device, precision=precision) as s:
        s.set_waves(np.deg2rad(heads_deg), amp=AMP, xw=(0.0, 0.0), rho=RHO, isol=ctl.isol)
        if ctl.irsp == 1:
            lid = np.concatenate([panel_geometry(lm, np.zeros(3))[0] for lm in _lids(indir, ctl)])
            s.set_lid(lid)
        if isx or isy:
            s.set_symmetry(isx, isy)
        if ctl.isolv == 2:
            s.set_solver("gmres")
        s.set_hydrostatics(deck.hydro)
        nfp = len(ctl.xfp)
        if nfp:
            s.set_field_points(ctl.xfp)
        for inft, value in frequency_list(ctl):
            v, wk, w1, _ = s.calwaveproperts(value, ctl.depth, inft)
            tp, wl = wave_quantities(inft, value, v, wk, w1)
            s.assemble()
            info = s.factor()
            if info != 0:
                say(f"hamsx: zero pivot {info} at frequency {value} (the reference continues; so does this host)")
            s.radiation_solver()
            s.diffraction_solver()
            amas, bdmp, exfc, _ = s.forces()
            try:
                dspl, _ = s.motions()
            except HamsxError as e:  # the reference's fixed point has no bound; this one stops at 1000 steps
                if e.code != HAMSX_ENOCONV:
                    raise
                say(f"hamsx: motions did not converge at frequency {value}; Motion.4 rows are NaN")
                dspl = np.full((6 * ctl.nbody, len(heads_deg)), np.nan + 0j)
            res.wk.append(wk)
            res.w1.append(w1)
            res.tp.append(tp)
            res.oufr.append(output_frequency(ctl.ouft, v, wk, w1, tp, wl))
            res.amas.append(amas)
            res.bdmp.append(bdmp)
            res.exfc.append(exfc)
            res.dspl.append(dspl)
            if nfp:
                vals, _, inc = s.field_values("auto", incident=True)
                lim = inft == FREQ_LIMIT
                nv = nondim_field(vals, ctl.xfp[:, 2], ctl.nbody, w1, AMP, RHO, ctl.refl, limit=lim)
                iv = nondim_field(np.concatenate([np.zeros((nfp, 6 * ctl.nbody)), inc], axis=1), ctl.xfp[:, 2],
                                  ctl.nbody, w1, AMP, RHO, ctl.refl, limit=lim)[:, 6 * ctl.nbody:]
                res.rad.append(nv[:, :6 * ctl.nbody])
                res.dif.append(nv[:, 6 * ctl.nbody:])
                res.inc.append(iv)
            say(f"hamsx: frequency {value} (type {inft}) done")
    if outdir is not None:
        res.files = write_outputs(outdir, ctl, deck, res, heads_deg)
    return res


def write_outputs(outdir, ctl, deck, res, heads_deg):
    """The HAMS tables (CreateHamsFiles, InputFiles.f90:296-309) and the WAMIT
    files (CreateWamitFiles, :311-334) from the loop's results."""
    hams = os.path.join(outdir, "Hams_format")
    wamit = os.path.join(outdir, "Wamit_format")
    os.makedirs(hams, exist_ok=True)
    os.makedirs(wamit, exist_ok=True)
    files = []
    for name, lines in W.hams_force_lines(res.wk, res.w1, res.amas, res.bdmp, res.exfc, ctl.nbody).items():
        files.append(os.path.join(hams, name))
        W.write_lines(files[-1], lines)
    out = {
        "AmssDamp.1": W.amssdamp_lines(res.oufr, res.w1, res.tp, res.amas, res.bdmp, RHO),
        "ExcForce.3": W.excforce_lines(res.oufr, res.tp, heads_deg, res.exfc, ctl.nbody, RHO, AMP),
        "Motion.4": W.motion_lines(res.oufr, res.tp, heads_deg, res.dspl, AMP, ctl.refl),
        "Hydrostat.hst": W.hydrostat_lines([h.crs for h in deck.hydro], RHO),
    }
    if len(ctl.xfp):
        out["PressureElevation.6p"] = W.pressure_elevation_lines(res.oufr, heads_deg, res.rad, res.dif)
        if ctl.nbody > 1:
            out["PressureElevationIncidence.6p"] = W.incidence_lines(res.oufr, heads_deg, res.inc)
    for name, lines in out.items():
        files.append(os.path.join(wamit, name))
        W.write_lines(files[-1], lines)
    return files


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    indir = argv[0] if len(argv) > 0 else "Input"
    outdir = argv[1] if len(argv) > 1 else "Output"
    res = run_deck(indir, outdir, log=lambda s: print(s, flush=True))
    print(f"hamsx: {len(res.w1)} frequencies, {len(res.files)} files written under {outdir}", flush=True)


if __name__ == "__main__":
    main()
