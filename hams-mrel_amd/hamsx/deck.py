"""Readers for HAMS-MREL input decks and writers for its HAMS-format outputs.

The on-disk formats are the reference's:
  * ``ControlFile.in`` fixed-column deck  -- ``src/InputFiles.f90:46-224``
  * ``HullMesh[_i].pnl`` panel meshes      -- ``src/HAMS_Prog.f90:483-512``,
    ``src/ReadPanelMeshMulti.f90:46-71`` (local -> global transform)
  * ``Hydrostatic[_i].in``                 -- ``src/HydroStaticMulti.f90:68-118``
  * ``Output/Hams_format/OAMASSk.txt`` etc -- ``src/PotentWavForceMulti.f90:119,202-203``
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np


def _ffloat(s: str) -> float:
    """Fortran list/fixed REAL field: accepts D exponents; a field without a
    decimal point under fW.D has an implied decimal (F-format rule)."""
    t = s.strip().replace("d", "e").replace("D", "E")
    if not t:
        return 0.0
    return float(t)


def _fixed_real(line: str, skip: int, width: int, count: int, implied: int) -> list[float]:
    out = []
    pos = skip
    for _ in range(count):
        fld = line[pos:pos + width]
        pos += width
        # formatted input ignores blanks inside a numeric field (BLANK='NULL')
        t = fld.replace(" ", "").replace("d", "e").replace("D", "E")
        if not t:
            out.append(0.0)
        elif "." in t or "e" in t.lower():
            out.append(float(t))
        else:
            out.append(float(t) / 10 ** implied)
    return out


def _fixed_int(line: str, skip: int, width: int) -> int:
    t = line[skip:skip + width].strip()
    return int(t) if t else 0


@dataclass
class ControlDeck:
    """The fields of ControlFile.in that the frequency-domain solve consumes."""
    depth: float
    sybo: int
    inft: int
    ouft: int
    wvnb: list            # input frequencies (type inft)
    nbody: int
    lcs: np.ndarray       # (nbody, 4): x0, y0, z0, rotation about z [rad]
    headings_deg: list
    xr: np.ndarray        # (nbody, 3) reference body centres
    refl: float
    isol: int
    irsp: int
    nthread: int
    isolv: int = 1
    xfp: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))  # field points (NFP x 3)


def read_control(path: str) -> ControlDeck:
    """ReadControlFile (``src/InputFiles.f90:46-224``), same column layout."""
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)
    nxt = lambda: next(it)
    nxt(); nxt()
    h = _fixed_real(nxt(), 14, 30, 1, 15)[0]
    nxt(); nxt()
    sybo = _fixed_int(nxt(), 27, 16)
    inft = _fixed_int(nxt(), 25, 16)
    ouft = _fixed_int(nxt(), 25, 17)
    npet = _fixed_int(nxt(), 26, 16)
    if npet > 0:
        wvnb = [_ffloat(x) for x in nxt().split()][:npet]
    else:
        wk1 = _fixed_real(nxt(), 27, 30, 1, 15)[0]
        dwk = _fixed_real(nxt(), 19, 30, 1, 15)[0]
        wvnb = [wk1 + i * dwk for i in range(abs(npet))]
    nxt(); nxt()
    body_check = nxt()
    nbody = _fixed_int(nxt(), 24, 16) if "multi" in body_check else 1
    lcs = np.zeros((nbody, 4))
    if nbody > 1:
        for b in range(nbody):
            lcs[b] = _fixed_real(nxt(), 26, 12, 4, 3)
        nxt(); nxt(); nxt()
    nbeta = _fixed_int(nxt(), 23, 16)
    if nbeta > 0:
        heads = [_ffloat(x) for x in nxt().split()][:nbeta]
    else:
        b1 = _fixed_real(nxt(), 20, 30, 1, 15)[0]
        db = _fixed_real(nxt(), 17, 30, 1, 15)[0]
        heads = [b1 + i * db for i in range(abs(nbeta))]
    nxt(); nxt()
    xr = np.zeros((nbody, 3))
    for b in range(nbody):
        xr[b] = _fixed_real(nxt(), 28, 12, 3, 3)
    refl = _fixed_real(nxt(), 26, 30, 1, 15)[0]
    isol = _fixed_int(nxt(), 26, 16)
    irsp = _fixed_int(nxt(), 23, 16)
    nthread = _fixed_int(nxt(), 23, 16)
    rest = list(it)
    k = 0
    isolv = 1
    # optional Solver_type line; when absent the reference steps back (backspace)
    if rest and "Solver_type" in rest[0]:
        try:
            isolv = _fixed_int(rest[0], 26, 16)
        except ValueError:
            isolv = 1
        if isolv not in (1, 2):
            isolv = 1
        k = 1
    # field points (:193-203): two lines, NFP in (27x,i16), then NFP list-directed x y z records
    xfp = np.zeros((0, 3))
    if k + 2 < len(rest):
        nfp = _fixed_int(rest[k + 2], 27, 16)
        if nfp > 0:
            xfp = np.array([[_ffloat(t) for t in rest[k + 3 + i].split()[:3]] for i in range(nfp)])
    return ControlDeck(h, sybo, inft, ouft, wvnb, nbody, lcs, heads, xr, refl, isol, irsp,
                       nthread, isolv, xfp)


@dataclass
class PanelMesh:
    """One body's panel mesh in global coordinates (nodes, 1-based connectivity)."""
    xyz: np.ndarray       # (nnode, 3)
    ncn: np.ndarray       # (nelem,) 3 or 4
    ncon: np.ndarray      # (nelem, 4) 1-based node ids (col 4 = 0 for triangles)
    isx: int = 0
    isy: int = 0

    @property
    def nelem(self) -> int:
        return len(self.ncn)

    @property
    def nnode(self) -> int:
        return len(self.xyz)


def place(mesh: PanelMesh, lcs) -> PanelMesh:
    """The LCS transform of ReadBodyMeshMulti (``src/ReadPanelMeshMulti.f90:57-59``)
    applied to a mesh read in local coordinates: rotation about z by lcs[3],
    then translation by (lcs[0], lcs[1]); z is kept."""
    c, s = math.cos(lcs[3]), math.sin(lcs[3])
    xyz = np.array(mesh.xyz, dtype=np.float64)
    x, y = xyz[:, 0].copy(), xyz[:, 1].copy()
    xyz[:, 0] = 0.0 + lcs[0] + c * x - s * y
    xyz[:, 1] = 0.0 + lcs[1] + s * x + c * y
    return PanelMesh(xyz, mesh.ncn.copy(), mesh.ncon.copy(), mesh.isx, mesh.isy)


def read_pnl(path: str, lcs=(0.0, 0.0, 0.0, 0.0)) -> PanelMesh:
    """HullMesh .pnl reader with the LCS transform of ReadBodyMeshMulti
    (``src/ReadPanelMeshMulti.f90:57-59``): rotation about z by lcs[3] then
    translation by (lcs[0], lcs[1]); z is kept as read."""
    with open(path) as f:
        lines = f.read().splitlines()
    nelem, nnode, isx, isy = (int(x) for x in lines[3].split()[:4])
    pos = 6
    xyz = np.zeros((nnode, 3))
    c, s = math.cos(lcs[3]), math.sin(lcs[3])
    for i in range(nnode):
        tok = lines[pos + i].split("!")[0].split()
        x, y, z = (_ffloat(t) for t in tok[1:4])
        xyz[i, 0] = 0.0 + lcs[0] + c * x - s * y
        xyz[i, 1] = 0.0 + lcs[1] + s * x + c * y
        xyz[i, 2] = z
    pos += nnode + 3
    ncn = np.zeros(nelem, dtype=np.int32)
    ncon = np.zeros((nelem, 4), dtype=np.int32)
    for e in range(nelem):
        tok = lines[pos + e].split("!")[0].split()
        k = int(tok[1])
        ncn[e] = k
        ncon[e, :k] = [int(t) for t in tok[2:2 + k]]
    return PanelMesh(xyz, ncn, ncon, isx, isy)


def write_pnl(path: str, mesh: PanelMesh, title: str = "Hull Mesh File") -> None:
    """Writes a mesh in the .pnl layout read by :func:`read_pnl`."""
    with open(path, "w") as f:
        f.write(f"    --------------{title}---------------\n \n")
        f.write("    # Number of Panels, Nodes, X-Symmetry and Y-Symmetry\n")
        f.write(f"{mesh.nelem:12d}{mesh.nnode:12d}{mesh.isx:12d}{mesh.isy:12d}\n \n")
        f.write("    # Start Definition of Node Coordinates     ! node_number   x   y   z\n")
        for i, p in enumerate(mesh.xyz):
            f.write(f"{i + 1:8d}" + "".join(f"{_num(x):>25s}" for x in p[:3]) + "\n")  # exact round trip
        f.write("   # End Definition of Node Coordinates\n \n")
        f.write("   # Start Definition of Node Relations   ! panel_number  number_of_vertices   Vertex1_ID   Vertex2_ID   Vertex3_ID   (Vertex4_ID)\n")
        for e in range(mesh.nelem):
            k = int(mesh.ncn[e])
            ids = " ".join(f"{int(v):8d}" for v in mesh.ncon[e, :k])
            f.write(f"{e + 1:8d}{k:6d} {ids}\n")
        f.write("   # End Definition of Node Relations\n")


def _num(x: float) -> str:
    """A real for a list-directed or F-format field: the shortest repr that
    reads back to the same double, with a decimal point."""
    t = repr(float(x))
    return t if ("." in t or "e" in t or "n" in t) else t + ".0"


def _f12(x: float) -> str:
    """A real for a 12-column F12.3 field (LCS, reference body centres,
    InputFiles.f90:135,164): the exact value when its shortest form fits in
    11 columns, else the fixed-point form with as many decimals as fit."""
    t = _num(x)
    if len(t) <= 11 and "e" not in t:
        return f"{t:>12s}"
    for d in range(10, 0, -1):
        t = f"{float(x):.{d}f}"
        if len(t) <= 11:
            return f"{t:>12s}"
    raise ValueError(f"{x} does not fit an F12.3 field")


def write_control(path: str, ctl: ControlDeck) -> None:
    """ControlFile.in in the layout ReadControlFile reads
    (``src/InputFiles.f90:46-224``): every fixed-column field at the column
    its READ format starts, frequencies and headings as list-directed
    records (positive counts), the optional Solver_type line when ISOLV = 2,
    and the field points."""
    L = ["   --------------HAMS Control file---------------", "",
         f"{'   Waterdepth':<14}{_num(ctl.depth)}", "",
         "   #Start Definition of Wave Frequencies",
         f"{'    0_inf_frequency_limits':<27}{ctl.sybo:16d}",
         f"{'    Input_frequency_type':<25}{ctl.inft:16d}",
         f"{'    Output_frequency_type':<25}{ctl.ouft:17d}",
         f"{'    Number_of_frequencies':<26}{len(ctl.wvnb):16d}",
         "    " + "  ".join(_num(v) for v in ctl.wvnb),
         "   #End Definition of Wave Frequencies", ""]
    if ctl.nbody > 1:
        L += ["   # Start of definition of bodies for multi-body interaction",
              f"{'    Number_of_bodies':<24}{ctl.nbody:16d}"]
        for b in range(ctl.nbody):
            L.append(f"{'    LCS_' + str(b + 1):<26}" + "".join(_f12(x) for x in ctl.lcs[b]))
        L += ["   # End of definition of bodies for multi-body interaction", ""]
    L += ["   #Start Definition of Wave Headings",
          f"{'    Number_of_headings':<23}{len(ctl.headings_deg):16d}",
          "    " + "  ".join(_num(v) for v in ctl.headings_deg),
          "   #End Definition of Wave Headings", ""]
    for b in range(ctl.nbody):
        name = "    Reference_body_center" + (f"_{b + 1}" if ctl.nbody > 1 else "")
        L.append(f"{name:<28}" + "".join(_f12(x) for x in ctl.xr[b]))
    L += [f"{'    Reference_body_length':<26}{_num(ctl.refl)}",
          f"{'    Wave_diffrac_solution':<26}{ctl.isol:16d}",
          f"{'    If_remove_irr_freq':<23}{ctl.irsp:16d}",
          f"{'    Number of threads':<23}{ctl.nthread:16d}"]
    if ctl.isolv == 2:
        L.append(f"{'    Solver_type':<26}{2:16d}")
    L += ["", "   #Start Definition of Pressure and/or Elevation (PE)",
          f"{'    Number_of_field_points':<27}{len(ctl.xfp):16d}"]
    L += ["    " + "  ".join(_num(x) for x in p) for p in np.asarray(ctl.xfp).reshape(-1, 3)]
    L += ["   #End Definition of Pressure and/or Elevation", "",
          "   ----------End HAMS Control file---------------"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_hydrostatic(path: str, h: "Hydrostatic") -> None:
    """Hydrostatic*.in in the layout ReadHydroStaticMulti reads
    (``src/HydroStaticMulti.f90:68-98``)."""
    L = [" Center of Gravity:", "  " + "  ".join(_num(x) for x in h.xg)]
    for title, m in (("Body Mass Matrix:", h.mass), ("External Linear Damping Matrix:", h.blnr),
                     ("External Quadratic Damping Matrix:", h.bqdr), ("Hydrostatic Restoring Matrix:", h.crs),
                     ("External Restoring Matrix:", h.kstf)):
        L.append(" " + title)
        L += ["  " + "   ".join(_num(x) for x in row) for row in np.asarray(m)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


@dataclass
class Hydrostatic:
    xg: np.ndarray
    mass: np.ndarray
    blnr: np.ndarray
    bqdr: np.ndarray
    crs: np.ndarray
    kstf: np.ndarray


def read_hydrostatic(path: str) -> Hydrostatic:
    """ReadHydroStaticMulti (``src/HydroStaticMulti.f90:68-98``)."""
    with open(path) as f:
        lines = f.read().splitlines()
    xg = np.array([_ffloat(t) for t in lines[1].split()[:3]])
    mats = []
    pos = 2
    for _ in range(5):
        pos += 1
        m = np.array([[_ffloat(t) for t in lines[pos + r].split()[:6]] for r in range(6)])
        mats.append(m)
        pos += 6
    return Hydrostatic(xg, *mats)


@dataclass
class Deck:
    """A complete multi-body input deck, combined as BuildCombinedArrays does
    (``src/CalGreenFuncMulti.f90:53-106``): bodies concatenated in order,
    connectivity offset by the nodes of the bodies before."""
    control: ControlDeck
    bodies: list = field(default_factory=list)
    hydro: list = field(default_factory=list)


def read_deck(indir: str) -> Deck:
    ctl = read_control(os.path.join(indir, "ControlFile.in"))
    d = Deck(ctl)
    for b in range(ctl.nbody):
        if ctl.nbody == 1:
            mp, hp = "HullMesh.pnl", "Hydrostatic.in"
            if not os.path.exists(os.path.join(indir, mp)):
                mp, hp = "HullMesh_1.pnl", "Hydrostatic_1.in"
        else:
            mp, hp = f"HullMesh_{b + 1}.pnl", f"Hydrostatic_{b + 1}.in"
        d.bodies.append(read_pnl(os.path.join(indir, mp), ctl.lcs[b]))
        d.hydro.append(read_hydrostatic(os.path.join(indir, hp)))
    return d


def read_hams_block_file(path: str) -> np.ndarray:
    """Reads an OAMASSk/ODAMPINGk/OEXFORk table (format F7.3,1X,F7.3,1X,6E14.x)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([_ffloat(t) for t in line.split()])
    return np.array(rows)
