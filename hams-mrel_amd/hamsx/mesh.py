"""Panel meshes: seeded synthetic bodies and the combined multi-body arrays.

The combination follows BuildCombinedArrays (``src/CalGreenFuncMulti.f90:53-106``):
bodies are concatenated in order and each body's 1-based connectivity is
offset by the node count of the bodies before it.  Panel geometry (centroid,
area, size, the six generalized normals about each body's reference point)
comes from ``hamsx_panel_geometry`` in the native library.

Synthetic bodies follow the reference's panel orientation: the normal that
CalTransNormals computes, ``(x4-x2) x (x3-x1)`` for quads and
``(x1-x2) x (x3-x2)`` for triangles (``src/NormalProcess.f90:74-124``),
points into the body, as it does on every panel of the reference's
``test/data/config1`` meshes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .deck import PanelMesh


def _raw_normal(p: np.ndarray) -> np.ndarray:
    if len(p) == 4:
        return np.cross(p[3] - p[1], p[2] - p[0])
    return np.cross(p[0] - p[1], p[2] - p[1])


def _orient(nodes: list, xyz: np.ndarray, inward: np.ndarray) -> list:
    n = _raw_normal(xyz[nodes])
    if float(np.dot(n, inward)) < 0.0:
        return [nodes[0]] + nodes[1:][::-1]
    return nodes


def _pack(xyz, panels) -> PanelMesh:
    ne = len(panels)
    ncn = np.array([len(p) for p in panels], dtype=np.int32)
    ncon = np.zeros((ne, 4), dtype=np.int32)
    for e, p in enumerate(panels):
        ncon[e, :len(p)] = np.asarray(p) + 1
    return PanelMesh(np.asarray(xyz, dtype=np.float64), ncn, ncon)


def cylinder(radius: float, draft: float, na: int, nz: int, nr: int, x0: float = 0.0,
             y0: float = 0.0) -> PanelMesh:
    """Truncated vertical cylinder: na x nz side quads, nr bottom rings (quads,
    with a centre ring of triangles).  Waterline at z = 0."""
    th = 2.0 * np.pi * np.arange(na) / na
    xyz = []
    for k in range(nz + 1):
        z = -draft * k / nz
        for a in range(na):
            xyz.append((x0 + radius * math.cos(th[a]), y0 + radius * math.sin(th[a]), z))
    ring = [list(range(nz * na, nz * na + na))]  # outer bottom ring = side bottom ring
    for m in range(1, nr):
        r = radius * (nr - m) / nr
        base = len(xyz)
        for a in range(na):
            xyz.append((x0 + r * math.cos(th[a]), y0 + r * math.sin(th[a]), -draft))
        ring.append(list(range(base, base + na)))
    centre = len(xyz)
    xyz.append((x0, y0, -draft))
    xyz = np.array(xyz)
    panels = []
    for k in range(nz):
        for a in range(na):
            b = (a + 1) % na
            p = [k * na + a, (k + 1) * na + a, (k + 1) * na + b, k * na + b]
            c = xyz[p].mean(axis=0)
            inward = np.array([x0 - c[0], y0 - c[1], 0.0])
            panels.append(_orient(p, xyz, inward))
    up = np.array([0.0, 0.0, 1.0])
    for m in range(nr - 1):
        for a in range(na):
            b = (a + 1) % na
            panels.append(_orient([ring[m][a], ring[m][b], ring[m + 1][b], ring[m + 1][a]], xyz, up))
    last = ring[-1]
    for a in range(na):
        b = (a + 1) % na
        panels.append(_orient([last[a], last[b], centre], xyz, up))
    return _pack(xyz, panels)


def waterplane_disk(radius: float, na: int, nr: int, x0: float = 0.0, y0: float = 0.0) -> PanelMesh:
    """Lid of a vertical cylinder for irregular-frequency removal: the
    waterplane disk at z = 0 in nr rings of na panels (triangles at the
    centre).  Only the panel centroids are used (as field points)."""
    th = 2.0 * np.pi * np.arange(na) / na
    xyz, rings = [], []
    for m in range(nr):
        r = radius * (nr - m) / nr
        rings.append(list(range(len(xyz), len(xyz) + na)))
        for a in range(na):
            xyz.append((x0 + r * math.cos(th[a]), y0 + r * math.sin(th[a]), 0.0))
    centre = len(xyz)
    xyz.append((x0, y0, 0.0))
    panels = []
    for m in range(nr - 1):
        for a in range(na):
            b = (a + 1) % na
            panels.append([rings[m][a], rings[m][b], rings[m + 1][b], rings[m + 1][a]])
    for a in range(na):
        panels.append([rings[-1][a], rings[-1][(a + 1) % na], centre])
    return _pack(np.array(xyz), panels)


def box(length: float, beam: float, draft: float, nl: int, nb: int, nt: int, x0: float = 0.0,
        y0: float = 0.0) -> PanelMesh:
    """Rectangular barge L x B x T: four walls and a bottom of uniform quads."""
    xyz, panels = [], []

    def grid(origin, u, v, nu, nv, inward):
        base = len(xyz)
        for j in range(nv + 1):
            for i in range(nu + 1):
                xyz.append(tuple(origin + u * i / nu + v * j / nv))
        arr = np.array(xyz)
        for j in range(nv):
            for i in range(nu):
                p = [base + j * (nu + 1) + i, base + j * (nu + 1) + i + 1,
                     base + (j + 1) * (nu + 1) + i + 1, base + (j + 1) * (nu + 1) + i]
                panels.append(_orient(p, arr, inward))

    hl, hb = length / 2.0, beam / 2.0
    o = np.array([x0, y0, 0.0])
    ex, ey, ez = np.eye(3)
    grid(o + [hl, -hb, -draft], 2 * hb * ey, draft * ez, nb, nt, -ex)
    grid(o + [-hl, -hb, -draft], 2 * hb * ey, draft * ez, nb, nt, ex)
    grid(o + [-hl, hb, -draft], 2 * hl * ex, draft * ez, nl, nt, -ey)
    grid(o + [-hl, -hb, -draft], 2 * hl * ex, draft * ez, nl, nt, ey)
    grid(o + [-hl, -hb, -draft], 2 * hl * ex, 2 * hb * ey, nl, nb, ez)
    return _pack(np.array(xyz), panels)


def cylinder_for_panels(radius: float, draft: float, target: int, x0=0.0, y0=0.0) -> PanelMesh:
    """Cylinder with roughly square panels and about ``target`` panels."""
    lo, hi = 1e-3, max(radius, draft)
    for _ in range(60):
        s = math.sqrt(lo * hi)
        na = max(8, round(2 * math.pi * radius / s))
        cnt = na * (max(1, round(draft / s)) + max(1, round(radius / s)))
        lo, hi = (s, hi) if cnt > target else (lo, s)
    s = math.sqrt(lo * hi)
    na = max(8, round(2 * math.pi * radius / s))
    return cylinder(radius, draft, na, max(1, round(draft / s)), max(1, round(radius / s)), x0, y0)


def box_for_panels(length, beam, draft, target: int, x0=0.0, y0=0.0) -> PanelMesh:
    area = 2 * (length + beam) * draft + length * beam
    s = math.sqrt(area / target)
    nl, nb, nt = (max(1, round(length / s)), max(1, round(beam / s)), max(1, round(draft / s)))
    return box(length, beam, draft, nl, nb, nt, x0, y0)


@dataclass
class Combined:
    """Combined multi-body arrays in the layout hamsx_set_mesh takes."""
    xyz: np.ndarray        # (nnode, 3) Fortran order
    ncn: np.ndarray        # (nelem,)
    ncon: np.ndarray       # (nelem, 4) Fortran order, 1-based
    cen: np.ndarray        # (nelem, 3) Fortran order
    area: np.ndarray
    pnsz: np.ndarray
    dxyz: np.ndarray       # (nelem, 6) Fortran order
    body_start: np.ndarray  # (nbody + 1,)

    @property
    def nelem(self) -> int:
        return len(self.ncn)

    @property
    def nnode(self) -> int:
        return len(self.xyz)

    @property
    def nbody(self) -> int:
        return len(self.body_start) - 1


def half_mesh(mesh: PanelMesh, axis: int) -> PanelMesh:
    """The panels of a mesh symmetric about the plane axis = 0 (0: x, 1: y)
    whose centroids lie on its positive side, nodes renumbered, with ISX /
    ISY set: the input the reference expects with a symmetry plane."""
    xyz = np.asarray(mesh.xyz)
    keep = [e for e in range(len(mesh.ncn)) if xyz[mesh.ncon[e, :mesh.ncn[e]] - 1].mean(axis=0)[axis] > 0]
    ncon = mesh.ncon[keep]
    used = np.unique(ncon[ncon > 0])
    remap = np.zeros(len(xyz) + 1, dtype=np.int32)
    remap[used] = np.arange(1, len(used) + 1)
    nc = np.where(ncon > 0, remap[ncon], 0).astype(np.int32)
    return PanelMesh(np.asfortranarray(xyz[used - 1]), mesh.ncn[keep].copy(), np.asfortranarray(nc),
                     1 if axis == 0 else 0, 1 if axis == 1 else 0)


def panel_geometry(mesh: PanelMesh, xr) -> tuple:
    """hamsx_panel_geometry on one body (NormalProcess.f90 semantics)."""
    lib = _lib.load()
    xyz = np.asfortranarray(mesh.xyz, dtype=np.float64)
    ncn = np.ascontiguousarray(mesh.ncn, dtype=np.int32)
    ncon = np.asfortranarray(mesh.ncon, dtype=np.int32)
    ne = mesh.nelem
    cen = np.zeros((ne, 3), order="F")
    area = np.zeros(ne)
    pnsz = np.zeros(ne)
    dxyz = np.zeros((ne, 6), order="F")
    xr = np.asarray(xr, dtype=np.float64)
    import ctypes as C
    _lib.check(lib.hamsx_panel_geometry(mesh.nnode, _lib.ptr(xyz), ne, _lib.ptr(ncn, C.c_int),
                                        _lib.ptr(ncon, C.c_int), _lib.ptr(xr), _lib.ptr(cen),
                                        _lib.ptr(area), _lib.ptr(pnsz), _lib.ptr(dxyz)),
               "hamsx_panel_geometry")
    return cen, area, pnsz, dxyz


def combine(bodies: list, xr: list | None = None) -> Combined:
    """BuildCombinedArrays: concatenate bodies; xr[b] is body b's reference point."""
    if xr is None:
        xr = [np.zeros(3)] * len(bodies)
    xyzs, ncns, ncons, cens, areas, pnszs, dxyzs, starts = [], [], [], [], [], [], [], [0]
    noff = 0
    for b, m in enumerate(bodies):
        cen, area, pnsz, dxyz = panel_geometry(m, xr[b])
        xyzs.append(m.xyz)
        ncns.append(m.ncn)
        nc = m.ncon.copy()
        nc[nc > 0] += noff
        ncons.append(nc)
        cens.append(cen)
        areas.append(area)
        pnszs.append(pnsz)
        dxyzs.append(dxyz)
        noff += m.nnode
        starts.append(starts[-1] + m.nelem)
    return Combined(np.asfortranarray(np.vstack(xyzs)), np.concatenate(ncns).astype(np.int32),
                    np.asfortranarray(np.vstack(ncons).astype(np.int32)), np.asfortranarray(np.vstack(cens)),
                    np.concatenate(areas), np.concatenate(pnszs), np.asfortranarray(np.vstack(dxyzs)),
                    np.array(starts, dtype=np.int32))
