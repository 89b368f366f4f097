"""The seeded synthetic workloads (cylinders and boxes) the benchmarks and
parity tests run on.  Bodies are generated with the reference's panel
orientation (see :mod:`hamsx.mesh`); the reference point of each body is its
waterline centre."""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import box_for_panels, cylinder, cylinder_for_panels, half_mesh, waterplane_disk


@dataclass
class Case:
    name: str
    bodies: list
    xr: list
    depth: float                  # < 0: infinite depth
    omegas: np.ndarray            # rad/s
    headings_deg: list
    rho: float = 1000.0
    amp: float = 1.0
    xw: tuple = (0.0, 0.0)
    meta: dict = field(default_factory=dict)

    @property
    def nelem(self) -> int:
        return sum(b.nelem for b in self.bodies)


def single_cylinder(na=48, nz=12, nr=6) -> Case:
    """One truncated cylinder R = 5 m, draft 10 m, 48 x 12 side quads and a
    6-ring bottom (864 panels), infinite depth, w = 1.0 rad/s, heading 0."""
    b = cylinder(5.0, 10.0, na, nz, nr)
    return Case("cylinder-1x864-deep", [b], [np.zeros(3)], -1.0, np.array([1.0]), [0.0])


def two_cylinders(panels_each=1984, nfreq=30) -> Case:
    """Two cylinders R = 5 m, draft 10 m (~2,000 panels each) at x = -20 / +20 m,
    depth 50 m, 30 frequencies uniform in [0.2, 2.5] rad/s, headings 0/45/90."""
    if panels_each == 1984:
        bodies = [cylinder(5.0, 10.0, 64, 21, 10, x0) for x0 in (-20.0, 20.0)]
    else:
        bodies = [cylinder_for_panels(5.0, 10.0, panels_each, x0) for x0 in (-20.0, 20.0)]
    xr = [np.array([x0, 0.0, 0.0]) for x0 in (-20.0, 20.0)]
    return Case("cylinders-2x1984-h50", bodies, xr, 50.0, np.linspace(0.2, 2.5, nfreq), [0.0, 45.0, 90.0])


def two_cylinders_half(nfreq=30) -> Case:
    """two_cylinders() through its symmetry plane y = 0 (ISY = 1): the halves
    y > 0 of both cylinders (992 panels each), the same depth, frequencies
    and headings."""
    full = two_cylinders(nfreq=nfreq)
    bodies = [half_mesh(b, 1) for b in full.bodies]
    return Case("cylinders-2x1984-h50-half-isy", bodies, full.xr, full.depth, full.omegas, full.headings_deg,
                meta={"symmetry": (0, 1)})


def five_cylinders(seed=2024, total=10000, nfreq=100) -> Case:
    """Five cylinders, R ~ U[3,6] m, draft ~ U[6,12] m, centres on a line 35 m
    apart +/- U[0,3] m, ~10,000 panels, depth 60 m, 100 frequencies in
    [0.1, 3.0] rad/s, heading 0."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(3.0, 6.0, 5)
    drafts = rng.uniform(6.0, 12.0, 5)
    xs = np.cumsum([0.0] + list(35.0 + rng.uniform(0.0, 3.0, 4)))
    xs = xs - xs.mean()
    bodies = [cylinder_for_panels(r, d, total // 5, x) for r, d, x in zip(radii, drafts, xs)]
    xr = [np.array([x, 0.0, 0.0]) for x in xs]
    return Case(f"cylinders-5x2000-h60-seed{seed}", bodies, xr, 60.0, np.linspace(0.1, 3.0, nfreq), [0.0],
                meta={"seed": seed})


def large_array(panels_each=1875, nfreq=50, grid=4) -> Case:
    """16 cylinders R = 4 m, draft 8 m on a 4 x 4 grid, 30 m spacing
    (about 30,000 hull panels), depth 80 m, 50 frequencies in [0.2, 2.0]
    rad/s, heading 0.  Each cylinder carries a waterplane lid for
    irregular-frequency removal, with panels about twice the hull's size
    (meta["lids"]).  grid < 4: the same layout on a grid x grid array."""
    bodies, xr, lids = [], [], []
    c0 = 15.0 * (grid - 1)
    for k in range(grid * grid):
        x0, y0 = 30.0 * (k % grid) - c0, 30.0 * (k // grid) - c0
        b = cylinder_for_panels(4.0, 8.0, panels_each, x0, y0)
        bodies.append(b)
        xr.append(np.array([x0, y0, 0.0]))
        na = max(8, int(round(2 * math.pi * 4.0 / (2.0 * math.sqrt(4.0 * 8.0 * 2 * math.pi / panels_each)))))
        lids.append(waterplane_disk(4.0, na, max(1, na // 6), x0, y0))
    return Case(f"array-{grid * grid}x{panels_each}-lid-h80", bodies, xr, 80.0, np.linspace(0.2, 2.0, nfreq), [0.0],
                meta={"lids": lids})


def random_request(rng: np.random.Generator, pmin=500, pmax=12000, total=None) -> Case:
    """One serving request: 1-9 cylinders or boxes (L, B ~ U[5,20] m, draft
    ~ U[4,10] m), 500-12,000 panels, w ~ U[0.1,3.0], heading ~ U[0,180) deg,
    depth infinite or U[30,200] m.  total: the panel count to aim at instead
    of a draw from [pmin, pmax] (the other draws are unchanged)."""
    nb = int(rng.integers(1, 10))
    drawn = int(rng.integers(pmin, pmax + 1))
    total = drawn if total is None else int(total)
    per = max(60, total // nb)
    bodies, xr = [], []
    for k in range(nb):
        x0, y0 = 40.0 * (k % 3), 40.0 * (k // 3)
        if rng.random() < 0.5:
            bodies.append(cylinder_for_panels(rng.uniform(3.0, 8.0), rng.uniform(4.0, 10.0), per, x0, y0))
        else:
            bodies.append(box_for_panels(rng.uniform(5, 20), rng.uniform(5, 20), rng.uniform(4, 10), per, x0, y0))
        xr.append(np.array([x0, y0, 0.0]))
    depth = -1.0 if rng.random() < 0.5 else float(rng.uniform(30.0, 200.0))
    return Case("request", bodies, xr, depth, np.array([rng.uniform(0.1, 3.0)]), [float(rng.uniform(0.0, 180.0))])
