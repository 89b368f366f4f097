"""Geometric symmetry (NSYS = 2) in the oracle: the half mesh solved with one
symmetry plane -- the reference's mirrored field points, RXY / SY / RX / RY
sign tables, two systems and the MXPOT recombination -- against the full
mesh solved directly.  The two are the same discretised problem (the full
mesh is the half and its mirror image), so they agree to rounding: 1e-13 in
deep water, and within the finite-depth Green function's rounding
sensitivity (1e-6) at finite depth.  A wrong sign in any table breaks the
agreement at O(1)."""
import numpy as np
import pytest

import oracle
from hamsx.deck import PanelMesh
from hamsx.mesh import cylinder
from test_oracle_config1 import combined_with_oracle_geometry


def half(mesh, axis):
    """The panels of mesh whose centroid lies on the positive side of the
    plane axis = 0 (0: x, 1: y), nodes renumbered."""
    xyz = np.asarray(mesh.xyz)
    keep = [e for e in range(len(mesh.ncn))
            if xyz[mesh.ncon[e, :mesh.ncn[e]] - 1].mean(axis=0)[axis] > 0]
    ncon = mesh.ncon[keep]
    used = np.unique(ncon[ncon > 0])
    remap = np.zeros(len(xyz) + 1, dtype=np.int32)
    remap[used] = np.arange(1, len(used) + 1)
    nc = np.where(ncon > 0, remap[ncon], 0).astype(np.int32)
    return PanelMesh(np.asfortranarray(xyz[used - 1]), mesh.ncn[keep].copy(), np.asfortranarray(nc))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth,w,tol", [(-1.0, 0.9, 1e-13), (8.0, 1.3, 1e-6)])
def test_half_mesh_with_symmetry_equals_full_mesh(axis, depth, w, tol):
    full = cylinder(2.0, 3.0, 16, 5, 3)
    mf = combined_with_oracle_geometry([full], [np.zeros(3)])
    mh = combined_with_oracle_geometry([half(full, axis)], [np.zeros(3)])
    assert 2 * mh.nelem == mf.nelem
    heads = np.deg2rad([0.0, 30.0, 90.0])
    af, bf, ef, _ = oracle.solve_frequency(mf, w, depth, heads)
    ah, bh, eh, info = oracle.solve_frequency_sym(mh, 1 if axis == 0 else 0, w, depth, heads)
    assert info == 0
    assert rel(ah, af) <= tol and rel(bh, bf) <= tol and rel(eh, ef) <= tol, (rel(ah, af), rel(bh, bf), rel(eh, ef))


def test_two_bodies_symmetric_about_y():
    """Two cylinders on the x axis, each symmetric about y = 0 (ISY = 1)."""
    bodies = [cylinder(2.0, 3.0, 16, 5, 3, x0=-6.0), cylinder(1.5, 2.5, 12, 4, 2, x0=6.0)]
    xr = [np.array([-6.0, 0, 0]), np.array([6.0, 0, 0])]
    mf = combined_with_oracle_geometry(bodies, xr)
    mh = combined_with_oracle_geometry([half(b, 1) for b in bodies], xr)
    heads = np.deg2rad([0.0, 45.0])
    af, bf, ef, _ = oracle.solve_frequency(mf, 1.1, -1.0, heads)
    ah, bh, eh, info = oracle.solve_frequency_sym(mh, 0, 1.1, -1.0, heads)
    assert info == 0
    assert rel(ah, af) <= 1e-13 and rel(bh, bf) <= 1e-13 and rel(eh, ef) <= 1e-13


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth", [-1.0, 8.0])
def test_field_values_half_mesh_equal_full_mesh(axis, depth):
    """Field points with one symmetry plane (the IS loops of
    CalPotentialMulti): the half mesh's potentials seen from the point and
    its mirror image give the full mesh's field values -- in the fluid, on
    the free surface, and at a panel centroid of the mirror half (free term
    2 pi found through the mirrored point)."""
    full = cylinder(2.0, 3.0, 16, 5, 3)
    mf = combined_with_oracle_geometry([full], [np.zeros(3)])
    mh = combined_with_oracle_geometry([half(full, axis)], [np.zeros(3)])
    heads = np.deg2rad([0.0, 50.0])
    w = 1.1
    mirror_cen = mf.cen[np.argmax(-mf.cen[:, axis])]          # a centroid on the mirror side
    xfp = np.array([[4.0, 1.5, -1.0], [-3.5, 2.5, 0.0], [1.0, -4.0, -2.5], [0.5, 0.3, -5.0], mirror_cen])
    phi_full, _ = oracle.solve_potentials(mf, w, depth, heads)
    vf, of = oracle.field_values(mf, phi_full, xfp, w, depth, heads, kind=0)
    *_, info, phi2 = oracle.solve_frequency_sym(mh, 1 if axis == 0 else 0, w, depth, heads, potentials=True)
    assert info == 0
    vh, oh = oracle.field_values_sym(mh, 1 if axis == 0 else 0, phi2, xfp, w, depth, heads, kind=0)
    assert list(of) == list(oh) and oh[-1] and not oh[0]
    tol = 1e-12 if depth < 0 else 1e-6
    assert rel(vh, vf) <= tol, rel(vh, vf)


def lid_half(lid_cen, axis):
    return lid_cen[lid_cen[:, axis] > 0]


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth,w,tol", [(-1.0, 1.6, 1e-12), (8.0, 1.3, 1e-6)])
def test_half_mesh_with_lid_and_symmetry_equals_full_mesh(axis, depth, w, tol):
    """IRSP = 1 with one symmetry plane (CALGREEN_MULTI_IRR's IS loops,
    ASSB_LEFT/RBC/DBC_IRR_MULTI with NSYS = 2): the half hull and half lid
    with the plane against the full hull and lid, each through the normal
    equations C = A^T A of its own system."""
    from hamsx.mesh import panel_geometry, waterplane_disk
    full = cylinder(2.0, 3.0, 16, 5, 3)
    lid, _, _, _ = panel_geometry(waterplane_disk(2.0, 16, 2), np.zeros(3))
    mf = combined_with_oracle_geometry([full], [np.zeros(3)])
    mh = combined_with_oracle_geometry([half(full, axis)], [np.zeros(3)])
    lh = lid_half(lid, axis)
    assert 2 * len(lh) == len(lid)
    heads = np.deg2rad([0.0, 30.0, 90.0])
    af, bf, ef, _ = oracle.solve_frequency_irr(mf, lid, w, depth, heads)
    ah, bh, eh, info = oracle.solve_frequency_sym(mh, 1 if axis == 0 else 0, w, depth, heads, lid_cen=lh)
    assert info == 0
    errs = rel(ah, af), rel(bh, bf), rel(eh, ef)
    assert max(errs) <= tol, errs
    a0, _, _, _ = oracle.solve_frequency_sym(mh, 1 if axis == 0 else 0, w, depth, heads)
    assert rel(a0, af) > 10 * tol          # the lid changes the answer
