"""GPU geometric symmetry (hamsx_set_symmetry, NSYS = 2: the mirrored field
points of CalGreenFuncMulti.f90:141-177, the two systems and right-hand sides
of AssbMatxMulti.f90:53-251, the MXPOT recombination :306-316 / :373-381 and
the force sums over both halves of PotentWavForceMulti.f90) against the
oracle's symmetric solve on the same half mesh, and against the GPU's own
solve of the full mesh.

Tolerances as in test_gpu_parity.py: deep water 1e-8 on the coefficients and
1e-10 on the potentials; finite depth max(1e-8, 4x the FMA / no-FMA distance
of the oracle on the same case)."""
import sys

import numpy as np
import pytest

import gates
import oracle
from hamsx._lib import HamsxError
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver
from test_oracle_symmetry import half

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def two_bodies(axis):
    """Two cylinders side by side along the other axis, each symmetric about
    the plane axis = 0."""
    if axis == 1:
        bodies = [cylinder(4.0, 8.0, 24, 6, 3, -12.0), cylinder(3.0, 6.0, 20, 5, 3, 12.0)]
        xr = [np.array([-12.0, 0, 0]), np.array([12.0, 0, 0])]
    else:
        bodies = [cylinder(4.0, 8.0, 24, 6, 3), cylinder(3.0, 6.0, 20, 5, 3)]
        for b, y0 in zip(bodies, (-12.0, 12.0)):
            b.xyz[:, 1] += y0
        xr = [np.array([0, -12.0, 0]), np.array([0, 12.0, 0])]
    return bodies, xr


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth,w", [(-1.0, 0.8), (30.0, 1.2)])
def test_symmetric_solve_matches_oracle(axis, depth, w):
    bodies, xr = two_bodies(axis)
    m = combine([half(b, axis) for b in bodies], xr)
    heads = np.deg2rad([0.0, 45.0, 120.0])
    isx, isy = (1, 0) if axis == 0 else (0, 1)
    say(f"[gpu] symmetry isx={isx} isy={isy} n={m.nelem} depth={depth} w={w}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_symmetry(isx, isy)
        am, bd, ex, info = s.solve_frequency(w, depth)
        assert info == 0
        s.calwaveproperts(w, depth)
        s.assemble()
        rad, dif = s.radiation_solver(), s.diffraction_solver()
    ao, bo, eo, io, phi = oracle.solve_frequency_sym(m, isx, w, depth, heads, potentials=True, nthreads=16)
    assert io == 0
    dr = gates.solve_draws(oracle.solve_frequency_sym, m, isx, w, depth, heads, nthreads=16) if depth > 0 else []
    nd = 6 * m.nbody
    ep = rel(np.concatenate([rad, dif], axis=1), phi)
    say(f"    MXPOT rel err {ep:.3e}")
    assert rad.shape == (m.nelem, nd, 2) and dif.shape == (m.nelem, len(heads), 2)
    gates.coef_gates((am, bd, ex), (ao, bo, eo), dr, 1e-8, say)
    if depth < 0:
        assert ep <= 1e-10


@pytest.mark.parametrize("axis", [0, 1])
def test_half_mesh_equals_full_mesh_on_gpu(axis):
    """The half mesh with the plane and the full mesh without it are the same
    discrete problem: deep water agrees to rounding."""
    bodies, xr = two_bodies(axis)
    heads = np.deg2rad([0.0, 60.0])
    say(f"[gpu] symmetry half vs full axis={axis}")
    with FrequencySolver(combine(bodies, xr)) as s:
        s.set_waves(heads)
        af, bf, ef, _ = s.solve_frequency(1.0, -1.0)
    with FrequencySolver(combine([half(b, axis) for b in bodies], xr)) as s:
        s.set_waves(heads)
        s.set_symmetry(*((1, 0) if axis == 0 else (0, 1)))
        ah, bh, eh, info = s.solve_frequency(1.0, -1.0)
    assert info == 0
    errs = rel(ah, af), rel(bh, bf), rel(eh, ef)
    say(f"    rel err {errs}")
    assert max(errs) <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_stream_symmetric_schedule_equals_one_stream(precision, monkeypatch):
    """The two systems of a symmetry plane factored and solved at once on two
    streams give the one-stream schedule's factors, potentials and
    coefficients bit for bit (HAMSX_SYM_STREAMS=0 selects one stream)."""
    bodies, xr = two_bodies(1)
    m = combine([half(b, 1) for b in bodies], xr)
    heads = np.deg2rad([0.0, 60.0])
    say(f"[gpu] symmetry two streams vs one n={m.nelem} {precision}")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HAMSX_SYM_STREAMS", mode)
        with FrequencySolver(m, precision=precision) as s:
            s.set_waves(heads)
            s.set_symmetry(0, 1)
            coef = s.solve_frequency(1.1, 30.0)
            s.calwaveproperts(1.1, 30.0)
            s.assemble()
            assert s.factor() == 0
            lu, ipiv = s.get_matrix("LU")
            out[mode] = coef[:3] + (lu, ipiv, s.radiation_solver(), s.diffraction_solver())
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a, b)


def test_symmetry_arguments_and_state():
    """Both planes are refused as the reference refuses them; a new mesh
    removes the plane."""
    bodies, xr = two_bodies(1)
    m = combine([half(b, 1) for b in bodies], xr)
    with FrequencySolver(m) as s:
        s.set_waves([0.0])
        with pytest.raises(HamsxError):
            s.set_symmetry(1, 1)
        s.set_symmetry(0, 1)
        a1, _, _, _ = s.solve_frequency(1.0, -1.0)
        s.set_mesh(m)
        assert s.nsys == 1
        s.set_waves([0.0])
        a0, _, _, _ = s.solve_frequency(1.0, -1.0)
        assert rel(a0, a1) > 1e-3      # the half mesh alone is a different body


@pytest.mark.parametrize("lid", [False, True])
@pytest.mark.parametrize("depth", [-1.0, 30.0])
def test_symmetric_gmres_and_mixed_match_fp64_lu(lid, depth):
    """The other solver paths under a symmetry plane, each running the two
    systems: GMRES (dense operator, tight tolerance) and the complex64 LU
    with FP64 refinement against the FP64 LU, with and without a lid."""
    bodies, xr = two_bodies(0)
    m = combine([half(b, 0) for b in bodies], xr)
    lh, _ = half_lid(0)
    heads = np.deg2rad([0.0, 45.0])
    w = 1.1
    say(f"[gpu] symmetry gmres / mixed lid={lid} depth={depth} n={m.nelem}")

    def run(precision="fp64", solver=None):
        with FrequencySolver(m, precision=precision) as s:
            s.set_waves(heads)
            s.set_symmetry(1, 0)
            if lid:
                s.set_lid(lh)
            if solver == "gmres":
                s.set_hmatrix(False)
                s.set_solver("gmres", tol=1e-11, restart=25, maxiter=400)
            out = s.solve_frequency(w, depth)
            its, res = s.refine_stats()
        assert out[3] == 0
        return out[:3], its, res

    ref, _, _ = run()
    gm, gits, gres = run(solver="gmres")
    mx, mits, mres = run(precision="mixed")
    eg = [rel(a, b) for a, b in zip(gm, ref)]
    em = [rel(a, b) for a, b in zip(mx, ref)]
    say(f"    gmres {gits} iterations, residual {gres:.1e}, rel err {max(eg):.2e}; "
        f"mixed {mits} steps, residual {mres:.1e}, rel err {max(em):.2e}")
    assert max(eg) <= 1e-8
    assert max(em) <= 1e-6



@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth", [-1.0, 30.0])
@pytest.mark.parametrize("chunk", [None, "2"])
def test_symmetric_field_values_match_oracle(axis, depth, chunk, monkeypatch):
    """hamsx_field_values with a symmetry plane (each point and its mirror
    image against the half mesh, PressureElevationMulti.f90:89-249) against
    the oracle's restatement fed the GPU's own potentials of both halves.
    Points in the fluid, on the free surface, on a panel of the mesh half and
    on a panel of the mirror half (found through the mirrored point);
    chunk = 2 splits the points into one-point passes."""
    if chunk:
        monkeypatch.setenv("HAMSX_FIELD_CHUNK", chunk)
    bodies, xr = two_bodies(axis)
    m = combine([half(b, axis) for b in bodies], xr)
    isx = 1 if axis == 0 else 0
    heads = np.deg2rad([0.0, 45.0, 120.0])
    w = 0.9
    on_half = m.cen[5]
    on_mirror = on_half.copy()
    on_mirror[axis] = -on_mirror[axis]
    xfp = np.array([[30.0, 25.0, -3.0], [-20.0, 5.0, 0.0], [2.0, -33.0, -9.0], on_half, on_mirror])
    say(f"[gpu] symmetric field values axis={axis} depth={depth} chunk={chunk}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_symmetry(isx, 1 - isx)
        s.calwaveproperts(w, depth)
        s.assemble()
        phi2 = np.concatenate([s.radiation_solver(), s.diffraction_solver()], axis=1)
        s.set_field_points(xfp)
        vals, ob = s.field_values("potential")
    ref, obr = oracle.field_values_sym(m, isx, phi2, xfp, w, depth, heads, kind=0)
    e = rel(vals, ref)
    floor = 0.0
    if depth > 0:
        rf, _ = oracle.field_values_sym(m, isx, phi2, xfp, w, depth, heads, kind=0, fma=True)
        floor = rel(rf, ref)
    say(f"    onbody {list(ob)}  rel err {e:.3e}  floor {floor:.1e}")
    assert list(ob) == list(obr) == [False, False, False, True, True]
    assert e <= max(1e-10, 4 * floor)


def half_lid(axis):
    """Waterplane lid points of the two cylinders of two_bodies(axis) on the
    positive side of the plane, and the full set."""
    from hamsx.mesh import panel_geometry, waterplane_disk
    pts = []
    for r, c in ((4.0, -12.0), (3.0, 12.0)):
        lid, _, _, _ = panel_geometry(waterplane_disk(r, 24 if r == 4.0 else 20, 3), np.zeros(3))
        lid = np.array(lid)
        lid[:, 1 - axis] += c
        pts.append(lid)
    full = np.vstack(pts)
    return full[full[:, axis] > 0], full


@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("depth,w", [(-1.0, 1.4), (30.0, 1.2)])
def test_symmetric_solve_with_lid_matches_oracle(axis, depth, w):
    """IRSP = 1 with a symmetry plane: half hull and half lid, each system
    through its own normal equations, against the oracle; the order of
    hamsx_set_lid and hamsx_set_symmetry does not matter (bit for bit)."""
    bodies, xr = two_bodies(axis)
    m = combine([half(b, axis) for b in bodies], xr)
    lh, _ = half_lid(axis)
    heads = np.deg2rad([0.0, 45.0])
    isx = 1 if axis == 0 else 0
    say(f"[gpu] symmetry + lid axis={axis} n={m.nelem} lid={len(lh)} depth={depth} w={w}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_lid(lh)
        s.set_symmetry(isx, 1 - isx)
        am, bd, ex, info = s.solve_frequency(w, depth)
        assert info == 0
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_symmetry(isx, 1 - isx)
        s.set_lid(lh)
        am2, bd2, ex2, _ = s.solve_frequency(w, depth)
    for a, b in ((am, am2), (bd, bd2), (ex, ex2)):
        np.testing.assert_array_equal(a, b)
    ref = oracle.solve_frequency_sym(m, isx, w, depth, heads, lid_cen=lh, nthreads=16)
    assert ref[3] == 0
    dr = (gates.solve_draws(oracle.solve_frequency_sym, m, isx, w, depth, heads, lid_cen=lh, nthreads=16)
          if depth > 0 else [])
    gates.coef_gates((am, bd, ex), ref, dr, 1e-8, say)


@pytest.mark.parametrize("axis", [0, 1])
def test_half_mesh_with_lid_equals_full_mesh_on_gpu(axis):
    bodies, xr = two_bodies(axis)
    lh, lf = half_lid(axis)
    heads = np.deg2rad([0.0, 60.0])
    say(f"[gpu] symmetry + lid half vs full axis={axis}")
    with FrequencySolver(combine(bodies, xr)) as s:
        s.set_waves(heads)
        s.set_lid(lf)
        af, bf, ef, _ = s.solve_frequency(1.4, -1.0)
    with FrequencySolver(combine([half(b, axis) for b in bodies], xr)) as s:
        s.set_waves(heads)
        s.set_lid(lh)
        s.set_symmetry(*((1, 0) if axis == 0 else (0, 1)))
        ah, bh, eh, info = s.solve_frequency(1.4, -1.0)
    assert info == 0
    errs = rel(ah, af), rel(bh, bf), rel(eh, ef)
    say(f"    rel err {errs}")
    assert max(errs) <= 1e-9


@pytest.mark.parametrize("quad", [False, True])
def test_symmetric_motions_equal_full_mesh(quad):
    """hamsx_motions under a symmetry plane (forces from both halves) against
    the full mesh's, deep water.  Linear damping: one complex solve, 1e-10.
    Quadratic damping: the fixed point stops at RERR <= 1e-6, and last-bit
    differences of the coefficients can move the stop by one step, so the
    bar is that tolerance."""
    from test_gpu_motions import buoy_matrices
    bodies, xr = two_bodies(1)
    mats = buoy_matrices(2, np.random.default_rng(11), quad)
    args = [mats[k] for k in ("mass", "blnr", "bqdr", "crs", "kstf")]
    heads = np.deg2rad([0.0, 60.0])
    say(f"[gpu] symmetric motions vs full mesh quadratic={quad}")
    out = []
    for m, sym in ((combine(bodies, xr), None), (combine([half(b, 1) for b in bodies], xr), (0, 1))):
        with FrequencySolver(m) as s:
            s.set_waves(heads)
            if sym:
                s.set_symmetry(*sym)
            s.set_body_matrices(*args)
            _, _, _, info = s.solve_frequency(0.9, -1.0)
            assert info == 0
            out.append(s.motions())
    (df, itf), (dh, ith) = out
    e = rel(dh, df)
    say(f"    rel err {e:.3e}  steps {list(itf)} / {list(ith)}")
    assert e <= (1e-6 if quad else 1e-10)
