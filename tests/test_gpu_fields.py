"""GPU parity of the field-point path (hamsx_set_field_points /
hamsx_field_values: CalPotentialMulti, CalPressureMulti, CalElevationMulti,
src/PressureElevationMulti.f90:55-348) against the oracle, and end to end
against the reference's shipped example2 free-surface elevations.

Tolerances: the oracle is fed the GPU's own potentials, so only the field
formula is compared -- max |difference| per mode column over max |value| of
the column within 1e-10 (deep water), or within 4x the distance between the
oracle with and without FMA contraction in finite depth (the same reference
rounding floor as the assembly parity tests)."""
import sys

import numpy as np
import pytest

import oracle
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver
from test_oracle_fields import check_example2

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pair():
    bodies = [cylinder(2.0, 3.0, 20, 5, 3, -6.0), cylinder(1.5, 2.5, 16, 5, 3, 6.0)]
    return combine(bodies, [np.array([-6.0, 0, 0]), np.array([6.0, 0, 0])])


def field_points(m, rng):
    """Centroids (on the body), points just off panels, fluid points near
    and far, and free-surface points (z = 0)."""
    cen = np.asarray(m.cen)
    nrm = np.asarray(m.dxyz)[:, :3]
    idx = rng.choice(m.nelem, 12, replace=False)
    pts = [cen[idx[:6]], cen[idx[6:]] - 0.05 * m.pnsz[idx[6:], None] * nrm[idx[6:]]]
    near = np.column_stack([rng.uniform(-12, 12, 10), rng.uniform(-6, 6, 10), rng.uniform(-3.5, -0.2, 10)])
    far = np.column_stack([rng.uniform(-300, 300, 6), rng.uniform(-300, 300, 6), rng.uniform(-9, -1, 6)])
    fs = np.column_stack([rng.uniform(-40, 40, 8), rng.uniform(-40, 40, 8), np.zeros(8)])
    keep = np.hypot(near[:, 0] + 6, near[:, 1]) > 2.5
    keep &= np.hypot(near[:, 0] - 6, near[:, 1]) > 2.0
    return np.vstack(pts + [near[keep], far, fs])


def col_err(got, ref):
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return np.max(np.abs(got - ref), axis=0) / scale


@pytest.mark.parametrize("depth,w", [(-1.0, 0.9), (-1.0, 2.2), (8.0, 1.1), (40.0, 0.6), (40.0, 1.8)])
def test_field_values_match_oracle(pair, depth, w):
    m = pair
    rng = np.random.default_rng(11)
    xfp = field_points(m, rng)
    heads = np.deg2rad([0.0, 45.0, 120.0])
    say(f"[gpu] field values n={m.nelem} points={len(xfp)} depth={depth} w={w}")
    with FrequencySolver(m) as s:
        s.set_waves(heads, amp=1.3, xw=(0.5, -1.0), rho=1025.0, isol=1)
        s.calwaveproperts(w, depth)
        s.assemble()
        assert s.factor() == 0
        phi = np.hstack([s.radiation_solver(), s.diffraction_solver()])
        s.set_field_points(xfp)
        got = {}
        for kind, code in (("potential", 0), ("pressure", 1), ("elevation", 2)):
            got[code], ob = s.field_values(kind)
        auto, _, inc = s.field_values("auto", incident=True)
    for code in (0, 1, 2):
        ref, ob_ref = oracle.field_values(m, phi, xfp, w, depth, heads, kind=code, amp=1.3, xw=(0.5, -1.0),
                                          rho=1025.0, isol=1)
        np.testing.assert_array_equal(ob, ob_ref)
        err = col_err(got[code], ref)
        floor = np.zeros_like(err)
        if depth > 0:
            ref_f, _ = oracle.field_values(m, phi, xfp, w, depth, heads, kind=code, amp=1.3, xw=(0.5, -1.0),
                                           rho=1025.0, isol=1, fma=True)
            floor = col_err(ref_f, ref)
        say(f"    kind {code}: max col err {err.max():.3e}  floor {floor.max():.3e}  on-body {int(ob.sum())}")
        assert np.all(err <= np.maximum(1e-10, 4 * floor)), (code, err, floor)
    # AUTO: pressure below the free surface, elevation on it
    below = np.abs(xfp[:, 2]) > 1e-6
    np.testing.assert_array_equal(auto[below], got[1][below])
    np.testing.assert_array_equal(auto[~below], got[2][~below])
    # the incident wave alone, as the elevation / pressure of VINP
    v, wk, _ = oracle.wave(w, depth)
    for b, be in enumerate(heads):
        for q, x in enumerate(xfp):
            f = np.zeros(2)
            oracle.lib().orc_vinp(x[0], x[1], x[2], 0.5, -1.0, be, v, w, wk, depth, 1.3, oracle._p(f))
            phi_i = complex(f[0], f[1])
            fac = 1j * w * 1025.0 if below[q] else 1j * w / oracle.G
            assert abs(inc[q, b] - fac * phi_i) <= 1e-12 * max(1.0, abs(fac * phi_i))


def test_field_points_on_the_body_return_the_solution(pair):
    """At the panel centroids (free term 2 pi) the field potential is the
    solved potential itself (plus VINP for diffraction when ISOL = 1)."""
    m = pair
    depth, w = 30.0, 1.3
    heads = np.deg2rad([30.0])
    say(f"[gpu] field values at the centroids n={m.nelem}")
    with FrequencySolver(m) as s:
        s.set_waves(heads, isol=2)
        s.calwaveproperts(w, depth)
        s.assemble()
        assert s.factor() == 0
        phi = np.hstack([s.radiation_solver(), s.diffraction_solver()])
        s.set_field_points(np.asarray(m.cen))
        val, ob = s.field_values("potential")
    assert ob.all()
    assert np.max(np.abs(val - phi)) <= 1e-9 * np.max(np.abs(phi))


@pytest.mark.parametrize("depth", [-1.0, 25.0])
def test_field_values_in_passes_and_grid_form_agree(pair, depth, monkeypatch):
    """Field points handled in passes of 7 (the pass loop of large point
    sets), and by the unbinned grid kernel, give the one-pass binned values."""
    m = pair
    rng = np.random.default_rng(5)
    xfp = field_points(m, rng)
    heads = np.deg2rad([0.0, 90.0])
    say(f"[gpu] field passes / grid form depth={depth} points={len(xfp)}")
    out = {}
    for tag, env in (("one", {}), ("passes", {"HAMSX_FIELD_CHUNK": "7"}), ("grid", {"HAMSX_ASSEMBLE_GRID": "1"})):
        for k in ("HAMSX_FIELD_CHUNK", "HAMSX_ASSEMBLE_GRID"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with FrequencySolver(m) as s:
            s.set_waves(heads)
            s.calwaveproperts(0.9, depth)
            s.assemble()
            s.set_field_points(xfp)
            out[tag] = s.field_values("potential")[0]
    scale = np.max(np.abs(out["one"]))
    assert np.max(np.abs(out["passes"] - out["one"])) <= 1e-12 * scale
    assert np.max(np.abs(out["grid"] - out["one"])) <= 1e-12 * scale


def test_field_values_need_points_and_reject_bad_kind(pair):
    from hamsx._lib import HamsxError
    with FrequencySolver(pair) as s:
        s.set_waves([0.0])
        s.calwaveproperts(1.0, -1.0)
        with pytest.raises(HamsxError):
            s.field_values("potential")
        s.set_field_points(np.array([[30.0, 0.0, -1.0]]))
        rc = s._lib.hamsx_field_values(s._h, 7, None, None, None)
        assert rc == -1


@pytest.fixture(scope="module")
def example2():
    from test_oracle_fields import load
    from hamsx.deck import PanelMesh
    z = load("example2.npz")
    bodies = [PanelMesh(np.asfortranarray(z["xyz"][b]), z["ncn"], np.asfortranarray(z["ncon"]))
              for b in range(len(z["xr"]))]
    return z, combine(bodies, list(z["xr"]))


@pytest.mark.parametrize("k", [0, 1])
def test_example2_field_elevations_on_gpu(example2, k):
    """The whole GPU path -- assembly, LU, solves, field values -- against
    the shipped example2 .6p elevations (both periods, 12 headings, 10
    bodies, 100 points), to the bounds the oracle meets."""
    z, m = example2
    w = 2 * np.pi / z["periods"][k]
    heads = np.deg2rad(z["headings_deg"])
    say(f"[gpu] example2 period {z['periods'][k]} n={m.nelem} points={len(z['xfp'])}")
    with FrequencySolver(m) as s:
        s.set_waves(heads, isol=int(z["isol"]))
        s.set_field_points(z["xfp"])
        _, _, _, info = s.solve_frequency(w, float(z["depth"]))
        assert info == 0
        vals, ob, inc = s.field_values("auto", incident=True)
    assert not ob.any()
    check_example2(z, k, vals, inc)
