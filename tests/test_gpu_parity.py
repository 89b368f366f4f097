"""GPU parity of the HIP path against the C oracle (FP64 restatement of the
reference, pinned to the reference's own Fortran in test_oracle_golden.py and
test_reference_spread.py).

Deep water: influence matrices within 1e-13 of the largest entry, entry by
entry; added mass / damping / excitation within 1e-8 of each table's largest
entry (plain FP64 path), 1e-6 (mixed precision).

Finite depth: the reference's PIDCOCK / EIGENE epsilon extrapolation moves
its own result under rounding changes alone (tests/test_reference_spread.py),
so the arrays are held to the gates of tests/gates.py: region by region,
entry by entry within 1e-13 where three rounding draws of the oracle agree
(EIGEN, LINTON), and the distribution of the distances within 4x the draws'
where they spread; coefficient tables within max(tol, 4x the draws' largest
distance), with median and rms within 4x the draws'.
"""
import sys

import numpy as np
import pytest

import gates
import oracle
from hamsx import configs
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_pair(na=24, nz=6, nr=3, gap=30.0):
    bodies = [cylinder(4.0, 8.0, na, nz, nr, -gap / 2), cylinder(3.0, 6.0, na, nz, nr, gap / 2)]
    xr = [np.array([-gap / 2, 0, 0]), np.array([gap / 2, 0, 0])]
    return combine(bodies, xr)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def mesh_small():
    return small_pair()


@pytest.mark.parametrize("depth,w", [(-1.0, 1.0), (-1.0, 0.35), (20.0, 0.8), (50.0, 1.5), (200.0, 2.5)])
def test_assembly_matches_oracle(mesh_small, depth, w):
    m = mesh_small
    say(f"[gpu] assembly n={m.nelem} depth={depth} w={w}")
    with FrequencySolver(m) as s:
        s.calwaveproperts(w, depth)
        s.assemble()
        A = s.get_matrix("A")
        S = s.get_matrix("S")
    Ao, So = oracle.assemble(m, w, depth, 16)
    if depth < 0:
        ra = gates.entry_gate("A", A, Ao, [])
        rs = gates.entry_gate("S", S, So, [])
    else:
        dr = gates.assembly_draws(m, w, depth)
        cls, nt = gates.regions(m.cen, m.cen, depth, with_nt=True)
        ra = gates.entry_gate("A", A, Ao, [d[0] for d in dr], cls, nt)
        rs = gates.entry_gate("S", S, So, [d[1] for d in dr], cls, nt)
    say(f"    A: {ra}\n    S: {rs}")


@pytest.mark.parametrize("depth,w", [(20.0, 0.8), (50.0, 1.5), (200.0, 2.5)])
def test_binned_assembly_equals_grid(mesh_small, depth, w, monkeypatch):
    """The region-binned finite-depth kernels give the unbinned kernel's
    entries bit for bit (same arithmetic per pair, only the scheduling
    differs), across all four R/h regions."""
    m = mesh_small
    say(f"[gpu] binned vs grid n={m.nelem} depth={depth} w={w}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("HAMSX_ASSEMBLE_GRID", mode)
        with FrequencySolver(m) as s:
            s.calwaveproperts(w, depth)
            s.assemble()
            out[mode] = (s.get_matrix("A"), s.get_matrix("S"))
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])


@pytest.mark.parametrize("w", [0.2, 1.39, 2.5])
def test_fast_limes_equals_ieee_division(w, monkeypatch):
    """The epsilon tables' fast reciprocal pass (the division sequence
    without its scale and fix-up steps, range-checked) gives the tables of
    IEEE division bit for bit: A and S of the two-cylinder workload (3,968
    panels at h = 50 m: every PIDCOCK and EIGENE series length, 15.7M
    entries) with and without HAMSX_LIMES_IEEE."""
    case = configs.two_cylinders()
    m = combine(case.bodies, case.xr)
    say(f"[gpu] fast Limes vs IEEE division n={m.nelem} depth={case.depth} w={w}")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HAMSX_LIMES_IEEE", mode)
        with FrequencySolver(m) as s:
            s.calwaveproperts(w, case.depth)
            s.assemble()
            out[mode] = (s.get_matrix("A"), s.get_matrix("S"))
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])


@pytest.mark.parametrize("depth,w", [(30.0, 1.2)])
def test_binned_assembly_equals_grid_with_lid(depth, w, monkeypatch):
    """With lid field points the binned assembly mixes single entries (lid
    rows, LINTON pairs) with hull-pair entries (both orientations per
    thread); C = A^T A over all field rows and the hull block of S still
    match the unbinned kernel bit for bit."""
    from hamsx.mesh import waterplane_disk
    hull = cylinder(4.0, 6.0, 24, 6, 3)
    m = combine([hull], [np.zeros(3)])
    lid = lid_centroids(waterplane_disk(4.0, 24, 3))
    say(f"[gpu] binned vs grid with lid n={m.nelem} lid={len(lid)} depth={depth} w={w}")
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("HAMSX_ASSEMBLE_GRID", mode)
        with FrequencySolver(m) as s:
            s.set_lid(lid)
            s.calwaveproperts(w, depth)
            s.assemble()
            out[mode] = (s.get_matrix("A"), s.get_matrix("S"))
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])


def test_lu_backward_error():
    m = small_pair(na=40, nz=8, nr=4)   # 2 x 480 panels: three tournament levels
    say(f"[gpu] LU n={m.nelem}")
    with FrequencySolver(m) as s:
        s.calwaveproperts(1.2, 40.0)
        s.assemble()
        A = s.get_matrix("A")
        info = s.factor()
        LU, ipiv = s.get_matrix("LU")
    assert info == 0
    n = m.nelem
    PA = A.copy()
    for k in range(n):
        p = ipiv[k] - 1
        if p != k:
            PA[[k, p], :] = PA[[p, k], :]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    err = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
    say(f"    ||PA - LU|| / ||A|| = {err:.3e}")
    assert err < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lu_outer_blocks_backward_error(precision):
    """From N = 6,144 on the LU works in outer blocks of four panels (the far
    columns get U = L11^-1 A12 for the block's 128 rows at its end and one
    rank-128 update).  N = 6,400 here: ||PA x - L (U x)|| / (||A|| ||x||) for
    random probes x, and the solves against A."""
    m = small_pair(na=80, nz=35, nr=5)
    assert m.nelem >= 6144
    say(f"[gpu] LU outer blocks n={m.nelem} {precision}")
    with FrequencySolver(m, precision=precision) as s:
        s.set_waves([0.0])
        s.calwaveproperts(1.0, -1.0)
        s.assemble()
        A = s.get_matrix("A")
        info = s.factor()
        LU, ipiv = s.get_matrix("LU")
    assert info == 0
    n = m.nelem
    perm = np.arange(n)
    for k in range(n):
        p = ipiv[k] - 1
        perm[[k, p]] = perm[[p, k]]
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
    ux = np.triu(LU) @ x
    lux = np.tril(LU, -1) @ ux + ux
    err = np.linalg.norm(A[perm] @ x - lux) / (np.linalg.norm(A) * np.linalg.norm(x))
    say(f"    probe backward error {err:.3e}")
    assert err < (1e-14 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("rows,outer_min", [(512, 6144), (1024, 6144), (256, 1024), (512, 1024)])
def test_lu_tuning_options(rows, outer_min):
    """The tuning options (selection blocks of 256/512/1024 rows, outer blocks
    from a given order on) give valid factorisations -- backward error on
    random probes as in test_lu_outer_blocks_backward_error -- and, for each
    option set, the look-ahead and one-stream schedules agree bit for bit.
    N = 2,240: blocks per tournament level 9, 2, 1 with 256-row blocks; 5, 1
    with 512; 3, 1 with 1,024."""
    m = small_pair(na=40, nz=24, nr=4)
    say(f"[gpu] LU options n={m.nelem} rows={rows} outer_min={outer_min}")
    out = []
    for la in (0, 1):
        with FrequencySolver(m) as s:
            s.set_options(lu_select_rows=rows, lu_outer_min_n=outer_min, lu_lookahead=la)
            got = s.options()
            assert {k: got[k] for k in ("lu_lookahead", "lu_select_rows", "lu_outer_min_n")} == \
                {"lu_lookahead": la, "lu_select_rows": rows, "lu_outer_min_n": outer_min}
            s.set_waves([0.0])
            s.calwaveproperts(1.0, 35.0)
            s.assemble()
            A = s.get_matrix("A")
            assert s.factor() == 0
            out.append(s.get_matrix("LU"))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    LU, ipiv = out[1]
    n = m.nelem
    perm = np.arange(n)
    for k in range(n):
        p = ipiv[k] - 1
        perm[[k, p]] = perm[[p, k]]
    rng = np.random.default_rng(6)
    x = rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4))
    ux = np.triu(LU) @ x
    lux = np.tril(LU, -1) @ ux + ux
    err = np.linalg.norm(A[perm] @ x - lux) / (np.linalg.norm(A) * np.linalg.norm(x))
    say(f"    probe backward error {err:.3e}")
    assert err < 1e-14


@pytest.mark.parametrize("precision,outer", [("fp64", False), ("mixed", False), ("fp64", True), ("mixed", True)])
def test_lu_lookahead_gives_the_same_factors(precision, outer, monkeypatch):
    """The look-ahead schedule (the next panel's columns first, the rest of
    the trailing update on a second stream overlapping the next pivot
    selection) computes every entry with the same operations as the one-stream
    schedule: factors, pivots and solutions agree bit for bit.  outer: N =
    6,320 (outer blocks of 128 columns from N = 6,144 on; N is a multiple of
    neither 128 nor 32, so the last outer block and the last panel are
    partial), in deep water."""
    m = small_pair(na=79, nz=35, nr=5) if outer else small_pair(na=40, nz=8, nr=4)
    if outer:
        assert m.nelem >= 6144 and m.nelem % 128 and m.nelem % 32
    depth = -1.0 if outer else 40.0
    say(f"[gpu] LU look-ahead vs one stream n={m.nelem} {precision}")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HAMSX_LU_LOOKAHEAD", mode)
        with FrequencySolver(m, precision=precision) as s:
            s.set_waves([0.0, 1.0])
            s.calwaveproperts(1.2, depth)
            s.assemble()
            assert s.factor() == 0
            lu, ipiv = s.get_matrix("LU")
            out[mode] = (lu, ipiv, s.radiation_solver(), s.diffraction_solver())
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("shape,precision", [((10, 3, 2), "fp64"), ((40, 8, 4), "fp64"), ((40, 8, 4), "mixed"),
                                              ((40, 24, 4), "fp64"), ((40, 24, 4), "mixed")])
def test_triangular_solves_match_lapack(shape, precision):
    """The solves (k_tri_inverse + k_trsm_super: 256-row blocks, 32-row
    diagonal blocks by their inverses) against LAPACK (scipy lu_factor /
    lu_solve) on the GPU's own A and right-hand sides from the GPU's S: N =
    100 (one partial block), 960 (three full blocks and a partial one) and
    2,240 (eight full, one partial), all 15 right-hand-side columns (two
    groups, one partial)."""
    m = small_pair(*shape)
    heads = np.deg2rad([0.0, 30.0, 90.0])
    say(f"[gpu] triangular solves n={m.nelem} {precision}")
    with FrequencySolver(m, precision=precision) as s:
        s.set_waves(heads)
        s.calwaveproperts(0.9, 35.0)
        s.assemble()
        A, S = s.get_matrix("A"), s.get_matrix("S")
        assert s.factor() == 0
        X = np.concatenate([s.radiation_solver(), s.diffraction_solver()], axis=1)
    Xo = oracle.solve_assembled(m, A, S, 0.9, 35.0, heads)[3]
    err = rel(X, Xo)
    say(f"    max |X - X_lapack| / max |X_lapack| = {err:.2e}")
    assert err < (1e-12 if precision == "fp64" else 1e-9)


def test_fp32_far_field_changes_only_the_eigen_region():
    """Mixed precision sums the EIGEN region's evanescent modes (R/h >= 0.5)
    in FP32: every other entry of A and S equals the FP64 assembly bit for
    bit, the EIGEN entries stay within 1e-6 of the largest EIGEN entry, and
    the coefficients within 1e-6 of the FP64 solve's (the mixed path's bar).
    Two cylinders 60 m apart at depth 40 m: most pairs are EIGEN pairs."""
    m = small_pair(na=24, nz=6, nr=3, gap=60.0)
    h, w = 40.0, 1.1
    heads = np.deg2rad([0.0, 45.0])
    rr = np.hypot(m.cen[:, None, 0] - m.cen[None, :, 0], m.cen[:, None, 1] - m.cen[None, :, 1])
    eig = np.abs(rr / h) >= 0.5
    say(f"[gpu] FP32 far field n={m.nelem}: {eig.mean():.2f} of the pairs in EIGEN")
    out = {}
    for prec in ("fp64", "mixed"):
        with FrequencySolver(m, precision=prec) as s:
            assert s.options()["fp32_farfield"] == 1
            s.set_waves(heads)
            s.calwaveproperts(w, h)
            s.assemble()
            out[prec] = (s.get_matrix("A"), s.get_matrix("S"), s.solve_frequency(w, h)[:3])
    for k in (0, 1):
        a, b = out["mixed"][k], out["fp64"][k]
        np.testing.assert_array_equal(a[~eig], b[~eig])
        d = np.max(np.abs(a[eig] - b[eig])) / np.max(np.abs(b[eig]))
        say(f"    {'AS'[k]}: EIGEN entries within {d:.1e} of the FP64 assembly's largest")
        assert 0 < d < 1e-6
    for name, a, b in zip(("amas", "bdmp", "exfc"), out["mixed"][2], out["fp64"][2]):
        d = np.max(np.abs(a - b)) / np.max(np.abs(b))
        say(f"    {name}: {d:.1e}")
        assert d < 1e-6


@pytest.mark.parametrize("depth", [-1.0, 30.0])
def test_frequency_solve_matches_oracle(mesh_small, depth):
    m = mesh_small
    heads = np.deg2rad([0.0, 45.0, 90.0])
    for w in (0.5, 1.3):
        say(f"[gpu] solve_frequency n={m.nelem} depth={depth} w={w}")
        with FrequencySolver(m) as s:
            s.set_waves(heads)
            am, bd, ex, info = s.solve_frequency(w, depth)
        ref = oracle.solve_frequency(m, w, depth, heads, nthreads=16)
        assert info == 0 and ref[3] == 0
        dr = gates.solve_draws(oracle.solve_frequency, m, w, depth, heads, nthreads=16) if depth > 0 else []
        gates.coef_gates((am, bd, ex), ref, dr, 1e-8, say)


def test_step_sequence_equals_fused_call(mesh_small):
    """The reference-named steps and the fused call give the same numbers."""
    m = mesh_small
    heads = np.deg2rad([0.0, 90.0])
    say(f"[gpu] step sequence n={m.nelem}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        a1, b1, e1, _ = s.solve_frequency(0.9, 35.0)
        s.calwaveproperts(0.9, 35.0)
        s.assemble()
        assert s.factor() == 0
        rad = s.radiation_solver()
        dif = s.diffraction_solver()
        a2, b2, e2, _ = s.forces()
    assert rad.shape == (m.nelem, 12) and dif.shape == (m.nelem, 2)
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(b1, b2)
    np.testing.assert_array_equal(e1, e2)


@pytest.mark.parametrize("depth", [-1.0, 45.0])
def test_mixed_precision_within_1e6(mesh_small, depth):
    m = mesh_small
    heads = np.deg2rad([0.0, 60.0])
    w = 1.1
    say(f"[gpu] mixed n={m.nelem} depth={depth}")
    with FrequencySolver(m, precision="mixed") as s:
        s.set_waves(heads)
        am, bd, ex, info = s.solve_frequency(w, depth)
        it, res = s.refine_stats()
    ref = oracle.solve_frequency(m, w, depth, heads, nthreads=16)
    dr = gates.solve_draws(oracle.solve_frequency, m, w, depth, heads, nthreads=16) if depth > 0 else []
    say(f"    mixed: refine iters {it} resid {res:.2e}")
    gates.coef_gates((am, bd, ex), ref, dr, 1e-6, say)


def test_mixed_refinement_stops_on_device(mesh_small):
    """Mixed precision: the device-side residual test stops the correction
    steps once max |R|/|X| <= tol; a fixed four-step run and a run allowed no
    correction bracket it (the latter reports HAMSX_ENOCONV)."""
    from hamsx._lib import HAMSX_ENOCONV, HamsxError
    m = mesh_small
    heads = np.deg2rad([0.0, 60.0])
    w, depth = 1.1, -1.0
    say(f"[gpu] mixed refinement n={m.nelem}")
    with FrequencySolver(m, precision="mixed") as s:
        s.set_waves(heads)
        auto = s.solve_frequency(w, depth)
        it, res = s.refine_stats()
        s.set_refinement(4, 0.0)
        full = s.solve_frequency(w, depth)
        it4, res4 = s.refine_stats()
        s.set_refinement(0, 1e-10)
        with pytest.raises(HamsxError) as e:
            s.solve_frequency(w, depth)
        it0, res0 = s.refine_stats()
    ao, bo, eo, _ = oracle.solve_frequency(m, w, depth, heads)
    say(f"    tol 1e-10: {it} steps, residual {res:.2e}; 4 fixed steps: residual {res4:.2e}; "
        f"no step: residual {res0:.2e}")
    assert e.value.code == HAMSX_ENOCONV and it0 == 0 and res0 > 1e-10
    assert 1 <= it <= 4 and res <= 1e-10 and it4 == 4 and res4 <= 1e-12
    for got in (auto, full):
        for g, o in zip(got[:3], (ao, bo, eo)):
            assert rel(g, o) <= 1e-8


def test_single_cylinder_config():
    case = configs.single_cylinder(na=24, nz=6, nr=3)
    m = combine(case.bodies, case.xr)
    heads = np.deg2rad(case.headings_deg)
    say(f"[gpu] single cylinder n={m.nelem}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        am, bd, ex, info = s.solve_frequency(float(case.omegas[0]), case.depth)
    ao, bo, eo, _ = oracle.solve_frequency(m, float(case.omegas[0]), case.depth, heads)
    assert max(rel(am, ao), rel(bd, bo), rel(ex, eo)) < 1e-8


def test_submitted_solves_in_flight_equal_blocking_solves(mesh_small):
    """Solves kept in flight on three contexts (three streams) give the same
    numbers bit for bit as the blocking call on one context, and a pending
    context refuses other work."""
    from hamsx._lib import HamsxError
    m = mesh_small
    heads = np.deg2rad([0.0, 30.0])
    cases = [(0.6, 40.0), (1.4, 40.0), (0.9, -1.0), (2.0, 25.0)]
    say(f"[gpu] submit/collect n={m.nelem}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        ref = [s.solve_frequency(w, h) for w, h in cases]
    ctxs = [FrequencySolver(m) for _ in range(3)]
    try:
        for c in ctxs:
            c.set_waves(heads)
        got = [None] * len(cases)
        ctxs[0].submit(*cases[0])
        ctxs[1].submit(*cases[1])
        ctxs[2].submit(*cases[2])
        with pytest.raises(HamsxError):
            ctxs[0].assemble()
        got[0] = ctxs[0].collect()
        ctxs[0].submit(*cases[3])
        got[1] = ctxs[1].collect()
        got[2] = ctxs[2].collect()
        got[3] = ctxs[0].collect()
    finally:
        for c in ctxs:
            c.close()
    for r, g in zip(ref, got):
        for a, b in zip(r, g):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision,outer,depth", [("fp64", 6144, 40.0), ("mixed", 6144, 40.0),
                                                   ("fp64", 256, -1.0), ("mixed", 256, 40.0)])
def test_frequency_group_equals_single_solves(mesh_small, precision, outer, depth):
    """A frequency group (one batched factorisation for several contexts'
    matrices) gives each context the coefficients and INFO of its own solve,
    bit for bit, in the per-panel and the outer-block LU (outer blocks from
    order 256 on); groups of one and of three, twice on the same contexts in
    another order.  Contexts of another precision are refused."""
    from hamsx._lib import HamsxError
    m = mesh_small
    heads = np.deg2rad([0.0, 45.0])
    ws = [0.5, 1.1, 1.7]
    say(f"[gpu] frequency group n={m.nelem} {precision} outer>={outer} depth={depth}")
    with FrequencySolver(m, precision=precision) as s:
        s.set_options(lu_outer_min_n=outer)
        s.set_waves(heads)
        ref = {w: s.solve_frequency(w, depth) for w in ws}
    ctxs = [FrequencySolver(m, precision=precision) for _ in ws]
    other = FrequencySolver(m, precision="fp64" if precision == "mixed" else "mixed")
    try:
        for c in ctxs + [other]:
            c.set_options(lu_outer_min_n=outer)
            c.set_waves(heads)
        FrequencySolver.submit_group(ctxs, ws, depth)
        got = {w: c.collect() for w, c in zip(ws, ctxs)}
        FrequencySolver.submit_group(ctxs[::-1], ws, depth)
        got2 = {w: c.collect() for w, c in zip(ws, ctxs[::-1])}
        FrequencySolver.submit_group(ctxs[1:2], ws[:1], depth)
        got1 = ctxs[1].collect()
        with pytest.raises(HamsxError):
            FrequencySolver.submit_group([ctxs[0], other], ws[:2], depth)
        with pytest.raises(HamsxError):
            FrequencySolver.submit_group([ctxs[0], ctxs[0]], ws[:2], depth)
    finally:
        for c in ctxs + [other]:
            c.close()
    for w in ws:
        assert ref[w][3] == 0
        for a, b, c in zip(ref[w], got[w], got2[w]):
            np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(a, c)
    for a, b in zip(ref[ws[0]], got1):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_frequency_group_with_mixed_heading_counts(mesh_small, precision):
    """Contexts of one group with different heading counts (so different
    right-hand-side counts) share the batched factorisation but solve on
    their own streams instead of in the batched solve; each still gets its
    own solve's coefficients bit for bit."""
    m = mesh_small
    heads = [np.deg2rad([0.0, 45.0]), np.deg2rad([0.0, 30.0, 90.0])]
    ws = [0.6, 1.3]
    say(f"[gpu] frequency group, heading counts 2 and 3, n={m.nelem} {precision}")
    ref = []
    for w, hd in zip(ws, heads):
        with FrequencySolver(m, precision=precision) as s:
            s.set_options(lu_outer_min_n=256)
            s.set_waves(hd)
            ref.append(s.solve_frequency(w, 40.0))
    ctxs = [FrequencySolver(m, precision=precision) for _ in ws]
    try:
        for c, hd in zip(ctxs, heads):
            c.set_options(lu_outer_min_n=256)
            c.set_waves(hd)
        FrequencySolver.submit_group(ctxs, ws, 40.0)
        got = [c.collect() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    for r, g in zip(ref, got):
        assert r[3] == 0
        for a, b in zip(r, g):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_frequency_group_default_outer_blocks(precision):
    """With the LU options left alone a group of two or more matrices factors
    in outer blocks from order 1024 on (single solves from 6144 on): at
    N = 1,728 the group's coefficients equal the single solves' to the LU's
    rounding (1e-10 of each table's largest entry)."""
    m = small_pair(na=48, nz=12, nr=6)
    assert m.nelem >= 1024
    heads = np.deg2rad([0.0, 90.0])
    ws = [0.6, 1.3]
    say(f"[gpu] default group vs single n={m.nelem} {precision}")
    with FrequencySolver(m, precision=precision) as s:
        s.set_waves(heads)
        ref = [s.solve_frequency(w, 40.0) for w in ws]
    ctxs = [FrequencySolver(m, precision=precision) for _ in ws]
    try:
        for c in ctxs:
            c.set_waves(heads)
        FrequencySolver.submit_group(ctxs, ws, 40.0)
        got = [c.collect() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    for r, g in zip(ref, got):
        assert r[3] == 0 and g[3] == 0
        errs = [rel(a, b) for a, b in zip(g[:3], r[:3])]
        say(f"    group vs single: {errs}")
        assert max(errs) < 1e-10


@pytest.mark.parametrize("depth", [-1.0, 35.0])
def test_gmres_matches_lu(depth):
    """Restarted GMRES with the block-diagonal preconditioner converges to the
    LU solution: at a tight tolerance to 1e-8 on every coefficient, at the
    reference's own tolerance (1e-4) as loosely as the reference's GMRES
    integration test allows (test/test_integration_gmres.f90:69,97)."""
    m = small_pair(na=28, nz=7, nr=3)
    heads = np.deg2rad([0.0, 45.0])
    say(f"[gpu] gmres vs lu n={m.nelem} depth={depth}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_hmatrix(False)  # dense operator: GMRES alone against the LU
        lu = s.solve_frequency(1.1, depth)
        s.set_solver("gmres", tol=1e-11, restart=25, maxiter=400)
        tight = s.solve_frequency(1.1, depth)
        it, res = s.refine_stats()
        s.set_solver("gmres")
        loose = s.solve_frequency(1.1, depth)
        it2, res2 = s.refine_stats()
    say(f"    tight: {it} iterations, residual {res:.2e}; reference tolerance: {it2} iterations, residual {res2:.2e}")
    assert res < 1e-11 and res2 < 1e-4
    for a, b, c in zip(tight[:3], lu[:3], loose[:3]):
        assert rel(a, b) < 1e-8
        assert rel(c, b) < 1e-3


@pytest.mark.parametrize("depth,tol", [(-1.0, 1e-6), (40.0, 1e-4)])
def test_hmatrix_matvec_matches_oracle(depth, tol):
    """The device H-matrix (ACA of the assembled matrix on the GPU) against
    the oracle's restatement built from the same matrix, and against the
    dense product."""
    import hmatrix as ohm
    m = small_pair()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((m.nelem, 3)) + 1j * rng.standard_normal((m.nelem, 3))
    say(f"[gpu] hmatrix n={m.nelem} depth={depth} tol={tol}")
    with FrequencySolver(m) as s:
        s.calwaveproperts(1.2, depth)
        s.assemble()
        A = s.get_matrix("A")
        s.set_hmatrix(True, 2.0, tol, 32, 50)
        y = s.hmatrix_matvec(x)
        nlr, nd, mr, ratio = s.hmatrix_stats()
    H = ohm.build(A, m.cen, 32, 2.0, tol, 50)
    yo = ohm.matvec(H, x)
    lr = [b for b in H[1] if b[4] == "lr"]
    omr = float(np.mean([b[5][0].shape[1] for b in lr])) if lr else 0.0
    ed, eo = rel(y, A @ x), rel(y, yo)
    say(f"    {nlr} low-rank (oracle {len(lr)}), {nd} dense, mean rank {mr:.2f} (oracle {omr:.2f}), "
        f"compression {ratio:.2f}; vs dense {ed:.2e}, vs oracle H {eo:.2e}")
    assert nlr == len(lr) and nlr > 0 and abs(mr - omr) <= 0.25
    assert ed <= 50 * tol and eo <= 10 * tol


@pytest.mark.parametrize("depth", [-1.0, 35.0])
def test_gmres_hmatrix_matches_lu(depth):
    """GMRES on the H-matrix operator: with a tight ACA tolerance it reaches
    the LU coefficients to 1e-6; with the reference's parameters (ACA 1e-4,
    GMRES 1e-4) it stays within the reference's own GMRES test bounds (added
    mass 1e-3, damping 5e-2: test/test_integration_gmres.f90:69,97)."""
    m = small_pair(na=28, nz=7, nr=3)
    heads = np.deg2rad([0.0, 45.0])
    say(f"[gpu] gmres + hmatrix vs lu n={m.nelem} depth={depth}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        lu = s.solve_frequency(1.1, depth)
        s.set_solver("gmres", tol=1e-10, restart=25, maxiter=400)
        s.set_hmatrix(True, 2.0, 1e-9, 32, 50)
        tight = s.solve_frequency(1.1, depth)
        st = s.hmatrix_stats()
        s.set_solver("gmres")
        s.set_hmatrix()
        ref = s.solve_frequency(1.1, depth)
        it, res = s.refine_stats()
        st2 = s.hmatrix_stats()
    et = [rel(a, b) for a, b in zip(tight[:3], lu[:3])]
    er = [rel(a, b) for a, b in zip(ref[:3], lu[:3])]
    say(f"    tight ACA: {st[0]} low-rank blocks, rank {st[2]:.1f}, err {et}; reference parameters: "
        f"{st2[0]} low-rank, rank {st2[2]:.1f}, {it} iterations, residual {res:.1e}, err {er}")
    assert st[0] > 0 and st2[0] > 0
    assert max(et) < 1e-6
    assert er[0] < 1e-3 and er[1] < 5e-2 and er[2] < 1e-3


@pytest.mark.parametrize("depth", [-1.0, 40.0])
def test_gmres_complex64_hmatrix_blocks(depth):
    """Mixed precision GMRES: the Arnoldi products read complex64 copies of
    the H-matrix blocks, while the Krylov vectors and every restart residual
    stay FP64 on the FP64 blocks.  It must reach the FP64 GMRES result on the
    same H-matrix (ACA 1e-9, GMRES 1e-10) to 1e-6 of each table's largest
    entry, and the dense FP64 LU to 1e-6."""
    m = small_pair(na=28, nz=7, nr=3)
    heads = np.deg2rad([0.0, 45.0])
    say(f"[gpu] gmres with complex64 H-matrix blocks n={m.nelem} depth={depth}")
    out = {}
    for prec in ("fp64", "mixed"):
        with FrequencySolver(m, precision=prec) as s:
            s.set_waves(heads)
            if prec == "fp64":
                out["lu"] = s.solve_frequency(1.1, depth)
            s.set_solver("gmres", tol=1e-10, restart=25, maxiter=400)
            s.set_hmatrix(True, 2.0, 1e-9, 32, 50)
            out[prec] = s.solve_frequency(1.1, depth)
            out[prec + "_it"] = s.refine_stats()
    e64 = [rel(a, b) for a, b in zip(out["mixed"][:3], out["fp64"][:3])]
    elu = [rel(a, b) for a, b in zip(out["mixed"][:3], out["lu"][:3])]
    say(f"    iterations fp64 {out['fp64_it']}, mixed {out['mixed_it']}; mixed vs FP64 GMRES {e64}, vs LU {elu}")
    assert max(e64) < 1e-6 and max(elu) < 1e-6
    assert out["mixed_it"][1] < 1e-9  # the final residual is the FP64 one


def lid_centroids(lid):
    from hamsx.mesh import panel_geometry
    cen, _, _, _ = panel_geometry(lid, np.zeros(3))
    return cen


@pytest.mark.parametrize("depth,w", [(-1.0, 1.6), (30.0, 1.2)])
def test_irregular_frequency_removal_matches_oracle(depth, w):
    """Hull + waterplane lid (IRSP = 1): the GPU normal-equation solve against
    the oracle's restatement of the reference's IRR path."""
    from hamsx.mesh import waterplane_disk
    hull = cylinder(4.0, 6.0, 24, 6, 3)
    m = combine([hull], [np.zeros(3)])
    lid = lid_centroids(waterplane_disk(4.0, 24, 3))
    heads = np.deg2rad([0.0, 60.0])
    say(f"[gpu] irr n={m.nelem} lid={len(lid)} depth={depth} w={w}")
    with FrequencySolver(m) as s:
        s.set_lid(lid)
        s.set_waves(heads)
        am, bd, ex, info = s.solve_frequency(w, depth)
        s.set_lid(None)
        am0, _, _, _ = s.solve_frequency(w, depth)
    ref = oracle.solve_frequency_irr(m, lid, w, depth, heads, nthreads=16)
    assert info == 0 and ref[3] == 0
    dr = gates.solve_draws(oracle.solve_frequency_irr, m, lid, w, depth, heads, nthreads=16) if depth > 0 else []
    res = gates.coef_gates((am, bd, ex), ref, dr, 1e-8, say)
    say(f"    without lid {rel(am0, ref[0]):.2e}")
    assert rel(am0, ref[0]) > 100 * max(res[0][0], 1e-12)   # the lid changes the answer


def test_irregular_frequency_example5_on_gpu():
    """The reference's example 5 (234 hull + 383 lid panels, h = 70 m) at two
    of its frequencies: GPU against the oracle."""
    import os
    from conftest import GOLDEN
    from hamsx.deck import PanelMesh
    z = np.load(os.path.join(GOLDEN, "example5.npz"))
    hull = PanelMesh(z["hull_xyz"], z["hull_ncn"], z["hull_ncon"])
    lid = lid_centroids(PanelMesh(z["lid_xyz"], z["lid_ncn"], z["lid_ncon"]))
    m = combine([hull], [z["xr"]])
    heads = np.deg2rad([0.0, 90.0])
    say(f"[gpu] example5 irr n={m.nelem} lid={len(lid)}")
    with FrequencySolver(m) as s:
        s.set_lid(lid)
        s.set_waves(heads)
        for w in (float(z["omegas"][0]), float(z["omegas"][-1])):
            am, bd, ex, info = s.solve_frequency(w, float(z["depth"]))
            ref = oracle.solve_frequency_irr(m, lid, w, float(z["depth"]), heads, nthreads=16)
            dr = gates.solve_draws(oracle.solve_frequency_irr, m, lid, w, float(z["depth"]), heads, nthreads=16)
            say(f"    w={w:.3f}")
            gates.coef_gates((am, bd, ex), ref, dr, 1e-8, say)
