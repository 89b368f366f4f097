"""Every benchmark workload at its stated size on the GPU, against the oracle
(or, for the serving and GMRES workloads, against the device's own plain
FP64 LU solve), with the gates of tests/gates.py.

  1. one cylinder, 864 panels, infinite depth, w = 1.0 rad/s, heading 0;
  2. two cylinders, 3,968 panels, depth 50 m, two of the 30 frequencies,
     headings 0/45/90;
  3. five cylinders, 10,026 panels, depth 60 m, one frequency: mixed
     precision against the GPU's FP64 solve and against the oracle;
  4. a seeded stream of six serving requests from 504 to 12,120 panels
     (deep and finite depth, 1-9 bodies): the same alone and in the mix;
  5. a 2 x 2 lidded cylinder array (2,448 hull + 264 lid panels), depth
     80 m: GMRES on the H-matrix (eta 2, ACA 1e-6) against the dense LU.

The oracle runs on the host's threads (16); the largest case takes a few
minutes of host time.  Bars: 1e-8 (FP64) and 1e-6 (mixed, GMRES) of each
table's largest entry, or the reference's own rounding spread where larger
(finite depth: tests/gates.py).  Per-coefficient relative errors (each
coefficient against its own magnitude, those above 1e-3 of the table's
largest) are printed beside them.
"""
import sys
import time

import numpy as np
import pytest

import gates
import oracle
from hamsx import configs
from hamsx.mesh import combine, panel_geometry
from hamsx.serving import Request, Server, poisson_arrivals
from hamsx.solver import FrequencySolver

pytestmark = pytest.mark.gpu
NT = 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def own_rel(c, co):
    """Largest |c - co| / |co| over the coefficients with |co| >= 1e-3 max |co|."""
    c, co = np.asarray(c), np.asarray(co)
    big = np.abs(co) >= 1e-3 * np.max(np.abs(co))
    return float(np.max(np.abs(c - co)[big] / np.abs(co)[big]))


def report_own(tag, got, ref):
    say(f"    {tag} per-coefficient relative (own magnitude): added mass {own_rel(got[0], ref[0]):.2e} "
        f"damping {own_rel(got[1], ref[1]):.2e} excitation {own_rel(got[2], ref[2]):.2e}")


def gpu_solve(m, w, h, heads, precision="fp64", matrices=False):
    with FrequencySolver(m, precision=precision) as s:
        s.set_waves(heads)
        t0 = time.perf_counter()
        out = s.solve_frequency(w, h)
        t = time.perf_counter() - t0
        assert out[3] == 0
        if matrices:
            s.calwaveproperts(w, h)
            s.assemble()
            out = out + (s.get_matrix("A"), s.get_matrix("S"))
    say(f"    GPU {precision} solve {t * 1e3:.0f} ms")
    return out


def test_config1_single_cylinder_864_deep():
    case = configs.single_cylinder()
    m = combine(case.bodies, case.xr)
    assert m.nelem == 864
    w, h, heads = float(case.omegas[0]), case.depth, np.deg2rad(case.headings_deg)
    say(f"[gpu] config 1: {case.name} N={m.nelem} w={w}")
    fp = gpu_solve(m, w, h, heads, matrices=True)
    mx = gpu_solve(m, w, h, heads, "mixed")
    Ao, So = oracle.assemble(m, w, h, NT)
    say(f"    A {gates.entry_gate('A', fp[4], Ao, [])}  S {gates.entry_gate('S', fp[5], So, [])}")
    ref = oracle.solve_assembled(m, Ao, So, w, h, heads)
    gates.coef_gates(fp, ref, [], 1e-8, say, "fp64 ")
    report_own("fp64", fp, ref)
    gates.coef_gates(mx, ref, [], 1e-6, say, "mixed ")
    report_own("mixed", mx, ref)


@pytest.mark.parametrize("k", [0, 15])
def test_config2_two_cylinders_3968_h50(k):
    case = configs.two_cylinders()
    m = combine(case.bodies, case.xr)
    assert m.nelem == 3968
    w, h, heads = float(case.omegas[k]), case.depth, np.deg2rad(case.headings_deg)
    say(f"[gpu] config 2: {case.name} N={m.nelem} depth={h} w={w:.3f} (frequency {k + 1} of 30)")
    fp = gpu_solve(m, w, h, heads, matrices=True)
    mx = gpu_solve(m, w, h, heads, "mixed")
    t0 = time.perf_counter()
    Ao, So = oracle.assemble(m, w, h, NT)
    dr = gates.assembly_draws(m, w, h, NT)
    say(f"    oracle: 4 assemblies in {time.perf_counter() - t0:.1f} s on {NT} threads")
    cls, nt = gates.regions(m.cen, m.cen, h, with_nt=True)
    say(f"    A {gates.entry_gate('A', fp[4], Ao, [d[0] for d in dr], cls, nt)}")
    say(f"    S {gates.entry_gate('S', fp[5], So, [d[1] for d in dr], cls, nt)}")
    got = fp[:4]  # keep the coefficients only
    del fp
    ref = oracle.solve_assembled(m, Ao, So, w, h, heads)
    rdr = [oracle.solve_assembled(mm, a, s, w, h, heads)
           for mm, (a, s) in zip((m, oracle.perturbed(m, 1), oracle.perturbed(m, 2)), dr)]
    gates.coef_gates(got, ref, rdr, 1e-8, say, "fp64 ")
    report_own("fp64", got, ref)
    gates.coef_gates(mx, ref, rdr, 1e-6, say, "mixed ")
    report_own("mixed", mx, ref)


def test_config3_five_cylinders_10026_mixed():
    case = configs.five_cylinders()
    m = combine(case.bodies, case.xr)
    assert m.nelem == 10026
    w, h, heads = float(case.omegas[31]), case.depth, np.deg2rad(case.headings_deg)
    say(f"[gpu] config 3: {case.name} N={m.nelem} depth={h} w={w:.3f}")
    fp = gpu_solve(m, w, h, heads, matrices=True)
    mx = gpu_solve(m, w, h, heads, "mixed")
    # mixed against the GPU's own FP64 solve: the same assembly, so only the
    # complex64 LU + FP64 refinement separates them
    gates.coef_gates(mx, fp, [], 1e-6, say, "mixed vs GPU fp64 ")
    report_own("mixed vs GPU fp64", mx, fp)
    rows = np.unique(np.linspace(0, m.nelem - 1, 384).astype(int))
    cls, nt = gates.regions(m.cen[rows], m.cen, h, with_nt=True)
    A_rows, S_rows = fp[4][rows], fp[5][rows]
    fp = fp[:4]
    t0 = time.perf_counter()
    variants = (m, oracle.perturbed(m, 2))  # the oracle, and a draw that moves R
    refs, ra, rs = [], [], []
    for q, mm in enumerate(variants):
        for fma in ((False, True) if q == 0 else (False,)):
            A, S = oracle.assemble(mm, w, h, NT, fma=fma)
            refs.append(oracle.solve_assembled(mm, A, S, w, h, heads))
            ra.append(A[rows])
            rs.append(S[rows])
            del A, S
    say(f"    oracle: 3 assemblies + LAPACK solves in {time.perf_counter() - t0:.1f} s on {NT} threads")
    say(f"    A rows {gates.entry_gate('A', A_rows, ra[0], ra[1:], cls, nt)}")
    say(f"    S rows {gates.entry_gate('S', S_rows, rs[0], rs[1:], cls, nt)}")
    ref, rdr = refs[0], refs[1:]
    gates.coef_gates(fp, ref, rdr, 1e-8, say, "fp64 ")
    report_own("fp64", fp, ref)
    gates.coef_gates(mx, ref, rdr, 1e-6, say, "mixed ")
    report_own("mixed", mx, ref)


def test_config4_serving_stream_504_to_12120_panels():
    rng = np.random.default_rng(11)
    reqs, sizes = [], []
    arr = poisson_arrivals(6, 8.0, np.random.default_rng(12))
    for k, total in enumerate((500, 2500, 5000, 7500, 10000, 12000)):
        case = configs.random_request(rng, total=total)
        mesh = combine(case.bodies, case.xr)
        sizes.append(mesh.nelem)
        reqs.append(Request(k, mesh, float(case.omegas[0]), case.depth, np.deg2rad(case.headings_deg),
                            arrival=float(arr[k])))
    say(f"[gpu] config 4: 6 seeded requests, {sizes} panels, depths {[round(r.depth, 1) for r in reqs]}")
    with Server(slots=4, precision="mixed") as srv:
        res = srv.run(reqs)
    assert sorted(r.rid for r in res) == list(range(6))
    overlap = sum(1 for x in res for y in res if x.rid < y.rid and x.admitted < y.done and y.admitted < x.done)
    lat = np.array([r.latency for r in res]) * 1e3
    say(f"    overlapping pairs {overlap}; latency p50 {np.percentile(lat, 50):.0f} ms max {lat.max():.0f} ms")
    assert overlap > 0
    for r in sorted(res, key=lambda x: x.rid):
        q = reqs[r.rid]
        assert r.info == 0
        alone = gpu_solve(q.mesh, q.w1, q.depth, q.headings_rad, "mixed")
        fp = gpu_solve(q.mesh, q.w1, q.depth, q.headings_rad, "fp64")
        got = (r.amas, r.bdmp, r.exfc)
        same = all(np.array_equal(g, a) for g, a in zip(got, alone[:3]))
        diff = max(float(np.max(np.abs(g - a)) / np.max(np.abs(a))) for g, a in zip(got, alone[:3]))
        say(f"    request {r.rid}: N={q.mesh.nelem} in the mix vs alone: {'bitwise' if same else f'{diff:.1e}'}")
        assert diff <= 1e-12
        gates.coef_gates(got, fp, [], 1e-6, say, f"request {r.rid} mixed vs fp64 ")


def test_config5_lidded_array_gmres_hmatrix():
    """GMRES (tolerance 1e-9) on the H-matrix of the normal matrix C = A^T A
    against the dense FP64 LU of C, both on the GPU, at the workload's ACA
    tolerance 1e-6 (rank <= 50) and at 1e-8 (rank <= 64).  Added mass and
    excitation are held to 5e-6 (ACA 1e-6; measured up to 1.9e-6) and 1e-6
    (ACA 1e-8; measured 2e-7) of their tables' largest entries.  Damping is
    held to the error the added mass implies for it, that bound x
    max(1, w max|AMAS| / max|BDMP|): at low frequency the damping is a small part of the
    potentials' effect, so the same potential error is a larger fraction of
    it -- why the reference's own GMRES test allows 5e-2 on damping against
    1e-3 on added mass (test/test_integration_gmres.f90:69,97)."""
    case = configs.large_array(600, grid=2)
    m = combine(case.bodies, case.xr)
    lid = np.concatenate([panel_geometry(lm, np.zeros(3))[0] for lm in case.meta["lids"]])
    heads = np.deg2rad(case.headings_deg)
    say(f"[gpu] config 5 (2 x 2): {case.name} N={m.nelem} lid={len(lid)} depth={case.depth}")

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    checks = []
    for w in (0.4, 1.4):
        with FrequencySolver(m) as s:
            s.set_lid(lid)
            s.set_waves(heads)
            lu = s.solve_frequency(w, case.depth)
            assert lu[3] == 0
            for aca, rank in ((1e-6, 50), (1e-8, 64)):
                s.set_solver("gmres", tol=1e-9, restart=25, maxiter=400)
                s.set_hmatrix(True, 2.0, aca, 32, rank)
                t0 = time.perf_counter()
                gm = s.solve_frequency(w, case.depth)
                t = time.perf_counter() - t0
                it, res = s.refine_stats()
                nlr, nd, mr, ratio = s.hmatrix_stats()
                errs = [rel(a, b) for a, b in zip(gm[:3], lu[:3])]
                bound = 5e-6 if aca >= 1e-6 else 1e-6
                bd_bound = bound * max(1.0, w * np.max(np.abs(lu[0])) / np.max(np.abs(lu[1])))
                say(f"    w={w} ACA {aca:.0e} rank <= {rank}: GMRES {it} iterations, residual {res:.1e}, "
                    f"{t * 1e3:.0f} ms; H-matrix {nlr} low-rank (mean rank {mr:.1f}), {nd} dense, compression "
                    f"{ratio:.1f}; vs LU: added mass {errs[0]:.2e} damping {errs[1]:.2e} (bound {bd_bound:.1e}) "
                    f"excitation {errs[2]:.2e}")
                report_own(f"ACA {aca:.0e} gmres+H vs LU", gm, lu)
                checks.append((gm[3] == 0 and nlr > 0, errs, bound, bd_bound))
    for ok, errs, bound, bd_bound in checks:
        assert ok
        assert errs[0] <= bound and errs[2] <= bound and errs[1] <= bd_bound


def test_config5_mixed_normal_matrix_in_complex64(monkeypatch):
    """Mixed precision on the lidded array: C = A^T A formed in complex64
    (csyrk on FP32 MFMA), its H-matrix blocks complex64 in the Arnoldi
    products, and every residual GMRES tests taken from A itself in FP64.
    Against the dense FP64 LU of the FP64 C, at the workload's ACA tolerance
    1e-6, with test_config5's bounds (added mass and excitation 5e-6 of
    their tables' largest entries, damping the bound the added mass implies
    for it); and against the same path with C formed in FP64
    (HAMSX_NORMAL_FP64=1)."""
    case = configs.large_array(600, grid=2)
    m = combine(case.bodies, case.xr)
    lid = np.concatenate([panel_geometry(lm, np.zeros(3))[0] for lm in case.meta["lids"]])
    heads = np.deg2rad(case.headings_deg)
    say(f"[gpu] config 5 mixed, C in complex64: {case.name} N={m.nelem} lid={len(lid)} depth={case.depth}")

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    for w in (0.4, 1.4):
        with FrequencySolver(m) as s:
            s.set_lid(lid)
            s.set_waves(heads)
            lu = s.solve_frequency(w, case.depth)
        assert lu[3] == 0
        got = {}
        for mode in ("0", "1"):
            monkeypatch.setenv("HAMSX_NORMAL_FP64", mode)
            with FrequencySolver(m, precision="mixed") as s:
                s.set_lid(lid)
                s.set_waves(heads)
                s.set_solver("gmres", tol=1e-9, restart=25, maxiter=400)
                s.set_hmatrix(True, 2.0, 1e-6, 32, 50)
                t0 = time.perf_counter()
                got[mode] = s.solve_frequency(w, case.depth)
                t = time.perf_counter() - t0
                it, res = s.refine_stats()
            errs = [rel(a, b) for a, b in zip(got[mode][:3], lu[:3])]
            say(f"    w={w} C in {'FP64' if mode == '1' else 'complex64'}: GMRES {it} iterations, residual {res:.1e}, "
                f"{t * 1e3:.0f} ms; vs LU: added mass {errs[0]:.2e} damping {errs[1]:.2e} excitation {errs[2]:.2e}")
            report_own(f"C {'FP64' if mode == '1' else 'complex64'} vs LU", got[mode], lu)
            assert got[mode][3] == 0 and res < 1e-9
            bd_bound = 5e-6 * max(1.0, w * np.max(np.abs(lu[0])) / np.max(np.abs(lu[1])))
            assert errs[0] <= 5e-6 and errs[2] <= 5e-6 and errs[1] <= bd_bound
        d = [rel(a, b) for a, b in zip(got["0"][:3], got["1"][:3])]
        say(f"    complex64 C vs FP64 C: {d}")
