"""Parity gates for results that the reference itself only defines up to its
rounding spread.

The reference's finite-depth Green function (FinGreen3D: PIDCOCK and EIGENE's
Wynn epsilon extrapolation, src/FinGreen3D.f90:413-506, :558-660) moves under
rounding changes alone: its own Fortran built with and without FMA contraction
disagrees by up to 1e-4 per Green value at depth 50 m and by up to 9e-2 at
depth 200 m (PIDCOCK), while the oracle and the reference built the same way
agree to 1e-15 (tests/test_reference_spread.py).  A single second build is
one draw from that spread.  The gates below take three draws -- the oracle
with FMA contraction, the oracle on centroids moved by one unit in the last
place in z, and in x and y (which moves R) -- and hold the result under test
to them:

  * entries of a region where every draw agrees with the oracle to 1e-13 of
    the array's largest entry (deep water, EIGEN, LINTON): every entry within
    1e-13 of it;
  * entries of a region where the draws spread (PIDCOCK, EIGENE): the
    distribution of the distances must match the draws' -- the median, the
    90th and 99th percentiles and the maximum each at most 4x the draws'
    (+1e-12), over the region and (median) over every series length NT
    with at least 256 entries.  A fault that moves a region's or a series
    length's entries systematically shifts a median or a percentile; a
    fault confined to under 1% of a region's entries must stay within 4x
    that region's own largest spread.
  The reference's own rounding does not allow a per-entry bound on these
  regions: entry by entry, a further valid rounding (centroids moved the other
  way, with FMA) lands beyond 4x the largest of the three draws on 5-12% of
  the PIDCOCK and EIGENE entries, and beyond 16x on 2-8%, because it and the
  draws flip different epsilon-table decisions; its percentiles and
  per-NT medians stay within 1.4x the draws' (tests/test_reference_spread.py).

Coefficient tables (added mass, damping, excitation) get the same treatment
on their scale: every coefficient within max(tol, 4x the largest draw
distance) of the table's largest entry, and the median and root mean square
of the distances at most 4x the draws' (+tol / 10).
"""
import numpy as np

LINTON, PIDCOCK, EIGENE, EIGEN = 0, 1, 2, 3
REGION_NAMES = ("LINTON", "PIDCOCK", "EIGENE", "EIGEN")


def regions(fcen, scen, h, with_nt=False):
    """R/h region of every (field row i, source column j) entry, with the
    kernels' expression |sqrt(dx^2 + dy^2) / h| (FinGreen3D.f90:183-191):
    LINTON < 5e-4 <= PIDCOCK < 0.05 <= EIGENE < 0.5 <= EIGEN.  with_nt: also
    the series length NT of PIDCOCK (INT(-1010.10 R/h + 100.50), :566) and
    EIGENE (INT(-88.89 R/h + 54.45), :421) entries (0 elsewhere)."""
    fcen = np.asarray(fcen)
    scen = np.asarray(scen)
    dx = scen[:, 0][None, :] - fcen[:, 0][:, None]
    dy = scen[:, 1][None, :] - fcen[:, 1][:, None]
    rh = np.abs(np.sqrt(dx * dx + dy * dy) / h)
    out = np.full(rh.shape, EIGEN, dtype=np.int8)
    out[rh < 0.5] = EIGENE
    out[rh < 0.05] = PIDCOCK
    out[rh < 5e-4] = LINTON
    if not with_nt:
        return out
    f = lambda x: np.float64(np.float32(x))  # the reference's REAL*4 literals
    nt = np.zeros(rh.shape, dtype=np.int16)
    e, p = out == EIGENE, out == PIDCOCK
    nt[e] = np.trunc(f(-88.89) * rh[e] + f(54.45)).astype(np.int16)
    nt[p] = np.trunc(f(-1010.10) * rh[p] + f(100.50)).astype(np.int16)
    return out, nt


def _draw_distance(Xo, draws, M):
    F = np.zeros(Xo.shape)
    for D in draws:
        F = np.maximum(F, np.abs(D - Xo) / M)
    return F


def entry_gate(name, X, Xo, draws, cls=None, nt=None, stable_tol=1e-13, spread_tol=1e-12, factor=4.0):
    """Assert the gates above for the array X against the oracle's Xo, with
    the rounding draws `draws` (arrays like Xo), region classes `cls` (None:
    one class) and series lengths `nt` (None: no per-length medians).
    Returns a per-region report (dict)."""
    X, Xo = np.asarray(X), np.asarray(Xo)
    M = float(np.max(np.abs(Xo)))
    d = np.abs(X - Xo) / M
    F = _draw_distance(Xo, draws, M)
    if cls is None:
        cls = np.zeros(Xo.shape, dtype=np.int8)
    report, errors = {}, []
    for r in np.unique(cls):
        sel = cls == r
        dr, Fr = d[sel], F[sel]
        key = REGION_NAMES[r] if cls.max() > 0 or r > 0 else "all"
        rep = {"entries": int(sel.sum()), "max": float(dr.max()), "draw_max": float(Fr.max())}
        if Fr.max() <= stable_tol:
            rep["rule"] = "stable"
            if dr.max() > stable_tol:
                errors.append(f"{name} {key}: max {dr.max():.3e} > {stable_tol:.0e} where every rounding draw "
                              f"agrees to {Fr.max():.3e}")
        else:
            rep["rule"] = "spread"
            for q in (50, 90, 99, 100):
                a, b = float(np.percentile(dr, q)), float(np.percentile(Fr, q))
                rep[f"p{q}"] = (a, b)
                if a > factor * b + spread_tol:
                    errors.append(f"{name} {key}: {q}th percentile {a:.3e} > {factor}x the draws' {b:.3e}")
            if nt is not None:
                ntr = nt[sel]
                worst = 0.0
                for k in np.unique(ntr):
                    b = ntr == k
                    if b.sum() < 256:
                        continue
                    a, f = float(np.median(dr[b])), float(np.median(Fr[b]))
                    worst = max(worst, a / (factor * f + spread_tol))
                    if a > factor * f + spread_tol:
                        errors.append(f"{name} {key} NT={k}: median {a:.3e} > {factor}x the draws' {f:.3e}")
                rep["nt_median_ratio"] = worst
        report[key] = rep
    assert not errors, "; ".join(errors)
    return report


def coef_gate(name, c, co, draws, tol, factor=4.0):
    """Assert the coefficient-table gates: every |c - co| <= max(tol,
    factor x the largest draw distance), and the median and the root mean
    square of |c - co| at most factor x the draws' + tol / 10, all relative to
    max |co|.  Returns (max, median, rms, draw max, draw median, draw rms)."""
    c, co = np.asarray(c), np.asarray(co)
    M = float(np.max(np.abs(co)))
    d = np.abs(c - co) / M
    F = _draw_distance(co, draws, M)
    bound = max(tol, factor * float(F.max()))
    med, fmed = float(np.median(d)), float(np.median(F))
    rms, frms = float(np.sqrt(np.mean(d * d))), float(np.sqrt(np.mean(F * F)))
    assert d.max() <= bound, f"{name}: max {d.max():.3e} > {bound:.3e} (draw max {F.max():.3e})"
    assert med <= factor * fmed + tol / 10, f"{name}: median {med:.3e} > {factor}x draw median {fmed:.3e} + {tol / 10:.0e}"
    assert rms <= factor * frms + tol / 10, f"{name}: rms {rms:.3e} > {factor}x draw rms {frms:.3e} + {tol / 10:.0e}"
    return float(d.max()), med, rms, float(F.max()), fmed, frms


def assembly_draws(m, w, h, nthreads=16):
    """The oracle's rounding draws of (A, S) for mesh m at (w, h): FMA
    contraction, centroids moved by one ulp in z, centroids moved by one ulp
    in x and y."""
    import oracle
    return [oracle.assemble(m, w, h, nthreads, fma=True), oracle.assemble(oracle.perturbed(m, 1), w, h, nthreads),
            oracle.assemble(oracle.perturbed(m, 2), w, h, nthreads)]


def solve_draws(solve, m, *args, **kw):
    """The same three draws of a coefficient solve: solve(mesh, *args, fma=,
    **kw) -> (amas, bdmp, exfc, ...)."""
    import oracle
    return [solve(m, *args, fma=True, **kw), solve(oracle.perturbed(m, 1), *args, **kw),
            solve(oracle.perturbed(m, 2), *args, **kw)]


def coef_gates(got, ref, draws, tol, say=print, label=""):
    """coef_gate on the added-mass, damping and excitation tables of a solve."""
    out = []
    for k, name in enumerate(("added mass", "damping", "excitation")):
        r = coef_gate(f"{label}{name}", got[k], ref[k], [d[k] for d in draws], tol)
        say(f"    {name:10s} max {r[0]:.2e} median {r[1]:.2e} rms {r[2]:.2e}   draws max {r[3]:.2e} "
            f"median {r[4]:.2e} rms {r[5]:.2e}")
        out.append(r)
    return out
