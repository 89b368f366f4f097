"""Writes the golden fixtures under tests/golden/ from the reference's own
Fortran routines (oracle/_ref/libhams_ref.so, built by `make -C oracle ref`
from /root/reference/src) and from the reference's shipped test data.

Run in the build container only:  python tests/golden/make_golden.py
Outputs (small .npz files, committed):
  special.npz     ERROR, E1XA, ENXA, EIX, IK01A, JY01B, StruveH0/H1, BesselJ0/J1, Limes
  dispersion.npz  DISPERSION roots and CalWaveProperts outputs
  green.npz       INFGREEN3D / FINGREEN3D over all four finite-depth regions and the
                  V = 0 / V = -1 limits
  newman.npz      SGLINTBD_MULTI_TRI/QUAD and NormalProcess geometry on body 1 of
                  test/data/config1
  incident.npz    DINP / VINP
  config1.npz     test/data/config1: input meshes and deck values, and the 18
                  expected_output tables (OAMASS1-6, ODAMPING1-6, OEXFOR1-6)
  example5.npz    examples/example5_single_body (irregular-frequency removal):
                  hull and waterplane meshes, the deck values and the 18 tables
                  of Output/Hams_format
  newman2.npz     SGLINTBD_MULTI_TRI2/QUAD2 (field-point variants) on body 1 of
                  test/data/config1 at points on, near and away from the panels
  example2.npz    examples/example2_multi_bodies_wave_fields (10 bodies, 100 field
                  points): meshes, deck values, and the shipped Wamit_format
                  PressureElevation.6p / PressureElevationIncidence.6p values
  motion.npz      the body matrices (Hydrostatic_*.in) and the shipped Motion.4
                  table of example 2
  writers.npz     text of shipped output files (config1's HAMS tables, the first
                  lines of example2's WAMIT files) for the writer round trips
  example1.npz    examples/example1_multi_bodies (5 buoys, depth 0.9 m, 80
                  frequencies and the two limits): one hull mesh, the LCS, deck
                  values, body matrices, the shipped Motion.4 table and the
                  OAMASS rows of the zero- and infinite-frequency limits

`python make_golden.py NAME ...` writes only the named fixtures.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "hams-mrel_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from hamsx.deck import read_control, read_hams_block_file, read_hydrostatic, read_pnl  # noqa: E402

REF_DATA = "/root/reference/test/data/config1"
d = C.c_double


def p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def newman2():
    """Field-point Newman integrals on body 1 of config1."""
    r = oracle.ref()
    rng = np.random.default_rng(2468)
    mesh = read_pnl(os.path.join(REF_DATA, "input", "HullMesh_1.pnl"))
    ne, nn = mesh.nelem, mesh.nnode
    xyz = np.asfortranarray(mesh.xyz)
    ncn = mesh.ncn.astype(np.int32)
    ncon = np.asfortranarray(mesh.ncon.astype(np.int32))
    cen = np.zeros((ne, 3), order="F")
    area, pnsz = np.zeros(ne), np.zeros(ne)
    dxyz = np.zeros((ne, 6), order="F")
    r.ref_geometry(nn, p(xyz), ne, p(ncn, C.c_int), p(ncon, C.c_int), p(np.zeros(3)), p(cen), p(area), p(pnsz),
                   p(dxyz))
    r.ref_set_mesh(nn, p(xyz), ne, p(ncn, C.c_int), p(ncon, C.c_int), p(cen))
    rows = []
    for j in list(range(0, ne, 5)) + [e for e in range(ne) if mesh.ncn[e] == 3][:20]:
        nrm = dxyz[j, :3]
        for kind in range(5):
            if kind == 0:
                xf = cen[j].copy()                                   # the panel's own centroid
            elif kind == 1:                                          # in the plane, off the centroid
                t = np.cross(nrm, [0.3, 0.5, 0.8])
                xf = cen[j] + 0.3 * pnsz[j] * t / np.linalg.norm(t)
            elif kind == 2:                                          # just off the plane
                xf = cen[j] + rng.choice([-1.0, 1.0]) * 2e-3 * pnsz[j] * nrm
            elif kind == 3:                                          # near, in the fluid
                xf = cen[j] + rng.uniform(-1.0, 1.0, 3) * pnsz[j]
            else:                                                    # far
                xf = cen[j] + rng.uniform(-30.0, 30.0, 3) * pnsz[j]
            rk = np.zeros(4)
            r.ref_sglint2(p(np.ascontiguousarray(xf)), j + 1, p(rk))
            rows.append([j, *xf, *rk])
    np.savez_compressed(os.path.join(HERE, "newman2.npz"), xyz=mesh.xyz, ncn=mesh.ncn, ncon=mesh.ncon,
                        rows=np.array(rows))


def _nums14(s):
    return [float(s[i:i + 14]) for i in range(0, len(s), 14) if s[i:i + 14].strip()]


def example2():
    """Field-point elevations of examples/example2_multi_bodies_wave_fields:
    per period, the radiation block (NFP points x NBODY lines of 12 values,
    OutputPressureElevation_RadiationMulti) then per heading the diffraction
    lines (OutputPressureElevation_DiffractionMulti), and the incidence file
    (OutputPressureElevation_IncidenceMulti: MOD, PHS, REL, IMG)."""
    ex = "/root/reference/examples/example2_multi_bodies_wave_fields"
    ctl = read_control(os.path.join(ex, "Input", "ControlFile.in"))
    nb, nfp, nper, nbeta = ctl.nbody, len(ctl.xfp), len(ctl.wvnb), len(ctl.headings_deg)
    bodies = [read_pnl(os.path.join(ex, "Input", f"HullMesh_{b + 1}.pnl"), ctl.lcs[b]) for b in range(nb)]
    lines = open(os.path.join(ex, "Output", "Wamit_format", "PressureElevation.6p")).read().splitlines()
    rad = np.zeros((nper, nfp, nb, 12))
    dif = np.zeros((nper, nbeta, nfp, 2))
    pos = 0
    for k in range(nper):
        for q in range(nfp):
            for b in range(nb):
                rad[k, q, b] = _nums14(lines[pos][24:])
                pos += 1
        for ib in range(nbeta):
            for q in range(nfp):
                t = lines[pos]
                dif[k, ib, q] = [float(t[38:52]), float(t[52:66])]
                pos += 1
    assert pos == len(lines)
    lines = open(os.path.join(ex, "Output", "Wamit_format", "PressureElevationIncidence.6p")).read().splitlines()
    inc = np.array([_nums14(t[38:]) for t in lines]).reshape(nper, nbeta, nfp, 4)
    np.savez_compressed(
        os.path.join(HERE, "example2.npz"), depth=ctl.depth, periods=np.array(ctl.wvnb), inft=ctl.inft,
        headings_deg=np.array(ctl.headings_deg), xr=ctl.xr, isol=ctl.isol, refl=ctl.refl, xfp=ctl.xfp,
        xyz=np.stack([m.xyz for m in bodies]), ncn=bodies[0].ncn, ncon=bodies[0].ncon, rad=rad, dif=dif, inc=inc)


def _motion4(path, nper, nbeta, nd):
    """Motion.4 (SolveMotionMulti, (2ES14.6,I6,4ES14.6)): per frequency and
    heading, rows MD = 1..nd of MAGN, PHS, NREL, NIMG."""
    rows = [ln for ln in open(path).read().splitlines() if ln.strip()]
    assert len(rows) == nper * nbeta * nd, (path, len(rows))
    vals = np.array([[float(t[34:48]), float(t[48:62]), float(t[62:76]), float(t[76:90])] for t in rows])
    mds = np.array([int(t[28:34]) for t in rows])
    assert np.all(mds.reshape(nper, nbeta, nd) == np.arange(1, nd + 1))
    return vals.reshape(nper, nbeta, nd, 4)


def _hydro(paths):
    hs = [read_hydrostatic(p) for p in paths]
    return np.array([[h.mass, h.blnr, h.bqdr, h.crs, h.kstf] for h in hs])   # nbody x 5 x 6 x 6


def motion():
    ex2 = "/root/reference/examples/example2_multi_bodies_wave_fields"
    c2 = read_control(os.path.join(ex2, "Input", "ControlFile.in"))
    np.savez_compressed(
        os.path.join(HERE, "motion.npz"),
        ex2_hydro=_hydro([os.path.join(ex2, "Input", f"Hydrostatic_{b + 1}.in") for b in range(c2.nbody)]),
        ex2_motion=_motion4(os.path.join(ex2, "Output", "Wamit_format", "Motion.4"), len(c2.wvnb),
                            len(c2.headings_deg), 6 * c2.nbody))


def example1():
    ex = "/root/reference/examples/example1_multi_bodies"
    ctl = read_control(os.path.join(ex, "Input", "ControlFile.in"))
    nb = ctl.nbody
    mesh = read_pnl(os.path.join(ex, "Input", "HullMesh_1.pnl"))
    for b in range(1, nb):   # the five hull files are the same mesh placed by the LCS
        other = read_pnl(os.path.join(ex, "Input", f"HullMesh_{b + 1}.pnl"))
        assert np.array_equal(other.xyz, mesh.xyz) and np.array_equal(other.ncon, mesh.ncon)
    np.savez_compressed(
        os.path.join(HERE, "example1.npz"), depth=ctl.depth, omegas=np.array(ctl.wvnb), inft=ctl.inft,
        headings_deg=np.array(ctl.headings_deg), xr=ctl.xr, lcs=ctl.lcs, isol=ctl.isol, refl=ctl.refl,
        xyz=mesh.xyz, ncn=mesh.ncn, ncon=mesh.ncon,
        hydro=_hydro([os.path.join(ex, "Input", f"Hydrostatic_{b + 1}.in") for b in range(nb)]),
        motion=_motion4(os.path.join(ex, "Output", "Wamit_format", "Motion.4"), len(ctl.wvnb),
                        len(ctl.headings_deg), 6 * nb),
        # 0_inf_frequency_limits = 1: the first two "frequencies" of the tables are the zero- and
        # infinite-frequency limits; OAMASS1..6 rows for them, (limit, body M, body N) order
        oamass_limits=np.array([read_hams_block_file(os.path.join(ex, "Output", "Hams_format", f"OAMASS{md}.txt"))
                                [:2 * nb * nb] for md in range(1, 7)]))


def writers():
    """Text of shipped output files, for the writer round trips: config1's
    18 HAMS tables whole, and the first lines of example2's WAMIT files."""
    cfg = os.path.join(REF_DATA, "expected_output")
    ex2 = "/root/reference/examples/example2_multi_bodies_wave_fields/Output/Wamit_format"
    out = {}
    for kind in ("OAMASS", "ODAMPING", "OEXFOR"):
        for md in range(1, 7):
            out[f"config1_{kind}{md}"] = np.array(open(os.path.join(cfg, f"{kind}{md}.txt")).read().splitlines())
    for name, count in (("AmssDamp.1", 360), ("ExcForce.3", 120), ("Motion.4", 120), ("Hydrostat.hst", 240),
                        ("PressureElevationIncidence.6p", 200)):
        lines = open(os.path.join(ex2, name)).read().splitlines()
        out["ex2_" + name.replace(".", "_")] = np.array(lines[:count])
    # .6p: the radiation lines of the first 10 points, and heading 0's diffraction lines (all points)
    lines = open(os.path.join(ex2, "PressureElevation.6p")).read().splitlines()
    out["ex2_6p_rad"] = np.array(lines[:100])
    out["ex2_6p_dif"] = np.array(lines[1000:1100])
    np.savez_compressed(os.path.join(HERE, "writers.npz"), **out)


def main():
    r = oracle.ref()
    rng = np.random.default_rng(12345)

    # ---- special functions ----
    xs = np.concatenate([rng.uniform(1e-4, 1.0, 60), rng.uniform(1.0, 9.0, 60), rng.uniform(9.0, 18.0, 40),
                         rng.uniform(18.0, 60.0, 40), [3.0, 3.5, 4.0, 9.0, 18.0, 35.0, 50.0]])
    out = {"x": xs}
    for name in ("error", "e1xa", "eix", "struve_h0", "struve_h1", "bessel_j0", "bessel_j1"):
        out[name] = np.array([getattr(r, "ref_" + name)(x) for x in xs])
    out["enxa"] = np.array([[r.ref_enxa(n, x) for x in xs] for n in range(1, 6)])
    k = np.zeros((len(xs), 2))
    j = np.zeros((len(xs), 4))
    for i, x in enumerate(xs):
        a, b = d(), d()
        r.ref_ik01a(x, C.byref(a), C.byref(b))
        k[i] = a.value, b.value
        q = [d() for _ in range(4)]
        r.ref_jy01b(x, *[C.byref(t) for t in q])
        j[i] = [t.value for t in q]
    out["ik01a"], out["jy01b"] = k, j
    seqs = []
    lim = []
    for n in (9, 20, 49, 98):
        terms = rng.normal(size=n) * np.exp(-0.08 * np.arange(n)) * (-1) ** np.arange(n)
        s = np.concatenate([[0.0], np.cumsum(terms)])[: n + 1]
        seqs.append(np.pad(s, (0, 100 - len(s))))
        lim.append(r.ref_limes(n, p(s.copy())))
    out["limes_seq"], out["limes_n"], out["limes"] = np.array(seqs), np.array([9, 20, 49, 98]), np.array(lim)
    np.savez_compressed(os.path.join(HERE, "special.npz"), **out)

    # ---- dispersion and wave properties ----
    ws = np.array([0.05, 0.1, 0.2, 0.5, 1.0, 1.7, 2.5, 3.0])
    hs = np.array([0.9, 14.5, 30.0, 50.0, 70.0, 200.0])
    roots = np.zeros((len(ws), len(hs), 200))
    for a, w in enumerate(ws):
        for b, h in enumerate(hs):
            r.ref_dispersion(w, h, 200, p(roots[a, b]))
    props = []
    for inft, val, h in [(1, 0.1, 50.0), (2, 0.2, 50.0), (3, 1.1, 50.0), (4, 6.0, 70.0), (5, 80.0, 30.0),
                         (1, 0.1, -1.0), (3, 1.1, -1.0), (4, 6.0, -1.0), (5, 80.0, -1.0), (2, 0.2, -1.0)]:
        v, wk, w1 = d(), d(), d()
        wv = np.zeros(200)
        r.ref_wave_properties(inft, val, h, C.byref(v), C.byref(wk), C.byref(w1), p(wv))
        props.append([inft, val, h, v.value, wk.value, w1.value])
    np.savez_compressed(os.path.join(HERE, "dispersion.npz"), w=ws, h=hs, roots=roots, props=np.array(props))

    # ---- Green functions ----
    rows = []
    for depth in (-1.0, 20.0, 50.0, 70.0, 200.0):
        for w in (0.1, 0.3, 1.0, 2.5):
            v = w * w / 9.80665
            wvn = np.zeros(200)
            if depth > 0:
                r.ref_dispersion(w, depth, 200, p(wvn))
            for t in range(60):
                src = rng.uniform([-15, -15, -12], [15, 15, -0.05])
                fld = rng.uniform([-15, -15, -12], [15, 15, -0.05])
                kind = t % 6
                if kind == 0:
                    fld[:2] = src[:2]                                      # R = 0
                elif kind == 1:
                    fld[:2] = src[:2] + rng.uniform(-1, 1, 2) * 2e-4 * abs(depth)  # LINTON
                elif kind == 2:
                    fld[:2] = src[:2] + rng.uniform(-1, 1, 2) * 0.03 * abs(depth)  # PIDCOCK
                elif kind == 3:
                    fld[:2] = src[:2] + rng.uniform(-1, 1, 2) * 0.3 * abs(depth)   # EIGENE
                flag = t % 2
                g = np.zeros(8)
                args = (src[0], fld[0], src[1], fld[1], src[2], fld[2], v)
                if depth < 0:
                    r.ref_infgreen3d(*args, flag, p(g))
                else:
                    r.ref_fingreen3d(*args, p(wvn), 200, depth, flag, p(g))
                rows.append(np.concatenate([[depth, w, flag], src, fld, g]))
    for depth in (-1.0, 50.0):     # V = 0 and V = -1 limits
        for v in (0.0, -1.0):
            for t in range(6):
                src = rng.uniform([-5, -5, -8], [5, 5, -0.1])
                fld = rng.uniform([-5, -5, -8], [5, 5, -0.1])
                flag = t % 2
                g = np.zeros(8)
                wvn = np.zeros(200)
                args = (src[0], fld[0], src[1], fld[1], src[2], fld[2], v)
                if depth < 0:
                    r.ref_infgreen3d(*args, flag, p(g))
                else:
                    r.ref_fingreen3d(*args, p(wvn), 200, depth, flag, p(g))
                rows.append(np.concatenate([[depth, v, 10 + flag], src, fld, g]))
    np.savez_compressed(os.path.join(HERE, "green.npz"), rows=np.array(rows))

    # ---- Newman integrals and geometry on body 1 of config1 ----
    mesh = read_pnl(os.path.join(REF_DATA, "input", "HullMesh_1.pnl"))
    ne, nn = mesh.nelem, mesh.nnode
    xyz = np.asfortranarray(mesh.xyz)
    ncn = mesh.ncn.astype(np.int32)
    ncon = np.asfortranarray(mesh.ncon.astype(np.int32))
    cen = np.zeros((ne, 3), order="F")
    area, pnsz = np.zeros(ne), np.zeros(ne)
    dxyz = np.zeros((ne, 6), order="F")
    xr = np.array([0.3, -0.2, -1.0])
    r.ref_geometry(nn, p(xyz), ne, p(ncn, C.c_int), p(ncon, C.c_int), p(xr), p(cen), p(area), p(pnsz), p(dxyz))
    r.ref_set_mesh(nn, p(xyz), ne, p(ncn, C.c_int), p(ncon, C.c_int), p(cen))
    pairs = [(i, i) for i in range(0, ne, 7)] + [tuple(rng.integers(0, ne, 2)) for _ in range(300)]
    pairs += [(i, j) for i in range(ne) for j in range(ne) if mesh.ncn[j] == 3][:40]
    nw = np.zeros((len(pairs), 4))
    for q, (i, j) in enumerate(pairs):
        s, dv = d(), np.zeros(3)
        r.ref_sglint(int(i) + 1, int(j) + 1, C.byref(s), p(dv))
        nw[q] = [s.value, *dv]
    np.savez_compressed(os.path.join(HERE, "newman.npz"), xyz=mesh.xyz, ncn=mesh.ncn, ncon=mesh.ncon, xr=xr,
                        cen=cen, area=area, pnsz=pnsz, dxyz=dxyz, pairs=np.array(pairs), sd=nw)

    # ---- incident wave ----
    rows = []
    for depth in (-1.0, 50.0, 600.0):
        for w in (0.3, 1.2):
            v = w * w / 9.80665
            if depth > 0:
                wvn = np.zeros(200)
                r.ref_dispersion(w, depth, 200, p(wvn))
                wk = wvn[0]
            else:
                wk = v
            for t in range(10):
                x, y, z = rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-10, -0.01)
                beta = rng.uniform(0, np.pi)
                dp = np.zeros(6)
                r.ref_dinp(x, y, z, 1.5, -2.0, beta, v, w, wk, depth, 1.3, p(dp))
                vi = np.zeros(2)
                r.ref_vinp(x, y, z, 1.5, -2.0, beta, v, w, wk, depth, 1.3, p(vi))
                rows.append([depth, w, v, wk, x, y, z, beta, *dp, vi[0], vi[1]])
    np.savez_compressed(os.path.join(HERE, "incident.npz"), rows=np.array(rows))

    # ---- config1 deck and expected outputs ----
    ctl = read_control(os.path.join(REF_DATA, "input", "ControlFile.in"))
    meshes = [read_pnl(os.path.join(REF_DATA, "input", f"HullMesh_{b + 1}.pnl"), ctl.lcs[b])
              for b in range(ctl.nbody)]
    tabs = {}
    for kind in ("OAMASS", "ODAMPING", "OEXFOR"):
        for md in range(1, 7):
            tabs[f"{kind}{md}"] = read_hams_block_file(os.path.join(REF_DATA, "expected_output", f"{kind}{md}.txt"))
    np.savez_compressed(
        os.path.join(HERE, "config1.npz"), depth=ctl.depth, omegas=np.array(ctl.wvnb), inft=ctl.inft,
        headings_deg=np.array(ctl.headings_deg), xr=ctl.xr, lcs=ctl.lcs, isol=ctl.isol,
        **{f"xyz{b}": m.xyz for b, m in enumerate(meshes)}, **{f"ncn{b}": m.ncn for b, m in enumerate(meshes)},
        **{f"ncon{b}": m.ncon for b, m in enumerate(meshes)}, **tabs)
    # ---- example 5: one body with a waterplane lid (If_remove_irr_freq = 1) ----
    ex5 = "/root/reference/examples/example5_single_body"
    hull = read_pnl(os.path.join(ex5, "Input", "HullMesh.pnl"))
    lid = read_pnl(os.path.join(ex5, "Input", "WaterPlaneMesh.pnl"))
    tabs5 = {}
    for kind in ("OAMASS", "ODAMPING", "OEXFOR"):
        for md in range(1, 7):
            tabs5[f"{kind}{md}"] = read_hams_block_file(os.path.join(ex5, "Output", "Hams_format", f"{kind}{md}.txt"))
    # ControlFile.in: depth 70, 10 frequencies 0.05 + 0.01 k rad/s (INFT 3),
    # 36 headings 10 k deg, reference point 0, Wave_diffrac_solution 1
    np.savez_compressed(
        os.path.join(HERE, "example5.npz"), depth=70.0, omegas=0.05 + 0.01 * np.arange(10),
        headings_deg=10.0 * np.arange(36), xr=np.zeros(3), isol=1, hull_xyz=hull.xyz, hull_ncn=hull.ncn,
        hull_ncon=hull.ncon, lid_xyz=lid.xyz, lid_ncn=lid.ncn, lid_ncon=lid.ncon, **tabs5)
    newman2()
    example2()
    motion()
    example1()
    writers()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            {"newman2": newman2, "example2": example2, "motion": motion, "example1": example1,
             "writers": writers}[name]()
    else:
        main()
