#!/usr/bin/env python3
"""Frequency-solve throughput of the MI355X panel solver on one GPU.

A step is one frequency solve: influence-matrix assembly (finite-depth or
deep-water Green function + Newman integrals for all N^2 panel pairs),
complex LU, the 6*NB radiation and NBETA diffraction solves, and the
added-mass / damping / excitation reductions, with one device
synchronisation (the forces come back to the host).  The mesh is resident in
HBM before the timed region starts.

Default workload: two truncated cylinders (R = 5 m, draft 10 m, 1,984 panels
each, centres 40 m apart on x), depth 50 m, 30 frequencies uniform in
[0.2, 2.5] rad/s, headings 0/45/90 deg, complex128 arithmetic.

Prints ONE JSON line.  Besides the common fields it carries
  * latency_ms: p50/p99 of the per-solve wall time;
  * stage_ms_per_solve: device time per stage from HIP events on the
    solver's own stream;
  * roofline: the kernel with the largest device time per solve against its
    bound (the finite-depth EIGENE region kernel: FP64 VALU), plus the LU
    factorisation against the MFMA peak;
  * cpu_baseline: the oracle (a C restatement of the reference's OpenMP
    Fortran loop) on a bounded row sample of the same assembly, plus LAPACK
    zgetrf/zgetrs on the full matrix, on the host's cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "hams-mrel_amd")]

import numpy as np  # noqa: E402

METRIC = "freq-solves/sec @4k-12k panels, mixed vs FP64 at 1e-6 rel; p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector and FP64 matrix (dense)
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second: 256 CUs x 4 SIMDs, 4 cycles each

# FP64 work and HBM traffic per panel pair (one A entry and one S entry) of
# each finite-depth region, summed over the region's kernels (hull-pair
# kernel + single-entry kernel), measured with rocprofv3 counters on the
# named workload (summary and commands:
# profiles/pmc_assembly_two_cylinders_fp64.md).  flops = lane-level FP64
# operations issued (FMA = 2, hardware transcendental = 1):
# SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F64 x 64 x the active-lane fraction
# SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU).  Traffic = FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0, the factors that make a 256 MiB device copy read 256 MiB
# and write 256 MiB.  The algorithmic bytes of every region are 32 per pair
# (A and S entries written once).  valu = wave-level VALU instructions
# (SQ_INSTS_VALU) per pair, for the fraction of the VALU issue slots busy.
REGION_COUNTERS = {
    "cylinders-2x1984-h50": {
        "asm_linton": {"flops_per_pair": 571398.4, "traffic_per_pair": (38857279 + 54625269) / 57728,
                       "valu_per_pair": 8.311e8 / 57728},
        "asm_pidcock": {"flops_per_pair": 113403.8, "traffic_per_pair": (7475511549 + 9363481643) / 1079424,
                        "valu_per_pair": 2.601e9 / 1079424},
        "asm_eigene": {"flops_per_pair": 43750.8, "traffic_per_pair": (6935846112 + 5435031093) / 6735360,
                       "valu_per_pair": 5.839e9 / 6735360},
        "asm_eigen": {"flops_per_pair": 1701.5, "traffic_per_pair": (30696267 + 252948651) / 7872512,
                      "valu_per_pair": 2.688e8 / 7872512},
    },
}
REGIONS = ("asm_linton", "asm_pidcock", "asm_eigene", "asm_eigen")


def region_ops(workload):
    """Algorithmic operations per entry of each finite-depth region, counted
    once on the CPU restatement of the reference's algorithm (region_ops.json,
    written by oracle/opcount.py: add/sub/mul/div/sqrt and each libm call = 1
    operation), or None when the file holds another workload."""
    try:
        with open(os.path.join(ROOT, "region_ops.json")) as f:
            d = json.load(f)
    except OSError:
        return None
    if d.get("workload") != workload:
        return None
    return {k: d["regions"][k[4:].upper()]["per_entry"]["ops"] for k in REGIONS if k[4:].upper() in d["regions"]}


def region_pairs(mesh, depth):
    """Pairs per finite-depth region (R/h < 5e-4, < 0.05, < 0.5, >= 0.5),
    classified with the kernels' own expression |sqrt(dx^2 + dy^2) / h|."""
    if depth < 0:
        return None
    c = mesh.cen
    out = np.zeros(4, dtype=np.int64)
    for i0 in range(0, mesh.nelem, 1024):
        dx = c[:, 0][None, :] - c[i0:i0 + 1024, 0][:, None]
        dy = c[:, 1][None, :] - c[i0:i0 + 1024, 1][:, None]
        rhp = np.abs(np.sqrt(dx * dx + dy * dy) / depth)
        out += [(rhp < 5e-4).sum(), ((rhp >= 5e-4) & (rhp < 0.05)).sum(), ((rhp >= 0.05) & (rhp < 0.5)).sum(),
                (rhp >= 0.5).sum()]
    return dict(zip(REGIONS, out.tolist()))


def roofline(case, mesh, ms, precision):
    """Roofline of the single kernel with the largest device time per solve.

    Candidates are kernels launched once per solve: the four finite-depth
    region kernels, or the deep-water grid assembly kernel.  The LU and the
    triangular solves are chains of hundreds of short launches; they are
    reported separately (roofline_lu)."""
    n = mesh.nelem
    if case.depth < 0:
        nbytes = 32.0 * n * n
        t = ms["assemble"]
        return {"kernel": "k_assemble<deep> (Havelock-fit Green function + Newman + A/S entries, one thread per "
                          "pair)", "bound": "valu", "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS, "launch_ms": t,
                "pairs": n * n, "algorithmic_bytes": nbytes, "achieved": None, "frac": None, "traffic": None}
    cands = [k for k in REGIONS if ms.get(k, 0.0) > 0.0]
    dom = max(cands, key=lambda k: ms[k])
    pairs = region_pairs(mesh, case.depth)[dom]
    cnt = REGION_COUNTERS.get(case.name, {}).get(dom)
    out = {"kernel": f"k_assemble_pairs<{dom[4:].upper()}> (finite-depth Green function region + Newman + A/S "
                     "entries; both orientations of a hull pair per thread)" if dom != "asm_linton" else
                     "k_assemble_list<LINTON> (finite-depth Green function region + Newman + A/S entries, one "
                     "thread per pair)",
           "bound": "valu", "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS, "launch_ms": ms[dom], "pairs": pairs,
           "algorithmic_bytes": 32.0 * pairs}
    if cnt:
        flops = cnt["flops_per_pair"] * pairs
        out.update({"achieved": flops / (ms[dom] * 1e-3) / 1e12,
                    "frac": flops / (ms[dom] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                    "traffic": cnt["traffic_per_pair"] * pairs, "flops_per_pair": cnt["flops_per_pair"],
                    "valu_issue_busy": cnt["valu_per_pair"] * pairs / (ms[dom] * 1e-3) / VALU_ISSUE_PER_S})
    else:
        out.update({"achieved": None, "frac": None, "traffic": None})
    alg = region_ops(case.name)
    if alg and dom in alg:
        # beside the issued rate above: the algorithm's own operation count
        # (a division or a libm call is one operation here, several issued
        # instructions on the GPU) at the measured launch time
        rate = alg[dom] * pairs / (ms[dom] * 1e-3) / 1e12
        out["algorithmic"] = {"ops_per_entry": alg[dom], "ops": alg[dom] * pairs, "achieved": rate,
                              "unit": "Tops/s", "frac": rate / PEAK_FP64_TFLOPS,
                              "issued_per_algorithmic": (cnt["flops_per_pair"] / alg[dom]) if cnt else None,
                              "source": "region_ops.json (oracle/opcount.py)"}
    return out


def solve_roofline(case, mesh, nbeta, ms_per_solve, lu_peak):
    """Whole-solve roofline: the time one solve would take with every part at
    its peak -- the assembly's FP64 work (the counter-measured operations per
    pair of each region, where measured) on the FP64 VALU, the LU's
    (8/3) N^3 on MFMA, the right-hand sides and the two triangular solves
    (8 N^2 per column each) on the FP64 VALU -- against the measured time per
    solve of the run (the throughput interval, solves in flight included)."""
    if case.depth < 0:
        return None
    cnt = REGION_COUNTERS.get(case.name)
    if not cnt:
        return None
    n = mesh.nelem
    pairs = region_pairs(mesh, case.depth)
    f_asm = sum(cnt[k]["flops_per_pair"] * pairs[k] for k in REGIONS)
    nrhs = 6 * mesh.nbody + nbeta
    f_lu = 8.0 * n ** 3 / 3.0
    f_vec = 3 * 8.0 * n * n * nrhs
    t_roof = (f_asm / PEAK_FP64_TFLOPS + f_lu / lu_peak + f_vec / PEAK_FP64_TFLOPS) / 1e9  # ms
    out = {"bound": "fp64 valu (assembly) + mfma (LU)", "flops": {"assembly": f_asm, "lu": f_lu, "rhs_solves": f_vec},
           "t_roof_ms": t_roof, "ms_per_solve": ms_per_solve, "frac": t_roof / ms_per_solve}
    alg = region_ops(case.name)
    if alg:
        a_asm = sum(alg[k] * pairs[k] for k in REGIONS if k in alg)
        t_alg = (a_asm / PEAK_FP64_TFLOPS + f_lu / lu_peak + f_vec / PEAK_FP64_TFLOPS) / 1e9
        out["algorithmic"] = {"assembly_ops": a_asm, "t_roof_ms": t_alg, "frac": t_alg / ms_per_solve}
    return out


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves per context before timing")
    ap.add_argument("--streams", type=int, default=None, help="frequency solves kept in flight (one context and HIP "
                    "stream each); default: the group size, or 4 slots for serving")
    ap.add_argument("--group", type=int, default=16, help="contexts per frequency group (one batched factorisation "
                    "for the group's matrices); 1: independent submissions")
    ap.add_argument("--stage-steps", type=int, default=3, help="isolated solves timed per stage for the roofline")
    ap.add_argument("--config", default="two-cylinders",
                    choices=["two-cylinders", "two-cylinders-half", "five-cylinders", "one-cylinder", "serving",
                             "large-array", "wave-field"])
    ap.add_argument("--requests", type=int, default=40, help="serving: requests in the Poisson stream")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: mean arrival rate, requests/s")
    ap.add_argument("--seed", type=int, default=7, help="serving: seed of the request stream")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default fp64 (serving: mixed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graphs", action="store_true", help="replay each context's recorded HIP graph of the solve "
                    "instead of launching it eagerly")
    ap.add_argument("--solver", default="lu", choices=["lu", "gmres"],
                    help="gmres: the reference's ISOLV = 2 path (GMRES(25), tolerance 1e-8, block-LU "
                         "preconditioner) on the H-matrix (eta 2, ACA 1e-6) instead of the dense LU; "
                         "independent submissions (no frequency groups)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision and five-cylinder runs that "
                    "the default FP64 line carries beside its value")
    ap.add_argument("--cpu-rows", type=int, default=256, help="field panels in the CPU baseline's assembly sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, host cores)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.solver == "gmres":
        args.group = 1
    if args.streams is None:
        args.streams = 4 if args.config == "serving" or args.group <= 1 else args.group
    args.precision_set = args.precision is not None
    if args.precision is None:
        args.precision = "fp64"
    return args


def copy_probe(mib):
    import torch
    x = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": x.numel() * 8}), flush=True)


def pick_case(name):
    from hamsx import configs
    if name == "two-cylinders":
        return configs.two_cylinders()
    if name == "five-cylinders":
        return configs.five_cylinders()
    if name == "two-cylinders-half":
        return configs.two_cylinders_half()
    return configs.single_cylinder()


def cpu_baseline(case, mesh, rows, threads):
    """Seconds per frequency solve on the host.

    With the reference checker built (oracle/_ref/libhams_ref.so, the
    reference's own Fortran compiled from its sources): the reference's own
    pair loop CALGREEN_MULTI (CalGreenFuncMulti.f90:112-183, its OpenMP loop,
    Green values stored in FP32 as the shipped build does) over all N^2
    pairs on `threads` threads -- kind "reference".  Otherwise the oracle's
    assembly (a C restatement of the same loop) on an evenly strided sample of
    `rows` field panels, scaled by N / rows -- kind "port".  Either way LAPACK
    zgetrf + zgetrs (scipy's OpenBLAS, for the reference's MKL) on a full
    N x N matrix with the same number of right-hand sides is added.  The
    right-hand sides and force sums are O(N^2) and O(N) and are left out
    (they favour the CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import scipy.linalg as sl
    n = mesh.nelem
    w = float(case.omegas[len(case.omegas) // 2])
    idx = np.unique(np.linspace(0, n - 1, min(rows, n)).astype(np.int32))
    t0 = time.perf_counter()
    oracle.assemble_rows(mesh, idx, w, case.depth, nthreads=threads)
    t_sample = time.perf_counter() - t0
    t_port = t_sample * n / len(idx)
    kind, t_asm = "port", t_port
    sample = (f"oracle assembly of {len(idx)} of {n} rows (all {n} columns) at w={w:.3f} rad/s, "
              f"{t_sample:.1f} s on {threads} threads, scaled to N rows = {t_port:.1f} s")
    if oracle.ref_available():
        t0 = time.perf_counter()
        chk = oracle.ref_calgreen(mesh, w, case.depth, threads)
        t_ref = time.perf_counter() - t0
        if np.isfinite(chk):
            kind, t_asm = "reference", t_ref
            sample = (f"the reference's CALGREEN_MULTI (its OpenMP pair loop, built from its sources) over all "
                      f"{n}x{n} pairs at w={w:.3f} rad/s: {t_ref:.1f} s on {threads} threads (the C port on a "
                      f"{len(idx)}-row sample, scaled: {t_port:.1f} s)")
    rng = np.random.default_rng(0)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 0.01 + 2 * math.pi * np.eye(n)
    A = np.asfortranarray(A)
    nr = 6 * mesh.nbody + len(case.headings_deg)
    B = np.asfortranarray(rng.standard_normal((n, nr)) + 0j)
    t0 = time.perf_counter()
    lu = sl.lu_factor(A, overwrite_a=True, check_finite=False)
    sl.lu_solve(lu, B, check_finite=False)
    t_lu = time.perf_counter() - t0
    per = t_asm + t_lu
    return {"value": 1.0 / per, "unit": "freq-solves/s", "cores": threads, "kind": kind,
            "sample": sample + f"; + scipy/OpenBLAS zgetrf+zgetrs N={n}, {nr} rhs = {t_lu:.2f} s",
            "seconds_per_solve": per, "port_seconds_per_solve": t_port + t_lu}


def serving(args):
    """Poisson stream of seeded uneven requests (1-9 cylinders/boxes, 500-12,000
    panels, w ~ U[0.1, 3], heading ~ U[0, 180), depth infinite or U[30, 200])
    through the continuous-batching server; value = completed requests per
    second of makespan, latency = completion minus arrival."""
    from hamsx import configs
    from hamsx.mesh import combine
    from hamsx.serving import Request, Server, poisson_arrivals
    rng = np.random.default_rng(args.seed)
    arr = poisson_arrivals(args.requests, args.rate, rng)
    reqs, sizes = [], []
    for k in range(args.requests):
        case = configs.random_request(rng)
        mesh = combine(case.bodies, case.xr)
        sizes.append(mesh.nelem)
        reqs.append(Request(k, mesh, float(case.omegas[0]), case.depth, np.deg2rad(case.headings_deg),
                            arrival=float(arr[k])))
    precision = args.precision if args.precision_set else "mixed"
    print(f"[bench] serving {args.requests} requests, {min(sizes)}-{max(sizes)} panels, rate {args.rate}/s, "
          f"{args.streams} slots, {precision}", file=sys.stderr, flush=True)
    def run(stream, tuning=None):
        with Server(slots=args.streams, precision=precision, tuning=tuning) as srv:
            warm = Request(-1, reqs[0].mesh, 1.0, reqs[0].depth, reqs[0].headings_rad)
            srv.run([warm])
            t0 = time.perf_counter()
            res = srv.run(stream)
            wall = time.perf_counter() - t0
        bad = [r.rid for r in res if r.info != 0]
        if bad:
            raise SystemExit(f"zero pivots in requests {bad}")
        lat = np.array([r.latency for r in res]) * 1e3
        span = max(r.done for r in res) - min(r.arrival for r in res)
        return res, lat, span, wall

    res, lat, span, wall = run(reqs)
    # capacity: the same requests all queued at once (the server never idles),
    # with the autotuner's per-size LU options and with the library defaults
    cap = {}
    for tag, tuning in (("tuned", None), ("defaults", {})):
        burst = [Request(r.rid, r.mesh, r.w1, r.depth, r.headings_rad, arrival=0.0) for r in reqs]
        cres, clat, cspan, _ = run(burst, tuning)
        cap[tag] = {"value": len(cres) / cspan, "unit": "freq-solves/s", "p50_ms": float(np.percentile(clat, 50)),
                    "p99_ms": float(np.percentile(clat, 99))}
        print(f"[bench] saturated, {tag} options: {cap[tag]['value']:.2f} solves/s", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": len(res) / span, "unit": "freq-solves/s", "n_gpus": 1, "steps": len(res),
        "warmup": 1, "ms_per_step": span / len(res) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64-lu+f64-refine" if precision == "mixed" else "f64",
        "data": f"synthetic seeded requests (seed {args.seed})",
        "config": {"workload": f"serving-poisson-{args.rate:g}hz-seed{args.seed}", "requests": len(res),
                   "panels_min": min(sizes), "panels_max": max(sizes), "panels_mean": float(np.mean(sizes)),
                   "slots": args.streams, "precision": precision, "wall_s": wall},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "capacity": cap,
    }
    print(json.dumps(out), flush=True)


def large_array(args):
    """16-cylinder array with lids (about 30,000 hull panels, depth 80 m):
    GMRES on the H-matrix operator (eta 2, ACA tolerance 1e-6, GMRES
    tolerance 1e-8) per frequency, against the dense FP64 LU at the first
    frequency.  One warm-up solve (pair binning, H-matrix partition) is
    untimed; then `steps` frequencies spread over [0.2, 2.0] rad/s."""
    from hamsx import configs
    from hamsx.mesh import combine, panel_geometry
    from hamsx.solver import FrequencySolver
    case = configs.large_array()
    mesh = combine(case.bodies, case.xr)
    lid = np.concatenate([panel_geometry(lm, np.zeros(3))[0] for lm in case.meta["lids"]])
    heads = np.deg2rad(case.headings_deg)
    steps = args.steps if "--steps" in sys.argv else 2
    omegas = [float(w) for w in np.linspace(case.omegas[0], case.omegas[-1], steps)]
    # mixed (the default here): complex64 H-matrix blocks in the Arnoldi
    # products, FP64 Krylov vectors and restart residuals, FP32 far field
    precision = args.precision if args.precision_set else "mixed"
    print(f"[bench] {case.name}: N={mesh.nelem} hull panels, {len(lid)} lid points, depth {case.depth}, "
          f"{steps} timed frequencies, {precision}", file=sys.stderr, flush=True)
    with FrequencySolver(mesh, precision=precision) as s:
        s.set_lid(lid)
        s.set_waves(heads)
        s.set_solver("gmres", tol=1e-8, restart=25, maxiter=400)
        s.set_hmatrix(True, 2.0, 1e-6, 32, 50)
        t0 = time.perf_counter()
        warm = s.solve_frequency(omegas[0], case.depth)
        t_warm = time.perf_counter() - t0
        print(f"[bench] warm-up solve {t_warm:.2f} s", file=sys.stderr, flush=True)
        s.set_profiling(True)
        s.stage_times(reset=True)
        lat, iters = [], []
        t0 = time.perf_counter()
        for w in omegas:
            t1 = time.perf_counter()
            _, _, _, info = s.solve_frequency(w, case.depth)
            lat.append(time.perf_counter() - t1)
            iters.append(s.refine_stats())
            if info != 0:
                raise SystemExit(f"INFO={info} at w={w}")
            print(f"[bench] w={w:.3f}: {lat[-1]:.2f} s, {iters[-1][0]} GMRES iterations", file=sys.stderr,
                  flush=True)
        elapsed = time.perf_counter() - t0
        stages = s.stage_times(reset=True)
        s.set_profiling(False)
        nlr, nd, mrank, ratio = s.hmatrix_stats()
        s.set_solver("lu")
        s.set_precision("fp64")
        t1 = time.perf_counter()
        lu = s.solve_frequency(omegas[0], case.depth)
        t_lu = time.perf_counter() - t1
    errs = [float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(warm[:3], lu[:3])]
    own = []
    for a, b in zip(warm[:3], lu[:3]):
        big = np.abs(b) >= 1e-3 * np.max(np.abs(b))
        own.append(float(np.max(np.abs(a - b)[big] / np.abs(b)[big])))
    ms = {k: v[0] / max(1, v[1]) for k, v in stages.items()}
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": steps / elapsed, "unit": "freq-solves/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if precision == "fp64" else "c64-hblocks+f64-gmres",
        "data": "synthetic cylinder array with lids (no reference files read)",
        "config": {"workload": case.name, "panels": mesh.nelem, "lid_points": int(len(lid)), "depth_m": case.depth,
                   "solver": "gmres(25) tol 1e-8 + block-LU preconditioner on the H-matrix (eta 2, ACA 1e-6)",
                   "precision": precision,
                   "frequencies": omegas},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "stage_ms_per_solve": ms,
        "gmres_iterations": [int(i[0]) for i in iters], "gmres_residual": [float(i[1]) for i in iters],
        "hmatrix": {"low_rank_blocks": nlr, "dense_blocks": nd, "mean_rank": mrank, "compression": ratio},
        "vs_dense_lu": {"amas": errs[0], "bdmp": errs[1], "exfc": errs[2], "lu_solve_s": t_lu,
                        "own_magnitude": {"amas": own[0], "bdmp": own[1], "exfc": own[2]},
                        "note": "at the first frequency, each table's largest difference over its largest entry; "
                                "own_magnitude: each coefficient above 1e-3 of its table's largest against itself"},
    }
    print(json.dumps(out), flush=True)


def wave_field_points():
    """A 64 x 64 free-surface grid on [-120, 120]^2 m around the two
    cylinders (points inside a hull waterline dropped) and a vertical line of
    32 pressure points between them, z in [-45, -1] m."""
    g = np.linspace(-120.0, 120.0, 64)
    X, Y = np.meshgrid(g, g, indexing="ij")
    fs = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
    keep = (np.hypot(fs[:, 0] + 20.0, fs[:, 1]) > 5.5) & (np.hypot(fs[:, 0] - 20.0, fs[:, 1]) > 5.5)
    line = np.column_stack([np.zeros(32), np.full(32, 3.0), np.linspace(-45.0, -1.0, 32)])
    return np.vstack([fs[keep], line])


def wave_field(args):
    """Two cylinders (the default workload) with wave-field output: each step
    is one frequency solve followed by the field values of every mode at
    every point (pressure below the free surface, elevation on it), copied to
    the host, on one context."""
    from hamsx import configs
    from hamsx.mesh import combine
    from hamsx.solver import FrequencySolver
    case = configs.two_cylinders()
    mesh = combine(case.bodies, case.xr)
    pts = wave_field_points()
    heads = np.deg2rad(case.headings_deg)
    omegas = [float(w) for w in case.omegas]
    steps = args.steps
    print(f"[bench] {case.name} + wave field: N={mesh.nelem} panels, {len(pts)} field points, {steps} frequencies",
          file=sys.stderr, flush=True)
    with FrequencySolver(mesh) as s:
        s.set_waves(heads, amp=case.amp, xw=case.xw, rho=case.rho)
        s.set_field_points(pts)
        for k in range(max(1, args.warmup)):
            s.solve_frequency(omegas[k % len(omegas)], case.depth)
            s.field_values("auto")
        s.set_profiling(True)
        s.stage_times(reset=True)
        lat = []
        t0 = time.perf_counter()
        for k in range(steps):
            t1 = time.perf_counter()
            _, _, _, info = s.solve_frequency(omegas[k % len(omegas)], case.depth)
            vals, onbody = s.field_values("auto")
            lat.append(time.perf_counter() - t1)
            if info != 0:
                raise SystemExit(f"zero pivot in solve {k}: INFO={info}")
        elapsed = time.perf_counter() - t0
        stages = s.stage_times(reset=True)
    ms = {k: v[0] / max(1, v[1]) for k, v in stages.items()}
    lat_ms = np.array(lat) * 1e3
    pairs = len(pts) * mesh.nelem
    out = {
        "metric": METRIC, "value": steps / elapsed, "unit": "freq-solves/s", "n_gpus": 1, "steps": steps,
        "warmup": max(1, args.warmup), "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic seeded cylinder meshes and field grid (no reference files read)",
        "config": {"workload": case.name + "+wave-field-4k", "panels": mesh.nelem, "field_points": int(len(pts)),
                   "depth_m": case.depth, "frequencies": len(omegas), "headings_deg": case.headings_deg,
                   "in_flight": 1},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "stage_ms_per_solve": ms,
        "field": {"ms_per_solve": ms["field"], "pairs": pairs,
                  "pairs_per_s": pairs / (ms["field"] * 1e-3) if ms["field"] > 0 else None},
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        base = cpu_baseline(case, mesh, args.cpu_rows, threads)
        # the oracle's CalPotentialMulti on a strided sample of the points (any potentials will do:
        # the cost is the pair loop), scaled to all points
        w = float(case.omegas[len(case.omegas) // 2])
        sample = pts[np.linspace(0, len(pts) - 1, 128).astype(int)]
        phi = np.ones((mesh.nelem, 6 * mesh.nbody + len(heads)), dtype=np.complex128, order="F")
        t0 = time.perf_counter()
        oracle.field_values(mesh, phi, sample, w, case.depth, heads, kind=2, nthreads=threads)
        t_field = (time.perf_counter() - t0) * len(pts) / len(sample)
        per = base["seconds_per_solve"] + t_field
        base.update({"value": 1.0 / per, "seconds_per_solve": per,
                     "sample": base["sample"] + f"; + oracle field values of 128 of {len(pts)} points, scaled = "
                                                f"{t_field:.1f} s"})
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    # Each context drives two HIP streams (its own and the LU look-ahead's, or
    # a symmetry plane's second system).  With several contexts in flight the
    # HIP runtime's default of four hardware queues makes streams share
    # queues, and a kernel waiting on an event in one stream then stalls the
    # others: measured 18.9 -> 19.8 solves/s and p99 307 -> 219 ms on the
    # default workload with eight queues.  Raised (never lowered) before the
    # runtime starts, to eight and no further: every queue that runs the
    # PIDCOCK kernel keeps 2.37 GB of scratch (4.5 kB per lane for every wave
    # slot of the GPU), and the GPU's scratch limit shared by all queues is
    # 32 GiB, which sixteen queues overrun (the dispatch then fails with
    # HSA_STATUS_ERROR_OUT_OF_RESOURCES).
    want = "8"
    try:
        have = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        have = 4
    if have < int(want):
        os.environ["GPU_MAX_HW_QUEUES"] = want
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs one GPU per process (asked for --gpus {args.gpus}); this build has no multi-GPU "
              "path -- run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == "serving":
        serving(args)
        return
    if args.config == "large-array":
        large_array(args)
        return
    if args.config == "wave-field":
        wave_field(args)
        return
    from hamsx.mesh import combine

    case = pick_case(args.config)
    mesh = combine(case.bodies, case.xr)
    n = mesh.nelem
    print(f"[bench] {case.name}: N={n} panels, depth {case.depth}, {len(case.omegas)} frequencies, "
          f"{len(case.headings_deg)} headings, precision {args.precision}, {max(1, args.streams)} solves in flight",
          file=sys.stderr, flush=True)
    group = args.group if args.group > 1 and "symmetry" not in case.meta else 1
    run = throughput(case, mesh, args.precision, args.streams, args.steps, args.warmup, args.stage_steps,
                     graphs=args.graphs, group=group, solver=args.solver)
    ms, elapsed, lat_ms = run["stages"], run["elapsed"], run["lat_ms"]
    print(f"[bench] timed {args.steps} solves in {elapsed:.2f} s", file=sys.stderr, flush=True)
    roof = roofline(case, mesh, ms, args.precision)
    lu_peak = PEAK_FP64_TFLOPS if args.precision == "fp64" else PEAK_FP32_MFMA_TFLOPS
    lu_ms = ms.get("factor_group", ms["factor"])
    roof_lu = roof_solve = None
    if lu_ms > 0.0:  # the dense LU path (GMRES factors no matrix)
        lu_tf = 8.0 * n ** 3 / 3.0 / (lu_ms * 1e-3) / 1e12
        roof_lu = {"kernel": "LU factorisation" + (f" (one batched launch sequence for a group of {group} matrices; "
                                                   "time per matrix)" if group > 1 else ""),
                   "bound": "mfma", "unit": "TFLOP/s", "achieved": lu_tf, "peak": lu_peak,
                   "frac": lu_tf / lu_peak, "launch_ms": lu_ms, "single_matrix_ms": ms["factor"],
                   "flops": 8.0 * n ** 3 / 3.0}
        roof_solve = solve_roofline(case, mesh, len(case.headings_deg), elapsed / args.steps * 1e3, lu_peak)
    vs_full = None
    if "symmetry" in case.meta:
        vs_full = half_vs_full(case, mesh, args.precision)
    out = {
        "metric": METRIC, "value": args.steps / elapsed, "unit": "freq-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "c64-lu+f64-refine",
        "data": "synthetic seeded cylinder meshes (no reference files read)",
        "config": {"workload": case.name, "panels": n, "symmetry": case.meta.get("symmetry"),
                   "depth_m": case.depth, "frequencies": len(case.omegas),
                   "headings_deg": case.headings_deg, "precision": args.precision,
                   "in_flight": max(1, args.streams), "group": group, "graphs": args.graphs,
                   "solver": args.solver},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "stage_ms_per_solve": ms,
        "roofline": roof,
        "roofline_lu": roof_lu,
        "roofline_solve": roof_solve,
    }
    if vs_full is not None:
        out["vs_full_mesh"] = vs_full
    if (args.precision == "fp64" and not args.no_mixed and args.solver == "lu"
            and args.config in ("two-cylinders", "five-cylinders")):
        # the headline: the mixed-precision path on the same workload, and both
        # paths on the five-cylinder (10,026-panel) workload
        out["mixed"] = mixed_vs_fp64(case, mesh, run, args.streams, args.steps, args.warmup, group)
        if args.config == "two-cylinders":
            c5 = pick_case("five-cylinders")
            m5 = combine(c5.bodies, c5.xr)
            steps5 = max(4, args.steps // 2)
            print(f"[bench] {c5.name}: N={m5.nelem}, fp64 and mixed, {steps5} solves each", file=sys.stderr, flush=True)
            r5 = throughput(c5, m5, "fp64", args.streams, steps5, args.warmup, 1, graphs=args.graphs, group=group)
            out["five_cylinders"] = {"workload": c5.name, "panels": m5.nelem, "depth_m": c5.depth,
                                     "fp64": summary(r5, steps5),
                                     "mixed": mixed_vs_fp64(c5, m5, r5, args.streams, steps5, args.warmup, group)}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"[bench] CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(case, mesh, args.cpu_rows, threads)
    print(json.dumps(out), flush=True)


def throughput(case, mesh, precision, streams, steps, warmup, stage_steps, graphs=False, group=1, solver="lu"):
    """`steps` frequency solves of case.omegas (in order, cycling) with
    `streams` solves in flight (one context and HIP stream each), after
    per-stage timing of `stage_steps` isolated solves and `warmup` untimed
    solves per context.  With group > 1 the contexts are submitted in
    frequency groups of `group` (hamsx_submit_frequency_group: one batched
    factorisation per group), and the factorisation's time per solve in a
    group is timed too.  Returns the stage times, the wall time, per-solve
    latencies (submit to collect) and every solve's coefficients in order."""
    from hamsx.solver import FrequencySolver
    heads = np.deg2rad(case.headings_deg)
    omegas = [float(w) for w in case.omegas]
    nctx = max(1, streams)
    solvers = []
    for _ in range(nctx):
        sv = FrequencySolver(mesh, precision=precision)
        sv.set_options(graphs=1 if graphs else 0)
        sv.set_waves(heads, amp=case.amp, xw=case.xw, rho=case.rho)
        if solver == "gmres":
            sv.set_solver("gmres", tol=1e-8, restart=25, maxiter=400)
            sv.set_hmatrix(True, 2.0, 1e-6, 32, 50)
        if "symmetry" in case.meta:
            sv.set_symmetry(*case.meta["symmetry"])
        solvers.append(sv)
    # per-stage device time of isolated solves (one context, nothing else on the GPU)
    solvers[0].solve_frequency(omegas[0], case.depth)
    solvers[0].set_profiling(True)
    solvers[0].stage_times(reset=True)
    for k in range(stage_steps):
        solvers[0].solve_frequency(omegas[(k * 7) % len(omegas)], case.depth)
    stages = solvers[0].stage_times(reset=True)
    solvers[0].set_profiling(False)
    if group > 1:
        return _throughput_groups(case, solvers, omegas, steps, warmup, stage_steps, group, stages)
    for k, sv in enumerate(solvers):
        for _ in range(warmup):  # through the timed path (the first submit at a depth runs eagerly)
            sv.submit(omegas[k % len(omegas)], case.depth)
            sv.collect()
    lat, coef = [], [None] * steps
    t_sub = [0.0] * nctx
    owner = [0] * nctx  # solve index each context holds
    submitted = done = 0
    t0 = time.perf_counter()
    while submitted < min(nctx, steps):
        t_sub[submitted] = time.perf_counter()
        owner[submitted] = submitted
        solvers[submitted].submit(omegas[submitted % len(omegas)], case.depth)
        submitted += 1
    while done < steps:
        k = done % nctx
        am, bd, ex, info = solvers[k].collect()
        lat.append(time.perf_counter() - t_sub[k])
        if info != 0:
            raise SystemExit(f"zero pivot in solve {done}: INFO={info}")
        coef[owner[k]] = (am, bd, ex)
        done += 1
        if submitted < steps:
            t_sub[k] = time.perf_counter()
            owner[k] = submitted
            solvers[k].submit(omegas[submitted % len(omegas)], case.depth)
            submitted += 1
    elapsed = time.perf_counter() - t0
    for sv in solvers:
        sv.close()
    return {"stages": {k: v[0] / max(1, v[1]) for k, v in stages.items()}, "elapsed": elapsed,
            "lat_ms": np.array(lat) * 1e3, "coef": coef}


def _throughput_groups(case, solvers, omegas, steps, warmup, stage_steps, group, stages):
    """The timed loop of throughput() for frequency groups: each group of
    contexts is submitted with one call and collected context by context;
    a finished group takes the next frequencies (the last one fewer if fewer
    remain)."""
    from hamsx.solver import FrequencySolver
    groups = [solvers[i:i + group] for i in range(0, len(solvers), group)]
    for _ in range(warmup):
        for g in groups:
            FrequencySolver.submit_group(g, [omegas[k % len(omegas)] for k in range(len(g))], case.depth)
            for sv in g:
                sv.collect()
    # the batched factorisation's device time, per solve of a full group
    g0 = groups[0]
    g0[0].set_profiling(True)
    g0[0].stage_times(reset=True)
    for k in range(stage_steps):
        FrequencySolver.submit_group(g0, [omegas[(k * 7 + i) % len(omegas)] for i in range(len(g0))], case.depth)
        for sv in g0:
            sv.collect()
    st = g0[0].stage_times(reset=True)
    g0[0].set_profiling(False)
    f_ms, f_cnt = st.get("factor", (0.0, 0))
    stages["factor_group"] = (f_ms / max(1, f_cnt) / len(g0), 1)
    lat, coef = [], [None] * steps
    t_sub = {}
    owner = {}
    submitted = done = 0
    queue = []
    t0 = time.perf_counter()

    def submit(gi):
        nonlocal submitted
        g = groups[gi][:min(len(groups[gi]), steps - submitted)]
        now = time.perf_counter()
        for i, sv in enumerate(g):
            t_sub[id(sv)] = now
            owner[id(sv)] = submitted + i
        FrequencySolver.submit_group(g, [omegas[(submitted + i) % len(omegas)] for i in range(len(g))], case.depth)
        submitted += len(g)
        queue.append(g)

    for gi in range(len(groups)):
        if submitted < steps:
            submit(gi)
    while queue:
        g = queue.pop(0)
        for sv in g:
            am, bd, ex, info = sv.collect()
            lat.append(time.perf_counter() - t_sub[id(sv)])
            if info != 0:
                raise SystemExit(f"zero pivot in solve {owner[id(sv)]}: INFO={info}")
            coef[owner[id(sv)]] = (am, bd, ex)
            done += 1
        if submitted < steps:
            submit(next(gi for gi, gg in enumerate(groups) if gg[0] is g[0]))
    elapsed = time.perf_counter() - t0
    for sv in solvers:
        sv.close()
    return {"stages": {k: v[0] / max(1, v[1]) for k, v in stages.items()}, "elapsed": elapsed,
            "lat_ms": np.array(lat) * 1e3, "coef": coef}


def summary(run, steps):
    return {"value": steps / run["elapsed"], "unit": "freq-solves/s",
            "latency_ms": {"p50": float(np.percentile(run["lat_ms"], 50)),
                           "p99": float(np.percentile(run["lat_ms"], 99))},
            "stage_ms_per_solve": {k: v for k, v in run["stages"].items() if v > 0}}


def mixed_vs_fp64(case, mesh, fp64_run, streams, steps, warmup, group=1):
    """The mixed-precision path (complex64 LU on MFMA + FP64 refinement) on the
    same workload and solve sequence as fp64_run: its throughput, latency, the
    ratio to the FP64 path, and the largest difference of every coefficient
    table from the FP64 path's over the run (relative to that table's largest
    entry at that frequency, and per coefficient against its own magnitude
    for those above 1e-3 of the largest)."""
    print(f"[bench] {case.name}: mixed precision, {steps} solves", file=sys.stderr, flush=True)
    mx = throughput(case, mesh, "mixed", streams, steps, warmup, 1, group=group)
    out = summary(mx, steps)
    out["ratio_vs_fp64"] = out["value"] / (steps / fp64_run["elapsed"])
    err = {"amas": 0.0, "bdmp": 0.0, "exfc": 0.0}
    own = dict(err)
    for a, b in zip(mx["coef"], fp64_run["coef"]):
        for key, x, y in zip(("amas", "bdmp", "exfc"), a, b):
            m = float(np.max(np.abs(y)))
            err[key] = max(err[key], float(np.max(np.abs(x - y))) / m)
            big = np.abs(y) >= 1e-3 * m
            own[key] = max(own[key], float(np.max(np.abs(x - y)[big] / np.abs(y)[big])))
    out["max_rel_err_vs_fp64"] = err
    out["max_rel_err_vs_fp64_own_magnitude"] = own
    return out


def half_vs_full(case, mesh, precision):
    """The same frequency on the whole mesh: the plane must not change the answer."""
    from hamsx import configs
    from hamsx.mesh import combine
    from hamsx.solver import FrequencySolver
    heads = np.deg2rad(case.headings_deg)
    w = float(case.omegas[0])
    full = configs.two_cylinders(nfreq=len(case.omegas))
    with FrequencySolver(combine(full.bodies, full.xr), precision=precision) as sv:
        sv.set_waves(heads, amp=case.amp, xw=case.xw, rho=case.rho)
        t1 = time.perf_counter()
        fa, fb, fe, _ = sv.solve_frequency(w, case.depth)
        t_full = time.perf_counter() - t1
    with FrequencySolver(mesh, precision=precision) as sv:
        sv.set_waves(heads, amp=case.amp, xw=case.xw, rho=case.rho)
        sv.set_symmetry(*case.meta["symmetry"])
        ha, hb, he, _ = sv.solve_frequency(w, case.depth)

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return {"w": w, "amas": rel(ha, fa), "bdmp": rel(hb, fb), "exfc": rel(he, fe), "full_mesh_solve_ms": t_full * 1e3}


if __name__ == "__main__":
    main()
