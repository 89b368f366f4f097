"""Algorithmic operation counts of the assembly per Green-function region,
counted once on the C restatement (hams_oracle.c) and written to
region_ops.json at the repository root, which bench.py reports beside the
counter-measured (issued) operation rate of the GPU kernels.

MEASUREMENT INFRASTRUCTURE, like the rest of oracle/: run by hand on the CPU
(`python oracle/opcount.py`), never by the product package.

Method.  `make -C oracle opcount` compiles hams_oracle.c at -O0 with no
inlining, no FMA contraction and C99 complex arithmetic expanded inline
(-fcx-limited-range), twice: with gcov coverage (exact per-line execution
counts) and to assembly (a per-line census of the scalar FP64 instructions
addsd/subsd/mulsd/divsd/sqrtsd and of the libm calls, from the .loc line
markers).  At -O0 every operation the source writes is executed as written
(no common-subexpression elimination), so

    ops = sum over source lines of  executions(line) x operations(line)

is the operation count of the algorithm as restated, with the conventions:
an add, subtract, multiply, divide or square root is one operation; a call
to exp, log, sin, cos, sinh, cosh, pow, atan2 or cabs is one operation
(reported separately as "libm", so other weights can be applied).  Comparisons,
fabs and integer work are not counted.

Sample.  The default workload's mesh (two cylinders of 1,984 panels, depth
50 m) is classified by the kernels' own region test on the centroid pair
(|sqrt(dx^2 + dy^2) / h| against 5e-4, 0.05, 0.5) and by FLAG (centroid
distance <= 50 x panel size, which adds the Newman integrals).  For each
(region, FLAG) class a seeded sample of entries (i, j) is evaluated with
orc_assemble_entry -- one A(i, j) and S(i, j) per call, the unit the GPU
kernels' pair counts use -- at several frequencies of the workload, in a
fresh process (so the class's coverage counts stand alone).  The per-entry
mean of each class is weighted by the class's share of the workload.
"""
from __future__ import annotations

import argparse
import collections
import gzip
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT_DIR = os.path.join(HERE, "opcount")
COV_SO = os.path.join(OUT_DIR, "liboracle_cov.so")
ASM = os.path.join(OUT_DIR, "hams_oracle.s")
SRC = os.path.join(HERE, "hams_oracle.c")
LIBM = ("exp", "log", "sin", "cos", "sinh", "cosh", "pow", "atan2", "atan", "tanh", "cabs", "log10")
REGIONS = ("LINTON", "PIDCOCK", "EIGENE", "EIGEN")
FP = {"addsd": "add_sub_mul", "subsd": "add_sub_mul", "mulsd": "add_sub_mul", "divsd": "div", "sqrtsd": "sqrt"}


def line_census(asm_path, src_name="hams_oracle.c"):
    """{line: Counter(kind -> n)} for the source file's lines."""
    files, cur_file, cur_line = {}, None, None
    census = collections.defaultdict(collections.Counter)
    for raw in open(asm_path):
        s = raw.strip()
        m = re.match(r'\.file\s+(\d+)\s+"([^"]+)"(?:\s+"([^"]+)")?', s)
        if m:
            files[int(m.group(1))] = m.group(3) or m.group(2)
            continue
        m = re.match(r"\.loc\s+(\d+)\s+(\d+)", s)
        if m:
            cur_file, cur_line = int(m.group(1)), int(m.group(2))
            continue
        if cur_line is None or not files.get(cur_file, "").endswith(src_name):
            continue
        op = s.split()[0] if s else ""
        if op in FP:
            census[cur_line][FP[op]] += 1
        elif op == "call":
            name = s.split()[1].split("@")[0]
            if name == "sqrt":
                census[cur_line]["sqrt"] += 1
            elif name in LIBM:
                census[cur_line]["libm"] += 1
                census[cur_line]["libm:" + name] += 1
    return census


def line_counts():
    """{line: executions} from the coverage data of the last run."""
    subprocess.run(["gcov", "--json-format", "--object-directory", OUT_DIR, SRC], cwd=OUT_DIR, check=True,
                   stdout=subprocess.DEVNULL)
    with gzip.open(os.path.join(OUT_DIR, "hams_oracle.gcov.json.gz"), "rt") as f:
        data = json.load(f)
    out = {}
    for fobj in data["files"]:
        if fobj["file"].endswith("hams_oracle.c"):
            for ln in fobj["lines"]:
                out[ln["line_number"]] = out.get(ln["line_number"], 0) + ln["count"]
    return out


def classify(mesh, depth):
    """Region index (0..3) and FLAG of every entry (i field row, j source column)."""
    c, n = mesh.cen, mesh.nelem
    reg = np.empty((n, n), dtype=np.int8)
    flag = np.empty((n, n), dtype=bool)
    for i0 in range(0, n, 512):
        dx = c[None, :, 0] - c[i0:i0 + 512, None, 0]
        dy = c[None, :, 1] - c[i0:i0 + 512, None, 1]
        dz = c[None, :, 2] - c[i0:i0 + 512, None, 2]
        rhp = np.abs(np.sqrt(dx * dx + dy * dy) / depth)
        reg[i0:i0 + 512] = np.select([rhp < 5e-4, rhp < 0.05, rhp < 0.5], [0, 1, 2], 3)
        flag[i0:i0 + 512] = np.sqrt(dx * dx + dy * dy + dz * dz) <= 50.0 * mesh.pnsz[None, :]
    return reg, flag


def _child(args):
    """Evaluate the sampled entries (or one expression) with the coverage
    build, in a fresh process."""
    sys.path[:0] = [HERE, os.path.join(ROOT, "hams-mrel_amd")]
    import oracle
    if args.expr:
        cov = oracle._load(COV_SO)  # noqa: F841 -- used by the expression
        eval(args.expr)
        return
    from hamsx import configs
    from hamsx.mesh import combine
    case = configs.two_cylinders()
    m = combine(case.bodies, case.xr)
    pairs = np.load(args.pairs)
    cov = oracle._load(COV_SO)
    for w in args.freqs:
        v, wk, wvn = oracle.wave(w, case.depth)
        a, s = np.zeros(2), np.zeros(2)
        for i, j in pairs:
            cov.orc_assemble_entry(int(i), int(j), m.nelem, m.nnode, oracle._p(m.xyz), oracle._p(m.ncn, oracle.C.c_int),
                                   oracle._p(m.ncon, oracle.C.c_int), oracle._p(m.cen), oracle._p(m.area),
                                   oracle._p(m.pnsz), oracle._p(m.dxyz), v, case.depth, oracle._p(wvn), oracle.NK,
                                   oracle._p(a), oracle._p(s))


def measure(child_args, census=None):
    """Operation totals (Counter by kind) of one fresh child process."""
    census = census if census is not None else line_census(ASM)
    gcda = os.path.join(OUT_DIR, "hams_oracle.gcda")
    if os.path.exists(gcda):
        os.remove(gcda)
    subprocess.run([sys.executable, os.path.abspath(__file__), "--child", *child_args], check=True)
    tot = collections.Counter()
    for ln, k in line_counts().items():
        for kind, c in census.get(ln, {}).items():
            tot[kind] += k * c
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-class", type=int, default=300, help="sampled entries per (region, FLAG) class")
    ap.add_argument("--freqs", type=float, nargs="+", default=[0.2, 0.99, 1.78, 2.5])
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default=os.path.join(ROOT, "region_ops.json"))
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--pairs")
    ap.add_argument("--expr", help="child: a Python expression to evaluate with the coverage build as `cov`")
    args = ap.parse_args()
    if args.child:
        _child(args)
        return
    if not (os.path.exists(COV_SO) and os.path.exists(ASM)):
        raise SystemExit("run `make -C oracle opcount` first")
    sys.path[:0] = [HERE, os.path.join(ROOT, "hams-mrel_amd")]
    from hamsx import configs
    from hamsx.mesh import combine
    case = configs.two_cylinders()
    m = combine(case.bodies, case.xr)
    reg, flag = classify(m, case.depth)
    census = line_census(ASM)
    rng = np.random.default_rng(args.seed)
    classes = {}
    for r, name in enumerate(REGIONS):
        for f in (0, 1):
            ij = np.argwhere((reg == r) & (flag == bool(f)))
            if len(ij) == 0:
                continue
            pick = ij[rng.choice(len(ij), size=min(args.per_class, len(ij)), replace=False)]
            path = os.path.join(OUT_DIR, "pairs.npy")
            np.save(path, pick.astype(np.int32))
            tot = measure(["--pairs", path, "--freqs", *map(str, args.freqs)], census)
            nent = len(pick) * len(args.freqs)
            per = {k: v / nent for k, v in sorted(tot.items())}
            per["ops"] = per.get("add_sub_mul", 0) + per.get("div", 0) + per.get("sqrt", 0) + per.get("libm", 0)
            classes[f"{name}/FLAG{f}"] = {"workload_entries": int(len(ij)), "sampled_entries": int(len(pick)),
                                          "per_entry": per}
            print(f"{name} FLAG={f}: {len(ij)} entries, {per['ops']:.1f} ops per entry "
                  f"({per.get('div', 0):.1f} div, {per.get('libm', 0):.1f} libm)", flush=True)
    regions = {}
    for name in REGIONS:
        cl = [v for k, v in classes.items() if k.startswith(name + "/")]
        ne = sum(c["workload_entries"] for c in cl)
        if not ne:
            continue
        keys = set().union(*(c["per_entry"] for c in cl))
        regions[name] = {"entries": ne, "per_entry": {k: sum(c["per_entry"].get(k, 0.0) * c["workload_entries"]
                                                             for c in cl) / ne for k in sorted(keys)}}
    out = {
        "workload": case.name, "panels": m.nelem, "depth_m": case.depth, "frequencies_rad_s": args.freqs,
        "unit": "operations per entry (one A(i,j) and one S(i,j))",
        "convention": "add, subtract, multiply, divide, square root = 1 operation each; one call of exp, log, sin, "
                      "cos, sinh, cosh, pow, atan2 or cabs = 1 operation (also listed as libm); comparisons, fabs "
                      "and integer work not counted",
        "method": "oracle/opcount.py: hams_oracle.c at -O0 without FMA contraction, gcov line counts x per-line "
                  "FP64 instruction census, over a seeded sample of each (region, FLAG) class",
        "seed": args.seed, "classes": classes, "regions": regions,
        "per_frequency_ops": sum(r["entries"] * r["per_entry"]["ops"] for r in regions.values()),
    }
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(f"wrote {args.out}: {out['per_frequency_ops']:.3e} operations per frequency", flush=True)


if __name__ == "__main__":
    main()
