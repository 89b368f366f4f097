"""The per-region operation counts (region_ops.json, written by
oracle/opcount.py from the C restatement) and the counting method itself.

The method -- gcov line executions x a per-line census of the FP64
instructions of the -O0 build -- is pinned on ENXA(4, 2.0)
(FinGrnExtSubs.f90:135-157 with E1XA, :198-213, on its x > 1 branch), whose
operations are counted by hand from the routine: E0 = exp(-x)/x (1 libm,
1 div); E1XA's two quartic polynomials (4 adds and 3 multiplies each) and
exp(-x)/x*es1/es2 (1 libm, 2 div, 1 multiply); three recurrence steps
(exp(-x) - x*e1)/(k - 1) (1 libm, 1 multiply, 2 subtractions, 1 div each).
A census that counted the untaken x <= 1 branch (11 more operations), missed
the loop's repetitions or a call, fails it.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import opcount  # noqa: E402


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "opcount"], check=True)


def test_census_counts_enxa_by_hand(built):
    tot = opcount.measure(["--expr", "cov.orc_enxa(4, 2.0)"])
    assert tot["add_sub_mul"] == 15 + 3 * 3
    assert tot["div"] == 1 + 2 + 3
    assert tot["libm"] == 1 + 1 + 3
    assert tot["libm:exp"] == 5
    assert tot.get("sqrt", 0) == 0


def test_region_ops_cover_the_workload():
    with open(os.path.join(ROOT, "region_ops.json")) as f:
        d = json.load(f)
    sys.path.insert(0, ROOT)
    import bench
    from hamsx import configs
    from hamsx.mesh import combine
    case = configs.two_cylinders()
    m = combine(case.bodies, case.xr)
    assert d["workload"] == case.name and d["panels"] == m.nelem
    pairs = bench.region_pairs(m, case.depth)
    for key, name in zip(bench.REGIONS, opcount.REGIONS):
        # the counter classifies entries with the kernels' own region test
        assert d["regions"][name]["entries"] == pairs[key]
        per = d["regions"][name]["per_entry"]
        assert per["ops"] == pytest.approx(per["add_sub_mul"] + per["div"] + per.get("sqrt", 0) + per["libm"])
    ops = [d["regions"][r]["per_entry"]["ops"] for r in opcount.REGIONS]
    # EIGEN (at most 10 terms, no extrapolation) is the cheapest region by far,
    # and the epsilon-extrapolated regions carry most of the divisions
    assert ops[3] < 0.1 * min(ops[:3])
    for r in ("PIDCOCK", "EIGENE"):
        per = d["regions"][r]["per_entry"]
        assert per["div"] > 0.25 * per["ops"]
    assert d["per_frequency_ops"] == pytest.approx(sum(d["regions"][r]["entries"] * d["regions"][r]["per_entry"]["ops"]
                                                       for r in opcount.REGIONS))
    assert np.isfinite(d["per_frequency_ops"])
