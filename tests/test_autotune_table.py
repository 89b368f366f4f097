"""The autotuner's table lookup (CPU): nearest bucket in log scale within the
same depth class and precision; no table, no options."""
from hamsx import autotune


def table():
    e = lambda n, d, p, rows: {"panels": n, "depth": d, "precision": p, "best": {"lu_select_rows": rows}}
    return {"entries": [e(1000, "deep", "fp64", 256), e(4000, "deep", "fp64", 512), e(12000, "deep", "fp64", 256),
                        e(4000, "finite", "fp64", 1024), e(4000, "deep", "mixed", 256)]}


def test_options_for_picks_the_nearest_bucket():
    t = table()
    assert autotune.options_for(t, 1900, -1.0, "fp64") == {"lu_select_rows": 256}   # log-nearer to 1,000
    assert autotune.options_for(t, 2100, -1.0, "fp64") == {"lu_select_rows": 512}   # log-nearer to 4,000
    assert autotune.options_for(t, 9000, -1.0, "fp64") == {"lu_select_rows": 256}
    assert autotune.options_for(t, 500, 60.0, "fp64") == {"lu_select_rows": 1024}   # finite depth
    assert autotune.options_for(t, 4000, -1.0, "mixed") == {"lu_select_rows": 256}
    assert autotune.options_for(t, 4000, 60.0, "mixed") == {}
    assert autotune.options_for({}, 4000, -1.0, "fp64") == {}


def test_committed_table_is_well_formed():
    t = autotune.load_table()
    if not t:
        return
    for e in t["entries"]:
        assert e["depth"] in ("deep", "finite") and e["precision"] in ("fp64", "mixed")
        assert e["best"] in [tr["options"] for tr in e["trials"]]
        assert e["solves_per_s"] == max(tr["solves_per_s"] for tr in e["trials"])
        assert e["best"]["lu_select_rows"] in (256, 512, 1024)
