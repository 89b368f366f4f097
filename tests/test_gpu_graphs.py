"""Recorded launch sequences (HIP graphs) of hamsx_submit_frequency and the
completion queue.

A context with HAMSX_OPT_GRAPHS = 1 solves the first frequency at a depth
eagerly, records the second with stream capture and replays the recording
for every later frequency (the frequency reaches the kernels through device
scalars written with the dispersion roots).  Each replay must equal, bit for
bit, the same solve launched eagerly on a context with graphs off -- replays
interleaved with eager work of the same context (a blocking solve at another
depth, which re-sorts the pair bins and so retires the recording) and with
host synchronisations between them."""
import numpy as np
import pytest

from hamsx.mesh import combine, cylinder
from hamsx.solver import CompletionQueue, FrequencySolver

pytestmark = pytest.mark.gpu


def pair(na=24, nz=6, nr=3):
    b = [cylinder(4.0, 8.0, na, nz, nr, -15.0), cylinder(3.0, 6.0, na, nz, nr, 15.0)]
    return combine(b, [np.array([-15.0, 0, 0]), np.array([15.0, 0, 0])])


def solve(s, w, h):
    s.submit(w, h)
    return s.collect()


@pytest.mark.parametrize("precision,depth", [("fp64", 40.0), ("mixed", 40.0), ("fp64", -1.0), ("mixed", -1.0)])
def test_graph_replays_equal_eager_launches(precision, depth):
    m = pair()
    print(f"[gpu] graph replays n={m.nelem} {precision} depth={depth}", flush=True)
    heads = np.deg2rad([0.0, 60.0])
    with FrequencySolver(m, precision=precision) as g, FrequencySolver(m, precision=precision) as e:
        assert e.options()["graphs"] == 0  # the default
        g.set_options(graphs=1)
        assert g.options()["graphs"] == 1
        for s in (g, e):
            s.set_waves(heads)
        seq = [0.5, 0.9, 1.3, 0.5, 2.0]
        for k, w in enumerate(seq):
            rg, re_ = solve(g, w, depth), solve(e, w, depth)
            for a, b in zip(rg, re_):
                np.testing.assert_array_equal(a, b)
        assert g.graph_stats() == (1, len(seq) - 1)  # recorded on the second solve, launched from then on
        assert e.graph_stats() == (0, 0)
        # eager work on the recording context: a blocking solve at another depth
        other = 25.0 if depth > 0 else 30.0
        np.testing.assert_array_equal(g.solve_frequency(1.1, other)[0], e.solve_frequency(1.1, other)[0])
        # back at the first depth: the pair bins were re-sorted, so the old
        # recording is not replayed; eager, recorded again, replayed
        for w in (0.7, 1.6, 2.3, 0.7):
            rg, re_ = solve(g, w, depth), solve(e, w, depth)
            for a, b in zip(rg, re_):
                np.testing.assert_array_equal(a, b)
        # the other depth's solve re-sorted the bins (finite depth) and allocated
        # its mode tables, so the first depth starts over: one eager solve, a
        # new recording, replays
        assert g.graph_stats() == (2, len(seq) - 1 + 3)


def test_graph_frequency_limits():
    """The zero- and infinite-frequency limits take their own branch on the
    host (no mode tables): they never replay an ordinary frequency's
    recording, and their own replays equal the eager solves."""
    m = pair()
    with FrequencySolver(m) as g, FrequencySolver(m) as e:
        g.set_options(graphs=1)
        for s in (g, e):
            s.set_waves(np.deg2rad([0.0]))
        for w in (0.8, 1.2, 0.0, -1.0, 0.0, 1.5, -1.0, 0.0):
            for a, b in zip(solve(g, w, 35.0), solve(e, w, 35.0)):
                np.testing.assert_array_equal(a, b)


def test_queue_posts_each_finished_solve():
    meshes = [pair(20, 5, 3), pair(24, 6, 3), pair(16, 4, 2)]
    with CompletionQueue() as q:
        solvers = [FrequencySolver(m) for m in meshes]
        try:
            for k, s in enumerate(solvers):
                s.set_options(graphs=k % 2)  # notices after a replayed and after an eager solve
                s.set_waves(np.deg2rad([0.0, 90.0]))
                s.set_queue(q, 10 + k)
            for rnd in range(3):
                for s in solvers:
                    s.submit(0.6 + 0.3 * rnd, 40.0)
                seen = []
                for _ in solvers:
                    tag = q.wait(60.0)
                    assert tag is not None
                    seen.append(tag)
                    res = solvers[tag - 10].collect()
                    with FrequencySolver(meshes[tag - 10]) as ref:
                        ref.set_waves(np.deg2rad([0.0, 90.0]))
                        want = ref.solve_frequency(0.6 + 0.3 * rnd, 40.0)
                    for a, b in zip(res, want):
                        np.testing.assert_array_equal(a, b)
                assert sorted(seen) == [10, 11, 12]
                assert q.wait(0.05) is None
        finally:
            for s in solvers:
                s.set_queue(None, 0)
                s.close()
