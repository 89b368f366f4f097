"""Continuous-batching server: every result is the same alone and in the
mix, admission respects the HBM budget, and slots keep meshes resident."""
import numpy as np
import pytest

from hamsx.mesh import box, combine, cylinder
from hamsx.serving import Request, Server, footprint_bytes
from hamsx.solver import FrequencySolver

pytestmark = pytest.mark.gpu


def meshes():
    a = combine([cylinder(4.0, 8.0, 24, 6, 3)], [np.zeros(3)])
    b = combine([cylinder(3.0, 6.0, 20, 5, 3, -12.0), box(8.0, 6.0, 4.0, 6, 4, 3, 12.0)],
                [np.array([-12.0, 0, 0]), np.array([12.0, 0, 0])])
    c = combine([box(10.0, 5.0, 5.0, 8, 4, 4)], [np.zeros(3)])
    return [a, b, c]


def requests():
    ms = meshes()
    spec = [(0, 0.8, 40.0, [0.0]), (1, 1.3, -1.0, [0.0, 45.0]), (2, 0.5, 25.0, [90.0]), (0, 1.9, 40.0, [30.0]),
            (1, 0.7, 60.0, [0.0, 120.0]), (2, 2.2, -1.0, [10.0]), (0, 1.1, 40.0, [0.0, 90.0, 180.0])]
    return [Request(k, ms[m], w, h, np.deg2rad(b), arrival=0.002 * k, mesh_key=m)
            for k, (m, w, h, b) in enumerate(spec)]


def alone(r, precision):
    with FrequencySolver(r.mesh, precision=precision) as s:
        s.set_waves(r.headings_rad)
        return s.solve_frequency(r.w1, r.depth)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_same_alone_and_in_mix(precision):
    reqs = requests()
    print(f"[gpu] serving {len(reqs)} requests, 3 slots, {precision}", flush=True)
    with Server(slots=3, precision=precision) as srv:
        res = srv.run(reqs)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    overlap = sum(1 for x in res for y in res if x.rid < y.rid and x.admitted < y.done and y.admitted < x.done)
    print(f"    overlapping pairs of requests: {overlap}", flush=True)
    assert overlap > 0
    for r in res:
        ref = alone(reqs[r.rid], precision)
        got = (r.amas, r.bdmp, r.exfc)
        for g, e in zip(got, ref[:3]):
            if precision == "fp64":
                np.testing.assert_array_equal(g, e)
            else:  # the refinement residual sums with atomics: equal to rounding
                assert np.max(np.abs(g - e)) <= 1e-12 * np.max(np.abs(e))
        assert r.info == 0


def test_admission_by_hbm_budget():
    reqs = requests()
    biggest = max(footprint_bytes(r.mesh.nelem, 6 * r.mesh.nbody + len(r.headings_rad), "fp64") for r in reqs)
    with Server(slots=3, precision="fp64", hbm_budget=biggest * 1.05) as srv:
        res = srv.run(reqs)
    assert len(res) == len(reqs)
    # only one request fits: one slot serves them all, one after another
    assert {r.slot for r in res} == {0}
    order = sorted(res, key=lambda r: r.admitted)
    for a, b in zip(order, order[1:]):
        assert a.done <= b.admitted
