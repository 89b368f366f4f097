"""The zero- and infinite-frequency limits (SYBO = 1, ImplementSubs.f90:45-70:
V = WK = W1 = 0 or -1, the Green functions' V = 0 / V = -1 branches,
InfGreen_Appr.f90:130-157, FinGreen3D.f90:100-150) in the oracle: in deep
water the solve at w -> 0 and w -> infinity tends to the limit solves.

In finite depth the reference's V = 0 branch (FinGreen3D.f90:100-123)
keeps only 1/R + 1/R1 -- the sea bed drops out -- so its zero-frequency
added mass is the deep-water one and the finite-depth solve does not tend to
it; the GPU tests hold that branch to the shipped example1 tables instead."""
import numpy as np
import pytest

import oracle
from hamsx.mesh import cylinder
from test_oracle_config1 import combined_with_oracle_geometry


@pytest.fixture(scope="module")
def body():
    return combined_with_oracle_geometry([cylinder(1.0, 2.0, 16, 6, 4)], [np.zeros(3)])


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_zero_frequency_limit_is_the_low_frequency_limit(body):
    a0, b0, e0, info = oracle.solve_frequency(body, 0.0, -1.0, [0.0])
    assert info == 0
    assert np.all(b0 == 0.0) and np.all(e0 == 0.0)     # W1 = 0 in every damping and excitation sum
    errs = [rel(oracle.solve_frequency(body, w, -1.0, [0.0])[0], a0) for w in (1e-2, 1e-3)]
    assert errs[1] < 1e-5 and errs[1] < 0.1 * errs[0], errs


def test_infinite_frequency_limit_is_the_high_frequency_limit(body):
    ai, bi, _, info = oracle.solve_frequency(body, -1.0, -1.0, [0.0])
    assert info == 0
    assert np.max(np.abs(bi)) == 0.0                    # real potentials: no damping
    errs = [rel(oracle.solve_frequency(body, w, -1.0, [0.0])[0], ai) for w in (30.0, 60.0)]
    assert errs[1] < 5e-3 and errs[1] < 0.5 * errs[0], errs


def test_limits_differ_from_each_other(body):
    a0 = oracle.solve_frequency(body, 0.0, -1.0, [0.0])[0]
    ai = oracle.solve_frequency(body, -1.0, -1.0, [0.0])[0]
    assert rel(ai, a0) > 0.1
