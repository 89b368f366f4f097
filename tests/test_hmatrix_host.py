"""H-matrix partition and ACA checks that need no GPU.

The library's partition (cluster tree by geometric bisection, eta-admissible
blocks; computed on the host) must equal the oracle's restatement of
HMatrix.f90:232-483 exactly, and cover the matrix once.  The oracle's ACA is
held to mathematics: an exactly rank-3 block is reproduced to machine
precision after three steps (a fourth step, if taken, only carries the
rounding remainder), and a smooth kernel between separated clusters
is approximated to its tolerance."""
import numpy as np
import pytest

import hmatrix as ohm
from hamsx.mesh import combine, cylinder
from hamsx.solver import hmatrix_partition


def meshes():
    a = combine([cylinder(4.0, 8.0, 24, 6, 3, -15.0), cylinder(3.0, 6.0, 24, 6, 3, 15.0)],
                [np.array([-15.0, 0, 0]), np.array([15.0, 0, 0])])
    b = combine([cylinder(3.0, 6.0, 20, 5, 3, x, y) for x in (-30.0, 0.0, 30.0) for y in (-30.0, 30.0)],
                [np.array([x, y, 0.0]) for x in (-30.0, 0.0, 30.0) for y in (-30.0, 30.0)])
    return [a, b]


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("leaf,eta", [(32, 2.0), (16, 1.0)])
def test_partition_equals_oracle_and_covers_once(k, leaf, eta):
    m = meshes()[k]
    perm, blocks = hmatrix_partition(m.cen, leaf, eta)
    operm, oblocks = ohm.partition(m.cen, leaf, eta)
    np.testing.assert_array_equal(perm, operm)
    assert [tuple(int(v) for v in b) for b in blocks] == [tuple(b) for b in oblocks]
    n = m.nelem
    assert sorted(perm.tolist()) == list(range(n))
    cover = np.zeros((n, n), dtype=np.int32)
    for rs, nr, cs, nc, _ in blocks:
        cover[rs:rs + nr, cs:cs + nc] += 1
    assert np.all(cover == 1)
    assert np.any(blocks[:, 4] == 1)  # some admissible blocks


def test_oracle_aca_exact_low_rank_and_smooth_kernel():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((60, 3)) + 1j * rng.standard_normal((60, 3))
    Y = rng.standard_normal((3, 45)) + 1j * rng.standard_normal((3, 45))
    A3 = X @ Y
    U, VH = ohm.aca(A3, rank_max=50, tol=1e-10)
    # three exact steps; the fourth picks up the rounding remainder and stops
    assert U.shape[1] in (3, 4)
    assert np.linalg.norm(U[:, 3:] @ VH[3:, :]) <= 1e-12 * np.linalg.norm(A3)
    assert np.linalg.norm(A3 - U @ VH) <= 1e-12 * np.linalg.norm(A3)
    # 1/|x - y| between two separated point clouds
    p = rng.uniform(-1, 1, (80, 3))
    q = rng.uniform(-1, 1, (70, 3)) + np.array([8.0, 0.0, 0.0])
    K = 1.0 / np.linalg.norm(p[:, None, :] - q[None, :, :], axis=2) + 0j
    for tol in (1e-4, 1e-8):
        U, VH = ohm.aca(K, rank_max=50, tol=tol)
        assert U.shape[1] < 30
        assert np.linalg.norm(K - U @ VH) <= 20 * tol * np.linalg.norm(K)
