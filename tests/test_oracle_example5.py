"""End-to-end pin of the oracle's irregular-frequency-removal path
(IRSP = 1: lid centroids as extra field points, C = A^T A, D = A^T B) against
the reference's shipped examples/example5_single_body outputs: one body of
234 panels with a 383-panel waterplane lid, depth 70 m, 10 frequencies
0.05-0.14 rad/s, 36 headings (tests/golden/example5.npz).

The shipped tables carry 5 (added mass, damping) and 6 (excitation) printed
digits and were written by the reference's default build (FP32 storage of
the Green-function values), which bounds how closely they pin an FP64
restatement -- as for test/data/config1."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from hamsx.deck import PanelMesh
from test_oracle_config1 import combined_with_oracle_geometry


@pytest.fixture(scope="module")
def example5():
    z = np.load(os.path.join(GOLDEN, "example5.npz"))
    hull = PanelMesh(z["hull_xyz"], z["hull_ncn"], z["hull_ncon"])
    lid = PanelMesh(z["lid_xyz"], z["lid_ncn"], z["lid_ncon"])
    return z, hull, lid


def tables(results, nbeta):
    am, bd, ex = {k: [] for k in range(6)}, {k: [] for k in range(6)}, {k: [] for k in range(6)}
    for amas, bdmp, exfc in results:
        for k in range(6):
            am[k].append(amas[k, :])
            bd[k].append(bdmp[k, :])
            for b in range(nbeta):
                ex[k].append([exfc[k, b].real, exfc[k, b].imag])
    return [{k: np.array(v) for k, v in t.items()} for t in (am, bd, ex)]


def rel_frob(got, ref):
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def example5_results(example5):
    """The oracle's IRSP = 1 solve at the ten frequencies: (amas, bdmp, exfc)."""
    z, hull, lid = example5
    m = combined_with_oracle_geometry([hull], [z["xr"]])
    lid_cen, _, _, _ = oracle.geometry(lid, z["xr"])
    heads = np.deg2rad(z["headings_deg"])
    res = []
    for w in z["omegas"]:
        amas, bdmp, exfc, info = oracle.solve_frequency_irr(m, lid_cen, float(w), float(z["depth"]), heads,
                                                            nthreads=8)
        assert info == 0
        res.append((amas, bdmp, exfc))
    return res


def test_oracle_irr_reproduces_example5(example5, example5_results):
    """Per table (Frobenius, relative): added mass <= 5e-5, excitation
    <= 1e-5, damping <= 2e-2.  The normal equations C = A^T A square the
    condition number, so the reference's FP32-stored Green values move its
    own printed result by about 1e-5 here (config1, solved directly, agrees
    to 5e-7); damping at 0.05-0.14 rad/s is five orders below added mass."""
    z, hull, lid = example5
    heads = np.deg2rad(z["headings_deg"])
    res = example5_results
    am, bd, ex = tables(res, len(heads))
    for name, tab, tol in (("OAMASS", am, 5e-5), ("ODAMPING", bd, 2e-2), ("OEXFOR", ex, 1e-5)):
        for k in range(6):
            assert rel_frob(tab[k], z[f"{name}{k + 1}"][:, 2:]) <= tol, (name, k + 1)


def test_example5_distinguishes_the_lid(example5):
    """The same solve without the lid misses the shipped added mass by about
    1.5%; with the lid it is at least 100 times closer."""
    z, hull, lid = example5
    m = combined_with_oracle_geometry([hull], [z["xr"]])
    lid_cen, _, _, _ = oracle.geometry(lid, z["xr"])
    w = float(z["omegas"][0])
    a_irr, _, _, _ = oracle.solve_frequency_irr(m, lid_cen, w, float(z["depth"]), [0.0], nthreads=8)
    a_plain, _, _, _ = oracle.solve_frequency(m, w, float(z["depth"]), [0.0], nthreads=8)
    ref = np.array([z[f"OAMASS{k + 1}"][0, 2:] for k in range(6)])
    e_irr, e_plain = rel_frob(a_irr, ref), rel_frob(a_plain, ref)
    assert e_plain > 5e-3 and e_irr * 100 < e_plain, (e_irr, e_plain)

