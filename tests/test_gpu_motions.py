"""GPU motion solve (hamsx_set_body_matrices / hamsx_motions:
SolveMotionMulti, src/SolveMotionMulti.f90:46-150) against the oracle on the
GPU's own coefficients, and end to end -- assembly, LU, solves, forces,
motions -- against the shipped Motion.4 of examples 1 and 2."""
import sys

import numpy as np
import pytest

import oracle
from hamsx.deck import PanelMesh, place
from hamsx.mesh import combine, cylinder
from hamsx.solver import FrequencySolver
from test_oracle_fields import load, motion_table

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def buoy_matrices(nb, rng, quad=True):
    """Plausible body matrices: diagonal mass and inertia, heave/roll/pitch
    restoring, small linear damping and (optionally) quadratic damping, with a
    few off-diagonal couplings."""
    mats = {k: [] for k in ("mass", "blnr", "bqdr", "crs", "kstf")}
    for _ in range(nb):
        m = np.diag([4.0e5, 4.0e5, 4.0e5, 6.0e6, 6.0e6, 5.0e6]) * rng.uniform(0.8, 1.2)
        m[0, 4] = m[4, 0] = -2.0e5
        m[1, 3] = m[3, 1] = 2.0e5
        c = np.zeros((6, 6))
        c[2, 2], c[3, 3], c[4, 4] = 7.7e5, 3.0e6, 3.0e6
        mats["mass"].append(m)
        mats["crs"].append(c)
        mats["kstf"].append(np.diag([2.0e4, 2.0e4, 0.0, 0.0, 0.0, 1.0e5]))
        mats["blnr"].append(np.diag([1.0e3, 1.0e3, 5.0e3, 0.0, 0.0, 0.0]))
        mats["bqdr"].append(np.diag([2.0e4, 2.0e4, 4.0e4, 1.0e5, 1.0e5, 0.0]) if quad else np.zeros((6, 6)))
    return mats


@pytest.mark.parametrize("depth,w,quad", [(-1.0, 0.7, True), (30.0, 1.1, True), (30.0, 0.5, False)])
def test_motions_match_oracle(depth, w, quad):
    """The device solve and fixed point against the oracle's on the GPU's
    coefficients.  Both stop at RERR <= 1e-6; taking the same number of steps
    they agree to rounding (1e-10 relative)."""
    bodies = [cylinder(5.0, 8.0, 24, 6, 3, -12.0), cylinder(4.0, 6.0, 20, 6, 3, 12.0)]
    m = combine(bodies, [np.array([-12.0, 0, -2.0]), np.array([12.0, 0, -2.0])])
    mats = buoy_matrices(2, np.random.default_rng(3), quad)
    heads = np.deg2rad([0.0, 60.0, 135.0])
    say(f"[gpu] motions n={m.nelem} depth={depth} w={w} quadratic={quad}")
    with FrequencySolver(m) as s:
        s.set_waves(heads)
        s.set_body_matrices(*(mats[k] for k in ("mass", "blnr", "bqdr", "crs", "kstf")))
        amas, bdmp, exfc, info = s.solve_frequency(w, depth)
        assert info == 0
        d, its = s.motions()
    ref, its_ref = oracle.motion(2, w, amas, bdmp, exfc, *(mats[k] for k in ("mass", "blnr", "bqdr", "crs", "kstf")))
    say(f"    steps gpu {list(its)} oracle {its_ref}   rel err {np.max(np.abs(d - ref)) / np.max(np.abs(ref)):.3e}")
    assert list(its) == list(its_ref)
    if not quad:
        assert all(i == 0 for i in its)
    assert np.max(np.abs(d - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_motions_need_body_matrices():
    from hamsx._lib import HamsxError
    m = combine([cylinder(5.0, 8.0, 16, 4, 2)], [np.zeros(3)])
    with FrequencySolver(m) as s:
        s.set_waves([0.0])
        s.calwaveproperts(1.0, -1.0)
        with pytest.raises(HamsxError):
            s.motions()


@pytest.fixture(scope="module")
def example1():
    z = load("example1.npz")
    local = PanelMesh(np.asfortranarray(z["xyz"]), z["ncn"], np.asfortranarray(z["ncon"]))
    bodies = [place(local, z["lcs"][b]) for b in range(len(z["xr"]))]
    return z, combine(bodies, list(z["xr"]))


def test_example1_motions_on_gpu(example1):
    """Five buoys at depth 0.9 m, 80 frequencies 0.2-14.8 rad/s, heading
    90 deg: the whole GPU path against the shipped Motion.4 (30 DOF per
    frequency).  Relative Frobenius over all frequencies, per mode."""
    z, m = example1
    nb = len(z["xr"])
    hy = z["hydro"]
    say(f"[gpu] example1 motions n={m.nelem} frequencies={len(z['omegas'])}")
    got = []
    with FrequencySolver(m) as s:
        s.set_waves(np.deg2rad(z["headings_deg"]), isol=int(z["isol"]))
        s.set_body_matrices(*[[hy[b, q] for b in range(nb)] for q in range(5)])
        for w in z["omegas"]:
            _, _, _, info = s.solve_frequency(float(w), float(z["depth"]))
            assert info == 0
            d, _ = s.motions()
            got.append(motion_table(d, refl=float(z["refl"])))
    got = np.array(got)[:, 0, :]                     # frequency x DOF
    ref = z["motion"][:, 0, :, 2] + 1j * z["motion"][:, 0, :, 3]
    errs = []
    for md in range(6 * nb):
        nrm = np.linalg.norm(ref[:, md])
        errs.append(np.linalg.norm(got[:, md] - ref[:, md]) / nrm if nrm > 1e-12 else 0.0)
    say(f"    per-DOF rel err: max {max(errs):.3e} median {np.median(errs):.3e}")
    say("    " + " ".join(f"{e:.1e}" for e in errs))
    tot = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"    overall {tot:.3e}")
    assert tot <= 1e-4


@pytest.mark.parametrize("k", [0, 1])
def test_example2_motions_on_gpu(k):
    """Ten bodies with quadratic damping, both periods, 12 headings."""
    z = load("example2.npz")
    mz = load("motion.npz")
    bodies = [PanelMesh(np.asfortranarray(z["xyz"][b]), z["ncn"], np.asfortranarray(z["ncon"]))
              for b in range(len(z["xr"]))]
    m = combine(bodies, list(z["xr"]))
    nb, hy = len(z["xr"]), mz["ex2_hydro"]
    w = 2 * np.pi / z["periods"][k]
    say(f"[gpu] example2 motions period {z['periods'][k]}")
    with FrequencySolver(m) as s:
        s.set_waves(np.deg2rad(z["headings_deg"]), isol=int(z["isol"]))
        s.set_body_matrices(*[[hy[b, q] for b in range(nb)] for q in range(5)])
        _, _, _, info = s.solve_frequency(w, float(z["depth"]))
        assert info == 0
        d, its = s.motions()
    got = motion_table(d, refl=float(z["refl"]))
    ref = mz["ex2_motion"][k][..., 2] + 1j * mz["ex2_motion"][k][..., 3]
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"    steps {list(its)}  rel err {e:.3e}")
    assert all(i > 0 for i in its)
    assert e <= 5e-6


@pytest.mark.parametrize("depth", [-1.0, 12.0])
@pytest.mark.parametrize("w", [0.0, -1.0])
def test_frequency_limits_match_oracle(depth, w):
    """The zero- (w = 0) and infinite- (w = -1) frequency limits: only the
    Rankine and image terms remain, so the GPU matches the oracle to
    rounding.  At the infinite limit in finite depth the reference's
    diffraction is NaN (DINP's cosh/sinh overflow at WK = 1e20, as example1's
    shipped OEXFOR rows show), and so is ours."""
    bodies = [cylinder(5.0, 8.0, 24, 6, 3, -12.0), cylinder(4.0, 6.0, 20, 6, 3, 12.0)]
    m = combine(bodies, [np.array([-12.0, 0, -2.0]), np.array([12.0, 0, -2.0])])
    say(f"[gpu] limit w={w} depth={depth}")
    with FrequencySolver(m) as s:
        s.set_waves([0.0])
        amas, bdmp, exfc, info = s.solve_frequency(w, depth)
    assert info == 0
    ao, bo, eo, _ = oracle.solve_frequency(m, w, depth, [0.0])
    assert np.max(np.abs(amas - ao)) <= 1e-12 * np.max(np.abs(ao))
    assert np.max(np.abs(bdmp)) == 0.0 and np.max(np.abs(bo)) == 0.0
    if w == -1.0 and depth > 0:
        assert np.all(np.isnan(exfc)) and np.all(np.isnan(eo))
    else:
        assert np.max(np.abs(exfc)) == 0.0 and np.max(np.abs(eo)) == 0.0


def test_example1_frequency_limits_on_gpu(example1):
    """0_inf_frequency_limits = 1 in example1: the shipped OAMASS rows of the
    two limits (25 body pairs each, 5 printed digits) against the GPU solve
    at w = 0 and w = -1 (depth 0.9 m: the reference's V = 0 branch has no
    sea bed).  Relative Frobenius per table <= 1e-4."""
    z, m = example1
    nb = len(z["xr"])
    say(f"[gpu] example1 limits n={m.nelem}")
    with FrequencySolver(m) as s:
        s.set_waves(np.deg2rad(z["headings_deg"]), isol=int(z["isol"]))
        res = [s.solve_frequency(w, float(z["depth"])) for w in (0.0, -1.0)]
    for k, (amas, bdmp, exfc, info) in enumerate(res):
        assert info == 0
        for md in range(6):
            ref = z["oamass_limits"][md][k * nb * nb:(k + 1) * nb * nb, 2:]
            got = np.array([amas[6 * M + md, 6 * N:6 * N + 6] for M in range(nb) for N in range(nb)])
            e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            say(f"    limit {k} OAMASS{md + 1}: rel err {e:.2e}")
            assert e <= 1e-4
