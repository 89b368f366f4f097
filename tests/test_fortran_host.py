"""The Fortran (iso_c_binding) host: hams-mrel_amd/fortran/hamsx_pnl_solve
reads a .pnl mesh and runs the reference driver's per-frequency sequence
through hamsx_mod's wrappers."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from hamsx.deck import read_pnl, write_pnl
from hamsx.mesh import combine, cylinder

DRIVER = os.path.join(ROOT, "hams-mrel_amd", "fortran", "hamsx_pnl_solve")


def write_case(tmp_path, depth, freqs, heads_deg, prec=0):
    pnl = tmp_path / "cyl.pnl"
    write_pnl(str(pnl), cylinder(4.0, 6.0, 20, 5, 3))
    inp = tmp_path / "driver.in"
    inp.write_text(f"'{pnl}'\n0.0 0.0 0.0\n{depth}\n{len(freqs)}\n{' '.join(map(str, freqs))}\n"
                   f"{len(heads_deg)}\n{' '.join(map(str, heads_deg))}\n{prec}\n")
    return pnl, inp


def parse(out, nbeta):
    toks = out.split()
    res, k = [], 0
    while k < len(toks):
        assert toks[k] == "W"
        w = float(toks[k + 1])
        k += 2
        assert toks[k] == "AMAS"
        am = np.array(toks[k + 1:k + 37], float).reshape(6, 6, order="F")
        k += 37
        assert toks[k] == "BDMP"
        bd = np.array(toks[k + 1:k + 37], float).reshape(6, 6, order="F")
        k += 37
        assert toks[k] == "EXFC"
        v = np.array(toks[k + 1:k + 1 + 12 * nbeta], float)
        ex = (v[0::2] + 1j * v[1::2]).reshape(6, nbeta, order="F")
        k += 1 + 12 * nbeta
        res.append((w, am, bd, ex))
    return res


def test_driver_built_and_reports_library_errors(tmp_path):
    """Without a GPU the driver stops with the library's own message."""
    assert os.access(DRIVER, os.X_OK), "build the Fortran host: make -C hams-mrel_amd"
    _, inp = write_case(tmp_path, 30.0, [1.0], [0.0])
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    p = subprocess.run([DRIVER, str(inp)], capture_output=True, text=True, timeout=120, env=env)
    if p.returncode == 0:
        pytest.skip("a GPU is visible despite HIP_VISIBLE_DEVICES=-1")
    assert "hamsx_create failed" in p.stdout + p.stderr


@pytest.mark.gpu
def test_fortran_host_equals_python_host(tmp_path):
    from hamsx.solver import FrequencySolver
    freqs, heads = [0.7, 1.6], [0.0, 60.0]
    for depth in (-1.0, 30.0):
        pnl, inp = write_case(tmp_path, depth, freqs, heads)
        print(f"[gpu] fortran host depth={depth}", flush=True)
        p = subprocess.run([DRIVER, str(inp)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        got = parse(p.stdout, len(heads))
        mesh = combine([read_pnl(str(pnl))], [np.zeros(3)])
        with FrequencySolver(mesh) as s:
            s.set_waves(np.deg2rad(heads))
            for (w, am, bd, ex), wq in zip(got, freqs):
                a2, b2, e2, info = s.solve_frequency(wq, depth)
                assert info == 0 and w == wq
                np.testing.assert_array_equal(am, a2)
                np.testing.assert_array_equal(bd, b2)
                np.testing.assert_array_equal(ex, e2)


@pytest.mark.gpu
@pytest.mark.parametrize("isx", [1, 0])
def test_fortran_host_half_mesh_with_symmetry(tmp_path, isx):
    """A half-mesh .pnl with ISX (or ISY) = 1 in its header: the Fortran host
    sets the symmetry plane and gives the Python host's numbers bit for bit."""
    from hamsx.deck import PanelMesh
    from hamsx.solver import FrequencySolver
    from test_oracle_symmetry import half
    freqs, heads = [0.9], [0.0, 30.0]
    h = half(cylinder(4.0, 6.0, 20, 5, 3), 0 if isx else 1)
    hm = PanelMesh(h.xyz, h.ncn, h.ncon, isx, 1 - isx)
    pnl = tmp_path / "half.pnl"
    write_pnl(str(pnl), hm)
    inp = tmp_path / "driver.in"
    inp.write_text(f"'{pnl}'\n0.0 0.0 0.0\n30.0\n1\n0.9\n2\n0.0 30.0\n0\n")
    print(f"[gpu] fortran host symmetry isx={isx}", flush=True)
    p = subprocess.run([DRIVER, str(inp)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    (w, am, bd, ex), = parse(p.stdout, len(heads))
    mesh = combine([read_pnl(str(pnl))], [np.zeros(3)])
    with FrequencySolver(mesh) as s:
        s.set_waves(np.deg2rad(heads))
        s.set_symmetry(isx, 1 - isx)
        a2, b2, e2, info = s.solve_frequency(freqs[0], 30.0)
    assert info == 0
    np.testing.assert_array_equal(am, a2)
    np.testing.assert_array_equal(bd, b2)
    np.testing.assert_array_equal(ex, e2)


@pytest.mark.gpu
@pytest.mark.parametrize("half", [False, True])
def test_reference_named_calls_equal_the_gpu_wrappers(tmp_path, half):
    """The loop body of HAMS_Prog.f90:629-724 under the reference's own names
    and argument lists (CALGREEN_MULTI, ASSB_LEFT_MULTI, ASSB_RBC_MULTI,
    RADIATION_SOLVER_MULTI, ASSB_DBC_MULTI, DIFFRACTION_SOLVER_MULTI; one
    heading per diffraction call) gives the *_GPU wrappers' potentials and
    coefficients bit for bit, with and without a symmetry plane."""
    pnl, inp = write_case(tmp_path, 30.0, [0.7, 1.6], [0.0, 60.0, 120.0])
    if half:  # the half y > 0 with ISY = 1 in the header: NSYS = 2
        from hamsx.deck import PanelMesh
        from test_oracle_symmetry import half as half_of
        h = half_of(cylinder(4.0, 6.0, 20, 5, 3), 1)
        write_pnl(str(pnl), PanelMesh(h.xyz, h.ncn, h.ncon, 0, 1))
    out = {}
    for style in (0, 1):
        text = inp.read_text().splitlines()[:8] + [str(style)]
        inp.write_text("\n".join(text) + "\n")
        print(f"[gpu] Fortran host style {style} half={half}", flush=True)
        p = subprocess.run([DRIVER, str(inp)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        toks = p.stdout.split()
        diffs = [float(toks[k + 1]) for k in range(len(toks)) if toks[k] == "MXPOT"]
        keep = [t for k, t in enumerate(toks) if t != "MXPOT" and (k == 0 or toks[k - 1] != "MXPOT")]
        out[style] = (parse(" ".join(keep), 3), diffs)
    assert out[1][1] == [0.0, 0.0]
    for a, b in zip(out[0][0], out[1][0]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
