"""Host-side mirror of HAMS-MREL's per-frequency solve on top of libhamsx.

The method names follow the reference's driver sequence
(``src/HAMS_Prog.f90:613-733``): ``calwaveproperts`` -> ``assemble``
(CALGREEN_MULTI + ASSB_LEFT_MULTI) -> ``factor`` (ZGETRF) ->
``radiation_solver`` / ``diffraction_solver`` (ASSB_RBC/DBC_MULTI + ZGETRS)
-> ``forces`` (RFORCE_MULTI / EFORCE_MULTI).  Errors raise
:class:`hamsx._lib.HamsxError`; a zero pivot is reported as INFO > 0 the way
ZGETRF reports it (the reference ignores it; here it is returned).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Combined


def wave_properties(inft: int, value: float, h: float):
    """CalWaveProperts (``src/ImplementSubs.f90:40-151``): returns v, wk, w1, wvn[200]."""
    lib = _lib.load()
    v, wk, w1 = C.c_double(), C.c_double(), C.c_double()
    wvn = np.zeros(_lib.HAMSX_NK)
    _lib.check(lib.hamsx_wave_properties(inft, value, h, C.byref(v), C.byref(wk), C.byref(w1), _lib.ptr(wvn)),
               "hamsx_wave_properties")
    return v.value, wk.value, w1.value, wvn


class FrequencySolver:
    """One device context holding a mesh resident in HBM."""

    def __init__(self, mesh: Combined, device: int = 0, precision: str = "fp64"):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.hamsx_create(C.byref(h), device), "hamsx_create")
        self._h = h
        self.set_precision(precision)
        self.nbeta = 0
        self.nfp = 0
        self.nsys = 1
        self.set_mesh(mesh)

    def set_mesh(self, mesh: Combined):
        """Upload a (new) mesh; the context keeps it resident until replaced."""
        m = mesh
        _lib.check(self._lib.hamsx_set_mesh(
            self._h, m.nelem, m.nnode, _lib.ptr(m.xyz), _lib.ptr(m.ncn, C.c_int), _lib.ptr(m.ncon, C.c_int),
            _lib.ptr(m.cen), _lib.ptr(m.area), _lib.ptr(m.pnsz), _lib.ptr(m.dxyz), m.nbody,
            _lib.ptr(m.body_start, C.c_int)), "hamsx_set_mesh")
        self.mesh = mesh
        self.nfp = 0
        self.nsys = 1

    # -- lifecycle -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.hamsx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration ---------------------------------------------------------
    def set_precision(self, precision: str):
        mode = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        _lib.check(self._lib.hamsx_set_precision(self._h, mode), "hamsx_set_precision")
        self.precision = precision

    def set_options(self, **opts):
        """LU tuning options by name (lu_lookahead, lu_select_rows,
        lu_outer_min_n), taking effect at the next factorisation."""
        for name, value in opts.items():
            _lib.check(self._lib.hamsx_set_option(self._h, _lib.OPTIONS[name], int(value)), "hamsx_set_option")

    def options(self) -> dict:
        out = {}
        for name, code in _lib.OPTIONS.items():
            v = C.c_int()
            _lib.check(self._lib.hamsx_get_option(self._h, code, C.byref(v)), "hamsx_get_option")
            out[name] = v.value
        return out

    def set_lid(self, lid_cen):
        """Irregular-frequency removal: lid_cen (L x 3) are the lid (waterplane)
        panel centroids; None or an empty array removes the lid."""
        if lid_cen is None or len(lid_cen) == 0:
            _lib.check(self._lib.hamsx_set_lid(self._h, 0, None), "hamsx_set_lid")
            return
        lc = np.asfortranarray(lid_cen, dtype=np.float64)
        _lib.check(self._lib.hamsx_set_lid(self._h, lc.shape[0], _lib.ptr(lc)), "hamsx_set_lid")

    def set_symmetry(self, isx: int = 0, isy: int = 0):
        """Geometric symmetry (ISX / ISY of ControlFile.in): the mesh is the
        half on the positive side of x = 0 (isx) or y = 0 (isy); solves then
        cover the whole body.  Both 0 removes the plane."""
        _lib.check(self._lib.hamsx_set_symmetry(self._h, int(isx), int(isy)), "hamsx_set_symmetry")
        self.nsys = 2 if (isx or isy) else 1

    def set_field_points(self, xfp):
        """Field points (P x 3) for pressure / elevation output (the XFP array
        of ControlFile.in); None or an empty array removes them."""
        if xfp is None or len(xfp) == 0:
            _lib.check(self._lib.hamsx_set_field_points(self._h, 0, None), "hamsx_set_field_points")
            self.nfp = 0
            return
        xf = np.asfortranarray(xfp, dtype=np.float64).reshape(-1, 3)
        _lib.check(self._lib.hamsx_set_field_points(self._h, xf.shape[0], _lib.ptr(xf)), "hamsx_set_field_points")
        self.nfp = xf.shape[0]

    def field_values(self, kind: str = "auto", incident: bool = False):
        """Values at the field points for every mode of the current frequency
        (CalPotentialMulti / CalPressureMulti / CalElevationMulti): kind
        'potential', 'pressure', 'elevation' or 'auto' (pressure below the
        free surface, elevation on it).  Returns (values P x (6 nbody +
        nbeta), onbody P bool) and, with incident=True, the incident-wave
        values P x nbeta as a third item."""
        code = {"potential": _lib.FIELD_POTENTIAL, "pressure": _lib.FIELD_PRESSURE,
                "elevation": _lib.FIELD_ELEVATION, "auto": _lib.FIELD_AUTO}[kind]
        nfp, nr = self.nfp, 6 * self.mesh.nbody + self.nbeta
        vals = np.zeros((nfp, nr), dtype=np.complex128, order="F")
        inc = np.zeros((nfp, max(1, self.nbeta)), dtype=np.complex128, order="F")
        ob = np.zeros(nfp, dtype=np.int32)
        _lib.check(self._lib.hamsx_field_values(self._h, code, _lib.ptr(vals),
                                                _lib.ptr(inc) if incident else None, _lib.ptr(ob, C.c_int)),
                   "hamsx_field_values")
        if incident:
            return vals, ob.astype(bool), inc[:, :self.nbeta]
        return vals, ob.astype(bool)

    def set_body_matrices(self, mass, blnr, bqdr, crs, kstf):
        """The motion equation's body matrices (ReadHydroStaticMulti): each
        argument is a list of nbody 6 x 6 arrays (rows as in Hydrostatic_*.in)."""
        def blocks(ms):
            if len(ms) != self.mesh.nbody:
                raise ValueError(f"need {self.mesh.nbody} 6 x 6 matrices")
            return np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=np.float64).T.ravel() for m in ms]))
        arrs = [blocks(m) for m in (mass, blnr, bqdr, crs, kstf)]
        _lib.check(self._lib.hamsx_set_body_matrices(self._h, *[_lib.ptr(a) for a in arrs]), "hamsx_set_body_matrices")

    def set_hydrostatics(self, hydro):
        """set_body_matrices from hamsx.deck.Hydrostatic records, one per body."""
        self.set_body_matrices([h.mass for h in hydro], [h.blnr for h in hydro], [h.bqdr for h in hydro],
                               [h.crs for h in hydro], [h.kstf for h in hydro])

    def motions(self):
        """SolveMotionMulti for every heading of the current frequency:
        (DSPL 6 nbody x nbeta complex, quadratic-damping steps per heading)."""
        nd = 6 * self.mesh.nbody
        d = np.zeros((nd, max(1, self.nbeta)), dtype=np.complex128, order="F")
        it = np.zeros(max(1, self.nbeta), dtype=np.int32)
        _lib.check(self._lib.hamsx_motions(self._h, _lib.ptr(d), _lib.ptr(it, C.c_int)), "hamsx_motions")
        return d[:, :self.nbeta], it[:self.nbeta]

    def set_solver(self, solver: str = "lu", tol: float = 1.0e-4, restart: int = 25, maxiter: int = 100):
        """'lu' (ZGETRF/ZGETRS) or 'gmres' (restarted GMRES with the per-body
        block-diagonal LU preconditioner; defaults are the reference's
        IterativeSolvers.f90:56-58)."""
        code = {"lu": _lib.SOLVER_LU, "gmres": _lib.SOLVER_GMRES}[solver]
        _lib.check(self._lib.hamsx_set_solver(self._h, code, tol, restart, maxiter), "hamsx_set_solver")
        self.solver = solver

    def set_waves(self, headings_rad, amp=1.0, xw=(0.0, 0.0), rho=1000.0, isol=1):
        b = np.ascontiguousarray(headings_rad, dtype=np.float64)
        self.nbeta = len(b)
        _lib.check(self._lib.hamsx_set_waves(self._h, len(b), _lib.ptr(b), amp, xw[0], xw[1], rho, isol),
                   "hamsx_set_waves")

    # -- the reference's per-frequency sequence --------------------------------
    def calwaveproperts(self, w1: float, h: float, inft: int = 3):
        v, wk, w, wvn = wave_properties(inft, w1, h)
        _lib.check(self._lib.hamsx_set_frequency(self._h, v, wk, w, h, _lib.ptr(wvn), len(wvn)),
                   "hamsx_set_frequency")
        return v, wk, w, wvn

    def assemble(self):
        _lib.check(self._lib.hamsx_assemble(self._h), "hamsx_assemble")

    def factor(self) -> int:
        return _lib.check(self._lib.hamsx_factor(self._h), "hamsx_factor", allow_info=True)

    def radiation_solver(self) -> np.ndarray:
        """MXPOT of the radiation modes: n x 6 nbody, or n x 6 nbody x 2 (the
        mesh half, then the mirror half) with a symmetry plane."""
        n, nd = self.mesh.nelem, 6 * self.mesh.nbody
        out = np.zeros((n, nd, self.nsys), dtype=np.complex128, order="F")
        _lib.check(self._lib.hamsx_solve_radiation(self._h, _lib.ptr(out)),
                   "hamsx_solve_radiation")
        return out[:, :, 0] if self.nsys == 1 else out

    def diffraction_solver(self) -> np.ndarray:
        """MXPOT of the headings: n x nbeta (x 2 with a symmetry plane)."""
        n = self.mesh.nelem
        out = np.zeros((n, max(1, self.nbeta), self.nsys), dtype=np.complex128, order="F")
        _lib.check(self._lib.hamsx_solve_diffraction(self._h, _lib.ptr(out)),
                   "hamsx_solve_diffraction")
        out = out[:, :self.nbeta]
        return out[:, :, 0] if self.nsys == 1 else out

    def forces(self):
        nd = 6 * self.mesh.nbody
        amas = np.zeros((nd, nd), order="F")
        bdmp = np.zeros((nd, nd), order="F")
        exfc = np.zeros((nd, max(1, self.nbeta)), dtype=np.complex128, order="F")
        info = _lib.check(self._lib.hamsx_forces(self._h, _lib.ptr(amas), _lib.ptr(bdmp),
                                                 _lib.ptr(exfc)), "hamsx_forces", allow_info=True)
        return amas, bdmp, exfc[:, :self.nbeta], info

    def solve_frequency(self, w1: float, h: float):
        """One whole frequency solve (one device synchronisation)."""
        nd = 6 * self.mesh.nbody
        amas = np.zeros((nd, nd), order="F")
        bdmp = np.zeros((nd, nd), order="F")
        exfc = np.zeros((nd, max(1, self.nbeta)), dtype=np.complex128, order="F")
        info = _lib.check(self._lib.hamsx_solve_frequency(self._h, w1, h, _lib.ptr(amas), _lib.ptr(bdmp),
                                                          _lib.ptr(exfc)),
                          "hamsx_solve_frequency", allow_info=True)
        return amas, bdmp, exfc[:, :self.nbeta], info

    # -- non-blocking form (several contexts in flight) ------------------------
    def submit(self, w1: float, h: float):
        """Enqueue one whole frequency solve on this context's stream."""
        _lib.check(self._lib.hamsx_submit_frequency(self._h, w1, h), "hamsx_submit_frequency")

    @staticmethod
    def submit_group(solvers, w1s, h: float):
        """Enqueue one frequency solve on each of `solvers` (at most sixteen
        contexts of one mesh size and precision, dense LU path), w1s[i] on
        solvers[i]: each assembles on its own stream, the group's matrices are
        factored by one batched launch sequence, and each solves on its own
        stream.  Collect each solver as after submit."""
        if not solvers:
            return
        lib = solvers[0]._lib
        arr = (C.c_void_p * len(solvers))(*[s._h for s in solvers])
        ws = np.ascontiguousarray(w1s, dtype=np.float64)
        if ws.shape != (len(solvers),):
            raise ValueError("one frequency per solver")
        _lib.check(lib.hamsx_submit_frequency_group(arr, len(solvers), _lib.ptr(ws), float(h)),
                   "hamsx_submit_frequency_group")

    def set_queue(self, queue, tag: int):
        """Post `tag` to `queue` (a CompletionQueue, or None to detach) when
        each submitted solve finishes."""
        _lib.check(self._lib.hamsx_set_queue(self._h, queue._q if queue is not None else None, int(tag)),
                   "hamsx_set_queue")

    def graph_stats(self):
        """(recorded, launched): HIP graphs recorded by this context and solves
        launched from one."""
        r, n = C.c_ulonglong(), C.c_ulonglong()
        _lib.check(self._lib.hamsx_graph_stats(self._h, C.byref(r), C.byref(n)), "hamsx_graph_stats")
        return r.value, n.value

    def query(self) -> bool:
        """True when the submitted solve has finished (or none is pending)."""
        return _lib.check(self._lib.hamsx_query(self._h), "hamsx_query", allow_info=True) == 1

    def collect(self):
        """Wait for the submitted solve: (amas, bdmp, exfc, info)."""
        nd = 6 * self.mesh.nbody
        amas = np.zeros((nd, nd), order="F")
        bdmp = np.zeros((nd, nd), order="F")
        exfc = np.zeros((nd, max(1, self.nbeta)), dtype=np.complex128, order="F")
        info = _lib.check(self._lib.hamsx_collect(self._h, _lib.ptr(amas), _lib.ptr(bdmp), _lib.ptr(exfc)),
                          "hamsx_collect", allow_info=True)
        return amas, bdmp, exfc[:, :self.nbeta], info

    def get_matrix(self, which: str):
        n = self.mesh.nelem
        code = {"A": 0, "S": 1, "LU": 2}[which]
        out = np.zeros((n, n), dtype=np.complex128, order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        _lib.check(self._lib.hamsx_get_matrix(self._h, code, _lib.ptr(out),
                                              _lib.ptr(ipiv, C.c_int) if code == 2 else None),
                   "hamsx_get_matrix")
        return (out, ipiv) if code == 2 else out

    def set_hmatrix(self, enable: bool = True, eta: float = 2.0, aca_tol: float = 1.0e-4, leaf: int = 32,
                    rank_max: int = 50):
        """H-matrix operator of the GMRES path (defaults: the reference's
        HMatrix.f90 parameters); enable=False uses the dense matrix."""
        _lib.check(self._lib.hamsx_set_hmatrix(self._h, int(bool(enable)), float(eta), float(aca_tol), int(leaf),
                                                int(rank_max)), "hamsx_set_hmatrix")

    def hmatrix_stats(self):
        """(low-rank blocks, dense blocks, mean rank, N^2 / stored entries) of the last build."""
        a, b, r, q = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        _lib.check(self._lib.hamsx_hmatrix_stats(self._h, C.byref(a), C.byref(b), C.byref(r), C.byref(q)),
                   "hamsx_hmatrix_stats")
        return a.value, b.value, r.value, q.value

    def hmatrix_matvec(self, x):
        """H x with the H-matrix of the assembled (unfactored) operator."""
        x = np.asfortranarray(x, dtype=np.complex128)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros_like(x, order="F")
        _lib.check(self._lib.hamsx_hmatrix_matvec(self._h, _lib.ptr(x), x.shape[1], _lib.ptr(y)),
                   "hamsx_hmatrix_matvec")
        return y

    def set_refinement(self, max_steps: int = 4, tol: float = 1.0e-10):
        """Mixed precision: at most max_steps FP64-residual / complex64
        correction steps, stopping once max |R|/|X| per column <= tol."""
        _lib.check(self._lib.hamsx_set_refinement(self._h, int(max_steps), float(tol)), "hamsx_set_refinement")

    def refine_stats(self):
        it, res = C.c_int(), C.c_double()
        _lib.check(self._lib.hamsx_refine_stats(self._h, C.byref(it), C.byref(res)), "hamsx_refine_stats")
        return it.value, res.value

    def set_profiling(self, on: bool = True):
        _lib.check(self._lib.hamsx_set_profiling(self._h, 1 if on else 0), "hamsx_set_profiling")

    def stage_times(self, reset: bool = False) -> dict:
        """Device milliseconds per stage (HIP events on the context stream)."""
        ms = np.zeros(len(_lib.STAGES))
        cnt = np.zeros(len(_lib.STAGES), dtype=np.int32)
        _lib.check(self._lib.hamsx_stage_times(self._h, 1 if reset else 0, _lib.ptr(ms), _lib.ptr(cnt, C.c_int)),
                   "hamsx_stage_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_lib.STAGES)}


class CompletionQueue:
    """Completion notices of submitted solves (hamsx_queue_*): contexts
    attached with FrequencySolver.set_queue post their tag when a solve has
    finished; wait() blocks for the next tag without polling."""

    def __init__(self):
        self._lib = _lib.load()
        q = C.c_void_p()
        _lib.check(self._lib.hamsx_queue_create(C.byref(q)), "hamsx_queue_create")
        self._q = q

    def wait(self, timeout_s: float | None = None):
        """The next posted tag, or None when timeout_s passes first."""
        tag = C.c_int()
        got = _lib.check(self._lib.hamsx_queue_wait(self._q, -1.0 if timeout_s is None else max(0.0, timeout_s),
                                                    C.byref(tag)), "hamsx_queue_wait", allow_info=True)
        return tag.value if got == 1 else None

    def close(self):
        if getattr(self, "_q", None):
            self._lib.hamsx_queue_destroy(self._q)
            self._q = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def hmatrix_partition(cen, leaf: int = 32, eta: float = 2.0):
    """The H-matrix partition of centroids cen (n x 3), computed by the
    library on the host: (perm, blocks) with blocks rows (rs, nr, cs, nc, aca),
    0-based, in creation order."""
    lib = _lib.load()
    c = np.asfortranarray(cen, dtype=np.float64)
    n = c.shape[0]
    perm = np.zeros(n, dtype=np.int32)
    nb = C.c_int()
    cap = 64 * n + 64
    blocks = np.zeros(5 * cap, dtype=np.int32)
    _lib.check(lib.hamsx_hmatrix_partition(n, _lib.ptr(c), int(leaf), float(eta), _lib.ptr(perm, C.c_int), cap,
                                           _lib.ptr(blocks, C.c_int), C.byref(nb)), "hamsx_hmatrix_partition")
    return perm, blocks[:5 * nb.value].reshape(nb.value, 5)
