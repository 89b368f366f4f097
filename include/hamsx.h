/*
 * hamsx.h -- C-ABI of the MI355X (gfx950) frequency-domain panel solver.
 *
 * The boundary replaces the per-frequency subroutine calls of HAMS-MREL's
 * driver loop (src/HAMS_Prog.f90:613-733, multi-body; :255-304 single body).
 * Plain pointers and sizes only; every array is caller-owned host memory in
 * Fortran column-major order, complex values are interleaved (re, im) doubles
 * (COMPLEX*16 layout), panel connectivity is 1-based as in the .pnl files.
 *
 * Return codes: 0 = ok; > 0 = LAPACK-style INFO (1-based index of a zero
 * pivot in the factorisation); < 0 = one of the HAMSX_E* codes below.
 * A context is bound to one device and one HIP stream; it is not re-entrant.
 * Many contexts may exist in one process.
 */
#ifndef HAMSX_H
#define HAMSX_H

#ifdef __cplusplus
extern "C" {
#endif

#define HAMSX_OK 0
#define HAMSX_EARG (-1)      /* bad argument (size, pointer, mode) */
#define HAMSX_EDEVICE (-2)   /* HIP runtime error */
#define HAMSX_ENOMEM (-3)    /* device allocation failed */
#define HAMSX_ESTATE (-4)    /* call out of sequence (e.g. factor before assemble) */
#define HAMSX_ENOCONV (-5)   /* iterative refinement did not reach its tolerance */

#define HAMSX_NK 200         /* roots of the dispersion relation kept (WaveDyn_mod NK) */

/* precision modes for hamsx_set_precision */
#define HAMSX_PREC_FP64 0    /* complex128 LU, the reference's arithmetic */
#define HAMSX_PREC_MIXED 1   /* complex64 LU + FP64 iterative refinement */

typedef struct hamsx_ctx hamsx_ctx;

/* Library/version string; never NULL. */
const char *hamsx_version(void);

/* Last error message recorded by the calling thread (empty if none). */
const char *hamsx_last_error(void);

/* Number of visible GPUs (0 if none or the runtime is missing). */
int hamsx_device_count(void);

/* ---- context ---------------------------------------------------------- */
int hamsx_create(hamsx_ctx **ctx, int device);
int hamsx_destroy(hamsx_ctx *ctx);
int hamsx_set_precision(hamsx_ctx *ctx, int mode);

/* Tuning options of a context, taking effect at its next factorisation (no
 * counterpart in the reference, whose ZGETRF has no such choices; picked per
 * problem size by hamsx/autotune.py):
 *   HAMSX_OPT_LU_LOOKAHEAD   -1 (default: on unless the environment sets
 *                            HAMSX_LU_LOOKAHEAD=0), 0 or 1 -- the trailing
 *                            update beyond the next panel on a second stream;
 *                            the factors are the same bit for bit either way;
 *   HAMSX_OPT_LU_SELECT_ROWS 256 (default), 512 or 1024 -- rows per
 *                            pivot-selection block (fewer tournament levels
 *                            with larger blocks; the pivots may differ);
 *   HAMSX_OPT_LU_OUTER_MIN_N 6144 (default) -- outer blocks of four panels
 *                            from this order on, <= 0 never; unless set, a
 *                            frequency group of two or more matrices uses
 *                            them from 1024 on (its selection steps serve
 *                            the whole group).
 * HAMSX_EARG for an unknown option or value. */
#define HAMSX_OPT_LU_LOOKAHEAD 1
#define HAMSX_OPT_LU_SELECT_ROWS 2
#define HAMSX_OPT_LU_OUTER_MIN_N 3
/*   HAMSX_OPT_GRAPHS         0 (default) or 1 -- hamsx_submit_frequency
 *                            records its launch sequence as a HIP graph on
 *                            the second solve at a depth and replays it for
 *                            every later frequency at that depth (dense LU
 *                            path, no lid, no symmetry plane, no stage
 *                            timing; otherwise, and with 0, it launches
 *                            eagerly).  Recordings use the one-stream LU
 *                            schedule; results are the same bit for bit. */
#define HAMSX_OPT_GRAPHS 4
/*   HAMSX_OPT_FP32_FARFIELD  1 (default) or 0 -- with HAMSX_PREC_MIXED, the
 *                            finite-depth far field (EIGEN region, R/h >= 0.5,
 *                            FinGreen3D.f90:281-362) sums its evanescent modes
 *                            in FP32 (K0/K1 by the Abramowitz & Stegun
 *                            polynomials); the propagating mode stays FP64.
 *                            No effect in FP64 mode. */
#define HAMSX_OPT_FP32_FARFIELD 5
int hamsx_set_option(hamsx_ctx *ctx, int option, int value);
int hamsx_get_option(hamsx_ctx *ctx, int option, int *value);

/* ---- host-side set-up routines (no device work) -------------------------
 * Panel centroid, area (two Heron triangles per quad), maximum edge length
 * and the six generalized normals about reference point xr[3].
 * Replaces CalPanelCentre/CalPanelArea/CalPanelChartLength/CalTransNormals/
 * CalRotNormals, src/NormalProcess.f90:46-441, as called by CalNormalsMulti,
 * src/ReadPanelMeshMulti.f90:150-156.
 *   node_xyz[nnode*3], ncn[nelem], ncon[nelem*4] (1-based);
 *   cen[nelem*3], area[nelem], pnsz[nelem], dxyz[nelem*6] are outputs. */
int hamsx_panel_geometry(int nnode, const double *node_xyz, int nelem, const int *ncn,
                         const int *ncon, const double *xr, double *cen, double *area,
                         double *pnsz, double *dxyz);

/* Wave numbers for a frequency given as input type inft (1 deep-water wave
 * number, 2 finite-depth wave number, 3 rad/s, 4 period, 5 wave length),
 * depth h (< 0: infinite).  Outputs v = w^2/g, wk, w1 and the NK dispersion
 * roots wvn[HAMSX_NK] (finite depth only; zeros otherwise).
 * inft = HAMSX_FREQ_LIMIT gives the frequency limits of SYBO = 1
 * (ImplementSubs.f90:45-70): value 0 -> v = wk = w1 = 0 (zero frequency),
 * value -1 -> v = wk = w1 = -1 (infinite frequency), no roots.
 * Replaces CalWaveProperts, src/ImplementSubs.f90:40-151 with DISPERSION,
 * src/WavDynSubs.f90:40-127. */
#define HAMSX_FREQ_LIMIT 0
int hamsx_wave_properties(int inft, double value, double h, double *v, double *wk,
                          double *w1, double *wvn);

/* ---- mesh (device-resident, frequency-invariant) ------------------------
 * The combined multi-body arrays of BuildCombinedArrays,
 * src/CalGreenFuncMulti.f90:53-106: nodes XYZ_GLOBAL_MULTI_COMB(nnode,3),
 * NCN/NCON_MULTI_COMB, centroids XYZ_GLOBAL_MULTI_COMB_P(nelem,3),
 * DS/PNSZ_MULTI_COMB, DXYZ_MULTI_COMB(nelem,6).  Bodies occupy consecutive
 * panel ranges: body b owns panels [body_start[b], body_start[b+1]) (0-based,
 * body_start has nbody+1 entries).  The mesh is the whole body (NSYS = 1)
 * unless hamsx_set_symmetry declares it the half on one side of a plane. */
int hamsx_set_mesh(hamsx_ctx *ctx, int nelem, int nnode, const double *node_xyz,
                   const int *ncn, const int *ncon, const double *cen, const double *area,
                   const double *pnsz, const double *dxyz, int nbody, const int *body_start);

/* Irregular-frequency removal (IRSP = 1): lid_cen[nlid*3] (column-major) are
 * the centroids of the waterplane (lid) panels of all bodies, used as extra
 * field points against the hull sources.  The solve then runs on the normal
 * equations C = A^T A, D = A^T B of the (N + nlid) x N system (plain
 * transposes), as CALGREEN_MULTI_IRR + ASSB_LEFT/RBC/DBC_IRR_MULTI +
 * RADIATION/DIFFRACTION_SOLVER_IRR_MULTI (src/CalGreenFuncMulti.f90:189-318,
 * src/AssbMatx_irrMulti.f90:53-585).  nlid = 0 removes the lid; a new
 * hamsx_set_mesh also removes it.  hamsx_get_matrix(which = 0) then returns C. */
int hamsx_set_lid(hamsx_ctx *ctx, int nlid, const double *lid_cen);

/* Geometric symmetry (ISX / ISY of the control file, NSYS = 2;
 * src/HAMS_Prog.f90:495-504, CalGreenFuncMulti.f90:141-177,
 * AssbMatxMulti.f90:53-383): the mesh set by hamsx_set_mesh is the half on
 * the positive side of the plane x = 0 (isx = 1) or y = 0 (isy = 1), and
 * every solve then works on the whole body: the mirrored panels' entries,
 * the two systems A1 = A + T2, A2 = A - T2 with their right-hand sides, and
 * forces summed over both halves.  isx = isy = 1 is refused as the reference
 * refuses it; isx = isy = 0 removes the plane, and so does a new
 * hamsx_set_mesh.  Every solver path runs both systems: the LU paths (FP64
 * or mixed) factor them at once on two streams, GMRES one after the other.
 * With a lid (hamsx_set_lid,
 * before or after; the half lid's points) each system is solved through its
 * own normal equations (ASSB_LEFT/RBC/DBC_IRR_MULTI with NSYS = 2).  With a
 * plane set, the mxpot outputs of hamsx_solve_radiation / _diffraction hold
 * both halves, MXPOT(:, :, 1) then MXPOT(:, :, 2), and hamsx_field_values
 * sums over both (PressureElevationMulti.f90:89-190). */
int hamsx_set_symmetry(hamsx_ctx *ctx, int isx, int isy);

/* ---- per-frequency state -------------------------------------------------
 * The scalars CalWaveProperts leaves in WaveDyn_mod (V, WK, W1, H, WVN). */
int hamsx_set_frequency(hamsx_ctx *ctx, double v, double wk, double w1, double h,
                        const double *wvn, int nk);

/* Incident-wave parameters read by ASSB_DBC_MULTI / EFORCE_MULTI from module
 * state: headings beta[nbeta] in radians, wave amplitude, reference point
 * XW(2), fluid density and ISOL (1: excitation includes the incident
 * potential, 2: diffraction potential only). */
int hamsx_set_waves(hamsx_ctx *ctx, int nbeta, const double *beta, double amp, double xw0,
                    double xw1, double rho, int isol);

/* CALGREEN_MULTI + ASSB_LEFT_MULTI (src/CalGreenFuncMulti.f90:112-183,
 * src/AssbMatxMulti.f90:53-105) without ZGETRF: builds the influence matrix
 * A(N,N) and the single-layer operator S(N,N) in device memory. */
int hamsx_assemble(hamsx_ctx *ctx);

/* ZGETRF of A (src/AssbMatxMulti.f90:110-114).  Returns INFO. */
int hamsx_factor(hamsx_ctx *ctx);

/* hamsx_factor on caller-owned device memory, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Factors the
 * n x n complex128 matrix a (device pointer, column-major, leading dimension
 * n, interleaved re/im doubles) in place with the kernels of hamsx_factor;
 * ipiv[n] (device int, 1-based LAPACK pivots) and info[0] (device int,
 * LAPACK INFO) are written on `stream` (a hipStream_t; NULL: the null
 * stream).  The call only enqueues work and returns 0, or a HAMSX_E* code
 * (HAMSX_EARG when n > HAMSX_NOALLOC_MAX_N).  Its workspace (the factored
 * diagonal block, the permutation and the pivot tournament's lists, about
 * 180 KB) is one fixed array per GPU, held by the library until the process
 * ends: calls that overlap on one GPU race on it. */
#define HAMSX_NOALLOC_MAX_N 32768
int hamsx_factor_noalloc(int n, double *a, int *ipiv, int *info, void *stream);

/* ASSB_RBC_MULTI + RADIATION_SOLVER_MULTI (src/AssbMatxMulti.f90:124-181,
 * :257-317): mxpot[nelem*6*nbody] (complex) receives the radiation
 * potentials; pass NULL to keep them on the device only. */
int hamsx_solve_radiation(hamsx_ctx *ctx, double *mxpot);

/* ASSB_DBC_MULTI + DIFFRACTION_SOLVER_MULTI (src/AssbMatxMulti.f90:187-251,
 * :323-383) for every heading set by hamsx_set_waves: mxpot[nelem*nbeta]
 * (complex) receives the diffraction potentials, or NULL. */
int hamsx_solve_diffraction(hamsx_ctx *ctx, double *mxpot);

/* RFORCE_MULTI / EFORCE_MULTI reductions (src/PotentWavForceMulti.f90:45-226):
 * amas, bdmp [(6nb)*(6nb)] column-major, AMAS_MULTI_COMB(KK,:,:) layout;
 * exfc [(6nb)*nbeta] complex, EXFC_MULTI_COMB(KK,:,:) transposed to
 * (dof, heading).  Any output pointer may be NULL. */
int hamsx_forces(hamsx_ctx *ctx, double *amas, double *bdmp, double *exfc);

/* One whole frequency solve: wave properties of w1 (rad/s) at the mesh depth
 * h, assembly, factorisation, radiation + all diffraction solves and the
 * force reductions, with a single device synchronisation.  w1 = 0 and
 * w1 = -1 solve the zero- and infinite-frequency limits (SYBO = 1): the
 * Green functions' V = 0 / V = -1 branches; at the limits the reference
 * reports the added mass only (bdmp and exfc are what its sums give: zero
 * at w1 = 0). */
int hamsx_solve_frequency(hamsx_ctx *ctx, double w1, double h, double *amas, double *bdmp,
                          double *exfc);

/* Non-blocking form of hamsx_solve_frequency for pipelines that keep
 * several solves in flight on several contexts (one stream each): submit
 * enqueues the whole solve and returns; query returns 1 when it has finished
 * (0 while running); collect waits for it, fills amas/bdmp/exfc as
 * hamsx_solve_frequency does and returns its INFO.  While a solve is pending
 * every other call on the context except hamsx_query, hamsx_collect and
 * hamsx_destroy returns HAMSX_ESTATE.  No counterpart in the reference,
 * whose frequency loop (src/HAMS_Prog.f90:613-733) is strictly serial; the
 * first submit at a new depth sorts the pairs and waits for the device once. */
int hamsx_submit_frequency(hamsx_ctx *ctx, double w1, double h);
int hamsx_query(hamsx_ctx *ctx);
int hamsx_collect(hamsx_ctx *ctx, double *amas, double *bdmp, double *exfc);

/* A group of 1..HAMSX_GROUP_MAX contexts of one device, one mesh size and
 * one precision (the dense LU path: no GMRES, lid or symmetry plane), each
 * solving its own frequency w1[i] at depth h: each context assembles on its
 * own stream, the group's matrices are factored by one launch sequence on
 * the first context's stream (every launch covers all of them, so the pivot
 * selection's dependent steps serve the whole group), and each context then
 * solves and reduces on its own stream.  Every context ends up pending, as
 * after hamsx_submit_frequency, and is collected the same way; each result
 * is the one its context gives alone under the same LU options, bit for bit
 * (a group of two or more takes outer blocks from order 1024 on unless
 * HAMSX_OPT_LU_OUTER_MIN_N was set, so its default factors differ from a
 * single solve's by rounding only).  No counterpart in the
 * reference, which factors each frequency's matrix by its own ZGETRF
 * (src/AssbMatxMulti.f90:110-114). */
#define HAMSX_GROUP_MAX 16
int hamsx_submit_frequency_group(hamsx_ctx **ctx, int count, const double *w1, double h);

/* Completion notices for hosts that keep several contexts in flight (no
 * counterpart in the reference): a context attached to a queue with a tag
 * (hamsx_set_queue; q NULL detaches) has a host function appended to each
 * submitted solve, which posts the tag when the solve's work has finished.
 * hamsx_queue_wait blocks until a tag is posted (timeout_s < 0: no limit)
 * and returns 1 with the tag, or 0 when the time runs out; the solve is then
 * collected with hamsx_collect as usual.  Destroy a queue only after every
 * context attached to it has been collected or destroyed. */
typedef struct hamsx_queue hamsx_queue;
int hamsx_queue_create(hamsx_queue **q);
int hamsx_queue_destroy(hamsx_queue *q);
int hamsx_set_queue(hamsx_ctx *ctx, hamsx_queue *q, int tag);
int hamsx_queue_wait(hamsx_queue *q, double timeout_s, int *tag);
/* Recorded launch sequences of this context (HAMSX_OPT_GRAPHS): how many
 * were recorded and how many solves were launched from one. */
int hamsx_graph_stats(hamsx_ctx *ctx, unsigned long long *records, unsigned long long *launches);

/* Parity/debug: copy A (which = 0), S (1) or the LU factors (2) to host
 * out[nelem*nelem] (complex); ipiv[nelem] (1-based) for which = 2, else NULL. */
int hamsx_get_matrix(hamsx_ctx *ctx, int which, double *out, int *ipiv);

/* Linear solver for the radiation and diffraction systems.  HAMSX_SOLVER_LU
 * (default): ZGETRF/ZGETRS as above.  HAMSX_SOLVER_GMRES: restarted GMRES
 * with the per-body block-diagonal LU preconditioner, x0 = 0, stopping at
 * relative residual tol or maxiter iterations (the reference's ISOLV = 2,
 * ZITER_SOLVE_MULTI with SETUP_BLOCK_PRECONDITIONER,
 * src/IterativeSolvers.f90:56-58 (restart 25, tol 1e-4, maxiter 100),
 * :190-587, called from src/AssbMatxMulti.f90:283-289, :349-357); the
 * operator is the H-matrix below, or the assembled matrix itself.  In
 * HAMSX_PREC_MIXED the products inside each Arnoldi cycle read complex64
 * copies of the H-matrix blocks, while the Krylov vectors, the
 * orthogonalisation and every restart residual (and so the stopping test)
 * stay FP64 on the FP64 blocks. */
#define HAMSX_SOLVER_LU 0
#define HAMSX_SOLVER_GMRES 1
int hamsx_set_solver(hamsx_ctx *ctx, int solver, double tol, int restart, int maxiter);

/* H-matrix operator of the GMRES path (the reference's HMATRIX_BUILD /
 * HMATVEC, src/HMatrix.f90:135-820, built at src/AssbMatxMulti.f90:288,
 * :356 for every ISOLV = 2 solve): cluster tree by geometric bisection with
 * at most `leaf` panels per leaf, blocks admissible when the distance of the
 * cluster centres exceeds eta times the larger diameter, admissible blocks
 * compressed by partially pivoted ACA to tolerance aca_tol and rank at most
 * rank_max (<= 64), the rest kept dense.  On by default with the reference's
 * parameters (enable 1, eta 2, aca_tol 1e-4, leaf 32, rank_max 50); enable 0
 * makes GMRES use the dense assembled matrix. */
int hamsx_set_hmatrix(hamsx_ctx *ctx, int enable, double eta, double aca_tol, int leaf, int rank_max);

/* After a GMRES solve with the H-matrix: low-rank and dense block counts,
 * the mean rank of the low-rank blocks and N^2 over the stored entries. */
int hamsx_hmatrix_stats(hamsx_ctx *ctx, int *nlowrank, int *ndense, double *mean_rank, double *compression);

/* y = H x for nrhs columns (n x nrhs complex, host arrays) with the
 * H-matrix of the current assembled (unfactored) operator, built if needed:
 * the product GMRES uses, for checks against the dense matrix. */
int hamsx_hmatrix_matvec(hamsx_ctx *ctx, const double *x, int nrhs, double *y);

/* The partition alone, on the host (no device): cen n x 3 column-major;
 * perm[k] = 0-based panel at cluster position k; blocks[5 k .. 5 k + 4] =
 * row start, rows, column start, columns (cluster order, 0-based) and 1 for
 * an ACA candidate / 0 for a dense block, in the reference's creation
 * order.  *nblocks receives the count; HAMSX_EARG when it exceeds
 * max_blocks. */
int hamsx_hmatrix_partition(int n, const double *cen, int leaf, double eta, int *perm, int max_blocks, int *blocks,
                            int *nblocks);

/* Mixed precision (HAMSX_PREC_MIXED) iterative refinement: after the
 * complex64 LU solve, FP64 residual + complex64 correction steps run until
 * max over the right-hand sides of |R|_max / |X|_max <= tol, at most
 * max_steps of them (defaults 4 and 1e-10).  The test runs on the device
 * (no host wait); hamsx_forces / hamsx_collect return HAMSX_ENOCONV when it
 * was not met.  tol = 0 runs exactly max_steps steps.  No counterpart in the reference, which has no
 * mixed-precision path. */
int hamsx_set_refinement(hamsx_ctx *ctx, int max_steps, double tol);

/* Statistics of the last solve: mixed precision -- correction steps taken
 * and the final max-norm relative residual; GMRES -- Arnoldi iterations and
 * the largest final relative residual 2-norm over the right-hand sides. */
int hamsx_refine_stats(hamsx_ctx *ctx, int *iters, double *resid);

/* ---- field points ----------------------------------------------------------
 * Pressure and free-surface elevation away from the hull panels:
 * CalPotentialMulti / CalPressureMulti / CalElevationMulti
 * (src/PressureElevationMulti.f90:55-348) for every point and every mode, as
 * the output drivers OutputPressureElevation_RadiationMulti / _DiffractionMulti
 * / _IncidenceMulti (:407-556) call them point by point.
 *
 * hamsx_set_field_points: xfp[nfp*3] column-major, the XFP(NFP,3) array that
 * ReadControlFile reads (src/InputFiles.f90:193-203).  nfp = 0 removes them;
 * a new hamsx_set_mesh also does. */
int hamsx_set_field_points(hamsx_ctx *ctx, int nfp, const double *xfp);

/* kinds for hamsx_field_values */
#define HAMSX_FIELD_POTENTIAL 0  /* CalPotentialMulti: POT */
#define HAMSX_FIELD_PRESSURE 1   /* CalPressureMulti: i w rho POT */
#define HAMSX_FIELD_ELEVATION 2  /* CalElevationMulti: i w / g POT */
#define HAMSX_FIELD_AUTO 3       /* the output drivers' choice per point: pressure
                                    where |z| > 1e-6, elevation on z = 0 */

/* Values of the current frequency's potentials (solved here if the solve has
 * not run yet) at the field points:
 *   values[nfp*(6nb + nbeta)] complex, column-major (point fastest): column
 *   6b + m is radiation mode m + 1 of body b (FACTOR_NORMAL on body b's panels,
 *   PHIDOF = b + 1), column 6nb + k the diffraction potential of heading k,
 *   including the incident potential when ISOL = 1;
 *   incident[nfp*nbeta] complex (may be NULL): the incident wave alone
 *   (CalPotentialIncMulti: VINP when ISOL = 1, else 0), same kind;
 *   onbody[nfp] (may be NULL): 1 where the point lies on a hull panel (free
 *   term 2 pi), 0 in the fluid (4 pi).
 * The reference's WAMIT scaling (WamitNondimensMulti, :352-402) is left to
 * the caller's writer.  NSYS = 1. */
int hamsx_field_values(hamsx_ctx *ctx, int kind, double *values, double *incident, int *onbody);

/* ---- motions ----------------------------------------------------------------
 * hamsx_set_body_matrices: the body matrices ReadHydroStaticMulti reads
 * (src/HydroStaticMulti.f90:68-98) -- mass MATX, external linear damping
 * BLNR, external quadratic damping BQDR, hydrostatic restoring CRS and
 * external restoring KSTF -- each as nbody blocks of 6 x 6 (body b's block at
 * [36 b], column-major), for the bodies of the current mesh (6 nbody <= 192).
 * A new hamsx_set_mesh drops them. */
int hamsx_set_body_matrices(hamsx_ctx *ctx, const double *mass, const double *blnr, const double *bqdr,
                            const double *crs, const double *kstf);

/* SolveMotionMulti (src/SolveMotionMulti.f90:46-150) for every heading of the
 * current frequency (solved here if it has not been):
 *   [-w^2 (MATX + AMAS) + i w (BDMP + BLNR + VDMP) + CRS + KSTF] DSPL = -Im(EXFC) - i Re(EXFC),
 * first with VDMP = 0, then -- when max |BQDR| >= 1e-6 -- with VDMP(i,j) =
 * BQDR(i,j) w |DSPL(j)| until sum_j |dDSPL_j| / |DSPL_j| <= 1e-6.
 * dspl[6nb*nbeta] complex (DOF fastest): DSPL before the WAMIT scaling of
 * Motion.4; iters[nbeta] (may be NULL): the quadratic-damping steps (0 when
 * none ran).  Returns HAMSX_ENOCONV, with the last iterates in dspl, when a
 * heading needs more than 1000 steps (the reference loops without a bound). */
int hamsx_motions(hamsx_ctx *ctx, double *dspl, int *iters);

/* Measurement hook (no counterpart in the reference, which has no timers):
 * with profiling on, every stage is bracketed by HIP events on the context's
 * stream.  hamsx_stage_times returns the device milliseconds accumulated per
 * stage since the last reset -- ms[0] assembly, ms[1] right-hand sides,
 * ms[2] factorisation, ms[3] solves (with refinement), ms[4] forces; within
 * the finite-depth assembly, ms[5..8] the region kernels LINTON, PIDCOCK,
 * EIGENE and EIGEN (R/h < 5e-4, < 0.05, < 0.5, >= 0.5); ms[9] the field-point
 * values (hamsx_field_values) -- and the number of launches of each in
 * count[HAMSX_NSTAGE]. */
#define HAMSX_NSTAGE 10
int hamsx_set_profiling(hamsx_ctx *ctx, int on);
int hamsx_stage_times(hamsx_ctx *ctx, int reset, double *ms, int *count);

#ifdef __cplusplus
}
#endif
#endif
