! hamsx_mod.f90 -- Fortran 2003 (iso_c_binding) binding of libhamsx.so, the
! MI355X (gfx950) frequency-domain panel solver, for a HAMS-MREL style host.
!
! Part 1 declares every C entry point of include/hamsx.h.  Part 2 wraps them
! in subroutines named after the per-frequency calls of HAMS-MREL's driver
! (src/HAMS_Prog.f90:613-733) and taking the driver's own arrays:
!
!   GPU_INIT                 replaces nothing (opens a device context)
!   SET_MESH_MULTI_GPU       after BuildCombinedArrays (CalGreenFuncMulti.f90:53-106)
!   CALWAVEPROPERTS_GPU      CalWaveProperts (ImplementSubs.f90:40-151)
!   CALGREEN_ASSB_LEFT_GPU   CALGREEN_MULTI + ASSB_LEFT_MULTI incl. ZGETRF
!                            (CalGreenFuncMulti.f90:112-183, AssbMatxMulti.f90:53-114)
!   RADIATION_SOLVER_GPU     ASSB_RBC_MULTI + RADIATION_SOLVER_MULTI (AssbMatxMulti.f90:124-317)
!   DIFFRACTION_SOLVER_GPU   ASSB_DBC_MULTI + DIFFRACTION_SOLVER_MULTI (AssbMatxMulti.f90:187-383)
!   FORCES_GPU               RFORCE_MULTI / EFORCE_MULTI sums (PotentWavForceMulti.f90:45-226)
!   PANEL_GEOMETRY_GPU       NormalProcess for one body (NormalProcess.f90:46-441)
!
! Part 3 keeps the reference's own names and argument lists for the calls of
! the multi-body loop (src/AssbMatxMulti.f90:53-383), acting on the context
! GPU_INIT opened last, so the loop body of src/HAMS_Prog.f90:629-724 runs
! unchanged:
!
!   CALGREEN_MULTI()                                   assembly on the device
!   ASSB_LEFT_MULTI(AMAT, IPIV, N, NSYS)               ZGETRF on the device
!   ASSB_RBC_MULTI(BRMAT, N, NSYS, JSTART, JEND)       (right-hand sides formed by the solver)
!   RADIATION_SOLVER_MULTI(AMAT, BRMAT, IPIV, MXPOT, N, NSYS)
!   ASSB_DBC_MULTI(BDMAT, N, NSYS)                     (right-hand side formed by the solver)
!   DIFFRACTION_SOLVER_MULTI(AMAT, BDMAT, IPIV, MXPOT, N, NSYS)   one heading per call
!
! AMAT, IPIV, BRMAT and BDMAT are the caller's storage in the reference; here
! the matrices, factors and right-hand sides stay in device memory and those
! arrays are left untouched.  MXPOT(N, 6 NBODY + 1, NSYS) receives the
! potentials as the reference fills it.
!
! The influence matrices and their factors stay in device memory; only the
! potentials (optional) and the force coefficients come back.  Every wrapper
! stops the program with the library's message on an error, the way the
! reference stops on input errors.
module hamsx_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: HAMSX_OK = 0, HAMSX_NK = 200
  integer(c_int), parameter, public :: HAMSX_PREC_FP64 = 0, HAMSX_PREC_MIXED = 1
  integer(c_int), parameter, public :: HAMSX_SOLVER_LU = 0, HAMSX_SOLVER_GMRES = 1
  integer(c_int), parameter, public :: HAMSX_FIELD_POTENTIAL = 0, HAMSX_FIELD_PRESSURE = 1, &
                                       HAMSX_FIELD_ELEVATION = 2, HAMSX_FIELD_AUTO = 3

  public :: hamsx_version, hamsx_last_error, hamsx_device_count, hamsx_create, hamsx_destroy, &
            hamsx_set_precision, hamsx_panel_geometry, hamsx_wave_properties, hamsx_set_mesh, &
            hamsx_set_frequency, hamsx_set_waves, hamsx_assemble, hamsx_factor, hamsx_solve_radiation, &
            hamsx_solve_diffraction, hamsx_forces, hamsx_solve_frequency, hamsx_submit_frequency, &
            hamsx_query, hamsx_collect, hamsx_get_matrix, hamsx_refine_stats, hamsx_set_profiling, &
            hamsx_set_solver, hamsx_set_lid, hamsx_set_refinement, hamsx_set_hmatrix, &
            hamsx_hmatrix_stats, hamsx_hmatrix_matvec, hamsx_hmatrix_partition, &
            hamsx_stage_times, hamsx_set_field_points, hamsx_field_values, hamsx_set_body_matrices, &
            hamsx_motions, hamsx_set_symmetry
  public :: GPU_INIT, GPU_FINALIZE, SET_MESH_MULTI_GPU, SET_WAVES_MULTI_GPU, CALWAVEPROPERTS_GPU, &
            CALGREEN_ASSB_LEFT_GPU, RADIATION_SOLVER_GPU, DIFFRACTION_SOLVER_GPU, FORCES_GPU, &
            PANEL_GEOMETRY_GPU, SET_SYMMETRY_GPU, hamsx_error_message
  public :: CALGREEN_MULTI, ASSB_LEFT_MULTI, ASSB_RBC_MULTI, RADIATION_SOLVER_MULTI, ASSB_DBC_MULTI, &
            DIFFRACTION_SOLVER_MULTI

  ! state of the reference-named calls (part 3): the active context, the mesh's
  ! body count, the headings and the diffraction potentials of all headings
  type(c_ptr), save :: g_ctx = c_null_ptr
  integer, save :: g_nbody = 0, g_nbeta = 0, g_nsys = 1, g_heading = 0
  complex(c_double_complex), allocatable, target, save :: g_dif(:, :, :)

  ! ---------------------------------------------------------------- part 1
  interface
    function hamsx_version() bind(C, name="hamsx_version") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function hamsx_last_error() bind(C, name="hamsx_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function hamsx_device_count() bind(C, name="hamsx_device_count") result(n)
      import :: c_int
      integer(c_int) :: n
    end function
    function hamsx_create(ctx, device) bind(C, name="hamsx_create") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
      integer(c_int) :: rc
    end function
    function hamsx_destroy(ctx) bind(C, name="hamsx_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function hamsx_set_precision(ctx, mode) bind(C, name="hamsx_set_precision") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: rc
    end function
    function hamsx_panel_geometry(nnode, node_xyz, nelem, ncn, ncon, xr, cen, area, pnsz, dxyz) &
        bind(C, name="hamsx_panel_geometry") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: nnode, nelem
      real(c_double), intent(in) :: node_xyz(*), xr(3)
      integer(c_int), intent(in) :: ncn(*), ncon(*)
      real(c_double), intent(out) :: cen(*), area(*), pnsz(*), dxyz(*)
      integer(c_int) :: rc
    end function
    function hamsx_wave_properties(inft, value, h, v, wk, w1, wvn) bind(C, name="hamsx_wave_properties") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: inft
      real(c_double), value :: value, h
      real(c_double), intent(out) :: v, wk, w1, wvn(*)
      integer(c_int) :: rc
    end function
    function hamsx_set_mesh(ctx, nelem, nnode, node_xyz, ncn, ncon, cen, area, pnsz, dxyz, nbody, body_start) &
        bind(C, name="hamsx_set_mesh") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nelem, nnode, nbody
      real(c_double), intent(in) :: node_xyz(*), cen(*), area(*), pnsz(*), dxyz(*)
      integer(c_int), intent(in) :: ncn(*), ncon(*), body_start(*)
      integer(c_int) :: rc
    end function
    function hamsx_set_frequency(ctx, v, wk, w1, h, wvn, nk) bind(C, name="hamsx_set_frequency") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: v, wk, w1, h
      real(c_double), intent(in) :: wvn(*)
      integer(c_int), value :: nk
      integer(c_int) :: rc
    end function
    function hamsx_set_waves(ctx, nbeta, beta, amp, xw0, xw1, rho, isol) bind(C, name="hamsx_set_waves") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbeta, isol
      real(c_double), intent(in) :: beta(*)
      real(c_double), value :: amp, xw0, xw1, rho
      integer(c_int) :: rc
    end function
    function hamsx_assemble(ctx) bind(C, name="hamsx_assemble") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function hamsx_factor(ctx) bind(C, name="hamsx_factor") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function hamsx_solve_radiation(ctx, mxpot) bind(C, name="hamsx_solve_radiation") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, mxpot
      integer(c_int) :: rc
    end function
    function hamsx_solve_diffraction(ctx, mxpot) bind(C, name="hamsx_solve_diffraction") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, mxpot
      integer(c_int) :: rc
    end function
    function hamsx_forces(ctx, amas, bdmp, exfc) bind(C, name="hamsx_forces") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, amas, bdmp, exfc
      integer(c_int) :: rc
    end function
    function hamsx_solve_frequency(ctx, w1, h, amas, bdmp, exfc) bind(C, name="hamsx_solve_frequency") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, amas, bdmp, exfc
      real(c_double), value :: w1, h
      integer(c_int) :: rc
    end function
    function hamsx_submit_frequency(ctx, w1, h) bind(C, name="hamsx_submit_frequency") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: w1, h
      integer(c_int) :: rc
    end function
    function hamsx_query(ctx) bind(C, name="hamsx_query") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function hamsx_collect(ctx, amas, bdmp, exfc) bind(C, name="hamsx_collect") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, amas, bdmp, exfc
      integer(c_int) :: rc
    end function
    function hamsx_get_matrix(ctx, which, out, ipiv) bind(C, name="hamsx_get_matrix") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, out, ipiv
      integer(c_int), value :: which
      integer(c_int) :: rc
    end function
    function hamsx_set_lid(ctx, nlid, lid_cen) bind(C, name="hamsx_set_lid") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nlid
      real(c_double), intent(in) :: lid_cen(*)
      integer(c_int) :: rc
    end function
    ! ISX / ISY of ReadControlFile: the mesh is the half on the positive side of the plane
    function hamsx_set_symmetry(ctx, isx, isy) bind(C, name="hamsx_set_symmetry") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: isx, isy
      integer(c_int) :: rc
    end function
    ! XFP(NFP,3) of ReadControlFile; pass any array with nfp = 0 to remove the points
    function hamsx_set_field_points(ctx, nfp, xfp) bind(C, name="hamsx_set_field_points") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfp
      real(c_double), intent(in) :: xfp(*)
      integer(c_int) :: rc
    end function
    ! values(NFP, 6*NBODY+NBETA); incident(NFP, NBETA) and onbody(NFP) may be c_null_ptr
    function hamsx_field_values(ctx, kind, values, incident, onbody) bind(C, name="hamsx_field_values") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      type(c_ptr), value :: values, incident, onbody
      integer(c_int) :: rc
    end function
    ! each argument: NBODY blocks of 6 x 6, e.g. MATX_MULTI(B,:,:) copied to MASS(:,:,B)
    function hamsx_set_body_matrices(ctx, mass, blnr, bqdr, crs, kstf) bind(C, name="hamsx_set_body_matrices") &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: mass(*), blnr(*), bqdr(*), crs(*), kstf(*)
      integer(c_int) :: rc
    end function
    ! dspl(6*NBODY, NBETA) complex; iters(NBETA) may be c_null_ptr
    function hamsx_motions(ctx, dspl, iters) bind(C, name="hamsx_motions") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, dspl, iters
      integer(c_int) :: rc
    end function
    function hamsx_set_solver(ctx, solver, tol, restart, maxiter) bind(C, name="hamsx_set_solver") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: solver, restart, maxiter
      real(c_double), value :: tol
      integer(c_int) :: rc
    end function
    function hamsx_set_hmatrix(ctx, enable, eta, aca_tol, leaf, rank_max) bind(C, name="hamsx_set_hmatrix") &
        result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable, leaf, rank_max
      real(c_double), value :: eta, aca_tol
      integer(c_int) :: rc
    end function
    function hamsx_hmatrix_stats(ctx, nlowrank, ndense, mean_rank, compression) &
        bind(C, name="hamsx_hmatrix_stats") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nlowrank, ndense
      real(c_double), intent(out) :: mean_rank, compression
      integer(c_int) :: rc
    end function
    function hamsx_hmatrix_matvec(ctx, x, nrhs, y) bind(C, name="hamsx_hmatrix_matvec") result(rc)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: x(*)
      integer(c_int), value :: nrhs
      complex(c_double_complex), intent(out) :: y(*)
      integer(c_int) :: rc
    end function
    function hamsx_hmatrix_partition(n, cen, leaf, eta, perm, max_blocks, blocks, nblocks) &
        bind(C, name="hamsx_hmatrix_partition") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n, leaf, max_blocks
      real(c_double), intent(in) :: cen(*)
      real(c_double), value :: eta
      integer(c_int), intent(out) :: perm(*), blocks(*), nblocks
      integer(c_int) :: rc
    end function
    function hamsx_set_refinement(ctx, max_steps, tol) bind(C, name="hamsx_set_refinement") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_steps
      real(c_double), value :: tol
      integer(c_int) :: rc
    end function
    function hamsx_refine_stats(ctx, iters, resid) bind(C, name="hamsx_refine_stats") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: resid
      integer(c_int) :: rc
    end function
    function hamsx_set_profiling(ctx, on) bind(C, name="hamsx_set_profiling") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      integer(c_int) :: rc
    end function
    function hamsx_stage_times(ctx, reset, ms, count) bind(C, name="hamsx_stage_times") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reset
      real(c_double), intent(out) :: ms(10)
      integer(c_int), intent(out) :: count(10)
      integer(c_int) :: rc
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! ---------------------------------------------------------------- part 2
  function hamsx_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = hamsx_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [max(n, 1)])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = chars(k)
    end do
  end function

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc < 0) then
      write(*, '(A,A,A,I0,A,A)') ' hamsx: ', where, ' failed (', rc, '): ', hamsx_error_message()
      stop 1
    end if
  end subroutine

  ! Open a context on GPU `device` (0-based) in the given precision mode.
  subroutine GPU_INIT(ctx, device, precision)
    type(c_ptr), intent(out) :: ctx
    integer, intent(in) :: device, precision
    call check(hamsx_create(ctx, int(device, c_int)), "hamsx_create")
    call check(hamsx_set_precision(ctx, int(precision, c_int)), "hamsx_set_precision")
    g_ctx = ctx
  end subroutine

  subroutine GPU_FINALIZE(ctx)
    type(c_ptr), intent(inout) :: ctx
    call check(hamsx_destroy(ctx), "hamsx_destroy")
    if (c_associated(ctx, g_ctx)) g_ctx = c_null_ptr
    ctx = c_null_ptr
  end subroutine

  ! NormalProcess for one body (CalPanelCentre/Area/ChartLength/Normals).
  subroutine PANEL_GEOMETRY_GPU(NNODE, XYZ, NELEM, NCN, NCON, XR, XYZ_P, DS, PNSZ, DXYZ_P)
    integer, intent(in) :: NNODE, NELEM, NCN(NELEM), NCON(NELEM, 4)
    real(c_double), intent(in) :: XYZ(NNODE, 3), XR(3)
    real(c_double), intent(out) :: XYZ_P(NELEM, 3), DS(NELEM), PNSZ(NELEM), DXYZ_P(NELEM, 6)
    integer(c_int) :: ncn_c(NELEM), ncon_c(NELEM, 4)
    ncn_c = int(NCN, c_int)
    ncon_c = int(NCON, c_int)
    call check(hamsx_panel_geometry(int(NNODE, c_int), XYZ, int(NELEM, c_int), ncn_c, ncon_c, XR, XYZ_P, DS, &
                                    PNSZ, DXYZ_P), "hamsx_panel_geometry")
  end subroutine

  ! The combined multi-body arrays of BuildCombinedArrays; NELEM_B(NB) are the
  ! panel counts of the bodies in the order they were concatenated.
  subroutine SET_MESH_MULTI_GPU(ctx, TNELEM, TNNODE, XYZ_COMB, NCN_COMB, NCON_COMB, XYZ_P_COMB, DS_COMB, &
                                  PNSZ_COMB, DXYZ_COMB, NB, NELEM_B)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: TNELEM, TNNODE, NB, NELEM_B(NB), NCN_COMB(TNELEM), NCON_COMB(TNELEM, 4)
    real(c_double), intent(in) :: XYZ_COMB(TNNODE, 3), XYZ_P_COMB(TNELEM, 3), DS_COMB(TNELEM), PNSZ_COMB(TNELEM), &
                                  DXYZ_COMB(TNELEM, 6)
    integer(c_int) :: starts(NB + 1), ncn_c(TNELEM), ncon_c(TNELEM, 4)
    integer :: b
    starts(1) = 0
    do b = 1, NB
      starts(b + 1) = starts(b) + int(NELEM_B(b), c_int)
    end do
    ncn_c = int(NCN_COMB, c_int)
    ncon_c = int(NCON_COMB, c_int)
    call check(hamsx_set_mesh(ctx, int(TNELEM, c_int), int(TNNODE, c_int), XYZ_COMB, ncn_c, ncon_c, XYZ_P_COMB, &
                              DS_COMB, PNSZ_COMB, DXYZ_COMB, int(NB, c_int), starts), "hamsx_set_mesh")
    g_nbody = NB
    g_nsys = 1
  end subroutine

  ! Headings (radians), AMP, XW(2), RHO and ISOL as read from ControlFile.in.
  subroutine SET_WAVES_MULTI_GPU(ctx, NBETA, WVHD, AMP, XW, RHO, ISOL)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: NBETA, ISOL
    real(c_double), intent(in) :: WVHD(NBETA), AMP, XW(2), RHO
    call check(hamsx_set_waves(ctx, int(NBETA, c_int), WVHD, AMP, XW(1), XW(2), RHO, int(ISOL, c_int)), &
               "hamsx_set_waves")
    g_nbeta = NBETA
  end subroutine

  ! INFT/OUFT as in ControlFile.in; returns V, WK, W1 and WVN(NK) and makes
  ! them the context's current frequency.
  subroutine CALWAVEPROPERTS_GPU(ctx, INFT, WFREQ, H, V, WK, W1, WVN)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: INFT
    real(c_double), intent(in) :: WFREQ, H
    real(c_double), intent(out) :: V, WK, W1, WVN(HAMSX_NK)
    call check(hamsx_wave_properties(int(INFT, c_int), WFREQ, H, V, WK, W1, WVN), "hamsx_wave_properties")
    call check(hamsx_set_frequency(ctx, V, WK, W1, H, WVN, HAMSX_NK), "hamsx_set_frequency")
  end subroutine

  ! Influence matrices on the device and their LU factors; INFO as ZGETRF.
  subroutine CALGREEN_ASSB_LEFT_GPU(ctx, INFO)
    type(c_ptr), intent(in) :: ctx
    integer, intent(out) :: INFO
    integer(c_int) :: rc
    call check(hamsx_assemble(ctx), "hamsx_assemble")
    rc = hamsx_factor(ctx)
    call check(rc, "hamsx_factor")
    INFO = int(rc)
  end subroutine

  ! MXPOT(TNELEM, 6*NB) receives the radiation potentials (MXPOT(TNELEM, 6*NB, 2)
  ! with a symmetry plane).
  subroutine RADIATION_SOLVER_GPU(ctx, MXPOT)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), intent(inout), target :: MXPOT(*)
    call check(hamsx_solve_radiation(ctx, c_loc(MXPOT(1))), "hamsx_solve_radiation")
  end subroutine

  ! MXPOT(TNELEM, NBETA) receives the diffraction potentials (MXPOT(TNELEM, NBETA, 2)
  ! with a symmetry plane).
  subroutine DIFFRACTION_SOLVER_GPU(ctx, MXPOT)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), intent(inout), target :: MXPOT(*)
    call check(hamsx_solve_diffraction(ctx, c_loc(MXPOT(1))), "hamsx_solve_diffraction")
  end subroutine

  ! ISX / ISY of the .pnl header (NSYS = 2 when one is 1): the mesh is the half
  ! on the positive side of the plane; the solver MXPOT arguments then hold
  ! MXPOT(TNELEM, ncol, 2).
  subroutine SET_SYMMETRY_GPU(ctx, ISX, ISY)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: ISX, ISY
    call check(hamsx_set_symmetry(ctx, int(ISX, c_int), int(ISY, c_int)), "hamsx_set_symmetry")
    g_nsys = merge(2, 1, ISX /= 0 .or. ISY /= 0)
  end subroutine

  ! AMAS/BDMP(6NB, 6NB) as AMAS_MULTI_COMB(KK,:,:); EXFC(6NB, NBETA).
  subroutine FORCES_GPU(ctx, AMAS, BDMP, EXFC)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out), target, contiguous :: AMAS(:, :), BDMP(:, :)
    complex(c_double_complex), intent(out), target, contiguous :: EXFC(:, :)
    call check(hamsx_forces(ctx, c_loc(AMAS), c_loc(BDMP), c_loc(EXFC)), "hamsx_forces")
  end subroutine

  ! ---------------------------------------------------------------- part 3
  ! CALGREEN_MULTI (CalGreenFuncMulti.f90:112-183): the Green functions and
  ! panel integrals of every pair, fused on the device with the influence
  ! matrix of ASSB_LEFT_MULTI (no CGRN / RKBN arrays).
  subroutine CALGREEN_MULTI()
    call check(hamsx_assemble(g_ctx), "hamsx_assemble")
    g_heading = 0
  end subroutine

  ! ASSB_LEFT_MULTI (AssbMatxMulti.f90:53-114): ZGETRF of the assembled
  ! matrix, on the device; INFO is ignored as the reference ignores it.
  subroutine ASSB_LEFT_MULTI(AMAT, IPIV, N, NSYS)
    integer, intent(in) :: N, NSYS
    complex(c_double_complex), intent(inout) :: AMAT(N, N, NSYS)
    integer, intent(inout) :: IPIV(N, NSYS)
    call check(hamsx_factor(g_ctx), "hamsx_factor")
  end subroutine

  ! ASSB_RBC_MULTI (AssbMatxMulti.f90:124-181): the radiation right-hand
  ! sides of panels JSTART..JEND; RADIATION_SOLVER_MULTI forms them on the device.
  subroutine ASSB_RBC_MULTI(BRMAT, N, NSYS, JSTART, JEND)
    integer, intent(in) :: N, NSYS, JSTART, JEND
    complex(c_double_complex), intent(inout) :: BRMAT(N, 6, NSYS)
  end subroutine

  ! RADIATION_SOLVER_MULTI (AssbMatxMulti.f90:257-317): MXPOT(:, 1:6 NBODY, :).
  subroutine RADIATION_SOLVER_MULTI(AMAT, BRMAT, IPIV, MXPOT, N, NSYS)
    integer, intent(in) :: N, NSYS
    complex(c_double_complex), intent(in) :: AMAT(N, N, NSYS), BRMAT(N, 6 * g_nbody, NSYS)
    integer, intent(in) :: IPIV(N, NSYS)
    complex(c_double_complex), intent(inout) :: MXPOT(N, 6 * g_nbody + 1, NSYS)
    complex(c_double_complex), allocatable, target :: x(:, :, :)
    if (NSYS /= g_nsys) call check(-1_c_int, "RADIATION_SOLVER_MULTI: NSYS differs from the symmetry set")
    allocate(x(N, 6 * g_nbody, NSYS))
    call check(hamsx_solve_radiation(g_ctx, c_loc(x)), "hamsx_solve_radiation")
    MXPOT(:, 1:6 * g_nbody, :) = x
  end subroutine

  ! ASSB_DBC_MULTI (AssbMatxMulti.f90:187-251): the diffraction right-hand
  ! side of the current heading; DIFFRACTION_SOLVER_MULTI forms it on the device.
  subroutine ASSB_DBC_MULTI(BDMAT, N, NSYS)
    integer, intent(in) :: N, NSYS
    complex(c_double_complex), intent(inout) :: BDMAT(N, NSYS)
  end subroutine

  ! DIFFRACTION_SOLVER_MULTI (AssbMatxMulti.f90:323-383): MXPOT(:, 6 NBODY + 1, :)
  ! for the next heading of the driver's loop DO II = 1, NBETA
  ! (HAMS_Prog.f90:700-727).  The first call after CALGREEN_MULTI solves every
  ! heading at once on the device; the following calls return the next one.
  subroutine DIFFRACTION_SOLVER_MULTI(AMAT, BDMAT, IPIV, MXPOT, N, NSYS)
    integer, intent(in) :: N, NSYS
    complex(c_double_complex), intent(in) :: AMAT(N, N, NSYS), BDMAT(N, NSYS)
    integer, intent(in) :: IPIV(N, NSYS)
    complex(c_double_complex), intent(inout) :: MXPOT(N, 6 * g_nbody + 1, NSYS)
    if (NSYS /= g_nsys) call check(-1_c_int, "DIFFRACTION_SOLVER_MULTI: NSYS differs from the symmetry set")
    if (g_heading == 0) then
      if (allocated(g_dif)) deallocate(g_dif)
      allocate(g_dif(N, max(g_nbeta, 1), NSYS))
      call check(hamsx_solve_diffraction(g_ctx, c_loc(g_dif)), "hamsx_solve_diffraction")
    end if
    if (g_heading >= g_nbeta) call check(-1_c_int, "DIFFRACTION_SOLVER_MULTI: more calls than headings")
    g_heading = g_heading + 1
    MXPOT(:, 6 * g_nbody + 1, :) = g_dif(:, g_heading, :)
  end subroutine

end module hamsx_mod
