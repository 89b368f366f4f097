! hamsx_pnl_solve.f90 -- a Fortran host for libhamsx: reads one body from a
! HullMesh .pnl file and runs the per-frequency sequence of HAMS-MREL's
! driver (src/HAMS_Prog.f90:613-733) through hamsx_mod's wrappers.
!
! Usage: hamsx_pnl_solve <input file>, the input file holding, list-directed:
!   'mesh.pnl'             panel mesh (format of src/ReadPanelMeshMulti.f90:46-71)
!   xr1 xr2 xr3            reference point of the normals (Hydrostatic XR)
!   h                      water depth (< 0: infinite)
!   nfreq                  then nfreq frequencies in rad/s (INFT = 3)
!   nbeta                  then nbeta headings in degrees
!   precision              0: FP64, 1: complex64 LU + FP64 refinement
!   style (optional)       0: hamsx_mod's *_GPU wrappers (default); 1: the
!                          reference-named calls CALGREEN_MULTI, ASSB_LEFT_MULTI,
!                          ASSB_RBC_MULTI, RADIATION_SOLVER_MULTI, ASSB_DBC_MULTI,
!                          DIFFRACTION_SOLVER_MULTI in the order of HAMS_Prog.f90:629-724,
!                          then a line "MXPOT" and the largest difference of their
!                          MXPOT from the *_GPU wrappers' potentials
! Prints, per frequency, lines "W", "AMAS", "BDMP" and "EXFC" followed by the
! values (ES25.16E3), for machine reading.
program hamsx_pnl_solve
  use, intrinsic :: iso_c_binding
  use hamsx_mod
  implicit none
  character(len=1024) :: inp, pnl
  integer :: u, nelem, nnode, isx, isy, i, k, nfreq, nbeta, prec, info, kf, idum, nsys, style, ios, ii
  real(c_double) :: xr(3), h, v, wk, w1, wvn(HAMSX_NK)
  real(c_double), allocatable :: xyz(:, :), cen(:, :), ds(:), pnsz(:), dxyz(:, :), freqs(:), beta(:)
  real(c_double), allocatable :: amas(:, :), bdmp(:, :)
  complex(c_double_complex), allocatable :: exfc(:, :), rad(:, :, :), dif(:, :, :)
  complex(c_double_complex), allocatable :: amat(:, :, :), brmat(:, :, :), bdmat(:, :), mxpot(:, :, :), &
                                              dif_ref(:, :, :), chk(:, :, :)
  integer, allocatable :: ipiv(:, :)
  real(c_double) :: dmax
  integer, allocatable :: ncn(:), ncon(:, :)
  type(c_ptr) :: ctx
  real(c_double), parameter :: pi = 3.141592653589793d0

  call get_command_argument(1, inp)
  open(newunit=u, file=trim(inp), status='old', action='read')
  read(u, *) pnl
  read(u, *) xr
  read(u, *) h
  read(u, *) nfreq
  allocate(freqs(nfreq))
  read(u, *) freqs
  read(u, *) nbeta
  allocate(beta(max(nbeta, 1)))
  if (nbeta > 0) read(u, *) beta(1:nbeta)
  read(u, *) prec
  read(u, *, iostat=ios) style
  if (ios /= 0) style = 0
  close(u)
  beta = beta * pi / 180.0d0

  ! .pnl: title, blank, comment, counts, blank, comment, nodes, 3 lines, panels
  open(newunit=u, file=trim(pnl), status='old', action='read')
  read(u, *)
  read(u, *)
  read(u, *)
  read(u, *) nelem, nnode, isx, isy
  read(u, *)
  read(u, *)
  allocate(xyz(nnode, 3), ncn(nelem), ncon(nelem, 4))
  do i = 1, nnode
    read(u, *) idum, xyz(i, 1), xyz(i, 2), xyz(i, 3)
  end do
  read(u, *)
  read(u, *)
  read(u, *)
  ncon = 0
  do i = 1, nelem
    read(u, *) idum, ncn(i)
    backspace(u)
    read(u, *) idum, ncn(i), (ncon(i, k), k = 1, ncn(i))
  end do
  close(u)

  allocate(cen(nelem, 3), ds(nelem), pnsz(nelem), dxyz(nelem, 6))
  call PANEL_GEOMETRY_GPU(nnode, xyz, nelem, ncn, ncon, xr, cen, ds, pnsz, dxyz)

  call GPU_INIT(ctx, 0, prec)
  call SET_MESH_MULTI_GPU(ctx, nelem, nnode, xyz, ncn, ncon, cen, ds, pnsz, dxyz, 1, [nelem])
  ! ISX / ISY = 1: the mesh is the half on the positive side of x = 0 / y = 0 (NSYS = 2)
  nsys = 1
  if (isx /= 0 .or. isy /= 0) then
    call SET_SYMMETRY_GPU(ctx, isx, isy)
    nsys = 2
  end if
  call SET_WAVES_MULTI_GPU(ctx, nbeta, beta, 1.0d0, [0.0d0, 0.0d0], 1000.0d0, 1)
  allocate(amas(6, 6), bdmp(6, 6), exfc(6, max(nbeta, 1)), rad(nelem, 6, nsys), dif(nelem, max(nbeta, 1), nsys))
  if (style == 1) allocate(amat(nelem, nelem, nsys), ipiv(nelem, nsys), brmat(nelem, 6, nsys), &
                           bdmat(nelem, nsys), mxpot(nelem, 7, nsys), dif_ref(nelem, max(nbeta, 1), nsys), &
                           chk(nelem, 6, nsys))
  do kf = 1, nfreq
    call CALWAVEPROPERTS_GPU(ctx, 3, freqs(kf), h, v, wk, w1, wvn)
    if (style == 1) then
      ! the reference's call sequence (HAMS_Prog.f90:629-724), names and arguments unchanged
      call CALGREEN_MULTI
      call ASSB_LEFT_MULTI(amat, ipiv, nelem, nsys)
      call ASSB_RBC_MULTI(brmat, nelem, nsys, 1, nelem)
      call RADIATION_SOLVER_MULTI(amat, brmat, ipiv, mxpot, nelem, nsys)
      rad = mxpot(:, 1:6, :)
      do ii = 1, nbeta
        call ASSB_DBC_MULTI(bdmat, nelem, nsys)
        call DIFFRACTION_SOLVER_MULTI(amat, bdmat, ipiv, mxpot, nelem, nsys)
        dif_ref(:, ii, :) = mxpot(:, 7, :)
      end do
      call FORCES_GPU(ctx, amas, bdmp, exfc)
      ! the same potentials through the *_GPU wrappers
      call RADIATION_SOLVER_GPU(ctx, chk)
      dmax = maxval(abs(rad - chk))
      call DIFFRACTION_SOLVER_GPU(ctx, dif)
      if (nbeta > 0) dmax = max(dmax, maxval(abs(dif(:, 1:nbeta, :) - dif_ref(:, 1:nbeta, :))))
      write(*, '(A)') 'MXPOT'
      write(*, '(ES25.16E3)') dmax
    else
      call CALGREEN_ASSB_LEFT_GPU(ctx, info)
      if (info /= 0) write(*, '(A,I0)') ' zero pivot, INFO = ', info
      call RADIATION_SOLVER_GPU(ctx, rad)
      call DIFFRACTION_SOLVER_GPU(ctx, dif)
      call FORCES_GPU(ctx, amas, bdmp, exfc)
    end if
    write(*, '(A)') 'W'
    write(*, '(ES25.16E3)') w1
    write(*, '(A)') 'AMAS'
    write(*, '(6ES25.16E3)') amas
    write(*, '(A)') 'BDMP'
    write(*, '(6ES25.16E3)') bdmp
    write(*, '(A)') 'EXFC'
    write(*, '(2ES25.16E3)') exfc(:, 1:max(nbeta, 1))
  end do
  call GPU_FINALIZE(ctx)
end program hamsx_pnl_solve
