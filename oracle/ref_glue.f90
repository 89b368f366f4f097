! ref_glue.f90 -- C-callable entry points onto HAMS-MREL's own Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's source
! files where they lie (WavDynMods, FinGrnExtSubs, FinGreen3D, InfGreen_Appr,
! SingularIntgrMulti, WavDynSubs, PatclVelct, NormalProcess, ImplementSubs,
! CalGreenFuncMulti)
! together with this file into oracle/_ref/libhams_ref.so.  It is used only to
! write and check golden vectors for the C oracle (tests/golden/make_golden.py,
! tests/test_oracle_golden.py).  Nothing in the product links it.
!
! Each wrapper only moves arguments into the reference's module state and
! calls the reference routine unchanged.

module ref_glue
   use iso_c_binding
   use Const_mod
   use WaveDyn_mod
   use PanelMesh_mod
   use INFG3D_Open
   use FinGreen3D_Open
   use FinGrnExtSubs
   use SingularIntgrMulti
   use PatcVelct
   use NormalProcess
   use ImplementSubs
   use HAMS_mod
   use Inerfs_mod
   use Potentials_mod
   use CalGreenFuncMulti
   implicit none

contains

   subroutine ref_infgreen3d(xf, xp, yf, yp, zf, zp, vv, flag, grn) bind(C, name='ref_infgreen3d')
      real(c_double), value :: xf, xp, yf, yp, zf, zp, vv
      integer(c_int), value :: flag
      complex(c_double_complex), intent(out) :: grn(4)
      complex*16 :: g(4)
      integer :: fl
      fl = flag
      call INFGREEN3D(xf, xp, yf, yp, zf, zp, vv, g, fl)
      grn = g
   end subroutine

   subroutine ref_fingreen3d(xf, xp, yf, yp, zf, zp, vv, wvn_, nk_, hh, flag, grn) &
         bind(C, name='ref_fingreen3d')
      real(c_double), value :: xf, xp, yf, yp, zf, zp, vv, hh
      integer(c_int), value :: nk_, flag
      real(c_double), intent(in) :: wvn_(nk_)
      complex(c_double_complex), intent(out) :: grn(4)
      complex*16 :: g(4)
      real*8 :: w(nk_)
      integer :: fl, n
      fl = flag
      n = nk_
      w = wvn_
      call FINGREEN3D(xf, xp, yf, yp, zf, zp, vv, w, n, hh, g, fl)
      grn = g
   end subroutine

   subroutine ref_dispersion(w, hh, nrt, wvn_) bind(C, name='ref_dispersion')
      real(c_double), value :: w, hh
      integer(c_int), value :: nrt
      real(c_double), intent(out) :: wvn_(nrt)
      real*8 :: out(nrt)
      integer :: n
      n = nrt
      call DISPERSION(out, n, w, hh)
      wvn_ = out
   end subroutine

   ! CalWaveProperts for one frequency value of input type inft (ImplementSubs.f90:40-151)
   subroutine ref_wave_properties(inft_, val, hh, v_out, wk_out, w1_out, wvn_out) &
         bind(C, name='ref_wave_properties')
      integer(c_int), value :: inft_
      real(c_double), value :: val, hh
      real(c_double), intent(out) :: v_out, wk_out, w1_out, wvn_out(200)
      INFT = inft_
      OUFT = 3
      SYBO = 0
      H = hh
      if (allocated(WVNB)) deallocate(WVNB)
      if (allocated(WVFQ)) deallocate(WVFQ)
      allocate(WVNB(1), WVFQ(1))
      WVNB(1) = val
      WVN = 0.d0
      call CalWaveProperts(1)
      v_out = V
      wk_out = WK
      w1_out = W1
      wvn_out = WVN
   end subroutine

   subroutine ref_ik01a(x, k0, k1) bind(C, name='ref_ik01a')
      real(c_double), value :: x
      real(c_double), intent(out) :: k0, k1
      real*8 :: a, b
      call IK01A(x, a, b)
      k0 = a
      k1 = b
   end subroutine

   subroutine ref_jy01b(x, j0, j1, y0, y1) bind(C, name='ref_jy01b')
      real(c_double), value :: x
      real(c_double), intent(out) :: j0, j1, y0, y1
      real*8 :: a, b, c, d
      call JY01B(x, a, b, c, d)
      j0 = a
      j1 = b
      y0 = c
      y1 = d
   end subroutine

   function ref_error(x) result(r) bind(C, name='ref_error')
      real(c_double), value :: x
      real(c_double) :: r
      real*8 :: e
      call ERROR(x, e)
      r = e
   end function

   function ref_e1xa(x) result(r) bind(C, name='ref_e1xa')
      real(c_double), value :: x
      real(c_double) :: r
      real*8 :: e
      call E1XA(x, e)
      r = e
   end function

   function ref_enxa(n, x) result(r) bind(C, name='ref_enxa')
      integer(c_int), value :: n
      real(c_double), value :: x
      real(c_double) :: r
      real*8 :: e
      integer :: nn
      nn = n
      call ENXA(nn, x, e)
      r = e
   end function

   function ref_eix(x) result(r) bind(C, name='ref_eix')
      real(c_double), value :: x
      real(c_double) :: r
      real*8 :: e
      call EIX(x, e)
      r = e
   end function

   function ref_limes(n, s) result(r) bind(C, name='ref_limes')
      integer(c_int), value :: n
      real(c_double), intent(inout) :: s(0:n)
      real(c_double) :: r
      real*8 :: ss(0:n), rl, er
      integer :: ip, kp, nn
      nn = n
      ss = s
      call Limes(nn, ss, rl, ip, kp, er)
      s = ss
      r = rl
   end function

   function ref_struve_h0(x) result(r) bind(C, name='ref_struve_h0')
      real(c_double), value :: x
      real(c_double) :: r
      r = StruveH0(x)
   end function

   function ref_struve_h1(x) result(r) bind(C, name='ref_struve_h1')
      real(c_double), value :: x
      real(c_double) :: r
      r = StruveH1(x)
   end function

   function ref_bessel_j0(x) result(r) bind(C, name='ref_bessel_j0')
      real(c_double), value :: x
      real(c_double) :: r
      r = BesselJ0(x)
   end function

   function ref_bessel_j1(x) result(r) bind(C, name='ref_bessel_j1')
      real(c_double), value :: x
      real(c_double) :: r
      r = BesselJ1(x)
   end function

   ! Mesh state for SGLINTBD_MULTI_*: node_xyz(nnode,3), ncn(nelem), ncon(nelem,4), cen(nelem,3)
   subroutine ref_set_mesh(nnode, xyz, nelem, ncn_, ncon_, cen) bind(C, name='ref_set_mesh')
      integer(c_int), value :: nnode, nelem
      real(c_double), intent(in) :: xyz(nnode, 3), cen(nelem, 3)
      integer(c_int), intent(in) :: ncn_(nelem), ncon_(nelem, 4)
      if (allocated(XYZ_GLOBAL_MULTI_COMB)) deallocate(XYZ_GLOBAL_MULTI_COMB)
      if (allocated(XYZ_GLOBAL_MULTI_COMB_P)) deallocate(XYZ_GLOBAL_MULTI_COMB_P)
      if (allocated(NCN_MULTI_COMB)) deallocate(NCN_MULTI_COMB)
      if (allocated(NCON_MULTI_COMB)) deallocate(NCON_MULTI_COMB)
      allocate(XYZ_GLOBAL_MULTI_COMB(nnode, 3), XYZ_GLOBAL_MULTI_COMB_P(nelem, 3))
      allocate(NCN_MULTI_COMB(nelem), NCON_MULTI_COMB(nelem, 4))
      XYZ_GLOBAL_MULTI_COMB = xyz
      XYZ_GLOBAL_MULTI_COMB_P = cen
      NCN_MULTI_COMB = ncn_
      NCON_MULTI_COMB = ncon_
      ISX = 0
      ISY = 0
   end subroutine

   ! The reference's own pair loop, CALGREEN_MULTI (CalGreenFuncMulti.f90:112-183,
   ! OpenMP over field panels), on the mesh of ref_set_mesh with panel sizes
   ! pnsz_: NSYS = 1, nthreads threads, the frequency (vv, wvn_, hh).  For the
   ! CPU baseline's timing.  Returns the sum of |CGRN(:,:,1,1)| as a checksum.
   function ref_calgreen(nthreads, nelem, pnsz_, vv, wvn_, nk_, hh) result(chk) bind(C, name='ref_calgreen')
      integer(c_int), value :: nthreads, nelem, nk_
      real(c_double), intent(in) :: pnsz_(nelem), wvn_(nk_)
      real(c_double), value :: vv, hh
      real(c_double) :: chk
      TNELEM = nelem
      NSYS = 1
      ISX = 0
      ISY = 0
      NTHREAD = nthreads
      V = vv
      H = hh
      WVN = 0.d0
      WVN(1:min(nk_, NK)) = wvn_(1:min(nk_, NK))
      if (allocated(PNSZ_MULTI_COMB)) deallocate(PNSZ_MULTI_COMB)
      allocate(PNSZ_MULTI_COMB(nelem))
      PNSZ_MULTI_COMB = pnsz_
      if (allocated(RKBN_MULTI_COMB)) deallocate(RKBN_MULTI_COMB)
      if (allocated(CGRN_MULTI_COMB)) deallocate(CGRN_MULTI_COMB)
      allocate(RKBN_MULTI_COMB(nelem, nelem, 1, 4), CGRN_MULTI_COMB(nelem, nelem, 1, 4))
      call CALGREEN_MULTI
      chk = sum(abs(CGRN_MULTI_COMB(:, :, 1, 1)))
      deallocate(RKBN_MULTI_COMB, CGRN_MULTI_COMB)
   end function

   ! Newman integral of source panel jel at the centroid of field panel iel (1-based)
   subroutine ref_sglint(iel, jel, sij, dij) bind(C, name='ref_sglint')
      integer(c_int), value :: iel, jel
      real(c_double), intent(out) :: sij, dij(3)
      real*8 :: s, d(3)
      integer :: i, j
      i = iel
      j = jel
      if (NCN_MULTI_COMB(j) .eq. 3) then
         call SGLINTBD_MULTI_TRI(1, i, j, s, d, 1)
      else
         call SGLINTBD_MULTI_QUAD(1, i, j, s, d, 1)
      end if
      sij = s
      dij = d
   end subroutine

   ! Newman integral of source panel jel (1-based) at an arbitrary field point
   ! xf: SGLINTBD_MULTI_TRI2 / _QUAD2 with IRR = 1 (hull panels), rk = (SIJ, DIJ)
   subroutine ref_sglint2(xf, jel, rk) bind(C, name='ref_sglint2')
      real(c_double), intent(in) :: xf(3)
      integer(c_int), value :: jel
      real(c_double), intent(out) :: rk(4)
      real*8 :: x(3), r(4)
      integer :: j
      x = xf
      j = jel
      if (NCN_MULTI_COMB(j) .eq. 3) then
         call SGLINTBD_MULTI_TRI2(x, j, r, 1)
      else
         call SGLINTBD_MULTI_QUAD2(x, j, r, 1)
      end if
      rk = r
   end subroutine

   subroutine ref_dinp(x, y, z, xw_, yw_, beta_, vv, w1_, wk_, hh, amp_, dpo) bind(C, name='ref_dinp')
      real(c_double), value :: x, y, z, xw_, yw_, beta_, vv, w1_, wk_, hh, amp_
      complex(c_double_complex), intent(out) :: dpo(3)
      complex*16 :: a, b, c
      V = vv
      W1 = w1_
      WK = wk_
      H = hh
      AMP = amp_
      call DINP(x, y, z, xw_, yw_, beta_, a, b, c)
      dpo(1) = a
      dpo(2) = b
      dpo(3) = c
   end subroutine

   subroutine ref_vinp(x, y, z, xw_, yw_, beta_, vv, w1_, wk_, hh, amp_, r) bind(C, name='ref_vinp')
      real(c_double), value :: x, y, z, xw_, yw_, beta_, vv, w1_, wk_, hh, amp_
      complex(c_double_complex), intent(out) :: r
      V = vv
      W1 = w1_
      WK = wk_
      H = hh
      AMP = amp_
      r = VINP(x, y, z, xw_, yw_, beta_)
   end subroutine

   ! Panel centroid, area, size and normals (NormalProcess.f90) for one body
   subroutine ref_geometry(nnode, xyz, nelem, ncn_, ncon_, xr_, cen, area, pnsz_, dxyz) &
         bind(C, name='ref_geometry')
      integer(c_int), value :: nnode, nelem
      real(c_double), intent(in) :: xyz(nnode, 3), xr_(3)
      integer(c_int), intent(in) :: ncn_(nelem), ncon_(nelem, 4)
      real(c_double), intent(out) :: cen(nelem, 3), area(nelem), pnsz_(nelem), dxyz(nelem, 6)
      real*8 :: x(nnode, 3), c(nelem, 3), a(nelem), p(nelem), d(nelem, 6), rg(3)
      integer :: nc(nelem), nn(nelem, 4), nt, ne
      x = xyz
      nc = ncn_
      nn = ncon_
      rg = xr_
      nt = nnode
      ne = nelem
      call CalPanelCentre(x, nt, ne, nc, nn, c)
      call CalPanelArea(x, nt, ne, nc, nn, a)
      call CalPanelChartLength(x, nt, ne, nc, nn, p)
      call CalTransNormals(x, nt, ne, nc, nn, d)
      call CalRotNormals(rg, c, d, ne)
      cen = c
      area = a
      pnsz_ = p
      dxyz = d
   end subroutine

end module ref_glue
