! ISO_C_BINDING shim over the reference's own hot-path kernels.
!
! TEST INFRASTRUCTURE ONLY: this file is compiled together with the unmodified
! reference sources (FinGreen3D.f90, FinGrnExtSubs.f90, InfGreen_Appr.f90,
! SingularIntgr.f90, WavDynSubs.f90, WavDynMods.f90) into oracle/_ref/ so that the
! tests can call the reference kernels element by element. Nothing in the product
! library links or loads it.
!
! Entry points (all bind(C), all scalar arguments by value):
!   ref_fingreen3d   -> FINGREEN3D       (src/FinGreen3D.f90:81)
!   ref_fingreen3d_r -> FINGREEN3D_R     (src/FinGreen3D.f90:160)
!   ref_infgreen3d   -> INFGREEN3D       (src/InfGreen_Appr.f90:104)
!   ref_dispersion   -> DISPERSION       (src/WavDynSubs.f90:40)
!   ref_sglintbd     -> SGLINTBD_QUAD/TRI(src/SingularIntgr.f90:252/:44)
!   ref_ik01a, ref_jy01b, ref_error, ref_e1xa, ref_enxa, ref_eix, ref_limes
!                    -> FinGrnExtSubs.f90 routines of the same names

module ref_shim
   use iso_c_binding
   use FinGrnExtSubs
   use FinGreen3D_Open
   use INFG3D_Open, only: INFGREEN3D
   use PanelMesh_mod
   use Inerfs_mod
   use SingularIntgr
   implicit none

   interface
      subroutine DISPERSION(WVN, NRT, W, H)
         integer, intent(in) :: NRT
         real(8), intent(in) :: W, H
         real(8), intent(out) :: WVN(1:NRT)
      end subroutine DISPERSION
   end interface

contains

   subroutine ref_fingreen3d(xf, xp, yf, yp, zf, zp, v, wvn, nk, h, grn, flag) bind(C, name="ref_fingreen3d")
      real(c_double), value :: xf, xp, yf, yp, zf, zp, v, h
      integer(c_int), value :: nk, flag
      real(c_double), intent(in) :: wvn(nk)
      complex(c_double_complex), intent(out) :: grn(4)
      call FINGREEN3D(xf, xp, yf, yp, zf, zp, v, wvn, int(nk), h, grn, int(flag))
   end subroutine

   subroutine ref_fingreen3d_r(rr, zf, zp, v, wvn, nk, h, grn, tag) bind(C, name="ref_fingreen3d_r")
      real(c_double), value :: rr, zf, zp, v, h
      integer(c_int), value :: nk, tag
      real(c_double), intent(in) :: wvn(nk)
      complex(c_double_complex), intent(out) :: grn(3)
      call FINGREEN3D_R(rr, zf, zp, v, wvn, int(nk), h, grn, int(tag))
   end subroutine

   subroutine ref_infgreen3d(xf, xp, yf, yp, zf, zp, v, grn, flag) bind(C, name="ref_infgreen3d")
      real(c_double), value :: xf, xp, yf, yp, zf, zp, v
      integer(c_int), value :: flag
      complex(c_double_complex), intent(out) :: grn(4)
      call INFGREEN3D(xf, xp, yf, yp, zf, zp, v, grn, int(flag))
   end subroutine

   subroutine ref_dispersion(wvn, nrt, w, h) bind(C, name="ref_dispersion")
      integer(c_int), value :: nrt
      real(c_double), value :: w, h
      real(c_double), intent(out) :: wvn(nrt)
      call DISPERSION(wvn, int(nrt), w, h)
   end subroutine

   ! Rankine panel integral of one source panel (ncn = 3 or 4 vertices, given
   ! directly as coordinates) seen from field point xf. Sets up the module state
   ! the reference routine reads (PanelMesh_mod XYZ/NCON/NCN/XYZ_P) for a
   ! one-panel mesh and calls it with IS=1, IEL=1, JEL=1, IRR=1, no symmetry.
   subroutine ref_sglintbd(ncn_in, verts, xf, sij, dij) bind(C, name="ref_sglintbd")
      integer(c_int), value :: ncn_in
      real(c_double), intent(in) :: verts(3, 4), xf(3)
      real(c_double), intent(out) :: sij, dij(3)
      integer :: k
      if (allocated(XYZ)) deallocate(XYZ)
      if (allocated(NCON)) deallocate(NCON)
      if (allocated(NCN)) deallocate(NCN)
      if (allocated(XYZ_P)) deallocate(XYZ_P)
      allocate(XYZ(4, 3), NCON(1, 4), NCN(1), XYZ_P(1, 3))
      do k = 1, 4
         XYZ(k, :) = verts(:, k)
         NCON(1, k) = k
      end do
      NCN(1) = ncn_in
      XYZ_P(1, :) = xf(:)
      ISX = 0
      ISY = 0
      if (ncn_in .eq. 3) then
         call SGLINTBD_TRI(1, 1, 1, sij, dij, 1)
      else
         call SGLINTBD_QUAD(1, 1, 1, sij, dij, 1)
      end if
   end subroutine

   subroutine ref_ik01a(x, bk0, bk1) bind(C, name="ref_ik01a")
      real(c_double), value :: x
      real(c_double), intent(out) :: bk0, bk1
      call IK01A(x, bk0, bk1)
   end subroutine

   subroutine ref_jy01b(x, bj0, bj1, by0, by1) bind(C, name="ref_jy01b")
      real(c_double), value :: x
      real(c_double), intent(out) :: bj0, bj1, by0, by1
      call JY01B(x, bj0, bj1, by0, by1)
   end subroutine

   subroutine ref_error(x, err) bind(C, name="ref_error")
      real(c_double), value :: x
      real(c_double), intent(out) :: err
      call ERROR(x, err)
   end subroutine

   subroutine ref_e1xa(x, e1) bind(C, name="ref_e1xa")
      real(c_double), value :: x
      real(c_double), intent(out) :: e1
      call E1XA(x, e1)
   end subroutine

   subroutine ref_enxa(n, x, en) bind(C, name="ref_enxa")
      integer(c_int), value :: n
      real(c_double), value :: x
      real(c_double), intent(out) :: en
      call ENXA(int(n), x, en)
   end subroutine

   subroutine ref_eix(x, ei) bind(C, name="ref_eix")
      real(c_double), value :: x
      real(c_double), intent(out) :: ei
      call EIX(x, ei)
   end subroutine

   ! Wynn epsilon limit of the partial sums s(0:n); s is overwritten.
   subroutine ref_limes(n, s, rlim) bind(C, name="ref_limes")
      integer(c_int), value :: n
      real(c_double), intent(inout) :: s(0:n)
      real(c_double), intent(out) :: rlim
      integer :: ip, kp
      real(8) :: err
      call Limes(int(n), s, rlim, ip, kp, err)
   end subroutine

end module ref_shim
