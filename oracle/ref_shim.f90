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
!   ref_set_mesh + ref_pairs_omp
!                    -> the body of the CALGREEN table-fill loop
!                       (CalGreenFunc.f90:58-115) over a list of (field panel,
!                       source panel) pairs, OpenMP-parallel like the reference;
!                       used to time the reference's CPU kernels

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

   ! mesh in the reference's module arrays (PanelMesh_mod), as ReadPanelMesh sets them
   subroutine ref_set_mesh(nelem_in, ntnd_in, xyz_in, ncn_in, ncon_in, xyz_p_in, pnsz_in) bind(C, name="ref_set_mesh")
      integer(c_int), value :: nelem_in, ntnd_in
      real(c_double), intent(in) :: xyz_in(3, ntnd_in), xyz_p_in(3, nelem_in), pnsz_in(nelem_in)
      integer(c_int), intent(in) :: ncn_in(nelem_in), ncon_in(4, nelem_in)
      integer :: k
      if (allocated(XYZ)) deallocate(XYZ)
      if (allocated(NCON)) deallocate(NCON)
      if (allocated(NCN)) deallocate(NCN)
      if (allocated(XYZ_P)) deallocate(XYZ_P)
      if (allocated(PNSZ)) deallocate(PNSZ)
      allocate(XYZ(ntnd_in, 3), NCON(nelem_in, 4), NCN(nelem_in), XYZ_P(nelem_in, 3), PNSZ(nelem_in))
      do k = 1, 3
         XYZ(:, k) = xyz_in(k, :)
         XYZ_P(:, k) = xyz_p_in(k, :)
      end do
      do k = 1, 4
         NCON(:, k) = ncon_in(k, :)
      end do
      NCN = ncn_in
      PNSZ = pnsz_in
      NELEM = nelem_in
      NTND = ntnd_in
      ISX = 0
      ISY = 0
      NSYS = 1
   end subroutine

   ! the CALGREEN loop body for np listed pairs (1-based panel numbers); returns
   ! the sum of all GRN(1) values so nothing is optimised away
   subroutine ref_pairs_omp(np, iel, jel, h, v, wvn, nk, nthreads, csum) bind(C, name="ref_pairs_omp")
      integer(c_int), value :: np, nk, nthreads
      integer(c_int), intent(in) :: iel(np), jel(np)
      real(c_double), value :: h, v
      real(c_double), intent(in) :: wvn(nk)
      complex(c_double_complex), intent(out) :: csum
      integer :: k, i, j, flag
      real(8) :: dist, sij, dij(3)
      complex(8) :: grn(4), acc
      acc = (0.0d0, 0.0d0)
!$OMP PARALLEL DO NUM_THREADS(nthreads) SCHEDULE(DYNAMIC,16) PRIVATE(k,i,j,flag,dist,sij,dij,grn) REDUCTION(+:acc)
      do k = 1, np
         i = iel(k)
         j = jel(k)
         dist = sqrt((XYZ_P(i, 1) - XYZ_P(j, 1))**2 + (XYZ_P(i, 2) - XYZ_P(j, 2))**2 + (XYZ_P(i, 3) - XYZ_P(j, 3))**2)
         if (dist .le. 50.d0 * PNSZ(j)) then
            flag = 1
         else
            flag = 0
         end if
         if (NCN(j) .eq. 3) then
            call SGLINTBD_TRI(1, i, j, sij, dij, 1)
         else
            call SGLINTBD_QUAD(1, i, j, sij, dij, 1)
         end if
         if (h .lt. 0.d0) then
            call INFGREEN3D(XYZ_P(j, 1), XYZ_P(i, 1), XYZ_P(j, 2), XYZ_P(i, 2), XYZ_P(j, 3), XYZ_P(i, 3), v, grn, flag)
         else
            call FINGREEN3D(XYZ_P(j, 1), XYZ_P(i, 1), XYZ_P(j, 2), XYZ_P(i, 2), XYZ_P(j, 3), XYZ_P(i, 3), v, wvn, &
                            int(nk), h, grn, flag)
         end if
         acc = acc + grn(1) + sij
      end do
!$OMP END PARALLEL DO
      csum = acc
   end subroutine

end module ref_shim
