! ISO_C_BINDING shim over the reference's multi-body table fill, element
! integration and force reductions.
!
! TEST INFRASTRUCTURE ONLY: compiled with the unmodified reference sources
! (CalGreenFuncMulti.f90, SingularIntgrMulti.f90, BodyIntgrMulti.f90,
! PatclVelct.f90, PotentWavForceMulti.f90 and the kernels they use) into
! oracle/_ref/ so that the tests can check the oracle's assembled matrices,
! right-hand sides and forces against the reference's own code. Nothing in the
! product library links or loads it.
!
! The reference's module state (PanelMesh_mod, WaveDyn_mod, Potentials_mod,
! Body_mod, HAMS_mod) is filled from the arguments; one symmetry system
! (NSYS = 1, ISX = ISY = 0), all arrays in the reference's Fortran layouts.
!
! Entry points (bind(C)):
!   ref_multi_set     geometry of all bodies (the *_MULTI_COMB arrays that
!                     BuildCombinedArrays, CalGreenFuncMulti.f90:53-106, packs)
!   ref_multi_freq    V, WK, W1, H, the 200 dispersion roots
!   ref_multi_calgreen -> CALGREEN_MULTI            (CalGreenFuncMulti.f90:112)
!   ref_multi_tables  the table it filled, RKBN/CGRN (stored at GRN_KIND)
!   ref_multi_left    AMAT by BODINT_MULTI_LEFT with the loop of ASSB_LEFT_MULTI
!                     (AssbMatxMulti.f90:53-108, without its ZGETRF)
!   ref_multi_rbc     BRMAT by RBC_MULTI_RIGHT with the loop of ASSB_RBC_MULTI
!                     (AssbMatxMulti.f90:124-181), body by body as HAMS_Prog.f90:633-642
!   ref_multi_dbc     BDMAT of one heading by DBC_MULTI_RIGHT / DINP with the loop
!                     of ASSB_DBC_MULTI (AssbMatxMulti.f90:187-251)
!   ref_multi_forces  RFORCE_MULTI / EFORCE_MULTI (PotentWavForceMulti.f90:138, :45)
!                     over every body pair / body as HAMS_Prog.f90:693-724 calls them,
!                     for given potentials; the records these routines write go to
!                     scratch files that are deleted when closed
module ref_shim_multi
   use iso_c_binding
   use HAMS_mod
   use Body_mod
   use Const_mod
   use WaveDyn_mod
   use PanelMesh_mod
   use Inerfs_mod
   use Potentials_mod
   use CalGreenFuncMulti, only: CALGREEN_MULTI
   use BodyIntgrMulti, only: BODINT_MULTI_LEFT, RBC_MULTI_RIGHT, DBC_MULTI_RIGHT
   use PotentWavForceMulti, only: RFORCE_MULTI, EFORCE_MULTI
   implicit none

   integer, allocatable :: bstart(:)     ! 1-based first panel of each body, bstart(nbody+1) = n + 1

contains

   subroutine ref_multi_set(nb, body_start, n, nnode, xyz, ncn, ncon, xyz_p, ds, pnsz, dxyz, hh, nthreads) &
         bind(C, name="ref_multi_set")
      integer(c_int), value :: nb, n, nnode, nthreads
      integer(c_int), intent(in) :: body_start(nb + 1), ncn(n), ncon(4, n)
      real(c_double), intent(in) :: xyz(3, nnode), xyz_p(3, n), ds(n), pnsz(n), dxyz(6, n)
      real(c_double), value :: hh
      integer :: k
      if (allocated(XYZ_GLOBAL_MULTI_COMB)) deallocate(XYZ_GLOBAL_MULTI_COMB)
      if (allocated(XYZ_GLOBAL_MULTI_COMB_P)) deallocate(XYZ_GLOBAL_MULTI_COMB_P)
      if (allocated(PNSZ_MULTI_COMB)) deallocate(PNSZ_MULTI_COMB)
      if (allocated(NCN_MULTI_COMB)) deallocate(NCN_MULTI_COMB)
      if (allocated(NCON_MULTI_COMB)) deallocate(NCON_MULTI_COMB)
      if (allocated(DS_MULTI_COMB)) deallocate(DS_MULTI_COMB)
      if (allocated(DXYZ_MULTI_COMB)) deallocate(DXYZ_MULTI_COMB)
      if (allocated(RKBN_MULTI_COMB)) deallocate(RKBN_MULTI_COMB)
      if (allocated(CGRN_MULTI_COMB)) deallocate(CGRN_MULTI_COMB)
      if (allocated(MXPOT_MULTI_COMB)) deallocate(MXPOT_MULTI_COMB)
      if (allocated(bstart)) deallocate(bstart)
      allocate(XYZ_GLOBAL_MULTI_COMB(nnode, 3), XYZ_GLOBAL_MULTI_COMB_P(n, 3), PNSZ_MULTI_COMB(n))
      allocate(NCN_MULTI_COMB(n), NCON_MULTI_COMB(n, 4), DS_MULTI_COMB(n), DXYZ_MULTI_COMB(n, 6))
      allocate(RKBN_MULTI_COMB(n, n, 1, 4), CGRN_MULTI_COMB(n, n, 1, 4), MXPOT_MULTI_COMB(n, 6 * nb + 1, 1))
      allocate(bstart(nb + 1))
      do k = 1, nnode
         XYZ_GLOBAL_MULTI_COMB(k, :) = xyz(:, k)
      end do
      do k = 1, n
         XYZ_GLOBAL_MULTI_COMB_P(k, :) = xyz_p(:, k)
         NCN_MULTI_COMB(k) = ncn(k)
         NCON_MULTI_COMB(k, :) = ncon(:, k)
         PNSZ_MULTI_COMB(k) = pnsz(k)
         DS_MULTI_COMB(k) = ds(k)
         DXYZ_MULTI_COMB(k, :) = dxyz(:, k)
      end do
      bstart = body_start + 1
      NBODY = nb
      TNELEM = n
      NSYS = 1
      ISYS = 0
      ISX = 0
      ISY = 0
      ISOL = 1
      IRSP = 0
      NTHREAD = nthreads
      H = hh
      XW = 0.d0
      AMP = 1.d0
   end subroutine

   subroutine ref_multi_freq(vv, wk_, w1_, wvn_) bind(C, name="ref_multi_freq")
      real(c_double), value :: vv, wk_, w1_
      real(c_double), intent(in) :: wvn_(NK)
      V = vv
      WK = wk_
      W1 = w1_
      TP = 2.d0 * PI / w1_
      WVN = wvn_
   end subroutine

   subroutine ref_multi_calgreen() bind(C, name="ref_multi_calgreen")
      call CALGREEN_MULTI
   end subroutine

   subroutine ref_multi_tables(n, rkbn, cgrn) bind(C, name="ref_multi_tables")
      integer(c_int), value :: n
      real(c_double), intent(out) :: rkbn(n, n, 4)
      complex(c_double_complex), intent(out) :: cgrn(n, n, 4)
      rkbn = real(RKBN_MULTI_COMB(:, :, 1, :), 8)
      cgrn = cmplx(CGRN_MULTI_COMB(:, :, 1, :), kind=8)
   end subroutine

   integer function near_flag(iel, jel)
      integer, intent(in) :: iel, jel
      real(8) :: dist
      dist = sqrt((XYZ_GLOBAL_MULTI_COMB_P(iel, 1) - XYZ_GLOBAL_MULTI_COMB_P(jel, 1))**2 &
                  + (XYZ_GLOBAL_MULTI_COMB_P(iel, 2) - XYZ_GLOBAL_MULTI_COMB_P(jel, 2))**2 &
                  + (XYZ_GLOBAL_MULTI_COMB_P(iel, 3) - XYZ_GLOBAL_MULTI_COMB_P(jel, 3))**2)
      if (dist .le. 50.d0 * PNSZ_MULTI_COMB(jel)) then
         near_flag = 1
      else
         near_flag = 0
      end if
   end function

   subroutine ref_multi_left(n, amat) bind(C, name="ref_multi_left")
      integer(c_int), value :: n
      complex(c_double_complex), intent(out) :: amat(n, n)
      integer :: iel, jel
      complex(8) :: tindp(4)
      amat = (0.d0, 0.d0)
      do iel = 1, n
         amat(iel, iel) = 2.d0 * PI
         do jel = 1, n
            tindp = (0.d0, 0.d0)
            call BODINT_MULTI_LEFT(1, iel, jel, tindp, near_flag(iel, jel))
            amat(iel, jel) = amat(iel, jel) + RXY_MULTI(1, 1) * tindp(1)
         end do
      end do
   end subroutine

   subroutine ref_multi_rbc(n, nb, brmat) bind(C, name="ref_multi_rbc")
      integer(c_int), value :: n, nb
      complex(c_double_complex), intent(out) :: brmat(n, 6 * nb)
      integer :: iel, jel, b, md
      complex(8) :: tinrd(4, 6, 4), btmp(6)
      brmat = (0.d0, 0.d0)
      do b = 1, nb
         do iel = 1, n
            btmp = (0.d0, 0.d0)
            do jel = bstart(b), bstart(b + 1) - 1
               tinrd = (0.d0, 0.d0)
               call RBC_MULTI_RIGHT(1, iel, jel, tinrd, near_flag(iel, jel))
               do md = 1, 6
                  btmp(md) = btmp(md) + tinrd(1, md, 1)
               end do
            end do
            do md = 1, 6
               brmat(iel, 6 * (b - 1) + md) = brmat(iel, 6 * (b - 1) + md) + RXY_MULTI(1, 1) * btmp(md)
            end do
         end do
      end do
   end subroutine

   subroutine ref_multi_dbc(n, beta_, bdmat) bind(C, name="ref_multi_dbc")
      integer(c_int), value :: n
      real(c_double), value :: beta_
      complex(c_double_complex), intent(out) :: bdmat(n)
      integer :: iel, jel
      complex(8) :: tinrd(4, 4), btmp
      BETA = beta_
      do iel = 1, n
         btmp = (0.d0, 0.d0)
         do jel = 1, n
            tinrd = (0.d0, 0.d0)
            call DBC_MULTI_RIGHT(1, iel, jel, tinrd, near_flag(iel, jel))
            btmp = btmp + tinrd(1, 1)
         end do
         bdmat(iel) = RXY_MULTI(1, 1) * btmp
      end do
   end subroutine

   subroutine open_sinks()
      integer :: u
      do u = 21, 26
         open(u, status='scratch')
      end do
      do u = 31, 36
         open(u, status='scratch')
      end do
      do u = 41, 46
         open(u, status='scratch')
      end do
      open(61, status='scratch')
      open(62, status='scratch')
   end subroutine

   subroutine close_sinks()
      integer :: u
      do u = 21, 26
         close(u)
      end do
      do u = 31, 36
         close(u)
      end do
      do u = 41, 46
         close(u)
      end do
      close(61)
      close(62)
   end subroutine

   ! mxpot(n, 6 nb + 1): the radiation potentials and the diffraction potential
   ! of heading beta_; amas, bdmp (6nb x 6nb), exfc (6nb)
   subroutine ref_multi_forces(n, nb, mxpot, beta_, amas, bdmp, exfc) bind(C, name="ref_multi_forces")
      integer(c_int), value :: n, nb
      complex(c_double_complex), intent(in) :: mxpot(n, 6 * nb + 1)
      real(c_double), value :: beta_
      real(c_double), intent(out) :: amas(6 * nb, 6 * nb), bdmp(6 * nb, 6 * nb)
      complex(c_double_complex), intent(out) :: exfc(6 * nb)
      integer :: bm, bn
      real(8) :: a6(6, 6), b6(6, 6)
      complex(8) :: e6(6)
      MXPOT_MULTI_COMB(:, :, 1) = mxpot
      BETA = beta_
      OUFR = W1
      call open_sinks()
      do bm = 1, nb
         do bn = 1, nb
            call RFORCE_MULTI(WK, W1, TP, a6, b6, bstart(bm), bstart(bm + 1) - 1, bn, bm, bn)
            amas(6 * (bm - 1) + 1:6 * bm, 6 * (bn - 1) + 1:6 * bn) = a6
            bdmp(6 * (bm - 1) + 1:6 * bm, 6 * (bn - 1) + 1:6 * bn) = b6
         end do
         call EFORCE_MULTI(WK, W1, TP, beta_, AMP, e6, bstart(bm), bstart(bm + 1) - 1, nb, bm)
         exfc(6 * (bm - 1) + 1:6 * bm) = e6
      end do
      call close_sinks()
   end subroutine

end module ref_shim_multi
