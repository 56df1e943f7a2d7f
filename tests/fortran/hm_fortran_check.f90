! A Fortran caller of the drop-in layer (hams-mrel_amd/fortran/hm_bem.f90) that
! calls the GPU path through the reference's own subroutine names and argument
! lists, as a HAMS-MREL driver would (HAMS_Prog.f90:271-300 for one body with
! irregular-frequency removal, :613-724 for several bodies). Test program of
! tests/test_gpu_fortran.py: it reads one problem from the stream file named by
! its first argument and writes what the calls returned to the second.
!
! input (stream, native endianness): int32 n, ni, nbody, ntnd, nbeta, npts, nsgl;
! int32 body_start(nbody+1), ncn(n), ncon(n,4), iel(nsgl), jel(nsgl), irr(nsgl);
! real64 xyz(ntnd,3), xyz_p(n,3), ds(n), pnsz(n), dxyz_p(n,6), ixyz_p(max(ni,1),3),
! h, beta(nbeta), rho, g, g_disp, amp, xw(2), w1, pts(npts,6) (xf,xp,yf,yp,zf,zp),
! flags(npts) as real64, lua(4,4) complex128 (a test matrix for ZGETRF/ZGETRS)
!
! output (stream): complex128 grn(4,npts); real64 sij(nsgl), dij(3,nsgl);
! complex128 mxpot(n, 6 nbody + nbeta) (radiation columns, then the diffraction
! potential of each heading); complex128 x(4) = lua \ (1,2,3,4)^T; int32 info;
! then the H-matrix route of ISOLV = 2 on the test kernel K(i,i) = 4 + i,
! K(i,j) = exp(0.3 i r_ij) / r_ij over xyz_p: complex128 y(n) = HMATVEC(K e),
! e(j) = 1 + i j / n; complex128 z(n,2) = ZITER_SOLVE_MULTI(2, ...) of
! K z = (e, conj(e)); int32 info of that solve
program hm_fortran_check
   use hm_bem
   use iso_c_binding, only: c_int
   implicit none
   integer :: n, ni, nbody, ntnd, nbeta, npts, nsgl, k, b, ih, m, tn, info, u, i, j, hinfo
   integer, allocatable :: body_start(:), ncn(:), ncon(:, :), iel(:), jel(:), irr(:), ipiv(:, :), ip4(:)
   real(8), allocatable :: xyz(:, :), xyz_p(:, :), ds(:), pnsz(:), dxyz_p(:, :), ixyz_p(:, :), beta_in(:), pts(:, :)
   real(8), allocatable :: flags(:), sij(:), dij(:, :), wvn(:)
   real(8) :: h, rho, g, g_disp, amp, xw(2), w1, v
   complex(8), allocatable :: grn(:, :), amat(:, :, :), cmat(:, :, :), brmat(:, :, :), drmat(:, :, :)
   complex(8), allocatable :: bdmat(:, :), ddmat(:, :), mxpot(:, :, :), out(:, :)
   complex(8) :: lua(4, 4), x(4)
   complex(8), allocatable :: kmat(:, :), hb(:, :), hz(:, :), hy(:)
   integer, allocatable :: bsizes(:)
   real(8) :: r
   character(len=512) :: fin, fout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
   read (u) n, ni, nbody, ntnd, nbeta, npts, nsgl
   allocate (body_start(nbody + 1), ncn(n), ncon(n, 4), iel(nsgl), jel(nsgl), irr(nsgl))
   read (u) body_start, ncn, ncon, iel, jel, irr
   allocate (xyz(ntnd, 3), xyz_p(n, 3), ds(n), pnsz(n), dxyz_p(n, 6), ixyz_p(max(ni, 1), 3), beta_in(nbeta))
   allocate (pts(npts, 6), flags(npts))
   read (u) xyz, xyz_p, ds, pnsz, dxyz_p, ixyz_p, h, beta_in, rho, g, g_disp, amp, xw, w1, pts, flags, lua
   close (u)

   call HM_SETUP(nbody, body_start, ntnd, xyz, ncn, ncon, xyz_p, ds, pnsz, dxyz_p, ni, ixyz_p, h, nbeta, beta_in, &
                 rho, g, g_disp, amp, xw)
   call HM_SET_FREQUENCY_F(w1)

   ! element kernels
   allocate (grn(4, npts), wvn(200))
   v = w1 * w1 / g
   if (h > 0.d0) then
      call HM_CHECK(hm_dispersion_f(wvn, w1, h, g_disp), 'dispersion')
   else
      wvn = 0.d0
   end if
   do k = 1, npts
      if (h > 0.d0) then
         call FINGREEN3D(pts(k, 1), pts(k, 2), pts(k, 3), pts(k, 4), pts(k, 5), pts(k, 6), v, wvn, 200, h, &
                         grn(:, k), nint(flags(k)))
      else
         call INFGREEN3D(pts(k, 1), pts(k, 2), pts(k, 3), pts(k, 4), pts(k, 5), pts(k, 6), v, grn(:, k), &
                         nint(flags(k)))
      end if
   end do
   allocate (sij(nsgl), dij(3, nsgl))
   do k = 1, nsgl
      if (ncn(jel(k)) == 4) then
         call SGLINTBD_QUAD(1, iel(k), jel(k), sij(k), dij(:, k), irr(k))
      else
         call SGLINTBD_TRI(1, iel(k), jel(k), sij(k), dij(:, k), irr(k))
      end if
   end do

   ! the frequency's chain in the reference driver's order
   m = 6 * nbody
   allocate (ipiv(n, 1), mxpot(n, m + 1, 1), out(n, m + nbeta))
   if (ni == 0) then
      allocate (amat(n, n, 1), brmat(n, m, 1), bdmat(n, 1))
      call CALGREEN_MULTI()
      call ASSB_LEFT_MULTI(amat, ipiv, n, 1)
      brmat = (0.d0, 0.d0)
      do b = 1, nbody
         call ASSB_RBC_MULTI(brmat(:, 6 * (b - 1) + 1:6 * b, :), n, 1, body_start(b) + 1, body_start(b + 1))
      end do
      call RADIATION_SOLVER_MULTI(amat, brmat, ipiv, mxpot, n, 1)
      out(:, 1:m) = mxpot(:, 1:m, 1)
      do ih = 1, nbeta
         BETA = beta_in(ih)
         call ASSB_DBC_MULTI(bdmat, n, 1)
         call DIFFRACTION_SOLVER_MULTI(amat, bdmat, ipiv, mxpot, n, 1)
         out(:, m + ih) = mxpot(:, m + 1, 1)
      end do
   else
      tn = n + ni
      allocate (amat(tn, tn, 1), cmat(n, n, 1), brmat(tn, 6, 1), drmat(n, 6, 1), bdmat(tn, 1), ddmat(n, 1))
      call CALGREEN_IRR()
      call ASSB_LEFT_IRR(amat, cmat, ipiv, n, tn, 1)
      call ASSB_RBC_IRR(brmat, drmat, amat, n, tn, 1)
      call RADIATION_SOLVER_IRR(cmat, drmat, ipiv, mxpot, n, 1)
      out(:, 1:m) = mxpot(:, 1:m, 1)
      do ih = 1, nbeta
         BETA = beta_in(ih)
         call ASSB_DBC_IRR(bdmat, ddmat, amat, n, tn, 1)
         call DIFFRACTION_SOLVER_IRR(cmat, ddmat, ipiv, mxpot, n, 1)
         out(:, m + ih) = mxpot(:, m + 1, 1)
      end do
   end if

   ! the H-matrix route of ISOLV = 2 (RADIATION_SOLVER_MULTI, AssbMatxMulti.f90:283-291:
   ! SETUP_BLOCK_PRECONDITIONER for several bodies, HMATRIX_BUILD, ZITER_SOLVE_MULTI,
   ! HMATRIX_DESTROY) on a smooth test kernel over the panel centroids
   allocate (kmat(n, n), hb(n, 2), hz(n, 2), hy(n), bsizes(nbody))
   do j = 1, n
      do i = 1, n
         if (i == j) then
            kmat(i, j) = (4.d0, 1.d0)
         else
            r = sqrt((xyz_p(i, 1) - xyz_p(j, 1))**2 + (xyz_p(i, 2) - xyz_p(j, 2))**2 + (xyz_p(i, 3) - xyz_p(j, 3))**2)
            kmat(i, j) = exp(cmplx(0.d0, 0.3d0 * r, kind=8)) / r
         end if
      end do
   end do
   do j = 1, n
      hb(j, 1) = cmplx(1.d0, dble(j) / dble(n), kind=8)
   end do
   hb(:, 2) = conjg(hb(:, 1))
   call HM_CHECK(hm_set_hmatrix(HM_CONTEXT(), 0_c_int, 1.d-8, 1.d-8, 50_c_int), 'hm_set_hmatrix')
   call HM_CHECK(hm_set_gmres(HM_CONTEXT(), 50_c_int, 1.d-10, 500_c_int, -1_c_int), 'hm_set_gmres')
   if (nbody > 1) then
      do b = 1, nbody
         bsizes(b) = body_start(b + 1) - body_start(b)
      end do
      call SETUP_BLOCK_PRECONDITIONER(n, kmat, nbody, bsizes)
   end if
   call HMATRIX_BUILD(n, kmat, xyz_p)
   call HMATVEC(n, hb(:, 1), hy)
   call ZITER_SOLVE_MULTI(2, n, kmat, hb, hz, 2, hinfo)
   call HMATRIX_DESTROY()

   ! LAPACK names
   allocate (ip4(4))
   call ZGETRF(4, 4, lua, 4, ip4, info)
   x = [(1.d0, 0.d0), (2.d0, 0.d0), (3.d0, 0.d0), (4.d0, 0.d0)]
   if (info == 0) call ZGETRS('N', 4, 1, lua, 4, ip4, x, 4, info)

   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   write (u) grn, sij, dij, out, x, info, hy, hz, hinfo
   close (u)

contains

   integer function hm_dispersion_f(wv, w, hh, gd)
      use iso_c_binding
      real(8), intent(out) :: wv(200)
      real(8), intent(in) :: w, hh, gd
      interface
         integer(c_int) function hm_dispersion(wvn, nrt, w, h, g_disp) bind(C, name="hm_dispersion")
            import :: c_int, c_double
            real(c_double), intent(out) :: wvn(*)
            integer(c_int), value :: nrt
            real(c_double), value :: w, h, g_disp
         end function
      end interface
      hm_dispersion_f = hm_dispersion(wv, 200, w, hh, gd)
   end function

end program hm_fortran_check
