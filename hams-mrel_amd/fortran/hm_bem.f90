! hm_bem.f90 -- Fortran side of the MI355X hot path: ISO_C_BINDING interfaces to
! include/hm_bem.h and wrappers that keep the reference subroutines' argument
! lists, so a HAMS-MREL driver loop can call the GPU path in place of
!   FINGREEN3D / INFGREEN3D            (FinGreen3D.f90:81, InfGreen_Appr.f90:104)
!   SGLINTBD_QUAD / SGLINTBD_TRI       (SingularIntgr.f90:252, :44)
!   CALGREEN(_MULTI)(_IRR)             (CalGreenFunc.f90:49/:124, CalGreenFuncMulti.f90:112/:189)
!   ASSB_LEFT(_MULTI), ASSB_RBC(_MULTI), ASSB_DBC(_MULTI)   (AssbMatx.f90:53-187, AssbMatxMulti.f90:53-187)
!   RADIATION/DIFFRACTION_SOLVER(_MULTI)                     (AssbMatx.f90:257/:318, AssbMatxMulti.f90:257/:323)
!   ASSB_LEFT/RBC/DBC_IRR(_MULTI), RADIATION/DIFFRACTION_SOLVER_IRR(_MULTI)
!                                      (AssbMatx_irr.f90:53-527, AssbMatx_irrMulti.f90:53-529)
!   RFORCE_MULTI / EFORCE_MULTI        (PotentWavForceMulti.f90:138, :45)
!   ZITER_SOLVE_MULTI, SETUP_BLOCK_PRECONDITIONER (IterativeSolvers.f90:574, :190)
!   HMATRIX_BUILD / HMATVEC / HMATRIX_DESTROY      (HMatrix.f90:135, :760, :852)
! The reference passes geometry and wave state through module variables
! (WavDynMods.f90); here the driver hands them over once with HM_SETUP, per
! frequency with HM_SET_FREQUENCY, and per heading through this module's BETA
! (radians; the driver's BETA = WVHD(II)*PI/180, HAMS_Prog.f90:696, sets it).
! NSYS (geometric symmetry) must be 1.
module hm_bem
   use iso_c_binding
   implicit none
   private

   integer, parameter, public :: HM_FP64 = 0, HM_FP32_IR = 1, HM_GMRES = 2

   interface
      integer(c_int) function hm_create(ctx) bind(C, name="hm_create")
         import :: c_int, c_ptr
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function hm_destroy(ctx) bind(C, name="hm_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function hm_set_constants(ctx, rho, g, g_disp, amp, xw, yw) bind(C, name="hm_set_constants")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: rho, g, g_disp, amp, xw, yw
      end function
      integer(c_int) function hm_set_geometry(ctx, nbody, body_start, ntnd, xyz, ncn, ncon, xyz_p, ds, pnsz, &
                                              dxyz_p, ni, ixyz_p) bind(C, name="hm_set_geometry")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nbody, ntnd, ni
         integer(c_int), intent(in) :: body_start(*), ncn(*), ncon(*)
         real(c_double), intent(in) :: xyz(*), xyz_p(*), ds(*), pnsz(*), dxyz_p(*), ixyz_p(*)
      end function
      integer(c_int) function hm_set_depth(ctx, h) bind(C, name="hm_set_depth")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: h
      end function
      integer(c_int) function hm_set_headings(ctx, nbeta, beta) bind(C, name="hm_set_headings")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nbeta
         real(c_double), intent(in) :: beta(*)
      end function
      integer(c_int) function hm_set_frequency(ctx, w1) bind(C, name="hm_set_frequency")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: w1
      end function
      integer(c_int) function hm_assemble(ctx) bind(C, name="hm_assemble")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function hm_factor(ctx, precision) bind(C, name="hm_factor")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: precision
      end function
      integer(c_int) function hm_solve(ctx, iters) bind(C, name="hm_solve")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: iters
      end function
      ! CalPotentialMulti / CalPressureMulti / CalElevationMulti (PressureElevationMulti.f90:55,
      ! :305, :329) for every mode and heading at nfp points after the solve
      integer(c_int) function hm_field_points(ctx, nfp, xfp, pot, prs, elv) bind(C, name="hm_field_points")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfp
         real(c_double), intent(in) :: xfp(*)
         complex(c_double_complex), intent(out) :: pot(*), prs(*), elv(*)
      end function
      ! SolveMotionMulti (SolveMotionMulti.f90:46) for one heading; the body matrices
      ! MATX_MULTI, CRS_MULTI, KSTF_MULTI are passed instead of read from the module
      integer(c_int) function hm_solve_motion(ctx, nbody, w1, amas, bdmp, blnr, bqdr, matx, crs, kstf, exfc, &
                                              dspl, iters) bind(C, name="hm_solve_motion")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: nbody
         real(c_double), value :: w1
         real(c_double), intent(in) :: amas(*), bdmp(*), blnr(*), bqdr(*), matx(*), crs(*), kstf(*)
         complex(c_double_complex), intent(in) :: exfc(*)
         complex(c_double_complex), intent(out) :: dspl(*)
         integer(c_int), intent(out) :: iters
      end function
      integer(c_int) function hm_forces(ctx, amas, bdmp, exfc) bind(C, name="hm_forces")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: amas(*), bdmp(*)
         complex(c_double_complex), intent(out) :: exfc(*)
      end function
      integer(c_int) function hm_get_system(ctx, amat, smat, rhs) bind(C, name="hm_get_system")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, amat, smat, rhs
      end function
      integer(c_int) function hm_get_normal_equations(ctx, cmat, drhs) bind(C, name="hm_get_normal_equations")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, cmat, drhs
      end function
      integer(c_int) function hm_sglintbd_batch(ctx, n, ncn, verts, xf, sij, dij) bind(C, name="hm_sglintbd_batch")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         integer(c_int), intent(in) :: ncn(*)
         real(c_double), intent(in) :: verts(*), xf(*)
         real(c_double), intent(out) :: sij(*), dij(*)
      end function
      integer(c_int) function hm_get_potentials(ctx, mxpot) bind(C, name="hm_get_potentials")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(out) :: mxpot(*)
      end function
      integer(c_int) function hm_fingreen3d_batch(ctx, n, xf, xp, yf, yp, zf, zp, v, wvn, nk, h, flag, grn) &
                                                  bind(C, name="hm_fingreen3d_batch")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: n, nk
         real(c_double), intent(in) :: xf(*), xp(*), yf(*), yp(*), zf(*), zp(*), wvn(*)
         real(c_double), value :: v, h
         integer(c_int), intent(in) :: flag(*)
         complex(c_double_complex), intent(out) :: grn(*)
      end function
      integer(c_int) function hm_infgreen3d_batch(ctx, n, xf, xp, yf, yp, zf, zp, v, flag, grn) &
                                                  bind(C, name="hm_infgreen3d_batch")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         real(c_double), intent(in) :: xf(*), xp(*), yf(*), yp(*), zf(*), zp(*)
         real(c_double), value :: v
         integer(c_int), intent(in) :: flag(*)
         complex(c_double_complex), intent(out) :: grn(*)
      end function
      integer(c_int) function hm_zgetrs(ctx, trans, n, nrhs, a, lda, ipiv, b, ldb, info) bind(C, name="hm_zgetrs")
         import :: c_int, c_ptr, c_char, c_double_complex
         type(c_ptr), value :: ctx
         character(kind=c_char), value :: trans
         integer(c_int), value :: n, nrhs, lda, ldb
         complex(c_double_complex), intent(in) :: a(*)
         integer(c_int), intent(in) :: ipiv(*)
         complex(c_double_complex), intent(inout) :: b(*)
         integer(c_int), intent(out) :: info
      end function
      integer(c_int) function hm_zgetrf(ctx, m, n, a, lda, ipiv, info) bind(C, name="hm_zgetrf")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: m, n, lda
         complex(c_double_complex), intent(inout) :: a(*)
         integer(c_int), intent(out) :: ipiv(*), info
      end function
      integer(c_int) function hm_set_gmres(ctx, restart, tol, maxiter, precond) bind(C, name="hm_set_gmres")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: restart, maxiter, precond
         real(c_double), value :: tol
      end function
      integer(c_int) function hm_set_hmatrix(ctx, enable, eps, eps_s, kmax) bind(C, name="hm_set_hmatrix")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable, kmax
         real(c_double), value :: eps, eps_s
      end function
      integer(c_int) function hm_set_pair_symmetry(ctx, enable) bind(C, name="hm_set_pair_symmetry")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
      end function
      integer(c_int) function hm_set_frequency_limit(ctx, which) bind(C, name="hm_set_frequency_limit")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: which
      end function
      integer(c_int) function hm_hmatrix_build(ctx, n, a, lda, nblocks, block_sizes) bind(C, name="hm_hmatrix_build")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: n, lda, nblocks
         complex(c_double_complex), intent(in) :: a(lda, *)
         integer(c_int), intent(in) :: block_sizes(*)
      end function
      integer(c_int) function hm_hmatrix_build_tree(ctx, n, a, lda, xyz_p, ldxyz, nblocks, block_sizes) &
            bind(C, name="hm_hmatrix_build_tree")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: n, lda, ldxyz, nblocks
         complex(c_double_complex), intent(in) :: a(lda, *)
         real(c_double), intent(in) :: xyz_p(ldxyz, 3)
         integer(c_int), intent(in) :: block_sizes(*)
      end function
      integer(c_int) function hm_hmatrix_release(ctx) bind(C, name="hm_hmatrix_release")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function hm_hmatrix_apply(ctx, mat, ncol, x, y) bind(C, name="hm_hmatrix_apply")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: mat, ncol
         complex(c_double_complex), intent(in) :: x(*)
         complex(c_double_complex), intent(out) :: y(*)
      end function
      integer(c_int) function hm_ziter_solve(ctx, isolv, n, a, lda, b, ldb, x, ldx, nrhs, nblocks, block_sizes, info) &
            bind(C, name="hm_ziter_solve")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: isolv, n, lda, ldb, ldx, nrhs, nblocks
         complex(c_double_complex), intent(in) :: a(lda, *), b(ldb, *)
         complex(c_double_complex), intent(out) :: x(ldx, *)
         integer(c_int), intent(in) :: block_sizes(*)
         integer(c_int), intent(out) :: info
      end function
      integer(c_int) function hm_batch_create(batch) bind(C, name="hm_batch_create")
         import :: c_ptr, c_int
         type(c_ptr) :: batch
      end function
      integer(c_int) function hm_batch_destroy(batch) bind(C, name="hm_batch_destroy")
         import :: c_ptr, c_int
         type(c_ptr), value :: batch
      end function
      integer(c_int) function hm_batch_set_constants(batch, rho, g, g_disp, amp, xw, yw) &
            bind(C, name="hm_batch_set_constants")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: batch
         real(c_double), value :: rho, g, g_disp, amp, xw, yw
      end function
      integer(c_int) function hm_batch_add_problem(batch, nbody, body_start, ntnd, xyz, ncn, ncon, xyz_p, ds, pnsz, &
                                                   dxyz_p, h, nbeta, beta) bind(C, name="hm_batch_add_problem")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: batch
         integer(c_int), value :: nbody, ntnd, nbeta
         integer(c_int), intent(in) :: body_start(*), ncn(*), ncon(*)
         real(c_double), intent(in) :: xyz(*), xyz_p(*), ds(*), pnsz(*), dxyz_p(*), beta(*)
         real(c_double), value :: h
      end function
      integer(c_int) function hm_batch_prepare(batch, precision, autotune, use_graph, nrefine) &
            bind(C, name="hm_batch_prepare")
         import :: c_ptr, c_int
         type(c_ptr), value :: batch
         integer(c_int), value :: precision, autotune, use_graph, nrefine
      end function
      integer(c_int) function hm_batch_solve(batch, w1, amas, bdmp, exfc) bind(C, name="hm_batch_solve")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: batch
         real(c_double), intent(in) :: w1(*)
         real(c_double), intent(out) :: amas(*), bdmp(*)
         complex(c_double_complex), intent(out) :: exfc(*)
      end function
   end interface

   public :: hm_create, hm_destroy, hm_set_constants, hm_set_geometry, hm_set_depth, hm_set_headings, &
             hm_set_frequency, hm_assemble, hm_factor, hm_solve, hm_forces, hm_field_points, hm_solve_motion, hm_get_system, &
             hm_get_potentials, &
             hm_fingreen3d_batch, hm_infgreen3d_batch, hm_zgetrf, hm_zgetrs
   public :: HM_CONTEXT, HM_SETUP, HM_SET_FREQUENCY_F, HM_CHECK
   public :: FINGREEN3D, INFGREEN3D
   public :: ASSB_LEFT_MULTI, ASSB_RBC_MULTI, ASSB_DBC_MULTI, RADIATION_SOLVER_MULTI, DIFFRACTION_SOLVER_MULTI
   public :: ASSB_LEFT, ASSB_RBC, ASSB_DBC, RADIATION_SOLVER, DIFFRACTION_SOLVER
   public :: ASSB_LEFT_IRR, ASSB_RBC_IRR, ASSB_DBC_IRR, RADIATION_SOLVER_IRR, DIFFRACTION_SOLVER_IRR
   public :: ASSB_LEFT_IRR_MULTI, ASSB_RBC_IRR_MULTI, ASSB_DBC_IRR_MULTI, RADIATION_SOLVER_IRR_MULTI, &
             DIFFRACTION_SOLVER_IRR_MULTI
   public :: CALGREEN, CALGREEN_MULTI, CALGREEN_IRR, CALGREEN_MULTI_IRR, SGLINTBD_QUAD, SGLINTBD_TRI
   public :: BETA, hm_get_normal_equations, hm_sglintbd_batch
   public :: RFORCE_EFORCE_MULTI
   public :: hm_set_gmres, hm_set_hmatrix, hm_set_pair_symmetry, hm_set_frequency_limit, hm_hmatrix_build, &
             hm_hmatrix_build_tree, hm_hmatrix_release, &
             hm_hmatrix_apply, hm_ziter_solve, hm_batch_create, hm_batch_destroy, hm_batch_set_constants, &
             hm_batch_add_problem, hm_batch_prepare, hm_batch_solve
   public :: ZITER_SOLVE_MULTI, SETUP_BLOCK_PRECONDITIONER, HMATRIX_BUILD, HMATVEC, HMATRIX_DESTROY
   public :: HM_SET_FREQUENCY_LIMIT_F

   type(c_ptr), save :: ctx_ = c_null_ptr
   ! SETUP_BLOCK_PRECONDITIONER's blocks, used by the next ZITER_SOLVE_MULTI
   integer(c_int), allocatable, save :: pblocks_(:)
   integer, save :: n_ = 0, ni_ = 0, nbody_ = 0, nbeta_ = 0
   integer, allocatable, save :: body_start_(:)
   complex(c_double_complex), allocatable, target, save :: rhs_(:, :), drhs_(:, :)
   ! host copies of the geometry for SGLINTBD_*: nodes (3, ntnd), connectivity
   ! (4, n), panel count per panel, hull and waterplane centroids (3, n / ni)
   real(8), allocatable, save :: xyz_(:, :), xyzp_(:, :), ixyzp_(:, :), beta_(:)
   integer, allocatable, save :: ncn_(:), ncon_(:, :)
   ! the current heading (radians) of ASSB_DBC*, as the reference's module BETA
   real(8), save :: BETA = 0.d0

contains

   ! the module's context, created on first use
   function HM_CONTEXT() result(ctx)
      type(c_ptr) :: ctx
      if (.not. c_associated(ctx_)) call HM_CHECK(hm_create(ctx_), 'hm_create')
      ctx = ctx_
   end function

   subroutine HM_CHECK(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      if (rc /= 0) then
         write (*, '(A,A,A,I0)') ' hm_bem: ', what, ' failed, code ', rc
         error stop 1
      end if
   end subroutine

   ! geometry, depth, headings and constants of the run: the module state the
   ! reference keeps in WavDynMods.f90 (the *_MULTI_COMB arrays, H, WVHD, RHO, G)
   subroutine HM_SETUP(NBODY, BODY_START, NTND, XYZ, NCN, NCON, XYZ_P, DS, PNSZ, DXYZ_P, NI, IXYZ_P, H, &
                       NBETA, BETA, RHO, G, G_DISP, AMP, XW)
      integer, intent(in) :: NBODY, BODY_START(NBODY + 1), NTND, NCN(:), NCON(:, :), NI, NBETA
      real(8), intent(in) :: XYZ(:, :), XYZ_P(:, :), DS(:), PNSZ(:), DXYZ_P(:, :), IXYZ_P(:, :), H
      real(8), intent(in) :: BETA(NBETA), RHO, G, G_DISP, AMP, XW(2)
      type(c_ptr) :: c
      integer(c_int), allocatable :: ncon_t(:, :)
      c = HM_CONTEXT()
      ! reference layout is (panel, k); the C-ABI is panel-major (k fastest)
      allocate (ncon_t(4, size(NCN)))
      ncon_t = transpose(NCON(:, 1:4))
      call HM_CHECK(hm_set_constants(c, RHO, G, G_DISP, AMP, XW(1), XW(2)), 'hm_set_constants')
      call HM_CHECK(hm_set_geometry(c, NBODY, BODY_START, NTND, transpose(XYZ), NCN, ncon_t, transpose(XYZ_P), &
                                    DS, PNSZ, transpose(DXYZ_P), NI, transpose(IXYZ_P)), 'hm_set_geometry')
      call HM_CHECK(hm_set_depth(c, H), 'hm_set_depth')
      call HM_CHECK(hm_set_headings(c, NBETA, BETA), 'hm_set_headings')
      n_ = size(NCN)
      ni_ = NI
      nbody_ = NBODY
      if (allocated(body_start_)) deallocate (body_start_)
      allocate (body_start_(NBODY + 1))
      body_start_ = BODY_START
      nbeta_ = NBETA
      if (allocated(xyz_)) deallocate (xyz_, xyzp_, ixyzp_, ncn_, ncon_, beta_)
      allocate (xyz_(3, NTND), xyzp_(3, n_), ixyzp_(3, max(NI, 1)), ncn_(n_), ncon_(4, n_), beta_(max(NBETA, 1)))
      xyz_ = transpose(XYZ(1:NTND, 1:3))
      xyzp_ = transpose(XYZ_P(1:n_, 1:3))
      ixyzp_ = 0.d0
      if (NI > 0) ixyzp_ = transpose(IXYZ_P(1:NI, 1:3))
      ncn_ = NCN
      ncon_ = ncon_t
      beta_ = 0.d0
      if (NBETA > 0) beta_(1:NBETA) = BETA(1:NBETA)
   end subroutine

   ! the heading of the reference's module BETA among those of HM_SETUP (1-based)
   integer function heading_index()
      integer :: k
      do k = 1, nbeta_
         if (abs(beta_(k) - BETA) <= 1.d-12 * max(1.d0, abs(BETA))) then
            heading_index = k
            return
         end if
      end do
      heading_index = 0
      call HM_CHECK(-104_c_int, 'ASSB_DBC*: BETA is not one of the headings given to HM_SETUP')
   end function

   ! CalWaveProperts(KK) for INFT=3 (ImplementSubs.f90:96): omega -> V, WVN
   subroutine HM_SET_FREQUENCY_F(W1)
      real(8), intent(in) :: W1
      call HM_CHECK(hm_set_frequency(HM_CONTEXT(), W1), 'hm_set_frequency')
   end subroutine

   ! CalWaveProperts(KK) with SYBO = 1 (ImplementSubs.f90:45-70): KK = 1 zero
   ! frequency, KK = 2 infinite frequency
   subroutine HM_SET_FREQUENCY_LIMIT_F(KK)
      integer, intent(in) :: KK
      call HM_CHECK(hm_set_frequency_limit(HM_CONTEXT(), int(KK - 1, c_int)), 'hm_set_frequency_limit')
   end subroutine

   ! ---- iterative solver and H-matrix with the reference argument lists -------
   ! SETUP_BLOCK_PRECONDITIONER (IterativeSolvers.f90:190): the blocks are kept
   ! and factored on the GPU by the next ZITER_SOLVE_MULTI
   subroutine SETUP_BLOCK_PRECONDITIONER(N, AMAT, NBLOCKS, BLOCK_SIZES)
      integer, intent(in) :: N, NBLOCKS
      complex(8), intent(in) :: AMAT(N, N)
      integer, intent(in) :: BLOCK_SIZES(NBLOCKS)
      if (allocated(pblocks_)) deallocate (pblocks_)
      allocate (pblocks_(NBLOCKS))
      pblocks_ = BLOCK_SIZES
      if (sum(BLOCK_SIZES) /= N .or. size(AMAT, 1) /= N) call HM_CHECK(-112_c_int, 'SETUP_BLOCK_PRECONDITIONER')
   end subroutine

   ! HMATRIX_BUILD (HMatrix.f90:135): the reference's cluster tree over XYZ_P,
   ! with the blocks of the last SETUP_BLOCK_PRECONDITIONER of this order as the
   ! preconditioner (none otherwise: the reference's single body)
   subroutine HMATRIX_BUILD(N, AMAT, XYZ_P)
      integer, intent(in) :: N
      complex(8), intent(in) :: AMAT(N, N)
      real(8), intent(in) :: XYZ_P(N, 3)
      integer(c_int) :: none(1)
      none = 0
      if (size(XYZ_P, 1) /= N) call HM_CHECK(-103_c_int, 'HMATRIX_BUILD: XYZ_P')
      if (allocated(pblocks_)) then
         if (sum(pblocks_) == N) then
            call HM_CHECK(hm_hmatrix_build_tree(HM_CONTEXT(), N, AMAT, N, XYZ_P, N, size(pblocks_), pblocks_), &
                          'HMATRIX_BUILD')
            return
         end if
      end if
      call HM_CHECK(hm_hmatrix_build_tree(HM_CONTEXT(), N, AMAT, N, XYZ_P, N, 0_c_int, none), 'HMATRIX_BUILD')
   end subroutine

   subroutine HMATVEC(N, X_IN, Y_OUT)
      integer, intent(in) :: N
      complex(8), intent(in) :: X_IN(N)
      complex(8), intent(out) :: Y_OUT(N)
      call HM_CHECK(hm_hmatrix_apply(HM_CONTEXT(), 0_c_int, 1_c_int, X_IN, Y_OUT), 'HMATVEC')
   end subroutine

   subroutine HMATRIX_DESTROY()
      call HM_CHECK(hm_hmatrix_release(HM_CONTEXT()), 'HMATRIX_DESTROY')
   end subroutine

   ! ZITER_SOLVE_MULTI (IterativeSolvers.f90:574): GMRES on the GPU, on the
   ! H-matrix when HMATRIX_BUILD has run for this matrix, else on AMAT with the
   ! blocks of SETUP_BLOCK_PRECONDITIONER; INFO = 0 converged, 1 not
   subroutine ZITER_SOLVE_MULTI(ISOLV_IN, N, AMAT, B, X, NRHS, INFO)
      integer, intent(in) :: ISOLV_IN, N, NRHS
      complex(8), intent(in) :: AMAT(N, N), B(N, NRHS)
      complex(8), intent(out) :: X(N, NRHS)
      integer, intent(out) :: INFO
      integer(c_int) :: inf, isolv, nblk
      integer(c_int) :: none(1)
      isolv = 1
      if (ISOLV_IN == 2) isolv = 2
      none(1) = 0
      nblk = 0
      if (allocated(pblocks_)) nblk = size(pblocks_)
      if (nblk > 0) then
         call HM_CHECK(hm_ziter_solve(HM_CONTEXT(), isolv, N, AMAT, N, B, N, X, N, NRHS, nblk, pblocks_, inf), &
                       'ZITER_SOLVE_MULTI')
      else
         call HM_CHECK(hm_ziter_solve(HM_CONTEXT(), isolv, N, AMAT, N, B, N, X, N, NRHS, 0_c_int, none, inf), &
                       'ZITER_SOLVE_MULTI')
      end if
      INFO = inf
   end subroutine

   ! ---- element-wise kernels with the reference argument lists ----------------
   subroutine FINGREEN3D(XF, XP, YF, YP, ZF, ZP, V, WVN, NK, H, GRN, FLAG)
      integer, intent(in) :: NK, FLAG
      real(8), intent(in) :: XF, XP, YF, YP, ZF, ZP, V, WVN(1:NK), H
      complex(8), intent(out) :: GRN(4)
      integer(c_int) :: f(1)
      f(1) = FLAG
      call HM_CHECK(hm_fingreen3d_batch(HM_CONTEXT(), 1, [XF], [XP], [YF], [YP], [ZF], [ZP], V, WVN, NK, H, f, GRN), &
                    'FINGREEN3D')
   end subroutine

   subroutine INFGREEN3D(XF, XP, YF, YP, ZF, ZP, V, GRN, FLAG)
      integer, intent(in) :: FLAG
      real(8), intent(in) :: XF, XP, YF, YP, ZF, ZP, V
      complex(8), intent(out) :: GRN(4)
      integer(c_int) :: f(1)
      f(1) = FLAG
      call HM_CHECK(hm_infgreen3d_batch(HM_CONTEXT(), 1, [XF], [XP], [YF], [YP], [ZF], [ZP], V, f, GRN), 'INFGREEN3D')
   end subroutine

   ! SGLINTBD_QUAD / SGLINTBD_TRI (SingularIntgr.f90:252, :44): the Rankine
   ! integral over hull panel JEL seen from the centroid of hull panel IEL
   ! (IRR = 1) or of waterplane panel IEL (IRR = 3). The waterplane panels are
   ! field points only on this path (IRR = 2, 4 take them as sources: refused).
   subroutine sglintbd_one(IS, IEL, JEL, SIJ, DIJ, IRR, nv)
      integer, intent(in) :: IS, IEL, JEL, IRR, nv
      real(8), intent(out) :: SIJ, DIJ(3)
      real(c_double) :: verts(12), xf(3), s(1), d(3)
      integer(c_int) :: cn(1)
      integer :: k
      if (IS /= 1) call HM_CHECK(-101_c_int, 'SGLINTBD: IS must be 1 (no symmetry)')
      if (IRR /= 1 .and. IRR /= 3) call HM_CHECK(-106_c_int, 'SGLINTBD: IRR 2/4 (waterplane source panels) not provided')
      if (JEL < 1 .or. JEL > n_) call HM_CHECK(-103_c_int, 'SGLINTBD: JEL out of range')
      if (ncn_(JEL) /= nv) call HM_CHECK(-103_c_int, 'SGLINTBD: panel JEL has another vertex count')
      if (IRR == 1) then
         if (IEL < 1 .or. IEL > n_) call HM_CHECK(-102_c_int, 'SGLINTBD: IEL out of range')
         xf = xyzp_(:, IEL)
      else
         if (IEL < 1 .or. IEL > ni_) call HM_CHECK(-102_c_int, 'SGLINTBD: IEL out of range')
         xf = ixyzp_(:, IEL)
      end if
      verts = 0.d0
      do k = 1, nv
         verts(3 * k - 2:3 * k) = xyz_(:, ncon_(k, JEL))
      end do
      cn(1) = nv
      call HM_CHECK(hm_sglintbd_batch(HM_CONTEXT(), 1_c_int, cn, verts, xf, s, d), 'SGLINTBD')
      SIJ = s(1)
      DIJ = d
   end subroutine

   subroutine SGLINTBD_QUAD(IS, IEL, JEL, SIJ, DIJ, IRR)
      integer, intent(in) :: IS, IEL, JEL, IRR
      real(8), intent(out) :: SIJ, DIJ(3)
      call sglintbd_one(IS, IEL, JEL, SIJ, DIJ, IRR, 4)
   end subroutine

   subroutine SGLINTBD_TRI(IS, IEL, JEL, SIJ, DIJ, IRR)
      integer, intent(in) :: IS, IEL, JEL, IRR
      real(8), intent(out) :: SIJ, DIJ(3)
      call sglintbd_one(IS, IEL, JEL, SIJ, DIJ, IRR, 3)
   end subroutine

   ! CALGREEN (CalGreenFunc.f90:49), CALGREEN_IRR (:124), CALGREEN_MULTI
   ! (CalGreenFuncMulti.f90:112), CALGREEN_MULTI_IRR (:189): the table fill is
   ! fused into the device assembly that the ASSB_LEFT* wrappers run, so the
   ! driver's calls return at once and leave no table on the host
   subroutine CALGREEN()
   end subroutine
   subroutine CALGREEN_IRR()
   end subroutine
   subroutine CALGREEN_MULTI()
   end subroutine
   subroutine CALGREEN_MULTI_IRR()
   end subroutine

   ! ---- the per-frequency chain with the reference argument lists --------------
   ! CALGREEN_MULTI + ASSB_LEFT_MULTI: assemble on the device and return the LU
   ! factors and pivots (ZGETRF, AssbMatxMulti.f90:112) as the reference does
   subroutine ASSB_LEFT_MULTI(AMAT, IPIV, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(out) :: IPIV(NELEM, NSYS)
      complex(8), intent(out), target :: AMAT(NELEM, NELEM, NSYS)
      integer(c_int) :: info, ncol
      type(c_ptr) :: c
      if (NSYS /= 1) call HM_CHECK(-101_c_int, 'ASSB_LEFT_MULTI: NSYS must be 1')
      c = HM_CONTEXT()
      ncol = 6 * nbody_ + nbeta_
      if (allocated(rhs_)) deallocate (rhs_)
      allocate (rhs_(NELEM, ncol))
      call HM_CHECK(hm_assemble(c), 'hm_assemble')
      call HM_CHECK(hm_get_system(c, c_loc(AMAT), c_null_ptr, c_loc(rhs_)), 'hm_get_system')
      call HM_CHECK(hm_zgetrf(c, NELEM, NELEM, AMAT, NELEM, IPIV, info), 'ZGETRF')
   end subroutine

   ! ASSB_RBC_MULTI: the 6 radiation columns of the body whose panels are
   ! NELEM_START..NELEM_END (the driver calls it once per body)
   subroutine ASSB_RBC_MULTI(BRMAT, NELEM, NSYS, NELEM_START, NELEM_END)
      integer, intent(in) :: NELEM, NSYS, NELEM_START, NELEM_END
      complex(8), intent(inout) :: BRMAT(NELEM, 6, NSYS)
      integer :: b
      do b = 0, nbody_ - 1
         if (body_start_(b + 1) == NELEM_START - 1) exit
      end do
      if (b >= nbody_ .or. body_start_(b + 2) /= NELEM_END .or. NSYS /= 1) &
         call HM_CHECK(-103_c_int, 'ASSB_RBC_MULTI: panel range is not one body')
      BRMAT(:, :, 1) = BRMAT(:, :, 1) + rhs_(:, 6 * b + 1:6 * b + 6)
   end subroutine

   ! ASSB_DBC_MULTI: diffraction right-hand side of the heading in BETA
   subroutine ASSB_DBC_MULTI(BDMAT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      complex(8), intent(out) :: BDMAT(NELEM, NSYS)
      if (NSYS /= 1) call HM_CHECK(-101_c_int, 'ASSB_DBC_MULTI: NSYS must be 1')
      BDMAT(:, 1) = rhs_(1:NELEM, 6 * nbody_ + heading_index())
   end subroutine

   ! ---- one body (AssbMatx.f90): the same chain with six modes -------------------
   subroutine ASSB_LEFT(AMAT, IPIV, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(out) :: IPIV(NELEM, NSYS)
      complex(8), intent(out), target :: AMAT(NELEM, NELEM, NSYS)
      call ASSB_LEFT_MULTI(AMAT, IPIV, NELEM, NSYS)
   end subroutine

   subroutine ASSB_RBC(BRMAT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      complex(8), intent(out) :: BRMAT(NELEM, 6, NSYS)
      if (NSYS /= 1 .or. nbody_ /= 1) call HM_CHECK(-101_c_int, 'ASSB_RBC: one body, NSYS = 1')
      BRMAT(:, :, 1) = rhs_(1:NELEM, 1:6)
   end subroutine

   subroutine ASSB_DBC(BDMAT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      complex(8), intent(out) :: BDMAT(NELEM, NSYS)
      call ASSB_DBC_MULTI(BDMAT, NELEM, NSYS)
   end subroutine

   subroutine RADIATION_SOLVER(AMAT, BRMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: AMAT(NELEM, NELEM, NSYS), BRMAT(NELEM, 6, NSYS)
      complex(8), intent(out) :: MXPOT(NELEM, 7, NSYS)
      call RADIATION_SOLVER_MULTI(AMAT, BRMAT, IPIV, MXPOT, NELEM, NSYS)
   end subroutine

   subroutine DIFFRACTION_SOLVER(AMAT, BDMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: AMAT(NELEM, NELEM, NSYS), BDMAT(NELEM, NSYS)
      complex(8), intent(inout) :: MXPOT(NELEM, 7, NSYS)
      call DIFFRACTION_SOLVER_MULTI(AMAT, BDMAT, IPIV, MXPOT, NELEM, NSYS)
   end subroutine

   ! ---- irregular-frequency removal (AssbMatx_irr.f90, AssbMatx_irrMulti.f90) ---
   ! ASSB_LEFT_IRR: AMAT (TNELEM x TNELEM, the first NELEM columns used: the hull
   ! rows and the waterplane rows), CMAT = A^T A factored (ZGETRF) with IPIV
   subroutine ASSB_LEFT_IRR(AMAT, CMAT, IPIV, NELEM, TNELEM, NSYS)
      integer, intent(in) :: NELEM, TNELEM, NSYS
      complex(8), intent(out) :: AMAT(TNELEM, TNELEM, NSYS)
      complex(8), intent(out), target :: CMAT(NELEM, NELEM, NSYS)
      integer, intent(out) :: IPIV(NELEM, NSYS)
      complex(c_double_complex), allocatable, target :: a(:, :)
      integer(c_int) :: info, ncol
      type(c_ptr) :: c
      if (NSYS /= 1) call HM_CHECK(-101_c_int, 'ASSB_LEFT_IRR: NSYS must be 1')
      if (NELEM /= n_ .or. TNELEM /= n_ + ni_ .or. ni_ == 0) call HM_CHECK(-102_c_int, 'ASSB_LEFT_IRR: NELEM, TNELEM')
      c = HM_CONTEXT()
      ncol = 6 * nbody_ + nbeta_
      if (allocated(rhs_)) deallocate (rhs_)
      if (allocated(drhs_)) deallocate (drhs_)
      allocate (rhs_(TNELEM, ncol), drhs_(NELEM, ncol), a(TNELEM, NELEM))
      call HM_CHECK(hm_assemble(c), 'hm_assemble')
      call HM_CHECK(hm_get_system(c, c_loc(a), c_null_ptr, c_loc(rhs_)), 'hm_get_system')
      call HM_CHECK(hm_get_normal_equations(c, c_loc(CMAT), c_loc(drhs_)), 'hm_get_normal_equations')
      AMAT = (0.d0, 0.d0)
      AMAT(:, 1:NELEM, 1) = a
      call HM_CHECK(hm_zgetrf(c, NELEM, NELEM, CMAT, NELEM, IPIV, info), 'ZGETRF')
   end subroutine

   ! ASSB_RBC_IRR: BRMAT (TNELEM x 6) and DRMAT = A^T BRMAT (NELEM x 6)
   subroutine ASSB_RBC_IRR(BRMAT, DRMAT, AMAT, NELEM, TNELEM, NSYS)
      integer, intent(in) :: NELEM, TNELEM, NSYS
      complex(8), intent(out) :: BRMAT(TNELEM, 6, NSYS), DRMAT(NELEM, 6, NSYS)
      complex(8), intent(in) :: AMAT(TNELEM, TNELEM, NSYS)
      if (NSYS /= 1 .or. nbody_ /= 1) call HM_CHECK(-101_c_int, 'ASSB_RBC_IRR: one body, NSYS = 1')
      BRMAT(:, :, 1) = rhs_(:, 1:6)
      DRMAT(:, :, 1) = drhs_(:, 1:6)
   end subroutine

   ! ASSB_DBC_IRR: BDMAT (TNELEM) and DDMAT = A^T BDMAT (NELEM) of the heading in BETA
   subroutine ASSB_DBC_IRR(BDMAT, DDMAT, AMAT, NELEM, TNELEM, NSYS)
      integer, intent(in) :: NELEM, TNELEM, NSYS
      complex(8), intent(out) :: BDMAT(TNELEM, NSYS), DDMAT(NELEM, NSYS)
      complex(8), intent(in) :: AMAT(TNELEM, TNELEM, NSYS)
      integer :: k
      if (NSYS /= 1) call HM_CHECK(-101_c_int, 'ASSB_DBC_IRR: NSYS must be 1')
      k = 6 * nbody_ + heading_index()
      BDMAT(:, 1) = rhs_(:, k)
      DDMAT(:, 1) = drhs_(:, k)
   end subroutine

   ! RADIATION_SOLVER_IRR / DIFFRACTION_SOLVER_IRR: ZGETRS with the factored CMAT
   subroutine RADIATION_SOLVER_IRR(CMAT, DRMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: CMAT(NELEM, NELEM, NSYS), DRMAT(NELEM, 6 * nbody_, NSYS)
      complex(8), intent(out) :: MXPOT(NELEM, 6 * nbody_ + 1, NSYS)
      call RADIATION_SOLVER_MULTI(CMAT, DRMAT, IPIV, MXPOT, NELEM, NSYS)
   end subroutine

   subroutine DIFFRACTION_SOLVER_IRR(CMAT, DDMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: CMAT(NELEM, NELEM, NSYS), DDMAT(NELEM, NSYS)
      complex(8), intent(inout) :: MXPOT(NELEM, 6 * nbody_ + 1, NSYS)
      call DIFFRACTION_SOLVER_MULTI(CMAT, DDMAT, IPIV, MXPOT, NELEM, NSYS)
   end subroutine

   ! the multi-body forms (NELEM_TOTAL hull panels, iNELEM_TOTAL waterplane panels)
   subroutine ASSB_LEFT_IRR_MULTI(AMAT, CMAT, IPIV, NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS)
      integer, intent(in) :: NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS
      complex(8), intent(out) :: AMAT(TNELEM, TNELEM, NSYS)
      complex(8), intent(out), target :: CMAT(NELEM_TOTAL, NELEM_TOTAL, NSYS)
      integer, intent(out) :: IPIV(NELEM_TOTAL, NSYS)
      if (INELEM_TOTAL /= ni_) call HM_CHECK(-102_c_int, 'ASSB_LEFT_IRR_MULTI: iNELEM_TOTAL')
      call ASSB_LEFT_IRR(AMAT, CMAT, IPIV, NELEM_TOTAL, TNELEM, NSYS)
   end subroutine

   subroutine ASSB_RBC_IRR_MULTI(BRMAT, DRMAT, AMAT, NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS, NELEM_START, NELEM_END)
      integer, intent(in) :: NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS, NELEM_START, NELEM_END
      complex(8), intent(inout) :: BRMAT(TNELEM, 6, NSYS), DRMAT(NELEM_TOTAL, 6, NSYS)
      complex(8), intent(in) :: AMAT(TNELEM, TNELEM, NSYS)
      integer :: b
      do b = 0, nbody_ - 1
         if (body_start_(b + 1) == NELEM_START - 1) exit
      end do
      if (b >= nbody_ .or. body_start_(b + 2) /= NELEM_END .or. NSYS /= 1) &
         call HM_CHECK(-103_c_int, 'ASSB_RBC_IRR_MULTI: panel range is not one body')
      BRMAT(:, :, 1) = BRMAT(:, :, 1) + rhs_(:, 6 * b + 1:6 * b + 6)
      DRMAT(:, :, 1) = DRMAT(:, :, 1) + drhs_(:, 6 * b + 1:6 * b + 6)
   end subroutine

   subroutine ASSB_DBC_IRR_MULTI(BDMAT, DDMAT, AMAT, NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS)
      integer, intent(in) :: NELEM_TOTAL, INELEM_TOTAL, TNELEM, NSYS
      complex(8), intent(out) :: BDMAT(TNELEM, NSYS), DDMAT(NELEM_TOTAL, NSYS)
      complex(8), intent(in) :: AMAT(TNELEM, TNELEM, NSYS)
      call ASSB_DBC_IRR(BDMAT, DDMAT, AMAT, NELEM_TOTAL, TNELEM, NSYS)
   end subroutine

   subroutine RADIATION_SOLVER_IRR_MULTI(CMAT, DRMAT, IPIV, MXPOT, NELEM_TOTAL, NSYS)
      integer, intent(in) :: NELEM_TOTAL, NSYS
      integer, intent(in) :: IPIV(NELEM_TOTAL, NSYS)
      complex(8), intent(in) :: CMAT(NELEM_TOTAL, NELEM_TOTAL, NSYS), DRMAT(NELEM_TOTAL, 6 * nbody_, NSYS)
      complex(8), intent(out) :: MXPOT(NELEM_TOTAL, 6 * nbody_ + 1, NSYS)
      call RADIATION_SOLVER_MULTI(CMAT, DRMAT, IPIV, MXPOT, NELEM_TOTAL, NSYS)
   end subroutine

   subroutine DIFFRACTION_SOLVER_IRR_MULTI(CMAT, DDMAT, IPIV, MXPOT, NELEM_TOTAL, NSYS)
      integer, intent(in) :: NELEM_TOTAL, NSYS
      integer, intent(in) :: IPIV(NELEM_TOTAL, NSYS)
      complex(8), intent(in) :: CMAT(NELEM_TOTAL, NELEM_TOTAL, NSYS), DDMAT(NELEM_TOTAL, NSYS)
      complex(8), intent(inout) :: MXPOT(NELEM_TOTAL, 6 * nbody_ + 1, NSYS)
      call DIFFRACTION_SOLVER_MULTI(CMAT, DDMAT, IPIV, MXPOT, NELEM_TOTAL, NSYS)
   end subroutine

   subroutine RADIATION_SOLVER_MULTI(AMAT, BRMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: AMAT(NELEM, NELEM, NSYS), BRMAT(NELEM, 6 * nbody_, NSYS)
      complex(8), intent(out) :: MXPOT(NELEM, 6 * nbody_ + 1, NSYS)
      integer(c_int) :: info
      MXPOT(:, 1:6 * nbody_, 1) = BRMAT(:, :, 1)
      call HM_CHECK(hm_zgetrs(HM_CONTEXT(), 'N', NELEM, 6 * nbody_, AMAT, NELEM, IPIV, MXPOT, NELEM, info), 'ZGETRS')
   end subroutine

   subroutine DIFFRACTION_SOLVER_MULTI(AMAT, BDMAT, IPIV, MXPOT, NELEM, NSYS)
      integer, intent(in) :: NELEM, NSYS
      integer, intent(in) :: IPIV(NELEM, NSYS)
      complex(8), intent(in) :: AMAT(NELEM, NELEM, NSYS), BDMAT(NELEM, NSYS)
      complex(8), intent(inout) :: MXPOT(NELEM, 6 * nbody_ + 1, NSYS)
      integer(c_int) :: info
      MXPOT(:, 6 * nbody_ + 1, 1) = BDMAT(:, 1)
      call HM_CHECK(hm_zgetrs(HM_CONTEXT(), 'N', NELEM, 1, AMAT, NELEM, IPIV, MXPOT(:, 6 * nbody_ + 1, 1), NELEM, &
                              info), 'ZGETRS')
   end subroutine

   ! RFORCE_MULTI + EFORCE_MULTI for all bodies and headings at once, from the
   ! device-resident solve (HM_FP64 or HM_FP32_IR): AMAS/BDMP (6nb x 6nb),
   ! EXFC (6nb x nbeta)
   subroutine RFORCE_EFORCE_MULTI(PRECISION, AMAS, BDMP, EXFC, ITERS)
      integer, intent(in) :: PRECISION
      real(8), intent(out) :: AMAS(6 * nbody_, 6 * nbody_), BDMP(6 * nbody_, 6 * nbody_)
      complex(8), intent(out) :: EXFC(6 * nbody_, nbeta_)
      integer, intent(out) :: ITERS
      type(c_ptr) :: c
      integer(c_int) :: it
      c = HM_CONTEXT()
      call HM_CHECK(hm_assemble(c), 'hm_assemble')
      call HM_CHECK(hm_factor(c, PRECISION), 'hm_factor')
      call HM_CHECK(hm_solve(c, it), 'hm_solve')
      call HM_CHECK(hm_forces(c, AMAS, BDMP, EXFC), 'hm_forces')
      ITERS = it
   end subroutine

end module hm_bem

! LAPACK-named external drop-ins: linking this object in place of MKL routes the
! reference's own ZGETRF/ZGETRS calls (AssbMatxMulti.f90:112,:279,:346) to the GPU.
subroutine ZGETRF(M, N, A, LDA, IPIV, INFO)
   use hm_bem, only: HM_CONTEXT, hm_zgetrf
   use iso_c_binding
   implicit none
   integer, intent(in) :: M, N, LDA
   complex(8), intent(inout) :: A(LDA, *)
   integer, intent(out) :: IPIV(*), INFO
   integer(c_int) :: rc, inf
   rc = hm_zgetrf(HM_CONTEXT(), M, N, A, LDA, IPIV, inf)
   INFO = inf
   if (rc < 0) INFO = rc
end subroutine

subroutine ZGETRS(TRANS, N, NRHS, A, LDA, IPIV, B, LDB, INFO)
   use hm_bem, only: HM_CONTEXT, hm_zgetrs
   use iso_c_binding
   implicit none
   character, intent(in) :: TRANS
   integer, intent(in) :: N, NRHS, LDA, LDB
   complex(8), intent(in) :: A(LDA, *)
   integer, intent(in) :: IPIV(*)
   complex(8), intent(inout) :: B(LDB, *)
   integer, intent(out) :: INFO
   integer(c_int) :: rc, inf
   rc = hm_zgetrs(HM_CONTEXT(), TRANS, N, NRHS, A, LDA, IPIV, B, LDB, inf)
   INFO = inf
   if (rc < 0) INFO = rc
end subroutine
