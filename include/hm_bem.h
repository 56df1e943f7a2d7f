/*
 * hm_bem.h -- C-ABI of the MI355X boundary-integral hot path (HAMS-MREL drop-in).
 *
 * Plain C types only: int, double, char, and opaque context pointers. Complex
 * arrays are interleaved (re, im) doubles in Fortran column-major order, i.e.
 * exactly the memory of a Fortran COMPLEX*16 / complex(c_double_complex) array.
 * Panel arrays are panel-major (xyz_p[3*i+k]); connectivity is 1-based as in the
 * reference's .pnl files. One context owns one HIP stream on the current device
 * and the device buffers of one problem; a context is single-threaded (the
 * reference's solver state is module-global too, IterativeSolvers.f90:61-74).
 *
 * Return codes: HM_OK (0); HM_ERR_ARG - k for a bad k-th argument (1-based);
 * HM_ERR_STATE when a call comes before the state it needs; HM_ERR_HIP - e for
 * HIP runtime error e; a positive value is a LAPACK-style INFO (zero pivot at
 * that column). Nothing in the library calls exit/abort; hm_error_string()
 * describes the last failure of a context.
 */
#ifndef HM_BEM_H
#define HM_BEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define HM_OK 0
#define HM_ERR_ARG (-100)
#define HM_ERR_STATE (-200)
#define HM_ERR_HIP (-1000)

/* precision of the dense solve */
#define HM_FP64 0      /* FP64 LU + FP64 triangular solves (plain path)          */
#define HM_FP32_IR 1   /* FP32 LU on MFMA, FP64 residuals, iterative refinement  */
#define HM_GMRES 2     /* restarted GMRES: FP32 MFMA products, FP64 residuals     */

typedef struct hm_ctx hm_ctx;

/* ---- context ---------------------------------------------------------------- */
int hm_create(hm_ctx **ctx);
int hm_destroy(hm_ctx *ctx);
const char *hm_error_string(const hm_ctx *ctx);

/* Physical constants. The reference hard-wires rho=1000, g=9.80665 and uses
 * g=9.807 inside DISPERSION (WavDynMods.f90:146, WavDynSubs.f90:58). */
int hm_set_constants(hm_ctx *ctx, double rho, double g, double g_disp, double amp, double xw, double yw);

/* Geometry of all bodies combined (the *_MULTI_COMB arrays that
 * BuildCombinedArrays packs, CalGreenFuncMulti.f90:53-106): n hull panels with
 * nodes xyz[3*ntnd], ncn[n], ncon[4*n] (1-based), centroids xyz_p[3n], areas
 * ds[n], longest edges pnsz[n], normals dxyz_p[6n]; body b owns panels
 * [body_start[b], body_start[b+1]). ni > 0 adds interior waterplane field points
 * ixyz_p[3*ni] for irregular-frequency removal (IRSP=1). */
int hm_set_geometry(hm_ctx *ctx, int nbody, const int *body_start, int ntnd, const double *xyz, const int *ncn,
                    const int *ncon, const double *xyz_p, const double *ds, const double *pnsz, const double *dxyz_p,
                    int ni, const double *ixyz_p);
/* Body frames: declares that the coordinates given to hm_set_geometry (nodes,
 * centroids, waterplane points) are each body's local coordinates, and places
 * body b at the horizontal offset (origin[2b], origin[2b+1]). The reference
 * reads body meshes the same way, local nodes placed by LCS_MULTI
 * (ReadPanelMeshMulti.f90:52-60; pass the rotation already applied to the
 * local coordinates). Pair geometry is then formed as the local difference
 * plus the difference of the placements, exactly the local difference within
 * one body: bodies with bitwise equal local hull panels have bitwise equal
 * diagonal blocks, on every path. wp_start[nbody+1] assigns the ni waterplane
 * points to bodies (NULL when ni = 0). Call after hm_set_geometry (which
 * clears the frames). No counterpart in the reference. */
int hm_set_body_origins(hm_ctx *ctx, int nbody, const double *origin, const int *wp_start);
/* rep[b] = the first body whose local hull panels equal body b's (b itself
 * when none does; every b without body frames). */
int hm_get_body_classes(hm_ctx *ctx, int nbody, int *rep);
/* H-matrix form (hm_set_hmatrix): bodies that share a representative (see
 * hm_get_body_classes) share its diagonal block: it is assembled, factored
 * and inverted once per frequency and copied; off-diagonal blocks between
 * such bodies at an equal offset share one cross approximation (enable = 1:
 * the blocks are the same bit for bit as with enable = 0, where every block is
 * computed). enable = 2 (the default): when every body is equal, also blocks at
 * offsets that a symmetry of the body maps onto each other (a rotation by a
 * multiple of 90 degrees or a reflection in an axis or a diagonal of its
 * frame that takes its hull panels onto themselves, to 1e-9 of its size):
 * such a block is the other's with rows and columns permuted by the
 * symmetry's panel map, and takes its cross approximation so permuted (the
 * same approximation error, other rounding and pivots than its own).
 * No counterpart in the reference. */
int hm_set_block_sharing(hm_ctx *ctx, int enable);
/* ok[8]: which horizontal maps take the hull panels of every body (all equal,
 * body frames on) onto themselves, in the order identity, rotations by 90,
 * 180 and 270 degrees, reflections in the x axis, the y axis, the diagonal
 * y = x and the diagonal y = -x (about the vertical axis through the frame's
 * origin); all 0 without body frames or with unequal bodies. No counterpart in
 * the reference. */
int hm_get_body_symmetries(hm_ctx *ctx, int *ok);
/* With the H-matrix form on: the diagonal blocks assembled per frequency, the
 * cross-approximation processes run (one per off-diagonal block and matrix;
 * blocks of equal bodies at an equal offset, o_b1 - o_b2 bit for bit, take one
 * process's factors) and their mean rank at the last assembly. No
 * counterpart in the reference. */
int hm_get_block_sharing(hm_ctx *ctx, int *diag_blocks, int *aca_processes, double *mean_rank_run);
/* H-matrix form with explicit preconditioner inverses: the Arnoldi steps apply
 * them by FP16 MFMA GEMMs with FP32 accumulation (enable = 1, the default) or
 * by FP32 ones (0); the update of the solution uses the FP32 inverses either
 * way. GMRES still stops on the FP64 residual; the FP16 rounding may cost one
 * more restart cycle. No counterpart in the reference. */
int hm_set_precond_fp16(hm_ctx *ctx, int enable);
/* Parity entry point of the FP16 complex GEMM that applies those inverses:
 * out (m x n) = a (m x k) * b (k x n), column-major, interleaved complex FP32
 * host arrays, rounded to FP16 and packed on the device, FP32 accumulation.
 * No counterpart in the reference. */
int hm_cgemm_f16_check(hm_ctx *ctx, int m, int n, int k, const float *a, const float *b, float *out);
/* Parity entry point of the FP64 complex GEMM of the H-matrix form's
 * residuals: out (m x n) = op(a) * b (k x n), op(a) = a (m x k) or, with
 * trans_a = 1, a^T (a k x m, not conjugated); column-major, interleaved
 * complex FP64 host arrays, each complex product formed from three real ones
 * (Re = ArBr - AiBi, Im = (Ar + Ai)(Br + Bi) - ArBr - AiBi). No counterpart in
 * the reference. */
int hm_cgemm_3m_check(hm_ctx *ctx, int trans_a, int m, int n, int k, const double *a, const double *b, double *out);
/* Water depth h (h < 0: infinite). Classifies the panel pairs by R/h regime. */
int hm_set_depth(hm_ctx *ctx, double h);
/* Wave headings [rad] of the diffraction problems (WVHD*PI/180, HAMS_Prog.f90:696). */
int hm_set_headings(hm_ctx *ctx, int nbeta, const double *beta);
/* Wave frequency omega [rad/s]: V = omega^2/g and the 200 dispersion roots
 * (CalWaveProperts with INFT=3, ImplementSubs.f90:96-105). */
int hm_set_frequency(hm_ctx *ctx, double w1);
/* Zero (which = 0) or infinite (which = 1) wave frequency: V = WK = W1 = 0 or -1
 * as CalWaveProperts sets them for SYBO = 1 (ImplementSubs.f90:45-70); the Green
 * functions take their V = 0 / V = -1 branches (FinGreen3D.f90:100-106,
 * InfGreen_Appr.f90:130-160) and the incident wave its 1e-20 / 1e+20
 * substitutes (PatclVelct.f90:91-99). The forces follow the reference's
 * formulas with W1 = 0 or -1 (zero damping and excitation at zero frequency). */
int hm_set_frequency_limit(hm_ctx *ctx, int which);

/* ---- the per-frequency chain (device resident) --------------------------------
 * hm_assemble   replaces CALGREEN_MULTI(_IRR) + ASSB_LEFT_MULTI's fill +
 *               ASSB_RBC_MULTI + ASSB_DBC_MULTI (+ the IRR normal equations
 *               C = A^T A, D = A^T B of AssbMatx_irr.f90:151-169,:300-318,:449-463)
 *               (CalGreenFuncMulti.f90:112, AssbMatxMulti.f90:53,:124,:187)
 * hm_factor     replaces ZGETRF on the assembled matrix (AssbMatxMulti.f90:112)
 * hm_solve      replaces ZGETRS with 6*NBODY + NBETA right-hand sides
 *               (RADIATION/DIFFRACTION_SOLVER_MULTI, AssbMatxMulti.f90:257,:323);
 *               *iters receives the refinement steps taken (0 for HM_FP64)
 * hm_forces     replaces RFORCE_MULTI / EFORCE_MULTI (PotentWavForceMulti.f90:138,:45):
 *               amas, bdmp: (6nb x 6nb) column-major [force dof, mode];
 *               exfc: complex (6nb x nbeta) column-major. */
int hm_assemble(hm_ctx *ctx);
int hm_factor(hm_ctx *ctx, int precision);
int hm_solve(hm_ctx *ctx, int *iters);
int hm_forces(hm_ctx *ctx, double *amas, double *bdmp, double *exfc);
/* Restarted GMRES (precision HM_GMRES), replacing ZITER_SOLVE_MULTI
 * (IterativeSolvers.f90:574) with its module parameters GMRES_RESTART, ITER_TOL,
 * ITER_MAXITER (:56-58). All right-hand sides advance together; each Krylov
 * step is one FP32 complex GEMM on MFMA and each restart recomputes the FP64
 * residual; converged when ||B - A X|| <= tol ||B|| in every column (:388).
 * precond: 1 block-diagonal LU with one block per body (SETUP_BLOCK_PRECONDITIONER
 * :190), applied by the blocks' triangular solves (APPLY_PRECONDITIONER :266);
 * 2 the same blocks applied as their explicit FP32 inverses (one batched GEMM
 * per Krylov step; the inverses, U^-1 L^-1 with the pivots applied to the
 * vector, are built after each factorisation); 0 none; -1 (default) as 2 for
 * several bodies, none for one.
 * Defaults: restart 50, tol 1e-6, maxiter 500 (total Krylov steps), precond -1.
 * hm_get_gmres_stats reports the last solve (converged = 0 is the reference's
 * INFO = 1). */
int hm_set_gmres(hm_ctx *ctx, int restart, double tol, int maxiter, int precond);
int hm_get_gmres_stats(hm_ctx *ctx, int *iters, int *cycles, int *converged, double *rel_residual);
/* H-matrix form of the system, replacing HMATRIX_BUILD / HMATVEC
 * (HMatrix.f90:135,:760) as ZITER_SOLVE_MULTI uses them. For two or more bodies
 * of equal panel counts without interior waterplane points (and kmax no larger
 * than a body), A and S are compressed as they are assembled; otherwise (one
 * body, bodies of unequal panel counts, waterplane points, or
 * hm_set_cluster_tree mode 1) the system matrix (A, or C = A^T A with
 * waterplane points) is assembled densely and compressed at hm_factor over the
 * reference's geometric cluster tree (see hm_hmatrix_build_tree), with the
 * dense right-hand sides, the block-diagonal preconditioner of the dense GMRES
 * and the FP32 tree products in its Krylov steps; hm_hmatrix_release then
 * frees it (factor again to solve). The equal-body form: one cluster per body: the diagonal blocks are
 * assembled densely, every off-diagonal block by adaptive cross approximation
 * with partial pivoting (ADD_LOWRANK_BLOCK :560) that evaluates only the rows
 * and columns it samples, stopping at |u_k||v_k| < eps |U V^T|_F or rank kmax
 * (ACA_TOL, RANK_MAX :59-60; defaults eps 1e-4, kmax 50). The right-hand sides
 * come from the H-matrix of S in FP64, compressed to its own tolerance eps_s
 * (the reference forms S W densely). Solves with precision HM_GMRES only;
 * hm_get_system is unavailable. Toggling the form reclassifies the pairs;
 * hm_set_geometry turns it off. hm_get_hmatrix_stats: blocks that stopped at
 * kmax before eps, ACA steps, mean and largest rank, and stored entries per
 * matrix over n^2, for the last assembly. */
int hm_set_hmatrix(hm_ctx *ctx, int enable, double eps, double eps_s, int kmax);
/* The cluster tree of the reference's HMATRIX_BUILD (BUILD_CLUSTER_TREE,
 * SPLIT_CLUSTER_ITERATIVE, BUILD_BLOCK_TREE_ITERATIVE, HMatrix.f90:203-497):
 * leaf = NLEAF_DEFAULT (32, at most 64 here), eta = ETA_DEFAULT (2.0, :57-58).
 * mode 0 (default): hm_set_hmatrix takes the body clusters where they apply;
 * mode 1: always the tree. hm_get_cluster_tree_stats reports the last tree
 * H-matrix: tree nodes, leaves, dense and low-rank blocks, and the admissible
 * blocks stored dense after their cross approximation (rank cap reached, or
 * rank (nr + nc) >= nr nc as in ADD_LOWRANK_BLOCK :731). */
int hm_set_cluster_tree(hm_ctx *ctx, int mode, int leaf, double eta);
int hm_get_cluster_tree_stats(hm_ctx *ctx, int *clusters, int *leaves, int *dense_blocks, int *lowrank_blocks,
                              int *dense_fallbacks);
/* No reference counterpart (a scheduling choice of this library): enable = 1
 * (the default) runs the cross approximation on a second stream of the context
 * beside the diagonal blocks and the block preconditioner; 0 runs it after them
 * on the context's stream, so that its kernels' timings are their own. */
int hm_set_hmatrix_overlap(hm_ctx *ctx, int enable);
int hm_get_hmatrix_stats(hm_ctx *ctx, int *unconverged, int *steps, double *mean_rank, int *max_rank,
                         double *stored_fraction);
/* The diagonal blocks of the H-matrix form assemble each PIDCOCK pair (i, j),
 * i < j, together with its mirror (j, i): the reference's series of the pair
 * is symmetric in the two depths, so one evaluation of four series (value, R-
 * and both depth derivatives) and four epsilon limits serves both (instead of
 * six and six). The mirrored entries equal a direct evaluation to rounding,
 * within the one-ulp sensitivity of the reference's epsilon algorithm (its
 * exact-repeat exit). enable = 0 assembles every pair directly (bit for bit
 * the dense path's blocks). Default on; toggling reclassifies the pairs. No
 * counterpart in the reference (it assembles every pair). */
int hm_set_pair_symmetry(hm_ctx *ctx, int enable);
/* y = H x in FP64 for the assembled H-matrix form of A (mat 0) or S (mat 1),
 * or for the H-matrix of hm_hmatrix_build / hm_hmatrix_build_tree or of the
 * tree form's last hm_factor (mat 0); x, y: n x ncol column-major
 * complex (HMATVEC, HMatrix.f90:760, for several columns). */
int hm_hmatrix_apply(hm_ctx *ctx, int mat, int ncol, const double *x, double *y);

/* Iterative solves on an assembled host matrix, with the reference's argument
 * lists around them (the Fortran wrappers ZITER_SOLVE_MULTI,
 * SETUP_BLOCK_PRECONDITIONER, HMATRIX_BUILD, HMATVEC, HMATRIX_DESTROY).
 * hm_hmatrix_build replaces HMATRIX_BUILD(N, AMAT, XYZ_P) (HMatrix.f90:135):
 * clusters are the nblocks equal blocks given (the bodies), the diagonal blocks
 * stay dense and every off-diagonal block is compressed by the cross
 * approximation of hm_set_hmatrix (eps, kmax) on the matrix's entries; the
 * block-diagonal FP32 preconditioner is factored with it. hm_hmatrix_release
 * replaces HMATRIX_DESTROY (:852). hm_ziter_solve replaces
 * ZITER_SOLVE_MULTI(ISOLV, N, AMAT, B, X, NRHS, INFO) (IterativeSolvers.f90:574):
 * isolv = 1 restarted GMRES (hm_set_gmres parameters) on the dense matrix a with
 * the block-diagonal preconditioner of SETUP_BLOCK_PRECONDITIONER(N, AMAT,
 * NBLOCKS, BLOCK_SIZES) (:190; nblocks = 0: none); isolv = 2 GMRES on the
 * H-matrix of the last hm_hmatrix_build (a and the blocks are then not read),
 * as ZGMRES_SOLVE uses HMATVEC once HMATRIX_BUILD has run (:325-343). info = 0
 * converged, 1 not converged (the reference's INFO). Both reuse the context's
 * matrix buffers: an assembled system is discarded. */
int hm_hmatrix_build(hm_ctx *ctx, int n, const double *a, int lda, int nblocks, const int *block_sizes);
/* HMATRIX_BUILD(N, AMAT, XYZ_P) (HMatrix.f90:135) as the reference builds it,
 * for any matrix: the cluster tree and block partition over the points XYZ_P
 * (column-major n x 3, leading dimension ldxyz), each admissible block
 * compressed by cross approximation of a's entries (eps and kmax of
 * hm_set_hmatrix), each inadmissible pair of leaves stored dense; with
 * nblocks > 0 the block-diagonal FP32 preconditioner over those blocks
 * (SETUP_BLOCK_PRECONDITIONER, IterativeSolvers.f90:190), with nblocks = 0 none
 * (the reference's single body, AssbMatx.f90:282). hm_ziter_solve(isolv = 2)
 * and hm_hmatrix_apply (mat 0) then use it. */
int hm_hmatrix_build_tree(hm_ctx *ctx, int n, const double *a, int lda, const double *xyz_p, int ldxyz, int nblocks,
                          const int *block_sizes);
/* The cluster tree and block partition hm_hmatrix_build_tree uses, on the host
 * (no device work): perm[k] = the point (0-based) at cluster position k; per
 * block 5 ints (first row and rows, first column and columns, in cluster
 * positions; 1 if admissible, 0 if stored dense). blocks = NULL returns the
 * block count in *nblocks only. */
int hm_cluster_tree_partition(int n, const double *xyz_p, int ldxyz, int leaf, double eta, int *perm, int max_blocks,
                              int *nblocks, int *blocks);
int hm_hmatrix_release(hm_ctx *ctx);
int hm_ziter_solve(hm_ctx *ctx, int isolv, int n, const double *a, int lda, const double *b, int ldb, double *x,
                   int ldx, int nrhs, int nblocks, const int *block_sizes, int *info);
/* set_frequency + assemble + factor + solve + forces */
int hm_solve_frequency(hm_ctx *ctx, double w1, int precision, double *amas, double *bdmp, double *exfc);
/* nfreq frequencies (the driver's frequency loop, HAMS_Prog.f90:613-724);
 * outputs stacked per frequency. With hm_set_workers(n > 1), n frequencies are
 * in flight at once: this context and n - 1 clones of its configuration (own
 * streams and device buffers, made at the first call and kept while the
 * configuration holds) each take the next frequency in a host thread of its
 * own; every frequency's outputs are those of hm_solve_frequency on it. */
int hm_run_frequencies(hm_ctx *ctx, int nfreq, const double *w1, int precision, double *amas, double *bdmp,
                       double *exfc);
/* frequencies in flight in hm_run_frequencies (1..8, default 1); clones
 * beyond n - 1 are released at once. No counterpart in the reference. */
int hm_set_workers(hm_ctx *ctx, int nworkers);

/* Kernel variant of the two epsilon-accelerated finite-depth regimes (EIGENE,
 * PIDCOCK; FinGreen3D.f90:413,:558): 0 one pair per lane (series arrays in
 * scratch memory); 1 (default) a group of 8 (EIGENE) / 16 (PIDCOCK) lanes per
 * pair with the series and the epsilon table in registers; 2 the same with 4 / 8
 * lanes. All compute the same algorithm; the lane groups form the partial sums
 * in a different order and take reciprocals to within two ulps, so they agree
 * with variant 0 to the conditioning of the epsilon algorithm. */
int hm_set_series_variant(hm_ctx *ctx, int variant);
/* Panels of the dense LU (ZGETRF's panel factorisation, lu_factor): 1 (the
 * default) one cooperative launch per 64-column panel where its grid fits the
 * device (the panel's rows in LDS across workgroups, one grid barrier a
 * column); 0 the per-column launches (pivot search, scale and rank-1 update
 * per column; FP32 matrices up to 2048 rows keep their one-launch register
 * panels either way). Same pivots and results; 0 exists for tools that do not
 * handle cooperative launches (rocprofv3 counter collection ends the process
 * with a segmentation fault inside the HIP runtime's cooperative launch). No
 * counterpart in the reference (MKL's ZGETRF). */
int hm_set_lu_panel(hm_ctx *ctx, int cooperative);

/* ---- learned surrogate of the evanescent-mode sums ----------------------------
 * Replaces, for 0.05 <= R/h < 0.5 and z, zeta in [-0.25 h, 0], the evanescent part of EIGENE
 * (FinGreen3D.f90:413-506): up to 50 K0/K1 terms and three Wynn
 * epsilon limits per pair. A 4-64-64-64-64-3 SiLU network of (R/h, (z+zeta)/h,
 * (z-zeta)/h, Kh) predicts the real regular part (G, dG/dR, dG/dz of the
 * evanescent sums minus 1/r + 1/r1, smooth for R/h >= 0.05); the propagating
 * mode (the whole imaginary part) stays exact FP64. Hidden and output layers use BF16 weights and activations with
 * FP32 accumulation; the first layer is FP32.
 *
 * hm_surrogate_train  generates nsamples training samples with the plain FP64
 *                     series on the GPU (h = 1, log-uniform R/h in [0.05, 0.5] and
 *                     Kh in [0.01, 50] over 2048 Kh values, z, zeta in
 *                     [-0.25 h, 0], denser near the surface), trains with Adam
 *                     (cosine learning rate from lr,
 *                     batch `batch`, `epochs` passes) by the library's own forward
 *                     and backward kernels (quant = 1: forward with BF16 weights
 *                     and activations as the inference runs it; 0: FP32), then
 *                     measures the error against the
 *                     plain series on an independent validation set (nsamples/5,
 *                     at least 2e5, on 512 other Kh values) and stores, per cell
 *                     of a 24 x 8 x 8 x 24 grid over the four inputs, 1.5 x the
 *                     largest relative error (against the size of the full Green
 *                     function component). stats[8]: final training loss
 *                     (normalized MSE), validation RMS error (normalized), largest
 *                     validation error, seconds, fraction of cells with bound
 *                     <= 1e-4, <= 1e-3, Adam steps, samples.
 * hm_surrogate_get/set  the model as hm_surrogate_blob_size() floats
 *                     (parameters, output scales, error map).
 * hm_set_surrogate    use it in hm_assemble for finite depth: an EIGENE pair
 *                     takes the network when its cell bound is <= tol, otherwise
 *                     the plain series (also always for points deeper than
 *                     0.25 h); R/h < 0.05 and R/h >= 0.5 always take the plain
 *                     series.
 * hm_surrogate_counts pairs assembled since hm_set_surrogate: [0] by the network,
 *                     [1] EIGENE pairs left to the plain series, [2] EIGEN pairs
 *                     (plain series).
 * hm_surrogate_eval   network and plain regular parts (h = 1) at n points, with
 *                     the cell bound of each point (t_net, t_plain: 3n; bound: n).
 * hm_surrogate_grad   loss and gradient of the loss (mean squared error of the
 *                     normalized outputs against scaled targets u) over one batch
 *                     for the current parameters; quant = 0 evaluates the network
 *                     without BF16 rounding. For checks against a CPU reference.
 * None of these has a counterpart in the reference, which has no surrogate. */
int hm_surrogate_blob_size(void);
int hm_surrogate_train(hm_ctx *ctx, long long nsamples, int epochs, int batch, double lr, unsigned long long seed,
                       int quant, double *stats);
int hm_surrogate_get(hm_ctx *ctx, float *blob, int n);
int hm_surrogate_set(hm_ctx *ctx, const float *blob, int n);
int hm_set_surrogate(hm_ctx *ctx, int enable, double tol);
int hm_surrogate_counts(hm_ctx *ctx, long long *counts);
int hm_surrogate_eval(hm_ctx *ctx, int n, const double *rh, const double *zs, const double *zd, const double *kh,
                      double *t_net, double *t_plain, double *bound);
int hm_surrogate_grad(hm_ctx *ctx, int batch, const float *x, const float *u, int quant, float *grad, double *loss);
/* network inference (the BF16 network on MFMA, as the assembly evaluates it)
 * for n normalised inputs x (4n) -> scaled outputs u (3n) */
int hm_surrogate_predict(hm_ctx *ctx, int n, const float *x, float *u);
/* Infinite-depth branch (no counterpart in the reference). The deep-water wave
 * part of INFGREEN3D (InfGreen_Appr.f90:104-183) is the Havelock function GF and
 * its KR-derivative GFh of (KR, K(z+zeta)) (HavelockGF :185): a second network
 * of the same architecture learns Re GF and Re GFh less their singular terms
 * (the log(d - v) of L0 and the (beta + KR)/(d - v) of Ls, d = K R1) from
 * (log KR, sin KR, cos KR, K(z+zeta)) over KR in [1e-3, 50], K(z+zeta) in
 * [-20, 0]; the imaginary parts are closed forms.
 * hm_surrogate_train_deep trains it (arguments and stats as hm_surrogate_train;
 * its error map: cells over the four inputs, against |GF| + 1/d
 * and the gradient size |GFh| + |GF| + 2/d + 1/d^2). hm_set_surrogate_deep: use it in the deep-water
 * assembly for the pairs in cells whose bound is <= tol; hm_surrogate_deep_count:
 * pairs it took since then. hm_surrogate_eval_deep: network and plain Re GF,
 * Re GFh (2 per point) at n points, with each point's cell bound. */
int hm_surrogate_train_deep(hm_ctx *ctx, long long nsamples, int epochs, int batch, double lr,
                            unsigned long long seed, int quant, double *stats);
int hm_set_surrogate_deep(hm_ctx *ctx, int enable, double tol);
int hm_surrogate_deep_count(hm_ctx *ctx, long long *taken);
int hm_surrogate_eval_deep(hm_ctx *ctx, int n, const double *kr, const double *kv, double *net, double *plain,
                           double *bound);

/* ---- batches of independent problems ------------------------------------------
 * The reference solves one problem at a time, calling the per-frequency chain
 * (assembly, ZGETRF/ZGETRS, forces; HAMS_Prog.f90:615-724) serially. A batch
 * keeps many problems (members; any panel counts, finite or infinite depth,
 * with or without interior waterplane points) resident in HBM and runs that chain for all of
 * them at once per call: one assembly launch per regime over the pairs of
 * every member (each wave's pairs from one member), grouped launches (each
 * member at its own order) for
 * the right-hand sides, the LU and the triangular solves, optionally captured
 * once in a HIP graph and replayed per frequency.
 * hm_batch_add_problem takes one member (arguments as hm_set_geometry,
 * hm_set_depth and hm_set_headings) and returns its index >= 0;
 * hm_batch_set_waterplane gives that member ni interior waterplane field points
 * (ixyz_p, 3 per point, as hm_set_geometry's): it then solves with
 * irregular-frequency removal, the least squares of ASSB_LEFT_IRR /
 * RADIATION_SOLVER_IRR (AssbMatx_irr.f90:53-180, :472) through the normal
 * equations C = A^T A, D = A^T B (plain transpose), as hm_assemble does.
 * hm_batch_prepare(precision HM_FP64, HM_FP32_IR or HM_GMRES, autotune,
 * use_graph, nrefine): allocates, classifies, times the series variants
 * (lane-group shapes) of the assembly when autotune != 0 and keeps the
 * fastest, and captures the chain when use_graph != 0. HM_FP32_IR factors in
 * FP32 and runs nrefine FP64-residual refinement sweeps inside the chain, then
 * more (at most 10) until every member's last correction is below 1e-13 of
 * its solution. HM_GMRES (members without waterplane points) captures the
 * assembly and the FP32 copies of A, then solves each member by the restarted
 * GMRES of hm_ziter_solve's isolv = 1 form (ITERATIVE_SOLVER_MULTI,
 * IterativeSolvers.f90:232-437: FP32 Krylov products, FP64 residuals, no
 * preconditioner) one member after another, and forms the forces.
 * hm_batch_solve(w1[member] > 0): outputs stacked by member in member order
 * (amas, bdmp: (6 nbody)^2 each; exfc: 6 nbody x nbeta complex). A positive
 * return is a zero pivot of some member (hm_batch_error_string names it).
 * hm_batch_info: members, largest order, chosen series variant, assembly
 * workgroups per frequency, tuned assembly time, the last refinement
 * correction, refinement sweeps run after the graph, device bytes held. */
typedef struct hm_batch hm_batch;
int hm_batch_create(hm_batch **batch);
int hm_batch_destroy(hm_batch *batch);
const char *hm_batch_error_string(const hm_batch *batch);
int hm_batch_set_constants(hm_batch *batch, double rho, double g, double g_disp, double amp, double xw, double yw);
int hm_batch_add_problem(hm_batch *batch, int nbody, const int *body_start, int ntnd, const double *xyz,
                         const int *ncn, const int *ncon, const double *xyz_p, const double *ds, const double *pnsz,
                         const double *dxyz_p, double h, int nbeta, const double *beta);
int hm_batch_set_waterplane(hm_batch *batch, int member, int ni, const double *ixyz_p);
int hm_batch_prepare(hm_batch *batch, int precision, int autotune, int use_graph, int nrefine);
int hm_batch_solve(hm_batch *batch, const double *w1, double *amas, double *bdmp, double *exfc);
int hm_batch_info(hm_batch *batch, int *members, int *nmax, int *series_variant, int *grid_blocks,
                  double *tuned_assembly_ms, double *last_correction, int *extra_sweeps, long long *device_bytes);
/* What hm_batch_prepare's autotuner chose: the series variant, the K chunk of
 * the grouped GEMM tiles (8, 16 or 32 complex columns staged through LDS per
 * step; default 16 without autotuning), the assembly time of the chosen
 * variant and the grouped LU's time at K chunks 8, 16, 32 (lu_ms[3]; zeros
 * without autotuning). No counterpart in the reference. */
int hm_batch_tuning(hm_batch *batch, int *series_variant, int *gemm_kchunk, double *assembly_ms, double *lu_ms);
/* HM_GMRES members: restart length, tolerance on ||B - A X|| / ||B|| per
 * column and iteration cap (before hm_batch_prepare; defaults 50, 1e-6, 500),
 * and after a solve the largest iteration and restart counts over the members,
 * whether every member converged and the largest final relative residual.
 * No counterpart in the reference (its GMRES is per problem). */
int hm_batch_set_gmres(hm_batch *batch, int restart, double tol, int maxiter);
int hm_batch_gmres_stats(hm_batch *batch, int *max_iterations, int *max_cycles, int *all_converged,
                         double *worst_rel_residual);

/* Stage timing: when enabled, a pair of HIP events is recorded on the context's
 * stream around each stage (no extra synchronisation); hm_get_timing waits for
 * the last one and accumulates elapsed milliseconds into ms[stage], calls into
 * counts[stage], for stages 0 classify, 1 assembly (all regimes), 2 RHS GEMM,
 * 3 IRR normal equations, 4 LU, 5 triangular solves + refinement, 6 forces,
 * 7..10 the assembly kernel of each finite-depth regime (Linton, Pidcock,
 * epsilon-accelerated, eigenfunction), 11 the surrogate assembly kernel, 12
 * the cross approximation of the H-matrix form (its stream, start to end), 13
 * its entry kernels (one count per launch, on that stream).
 * hm_set_timing resets the totals. */
int hm_set_timing(hm_ctx *ctx, int enable);
int hm_get_timing(hm_ctx *ctx, double *ms, int *counts, int nstage);

/* Pair classification of the current geometry and depth: counts[regime*128 + len]
 * = number of (field point, source panel) pairs in regime 0 Linton, 1 Pidcock,
 * 2 epsilon-accelerated, 3 eigenfunction with series length len (0 when the
 * regime has no fixed length); nkeys >= 512. All zero for infinite depth. */
int hm_get_pair_histogram(hm_ctx *ctx, long long *counts, int nkeys);

/* Host copies of the assembled system (after hm_assemble, before hm_factor):
 * amat, smat: (n+ni) x n complex; rhs: (n+ni) x (6nb+nbeta) complex. Any may be
 * NULL. hm_get_potentials: n x (6nb+nbeta) complex after hm_solve (MXPOT). */
int hm_get_system(hm_ctx *ctx, double *amat, double *smat, double *rhs);
int hm_get_potentials(hm_ctx *ctx, double *mxpot);
/* With waterplane panels (irregular-frequency removal), after hm_assemble: the
 * normal equations the solve uses, CMAT = A^T A (n x n, AssbMatx_irr.f90:153-169)
 * and A^T of the right-hand sides (n x (6nb+nbeta): DRMAT, DDMAT,
 * AssbMatx_irr.f90:300-318, :448-462), column-major complex; either may be NULL. */
int hm_get_normal_equations(hm_ctx *ctx, double *cmat, double *drhs);

/* ---- element-wise kernels (parity entry points, host arrays) ----------------
 * FINGREEN3D(XF,XP,YF,YP,ZF,ZP,V,WVN,NK,H,GRN,FLAG)  FinGreen3D.f90:81
 * INFGREEN3D(XF,XP,YF,YP,ZF,ZP,V,GRN,FLAG)           InfGreen_Appr.f90:104
 * evaluated for n pairs at once; grn is 4 complex per pair (8n doubles). */
int hm_fingreen3d_batch(hm_ctx *ctx, int n, const double *xf, const double *xp, const double *yf, const double *yp,
                        const double *zf, const double *zp, double v, const double *wvn, int nk, double h,
                        const int *flag, double *grn);
int hm_infgreen3d_batch(hm_ctx *ctx, int n, const double *xf, const double *xp, const double *yf, const double *yp,
                        const double *zf, const double *zp, double v, const int *flag, double *grn);
/* No counterpart in the reference: hm_fingreen3d_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same arguments, results bit for bit the
 * same; the inputs and outputs are staged in a fixed device array that each GPU
 * holds from the first context created on it until the process ends
 * (30,410,496 bytes), so n is at most 262144 pairs per call (larger n returns
 * HM_ERR_ARG - 2). Calls that overlap on one GPU share that array and race. */
int hm_fingreen3d_batch_noalloc(hm_ctx *ctx, int n, const double *xf, const double *xp, const double *yf,
                                const double *yp, const double *zf, const double *zp, double v, const double *wvn,
                                int nk, double h, const int *flag, double *grn);
/* SGLINTBD_QUAD/TRI(IS,IEL,JEL,SIJ,DIJ,IRR)  SingularIntgr.f90:252/:44, with the
 * source panel's vertices verts[12*i] (ncn[i] of them) and field point xf[3*i]. */
int hm_sglintbd_batch(hm_ctx *ctx, int n, const int *ncn, const double *verts, const double *xf, double *sij,
                      double *dij);
/* CalPotentialMulti / CalPressureMulti / CalElevationMulti(XP,RDFLG,MD,...)
 * PressureElevationMulti.f90:55, :305, :329, for every radiation mode
 * (column 6 b + d: body b, degree of freedom d) and heading at once, after
 * hm_solve: potential, pressure (i w rho phi) and free-surface elevation
 * (i w / g phi) at nfp field points xfp (3 per point), each an nfp x (6 nbody +
 * nbeta) column-major complex array (any may be NULL). The free term is 2 pi on a
 * hull panel and 4 pi elsewhere; the diffraction columns include the incident
 * potential (ISOL = 1). */
int hm_field_points(hm_ctx *ctx, int nfp, const double *xfp, double *pot, double *prs, double *elv);
/* SolveMotionMulti(W1,TP,OUFR,BETA,AMP,AMAS,BDMP,BLNR,BQDR,EXFC,DSPL,NBODY)
 * SolveMotionMulti.f90:46, with the body matrices the reference keeps in its
 * module (MATX_MULTI, CRS_MULTI, KSTF_MULTI) passed explicitly: amas, bdmp
 * (6nbody x 6nbody), blnr, bqdr, matx, crs, kstf (nbody x 6 x 6, Fortran order),
 * exfc (6nbody complex, one heading) -> dspl (6nbody complex). With quadratic
 * damping the fixed-point iterations taken are returned in *iters (0 without);
 * the solves are the FP64 LU on the device; a zero pivot returns its column. */
int hm_solve_motion(hm_ctx *ctx, int nbody, double w1, const double *amas, const double *bdmp, const double *blnr,
                    const double *bqdr, const double *matx, const double *crs, const double *kstf, const double *exfc,
                    double *dspl, int *iters);
/* IK01A(X,BI0,DI0,BI1,DI1,BK0,DK0,BK1,DK1)  FinGrnExtSubs.f90:538, the K0 and K1
 * outputs, at n points: mode 0 evaluates as the reference does (the plain
 * path's kernels), mode 1 as the mixed-precision kernels do (the same series
 * as fixed-degree polynomials, equal to rounding). */
int hm_ik01a_batch(hm_ctx *ctx, int n, const double *x, double *bk0, double *bk1, int mode);
/* DISPERSION(WVN,NRT,W,H)  WavDynSubs.f90:40 (host; g_disp = 9.807 in the reference) */
int hm_dispersion(double *wvn, int nrt, double w, double h, double g_disp);

/* ---- dense complex LU with LAPACK argument lists (host arrays) --------------
 * ZGETRF(M,N,A,LDA,IPIV,INFO) and ZGETRS(TRANS,N,NRHS,A,LDA,IPIV,B,LDB,INFO)
 * as called at AssbMatxMulti.f90:112,:279,:346 (square, TRANS='N'); ipiv is
 * 1-based. The factorisation runs in FP64 on the GPU. */
int hm_zgetrf(hm_ctx *ctx, int m, int n, double *a, int lda, int *ipiv, int *info);
int hm_zgetrs(hm_ctx *ctx, char trans, int n, int nrhs, const double *a, int lda, const int *ipiv, double *b, int ldb,
              int *info);

#ifdef __cplusplus
}
#endif
#endif
