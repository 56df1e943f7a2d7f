/*
 * hm_oracle.h -- CPU restatement of the HAMS-MREL boundary-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the tests, the smoke
 * entry and the CPU-baseline leg of the benchmark compare the HIP path against.
 * The product library never links, loads or calls it.
 *
 * Conventions:
 *   - panel arrays are panel-major: xyz_p[3*i+k], dxyz_p[6*i+k], ncon[4*i+k]
 *     (ncon holds 1-based node numbers, as in the reference .pnl files);
 *   - node arrays are node-major: xyz[3*n+k];
 *   - complex matrices are column-major (Fortran layout), interleaved re/im.
 */
#ifndef HM_ORACLE_H
#define HM_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- special functions (src/FinGrnExtSubs.f90) ---------------------------- */
double hmo_error(double x);                       /* ERROR  :62  */
double hmo_e1xa(double x);                        /* E1XA   :198 */
double hmo_enxa(int n, double x);                 /* ENXA   :135 */
double hmo_eix(double x);                         /* EIX    :254 */
void   hmo_ik01a(double x, double *bk0, double *bk1);                 /* IK01A :538 */
void   hmo_jy01b(double x, double *j0, double *j1, double *y0, double *y1); /* JY01B :663 */
double hmo_limes(int n, double *s);               /* Limes  :370 (s overwritten) */

/* ---- dispersion (src/WavDynSubs.f90:40) ------------------------------------ */
void hmo_dispersion(double *wvn, int nrt, double w, double h, double g_disp);

/* ---- Green functions ---------------------------------------------------------
 * grn[0] = G, grn[1..3] = dG/dx, dG/dy, dG/dz at the source point (xf,yf,zf),
 * field point (xp,yp,zp). flag=1 adds the image 1/r1, flag=0 adds 1/r + 1/r1. */
void hmo_fingreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, const double *wvn, int nk, double h,
                    double complex grn[4], int flag);            /* FinGreen3D.f90:81  */
void hmo_fingreen3d_r(double rr, double zf, double zp, double v, const double *wvn,
                      int nk, double h, double complex grn[3], int tag); /* :160 */
void hmo_infgreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, double complex grn[4], int flag);  /* InfGreen_Appr.f90:104 */

/* ---- Rankine panel integral (src/SingularIntgr.f90:44 / :252) -------------- */
void hmo_sglintbd(int ncn, const double verts[12], const double xf[3],
                  double *sij, double dij[3]);

/* ---- operation counter used to define the algorithmic flop count of a Green
 * function evaluation (adds, multiplies, divides, square roots count 1; every
 * libm transcendental call counts HMO_TRANS_FLOPS). */
#define HMO_TRANS_FLOPS 20.0
void   hmo_flops_reset(void);
double hmo_flops_get(void);

/* ---- panel geometry (src/NormalProcess.f90) --------------------------------- */
void hmo_panel_geometry(int nelem, const int *ncn, const int *ncon, const double *xyz,
                        const double rg[3], double *xyz_p, double *ds, double *pnsz,
                        double *dxyz_p);

/* ---- one panel-frequency system ---------------------------------------------- */
typedef struct {
    int n;              /* hull panels (all bodies)                         */
    int ni;             /* interior waterplane panels (0: no IRR)           */
    int nbody;
    const int *body_start;     /* nbody+1 offsets into the hull panels       */
    const double *xyz;         /* hull nodes                                 */
    const int *ncn, *ncon;     /* hull connectivity (1-based)                */
    const double *xyz_p, *ds, *pnsz, *dxyz_p;   /* hull panel geometry       */
    const double *ixyz_p;      /* waterplane field points (ni*3)            */
    double h;           /* depth; < 0 means infinite depth                  */
    double rho, g, g_disp;
    double xw[2];       /* wave reference point                             */
    double amp;
    int nbeta;
    const double *beta;        /* headings [rad]                              */
} hmo_problem;

typedef struct {
    double *amas;       /* (6nb x 6nb), column-major, [row=force dof, col=mode] */
    double *bdmp;
    double complex *exfc;  /* (6nb x nbeta) column-major                      */
    double complex *amat;  /* optional: (n+ni) x n assembled matrix           */
    double complex *smat;  /* optional: (n+ni) x n single-layer matrix        */
    double complex *rhs;   /* optional: (n+ni) x (6nb+nbeta) right-hand sides */
    double complex *mxpot; /* optional: n x (6nb+nbeta) potentials            */
} hmo_result;

/* Full per-frequency chain: wave properties, Green tables, assembly (with the
 * IRR normal equations when ni>0), LU solve, force reductions. Returns the
 * LAPACK-style INFO of the factorisation (0 = ok). */
int hmo_solve_frequency(const hmo_problem *p, double w1, hmo_result *out);
/* Field-point potential, pressure and elevation (CalPotentialMulti,
 * CalPressureMulti, CalElevationMulti: PressureElevationMulti.f90:55-348) of
 * every radiation mode and heading at nfp points xfp (3 per point), after the
 * frequency solve; outputs nfp x (6 nbody + nbeta), column-major. */
int hmo_field_points(const hmo_problem *p, double w1, int nfp, const double *xfp, double complex *pot,
                     double complex *prs, double complex *elv);
/* Motion response of one heading (SolveMotionMulti, SolveMotionMulti.f90:46-101):
 * amas, bdmp (6nbody x 6nbody), body matrices nbody x 6 x 6 (Fortran order),
 * exfc -> dspl (6nbody); iters = quadratic-damping iterations. */
int hmo_solve_motion(int nbody, double w1, const double *amas, const double *bdmp, const double *blnr,
                     const double *bqdr, const double *matx, const double *crs, const double *kstf,
                     const double complex *exfc, double complex *dspl, int *iters);
/* The same chain at zero (which = 0) or infinite (which = 1) wave frequency. */
int hmo_solve_limit(const hmo_problem *p, int which, hmo_result *out);

/* Textbook partial-pivot LU (pivot by max |re|+|im|, as izamax) + solve. */
int  hmo_zgetrf(int n, double complex *a, int lda, int *ipiv);
void hmo_zgetrs(int n, int nrhs, const double complex *a, int lda, const int *ipiv,
                double complex *b, int ldb);

#ifdef __cplusplus
}
#endif
#endif
