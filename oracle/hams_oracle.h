/*
 * hams_oracle.h -- CPU restatement of HAMS-MREL's frequency-domain panel solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Array conventions match the product C-ABI (include/hamsx.h): Fortran
 * column-major, complex numbers interleaved (re, im), connectivity 1-based.
 *   node_xyz[n + nnode*k]   cen[e + nelem*k]   dxyz[e + nelem*m]   ncon[e + nelem*c]
 *   A[i + N*j] (complex)    rhs[i + N*r] (complex)
 */
#ifndef HAMS_ORACLE_H
#define HAMS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar special functions (FinGrnExtSubs.f90, InfGreen_Appr.f90) ---- */
double orc_error(double x);                       /* erf, ERROR  FinGrnExtSubs.f90:62-93 */
double orc_e1xa(double x);                        /* E1XA        FinGrnExtSubs.f90:198-213 */
double orc_enxa(int n, double x);                 /* ENXA        FinGrnExtSubs.f90:135-157 */
double orc_eix(double x);                         /* EIX         FinGrnExtSubs.f90:254-282 */
void   orc_ik01a(double x, double *k0, double *k1);                       /* FinGrnExtSubs.f90:538-628 */
void   orc_jy01b(double x, double *j0, double *j1, double *y0, double *y1); /* FinGrnExtSubs.f90:663-705 */
double orc_limes(int n, double *s, int *i_pade, int *k_pade, double *err); /* FinGrnExtSubs.f90:370-503 */
double orc_struve_h0(double x);                   /* InfGreen_Appr.f90:765-818 */
double orc_struve_h1(double x);                   /* InfGreen_Appr.f90:821-872 */
double orc_bessel_j0(double x);                   /* InfGreen_Appr.f90:583-626 */
double orc_bessel_j1(double x);                   /* InfGreen_Appr.f90:629-672 */

/* ---- wave setup ---- */
void orc_dispersion(double w, double h, int nrt, double *wvn);           /* WavDynSubs.f90:40-127 */

/* ---- Green functions: grn[8] = 4 complex (G, dG/dx, dG/dy, dG/dz) ---- */
void orc_infgreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, int flag, double *grn);                     /* InfGreen_Appr.f90:104-186 */
void orc_fingreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, const double *wvn, int nk, double h, int flag,
                    double *grn);                                          /* FinGreen3D.f90:81-155 */

/* ---- Newman panel integral: xv[k*3+d] vertex k coordinate d (ncn = 3 or 4) ---- */
void orc_sglint(int ncn, const double *xv, const double *xf, double *sij, double *dij);

/* ---- incident wave: DINP / VINP (PatclVelct.f90:42-123) ---- */
void orc_dinp(double x, double y, double z, double xw, double yw, double beta,
              double v, double w1, double wk, double h, double amp, double *dpo /*6*/);
void orc_vinp(double x, double y, double z, double xw, double yw, double beta,
              double v, double w1, double wk, double h, double amp, double *phi /*2*/);

/* ---- geometry: NormalProcess.f90 ---- */
void orc_panel_geometry(int nnode, const double *node_xyz, int nelem, const int *ncn,
                        const int *ncon, const double *xr, double *cen, double *area,
                        double *pnsz, double *dxyz);

/* ---- influence matrices (CALGREEN_MULTI + ASSB_LEFT_MULTI + S operator) ----
 * A[i + N*j] = 2*pi*delta_ij + double-layer term;  S[i + N*j] = single-layer term.   */
void orc_assemble(int nelem, int nnode, const double *node_xyz, const int *ncn,
                  const int *ncon, const double *cen, const double *area,
                  const double *pnsz, const double *dxyz, double v, double h,
                  const double *wvn, int nk, double *A, double *S, int nthreads);

/* One (i, j) entry pair of the above (0-based), for spot checks on big meshes. */
void orc_assemble_entry(int i, int j, int nelem, int nnode, const double *node_xyz,
                        const int *ncn, const int *ncon, const double *cen,
                        const double *area, const double *pnsz, const double *dxyz,
                        double v, double h, const double *wvn, int nk, double *a_ij,
                        double *s_ij);

/* Rows rows[0..nrows) (0-based) of A and S into A[r + nrows*j], S[r + nrows*j]. */
void orc_assemble_rows(int nrows, const int *rows, int nelem, int nnode, const double *node_xyz,
                       const int *ncn, const int *ncon, const double *cen, const double *area,
                       const double *pnsz, const double *dxyz, double v, double h,
                       const double *wvn, int nk, double *A, double *S, int nthreads);

/* ---- right-hand sides: ASSB_RBC_MULTI / ASSB_DBC_MULTI ----
 * rhs[:, 0:6*nbody] radiation, rhs[:, 6*nbody + b] diffraction for heading beta[b]. */
void orc_rhs(int nelem, const double *S, const double *cen, const double *dxyz,
             int nbody, const int *body_start, int nbeta, const double *beta,
             double v, double w1, double wk, double h, double amp, double xw0,
             double xw1, double *rhs);

/* ---- dense complex LU (ZGETRF semantics, IZAMAX = |re|+|im|) and solve ---- */
int  orc_zgetrf(int n, double *A, int *ipiv);                 /* returns INFO */
void orc_zgetrs(int n, const double *LU, const int *ipiv, int nrhs, double *B);

/* ---- force reductions: RFORCE_MULTI / EFORCE_MULTI ----
 * amas, bdmp: (6nb x 6nb) column-major; exfc: (6nb x nbeta) complex column-major. */
void orc_forces(int nelem, const double *phi, const double *cen, const double *area,
                const double *dxyz, int nbody, const int *body_start, int nbeta,
                const double *beta, double v, double w1, double wk, double h,
                double amp, double xw0, double xw1, double rho, int isol,
                double *amas, double *bdmp, double *exfc);

/* ---- one whole frequency solve; returns LU INFO ---- */
int orc_solve_frequency(int nelem, int nnode, const double *node_xyz, const int *ncn,
                        const int *ncon, const double *cen, const double *area,
                        const double *pnsz, const double *dxyz, int nbody,
                        const int *body_start, double w1, double h, int nbeta,
                        const double *beta, double amp, double xw0, double xw1,
                        double rho, int isol, double *amas, double *bdmp, double *exfc,
                        int nthreads);

/* IRSP = 1: nlid lid centroids lid_cen[nlid*3] (column-major) as extra field
 * points; C = A^T A, D = A^T B, LU.  Outputs as orc_solve_frequency. */
int orc_solve_frequency_irr(int nelem, int nnode, const double *node_xyz, const int *ncn,
                            const int *ncon, const double *cen, const double *area,
                            const double *pnsz, const double *dxyz, int nbody,
                            const int *body_start, int nlid, const double *lid_cen, double w1,
                            double h, int nbeta, const double *beta, double amp, double xw0,
                            double xw1, double rho, int isol, double *amas, double *bdmp,
                            double *exfc, int nthreads);


/* ---- potentials of one frequency: phi[i + nelem*col] (MXPOT); returns INFO ---- */
int orc_solve_potentials(int nelem, int nnode, const double *node_xyz, const int *ncn,
                         const int *ncon, const double *cen, const double *area,
                         const double *pnsz, const double *dxyz, int nbody,
                         const int *body_start, double w1, double h, int nbeta,
                         const double *beta, double amp, double xw0, double xw1, double *phi,
                         int nthreads);

/* ---- field points: CalPotentialMulti / CalPressureMulti / CalElevationMulti
 * (PressureElevationMulti.f90:55-348).  xfp[p + nfp*d]; kind 0 potential,
 * 1 pressure, 2 elevation; out[p + nfp*col] complex for the 6 nbody radiation
 * modes and the nbeta headings; onbody[p] (may be NULL) the free-term test. */
int orc_on_body(const double *xp, int nelem, int nnode, const double *node_xyz, const int *ncn,
                const int *ncon, const double *pnsz, const double *dxyz);
void orc_field_values(int nelem, int nnode, const double *node_xyz, const int *ncn,
                      const int *ncon, const double *cen, const double *area,
                      const double *pnsz, const double *dxyz, int nbody, const int *body_start,
                      double w1, double h, int nbeta, const double *beta, double amp, double xw0,
                      double xw1, double rho, int isol, const double *phi, int nfp,
                      const double *xfp, int kind, double *out, int *onbody, int nthreads);
/* the same with one symmetry plane (isx = 1: x = 0, else y = 0) and the
 * symmetric solve's potentials phi2 (nelem x nr x 2) */
void orc_field_values_sym(int nelem, int nnode, const double *node_xyz, const int *ncn,
                          const int *ncon, const double *cen, const double *area,
                          const double *pnsz, const double *dxyz, int nbody, const int *body_start,
                          int isx, double w1, double h, int nbeta, const double *beta, double amp,
                          double xw0, double xw1, double rho, int isol, const double *phi2, int nfp,
                          const double *xfp, int kind, double *out, int *onbody, int nthreads);


/* ---- motions: SolveMotionMulti (SolveMotionMulti.f90:46-150) for one heading;
 * body matrices nbody x (6 x 6) column-major; returns the quadratic-damping
 * iterations (0: none needed, -1: not converged in 1000). ---- */
int orc_motion(int nbody, double w1, const double *amas, const double *bdmp, const double *exfc,
               const double *mass, const double *blnr, const double *bqdr, const double *crs,
               const double *kstf, double *dspl);


/* ---- NSYS = 2 (one symmetry plane: isx = 1 mirrors x, else y) on the half
 * mesh; outputs as orc_solve_frequency; phi2 (may be NULL): N x (6nb+nbeta) x 2
 * complex, the potentials of the mesh half and of its mirror. ---- */
int orc_solve_frequency_sym(int nelem, int nnode, const double *node_xyz, const int *ncn,
                            const int *ncon, const double *cen, const double *area,
                            const double *pnsz, const double *dxyz, int nbody,
                            const int *body_start, int isx, int nlid, const double *lid_cen,
                            double w1, double h, int nbeta, const double *beta, double amp,
                            double xw0, double xw1, double rho, int isol, double *amas,
                            double *bdmp, double *exfc, double *phi2, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
