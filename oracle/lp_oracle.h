/*
 * lp_oracle.h -- CPU FP64 restatement of the reference's pseudo-spectral
 * Landau free-energy path (JKGroupCUDA/SPM-IQC).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against, and the CPU baseline that bench.py
 * times beside it.  Nothing in the product (spm-iqc_amd/) links or calls it.
 *
 * Fields use the reference's storage: `double[2]` interleaved complex values
 * of which only the real part carries a real field, N^3 points in x-major
 * order idx = (i*N + j)*N + k (Initialization.cpp:109-120).  One process
 * holds the whole grid (numprocs = 1, local_n0 = N); loops over pencils and
 * over fixed blocks of points run on OpenMP threads, and every reduction sums
 * fixed-size blocks in a fixed order, so results do not depend on the thread
 * count.
 */
#ifndef LP_ORACLE_H
#define LP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef double lpo_cplx[2];
typedef struct lpo_ctx lpo_ctx;

/* LPsystParameter + get_kindex_all + get_ksquare_local + get_epmckt2_local +
 * get_ikt2_local (Initialization.cpp:7-42, 95-174).  The reciprocal metric is
 * k_i = kscale * n_i / L; the reference hard-codes kscale = 1
 * (ProjMatrix = 1/L, Initialization.cpp:35-38). */
lpo_ctx *lpo_create(int N, double L, double kscale, double Tau, double Gamma,
                    double dt, int nthreads);
void lpo_destroy(lpo_ctx *c);
int lpo_N(const lpo_ctx *c);
double lpo_get_GK(double k2);                              /* Initialization.cpp:150-166 */
void lpo_tables(const lpo_ctx *c, double *ksq, double *G, double *ikt2); /* copies of the R-point tables */

/* My_FFT.cpp:4-21 (c2c; forward normalised by 1/N^3, inverse unnormalised) */
void lpo_my_fft(lpo_ctx *c, const lpo_cplx *x_R, lpo_cplx *x_F);
void lpo_my_ifft(lpo_ctx *c, const lpo_cplx *x_F, lpo_cplx *x_R);

/* BasicOperators.cpp:121-172 */
double lpo_inpfp(lpo_ctx *c, const lpo_cplx *p, const lpo_cplx *q);
double lpo_nrmp(lpo_ctx *c, const lpo_cplx *p);
double lpo_mean(lpo_ctx *c, const lpo_cplx *p);

/* LPsrc.cpp:8-38: energy of (u_R, u_F). */
double lpo_get_Ene(lpo_ctx *c, const lpo_cplx *u_R, const lpo_cplx *u_F);

/* LPsrc.cpp:40-117: semi-implicit gradient flow on (u_R, u_F) in place.
 * Stops when res <= res_tol or after max_iter iterations (the reference uses
 * 1e-16 and 100000).  Per iteration it records the pre-step energy E, the
 * residual res = max|Re(u_F' - u_F)| and, if gnorm != NULL, nrmp(ngrad_cam(u_R))
 * of the pre-step state.  Returns the number of iterations done. */
int lpo_get_min_u_F(lpo_ctx *c, lpo_cplx *u_R, lpo_cplx *u_F, int max_iter,
                    double res_tol, double *E, double *res, double *gnorm);

/* LPsrc.cpp:119-173: gradr = -grad F(phir). */
void lpo_ngrad_cam(lpo_ctx *c, const lpo_cplx *phir, lpo_cplx *gradr);

/* LOBPCG.cpp:113-199 */
void lpo_hv_cam(lpo_ctx *c, const lpo_cplx *phir, const lpo_cplx *v, lpo_cplx *hv);
void lpo_precond_camnew(lpo_ctx *c, lpo_cplx *w);
double lpo_rayleighq(lpo_ctx *c, const lpo_cplx *x, const lpo_cplx *v); /* BasicOperators.cpp:174-185 */

/* LOBPCG.cpp:10-109 with the random block supplied by the caller:
 * start holds k*R doubles (the reference draws them from Eigen's
 * MatrixXd::Random, LOBPCG.cpp:32-42).  V receives k fields of R points. */
void lpo_lobpcg_initialization(lpo_ctx *c, int k, const double *start, lpo_cplx *V);

/* LOBPCG.cpp:207-487: maxiter iterations on V (k fields, in place).
 * ritz (k doubles, may be NULL) receives the lowest k Ritz values of the last
 * Rayleigh-Ritz step, which the reference prints every 10 iterations
 * (LOBPCG.cpp:406-410).  Returns the number of iterations done. */
int lpo_lobpcg_inip(lpo_ctx *c, const lpo_cplx *x, lpo_cplx *V, int k, int maxiter, double *ritz);

/* SPM.cpp:15-263 */
void lpo_get_Tv(lpo_ctx *c, const lpo_cplx *Phi0, const lpo_cplx *Phi1, lpo_cplx *v);
void lpo_Natural_Forces(lpo_ctx *c, lpo_cplx *x, const lpo_cplx *v);
void lpo_MinMax_Forces(lpo_ctx *c, lpo_cplx *x, const lpo_cplx *v);
/* Returns the number of iterations; *res_out receives the last residual. */
int lpo_Spring_Pairing(lpo_ctx *c, lpo_cplx *Phi0, lpo_cplx *Phi1, int maxiter,
                       double tol, double *res_out);

/* HISD.cpp:86-485 (hiosd_lobpcg_sieh): x and V (k fields) in place; returns
 * the number of iterations; *res_out receives the last nrmp(ngrad_cam(x)). */
int lpo_hiosd_lobpcg_sieh(lpo_ctx *c, lpo_cplx *x, lpo_cplx *V, int k, double dt, double betat, double betau,
                          double epsf, int maxiter, double *res_out);
/* GOSD.cpp:123-531 (hiosd_lobpcg_sieh_V0): the same, with the sweep kept
 * orthogonal to the k0 fixed fields V0 (orthonormal, read only). */
int lpo_hiosd_lobpcg_sieh_V0(lpo_ctx *c, lpo_cplx *x, lpo_cplx *V, int k, const lpo_cplx *V0, int k0, double dt,
                             double betat, double betau, double epsf, int maxiter, double *res_out);

/* Symmetric eigensolver for the small Rayleigh-Ritz problem (cyclic Jacobi).
 * A is n*n row-major, overwritten; w receives eigenvalues ascending, Z the
 * matching eigenvectors as columns (row-major n*n). */
void lpo_sym_eig(int n, double *A, double *w, double *Z);

#ifdef __cplusplus
}
#endif
#endif
