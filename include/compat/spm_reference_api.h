/*
 * spm_reference_api.h -- the reference's C++ interface (JKGroupCUDA/SPM-IQC,
 * include/*.h) over the MI355X library.  Same names, parameter types and
 * call order; every fftw_complex field pointer is a DEVICE pointer to
 * CplxDofs_local = N^3 interleaved complex doubles, of which only the real
 * part carries a field, as in the reference.  One process holds the whole
 * grid: numprocs = 1, myid = 0, local_n0 = N, local_0_start = 0.
 *
 * Global state of Data.h:4-44; functions of Initialization.h:8-23,
 * My_FFT.h:10-14, LPsrc.h:5-7, LOBPCG.h:7-15, SPM.h:7-10 (+ MinMax_Forces,
 * SPM.cpp:103), BasicOperators.h:6-50 (the operators the drivers use),
 * Memfree.h:6-14 and File_Op.h:7-13.  Errors abort the process with a
 * message (the reference has no error returns either).
 */
#ifndef SPM_REFERENCE_API_H
#define SPM_REFERENCE_API_H

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

typedef double fftw_complex[2];
typedef struct spm_compat_plan *fftw_plan; /* plans live inside the device context */

/* --- Data.h:4-28 */
extern int *NCpt, *K, dimCpt, dimPhy, max_iter;
extern double L;
extern int **kindex;
extern double **ProjMatrix;
extern double Gamma, Tau, q, dt;
extern fftw_complex *u_F, *u_R; /* host: whole-grid fields for input/output */
extern fftw_complex *data;      /* device scratch field */
extern fftw_complex *u_R_local, *u_F_local, *u2_R_local, *u3_R_local, *u4_R_local;
extern fftw_complex *nln_R_local, *nln_F_local, *Laplas_R_local, *Laplas_F_local;
extern fftw_complex *fftw_Ctmp_local, *fftw_Rtmp_local;
extern fftw_complex **Phi0, **dPhi, *tv;
extern int **kindex_local;
extern double **PK_local, *k_square_local, *epmckt2_local, *ikt2_local; /* host tables */
extern int myid, numprocs;
extern int N, CplxDofs;
extern double PI;
extern int CplxDofs_local;
extern ptrdiff_t alloc_local, local_n0, local_0_start;
extern int Ini_Phase;
extern fftw_plan plan1, plan2;

/* Data.h:31-43 */
struct Option {
    int k;
    int maxiter;
    int outputp;
    int outputd;
    double dt;
    double epsf;
    double betat;
    double betau;
    fftw_complex **V; /* k device fields */
};
extern Option *option1, *option2;

/* --- Initialization.h:8-23 */
void memAllocation_all();
void memAllocation_local();
void LPsystParameter();
void get_kindex_all();
void get_ikt2_local();
void get_ksquare_local();
void get_epmckt2_local();
double get_GK(double k2);
void get_initial_value(fftw_complex *v, int choice, double angle); /* host v (whole grid) */
void Distribute_C(int **A, int **a);
void Distribute_C(fftw_complex *A, fftw_complex *a); /* host whole grid -> device local */
/* Device local -> host whole grid: the single-process counterpart of the
 * drivers' MPI_Gather (main_SPM.cpp:133, grad.cpp:132-133). */
void Collect_C(fftw_complex *a, fftw_complex *A);

/* --- My_FFT.h:10-14 */
void my_fft(fftw_complex *x_R_local, fftw_complex *x_F_local);
void my_ifft(fftw_complex *x_F_local, fftw_complex *x_R_local);
void My_mpi_fft_set_up();

/* --- LPsrc.h:5-7 */
double get_Ene();
void get_min_u_F();
void ngrad_cam(fftw_complex *phir, fftw_complex *gradr);

/* --- LOBPCG.h:7-15 */
void hv_cam(fftw_complex *phir, fftw_complex *v, fftw_complex *hv);
void precond_camnew(fftw_complex *w);
void hiosd_lobpcg_initialization(Option *option, int ki);
void hiosd_lobpcg_inip(fftw_complex *x, Option *option);
void copy_option_V(Option *option1, Option *option2, int k);
/* HISD.cpp:13-485 (include/HISD.h) */
void critical_point_HISD();
void critical_point_HISD_v();
void hiosd_lobpcg_sieh(fftw_complex *x, Option *option);
/* GOSD.cpp:13-531 (include/GOSD.h); V0: k0 device fields */
void critical_point();
void critical_point_v();
void hiosd_lobpcg_sieh_V0(fftw_complex *x, Option *option, fftw_complex **V0, int k0);
void copy_option_V(fftw_complex *v, Option *option2, int k);
void copy_option_V(fftw_complex **V, Option *option2, int k);

/* --- SPM.h:7-10 and SPM.cpp:103 */
void get_Tv(fftw_complex **Phi, fftw_complex *v);
void Natural_Forces(fftw_complex *x, fftw_complex *v);
void MinMax_Forces(fftw_complex *x, fftw_complex *v);
void Spring_Pairing(fftw_complex **Phi, double maxiter, double tol);

/* --- BasicOperators.h:6-50, all but the three functions that pass Eigen's
 * MatrixXd (get_mat, sort_eig; their work happens inside spm_lobpcg).  double*
 * fields are device arrays of CplxDofs_local reals; the Print functions copy
 * their field to the host to print it. */
void FuncsLinear1Cplx(fftw_complex *rslt, int n, const double a1, const fftw_complex *F1);
void FuncsLinear2Cplx(fftw_complex *rslt, int n, const double a1, const fftw_complex *F1, const double a2,
                      const fftw_complex *F2);
void FuncCplxAddAConst(fftw_complex *rslt, int n, const double a);
void FuncRealAddAConst(fftw_complex *rslt, int n, const double a);
void Print(fftw_complex *x, int n);
void Print(double *x, int n);
void Print(fftw_complex *x);
void Print(double *x);
void Cpy(fftw_complex *a, fftw_complex *b, int n);
void norm_fft(fftw_complex *x, int n);
double Max(double *x, int n);
double mean_Cplx(fftw_complex *src, int n);
void FFT_dot_constant(fftw_complex *rslt, int n, double a1);
void Dot_constant(double *rslt, int n, double a1);
void setCplxZero(fftw_complex rslt);
void setCplxZero_V(fftw_complex **W, int k, int cplxDofs);
void setCplxZero_v(fftw_complex *x, int cplxDofs);
void Orthoi_v2V(fftw_complex *v, fftw_complex **V, int k);
void Orthoi_Vi(fftw_complex **V, int k);
void get_Merge(double **UU, fftw_complex **V, fftw_complex **W, fftw_complex **Vp, int k);
int get_nozero(int *emp, int k);
void Printf_begin_end(fftw_complex *x);
void PV_Cplx(fftw_complex *v, fftw_complex **V, fftw_complex *tmpp, int k);
void PV_Cplx(fftw_complex *v, fftw_complex *V, fftw_complex *tmpp, int k);
double inpfp(fftw_complex *p, fftw_complex *q);
double inpfp(double *p, double *q);
double nrmp(fftw_complex *p);
double nrmp(double *p);
double rayleighq(fftw_complex *x, fftw_complex *v);

/* --- Memfree.h:6-14 */
void Memfree_global_kspace();
void Memfree_global_data();
void Memfree_local_kspace();
void Memfree_local_data();
void Memfree_Option(Option *option);

/* --- File_Op.h:7-13 (host side; Option::V is copied through the host) */
void Save_Option_V(Option *option, char *filename);
void Save_phi(fftw_complex *phi, char *filename); /* host whole-grid field */
void load_phi(fftw_complex *x, char *filename);
void load_Option_V(Option *option, char *filename);

/* Seconds since an arbitrary origin (stands in for MPI_Wtime in drivers). */
double spm_compat_wtime();
/* The C-ABI context behind the globals (valid after My_mpi_fft_set_up). */
struct spm_ctx *spm_compat_context();

#endif
