// compat.cpp -- the reference's C++ interface (include/compat/spm_reference_api.h)
// implemented over the C ABI of libspm_iqc.so.
//
// The globals of Data.h point at DEVICE buffers of N^3 interleaved complex
// doubles (the reference's fftw_complex layout).  Each solver call takes the
// real part of its fields into the library's real layout, runs there, and
// writes the result back with a zero imaginary part.  The device context is
// created by My_mpi_fft_set_up and re-created whenever N, L, Tau, Gamma or dt
// change, because the reference reads those globals at every call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/compat/spm_reference_api.h"
#include "../../include/spm_iqc.h"

// ----------------------------------------------------------------- globals (Data.cpp:1-29)
int *NCpt, *K, dimCpt, dimPhy, max_iter;
double L;
int **kindex;
double **ProjMatrix;
double Gamma, Tau, q, dt;
fftw_complex *u_F, *u_R;
fftw_complex *data;
fftw_complex *u_R_local, *u_F_local, *u2_R_local, *u3_R_local, *u4_R_local;
fftw_complex *nln_R_local, *nln_F_local, *Laplas_R_local, *Laplas_F_local;
fftw_complex *fftw_Ctmp_local, *fftw_Rtmp_local;
int **kindex_local;
double **PK_local, *k_square_local, *epmckt2_local, *ikt2_local;
int myid, numprocs;
int N, CplxDofs;
double PI;
int CplxDofs_local;
ptrdiff_t alloc_local, local_n0, local_0_start;
int Ini_Phase;
fftw_complex **Phi0, **dPhi;
fftw_complex *tv;
fftw_plan plan1, plan2;
Option *option1, *option2;

// ----------------------------------------------------------------- internals
namespace {

struct cd_host {
    double re, im;
};

spm_ctx *g_ctx = nullptr;
double g_key[5];                   // N, L, Tau, Gamma, dt of g_ctx
double *g_re[4] = {nullptr, nullptr, nullptr, nullptr}; // device real scratch fields
double *g_ones = nullptr;
double *g_blk = nullptr;           // LOBPCG block (k real fields)
int g_blk_k = 0;
int *g_kflat = nullptr, *g_kflat_local = nullptr;
double *g_pkflat = nullptr;

[[noreturn]] void die(const char *what, int rc)
{
    std::fprintf(stderr, "spm compat: %s failed (%d): %s\n", what, rc, spm_last_error());
    std::abort();
}
#define CK(x)                                                                                       \
    do {                                                                                            \
        int _r = (x);                                                                               \
        if (_r) die(#x, _r);                                                                        \
    } while (0)
void hck(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "spm compat: %s: %s\n", what, hipGetErrorString(e));
        std::abort();
    }
}

size_t R() { return (size_t)CplxDofs_local; }

void free_scratch()
{
    for (double *&p : g_re)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
    if (g_ones) (void)hipFree(g_ones);
    if (g_blk) (void)hipFree(g_blk);
    g_ones = g_blk = nullptr;
    g_blk_k = 0;
}

// The device context for the current globals.
spm_ctx *ctx()
{
    const double key[5] = {(double)N, L, Tau, Gamma, dt};
    if (g_ctx && std::memcmp(key, g_key, sizeof key) == 0) return g_ctx;
    if (g_ctx) {
        spm_destroy(g_ctx);
        g_ctx = nullptr;
        free_scratch();
    }
    spm_params p{};
    p.N = N;
    p.precision = SPM_FP64;
    p.L = L;
    // ProjMatrix[j][j] = kscale / L (Initialization.cpp:35-38 has kscale = 1)
    p.kscale = ProjMatrix ? ProjMatrix[0][0] * L : 1.0;
    p.Tau = Tau;
    p.Gamma = Gamma;
    p.dt = dt;
    CK(spm_create(&p, &g_ctx));
    CK(spm_set_stream(g_ctx, nullptr)); // legacy stream: ordered with hipMemcpy
    std::memcpy(g_key, key, sizeof key);
    const size_t b = sizeof(double) * (size_t)N * N * N;
    for (double *&p2 : g_re) hck(hipMalloc(&p2, b), "hipMalloc scratch");
    std::vector<double> ones((size_t)N * N * N, 1.0);
    hck(hipMalloc(&g_ones, b), "hipMalloc ones");
    hck(hipMemcpy(g_ones, ones.data(), b, hipMemcpyHostToDevice), "upload ones");
    return g_ctx;
}

double *block(int k)
{
    if (g_blk_k < k) {
        if (g_blk) (void)hipFree(g_blk);
        hck(hipMalloc(&g_blk, sizeof(double) * k * R()), "hipMalloc block");
        g_blk_k = k;
    }
    return g_blk;
}

void to_real(const fftw_complex *z, double *x) { CK(spm_cplx_to_real(ctx(), z, x)); }
void to_cplx(const double *x, fftw_complex *z) { CK(spm_real_to_cplx(ctx(), x, z)); }

fftw_complex *dev_field()
{
    fftw_complex *p;
    hck(hipMalloc(&p, sizeof(fftw_complex) * R()), "hipMalloc field");
    hck(hipMemset(p, 0, sizeof(fftw_complex) * R()), "hipMemset field");
    return p;
}

// a1 F1 + a2 F2 on n complex values = 2n doubles, done as two halves of the
// context's field length (n must equal CplxDofs_local)
void lin2_cplx(fftw_complex *r, double a1, const fftw_complex *F1, double a2, const fftw_complex *F2, int n)
{
    if ((size_t)n != R()) {
        std::fprintf(stderr, "spm compat: vector length %d != CplxDofs_local %d\n", n, CplxDofs_local);
        std::abort();
    }
    for (int h = 0; h < 2; ++h)
        CK(spm_lin2(ctx(), (double *)r + h * R(), a1, (const double *)F1 + h * R(), a2,
                    (const double *)F2 + h * R()));
}

FILE *open_or_die(const char *name, const char *mode)
{
    FILE *fp = std::fopen(name, mode);
    if (!fp) {
        std::fprintf(stderr, "spm compat: cannot open %s\n", name);
        std::abort();
    }
    return fp;
}

} // namespace

// ----------------------------------------------------------------- Initialization
// Initialization.cpp:7-42.  N may be overridden with the environment variable
// SPM_N (the reference hard-codes 256); L, Tau, Gamma, dt and ProjMatrix can be
// changed through the globals before My_mpi_fft_set_up.
void LPsystParameter()
{
    Ini_Phase = 0;
    N = 256;
    if (const char *e = std::getenv("SPM_N")) N = std::atoi(e);
    CplxDofs = N * N * N;
    PI = 3.14159265358979311599796346854418516;
    dimCpt = 3, dimPhy = 3, max_iter = 100000, L = 26;
    dt = 0.3;
    Tau = -0.0061;
    Gamma = 0.30;
    if (myid == 0) {
        std::printf("\nset parameter: Tau = %f, Gamma = %f \n", Tau, Gamma);
        std::printf("\nset parameter: L = %f, N = %d \n", L, N);
    }
    q = 0;
    ProjMatrix = (double **)std::malloc(sizeof(double *) * dimPhy);
    for (int i = 0; i < dimPhy; i++) ProjMatrix[i] = (double *)std::calloc(dimCpt, sizeof(double));
    for (int j = 0; j < dimCpt; j++) ProjMatrix[j][j] = 1.0 / L;
    NCpt = (int *)std::malloc(sizeof(int) * dimCpt);
    for (int i = 0; i < dimCpt; i++) NCpt[i] = N;
}

// Initialization.cpp:80-91 (one block instead of N^3 small allocations)
void memAllocation_all()
{
    K = (int *)std::malloc(sizeof(int) * N);
    g_kflat = (int *)std::malloc(sizeof(int) * 3 * (size_t)CplxDofs);
    kindex = (int **)std::malloc(sizeof(int *) * (size_t)CplxDofs);
    for (size_t i = 0; i < (size_t)CplxDofs; i++) kindex[i] = g_kflat + 3 * i;
    u_F = (fftw_complex *)std::calloc((size_t)CplxDofs, sizeof(fftw_complex));
    u_R = (fftw_complex *)std::calloc((size_t)CplxDofs, sizeof(fftw_complex));
}

// Initialization.cpp:95-121
void get_kindex_all()
{
    for (int i = 0; i < N; i++) K[i] = i < N / 2 ? i : i - N;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
            for (int k = 0; k < N; k++) {
                const size_t l = ((size_t)i * N + j) * N + k;
                kindex[l][0] = K[i];
                kindex[l][1] = K[j];
                kindex[l][2] = K[k];
            }
}

// Initialization.cpp:44-78: field buffers on the device, k-space tables on the host
void memAllocation_local()
{
    const size_t n = R();
    g_kflat_local = (int *)std::malloc(sizeof(int) * 3 * n);
    g_pkflat = (double *)std::malloc(sizeof(double) * 3 * n);
    kindex_local = (int **)std::malloc(sizeof(int *) * n);
    PK_local = (double **)std::malloc(sizeof(double *) * n);
    for (size_t i = 0; i < n; i++) {
        kindex_local[i] = g_kflat_local + 3 * i;
        PK_local[i] = g_pkflat + 3 * i;
    }
    k_square_local = (double *)std::calloc(n, sizeof(double));
    epmckt2_local = (double *)std::calloc(n, sizeof(double));
    ikt2_local = (double *)std::calloc(n, sizeof(double));
    u_F_local = dev_field();
    u_R_local = dev_field();
    u2_R_local = dev_field();
    u3_R_local = dev_field();
    u4_R_local = dev_field();
    nln_R_local = dev_field();
    nln_F_local = dev_field();
    Laplas_R_local = dev_field();
    Laplas_F_local = dev_field();
    fftw_Ctmp_local = dev_field();
    fftw_Rtmp_local = dev_field();
    Phi0 = (fftw_complex **)std::malloc(sizeof(fftw_complex *) * 2);
    for (int i = 0; i < 2; i++) Phi0[i] = dev_field();
    tv = dev_field();
}

// Initialization.cpp:123-129
void get_ikt2_local()
{
    for (int i = 0; i < CplxDofs_local; i++) ikt2_local[i] = 1.0 / (epmckt2_local[i] + std::fabs(Tau) + 0.02);
}

// Initialization.cpp:132-148 (k_square_local starts from zero)
void get_ksquare_local()
{
    for (int l = 0; l < CplxDofs_local; l++) {
        k_square_local[l] = 0.0;
        for (int i = 0; i < dimPhy; i++) {
            double tmp = 0.0;
            for (int j = 0; j < dimCpt; j++) tmp = tmp + ProjMatrix[i][j] * kindex_local[l][j];
            PK_local[l][i] = tmp;
            k_square_local[l] = k_square_local[l] + std::pow(PK_local[l][i], 2);
        }
    }
}

double get_GK(double k2) { return spm_get_GK(k2); }

// Initialization.cpp:168-174
void get_epmckt2_local()
{
    for (int i = 0; i < CplxDofs_local; i++) epmckt2_local[i] = get_GK(k_square_local[i]);
}

// Initialization.cpp:176-205 (one rank: the local slab is the whole grid)
void Distribute_C(fftw_complex *A, fftw_complex *a)
{
    hck(hipMemcpy(a, A + (size_t)local_0_start * N * N, sizeof(fftw_complex) * R(), hipMemcpyHostToDevice),
        "Distribute_C");
}

void Distribute_C(int **A, int **a)
{
    for (size_t l = 0; l < R(); ++l)
        for (int d = 0; d < 3; ++d) a[l][d] = A[(size_t)local_0_start * N * N + l][d];
}

void Collect_C(fftw_complex *a, fftw_complex *A)
{
    hck(hipMemcpy(A, a, sizeof(fftw_complex) * R(), hipMemcpyDeviceToHost), "Collect_C");
}

// Initialization.cpp:242-338 (host; u_F is the whole-grid host field)
void get_initial_value(fftw_complex *v, int choice, double angle)
{
    if (myid == 0) std::printf("\nchoice = %d \t angle = %.3e ", choice, angle);
    static const char *names[] = {"DIS", "BCC", "FCC", "HPC", "LAM", "DG"};
    static const double lens[] = {1.0, std::sqrt(2.0), std::sqrt(3.0), std::sqrt(2.0), 1.0, std::sqrt(6.0)};
    if (choice < 0 || choice > 5) {
        std::fprintf(stderr, "spm compat: get_initial_value: choice %d not in 0..5\n", choice);
        std::abort();
    }
    if (myid == 0) std::printf("\n--------------------%s----------------------------\n\n", names[choice]);
    const double l = lens[choice];
    for (int i = 0; i < CplxDofs; i++) v[i][0] = v[i][1] = 0.0;
    const std::string path = std::string("./initData/") + names[choice];
    FILE *fp = open_or_die(path.c_str(), "r");
    int n = 0;
    if (std::fscanf(fp, "%d", &n) != 1) n = 0;
    std::vector<int> f(4 * (size_t)n, 0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 4; j++)
            if (std::fscanf(fp, "%d", &f[4 * i + j]) != 1) f[4 * i + j] = 0;
    std::fclose(fp);
    for (int i = 0; i < n; i++) {
        int k1 = (int)std::round(1.0 * f[4 * i + 0] / l * L);
        int k2 = (int)std::round(1.0 * f[4 * i + 1] / l * L);
        int k3 = (int)std::round(1.0 * f[4 * i + 2] / l * L);
        if (k1 < 0) k1 = N + k1;
        if (k2 < 0) k2 = N + k2;
        if (k3 < 0) k3 = N + k3;
        // the reference writes outside u_F when a mode does not fit the grid (N too
        // small for L); here that is an error
        if (k1 < 0 || k1 >= N || k2 < 0 || k2 >= N || k3 < 0 || k3 >= N) {
            std::fprintf(stderr, "spm compat: get_initial_value: mode %d of %s lies outside the %d^3 grid\n", i,
                         path.c_str(), N);
            std::abort();
        }
        v[((size_t)k1 * N + k2) * N + k3][0] = f[4 * i + 3] * 0.1;
    }
}

// ----------------------------------------------------------------- My_FFT (My_FFT.cpp)
void My_mpi_fft_set_up()
{
    alloc_local = CplxDofs;
    local_n0 = N;
    local_0_start = 0;
    CplxDofs_local = (int)(local_n0 * N * N);
    numprocs = numprocs > 0 ? numprocs : 1;
    ctx();
    data = dev_field();
    plan1 = plan2 = nullptr;
}

void my_fft(fftw_complex *x_R_local, fftw_complex *x_F_local) { CK(spm_my_fft(ctx(), x_R_local, x_F_local)); }
void my_ifft(fftw_complex *x_F_local, fftw_complex *x_R_local) { CK(spm_my_ifft(ctx(), x_F_local, x_R_local)); }

// ----------------------------------------------------------------- LPsrc
// get_Ene (LPsrc.cpp:8-38): the density term of u_R_local, the non-local term
// from the carried spectrum u_F_local (LPsrc.cpp:27-35), as the reference reads it.
double get_Ene()
{
    double E = 0.0;
    to_real(u_R_local, g_re[0]);
    CK(spm_energy_uf(ctx(), g_re[0], u_F_local, &E));
    return E;
}

// get_min_u_F (LPsrc.cpp:40-117) on (u_R_local, u_F_local)
void get_min_u_F()
{
    to_real(u_R_local, g_re[0]);
    std::vector<double> hist(3 * (size_t)(max_iter > 0 ? max_iter : 1));
    int it = 0;
    CK(spm_gradflow(ctx(), g_re[0], max_iter, 1e-16, 0.0, 1000, hist.data(), &it));
    if (myid == 0)
        for (int i = 0; i < it; i += 10)
            std::printf("\nmyid : %d ,\t iter : %d \t res: %.3e \t E: %.3e \t", myid, i, hist[3 * i + 2], hist[3 * i]);
    to_cplx(g_re[0], u_R_local);
    my_fft(u_R_local, u_F_local);
}

void ngrad_cam(fftw_complex *phir, fftw_complex *gradr)
{
    to_real(phir, g_re[0]);
    CK(spm_ngrad(ctx(), g_re[0], g_re[1], nullptr));
    to_cplx(g_re[1], gradr);
}

// ----------------------------------------------------------------- LOBPCG
void hv_cam(fftw_complex *phir, fftw_complex *v, fftw_complex *hv)
{
    to_real(phir, g_re[0]);
    to_real(v, g_re[1]);
    CK(spm_hv(ctx(), g_re[0], g_re[1], g_re[2]));
    to_cplx(g_re[2], hv);
}

void precond_camnew(fftw_complex *w)
{
    to_real(w, g_re[0]);
    CK(spm_precond(ctx(), g_re[0]));
    to_cplx(g_re[0], w);
}

// LOBPCG.cpp:10-109; the random block comes from a fixed-seed generator
// (the reference uses Eigen's MatrixXd::Random)
void hiosd_lobpcg_initialization(Option *option, int ki)
{
    option->k = ki;
    option->maxiter = 21;
    option->dt = 0.1;
    option->epsf = 1e-7;
    option->betat = 1.0;
    option->betau = 1.0;
    option->outputp = 1;
    option->outputd = 1;
    option->V = (fftw_complex **)std::malloc(sizeof(fftw_complex *) * ki);
    for (int i = 0; i < ki; i++) option->V[i] = dev_field();
    std::mt19937_64 gen(20240607u);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> st((size_t)ki * R());
    for (double &x : st) x = U(gen);
    double *blk = block(ki);
    hck(hipMemcpy(blk, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice), "upload start");
    CK(spm_lobpcg_init(ctx(), blk, blk, ki));
    for (int i = 0; i < ki; i++) to_cplx(blk + i * R(), option->V[i]);
}

// LOBPCG.cpp:207-487
void hiosd_lobpcg_inip(fftw_complex *x, Option *option)
{
    const int k = option->k;
    to_real(x, g_re[0]);
    double *blk = block(k);
    for (int i = 0; i < k; i++) to_real(option->V[i], blk + i * R());
    std::vector<double> ritz(k);
    CK(spm_lobpcg(ctx(), g_re[0], blk, k, option->maxiter, ritz.data()));
    for (int i = 0; i < k; i++) to_cplx(blk + i * R(), option->V[i]);
    if (myid == 0) {
        std::printf("\n%d ", option->maxiter - 1);
        for (int i = 0; i < k; i++) std::printf(" %.10g", ritz[i]);
        std::printf("\n\n");
    }
}

// HISD.cpp:86-485, on the device through spm_hisd.  As the reference, it
// leaves x in u_R_local and its spectrum in u_F_local (:446-448).
void hiosd_lobpcg_sieh(fftw_complex *x, Option *option)
{
    const int k = option->k;
    to_real(x, g_re[0]);
    double *blk = block(k);
    for (int i = 0; i < k; i++) to_real(option->V[i], blk + i * R());
    int it = 0;
    double res = 0.0;
    CK(spm_hisd(ctx(), g_re[0], blk, k, option->dt, option->betat, option->betau, option->epsf, option->maxiter, &it,
                &res));
    to_cplx(g_re[0], x);
    for (int i = 0; i < k; i++) to_cplx(blk + i * R(), option->V[i]);
    if (x != u_R_local) Cpy(x, u_R_local, CplxDofs_local);
    my_fft(u_R_local, u_F_local);
    if (myid == 0) std::printf("\n %d \t res = %.3e \n", it, res);
}

// HISD.cpp:13-30: index-1 saddle search from u_R_local + 0.01 V_1 of option1
void critical_point_HISD()
{
    const int k = 1;
    hiosd_lobpcg_initialization(option2, k);
    option2->dt = 0.8;
    option2->betat = 8.0;
    option2->betau = 0.02;
    option2->epsf = 1e-7;
    option2->maxiter = 10000;
    option2->outputp = 100;
    option2->outputd = 1000;
    FuncsLinear2Cplx(u_R_local, CplxDofs_local, 1.0, u_R_local, 0.01, option1->V[k - 1]);
    copy_option_V(option1, option2, k);
    hiosd_lobpcg_sieh(u_R_local, option2);
}

// HISD.cpp:32-82: the same with betau = 0.6.  The reference also reads the FCC
// seed into a field it then leaves unused; the read (and its file) is kept.
void critical_point_HISD_v()
{
    const int k = 1;
    hiosd_lobpcg_initialization(option2, k);
    option2->dt = 0.8;
    option2->betat = 8.0;
    option2->betau = 0.6;
    option2->epsf = 1e-7;
    option2->maxiter = 10000;
    option2->outputp = 100;
    option2->outputd = 1000;
    std::vector<fftw_complex> v((size_t)CplxDofs);
    get_initial_value(v.data(), 2, 0.0);
    FuncsLinear2Cplx(u_R_local, CplxDofs_local, 1.0, u_R_local, 0.01, option1->V[k - 1]);
    copy_option_V(option1, option2, k);
    hiosd_lobpcg_sieh(u_R_local, option2);
}

// GOSD.cpp:123-531 through spm_gosd: hiosd_lobpcg_sieh with the sweep kept
// orthogonal to the k0 device fields V0 (orthonormal); x is left in u_R_local
// and its spectrum in u_F_local, as in hiosd_lobpcg_sieh.
void hiosd_lobpcg_sieh_V0(fftw_complex *x, Option *option, fftw_complex **V0, int k0)
{
    const int k = option->k;
    to_real(x, g_re[0]);
    double *blk = block(k);
    for (int i = 0; i < k; i++) to_real(option->V[i], blk + i * R());
    double *v0 = nullptr;
    if (k0 > 0) {
        hck(hipMalloc(&v0, sizeof(double) * k0 * R()), "hipMalloc V0");
        for (int i = 0; i < k0; i++) to_real(V0[i], v0 + i * R());
    }
    int it = 0;
    double res = 0.0;
    const int rc = spm_gosd(ctx(), g_re[0], blk, k, v0, k0, option->dt, option->betat, option->betau, option->epsf,
                            option->maxiter, &it, &res);
    if (v0) (void)hipFree(v0); // spm_gosd has synchronised its stream
    CK(rc);
    to_cplx(g_re[0], x);
    for (int i = 0; i < k; i++) to_cplx(blk + i * R(), option->V[i]);
    if (x != u_R_local) Cpy(x, u_R_local, CplxDofs_local);
    my_fft(u_R_local, u_F_local);
    if (myid == 0) std::printf("\n %d \t res = %.3e \n", it, res);
}

// the real part of src into dst (the drivers' [i][0] = ..., [i][1] = 0 loops)
static void real_part_into(fftw_complex *dst, const fftw_complex *src)
{
    to_real(src, g_re[1]);
    to_cplx(g_re[1], dst);
}

// GOSD.cpp:13-58 and 60-120: V0 = the first k0 = 3 modes of option1, and the
// saddle search from u_R_local - 0.1 V_3 along V_3 (critical_point) or from
// u_R_local + 0.1 v along v, v the HEX seed in real space (critical_point_v).
static void gosd_driver(bool from_hex)
{
    const int k0 = 3;
    hiosd_lobpcg_initialization(option2, 1);
    option2->dt = 0.8;
    option2->betat = 8.0;
    option2->betau = from_hex ? 0.4 : 1.3;
    option2->epsf = from_hex ? 1e-7 : 1e-9;
    option2->maxiter = from_hex ? 20000 : 40000;
    option2->outputp = 100;
    option2->outputd = 1000;
    if (from_hex) {
        std::vector<fftw_complex> v((size_t)CplxDofs);
        get_initial_value(v.data(), 3, 0.0);
        fftw_complex *vF = dev_field(), *vR = dev_field();
        Distribute_C(v.data(), vF);
        my_ifft(vF, vR); // the driver's Cpy, fftw_execute(plan2), Cpy (My_FFT.cpp:16-21)
        FuncsLinear2Cplx(u_R_local, CplxDofs_local, 1.0, u_R_local, 0.1, vR);
        real_part_into(option2->V[0], vR);
        (void)hipFree(vF);
        (void)hipFree(vR);
    } else {
        FuncsLinear2Cplx(u_R_local, CplxDofs_local, 1.0, u_R_local, -0.1, option1->V[k0]);
        real_part_into(option2->V[0], option1->V[k0]);
    }
    fftw_complex *V0[k0];
    for (int i = 0; i < k0; i++) {
        V0[i] = dev_field();
        real_part_into(V0[i], option1->V[i]);
    }
    hiosd_lobpcg_sieh_V0(u_R_local, option2, V0, k0);
    for (int i = 0; i < k0; i++) (void)hipFree(V0[i]);
}

void critical_point() { gosd_driver(false); }
void critical_point_v() { gosd_driver(true); }

void copy_option_V(Option *o1, Option *o2, int k)
{
    for (int i = 0; i < k; i++) {
        to_real(o1->V[i], g_re[0]);
        to_cplx(g_re[0], o2->V[i]);
    }
}
void copy_option_V(fftw_complex **V, Option *o2, int k)
{
    for (int i = 0; i < k; i++) {
        to_real(V[i], g_re[0]);
        to_cplx(g_re[0], o2->V[i]);
    }
}
void copy_option_V(fftw_complex *v, Option *o2, int k)
{
    to_real(v, g_re[0]);
    to_cplx(g_re[0], o2->V[k]);
}

// ----------------------------------------------------------------- SPM
void get_Tv(fftw_complex **Phi, fftw_complex *v)
{
    to_real(Phi[0], g_re[0]);
    to_real(Phi[1], g_re[1]);
    CK(spm_get_tv(ctx(), g_re[0], g_re[1], g_re[2]));
    to_cplx(g_re[2], v);
}

void Natural_Forces(fftw_complex *x, fftw_complex *v)
{
    to_real(x, g_re[0]);
    to_real(v, g_re[1]);
    CK(spm_natural_force(ctx(), g_re[0], g_re[1]));
    to_cplx(g_re[0], x);
}

void MinMax_Forces(fftw_complex *x, fftw_complex *v)
{
    to_real(x, g_re[0]);
    to_real(v, g_re[1]);
    CK(spm_minmax_force(ctx(), g_re[0], g_re[1]));
    to_cplx(g_re[0], x);
}

// SPM.cpp:170-263; leaves the final tangent in the global tv as the reference does
void Spring_Pairing(fftw_complex **Phi, double maxiter, double tol)
{
    to_real(Phi[0], g_re[0]);
    to_real(Phi[1], g_re[1]);
    const int mi = (int)maxiter;
    std::vector<double> hist(mi > 0 ? mi : 1);
    int it = 0;
    double res = 0.0;
    CK(spm_spring_pairing(ctx(), g_re[0], g_re[1], mi, tol, &it, &res, hist.data()));
    if (myid == 0)
        for (int i = 10; i <= it; i += 10) std::printf("\n iter = %d , res =  %.3e \n", i, hist[i - 1]);
    to_cplx(g_re[0], Phi[0]);
    to_cplx(g_re[1], Phi[1]);
    get_Tv(Phi, tv);
}

// ----------------------------------------------------------------- BasicOperators
void FuncsLinear1Cplx(fftw_complex *rslt, int n, const double a1, const fftw_complex *F1)
{
    lin2_cplx(rslt, a1, F1, 0.0, F1, n);
}
void FuncsLinear2Cplx(fftw_complex *rslt, int n, const double a1, const fftw_complex *F1, const double a2,
                      const fftw_complex *F2)
{
    lin2_cplx(rslt, a1, F1, a2, F2, n);
}
void Cpy(fftw_complex *a, fftw_complex *b, int n)
{
    hck(hipMemcpy(b, a, sizeof(fftw_complex) * (size_t)n, hipMemcpyDeviceToDevice), "Cpy");
}
void FFT_dot_constant(fftw_complex *rslt, int n, double a1) { lin2_cplx(rslt, a1, rslt, 0.0, rslt, n); }
void setCplxZero_v(fftw_complex *x, int cplxDofs)
{
    hck(hipMemset(x, 0, sizeof(fftw_complex) * (size_t)cplxDofs), "setCplxZero_v");
}
double mean_Cplx(fftw_complex *src, int n)
{
    if ((size_t)n != R()) die("mean_Cplx: n != CplxDofs_local", SPM_EINVAL);
    double m = 0.0;
    to_real(src, g_re[0]);
    CK(spm_inpfp(ctx(), g_re[0], g_ones, &m));
    return m;
}
double inpfp(fftw_complex *p, fftw_complex *qv)
{
    double s = 0.0;
    to_real(p, g_re[0]);
    to_real(qv, g_re[1]);
    CK(spm_inpfp(ctx(), g_re[0], g_re[1], &s));
    return s;
}
double nrmp(fftw_complex *p) { return std::sqrt(inpfp(p, p)); }
double rayleighq(fftw_complex *x, fftw_complex *v)
{
    double s = 0.0;
    to_real(x, g_re[0]);
    to_real(v, g_re[1]);
    CK(spm_rayleighq(ctx(), g_re[0], g_re[1], &s));
    return s;
}

// The rest of BasicOperators.h:6-50 (everything that does not pass Eigen types).
// Fields are device arrays: fftw_complex* of n complex values, double* of n reals.
namespace {

__global__ void k_add_const(double *z, long n, double a, int both)
{ // FuncRealAddAConst (both = 0) / FuncCplxAddAConst (both = 1), BasicOperators.cpp:28-43
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        z[2 * i] += a;
        if (both) z[2 * i + 1] += a;
    }
}

__global__ void k_div(double *x, long n, double d)
{ // norm_fft's x / CplxDofs (BasicOperators.cpp:103-109)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = x[i] / d;
}

__global__ void k_mul(double *x, long n, double a)
{ // Dot_constant (BasicOperators.cpp:197-203)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = a * x[i];
}

constexpr int MAX_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_max(const double *x, long n, double *part)
{ // Max (BasicOperators.cpp:110-119): s starts at 0 and takes any larger x
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (s < x[i]) s = x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w /= 2) {
        if ((int)threadIdx.x < w && red[threadIdx.x] < red[threadIdx.x + w]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

void launched(const char *what) { hck(hipGetLastError(), what); }
unsigned blocks(long n) { return (unsigned)std::max<long>(1, std::min<long>(2048, (n + 255) / 256)); }

std::vector<cd_host> host_copy(const fftw_complex *x, long n)
{
    std::vector<cd_host> h(n);
    hck(hipMemcpy(h.data(), x, sizeof(fftw_complex) * n, hipMemcpyDeviceToHost), "Print copy");
    return h;
}

} // namespace

void FuncCplxAddAConst(fftw_complex *rslt, int n, const double a)
{
    hipLaunchKernelGGL(k_add_const, dim3(blocks(n)), dim3(256), 0, nullptr, (double *)rslt, (long)n, a, 1);
    launched("FuncCplxAddAConst");
}
void FuncRealAddAConst(fftw_complex *rslt, int n, const double a)
{
    hipLaunchKernelGGL(k_add_const, dim3(blocks(n)), dim3(256), 0, nullptr, (double *)rslt, (long)n, a, 0);
    launched("FuncRealAddAConst");
}
void norm_fft(fftw_complex *x, int n)
{
    hipLaunchKernelGGL(k_div, dim3(blocks(2L * n)), dim3(256), 0, nullptr, (double *)x, 2L * n, 1.0 * CplxDofs);
    launched("norm_fft");
}
void Dot_constant(double *rslt, int n, double a1)
{
    hipLaunchKernelGGL(k_mul, dim3(blocks(n)), dim3(256), 0, nullptr, rslt, (long)n, a1);
    launched("Dot_constant");
}
double Max(double *x, int n)
{
    double *part = nullptr;
    hck(hipMalloc(&part, sizeof(double) * MAX_BLOCKS), "hipMalloc Max");
    const unsigned nb = std::min<unsigned>(MAX_BLOCKS, blocks(n));
    hipLaunchKernelGGL(k_max, dim3(nb), dim3(256), 0, nullptr, x, (long)n, part);
    launched("Max");
    std::vector<double> h(nb);
    hck(hipMemcpy(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost), "Max read");
    (void)hipFree(part);
    double s = 0.0;
    for (double v : h)
        if (s < v) s = v;
    return s;
}
void setCplxZero(fftw_complex rslt) { hck(hipMemset(rslt, 0, sizeof(fftw_complex)), "setCplxZero"); }
void setCplxZero_V(fftw_complex **W, int k, int cplxDofs)
{
    for (int i = 0; i < k; i++) setCplxZero_v(W[i], cplxDofs);
}
// BasicOperators.cpp:229-241: v -= <v, V_i> V_i for i = 0..k-1 in turn, imaginary part 0
void Orthoi_v2V(fftw_complex *v, fftw_complex **V, int k)
{
    to_real(v, g_re[0]);
    for (int i = 0; i < k; i++) {
        double t = 0.0;
        to_real(V[i], g_re[1]);
        CK(spm_inpfp(ctx(), g_re[0], g_re[1], &t));
        CK(spm_lin2(ctx(), g_re[0], 1.0, g_re[0], -t, g_re[1]));
    }
    to_cplx(g_re[0], v);
}
// BasicOperators.cpp:243-255: V_k against V_0..V_{k-1}
void Orthoi_Vi(fftw_complex **V, int k) { Orthoi_v2V(V[k], V, k); }
// BasicOperators.cpp:344-358: tmpp = sum_i <v, V_i> V_i
void PV_Cplx(fftw_complex *v, fftw_complex **V, fftw_complex *tmpp, int k)
{
    to_real(v, g_re[0]);
    CK(spm_lin2(ctx(), g_re[2], 0.0, g_re[0], 0.0, g_re[0])); // tmpp = 0 (setCplxZero_v)
    for (int i = 0; i < k; i++) {
        double t = 0.0;
        to_real(V[i], g_re[1]);
        CK(spm_inpfp(ctx(), g_re[0], g_re[1], &t));
        CK(spm_lin2(ctx(), g_re[2], 1.0, g_re[2], t, g_re[1]));
    }
    to_cplx(g_re[2], tmpp);
}
// BasicOperators.cpp:361-371: tmpp = <v, V> V
void PV_Cplx(fftw_complex *v, fftw_complex *V, fftw_complex *tmpp, int k)
{
    (void)k;
    PV_Cplx(v, &V, tmpp, 1);
}
// BasicOperators.cpp:145-172 on device arrays of CplxDofs_local reals
double inpfp(double *p, double *qv)
{
    double s = 0.0;
    CK(spm_inpfp(ctx(), p, qv, &s));
    return s;
}
double nrmp(double *p) { return std::sqrt(inpfp(p, p)); }
// BasicOperators.cpp:257-281: the real parts of V, W, Vp as rows 0..3k-1 of UU
void get_Merge(double **UU, fftw_complex **V, fftw_complex **W, fftw_complex **Vp, int k)
{
    for (int i = 0; i < k; i++) to_real(V[i], UU[i]);
    for (int i = 0; i < k; i++) to_real(W[i], UU[k + i]);
    for (int i = 0; i < k; i++) to_real(Vp[i], UU[2 * k + i]);
}
// BasicOperators.cpp:283-294 (host flags)
int get_nozero(int *emp, int k)
{
    int iter = 0;
    for (int i = 0; i < 3 * k; i++)
        if (emp[i] == 1) iter = iter + 1;
    return iter;
}
// Print / Printf_begin_end (BasicOperators.cpp:45-92, 337-341): debug output of device fields
void Print(fftw_complex *x, int n)
{
    const std::vector<cd_host> h = host_copy(x, n);
    for (int i = 0; i < n; i++) {
        std::printf("%.3e\t", h[i].re);
        if ((i + 1) % N == 0) std::printf("\n");
    }
}
void Print(double *x, int n)
{
    std::vector<double> h(n);
    hck(hipMemcpy(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost), "Print copy");
    for (int i = 0; i < n; i++) {
        std::printf("%.3e\t", h[i]);
        if ((i + 1) % N == 0) std::printf("\n");
    }
}
void Print(fftw_complex *x)
{
    const std::vector<cd_host> h = host_copy(x, (long)N * N);
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) std::printf("(%.3e,%.3e)\t", h[i * N + j].re, h[i * N + j].im);
        std::printf("\n");
    }
}
void Print(double *x)
{
    std::vector<double> h((size_t)N * N);
    hck(hipMemcpy(h.data(), x, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "Print copy");
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) std::printf("%.3e\t", h[i * N + j]);
        std::printf("\n");
    }
}
void Printf_begin_end(fftw_complex *x)
{
    const std::vector<cd_host> h = host_copy(x, CplxDofs_local);
    std::printf("\n begin:  %.3e \t", h[0].re);
    std::printf("end:  %.3e \n", h[CplxDofs_local - 1].re);
}

// ----------------------------------------------------------------- Memfree (Memfree.cpp:5-73)
void Memfree_global_kspace()
{
    std::free(K);
    std::free(kindex);
    std::free(g_kflat);
    K = nullptr;
    kindex = nullptr;
    g_kflat = nullptr;
}

void Memfree_global_data()
{
    std::free(u_R);
    std::free(u_F);
    std::free(NCpt);
    for (int i = 0; i < dimPhy; i++) std::free(ProjMatrix[i]);
    std::free(ProjMatrix);
    u_R = u_F = nullptr;
    ProjMatrix = nullptr;
}

void Memfree_local_kspace()
{
    std::free(PK_local);
    std::free(kindex_local);
    std::free(g_kflat_local);
    std::free(g_pkflat);
    std::free(k_square_local);
    std::free(epmckt2_local);
    std::free(ikt2_local);
}

void Memfree_local_data()
{
    fftw_complex *fs[] = {u_R_local, u_F_local, u2_R_local, u3_R_local, u4_R_local, nln_R_local, nln_F_local,
                          Laplas_R_local, Laplas_F_local, fftw_Ctmp_local, fftw_Rtmp_local, tv, data};
    for (fftw_complex *p : fs)
        if (p) (void)hipFree(p);
    for (int i = 0; i < 2; i++) (void)hipFree(Phi0[i]);
    std::free(Phi0);
    free_scratch();
    if (g_ctx) spm_destroy(g_ctx);
    g_ctx = nullptr;
}

void Memfree_Option(Option *option)
{
    for (int i = 0; i < option->k; i++) (void)hipFree(option->V[i]);
    std::free(option->V);
    std::free(option);
}

// ----------------------------------------------------------------- File_Op (File_Op.cpp:8-89)
void Save_Option_V(Option *option, char *filename)
{
    std::printf("\n------------Save---V----------------------------\n");
    std::vector<fftw_complex> v(R());
    FILE *fp = open_or_die(filename, "w");
    for (int i = 0; i < option->k; i++) {
        hck(hipMemcpy(v.data(), option->V[i], sizeof(fftw_complex) * R(), hipMemcpyDeviceToHost), "Save_Option_V");
        for (size_t j = 0; j < R(); j++) std::fprintf(fp, "%.16e \n", v[j][0]);
    }
    std::fclose(fp);
}

void Save_phi(fftw_complex *phi, char *filename)
{
    std::printf("\n--------------save-----------------\n");
    FILE *fp = open_or_die(filename, "w");
    for (int i = 0; i < CplxDofs; i++) std::fprintf(fp, "%.16e \n", phi[i][0]);
    std::fclose(fp);
}

void load_phi(fftw_complex *x, char *filename)
{
    std::printf("\n -----------loading------------------");
    FILE *fp = open_or_die(filename, "r");
    for (int i = 0; i < CplxDofs; i++) {
        if (std::fscanf(fp, "%lf", &(x[i][0])) != 1) die("load_phi: short file", SPM_EINVAL);
        x[i][1] = 0.0;
    }
    std::fclose(fp);
}

void load_Option_V(Option *option, char *filename)
{
    std::printf("\n-----------------load-V----------------------");
    std::vector<fftw_complex> v(R());
    FILE *fp = open_or_die(filename, "r");
    for (int i = 0; i < option->k; i++) {
        for (size_t j = 0; j < R(); j++) {
            if (std::fscanf(fp, "%lf", &(v[j][0])) != 1) die("load_Option_V: short file", SPM_EINVAL);
            v[j][1] = 0.0;
        }
        hck(hipMemcpy(option->V[i], v.data(), sizeof(fftw_complex) * R(), hipMemcpyHostToDevice), "load_Option_V");
    }
    std::fclose(fp);
}

double spm_compat_wtime()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

spm_ctx *spm_compat_context() { return g_ctx; }
