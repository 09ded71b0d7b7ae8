// spm_capi.hip -- the C ABI (include/spm_iqc.h): context set-up, tables,
// generic vector kernels and reductions, and the host side of each entry
// point of the reference's path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace spm;

static thread_local std::string g_err;

int spm_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

// ======================================================================
// generic kernels
// ======================================================================

// out[q] = scale * sum_b part[q*nblk + b] (or max over b if bit q of maxmask)
__global__ void __launch_bounds__(256) k_reduce(const double *__restrict__ part, long nblk, unsigned maxmask,
                                               double scale, double *__restrict__ out)
{
    __shared__ double red[16];
    const int q = blockIdx.x;
    const bool mx = (maxmask >> q) & 1u;
    double s = mx ? 0.0 : 0.0;
    for (long b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double x = part[(long)q * nblk + b];
        s = mx ? fmax(s, x) : s + x;
    }
    double v[1] = {s};
    if (mx) block_reduce<1, 1u>(v, red);
    else block_reduce<1, 0u>(v, red);
    if (threadIdx.x == 0) out[q] = mx ? v[0] : v[0] * scale;
}

__global__ void k_ctl_init(Ctl *ctl, int max_iter, double res_tol, double gn_tol)
{
    ctl->iter = 0;
    ctl->stop = max_iter <= 0 ? 1 : 0;
    ctl->max_iter = max_iter;
    ctl->flags = 0;
    ctl->res_tol = res_tol;
    ctl->gn_tol = gn_tol;
}

static constexpr int DOT_BLOCKS = 1024;

template <typename T>
__global__ void __launch_bounds__(256) k_dot(const T *__restrict__ a, const T *__restrict__ b, long n,
                                            double *__restrict__ part)
{
    __shared__ double red[16];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    double v[1] = {s};
    block_reduce<1, 0u>(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// get_Ene's two sums as the reference forms them (LPsrc.cpp:11-35): the density
// -Tau/2 phi^2 - Gamma/3 phi^3 + phi^4/4 of the real field, and G |u_F|^2 over
// the full complex spectrum u_F (N^3 entries, x-major), with G looked up by |n|^2.
template <typename T>
__global__ void __launch_bounds__(256) k_energy_uf(const T *__restrict__ phi, const C2<T> *__restrict__ uF,
                                                  const double *__restrict__ G, int N, long n, double Tau,
                                                  double Gamma, double *__restrict__ part)
{
    __shared__ double red[2 * 16];
    double s[2] = {0.0, 0.0};
    for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) {
        const double u = (double)phi[l];
        const double u2 = u * u, u3 = u2 * u, u4 = u3 * u;
        s[0] += -0.5 * Tau * u2 - Gamma * u3 / 3 + 0.25 * u4;
        const int k = (int)(l % N), j = (int)((l / N) % N), i = (int)(l / ((long)N * N));
        const int a = kfold(i, N), b = kfold(j, N), cz = kfold(k, N);
        const double re = (double)uF[l].x, im = (double)uF[l].y;
        s[1] += (re * re + im * im) * G[a * a + b * b + cz * cz];
    }
    block_reduce<2, 0u>(s, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s[0];
        part[gridDim.x + blockIdx.x] = s[1];
    }
}

template <typename T>
__global__ void k_lin2(T *r, double a1, const T *F1, double a2, const T *F2, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r[i] = (T)(a1 * (double)F1[i] + a2 * (double)F2[i]);
}

template <typename T> __global__ void k_cplx_to_real(const C2<T> *__restrict__ z, T *__restrict__ x, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = z[i].x;
}

template <typename T> __global__ void k_real_to_cplx(const T *__restrict__ x, C2<T> *__restrict__ z, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        z[i] = C2<T>{x[i], (T)0};
}

namespace spm {

void launch_reduce(Ctx &c, const double *part, long nblk, int nval, unsigned maxmask, double scale, double *out)
{
    hipLaunchKernelGGL(k_reduce, dim3(nval), dim3(256), 0, c.stream, part, nblk, maxmask, scale, out);
}

#define SPM_FOR_EACH_N(X) X(4) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#define SPM_DECL_N(n) GridIface *make_grid_n##n(int, int); int grid_shapes_n##n(const int (**)[3]);
SPM_FOR_EACH_N(SPM_DECL_N)
#undef SPM_DECL_N

int grid_shapes(int N, const int (**s)[3])
{
    switch (N) {
#define SPM_CASE_N(n) case n: return grid_shapes_n##n(s);
        SPM_FOR_EACH_N(SPM_CASE_N)
#undef SPM_CASE_N
    default: return 0;
    }
}

GridIface *make_grid(int N, int prec, int shape)
{
    switch (N) {
#define SPM_CASE_N(n) case n: return make_grid_n##n(prec, shape);
        SPM_FOR_EACH_N(SPM_CASE_N)
#undef SPM_CASE_N
    default: return nullptr;
    }
}
#undef SPM_FOR_EACH_N

} // namespace spm

// ======================================================================
// host helpers
// ======================================================================

// Initialization.cpp:150-166
extern "C" double spm_get_GK(double k2)
{
    double ph = 0.568809;
    double d[5] = {1.969182919080340, -14.949858723012090, 37.445683882581790, -38.839451762550148,
                   13.999261011773505};
    double d0 = d[0], d2 = d[1], d4 = d[2], d6 = d[3], d8 = d[4];
    double s2 = ph * ph;
    double k4 = k2 * k2, k6 = k4 * k2, k8 = k6 * k2;
    return exp(-0.5 * k2 / s2) * (d0 + d2 * k2 + d4 * k4 + d6 * k6 + d8 * k8) + 0.08;
}

void spm_host_twiddles(int N, std::vector<double> &twd, std::vector<float> &twf)
{
    twd.resize(2 * N);
    twf.resize(2 * N);
    for (int e = 0; e < N; ++e) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * e / N;
        twd[2 * e] = (double)cosl(a);
        twd[2 * e + 1] = (double)(-sinl(a));
        twf[2 * e] = (float)cosl(a);
        twf[2 * e + 1] = (float)(-sinl(a));
    }
}

double spm_host_k2(double kscale, double L, int n2)
{
    const double P = kscale / L;
    return (double)((long double)P * (long double)P * (long double)n2);
}

int spm_dev_alloc(spm_ctx *h, void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return spm_fail(SPM_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    h->dev_bytes += bytes;
    return SPM_OK;
}

static void free_all(spm_ctx *h)
{
    Ctx &c = h->c;
    void *ps[] = {c.tw, c.Gtab, c.Itab, c.Btab, c.GDtab, c.W, c.U, c.part, c.scal, c.ctl, c.hist, c.counter, c.prm};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (h->h_pin) (void)hipHostFree(h->h_pin);
    if (h->gf_exec) (void)hipGraphExecDestroy(h->gf_exec);
    if (h->ev_in) (void)hipEventDestroy(h->ev_in);
    if (h->ev_out) (void)hipEventDestroy(h->ev_out);
    for (void *p : {h->ws, (void *)h->bscr, h->specws})
        if (p) (void)hipFree(p);
    if (h->own) (void)hipStreamDestroy(h->own);
    delete c.g;
}

static bool supported_N(int N)
{
    switch (N) {
    case 4: case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
        return true;
    default: return false;
    }
}

int spm_check_params(const spm_params *p, const char *who)
{
    const std::string w(who);
    if (!supported_N(p->N)) return spm_fail(SPM_ENOTSUP, w + ": unsupported N " + std::to_string(p->N));
    if (p->precision != SPM_FP64 && p->precision != SPM_FP32)
        return spm_fail(SPM_EINVAL, w + ": precision must be SPM_FP64 or SPM_FP32");
    if (!(p->L > 0) || !(p->kscale > 0) || !std::isfinite(p->Tau) || !std::isfinite(p->Gamma) || !(p->dt > 0))
        return spm_fail(SPM_EINVAL, w + ": L, kscale, dt must be > 0 and Tau, Gamma finite");
    if (p->step_passes != 0 && p->step_passes != 2 && p->step_passes != 4)
        return spm_fail(SPM_EINVAL, w + ": step_passes must be 0, 2 or 4");
    if (p->step_passes == 2 && p->N != 64 && p->N != 128 && p->N != 256)
        return spm_fail(SPM_ENOTSUP, w + ": the two-pass gradient-flow step exists for N = 64, 128 and 256 only");
    const int sv = p->step_variant;
    if (sv != 0 && sv != -1 &&
        !((sv & ~7) == SPM_STEP_EXPLICIT && (sv & SPM_STEP_U_MASK) != SPM_STEP_U_MASK))
        return spm_fail(SPM_EINVAL, w + ": step_variant must be 0, -1 or SPM_STEP_EXPLICIT | SPM_STEP_* flags");
    return SPM_OK;
}

// The built-in two-pass step variant, from spm_tune_step on MI355X (us per step,
// preloaded against gathered): FP64 at 256^3 preloads U (191 against 211: its
// pass 2 holds 2 waves per SIMD either way); FP32 at 256^3 and 128^3 preloads
// (111 against 114, 28.3 against 28.9); FP64 at 128^3 and both precisions at
// 64^3 gather U, 4 elements' loads at a time (FP64 36.1 against 37.6 at 128^3).
// The non-temporal hints follow the size rule (one FP64 256^3 field: W + U =
// 270 MB, over the last-level cache, 191.2-191.4 against 191.5-192.5 cached).
static int builtin_step_variant(int N, int prec, int nmem)
{
    (void)nmem;
    if (N == 256 || (prec == SPM_FP32 && N == 128)) return 0;
    return SPM_STEP_GATHER_U;
}

static int ensure_hist(spm_ctx *h, long n)
{
    Ctx &c = h->c;
    if (c.hist_cap >= n) return SPM_OK;
    if (c.hist) {
        HIPCHK(hipStreamSynchronize(c.stream));
        HIPCHK(hipFree(c.hist));
        h->dev_bytes -= sizeof(double) * 3 * c.hist_cap;
        c.hist = nullptr;
    }
    long cap = 1024;
    while (cap < n) cap *= 2;
    int rc = spm_dev_alloc(h, (void **)&c.hist, sizeof(double) * 3 * cap);
    if (rc) return rc;
    c.hist_cap = cap;
    return SPM_OK;
}

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

const char *spm_last_error(void) { return g_err.c_str(); }

int spm_create(const spm_params *p, spm_ctx **out)
{
    if (!p || !out) return spm_fail(SPM_EINVAL, "spm_create: null argument");
    return spm_ctx_create(p, 1, 1024, out);
}

} // extern "C"

int spm_ctx_create(const spm_params *p, int nmem, long hist_cap, spm_ctx **out)
{
    if (!p || !out) return spm_fail(SPM_EINVAL, "spm_create: null argument");
    *out = nullptr;
    if (nmem < 1) return spm_fail(SPM_EINVAL, "spm_create: batch capacity must be >= 1");
    RCHK(spm_check_params(p, "spm_create"));
    spm_ctx *h = new spm_ctx();
    Ctx &c = h->c;
    c.N = p->N;
    c.prec = p->precision;
    c.esz = c.prec == SPM_FP64 ? 8 : 4;
    c.NH = c.N / 2 + 1;
    const int align = c.prec == SPM_FP64 ? 8 : 16; // 128-byte rows
    c.NHP = (c.NH + align - 1) / align * align;
    c.R = (long)c.N * c.N * c.N;
    c.C = (long)c.N * c.N * c.NHP;
    c.L = p->L;
    c.kscale = p->kscale;
    c.Tau = p->Tau;
    c.Gamma = p->Gamma;
    c.dt = p->dt;
    int shape = p->shape;
    if (shape < 0) { // the tuned shape for this size, precision and batch (measured on a miss)
        const int rc = spm_resolve_shape(p, nmem, &shape);
        if (rc) {
            delete h;
            return rc;
        }
    }
    const int (*shapes)[3] = nullptr;
    if (shape >= grid_shapes(c.N, &shapes)) {
        delete h;
        return spm_fail(SPM_EINVAL, "spm_create: no FFT pass shape " + std::to_string(shape) + " for N " +
                                        std::to_string(c.N));
    }
    c.shape = shape;
    c.gf2 = p->step_passes == 2 || (p->step_passes == 0 && (c.N == 64 || c.N == 128 || c.N == 256));
    if (p->step_variant == 0 || !c.gf2) {
        c.sjv = builtin_step_variant(c.N, c.prec, nmem);
    } else if (p->step_variant < 0) { // the tuned variant for this size, precision and batch
        const int rc = spm_resolve_step(p, nmem, &c.sjv);
        if (rc) {
            delete h;
            return rc;
        }
    } else {
        c.sjv = p->step_variant & ~SPM_STEP_EXPLICIT;
    }
    c.nmem = nmem;
    c.g = make_grid(c.N, c.prec, shape);
    if (!c.g) {
        delete h;
        return spm_fail(SPM_ENOTSUP, "spm_create: no kernels for this N");
    }
    c.g->geometry(c.nb_rows, c.nb_j, c.nb_i, c.NHP);
    int rc = SPM_OK;
    auto bail = [&](int code) {
        if (h->own) (void)hipStreamSynchronize(h->own); // uploads in flight read host buffers of this frame
        free_all(h);
        delete h;
        return code;
    };
    if (hipError_t e = hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking); e != hipSuccess)
        return bail(spm_fail(SPM_EHIP, std::string("spm_create: hipStreamCreate: ") + hipGetErrorString(e)));
    c.stream = h->own;
    // Every upload and fill below is ordered on the context's own stream, which
    // is synchronised before returning: the legacy stream (hipMemcpy, hipMemset)
    // is not ordered with the non-blocking streams the context's work runs on,
    // so a fill queued there could land after the first kernels.  The host
    // sources stay alive until that synchronisation.
    // twiddles exp(-2 pi i e / N)
    std::vector<double> twd;
    std::vector<float> twf;
    spm_host_twiddles(c.N, twd, twf);
    if ((rc = spm_dev_alloc(h, &c.tw, 2 * c.N * c.esz))) return bail(rc);
    if (hipMemcpyAsync(c.tw, c.prec == SPM_FP64 ? (void *)twd.data() : (void *)twf.data(), 2 * c.N * c.esz,
                       hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return bail(spm_fail(SPM_EHIP, "spm_create: twiddle upload failed"));
    // G and preconditioner tables over |n|^2 (get_ksquare_local, get_epmckt2_local,
    // get_ikt2_local; Initialization.cpp:123-174)
    const int h2 = c.N / 2;
    c.ntab = 3 * h2 * h2 + 1;
    std::vector<double> Gt(c.ntab), It(c.ntab), Bt(c.ntab), GDd(2 * c.ntab);
    std::vector<float> GDf(2 * c.ntab);
    {
        for (int n2 = 0; n2 < c.ntab; ++n2) {
            Gt[n2] = spm_get_GK(spm_host_k2(c.kscale, c.L, n2));
            It[n2] = 1.0 / (Gt[n2] + std::fabs(c.Tau) + 0.02);
            Bt[n2] = 1.0 / (1 + 1.0 * Gt[n2]);
            GDd[2 * n2] = Gt[n2];
            GDd[2 * n2 + 1] = 1.0 / (1 + c.dt * Gt[n2]);
            GDf[2 * n2] = (float)GDd[2 * n2];
            GDf[2 * n2 + 1] = (float)GDd[2 * n2 + 1];
        }
        if ((rc = spm_dev_alloc(h, (void **)&c.Gtab, sizeof(double) * c.ntab))) return bail(rc);
        if ((rc = spm_dev_alloc(h, (void **)&c.Itab, sizeof(double) * c.ntab))) return bail(rc);
        if ((rc = spm_dev_alloc(h, &c.GDtab, 2 * c.esz * c.ntab))) return bail(rc);
        if ((rc = spm_dev_alloc(h, (void **)&c.Btab, sizeof(double) * c.ntab))) return bail(rc);
        if (hipMemcpyAsync(c.Gtab, Gt.data(), sizeof(double) * c.ntab, hipMemcpyHostToDevice, c.stream) !=
                hipSuccess ||
            hipMemcpyAsync(c.Itab, It.data(), sizeof(double) * c.ntab, hipMemcpyHostToDevice, c.stream) !=
                hipSuccess ||
            hipMemcpyAsync(c.Btab, Bt.data(), sizeof(double) * c.ntab, hipMemcpyHostToDevice, c.stream) !=
                hipSuccess ||
            hipMemcpyAsync(c.GDtab, c.prec == SPM_FP64 ? (const void *)GDd.data() : (const void *)GDf.data(),
                           2 * c.esz * c.ntab, hipMemcpyHostToDevice, c.stream) != hipSuccess)
            return bail(spm_fail(SPM_EHIP, "spm_create: table upload failed"));
    }
    const size_t spec = (size_t)c.C * 2 * c.esz;
    if ((rc = spm_dev_alloc(h, &c.W, spec * nmem))) return bail(rc);
    if ((rc = spm_dev_alloc(h, &c.U, spec * nmem))) return bail(rc);
    long mb = std::max(std::max(c.nb_rows, c.nb_j), std::max(c.nb_i, (long)DOT_BLOCKS));
    c.part_slot = std::max(8 * mb, c.g->two_pass_partials()) * nmem; // every member's share of a slot
    c.nslots = 8;
    if ((rc = spm_dev_alloc(h, (void **)&c.part, sizeof(double) * c.part_slot * c.nslots))) return bail(rc);
    if ((rc = spm_dev_alloc(h, (void **)&c.scal, sizeof(double) * SC_COUNT))) return bail(rc);
    if ((rc = spm_dev_alloc(h, (void **)&c.ctl, sizeof(Ctl) * nmem))) return bail(rc);
    std::vector<double> prm(2 * nmem);
    if (nmem > 1) { // every member starts with the context's (Tau, Gamma)
        if ((rc = spm_dev_alloc(h, (void **)&c.prm, sizeof(double) * 2 * nmem))) return bail(rc);
        for (int m = 0; m < nmem; ++m) {
            prm[2 * m] = c.Tau;
            prm[2 * m + 1] = c.Gamma;
        }
        if (hipMemcpyAsync(c.prm, prm.data(), sizeof(double) * 2 * nmem, hipMemcpyHostToDevice, c.stream) !=
            hipSuccess)
            return bail(spm_fail(SPM_EHIP, "spm_create: parameter upload failed"));
    }
    if ((rc = spm_dev_alloc(h, (void **)&c.counter, 64 * sizeof(unsigned)))) return bail(rc);
    if (hipMemsetAsync(c.counter, 0, 64 * sizeof(unsigned), c.stream) != hipSuccess)
        return bail(spm_fail(SPM_EHIP, "spm_create: memset"));
    if (hipHostMalloc((void **)&h->h_pin, 4096, hipHostMallocDefault) != hipSuccess)
        return bail(spm_fail(SPM_ENOMEM, "spm_create: pinned scratch"));
    if (nmem == 1) {
        if ((rc = ensure_hist(h, hist_cap))) return bail(rc);
    } else { // fixed: a captured step graph holds the address
        if ((rc = spm_dev_alloc(h, (void **)&c.hist, sizeof(double) * 3 * hist_cap * nmem))) return bail(rc);
        c.hist_cap = hist_cap;
    }
    if (hipMemsetAsync(c.scal, 0, sizeof(double) * SC_COUNT, c.stream) != hipSuccess)
        return bail(spm_fail(SPM_EHIP, "spm_create: memset"));
    for (int m = 0; m < nmem; ++m) // every member starts stopped
        hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, c.stream, c.ctl + m, 0, 0.0, 0.0);
    if ((rc = spm_launch_status("spm_create"))) return bail(rc);
    if (hipStreamSynchronize(c.stream) != hipSuccess) return bail(spm_fail(SPM_EHIP, "spm_create: sync"));
    *out = h;
    return SPM_OK;
}

void spm_ctl_init(spm_ctx *h, int m, int max_iter, double res_tol, double gn_tol)
{
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, h->c.stream, h->c.ctl + m, max_iter, res_tol, gn_tol);
}

void spm_ctx_free(spm_ctx *h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->c.stream);
    free_all(h);
    delete h;
}

extern "C" {

int spm_destroy(spm_ctx *h)
{
    spm_ctx_free(h);
    return SPM_OK;
}

int spm_set_stream(spm_ctx *h, void *s)
{
    if (!h) return spm_fail(SPM_EINVAL, "null context");
    h->c.stream = (hipStream_t)s; // NULL: the legacy default stream
    return SPM_OK;
}

int spm_get_layout(const spm_ctx *h, spm_layout *o)
{
    if (!h || !o) return spm_fail(SPM_EINVAL, "null argument");
    const Ctx &c = h->c;
    o->N = c.N;
    o->nh = c.NH;
    o->nhp = c.NHP;
    o->precision = c.prec;
    o->R = c.R;
    o->C = c.C;
    o->real_bytes = (size_t)c.R * c.esz;
    o->spec_bytes = (size_t)c.C * 2 * c.esz;
    o->ctx_bytes = h->dev_bytes;
    return SPM_OK;
}


} // extern "C"

int spm_launch_status(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return spm_fail(SPM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return SPM_OK;
}

extern "C" {

int spm_my_fft(spm_ctx *h, const void *x_R, void *x_F)
{
    CHECK_CTX(h);
    if (!x_R || !x_F) return spm_fail(SPM_EINVAL, "spm_my_fft: null field");
    c.g->c2c(c, x_R, x_F, -1);
    return spm_launch_status("spm_my_fft");
}

int spm_my_ifft(spm_ctx *h, const void *x_F, void *x_R)
{
    CHECK_CTX(h);
    if (!x_R || !x_F) return spm_fail(SPM_EINVAL, "spm_my_ifft: null field");
    c.g->c2c(c, x_F, x_R, +1);
    return spm_launch_status("spm_my_ifft");
}

int spm_r2c(spm_ctx *h, const void *x, void *S)
{
    CHECK_CTX(h);
    if (!x || !S) return spm_fail(SPM_EINVAL, "spm_r2c: null field");
    c.g->r2c(c, x, S);
    return spm_launch_status("spm_r2c");
}

int spm_c2r(spm_ctx *h, const void *S, void *x)
{
    CHECK_CTX(h);
    if (!x || !S) return spm_fail(SPM_EINVAL, "spm_c2r: null field");
    c.g->c2r(c, S, x);
    return spm_launch_status("spm_c2r");
}

} // extern "C"

// Capture and instantiate the GF_GRAPH_STEPS-step graph (once per history
// buffer) and upload it to the device, so that the first replay costs what
// later ones do.
static int gf_graph_ready(spm_ctx *h)
{
    Ctx &c = h->c;
    if (h->gf_exec && h->gf_exec_hist == c.hist) return SPM_OK;
    if (h->gf_exec) {
        HIPCHK(hipGraphExecDestroy(h->gf_exec));
        h->gf_exec = nullptr;
    }
    hipStream_t caller = c.stream;
    hipGraph_t g = nullptr;
    c.stream = h->own;
    hipError_t e = hipStreamBeginCapture(h->own, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        for (int s = 0; s < GF_GRAPH_STEPS; ++s) c.g->gf_step(c);
        e = hipStreamEndCapture(h->own, &g);
    }
    c.stream = caller;
    if (e == hipSuccess) e = hipGraphInstantiate(&h->gf_exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e == hipSuccess) e = hipGraphUpload(h->gf_exec, h->own);
    if (e != hipSuccess) {
        if (h->gf_exec) (void)hipGraphExecDestroy(h->gf_exec);
        h->gf_exec = nullptr;
        return spm_fail(SPM_EHIP, std::string("gradient-flow graph: ") + hipGetErrorString(e));
    }
    h->gf_exec_hist = c.hist;
    return SPM_OK;
}

int spm_gf_enqueue(spm_ctx *h, long n)
{
    Ctx &c = h->c;
    if (n < GF_GRAPH_STEPS) {
        for (long s = 0; s < n; ++s) c.g->gf_step(c);
        return spm_launch_status("gradient-flow steps");
    }
    hipStream_t caller = c.stream;
    RCHK(gf_graph_ready(h));
    if (!h->ev_in) HIPCHK(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
    if (!h->ev_out) HIPCHK(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
    if (caller != h->own) {
        HIPCHK(hipEventRecord(h->ev_in, caller));
        HIPCHK(hipStreamWaitEvent(h->own, h->ev_in, 0));
    }
    long s = 0;
    for (; s + GF_GRAPH_STEPS <= n; s += GF_GRAPH_STEPS) HIPCHK(hipGraphLaunch(h->gf_exec, h->own));
    c.stream = h->own;
    for (; s < n; ++s) c.g->gf_step(c);
    c.stream = caller;
    RCHK(spm_launch_status("gradient-flow steps"));
    if (caller != h->own) {
        HIPCHK(hipEventRecord(h->ev_out, h->own));
        HIPCHK(hipStreamWaitEvent(caller, h->ev_out, 0));
    }
    return SPM_OK;
}

int spm_read_scalars(spm_ctx *h, const double *dev, int n, double *out)
{
    Ctx &c = h->c;
    HIPCHK(hipMemcpyAsync(h->h_pin, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    std::memcpy(out, h->h_pin, sizeof(double) * n);
    return SPM_OK;
}

extern "C" {

int spm_energy(spm_ctx *h, const void *phi, double *E)
{
    CHECK_CTX(h);
    if (!phi || !E) return spm_fail(SPM_EINVAL, "spm_energy: null argument");
    c.g->energy(c, phi);
    int rc = spm_launch_status("spm_energy");
    if (rc) return rc;
    double v[2];
    if ((rc = spm_read_scalars(h, c.scal + SC_E1, 2, v))) return rc;
    *E = 0.5 * v[0] + v[1]; // E1 + E2 (LPsrc.cpp:35-36)
    return SPM_OK;
}

int spm_energy_uf(spm_ctx *h, const void *phi, const void *u_F, double *E)
{
    CHECK_CTX(h);
    if (!phi || !u_F || !E) return spm_fail(SPM_EINVAL, "spm_energy_uf: null argument");
    double *part = c.slot(7);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_energy_uf<double>, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, (const double *)phi,
                           (const C2<double> *)u_F, c.Gtab, c.N, c.R, c.Tau, c.Gamma, part);
    else
        hipLaunchKernelGGL(k_energy_uf<float>, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, (const float *)phi,
                           (const C2<float> *)u_F, c.Gtab, c.N, c.R, c.Tau, c.Gamma, part);
    launch_reduce(c, part, DOT_BLOCKS, 2, 0u, 1.0, c.scal + SC_TMP0);
    int rc = spm_launch_status("spm_energy_uf");
    if (rc) return rc;
    double v[2];
    if ((rc = spm_read_scalars(h, c.scal + SC_TMP0, 2, v))) return rc;
    *E = 0.5 * v[1] + v[0] / (double)c.R; // E1 + E2 (LPsrc.cpp:24, 35-36)
    return SPM_OK;
}

int spm_ngrad(spm_ctx *h, const void *phi, void *grad, double *nrm)
{
    CHECK_CTX(h);
    if (!phi || !grad) return spm_fail(SPM_EINVAL, "spm_ngrad: null field");
    if (phi == grad) return spm_fail(SPM_EINVAL, "spm_ngrad: phi and grad must differ");
    c.g->ngrad(c, phi, grad);
    int rc = spm_launch_status("spm_ngrad");
    if (rc || !nrm) return rc;
    double g2;
    if ((rc = spm_read_scalars(h, c.scal + SC_GN2, 1, &g2))) return rc;
    *nrm = std::sqrt(g2);
    return SPM_OK;
}

int spm_hv(spm_ctx *h, const void *phi, const void *v, void *hv)
{
    CHECK_CTX(h);
    if (!phi || !v || !hv) return spm_fail(SPM_EINVAL, "spm_hv: null field");
    if (hv == v || hv == phi) return spm_fail(SPM_EINVAL, "spm_hv: hv must not alias phi or v");
    c.g->hv(c, phi, v, hv);
    return spm_launch_status("spm_hv");
}

int spm_precond(spm_ctx *h, void *w)
{
    CHECK_CTX(h);
    if (!w) return spm_fail(SPM_EINVAL, "spm_precond: null field");
    c.g->precond(c, w);
    return spm_launch_status("spm_precond");
}

int spm_gradflow(spm_ctx *h, void *phi, int max_iter, double res_tol, double gn_tol, int check_every,
                 double *hist, int *iters)
{
    CHECK_CTX(h);
    if (!phi) return spm_fail(SPM_EINVAL, "spm_gradflow: null field");
    if (max_iter < 0 || check_every < 1) return spm_fail(SPM_EINVAL, "spm_gradflow: max_iter >= 0, check_every >= 1");
    int rc = ensure_hist(h, max_iter > 0 ? max_iter : 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, c.stream, c.ctl, max_iter, res_tol, gn_tol);
    if (max_iter > 0) {
        c.g->gf_init(c, phi);
        long queued = 0;
        while (queued < max_iter) {
            const long n = std::min<long>(check_every, max_iter - queued);
            if ((rc = spm_gf_enqueue(h, n))) return rc;
            queued += n;
            HIPCHK(hipMemcpyAsync(h->h_pin, c.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (((Ctl *)h->h_pin)->stop) break;
        }
        c.g->gf_read_member(c, 0, phi);
    }
    HIPCHK(hipMemcpyAsync(h->h_pin, c.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    const int it = ((Ctl *)h->h_pin)->iter;
    if (iters) *iters = it;
    if (hist && it > 0) {
        HIPCHK(hipMemcpyAsync(hist, c.hist, sizeof(double) * 3 * it, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    return spm_launch_status("spm_gradflow");
}

int spm_gf_begin(spm_ctx *h, const void *phi, int max_iter, double res_tol, double gn_tol)
{
    CHECK_CTX(h);
    if (!phi) return spm_fail(SPM_EINVAL, "spm_gf_begin: null field");
    if (max_iter < 0) return spm_fail(SPM_EINVAL, "spm_gf_begin: max_iter >= 0");
    int rc = ensure_hist(h, max_iter > 0 ? max_iter : 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, c.stream, c.ctl, max_iter, res_tol, gn_tol);
    if (max_iter > 0) c.g->gf_init(c, phi);
    RCHK(spm_launch_status("spm_gf_begin"));
    // the step graph is captured here, not at the first spm_gf_steps call, so
    // that timing around spm_gf_steps sees replays only
    if (max_iter >= GF_GRAPH_STEPS) RCHK(gf_graph_ready(h));
    return SPM_OK;
}

int spm_gf_steps(spm_ctx *h, int nsteps)
{
    CHECK_CTX(h);
    if (nsteps < 0) return spm_fail(SPM_EINVAL, "spm_gf_steps: nsteps >= 0");
    return spm_gf_enqueue(h, nsteps);
}

int spm_gf_end(spm_ctx *h, void *phi, double *hist, int *iters)
{
    CHECK_CTX(h);
    if (phi) c.g->gf_read_member(c, 0, phi);
    HIPCHK(hipMemcpyAsync(h->h_pin, c.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    const int it = ((Ctl *)h->h_pin)->iter;
    if (iters) *iters = it;
    if (hist && it > 0) {
        HIPCHK(hipMemcpyAsync(hist, c.hist, sizeof(double) * 3 * it, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    return spm_launch_status("spm_gf_end");
}

int spm_inpfp(spm_ctx *h, const void *p, const void *q, double *out)
{
    CHECK_CTX(h);
    if (!p || !q || !out) return spm_fail(SPM_EINVAL, "spm_inpfp: null argument");
    double *part = c.slot(7);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_dot<double>, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, (const double *)p,
                           (const double *)q, c.R, part);
    else
        hipLaunchKernelGGL(k_dot<float>, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, (const float *)p,
                           (const float *)q, c.R, part);
    launch_reduce(c, part, DOT_BLOCKS, 1, 0u, 1.0 / (double)c.R, c.scal + SC_TMP0);
    int rc = spm_launch_status("spm_inpfp");
    if (rc) return rc;
    return spm_read_scalars(h, c.scal + SC_TMP0, 1, out);
}

int spm_nrmp(spm_ctx *h, const void *p, double *out)
{
    int rc = spm_inpfp(h, p, p, out);
    if (rc == SPM_OK) *out = std::sqrt(*out);
    return rc;
}

int spm_lin2(spm_ctx *h, void *r, double a1, const void *F1, double a2, const void *F2)
{
    CHECK_CTX(h);
    if (!r || !F1 || !F2) return spm_fail(SPM_EINVAL, "spm_lin2: null field");
    const int nb = 2048;
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_lin2<double>, dim3(nb), dim3(256), 0, c.stream, (double *)r, a1, (const double *)F1,
                           a2, (const double *)F2, c.R);
    else
        hipLaunchKernelGGL(k_lin2<float>, dim3(nb), dim3(256), 0, c.stream, (float *)r, a1, (const float *)F1, a2,
                           (const float *)F2, c.R);
    return spm_launch_status("spm_lin2");
}

int spm_cplx_to_real(spm_ctx *h, const void *z, void *x)
{
    CHECK_CTX(h);
    if (!z || !x) return spm_fail(SPM_EINVAL, "null field");
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_cplx_to_real<double>, dim3(2048), dim3(256), 0, c.stream, (const C2<double> *)z,
                           (double *)x, c.R);
    else
        hipLaunchKernelGGL(k_cplx_to_real<float>, dim3(2048), dim3(256), 0, c.stream, (const C2<float> *)z,
                           (float *)x, c.R);
    return spm_launch_status("spm_cplx_to_real");
}

int spm_real_to_cplx(spm_ctx *h, const void *x, void *z)
{
    CHECK_CTX(h);
    if (!z || !x) return spm_fail(SPM_EINVAL, "null field");
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_real_to_cplx<double>, dim3(2048), dim3(256), 0, c.stream, (const double *)x,
                           (C2<double> *)z, c.R);
    else
        hipLaunchKernelGGL(k_real_to_cplx<float>, dim3(2048), dim3(256), 0, c.stream, (const float *)x,
                           (C2<float> *)z, c.R);
    return spm_launch_status("spm_real_to_cplx");
}

} // extern "C"
