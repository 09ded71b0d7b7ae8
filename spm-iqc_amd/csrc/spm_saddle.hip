// spm_saddle.hip -- the Spring Pair Method for an index-1 saddle
// (get_Tv, Natural_Forces, MinMax_Forces, Spring_Pairing; SPM.cpp:15-263)
// on two device-resident replicas.  Scalars that the reference gathers with
// MPI_Allreduce (the tangent norm, <f, v>) stay on the device; the host reads
// only the residual, once per iteration.
#include <cmath>
#include <cstring>

#include "block_ops.h"

using namespace spm;

namespace {

constexpr int DOTB = 512;

// out = sum(part) / R   (inpfp's weight, BasicOperators.cpp:132-143; nrmp^2 of get_Tv)
__global__ void __launch_bounds__(256) k_dot2_fin(const double *__restrict__ part, int nb, double R, double *out)
{
    __shared__ double red[16];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
    double v[1] = {s};
    block_reduce<1, 0u>(v, red);
    if (threadIdx.x == 0) *out = v[0] / R;
}

// Sums for <f, v> without forming f = ngrad_cam(x) (LPsrc.cpp:147-169):
// f = -(G*x) + nl(x) - mean(nl), nl = Tau x + Gamma x^2 - x^3, and
// <G*x, v> = <x, G*v> for the even kernel G (Parseval), so with vt = G*v
// <f, v> = -mean(x vt) + mean(nl v) - mean(nl) mean(v).
template <typename T>
__global__ void __launch_bounds__(256) k_fv_sums(const T *__restrict__ x, const T *__restrict__ v,
                                                const T *__restrict__ vt, long n, double Tau, double Gamma,
                                                double *__restrict__ part)
{
    __shared__ double red[4 * 16];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double u = (double)x[i], vi = (double)v[i];
        const double nl = Tau * u + Gamma * u * u - u * u * u;
        s[0] += u * (double)vt[i];
        s[1] += nl * vi;
        s[2] += nl;
        s[3] += vi;
    }
    block_reduce<4, 0u>(s, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[(long)q * gridDim.x + blockIdx.x] = s[q];
}

__global__ void __launch_bounds__(256) k_fv_fin(const double *__restrict__ part, int nb, double R, double *out)
{
    __shared__ double red[4 * 16];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += part[(long)q * nb + b];
    block_reduce<4, 0u>(s, red);
    if (threadIdx.x == 0) *out = (-s[0] + s[1]) / R - (s[2] / R) * (s[3] / R);
}

// get_Tv (SPM.cpp:15-32), in two kernels: v = Phi1 - Phi0 (with its norm), v /= nrmp(v)
template <typename T>
__global__ void __launch_bounds__(256) k_tv_diff(const T *__restrict__ p0, const T *__restrict__ p1, T *__restrict__ v,
                                                long n, double *__restrict__ part)
{
    __shared__ double red[16];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const T d = p1[i] - p0[i];
        v[i] = d;
        s += (double)d * (double)d;
    }
    double q[1] = {s};
    block_reduce<1, 0u>(q, red);
    if (threadIdx.x == 0) part[blockIdx.x] = q[0];
}

template <typename T> __global__ void k_div_by_sqrt(T *v, long n, const double *d2)
{
    const double d = sqrt(*d2);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[i] = (T)((double)v[i] / d);
}

// The spring half-step of SPM.cpp:203-216 (and 233-246), replicas updated in
// the reference's order: Phi0 += dt ks (Phi1 - Phi0); Phi1 += dt ks (Phi0' - Phi1).
template <typename T> __global__ void k_spring(T *p0, T *p1, long n, double dt, double ks)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double a = (double)p0[i], b = (double)p1[i];
        const T a1 = (T)(1 * a + dt * (ks * b + -ks * a));
        const T b1 = (T)(1 * b + dt * (ks * (double)a1 + -ks * b));
        p0[i] = a1;
        p1[i] = b1;
    }
}

__global__ void k_copy_scalar(const double *src, double *dst) { *dst = *src; }

template <typename T> int launch_tv(spm_ctx *h, const void *p0, const void *p1, void *v)
{
    Ctx &c = h->c;
    double *part = c.slot(6);
    hipLaunchKernelGGL(k_tv_diff<T>, dim3(DOTB), dim3(256), 0, c.stream, (const T *)p0, (const T *)p1, (T *)v, c.R,
                       part);
    hipLaunchKernelGGL(k_dot2_fin, dim3(1), dim3(256), 0, c.stream, part, DOTB, (double)c.R, c.scal + SC_TV2);
    hipLaunchKernelGGL(k_div_by_sqrt<T>, dim3(2048), dim3(256), 0, c.stream, (T *)v, c.R, c.scal + SC_TV2);
    return spm_launch_status("get_Tv");
}

int get_tv(spm_ctx *h, const void *p0, const void *p1, void *v)
{
    return h->c.prec == SPM_FP64 ? launch_tv<double>(h, p0, p1, v) : launch_tv<float>(h, p0, p1, v);
}

// Natural_Forces (coef 1) / MinMax_Forces (coef 2) on x in place, given the
// tangent v and vt = G*v (one G convolution per tangent serves every force
// with that tangent).
int forces(spm_ctx *h, void *x, const void *v, const void *vt, double coef)
{
    Ctx &c = h->c;
    double *part = c.slot(6);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_fv_sums<double>, dim3(DOTB), dim3(256), 0, c.stream, (const double *)x,
                           (const double *)v, (const double *)vt, c.R, c.Tau, c.Gamma, part);
    else
        hipLaunchKernelGGL(k_fv_sums<float>, dim3(DOTB), dim3(256), 0, c.stream, (const float *)x,
                           (const float *)v, (const float *)vt, c.R, c.Tau, c.Gamma, part);
    hipLaunchKernelGGL(k_fv_fin, dim3(1), dim3(256), 0, c.stream, part, DOTB, (double)c.R, c.scal + SC_FV);
    c.g->force(c, x, v, c.scal + SC_FV, coef);
    return spm_launch_status("forces");
}

int spring(spm_ctx *h, void *p0, void *p1, double dt, double ks)
{
    Ctx &c = h->c;
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_spring<double>, dim3(2048), dim3(256), 0, c.stream, (double *)p0, (double *)p1, c.R,
                           dt, ks);
    else
        hipLaunchKernelGGL(k_spring<float>, dim3(2048), dim3(256), 0, c.stream, (float *)p0, (float *)p1, c.R, dt,
                           ks);
    return spm_launch_status("spring");
}

int ensure_fields(spm_ctx *h, int n)
{
    const size_t bytes = (size_t)n * h->c.R * h->c.esz;
    if (h->ws_bytes >= bytes) return SPM_OK;
    if (h->ws) {
        HIPCHK(hipStreamSynchronize(h->c.stream));
        HIPCHK(hipFree(h->ws));
        h->dev_bytes -= h->ws_bytes;
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    RCHK(spm_dev_alloc(h, &h->ws, bytes));
    h->ws_bytes = bytes;
    return SPM_OK;
}

} // namespace

extern "C" {

int spm_get_tv(spm_ctx *h, const void *phi0, const void *phi1, void *v)
{
    CHECK_CTX(h);
    if (!phi0 || !phi1 || !v) return spm_fail(SPM_EINVAL, "spm_get_tv: null field");
    return get_tv(h, phi0, phi1, v);
}

int spm_natural_force(spm_ctx *h, void *x, const void *v)
{
    CHECK_CTX(h);
    if (!x || !v || x == v) return spm_fail(SPM_EINVAL, "spm_natural_force: bad fields");
    RCHK(ensure_fields(h, 1));
    c.g->gconv(c, v, h->ws);
    return forces(h, x, v, h->ws, 1.0);
}

int spm_minmax_force(spm_ctx *h, void *x, const void *v)
{
    CHECK_CTX(h);
    if (!x || !v || x == v) return spm_fail(SPM_EINVAL, "spm_minmax_force: bad fields");
    RCHK(ensure_fields(h, 1));
    c.g->gconv(c, v, h->ws);
    return forces(h, x, v, h->ws, 2.0);
}

int spm_spring_pairing(spm_ctx *h, void *phi0, void *phi1, int maxiter, double tol, int *iters, double *res_out,
                       double *res_hist)
{
    CHECK_CTX(h);
    if (!phi0 || !phi1 || phi0 == phi1) return spm_fail(SPM_EINVAL, "spm_spring_pairing: bad replicas");
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_spring_pairing: maxiter >= 0");
    RCHK(ensure_fields(h, 3));
    void *tv = h->ws;
    void *f = (char *)h->ws + (size_t)c.R * c.esz;
    void *vt = (char *)h->ws + 2 * (size_t)c.R * c.esz; // G * tv
    void *Phi[2] = {phi0, phi1};
    const double dt = 0.001, ks = 1; // SPM.cpp:177-178
    double res = 1e-6;               // SPM.cpp:180
    int iter = 0;
    while (res > tol && iter < maxiter) {
        iter = iter + 1;
        RCHK(get_tv(h, Phi[0], Phi[1], tv));
        c.g->gconv(c, tv, vt);
        for (int i = 0; i < 2; ++i) RCHK(forces(h, Phi[i], tv, vt, 1.0));
        RCHK(spring(h, Phi[0], Phi[1], dt, ks));
        RCHK(get_tv(h, Phi[0], Phi[1], tv));
        c.g->gconv(c, tv, vt);
        for (int i = 0; i < 2; ++i) {
            RCHK(forces(h, Phi[i], tv, vt, 2.0));
            c.g->ngrad(c, Phi[i], f);
            hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, c.stream, c.scal + SC_GN2, c.scal + SC_RES0 + i);
        }
        RCHK(spring(h, Phi[0], Phi[1], dt, ks));
        double g2[2];
        RCHK(spm_read_scalars(h, c.scal + SC_RES0, 2, g2));
        res = 0.0;
        for (int i = 0; i < 2; ++i) {
            const double resi = std::sqrt(g2[i]);
            if (res < resi) res = resi;
        }
        if (res_hist) res_hist[iter - 1] = res;
    }
    RCHK(get_tv(h, Phi[0], Phi[1], tv));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (iters) *iters = iter;
    if (res_out) *res_out = res;
    return spm_launch_status("spm_spring_pairing");
}

} // extern "C"
