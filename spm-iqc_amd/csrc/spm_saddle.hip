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
// <f, v> = -mean(x vt) + mean(nl v) - mean(nl) mean(v): the sums of x vt, nl v,
// nl and v.  (Spring_Pairing gets the same sums of both replicas from the
// output pass of its G convolution, OpFvOut.)
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

// From NX replicas' sums (per replica x vt, nl v, nl; then v; part[q nb + b]),
// out[r] = <f_r, v> / d2 (d2 = nrmp(v)^2 of an unnormalised tangent v, or 1
// when d2 is null): the force term <f, t> t along the unit tangent t = v / nrmp(v)
// is then out * v.
template <int NX>
__global__ void __launch_bounds__(256) k_fv_fin(const double *__restrict__ part, int nb, double R,
                                               const double *d2, double *out0, double *out1)
{
    constexpr int NS = 3 * NX + 1;
    __shared__ double red[NS * 16];
    double s[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NS; ++q) s[q] += part[(long)q * nb + b];
    block_reduce<NS, 0u>(s, red);
    if (threadIdx.x == 0) {
        const double w = d2 ? *d2 : 1.0, mv = s[3 * NX] / R;
        *out0 = ((-s[0] + s[1]) / R - (s[2] / R) * mv) / w;
        if (NX == 2) *out1 = ((-s[3] + s[4]) / R - (s[5] / R) * mv) / w;
    }
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

__global__ void k_copy_scalar(const double *src, double *dst) { *dst = *src; }

// get_Tv: v = (Phi1 - Phi0) / nrmp(Phi1 - Phi0)
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

// Spring_Pairing's tangent (after the spring half-steps when spring): d = Phi1 -
// Phi0 unnormalised, scal[SC_TV2] = nrmp(d)^2 and the force scalars
// scal[SC_FV], scal[SC_FV1] = <f_r, d> / nrmp(d)^2 of both replicas, from one
// G-convolution round trip of d (its passes carry the spring step and the sums).
int tangent(spm_ctx *h, void *p0, void *p1, void *d, bool spring, double dt, double ks)
{
    Ctx &c = h->c;
    double *pt = c.slot(5), *pf = c.slot(6);
    const long nb = c.g->spm_tangent(c, p0, p1, d, spring, dt, ks, pt, pf);
    if (7 * nb > c.part_slot) return spm_fail(SPM_EINVAL, "spm tangent: partial sums exceed a slot");
    launch_reduce(c, pt, nb, 1, 0u, 1.0 / (double)c.R, c.scal + SC_TV2);
    hipLaunchKernelGGL(k_fv_fin<2>, dim3(1), dim3(256), 0, c.stream, pf, (int)nb, (double)c.R, c.scal + SC_TV2,
                       c.scal + SC_FV, c.scal + SC_FV1);
    return spm_launch_status("spm tangent");
}

// <f, v> of x along the unit tangent v into scal[SC_FV], given vt = G*v.
int fv_scalar(spm_ctx *h, const void *x, const void *v, const void *vt)
{
    Ctx &c = h->c;
    double *part = c.slot(6);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_fv_sums<double>, dim3(DOTB), dim3(256), 0, c.stream, (const double *)x,
                           (const double *)v, (const double *)vt, c.R, c.Tau, c.Gamma, part);
    else
        hipLaunchKernelGGL(k_fv_sums<float>, dim3(DOTB), dim3(256), 0, c.stream, (const float *)x,
                           (const float *)v, (const float *)vt, c.R, c.Tau, c.Gamma, part);
    hipLaunchKernelGGL(k_fv_fin<1>, dim3(1), dim3(256), 0, c.stream, part, DOTB, (double)c.R, nullptr,
                       c.scal + SC_FV, nullptr);
    return spm_launch_status("forces");
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

// Natural_Forces (coef 1) / MinMax_Forces (coef 2) on x in place along the
// unit tangent v: one G convolution of v, the <f, v> sums, one force round trip.
int unit_force(spm_ctx *h, void *x, const void *v, double coef)
{
    Ctx &c = h->c;
    RCHK(ensure_fields(h, 1));
    c.g->gconv(c, v, h->ws);
    RCHK(fv_scalar(h, x, v, h->ws));
    c.g->force(c, x, v, c.scal + SC_FV, coef);
    return spm_launch_status("forces");
}

} // namespace

extern "C" {

int spm_get_tv(spm_ctx *h, const void *phi0, const void *phi1, void *v)
{
    CHECK_CTX(h);
    if (!phi0 || !phi1 || !v) return spm_fail(SPM_EINVAL, "spm_get_tv: null field");
    return h->c.prec == SPM_FP64 ? launch_tv<double>(h, phi0, phi1, v) : launch_tv<float>(h, phi0, phi1, v);
}

int spm_natural_force(spm_ctx *h, void *x, const void *v)
{
    CHECK_CTX(h);
    if (!x || !v || x == v) return spm_fail(SPM_EINVAL, "spm_natural_force: bad fields");
    return unit_force(h, x, v, 1.0);
}

int spm_minmax_force(spm_ctx *h, void *x, const void *v)
{
    CHECK_CTX(h);
    if (!x || !v || x == v) return spm_fail(SPM_EINVAL, "spm_minmax_force: bad fields");
    return unit_force(h, x, v, 2.0);
}

// One iteration is, in the reference's order (SPM.cpp:183-254):
//   get_Tv; Natural_Forces on both replicas; spring; get_Tv;
//   MinMax_Forces and ngrad_cam on both replicas; spring.
// Here the tangent d = Phi1 - Phi0 stays unnormalised: the force of replica r
// along t = d / nrmp(d) is <f_r, t> t = (<f_r, d> / nrmp(d)^2) d, so the
// division by nrmp(d) moves into the two force scalars.  A force needs
// <f_r, d>, which takes G*d; one round trip of d gives both replicas' sums (the
// forces on one replica do not touch the other), and its input pass also does
// the spring step that precedes each get_Tv.  So the spring at the end of an
// iteration also forms the next iteration's tangent.  ngrad_cam keeps only its
// norm (the residual), and its input transform is the force round trip's
// output pass.  The host reads the residual once per iteration.
int spm_spring_pairing(spm_ctx *h, void *phi0, void *phi1, int maxiter, double tol, int *iters, double *res_out,
                       double *res_hist)
{
    CHECK_CTX(h);
    if (!phi0 || !phi1 || phi0 == phi1) return spm_fail(SPM_EINVAL, "spm_spring_pairing: bad replicas");
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_spring_pairing: maxiter >= 0");
    RCHK(ensure_fields(h, 1));
    void *d = h->ws; // Phi1 - Phi0
    void *Phi[2] = {phi0, phi1};
    const double dt = 0.001, ks = 1; // SPM.cpp:177-178
    double res = 1e-6;               // SPM.cpp:180
    int iter = 0;
    if (res > tol && maxiter > 0) RCHK(tangent(h, Phi[0], Phi[1], d, false, dt, ks));
    while (res > tol && iter < maxiter) {
        iter = iter + 1;
        for (int i = 0; i < 2; ++i) c.g->force(c, Phi[i], d, c.scal + (i ? SC_FV1 : SC_FV), 1.0);
        RCHK(tangent(h, Phi[0], Phi[1], d, true, dt, ks));
        for (int i = 0; i < 2; ++i) {
            c.g->force_ngrad(c, Phi[i], d, c.scal + (i ? SC_FV1 : SC_FV), 2.0);
            hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(1), 0, c.stream, c.scal + SC_GN2, c.scal + SC_RES0 + i);
        }
        RCHK(tangent(h, Phi[0], Phi[1], d, true, dt, ks)); // and the next iteration's tangent
        double g2[2];
        RCHK(spm_read_scalars(h, c.scal + SC_RES0, 2, g2));
        res = 0.0;
        for (int i = 0; i < 2; ++i) {
            const double resi = std::sqrt(g2[i]);
            if (res < resi) res = resi;
        }
        if (res_hist) res_hist[iter - 1] = res;
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    if (iters) *iters = iter;
    if (res_out) *res_out = res;
    return spm_launch_status("spm_spring_pairing");
}

} // extern "C"
