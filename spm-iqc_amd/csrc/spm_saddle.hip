// spm_saddle.hip -- the Spring Pair Method for an index-1 saddle
// (get_Tv, Natural_Forces, MinMax_Forces, Spring_Pairing; SPM.cpp:15-263)
// on two device-resident replicas.  Scalars that the reference gathers with
// MPI_Allreduce (the tangent norm, <f, v>) stay on the device; the host reads
// only the residual, once per iteration.
#include <algorithm>
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

// The spring half-steps of SPM.cpp:203-216 (and 233-246), replicas updated in
// the reference's order: Phi0 += dt ks (Phi1 - Phi0); Phi1 += dt ks (Phi0' - Phi1).
// The get_Tv that follows (SPM.cpp:218, and 185 of the next iteration) rides
// along: d = Phi1' - Phi0', left unnormalised.  So do the real-space sums of the
// next two force scalars: d^2; per replica nl_r d and nl_r
// (nl = Tau x + Gamma x^2 - x^3); d; and Phi0'^2 (the scale of the replicas).
constexpr int SPRING_SUMS = 7;
template <typename T>
__global__ void __launch_bounds__(256) k_spring_sums(T *p0, T *p1, T *__restrict__ d, long n, double dt, double ks,
                                                    double Tau, double Gamma, double *__restrict__ part)
{
    __shared__ double red[SPRING_SUMS * 16];
    double s[SPRING_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double a = (double)p0[i], b = (double)p1[i];
        const T a1 = (T)(1 * a + dt * (ks * b + -ks * a));
        const T b1 = (T)(1 * b + dt * (ks * (double)a1 + -ks * b));
        p0[i] = a1;
        p1[i] = b1;
        const T di = b1 - a1;
        d[i] = di;
        const double dd = (double)di, u0 = (double)a1, u1 = (double)b1;
        const double n0 = Tau * u0 + Gamma * u0 * u0 - u0 * u0 * u0;
        const double n1 = Tau * u1 + Gamma * u1 * u1 - u1 * u1 * u1;
        s[0] += dd * dd;
        s[1] += n0 * dd;
        s[2] += n0;
        s[3] += n1 * dd;
        s[4] += n1;
        s[5] += dd;
        s[6] += u0 * u0;
    }
    block_reduce<SPRING_SUMS, 0u>(s, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < SPRING_SUMS; ++q) part[(long)q * gridDim.x + blockIdx.x] = s[q];
}

// The force scalars after a spring step: tv2 = nrmp(d)^2 and, for both
// replicas, <f_r, d> / nrmp(d)^2 (k_fv_fin's formula).  Its G-convolution terms
// <x_r, G*d> come from the quadratic forms Q_pq = <x_p, G*x_q> of the replicas
// before the spring step (pq: sums of the force passes, SOpForceSpec), because
// that step is linear: with a = dt ks,
//   x0' = (1 - a) x0 + a x1,  x1' = a (1 - a) x0 + (1 - a + a^2) x1,
//   d' = x1' - x0' = (1 - a)^2 (x1 - x0).
// These combinations cancel: their rounding error is about eps |x|^2 |G|
// whatever |d|, so the force term's error grows like 1 / |d| as the replicas
// close.  drel = |d|^2 / mean(x0^2) tells the host when to take the terms from a
// G convolution of d instead (first_tangent), whose error is relative to |d|.
__global__ void __launch_bounds__(256) k_spring_fin(const double *__restrict__ ps, int nbs,
                                                   const double *__restrict__ pq, int nbq, double R, double a,
                                                   double *tv2, double *out0, double *out1, double *drel)
{
    constexpr int NS = SPRING_SUMS + 3;
    __shared__ double red[NS * 16];
    double s[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = 0.0;
    for (int b = threadIdx.x; b < nbs; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < SPRING_SUMS; ++q) s[q] += ps[(long)q * nbs + b];
    for (int b = threadIdx.x; b < nbq; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 3; ++q) s[SPRING_SUMS + q] += pq[(long)q * nbq + b];
    block_reduce<NS, 0u>(s, red);
    if (threadIdx.x == 0) {
        const double Q00 = s[SPRING_SUMS], Q01 = s[SPRING_SUMS + 1], Q11 = s[SPRING_SUMS + 2];
        const double A[2][2] = {{1 - a, a}, {a * (1 - a), 1 - a + a * a}};
        const double D0 = -(1 - a) * (1 - a), D1 = (1 - a) * (1 - a);
        const double d2 = s[0] / R, mv = s[5] / R;
        *tv2 = d2;
        *drel = d2 / (s[6] / R);
        double *out[2] = {out0, out1};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const double xgd = A[r][0] * (D0 * Q00 + D1 * Q01) + A[r][1] * (D0 * Q01 + D1 * Q11);
            *out[r] = (-xgd + s[1 + 2 * r] / R - (s[2 + 2 * r] / R) * mv) / d2;
        }
    }
}


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

// Spring_Pairing's first tangent: d = Phi1 - Phi0 unnormalised,
// scal[SC_TV2] = nrmp(d)^2 and the force scalars scal[SC_FV], scal[SC_FV1] =
// <f_r, d> / nrmp(d)^2 of both replicas, from one G-convolution round trip of d
// (its passes form d and the sums).
int first_tangent(spm_ctx *h, const void *p0, const void *p1, void *d)
{
    Ctx &c = h->c;
    double *pt = c.slot(5), *pf = c.slot(6);
    const long nb = c.g->spm_tangent(c, p0, p1, d, pt, pf);
    if (7 * nb > c.part_slot) return spm_fail(SPM_EINVAL, "spm tangent: partial sums exceed a slot");
    launch_reduce(c, pt, nb, 1, 0u, 1.0 / (double)c.R, c.scal + SC_TV2);
    hipLaunchKernelGGL(k_fv_fin<2>, dim3(1), dim3(256), 0, c.stream, pf, (int)nb, (double)c.R, c.scal + SC_TV2,
                       c.scal + SC_FV, c.scal + SC_FV1);
    return spm_launch_status("spm tangent");
}

// The spring step of a replica pair and the next tangent d with its force
// scalars; fs holds the Q sums of the force pair just done.
int spring_tangent(spm_ctx *h, void *p0, void *p1, void *d, double dt, double ks, const ForceSpec &fs)
{
    Ctx &c = h->c;
    double *ps = c.slot(5);
    // 2048 blocks (8 per CU) keep enough loads in flight
    const int nbs = (int)std::min<long>(2048, c.part_slot / SPRING_SUMS);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_spring_sums<double>, dim3(nbs), dim3(256), 0, c.stream, (double *)p0, (double *)p1,
                           (double *)d, c.R, dt, ks, c.Tau, c.Gamma, ps);
    else
        hipLaunchKernelGGL(k_spring_sums<float>, dim3(nbs), dim3(256), 0, c.stream, (float *)p0, (float *)p1,
                           (float *)d, c.R, dt, ks, c.Tau, c.Gamma, ps);
    hipLaunchKernelGGL(k_spring_fin, dim3(1), dim3(256), 0, c.stream, ps, nbs, fs.part, (int)fs.nb, (double)c.R,
                       dt * ks, c.scal + SC_TV2, c.scal + SC_FV, c.scal + SC_FV1, c.scal + SC_DREL);
    return spm_launch_status("spring");
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

int ensure_ws(spm_ctx *h, size_t bytes)
{
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
    RCHK(ensure_ws(h, (size_t)c.R * c.esz));
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
// division by nrmp(d) moves into the two force scalars.  <f_r, d> needs
// <x_r, G*d>.  Every get_Tv but the first follows a spring step, which is
// linear, so those terms are combinations of <x_p, G*x_q> of the replicas the
// force pair has just produced; the force passes sum these in spectral space
// (Parseval) from the spectra they already hold, and the spring kernel forms d
// and the real-space sums.  So no transform of d is needed after the first.
// ngrad_cam keeps only its norm (the residual), and its input transform is the
// force round trip's output pass.  The host reads the residual once per iteration,
// with |d|^2 / mean(Phi0^2): below CLOSE_FP64 (CLOSE_FP32), where the spectral
// combinations have lost too many digits to cancellation, the next iteration's
// tangents come from a G convolution of d itself, as the first one does.
int spm_spring_pairing(spm_ctx *h, void *phi0, void *phi1, int maxiter, double tol, int *iters, double *res_out,
                       double *res_hist)
{
    CHECK_CTX(h);
    if (!phi0 || !phi1 || phi0 == phi1) return spm_fail(SPM_EINVAL, "spm_spring_pairing: bad replicas");
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_spring_pairing: maxiter >= 0");
    const size_t dbytes = (size_t)c.R * c.esz;
    RCHK(ensure_ws(h, dbytes + (size_t)c.C * 2 * c.esz));
    void *d = h->ws;                    // Phi1 - Phi0
    void *keep = (char *)h->ws + dbytes; // spectrum of the first replica after its force
    void *Phi[2] = {phi0, phi1};
    const double dt = 0.001, ks = 1; // SPM.cpp:177-178
    double res = 1e-6;               // SPM.cpp:180
    int iter = 0;
    ForceSpec f0, f1;
    f0.keep = keep;
    f1.other = keep;
    f1.part = c.slot(7);
    f0.gn2 = c.scal + SC_RES0; // each replica's residual^2 straight from its ngrad
    f1.gn2 = c.scal + SC_RES1;
    // |d|^2 / mean(Phi0^2) below which the tangent terms come from d's own transform:
    // the spectral form then keeps about 10 digits of the force scalar in FP64 and 5 in FP32
    const double close = c.prec == SPM_FP64 ? 1e-12 : 1e-4;
    bool direct = true; // until the first read-back gives the ratio
    if (res > tol && maxiter > 0) RCHK(first_tangent(h, Phi[0], Phi[1], d));
    while (res > tol && iter < maxiter) {
        iter = iter + 1;
        c.g->force(c, Phi[0], d, c.scal + SC_FV, 1.0, &f0);
        c.g->force(c, Phi[1], d, c.scal + SC_FV1, 1.0, &f1);
        RCHK(spring_tangent(h, Phi[0], Phi[1], d, dt, ks, f1));
        if (direct) RCHK(first_tangent(h, Phi[0], Phi[1], d));
        for (int i = 0; i < 2; ++i) c.g->force_ngrad(c, Phi[i], d, c.scal + (i ? SC_FV1 : SC_FV), 2.0, i ? &f1 : &f0);
        RCHK(spring_tangent(h, Phi[0], Phi[1], d, dt, ks, f1)); // and the next iteration's tangent
        double g2[4]; // SC_RES0, SC_RES1, SC_FV1, SC_DREL
        RCHK(spm_read_scalars(h, c.scal + SC_RES0, 4, g2));
        direct = !(g2[3] >= close); // a NaN ratio (coinciding or zero replicas) as well
        if (direct) RCHK(first_tangent(h, Phi[0], Phi[1], d));
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
