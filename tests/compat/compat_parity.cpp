// compat_parity.cpp -- test program: the reference-signature layer (device
// fields, include/compat/) against the CPU oracle (oracle/lp_oracle.h) on the
// same seeded field.  Prints one JSON object of maximum errors.
//   compat_parity N
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/compat/spm_reference_api.h"
#include "../../oracle/lp_oracle.h"

struct cd {
    double re, im;
};
typedef std::vector<cd> field;
#define FP(v) ((fftw_complex *)(v).data())

static field down(const fftw_complex *d)
{
    field h(CplxDofs_local);
    if (hipMemcpy(h.data(), d, sizeof(fftw_complex) * h.size(), hipMemcpyDeviceToHost) != hipSuccess) std::abort();
    return h;
}
static void up(fftw_complex *d, const field &h)
{
    if (hipMemcpy(d, h.data(), sizeof(fftw_complex) * h.size(), hipMemcpyHostToDevice) != hipSuccess) std::abort();
}
static double maxdiff_re(const field &a, const field &b)
{
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::fabs(a[i].re - b[i].re));
    return m;
}
static double maxdiff_c(const field &a, const field &b)
{
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i)
        m = std::fmax(m, std::fmax(std::fabs(a[i].re - b[i].re), std::fabs(a[i].im - b[i].im)));
    return m;
}

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 16;
    numprocs = 1;
    myid = 0;
    LPsystParameter();
    N = n;
    CplxDofs = N * N * N;
    memAllocation_all();
    get_kindex_all();
    My_mpi_fft_set_up();
    memAllocation_local();
    Distribute_C(kindex, kindex_local);
    get_ksquare_local();
    get_epmckt2_local();
    get_ikt2_local();
    const size_t R = CplxDofs_local;
    lpo_ctx *o = lpo_create(N, L, 1.0, Tau, Gamma, dt, 0);
    // seeded field: two modes plus a small deterministic ripple
    field phi(R), v(R), z(R);
    unsigned s = 12345u;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t l = ((size_t)i * N + j) * N + k;
                s = s * 1664525u + 1013904223u;
                const double r = ((s >> 8) / 16777216.0 - 0.5);
                phi[l].re = 0.1 * std::cos(2 * M_PI * (2 * i + j) / N) + 0.05 * std::cos(2 * M_PI * (3 * k) / N + 0.4) +
                            0.01 * r;
                phi[l].im = 0.0;
                s = s * 1664525u + 1013904223u;
                v[l].re = ((s >> 8) / 16777216.0 - 0.5);
                v[l].im = 0.0;
                z[l].re = phi[l].re;
                z[l].im = v[l].re;
            }
    double e_fft, e_ifft, e_E, e_grad, e_hv, e_pc, e_gf, e_tv, e_nat, e_mm, e_sp, e_inp, e_file;
    {   // My_FFT on a complex field
        up(data, z);
        my_fft(data, fftw_Ctmp_local);
        field ref(R);
        lpo_my_fft(o, FP(z), FP(ref));
        e_fft = maxdiff_c(down(fftw_Ctmp_local), ref);
        my_ifft(data, fftw_Ctmp_local);
        lpo_my_ifft(o, FP(z), FP(ref));
        e_ifft = maxdiff_c(down(fftw_Ctmp_local), ref);
    }
    {   // get_Ene
        up(u_R_local, phi);
        my_fft(u_R_local, u_F_local);
        const double E = get_Ene();
        field F(R);
        lpo_my_fft(o, FP(phi), FP(F));
        const double Er = lpo_get_Ene(o, FP(phi), FP(F));
        e_E = std::fabs(E - Er) / std::fabs(Er);
    }
    {   // ngrad_cam, hv_cam, precond_camnew
        field ref(R);
        up(data, phi);
        ngrad_cam(data, fftw_Ctmp_local);
        lpo_ngrad_cam(o, FP(phi), FP(ref));
        e_grad = maxdiff_re(down(fftw_Ctmp_local), ref);
        up(fftw_Rtmp_local, v);
        hv_cam(data, fftw_Rtmp_local, fftw_Ctmp_local);
        lpo_hv_cam(o, FP(phi), FP(v), FP(ref));
        e_hv = maxdiff_re(down(fftw_Ctmp_local), ref);
        precond_camnew(fftw_Rtmp_local);
        ref = v;
        lpo_precond_camnew(o, FP(ref));
        e_pc = maxdiff_re(down(fftw_Rtmp_local), ref);
        e_inp = std::fabs(inpfp(data, fftw_Rtmp_local) - lpo_inpfp(o, FP(phi), FP(ref)));
    }
    {   // get_min_u_F, 40 iterations
        max_iter = 40;
        up(u_R_local, phi);
        my_fft(u_R_local, u_F_local);
        get_min_u_F();
        field uR = phi, uF(R);
        lpo_my_fft(o, FP(uR), FP(uF));
        lpo_get_min_u_F(o, FP(uR), FP(uF), 40, 1e-16, nullptr, nullptr, nullptr);
        e_gf = maxdiff_re(down(u_R_local), uR);
    }
    {   // SPM pieces and 4 Spring_Pairing iterations
        field p0 = phi, p1 = phi, tvr(R);
        for (size_t l = 0; l < R; ++l) p1[l].re += 0.001 * v[l].re;
        up(Phi0[0], p0);
        up(Phi0[1], p1);
        get_Tv(Phi0, tv);
        lpo_get_Tv(o, FP(p0), FP(p1), FP(tvr));
        e_tv = maxdiff_re(down(tv), tvr);
        field x = p0;
        up(data, x);
        Natural_Forces(data, tv);
        lpo_Natural_Forces(o, FP(x), FP(tvr));
        e_nat = maxdiff_re(down(data), x);
        x = p1;
        up(data, x);
        MinMax_Forces(data, tv);
        lpo_MinMax_Forces(o, FP(x), FP(tvr));
        e_mm = maxdiff_re(down(data), x);
        Spring_Pairing(Phi0, 4, 1e-30);
        lpo_Spring_Pairing(o, FP(p0), FP(p1), 4, 1e-30, nullptr);
        e_sp = std::fmax(maxdiff_re(down(Phi0[0]), p0), maxdiff_re(down(Phi0[1]), p1));
    }
    {   // File_Op round trip
        for (size_t l = 0; l < R; ++l) u_R[l][0] = phi[l].re, u_R[l][1] = 0.0;
        char name[] = "compat_parity_phi.txt";
        Save_phi(u_R, name);
        for (size_t l = 0; l < R; ++l) u_R[l][0] = 0.0;
        load_phi(u_R, name);
        std::remove(name);
        e_file = 0;
        for (size_t l = 0; l < R; ++l) e_file = std::fmax(e_file, std::fabs(u_R[l][0] - phi[l].re));
    }
    // LOBPCG at phi = 0: the lowest eigenvalue is min over k != 0 of G(k) - Tau
    setCplxZero_v(u_R_local, CplxDofs_local);
    Option *opt = (Option *)std::malloc(sizeof(Option));
    hiosd_lobpcg_initialization(opt, 2);
    for (int r = 0; r < 12; ++r) hiosd_lobpcg_inip(u_R_local, opt);
    const double lam = rayleighq(u_R_local, opt->V[0]);
    double lam_ref = 1e300;
    for (size_t l = 1; l < R; ++l) lam_ref = std::fmin(lam_ref, epmckt2_local[l] - Tau);
    Memfree_Option(opt);
    double e_hx, e_hv1, e_hres;
    {   // hiosd_lobpcg_sieh, 5 steps from phi with one unit mode v
        field vu = v;
        double nv = 0;
        for (size_t l = 0; l < R; ++l) nv += vu[l].re * vu[l].re;
        nv = std::sqrt(nv / R);
        for (size_t l = 0; l < R; ++l) vu[l].re /= nv;
        Option *h = (Option *)std::malloc(sizeof(Option));
        hiosd_lobpcg_initialization(h, 1);
        h->dt = 0.8;
        h->betat = 8.0;
        h->betau = 0.02;
        h->epsf = 1e-30;
        h->maxiter = 5;
        up(h->V[0], vu);
        up(u_R_local, phi);
        hiosd_lobpcg_sieh(u_R_local, h);
        field x = phi, vr = vu;
        double res = 0;
        lpo_hiosd_lobpcg_sieh(o, FP(x), FP(vr), 1, 0.8, 8.0, 0.02, 1e-30, 5, &res);
        e_hx = maxdiff_re(down(u_R_local), x);
        const field vg = down(h->V[0]);
        double dot = 0, ng = 0, nr = 0;
        for (size_t l = 0; l < R; ++l) dot += vg[l].re * vr[l].re, ng += vg[l].re * vg[l].re, nr += vr[l].re * vr[l].re;
        e_hv1 = 1.0 - std::fabs(dot) / std::sqrt(ng * nr);
        // the layer leaves x in u_R_local and its spectrum in u_F_local; its
        // residual is nrmp(ngrad_cam(x)), recomputed here from the result
        ngrad_cam(u_R_local, fftw_Ctmp_local);
        e_hres = std::fabs(nrmp(fftw_Ctmp_local) - res) / res;
        Memfree_Option(h);
    }
    double e_gx, e_gv, e_gperp;
    {   // hiosd_lobpcg_sieh_V0, 3 steps with a null space of 2 orthonormal mean-free fields
        const int k0 = 2;
        std::vector<field> w(k0, field(R));
        unsigned t = 777u;
        for (int a = 0; a < k0; ++a) {
            double m = 0;
            for (size_t l = 0; l < R; ++l) {
                t = t * 1664525u + 1013904223u;
                w[a][l].re = (t >> 8) / 16777216.0 - 0.5;
                m += w[a][l].re;
            }
            for (size_t l = 0; l < R; ++l) w[a][l].re -= m / R;
            for (int b = 0; b < a; ++b) {  // Gram-Schmidt under the mean of products
                double d = 0;
                for (size_t l = 0; l < R; ++l) d += w[a][l].re * w[b][l].re;
                for (size_t l = 0; l < R; ++l) w[a][l].re -= d / R * w[b][l].re;
            }
            double nn = 0;
            for (size_t l = 0; l < R; ++l) nn += w[a][l].re * w[a][l].re;
            for (size_t l = 0; l < R; ++l) w[a][l].re /= std::sqrt(nn / R);
        }
        field vu = v, w2(2 * R);
        double nv = 0;
        for (size_t l = 0; l < R; ++l) nv += vu[l].re * vu[l].re;
        for (size_t l = 0; l < R; ++l) vu[l].re /= std::sqrt(nv / R);
        for (int a = 0; a < k0; ++a)
            for (size_t l = 0; l < R; ++l) w2[a * R + l] = w[a][l];
        fftw_complex *V0[k0];
        for (int a = 0; a < k0; ++a) {
            if (hipMalloc(&V0[a], sizeof(fftw_complex) * R) != hipSuccess) std::abort();
            up(V0[a], w[a]);
        }
        Option *h = (Option *)std::malloc(sizeof(Option));
        hiosd_lobpcg_initialization(h, 1);
        h->dt = 0.8;
        h->betat = 8.0;
        h->betau = 1.3;
        h->epsf = 1e-30;
        h->maxiter = 3;
        up(h->V[0], vu);
        up(u_R_local, phi);
        hiosd_lobpcg_sieh_V0(u_R_local, h, V0, k0);
        field x = phi, vr = vu;
        double res = 0;
        lpo_hiosd_lobpcg_sieh_V0(o, FP(x), FP(vr), 1, FP(w2), k0, 0.8, 8.0, 1.3, 1e-30, 3, &res);
        e_gx = maxdiff_re(down(u_R_local), x);
        const field vg = down(h->V[0]);
        double dot = 0, ng = 0, nr = 0;
        for (size_t l = 0; l < R; ++l) dot += vg[l].re * vr[l].re, ng += vg[l].re * vg[l].re, nr += vr[l].re * vr[l].re;
        e_gv = 1.0 - std::fabs(dot) / std::sqrt(ng * nr);
        e_gperp = 0;
        for (int a = 0; a < k0; ++a) {
            double d = 0;
            for (size_t l = 0; l < R; ++l) d += vg[l].re * w[a][l].re;
            e_gperp = std::fmax(e_gperp, std::fabs(d / R));
            (void)hipFree(V0[a]);
        }
        Memfree_Option(h);
    }
    std::printf("{\"fft\": %.3e, \"ifft\": %.3e, \"energy_rel\": %.3e, \"ngrad\": %.3e, \"hv\": %.3e, "
                "\"precond\": %.3e, \"inpfp\": %.3e, \"gradflow\": %.3e, \"tv\": %.3e, \"natural\": %.3e, "
                "\"minmax\": %.3e, \"spring\": %.3e, \"file\": %.3e, \"lobpcg_lowest\": %.3e, "
                "\"hisd_x\": %.3e, \"hisd_v\": %.3e, \"hisd_res_rel\": %.3e, \"gosd_x\": %.3e, \"gosd_v\": %.3e, "
                "\"gosd_perp\": %.3e}\n",
                e_fft, e_ifft, e_E, e_grad, e_hv, e_pc, e_inp, e_gf, e_tv, e_nat, e_mm, e_sp, e_file,
                std::fabs(lam - lam_ref), e_hx, e_hv1, e_hres, e_gx, e_gv, e_gperp);
    lpo_destroy(o);
    Memfree_global_kspace();
    Memfree_global_data();
    Memfree_local_kspace();
    Memfree_local_data();
    return 0;
}
