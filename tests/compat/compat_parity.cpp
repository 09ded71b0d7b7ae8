// compat_parity.cpp -- test program: the reference-signature layer (device
// fields, include/compat/) against the CPU oracle (oracle/lp_oracle.h) on the
// same seeded field.  Prints one JSON object of maximum errors.
//   compat_parity N                  every function of the layer
//   compat_parity seeds N L          get_initial_value of each initData file in ./initData
//                                    (host only: prints the nonzero spectrum entries)
//   compat_parity drivers NAME N     one HiSD / GOSD driver (HISD.cpp:13-83, GOSD.cpp:13-120)
//                                    in the {110} cell, against the oracle's solver run with
//                                    that driver's parameters from the same inputs
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
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
// get_Ene of a device field through the layer (u_F_local = its transform)
static double get_Ene_of(fftw_complex *x)
{
    Cpy(x, u_R_local, CplxDofs_local);
    my_fft(u_R_local, u_F_local);
    return get_Ene();
}
static double lpo_get_Ene_of(lpo_ctx *o, const std::vector<cd> &x)
{
    std::vector<cd> F(x.size());
    lpo_my_fft(o, (const lpo_cplx *)x.data(), (lpo_cplx *)F.data());
    return lpo_get_Ene(o, (const lpo_cplx *)x.data(), (const lpo_cplx *)F.data());
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

static void setup(int n, double box)
{
    numprocs = 1;
    myid = 0;
    LPsystParameter();
    N = n;
    if (box > 0) { // a box of side box in the reference's metric (ProjMatrix = 1 / L)
        L = box;
        for (int d = 0; d < dimCpt; ++d) ProjMatrix[d][d] = 1.0 / L;
    }
    CplxDofs = N * N * N;
    memAllocation_all();
    get_kindex_all();
    My_mpi_fft_set_up();
    memAllocation_local();
    Distribute_C(kindex, kindex_local);
    get_ksquare_local();
    get_epmckt2_local();
    get_ikt2_local();
}

// get_initial_value (Initialization.cpp:242-338) of every selectable seed file
static int seeds_main(int n, double box)
{
    numprocs = 1;
    myid = 0;
    LPsystParameter();
    N = n;
    L = box;
    CplxDofs = N * N * N;
    std::vector<cd> v(CplxDofs);
    std::string out = "{";
    char buf[128];
    for (int choice = 0; choice <= 5; ++choice) {
        get_initial_value(FP(v), choice, 0.0);
        std::snprintf(buf, sizeof buf, "%s\"%d\": [", choice ? ", " : "", choice);
        out += buf;
        bool first = true;
        for (int l = 0; l < CplxDofs; ++l)
            if (v[l].re != 0.0 || v[l].im != 0.0) {
                std::snprintf(buf, sizeof buf, "%s[%d, %.17g, %.17g]", first ? "" : ", ", l, v[l].re, v[l].im);
                out += buf;
                first = false;
            }
        out += "]";
    }
    std::printf("\n%s}\n", out.c_str());
    return 0;
}

// Max-norm of a - b (real parts) after removing its least-squares component along
// the three translation modes d_x b, d_y b, d_z b of the periodic box (spectral
// derivatives; the Nyquist index has no derivative).  Translations are exact
// symmetries of the free energy, so critical points are unique only up to them.
static double maxdiff_mod_translation(lpo_ctx *o, const field &a, const field &b)
{
    const size_t R = a.size();
    field F(R), T[3];
    lpo_my_fft(o, FP(b), FP(F));
    for (int ax = 0; ax < 3; ++ax) {
        field D(R);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    const int m = ax == 0 ? i : ax == 1 ? j : k;
                    const double n = 2 * m < N ? m : 2 * m == N ? 0.0 : m - N;
                    const size_t l = ((size_t)i * N + j) * N + k;
                    D[l] = cd{-n * F[l].im, n * F[l].re}; // i n F
                }
        T[ax].resize(R);
        lpo_my_ifft(o, FP(D), FP(T[ax]));
    }
    std::vector<double> d(R);
    for (size_t l = 0; l < R; ++l) d[l] = a[l].re - b[l].re;
    double G[3][3] = {}, r[3] = {};
    for (int p = 0; p < 3; ++p) {
        for (size_t l = 0; l < R; ++l) r[p] += T[p][l].re * d[l];
        for (int q = 0; q < 3; ++q)
            for (size_t l = 0; l < R; ++l) G[p][q] += T[p][l].re * T[q][l].re;
    }
    // solve G c = r (symmetric positive semidefinite: Gaussian elimination with a
    // relative pivot threshold, so a direction with no derivative drops out)
    double c[3] = {0, 0, 0};
    int piv[3] = {0, 0, 0};
    const double scale = std::fmax(std::fmax(G[0][0], G[1][1]), G[2][2]);
    for (int p = 0; p < 3; ++p) {
        piv[p] = G[p][p] > 1e-12 * scale;
        if (!piv[p]) continue;
        for (int q = p + 1; q < 3; ++q) {
            const double f = G[q][p] / G[p][p];
            for (int s = p; s < 3; ++s) G[q][s] -= f * G[p][s];
            r[q] -= f * r[p];
        }
    }
    for (int p = 2; p >= 0; --p) {
        if (!piv[p]) continue;
        double s = r[p];
        for (int q = p + 1; q < 3; ++q) s -= G[p][q] * c[q];
        c[p] = s / G[p][p];
    }
    double m = 0;
    for (size_t l = 0; l < R; ++l)
        m = std::fmax(m, std::fabs(d[l] - c[0] * T[0][l].re - c[1] * T[1][l].re - c[2] * T[2][l].re));
    return m;
}

// One driver of HISD.cpp / GOSD.cpp in the {110} cell (L = sqrt 2): u_R_local is
// the index-1 saddle S that main_SPM's Spring_Pairing reaches from phi = 0, and
// option1 holds the lowest 4 Hessian modes at S (unstable mode, 3 translations).
// The oracle runs the same solver with the driver's parameters from the same
// perturbed field and directions.
static int drivers_main(const char *name, int n)
{
    setup(n, std::sqrt(2.0));
    const size_t R = CplxDofs_local;
    lpo_ctx *o = lpo_create(N, L, 1.0, Tau, Gamma, dt, 0);
    option1 = (Option *)std::malloc(sizeof(Option)); // as main_SPM.cpp:82 and grad.cpp:89 do
    option2 = (Option *)std::malloc(sizeof(Option));
    setCplxZero_v(u_R_local, CplxDofs_local);
    hiosd_lobpcg_initialization(option1, 2);
    for (int r = 0; r < 10; ++r) hiosd_lobpcg_inip(u_R_local, option1);
    Cpy(u_R_local, Phi0[0], CplxDofs_local);
    FuncsLinear2Cplx(Phi0[1], CplxDofs_local, 1.0, u_R_local, 0.001, option1->V[0]);
    Spring_Pairing(Phi0, 30000, 1e-9);
    Cpy(Phi0[1], u_R_local, CplxDofs_local);
    Memfree_Option(option1); // frees the Option as well (Memfree.cpp:65-73)
    option1 = (Option *)std::malloc(sizeof(Option));
    hiosd_lobpcg_initialization(option1, 4);
    for (int r = 0; r < 10; ++r) hiosd_lobpcg_inip(u_R_local, option1);
    const field S = down(u_R_local);
    std::vector<field> V1(4);
    for (int i = 0; i < 4; ++i) V1[i] = down(option1->V[i]);
    // the oracle's inputs, as the driver forms them
    field x = S, v(R), v0(3 * R);
    int k0 = 0, maxiter = 0;
    double betau = 0, epsf = 0;
    const std::string d(name);
    if (d == "critical_point_HISD" || d == "critical_point_HISD_v") { // HISD.cpp:15-29, 34-82
        for (size_t l = 0; l < R; ++l) x[l].re += 0.01 * V1[0][l].re, x[l].im += 0.01 * V1[0][l].im;
        v = V1[0];
        betau = d == "critical_point_HISD" ? 0.02 : 0.6;
        epsf = 1e-7;
        maxiter = 10000;
    } else if (d == "critical_point") { // GOSD.cpp:13-58
        k0 = 3;
        for (size_t l = 0; l < R; ++l) x[l].re -= 0.1 * V1[3][l].re, x[l].im -= 0.1 * V1[3][l].im;
        v = V1[3];
        betau = 1.3;
        epsf = 1e-9;
        maxiter = 40000;
    } else if (d == "critical_point_v") { // GOSD.cpp:60-120: the HEX seed file, inverse-transformed
        k0 = 3;
        std::vector<cd> hs(CplxDofs);
        get_initial_value(FP(hs), 3, 0.0);
        field vr(R);
        lpo_my_ifft(o, FP(hs), FP(vr));
        for (size_t l = 0; l < R; ++l) x[l].re += 0.1 * vr[l].re, x[l].im += 0.1 * vr[l].im;
        v = vr;
        betau = 0.4;
        epsf = 1e-7;
        maxiter = 20000;
    } else {
        std::fprintf(stderr, "unknown driver %s\n", name);
        return 2;
    }
    for (size_t l = 0; l < R; ++l) v[l].im = 0.0;
    for (int a = 0; a < k0; ++a)
        for (size_t l = 0; l < R; ++l) v0[a * R + l] = cd{V1[a][l].re, 0.0};
    double res_o = 0;
    const int it_o = k0 ? lpo_hiosd_lobpcg_sieh_V0(o, FP(x), FP(v), 1, FP(v0), k0, 0.8, 8.0, betau, epsf, maxiter, &res_o)
                        : lpo_hiosd_lobpcg_sieh(o, FP(x), FP(v), 1, 0.8, 8.0, betau, epsf, maxiter, &res_o);
    if (d == "critical_point_HISD") critical_point_HISD();
    else if (d == "critical_point_HISD_v") critical_point_HISD_v();
    else if (d == "critical_point") critical_point();
    else critical_point_v();
    const field xg = down(u_R_local);
    const double ex = maxdiff_re(xg, x), ext = maxdiff_mod_translation(o, xg, x);
    // the oracle at the GPU's end point: its residual, and |H dx| for the difference
    // dx of the two end points (H at the oracle's), which the two stopping tests
    // bound by |ngrad(x_gpu)| + |ngrad(x_oracle)| to first order
    field gg(R), dx(R), hdx(R);
    for (size_t l = 0; l < R; ++l) dx[l] = cd{xg[l].re - x[l].re, 0.0};
    lpo_ngrad_cam(o, FP(xg), FP(gg));
    const double res_og = lpo_nrmp(o, FP(gg));
    lpo_hv_cam(o, FP(x), FP(dx), FP(hdx));
    const double nhdx = lpo_nrmp(o, FP(hdx)), ndx = lpo_nrmp(o, FP(dx));
    ngrad_cam(u_R_local, fftw_Ctmp_local);
    const double res = nrmp(fftw_Ctmp_local);
    const field vg = down(option2->V[0]);
    double dot = 0, ng = 0, nr = 0;
    for (size_t l = 0; l < R; ++l) dot += vg[l].re * v[l].re, ng += vg[l].re * vg[l].re, nr += v[l].re * v[l].re;
    std::printf("\n{\"driver\": \"%s\", \"iters_oracle\": %d, \"res_oracle\": %.3e, \"res\": %.3e, "
                "\"epsf\": %.1e, \"maxiter\": %d, \"x\": %.3e, \"x_mod_translation\": %.3e, \"dx\": %.3e, "
                "\"hdx\": %.3e, \"res_oracle_at_gpu\": %.3e, \"v\": %.3e, "
                "\"E\": %.15e, \"E_oracle\": %.15e, \"E_S\": %.15e}\n",
                name, it_o, res_o, res, epsf, maxiter, ex, ext, ndx, nhdx, res_og, 1.0 - std::fabs(dot) / std::sqrt(ng * nr),
                get_Ene_of(u_R_local),
                lpo_get_Ene_of(o, x), lpo_get_Ene_of(o, S));
    lpo_destroy(o);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::string(argv[1]) == "seeds")
        return seeds_main(argc > 2 ? std::atoi(argv[2]) : 16, argc > 3 ? std::atof(argv[3]) : 26.0);
    if (argc > 3 && std::string(argv[1]) == "drivers") return drivers_main(argv[2], std::atoi(argv[3]));
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
    double e_Euf;
    {   // get_Ene reads u_F_local (LPsrc.cpp:27-35): here a spectrum that is not the
        // transform of u_R_local (the transform plus 0.3 that of phi + i v, not Hermitian)
        up(u_R_local, phi);
        field F(R), F2(R);
        lpo_my_fft(o, FP(phi), FP(F));
        lpo_my_fft(o, FP(z), FP(F2));
        for (size_t l = 0; l < R; ++l) F[l].re += 0.3 * F2[l].re, F[l].im += 0.3 * F2[l].im;
        up(u_F_local, F);
        const double E = get_Ene();
        const double Er = lpo_get_Ene(o, FP(phi), FP(F));
        e_Euf = std::fabs(E - Er) / std::fabs(Er);
    }
    // BasicOperators.h:6-50 against host loops of the reference's formulas
    double e_bo = 0.0;
    auto dnew = [&](const field &h) {
        fftw_complex *d = nullptr;
        if (hipMalloc(&d, sizeof(fftw_complex) * R) != hipSuccess) std::abort();
        up(d, h);
        return d;
    };
    auto dot = [&](const field &a, const field &b) {
        double t = 0;
        for (size_t l = 0; l < R; ++l) t += a[l].re * b[l].re;
        return t / R;
    };
    auto bump = [&](double e) { e_bo = std::fmax(e_bo, e); };
    {
        field A = phi, B = v, C(R);
        for (size_t l = 0; l < R; ++l) C[l] = cd{std::sin(0.1 * l), 0.25 * std::cos(0.3 * l)};
        fftw_complex *dA = dnew(A), *dB = dnew(B), *dC = dnew(C);
        FuncRealAddAConst(dC, (int)R, 0.7); // BasicOperators.cpp:39-43
        for (auto &c : C) c.re += 0.7;
        bump(maxdiff_c(down(dC), C));
        FuncCplxAddAConst(dC, (int)R, -0.2); // :28-37
        for (auto &c : C) c.re += -0.2, c.im += -0.2;
        bump(maxdiff_c(down(dC), C));
        norm_fft(dC, (int)R); // :103-109
        for (auto &c : C) c.re = c.re / (1.0 * CplxDofs), c.im = c.im / (1.0 * CplxDofs);
        bump(maxdiff_c(down(dC), C) * CplxDofs);
        setCplxZero(dC[5]); // :205-208
        C[5] = cd{0.0, 0.0};
        bump(maxdiff_c(down(dC), C));
        double *dr = nullptr, *dq = nullptr;
        if (hipMalloc(&dr, sizeof(double) * R) != hipSuccess || hipMalloc(&dq, sizeof(double) * R) != hipSuccess)
            std::abort();
        std::vector<double> hr(R), hq(R);
        for (size_t l = 0; l < R; ++l) hr[l] = A[l].re, hq[l] = B[l].re;
        hipMemcpy(dr, hr.data(), sizeof(double) * R, hipMemcpyHostToDevice);
        hipMemcpy(dq, hq.data(), sizeof(double) * R, hipMemcpyHostToDevice);
        Dot_constant(dr, (int)R, 1.7); // :197-203
        double mx = 0.0;
        for (size_t l = 0; l < R; ++l) {
            hr[l] = 1.7 * hr[l];
            if (mx < hr[l]) mx = hr[l]; // Max (:110-119)
        }
        std::vector<double> back(R);
        hipMemcpy(back.data(), dr, sizeof(double) * R, hipMemcpyDeviceToHost);
        for (size_t l = 0; l < R; ++l) bump(std::fabs(back[l] - hr[l]));
        bump(std::fabs(Max(dr, (int)R) - mx));
        double ip = 0;
        for (size_t l = 0; l < R; ++l) ip += hr[l] * hq[l];
        bump(std::fabs(inpfp(dr, dq) - ip / CplxDofs)); // inpfp(double *, double *) (:145-156)
        bump(std::fabs(nrmp(dq) - std::sqrt(dot(B, B))));
        // Orthoi_v2V (:229-241), in order, imaginary part 0
        field X = C;
        fftw_complex *dX = dnew(X);
        fftw_complex *Vv[2] = {dA, dB};
        Orthoi_v2V(dX, Vv, 2);
        for (const field *Vi : {&A, &B}) {
            const double t = dot(X, *Vi);
            for (size_t l = 0; l < R; ++l) X[l].re -= t * (*Vi)[l].re, X[l].im = 0.0;
        }
        bump(maxdiff_c(down(dX), X));
        // Orthoi_Vi (:243-255): the third of (A, B, C) against the first two
        field Y = C;
        fftw_complex *dY = dnew(Y);
        fftw_complex *Vs[3] = {dA, dB, dY};
        Orthoi_Vi(Vs, 2);
        for (const field *Vi : {&A, &B}) {
            const double t = dot(*Vi, Y);
            for (size_t l = 0; l < R; ++l) Y[l].re -= t * (*Vi)[l].re, Y[l].im = 0.0;
        }
        bump(maxdiff_c(down(dY), Y));
        // PV_Cplx, both overloads (:344-371)
        fftw_complex *dT = dnew(C);
        PV_Cplx(dC, Vv, dT, 2);
        field T(R, cd{0.0, 0.0});
        for (const field *Vi : {&A, &B}) {
            const double t = dot(C, *Vi);
            for (size_t l = 0; l < R; ++l) T[l].re += t * (*Vi)[l].re;
        }
        bump(maxdiff_c(down(dT), T));
        PV_Cplx(dC, dB, dT, 1);
        const double tb = dot(C, B);
        for (size_t l = 0; l < R; ++l) T[l] = cd{tb * B[l].re, 0.0};
        bump(maxdiff_c(down(dT), T));
        // setCplxZero_V (:210-219), get_Merge (:257-281), get_nozero (:283-294)
        double *UU[3] = {dr, dq, nullptr};
        if (hipMalloc(&UU[2], sizeof(double) * R) != hipSuccess) std::abort();
        fftw_complex *one[1] = {dA}, *two[1] = {dB}, *three[1] = {dC};
        get_Merge(UU, one, two, three, 1);
        for (int a = 0; a < 3; ++a) {
            const field &src = a == 0 ? A : (a == 1 ? B : C);
            hipMemcpy(back.data(), UU[a], sizeof(double) * R, hipMemcpyDeviceToHost);
            for (size_t l = 0; l < R; ++l) bump(std::fabs(back[l] - src[l].re));
        }
        int emp[6] = {1, 0, 1, 1, 0, 1};
        bump(std::fabs(get_nozero(emp, 2) - 4.0));
        setCplxZero_V(Vs, 3, (int)R);
        for (fftw_complex *p2 : Vs) bump(maxdiff_c(down(p2), field(R, cd{0.0, 0.0})));
        for (void *p2 : {(void *)dC, (void *)dX, (void *)dT, (void *)dr, (void *)dq, (void *)UU[2]}) hipFree(p2);
        hipFree(dA);
        hipFree(dB);
        hipFree(dY);
    }
    // copy_option_V, the three overloads (LOBPCG.cpp:489-520): real parts, imaginary part 0
    double e_copy = 0.0;
    {
        Option *o1 = (Option *)std::malloc(sizeof(Option)), *o2 = (Option *)std::malloc(sizeof(Option));
        hiosd_lobpcg_initialization(o1, 2);
        hiosd_lobpcg_initialization(o2, 3);
        field C(R);
        for (size_t l = 0; l < R; ++l) C[l] = cd{std::cos(0.07 * l), std::sin(0.05 * l)};
        up(o1->V[0], z);
        up(o1->V[1], C);
        auto re0 = [&](const field &h) {
            field r(R);
            for (size_t l = 0; l < R; ++l) r[l] = cd{h[l].re, 0.0};
            return r;
        };
        copy_option_V(o1, o2, 2);
        e_copy = std::fmax(maxdiff_c(down(o2->V[0]), re0(z)), maxdiff_c(down(o2->V[1]), re0(C)));
        setCplxZero_V(o2->V, 3, (int)R);
        fftw_complex *W[2] = {o1->V[1], o1->V[0]};
        copy_option_V(W, o2, 2);
        e_copy = std::fmax(e_copy, std::fmax(maxdiff_c(down(o2->V[0]), re0(C)), maxdiff_c(down(o2->V[1]), re0(z))));
        copy_option_V(o1->V[1], o2, 2);
        e_copy = std::fmax(e_copy, maxdiff_c(down(o2->V[2]), re0(C)));
        Memfree_Option(o1);
        Memfree_Option(o2);
    }
    std::printf("{\"energy_uf_rel\": %.3e, \"basic_ops\": %.3e, \"copy_option_V\": %.3e, ", e_Euf, e_bo, e_copy);
    std::printf("\"fft\": %.3e, \"ifft\": %.3e, \"energy_rel\": %.3e, \"ngrad\": %.3e, \"hv\": %.3e, "
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
