// grid_impl.h -- GridIface for one (precision, N): the pass sequences of the
// transforms and operators of the reference's path, launched on ctx.stream.
#pragma once
#include "gf_split.h"
#include "spm_ctx.h"

#include <algorithm>

namespace spm {

// TR, TJ, TI: threads per pencil of the row, j-axis and i-axis passes.
template <typename T, int N, int TR = DefTPP<N>::value, int TJ = DefTPP<N>::value, int TI = DefTPP<N>::value>
struct GridImpl final : GridIface {
    using GR = Geo<T, N, TR>;
    using GJ = Geo<T, N, TJ>;
    using GI = Geo<T, N, TI>;
    using CT = C2<T>;
    static constexpr long NPAIRS = (long)N * N / 2; // row pairs of one field

    void geometry(long &nb_rows, long &nb_j, long &nb_i, int NHP) override
    {
        nb_rows = (NPAIRS + GR::PP - 1) / GR::PP;
        nb_j = (long)N * ((NHP + GJ::B - 1) / GJ::B);
        nb_i = ((long)N * NHP + GI::B - 1) / GI::B;
    }
    long two_pass_partials() override
    {
        if constexpr (N == 64 || N == 128 || N == 256)
            return std::max<long>(8 * SJ<T, N>::R_BLOCKS, 3 * SJ<T, N>::C_BLOCKS);
        else
            return 0;
    }

    // ---------------------------------------------------------- launchers
    // nm > 1: nm batch members (blockIdx.y), member m at in + m * in_bs, out + m * out_bs
    template <bool IN_SPEC, bool OUT_SPEC, class Op, class Pro = ProNone>
    static void rows(Ctx &c, const void *in, void *out, Op op, double *part, const Ctl *ctl = nullptr,
                     Pro pro = Pro{}, int nm = 1, long in_bs = 0, long out_bs = 0)
    {
        const long nb = (NPAIRS + GR::PP - 1) / GR::PP;
        hipLaunchKernelGGL((k_rows<T, N, TR, IN_SPEC, OUT_SPEC, Op, Pro>), dim3((unsigned)nb, (unsigned)nm),
                           dim3(GR::ROW_THREADS), 0, c.stream, in, (CT *)out, c.NHP, NPAIRS, (const CT *)c.tw, op,
                           part, ctl, pro, in_bs, out_bs);
    }
    // pencils along j inside each i plane; row length nhp, valid nh
    template <bool FWD, bool INV, bool STORE, class SOp, class Pro = ProNone>
    static void cols_j(Ctx &c, const void *src, void *dst, SOp op, double *part, const Ctl *ctl = nullptr,
                       int nh = -1, int nhp = -1, Pro pro = Pro{}, int nm = 1, long bs = 0)
    {
        if (nh < 0) { nh = c.NH; nhp = c.NHP; }
        const long nb = (long)N * ((nhp + GJ::B - 1) / GJ::B);
        hipLaunchKernelGGL((k_cols<T, N, TJ, FWD, INV, STORE, SOp, Pro>), dim3((unsigned)nb, (unsigned)nm),
                           dim3(GJ::COL_THREADS), 0, c.stream, (const CT *)src, (CT *)dst, (long)nhp, (long)nhp,
                           (long)N * nhp, nh, nhp, 0, (const CT *)c.tw, op, part, ctl, pro, bs);
    }
    // pencils along i over the whole array
    template <bool FWD, bool INV, bool STORE, class SOp>
    static void cols_i(Ctx &c, const void *src, void *dst, SOp op, double *part, const Ctl *ctl = nullptr,
                       int nh = -1, int nhp = -1, int nm = 1, long bs = 0)
    {
        if (nh < 0) { nh = c.NH; nhp = c.NHP; }
        const long ncols = (long)N * nhp;
        const long nb = (ncols + GI::B - 1) / GI::B;
        hipLaunchKernelGGL((k_cols<T, N, TI, FWD, INV, STORE, SOp, ProNone>), dim3((unsigned)nb, (unsigned)nm),
                           dim3(GI::COL_THREADS), 0, c.stream, (const CT *)src, (CT *)dst, ncols, ncols, 0L, nh, nhp,
                           1, (const CT *)c.tw, op, part, ctl, ProNone{}, bs);
    }

    // two-pass gradient-flow step (N = 64, 128, 256): pass 1 over rows, pass 2 over columns
    using SJT = SJ<T, N>;
    template <bool IN_SPEC, bool OUT_SPEC, class Op, class Pro = ProNone>
    static void sj_rows(Ctx &c, const void *in, void *out, Op op, double *part, const Ctl *ctl, Pro pro = Pro{},
                        int nm = 1, long in_bs = 0, long out_bs = 0)
    {
        hipLaunchKernelGGL((k_sj_rows<T, N, IN_SPEC, OUT_SPEC, Op, Pro>),
                           dim3((unsigned)SJT::R_BLOCKS, (unsigned)nm), dim3(SJT::R_THREADS), 0, c.stream, in, out,
                           (const CT *)c.tw, op, part, ctl, pro, in_bs, out_bs);
    }
    template <bool FWD, bool INV, class SOp, class Pro = ProNone>
    static void sj_cols(Ctx &c, const void *src, void *dst, SOp op, double *part, const Ctl *ctl, Pro pro = Pro{},
                        int nm = 1, long bs = 0)
    {
        hipLaunchKernelGGL((k_sj_cols<T, N, FWD, INV, SOp, Pro>), dim3((unsigned)SJT::C_BLOCKS, (unsigned)nm),
                           dim3(SJT::C_THREADS), 0, c.stream, (const CT *)src, (CT *)dst, (const CT *)c.tw, op, part,
                           ctl, pro, bs);
    }

    // A transform round trip r2c -> spectral op -> c2r as pass 1, pass 2, pass 1
    // (N = 64, 128, 256 in two-pass mode): six array passes instead of ten.  opin runs on
    // the input field (its sums into pin), opout on the output (sums into pout);
    // mid (host) runs between the first pass and the second (scalar reductions).
    template <class OpIn, class SOp, class OpOut, class Mid>
    static void sj_round_trip(Ctx &c, const void *in, OpIn opin, double *pin, Mid mid, SOp sop, OpOut opout,
                              double *pout)
    {
        sj_rows<false, true>(c, in, c.W, opin, pin, nullptr);
        mid();
        sj_cols<true, true>(c, c.W, c.W, sop, nullptr, nullptr);
        sj_rows<true, false>(c, c.W, nullptr, opout, pout, nullptr);
    }
    // The same with the input pass's scalar (its partials summed into *out, scaled
    // by 1/R) formed by block 0 of the second pass instead of a launch of its own.
    template <class OpIn, class SOp, class OpOut>
    static void sj_round_trip_sum(Ctx &c, const void *in, OpIn opin, double *out, SOp sop, OpOut opout, double *pout)
    {
        sj_rows<false, true>(c, in, c.W, opin, c.slot(0), nullptr);
        sj_cols<true, true>(c, c.W, c.W, sop, nullptr, nullptr, ProSum{c.slot(0), SJT::R_BLOCKS, 1.0 / (double)c.R, out});
        sj_rows<true, false>(c, c.W, nullptr, opout, pout, nullptr);
    }
    static bool two_pass(const Ctx &c) { return (N == 64 || N == 128 || N == 256) && c.gf2; }

    // ---------------------------------------------------------- transforms
    void r2c(Ctx &c, const void *x, void *S) override
    {
        rows<false, true>(c, x, S, OpNone{}, nullptr);
        cols_j<true, false, true>(c, S, S, SOpNone{}, nullptr);
        cols_i<true, false, true>(c, S, S, SOpScale<T>{1.0 / (double)c.R}, nullptr);
    }

    void c2r(Ctx &c, const void *S, void *x) override { c2r_ws(c, S, x, c.W); }

    static void c2r_ws(Ctx &c, const void *S, void *x, void *ws)
    {
        cols_i<false, true, true>(c, S, ws, SOpNone{}, nullptr);
        cols_j<false, true, true>(c, ws, ws, SOpNone{}, nullptr);
        rows<true, false>(c, ws, nullptr, OpStore<T>{(T *)x}, nullptr);
    }

    void c2c(Ctx &c, const void *in, void *out, int dir) override
    {
        const long nrows = (long)N * N;
        const long nb = (nrows + GR::PP - 1) / GR::PP;
        if (dir < 0) {
            hipLaunchKernelGGL((k_rows_c2c<T, N, TR, -1>), dim3((unsigned)nb), dim3(GR::ROW_THREADS), 0, c.stream,
                               (const CT *)in, (CT *)out, nrows, 1.0, (const CT *)c.tw);
            cols_j<true, false, true>(c, out, out, SOpNone{}, nullptr, nullptr, N, N);
            cols_i<true, false, true>(c, out, out, SOpScale<T>{1.0 / (double)c.R}, nullptr, nullptr, N, N);
        } else {
            hipLaunchKernelGGL((k_rows_c2c<T, N, TR, +1>), dim3((unsigned)nb), dim3(GR::ROW_THREADS), 0, c.stream,
                               (const CT *)in, (CT *)out, nrows, 1.0, (const CT *)c.tw);
            cols_j<false, true, true>(c, out, out, SOpNone{}, nullptr, nullptr, N, N);
            cols_i<false, true, true>(c, out, out, SOpNone{}, nullptr, nullptr, N, N);
        }
    }

    // ---------------------------------------------------------- operators
    // get_Ene (LPsrc.cpp:8-38) of a real field, without its two FFTs: the
    // local term is the mean of the density, the non-local term the weighted
    // spectral sum of G |phi_hat|^2.
    void energy(Ctx &c, const void *phi) override
    {
        rows<false, true>(c, phi, c.W, OpDensity<T>{c.Tau, c.Gamma}, c.slot(0));
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, false, false>(c, c.W, c.W, SOpEnergy<T, N>{1.0 / (double)c.R, c.Gtab}, c.slot(1));
        launch_reduce(c, c.slot(0), c.nb_rows, 1, 0u, 1.0 / (double)c.R, c.scal + SC_E2);
        launch_reduce(c, c.slot(1), c.nb_i, 1, 0u, 1.0, c.scal + SC_E1);
    }

    // ngrad_cam (LPsrc.cpp:119-173)
    void ngrad(Ctx &c, const void *phi, void *grad) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip_sum(c, phi, OpNgradIn<T>{c.Tau, c.Gamma}, c.scal + SC_MEAN_F,
                                  SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab},
                                  OpNgradOut<T>{c.Tau, c.Gamma, (const T *)phi, c.scal + SC_MEAN_F, (T *)grad},
                                  c.slot(1));
                launch_reduce(c, c.slot(1), SJT::R_BLOCKS, 1, 0u, 1.0 / (double)c.R, c.scal + SC_GN2);
                return;
            }
        }
        rows<false, true>(c, phi, c.W, OpNgradIn<T>{c.Tau, c.Gamma}, c.slot(0));
        launch_reduce(c, c.slot(0), c.nb_rows, 1, 0u, 1.0 / (double)c.R, c.scal + SC_MEAN_F);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr,
                          OpNgradOut<T>{c.Tau, c.Gamma, (const T *)phi, c.scal + SC_MEAN_F, (T *)grad}, c.slot(1));
        launch_reduce(c, c.slot(1), c.nb_rows, 1, 0u, 1.0 / (double)c.R, c.scal + SC_GN2);
    }

    // hv_cam (LOBPCG.cpp:113-166)
    void hv(Ctx &c, const void *phi, const void *v, void *hvo) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip_sum(c, v, OpHvIn<T>{c.Tau, c.Gamma, (const T *)phi}, c.scal + SC_MEAN_CV,
                                  SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab},
                                  OpHvOut<T>{c.Tau, c.Gamma, (const T *)phi, (const T *)v, c.scal + SC_MEAN_CV, (T *)hvo},
                                  nullptr);
                return;
            }
        }
        rows<false, true>(c, v, c.W, OpHvIn<T>{c.Tau, c.Gamma, (const T *)phi}, c.slot(0));
        launch_reduce(c, c.slot(0), c.nb_rows, 1, 0u, 1.0 / (double)c.R, c.scal + SC_MEAN_CV);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr,
                          OpHvOut<T>{c.Tau, c.Gamma, (const T *)phi, (const T *)v, c.scal + SC_MEAN_CV, (T *)hvo},
                          nullptr);
    }

    void hv_dot(Ctx &c, const void *phi, const void *v, void *hvo, double *dot) override
    {
        const OpHvOut<T, true> out{c.Tau, c.Gamma, (const T *)phi, (const T *)v, c.scal + SC_MEAN_CV, (T *)hvo};
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip_sum(c, v, OpHvIn<T>{c.Tau, c.Gamma, (const T *)phi}, c.scal + SC_MEAN_CV,
                                  SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab}, out, c.slot(1));
                launch_reduce(c, c.slot(1), SJT::R_BLOCKS, 1, 0u, 1.0 / (double)c.R, dot);
                return;
            }
        }
        rows<false, true>(c, v, c.W, OpHvIn<T>{c.Tau, c.Gamma, (const T *)phi}, c.slot(0));
        launch_reduce(c, c.slot(0), c.nb_rows, 1, 0u, 1.0 / (double)c.R, c.scal + SC_MEAN_CV);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, out, c.slot(1));
        launch_reduce(c, c.slot(1), c.nb_rows, 1, 0u, 1.0 / (double)c.R, dot);
    }

    // precond_camnew (LOBPCG.cpp:168-199): w <- IFFT(FFT(w)/N^3 / (G+|Tau|+0.02))
    // minus its mean (= the same with the DC set to zero).
    template <class OpIn> void precond_in(Ctx &c, const void *in, OpIn opin, void *w)
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip(c, in, opin, nullptr, [] {}, SOpMulTab<T, N>{1.0 / (double)c.R, c.Itab},
                              OpStore<T>{(T *)w}, nullptr);
                return;
            }
        }
        rows<false, true>(c, in, c.W, opin, nullptr);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{1.0 / (double)c.R, c.Itab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, OpStore<T>{(T *)w}, nullptr);
    }
    void precond(Ctx &c, void *w) override { precond_in(c, w, OpNone{}, w); }
    void precond_resid(Ctx &c, const void *u, const void *v, double alpha, void *w) override
    {
        precond_in(c, u, OpResid<T>{alpha, (const T *)v}, w);
    }

    bool hv_many(Ctx &c, const void *phi, const void *v0, void *out0, int n, void *spec, double *means,
                 double *dots) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c) && n >= 1 && n <= BATCH_OPS) {
                const double sc = 1.0 / (double)c.R;
                const long fs = c.R;
                sj_rows<false, true>(c, v0, spec, OpHvIn<T>{c.Tau, c.Gamma, (const T *)phi}, c.slot(0), nullptr,
                                     ProNone{}, n, fs, c.C);
                sj_cols<true, true>(c, spec, spec, SOpMulTab<T, N>{sc, c.Gtab}, nullptr, nullptr,
                                    ProSum{c.slot(0), SJT::R_BLOCKS, sc, means}, n, c.C);
                if (dots) {
                    OpHvOutM<T, true> out{};
                    out.Tau = c.Tau;
                    out.Gamma = c.Gamma;
                    out.phi = (const T *)phi;
                    out.v = (const T *)v0;
                    out.hv = (T *)out0;
                    out.fs = fs;
                    out.means = means;
                    sj_rows<true, false>(c, spec, nullptr, out, c.slot(1), nullptr, ProNone{}, n, c.C, 0);
                    launch_reduce(c, c.slot(1), SJT::R_BLOCKS, n, 0u, sc, dots);
                } else {
                    OpHvOutM<T, false> out{};
                    out.Tau = c.Tau;
                    out.Gamma = c.Gamma;
                    out.phi = (const T *)phi;
                    out.v = (const T *)v0;
                    out.hv = (T *)out0;
                    out.fs = fs;
                    out.means = means;
                    sj_rows<true, false>(c, spec, nullptr, out, nullptr, nullptr, ProNone{}, n, c.C, 0);
                }
                return true;
            }
        }
        return false;
    }

    bool precond_resid_many(Ctx &c, const void *u0, const void *v0, const double *alphas, void *w0, int n,
                            void *spec) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c) && n >= 1 && n <= BATCH_OPS) {
                const long fs = c.R;
                OpResidM<T> in{};
                in.v = (const T *)v0;
                in.fs = fs;
                in.alphas = alphas;
                OpStoreM<T> out{};
                out.out = (T *)w0;
                out.fs = fs;
                sj_rows<false, true>(c, u0, spec, in, nullptr, nullptr, ProNone{}, n, fs, c.C);
                sj_cols<true, true>(c, spec, spec, SOpMulTab<T, N>{1.0 / (double)c.R, c.Itab}, nullptr, nullptr,
                                    ProNone{}, n, c.C);
                sj_rows<true, false>(c, spec, nullptr, out, nullptr, nullptr, ProNone{}, n, c.C, 0);
                return true;
            }
        }
        return false;
    }

    void gconv(Ctx &c, const void *in, void *out) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip(c, in, OpNone{}, nullptr, [] {}, SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab},
                              OpStore<T>{(T *)out}, nullptr);
                return;
            }
        }
        rows<false, true>(c, in, c.W, OpNone{}, nullptr);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{1.0 / (double)c.R, c.Gtab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, OpStore<T>{(T *)out}, nullptr);
    }

    void force(Ctx &c, void *x, const void *v, const double *s, double coef, ForceSpec *fs) override
    {
        force_any(c, x, v, s, coef, fs, false);
    }
    long spm_tangent(Ctx &c, const void *p0, const void *p1, void *d, double *pt, double *pf) override
    {
        const OpTangentIn<T> in{(const T *)p1, (T *)d};
        const OpFvOut<T> out{c.Tau, c.Gamma, (const T *)p0, (const T *)p1, (const T *)d};
        const SOpMulTab<T, N> g{1.0 / (double)c.R, c.Gtab};
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip(c, p0, in, pt, [] {}, g, out, pf);
                return SJT::R_BLOCKS;
            }
        }
        rows<false, true>(c, p0, c.W, in, pt);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, g, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, out, pf);
        return c.nb_rows;
    }

    void hisd_step(Ctx &c, void *x, const void *tmpp, double bta) override
    {
        const OpHisdIn<T> in{c.Tau, c.Gamma, bta, (const T *)tmpp};
        const SOpHisdStep<T, N> st{1.0 / (double)c.R, bta, c.Gtab};
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_round_trip(c, x, in, nullptr, [] {}, st, OpStore<T>{(T *)x}, nullptr);
                return;
            }
        }
        rows<false, true>(c, x, c.W, in, nullptr);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, st, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, OpStore<T>{(T *)x}, nullptr);
    }

    void force_ngrad(Ctx &c, void *x, const void *v, const double *s, double coef, ForceSpec *fs) override
    {
        force_any(c, x, v, s, coef, fs, true);
    }

    // One SPM force step on x (and, with ngrad, ngrad_cam's norm at the result:
    // the output pass of the force round trip is also the input pass of
    // ngrad's, one pass over x fewer).  sop is the force's spectral step.
    template <class SOp>
    void force_with(Ctx &c, void *x, const void *v, const double *s, double coef, SOp sop, double *sp, bool ngrad,
                    double *gn2)
    {
        const double sc = 1.0 / (double)c.R;
        const OpForceIn<T> fin{c.Tau, c.Gamma, coef, (const T *)v, s};
        const OpStoreNgradIn<T> mid{c.Tau, c.Gamma, (T *)x};
        const OpNgradOut<T> nout{c.Tau, c.Gamma, (const T *)x, c.scal + SC_MEAN_F, nullptr};
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (two_pass(c)) {
                sj_rows<false, true>(c, x, c.W, fin, nullptr, nullptr);
                sj_cols<true, true>(c, c.W, c.W, sop, sp, nullptr);
                if (!ngrad) {
                    sj_rows<true, false>(c, c.W, nullptr, OpStore<T>{(T *)x}, nullptr, nullptr);
                    return;
                }
                sj_rows<true, true>(c, c.W, c.W, mid, c.slot(0), nullptr);
                sj_cols<true, true>(c, c.W, c.W, SOpMulTab<T, N>{sc, c.Gtab}, nullptr, nullptr,
                                    ProSum{c.slot(0), SJT::R_BLOCKS, sc, c.scal + SC_MEAN_F});
                sj_rows<true, false>(c, c.W, nullptr, nout, c.slot(1), nullptr);
                launch_reduce(c, c.slot(1), SJT::R_BLOCKS, 1, 0u, sc, gn2);
                return;
            }
        }
        rows<false, true>(c, x, c.W, fin, nullptr);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, sop, sp);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        if (!ngrad) {
            rows<true, false>(c, c.W, nullptr, OpStore<T>{(T *)x}, nullptr);
            return;
        }
        rows<true, true>(c, c.W, c.W, mid, c.slot(0));
        launch_reduce(c, c.slot(0), c.nb_rows, 1, 0u, sc, c.scal + SC_MEAN_F);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr);
        cols_i<true, true, true>(c, c.W, c.W, SOpMulTab<T, N>{sc, c.Gtab}, nullptr);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr);
        rows<true, false>(c, c.W, nullptr, nout, c.slot(1));
        launch_reduce(c, c.slot(1), c.nb_rows, 1, 0u, sc, gn2);
    }

    void force_any(Ctx &c, void *x, const void *v, const double *s, double coef, ForceSpec *fs, bool ngrad)
    {
        const double sc = 1.0 / (double)c.R;
        double *gn2 = fs && fs->gn2 ? fs->gn2 : c.scal + SC_GN2;
        if (!fs) return force_with(c, x, v, s, coef, SOpMulTab<T, N>{sc, c.Btab}, nullptr, ngrad, gn2);
        if (!fs->other)
            return force_with(c, x, v, s, coef, SOpForceSpec<T, N, false>{sc, c.Btab, c.Gtab, (CT *)fs->keep, nullptr},
                              nullptr, ngrad, gn2);
        fs->nb = ((long)N * c.NHP + GI::B - 1) / GI::B; // blocks of the i pass
        if constexpr (N == 64 || N == 128 || N == 256)
            if (two_pass(c)) fs->nb = SJT::C_BLOCKS;
        force_with(c, x, v, s, coef, SOpForceSpec<T, N, true>{sc, c.Btab, c.Gtab, nullptr, (const CT *)fs->other},
                   fs->part, ngrad, gn2);
    }

    // ---------------------------------------------------------- gradient flow
    // State: U = u_F (normalised spectrum of phi).  W carries the spectrum of
    // f(phi) between passes (both in the two-pass layouts when c.gf2).  Slot 2
    // holds the E2 sums of the current phi (row pass), slot 3 the E1 /
    // ||grad F||^2 / residual sums (i pass).
    void gf_init(Ctx &c, const void *phi) override { gf_init_member(c, 0, phi, c.Tau, c.Gamma); }

    // Member m of a batch (m = 0 for a single field): its U and W are the m-th
    // spectra, its partial sums the m-th share of each slot, its control block ctl[m].
    void gf_init_member(Ctx &c, int m, const void *phi, double Tau, double Gamma) override
    {
        CT *Um = (CT *)c.U + (long)m * c.C, *Wm = (CT *)c.W + (long)m * c.C;
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (c.gf2) { // U and W in the two-pass layout; slot 2 holds the density partials of phi
                using S = SJT;
                sj_rows<false, true>(c, phi, Wm, OpNone{}, nullptr, c.ctl + m);
                sj_cols<true, false>(c, Wm, Um, SOpScale<T>{1.0 / (double)c.R}, nullptr, c.ctl + m);
                sj_rows<false, true>(c, phi, Wm, OpGradFlowRow<T>{Tau, Gamma, nullptr, nullptr, 0},
                                     c.slot(2) + (long)m * S::R_BLOCKS, c.ctl + m);
                return;
            }
        }
        r2c(c, phi, Um);
        rows<false, true>(c, phi, Wm, OpGradFlowRow<T>{Tau, Gamma, nullptr, nullptr, 0}, c.slot(2) + (long)m * c.nb_rows,
                          c.ctl + m);
        cols_j<true, false, true>(c, Wm, Wm, SOpNone{}, nullptr, c.ctl + m);
    }

    // phi of member m from its state spectrum (its W is the scratch)
    void gf_read_member(Ctx &c, int m, void *phi) override
    {
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (c.gf2) {
                CT *Um = (CT *)c.U + (long)m * c.C, *Wm = (CT *)c.W + (long)m * c.C;
                sj_cols<false, true>(c, Um, Wm, SOpNone{}, nullptr, nullptr);
                sj_rows<true, false>(c, Wm, nullptr, OpStore<T>{(T *)phi}, nullptr, nullptr);
                return;
            }
        }
        c2r_ws(c, (const CT *)c.U + (long)m * c.C, phi, (CT *)c.W + (long)m * c.C);
    }

    // One get_min_u_F iteration (LPsrc.cpp:52-111).  Kernels return at once
    // when ctl->stop is set, so a fixed sequence of steps can be queued (or
    // replayed from a graph) past convergence.
    // Two passes (c.gf2, N = 64, 128, 256; gf_split.h): pass 2 (j stage B, i FFT,
    // update, inverse) then pass 1 (A^-1, c2r, f(phi), r2c, A).  Block 0 of pass 2
    // sums the density partials of the previous pass 1 into slot 4; block 0 of
    // pass 1 closes the iteration from the pass-2 partials (slot 3).
    // Four passes otherwise: i pass, j inverse (reduces), rows (closes), j forward.
    // Batched (c.nmem members): every launch covers all members, each with its
    // own (Tau, Gamma) from c.prm; members whose ctl[m].stop is set skip out.
    void gf_step(Ctx &c) override
    {
        const int kf = (int)std::min<long>(64, c.nb_j);
        const int nm = c.nmem;
        const long bs = c.C;
        // U with non-temporal hints when W + U overflow the 256 MB last-level cache,
        // unless the step variant fixes the hint
        const int um = (c.sjv & SPM_STEP_U_MASK) >> 1;
        const bool nt = um == 0 ? 2.0 * (double)bs * sizeof(CT) * nm > 200e6 : um == 2;
        if constexpr (N == 64 || N == 128 || N == 256) {
            if (c.gf2) {
                using S = SJT;
                const ProSJDensity pd{c.slot(2), S::R_BLOCKS, c.slot(4)};
                auto pass2 = [&](auto op) { sj_cols<true, true>(c, c.W, c.W, op, c.slot(3), c.ctl, pd, nm, bs); };
                const double sc = 1.0 / (double)c.R;
                const CT *GD = (const CT *)c.GDtab;
                CT *U = (CT *)c.U;
                // FP32 takes the hint as a template constant, FP64 as a flag (one kernel)
                const bool nt32 = sizeof(T) == 4 && nt;
                if (c.sjv & SPM_STEP_GATHER_U) {
                    if (nt32)
                        pass2(SOpGradFlowStep<T, N, false, sizeof(T) == 4>{sc, c.dt, GD, U, bs, nt});
                    else
                        pass2(SOpGradFlowStep<T, N, false, false>{sc, c.dt, GD, U, bs, nt});
                } else {
                    if (nt32)
                        pass2(SOpGradFlowStep<T, N, true, sizeof(T) == 4>{sc, c.dt, GD, U, bs, nt});
                    else
                        pass2(SOpGradFlowStep<T, N, true, false>{sc, c.dt, GD, U, bs, nt});
                }
                sj_rows<true, true>(c, c.W, c.W, OpGradFlowRow<T>{c.Tau, c.Gamma, nullptr, c.prm, 0}, c.slot(2), c.ctl,
                                    ProSJFinalize{c.slot(3), S::C_BLOCKS, c.slot(4), (double)c.R, c.ctl, c.hist,
                                                  3 * c.hist_cap},
                                    nm, bs, bs);
                return;
            }
        }
        if (sizeof(T) == 4 && nt)
            cols_i<true, true, true>(c, c.W, c.W,
                                     SOpGradFlowStep<T, N, true, sizeof(T) == 4>{1.0 / (double)c.R, c.dt,
                                                                                 (const CT *)c.GDtab, (CT *)c.U, bs, nt},
                                     c.slot(3), c.ctl, -1, -1, nm, bs);
        else
            cols_i<true, true, true>(c, c.W, c.W,
                                     SOpGradFlowStep<T, N, true, false>{1.0 / (double)c.R, c.dt, (const CT *)c.GDtab,
                                                                        (CT *)c.U, bs, nt},
                                     c.slot(3), c.ctl, -1, -1, nm, bs);
        cols_j<false, true, true>(c, c.W, c.W, SOpNone{}, nullptr, c.ctl, -1, -1,
                                  ProGradFlowReduce{c.slot(3), c.nb_i, c.slot(2), c.nb_rows, c.slot(4), kf}, nm, bs);
        rows<true, true>(c, c.W, c.W, OpGradFlowRow<T>{c.Tau, c.Gamma, nullptr, c.prm, 0}, c.slot(2), c.ctl,
                         ProGradFlowFinalize{c.slot(4), kf, (double)c.R, c.ctl, c.hist, 3 * c.hist_cap}, nm, bs, bs);
        cols_j<true, false, true>(c, c.W, c.W, SOpNone{}, nullptr, c.ctl, -1, -1, ProNone{}, nm, bs);
    }
};

} // namespace spm
