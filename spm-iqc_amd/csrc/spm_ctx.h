// spm_ctx.h -- device-resident context of one N^3 grid (host-side C++).
//
// The reference keeps its state in process globals (Data.h:4-28, allocated by
// memAllocation_local, Initialization.cpp:44-78).  Here one context owns the
// device buffers for one grid size, model parameters and precision; every
// operation runs on the context's HIP stream.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "grid_ops.h"
#include "spm_iqc.h"

namespace spm {

struct GridIface;

enum ScalarSlot { // device scalars (ctx.scal)
    SC_MEAN_F = 0, SC_GN2, SC_MEAN_CV, SC_E1, SC_E2, SC_TMP0, SC_TMP1, SC_FV, SC_TV2, SC_RES0, SC_RES1, SC_FV1, SC_DREL,
    SC_COUNT = 64
};

struct Ctx {
    // grid and model (LPsystParameter, Initialization.cpp:7-42)
    int N = 0, NH = 0, NHP = 0, prec = 0; // prec 0 = FP64, 1 = FP32
    long R = 0, C = 0;                    // real points, half-spectrum entries (N*N*NHP)
    size_t esz = 8;                       // sizeof(real)
    double L = 26, kscale = 1, Tau = -0.0061, Gamma = 0.30, dt = 0.3;
    hipStream_t stream = nullptr;
    GridIface *g = nullptr;
    // tables
    void *tw = nullptr;      // exp(-2 pi i e / N), e < N, in the context precision
    double *Gtab = nullptr;  // get_GK(kscale^2 n2 / L^2), n2 = |n|^2 = 0..3(N/2)^2
    double *Itab = nullptr;  // 1 / (G + |Tau| + 0.02)  (get_ikt2_local)
    double *Btab = nullptr;  // 1 / (1 + G): the SPM force denominator with beta = 1 (SPM.cpp:84-85)
    void *GDtab = nullptr;   // (G, 1 / (1 + dt G)) pairs in the context precision (LPsrc.cpp:67-68)
    int ntab = 0;
    // work buffers
    void *W = nullptr;       // spectrum-sized work array
    void *U = nullptr;       // gradient-flow state spectrum
    double *part = nullptr;  // block partial sums: nslots slots of part_slot doubles
    long part_slot = 0;
    int nslots = 0;
    double *scal = nullptr;  // SC_COUNT device scalars
    Ctl *ctl = nullptr;
    unsigned *counter = nullptr; // arrival counters of last-block reductions (kept at 0 between launches)
    double *hist = nullptr;  // 3 doubles per iteration: E, ||grad F||, res
    long hist_cap = 0;
    // block counts of the passes over one field
    long nb_rows = 0, nb_j = 0, nb_i = 0;
    // batches of independent gradient flows: nmem members, member m's U and W
    // at U + m C, W + m C, (Tau, Gamma) at prm[2m] (nullptr: the context's),
    // control block ctl[m], history hist + 3 m hist_cap
    int nmem = 1;
    double *prm = nullptr;
    int shape = 0; // FFT pass shape (index into grid_shapes(N))
    // two-pass step variant (SPM_STEP_* flags without SPM_STEP_EXPLICIT): bit 0
    // gathers U in pass 2 where it is used, bits 1-2 the cache hint of U
    int sjv = 0;
    int gf2 = 0;   // N = 256: gradient-flow step in two passes (U and W in the two-pass layout)

    double *slot(int s) { return part + (long)s * part_slot; }
};

// Transform and operator passes specialised for (precision, N).
// Spring_Pairing's force pair (SOpForceSpec): the first replica's force stores
// the spectrum of its result in keep; the second reads it (other) and writes the
// partial sums of Q_00, Q_01, Q_11 into part, nb blocks of 3 (set by the call).
struct ForceSpec {
    void *keep = nullptr;
    const void *other = nullptr;
    double *part = nullptr;
    long nb = 0;
    double *gn2 = nullptr; // force_ngrad: where the norm^2 goes (default scal[SC_GN2])
};

struct GridIface {
    virtual ~GridIface() = default;
    virtual void geometry(long &nb_rows, long &nb_j, long &nb_i, int NHP) = 0;
    // partial sums one field writes in a two-pass launch (0: no two-pass kernels for this N)
    virtual long two_pass_partials() = 0;
    // S = FFT(x) / N^3 into the half spectrum S
    virtual void r2c(Ctx &c, const void *x, void *S) = 0;
    // x = IFFT(S) (unnormalised); S is preserved (W is used)
    virtual void c2r(Ctx &c, const void *S, void *x) = 0;
    // full complex N^3 arrays (the reference's fftw_complex layout):
    // dir < 0: out = FFT(in) / N^3; dir > 0: out = IFFT(in); in may equal out
    virtual void c2c(Ctx &c, const void *in, void *out, int dir) = 0;
    // E1 (before the 1/2) and E2 sums of get_Ene into scal[SC_E1], scal[SC_E2]
    virtual void energy(Ctx &c, const void *phi) = 0;
    virtual void ngrad(Ctx &c, const void *phi, void *grad) = 0; // scal[SC_GN2] = nrmp^2
    virtual void hv(Ctx &c, const void *phi, const void *v, void *hv) = 0;
    // hv, and *dot (device) = inpfp(v, hv)
    virtual void hv_dot(Ctx &c, const void *phi, const void *v, void *hv, double *dot) = 0;
    virtual void precond(Ctx &c, void *w) = 0;
    // w = precond(u - alpha v)  (w may not alias u or v)
    virtual void precond_resid(Ctx &c, const void *u, const void *v, double alpha, void *w) = 0;
    // hv_cam / precond_camnew of n fields at once (n <= BATCH_OPS fields at a stride of
    // R values from v0, u0, ...), each a member of one launch per pass; spec: n
    // spectra of scratch; means: n device doubles; dots (may be NULL): <v_i, H v_i>;
    // alphas: n device doubles.  False where the operators have no batched form
    // (the axis-by-axis passes): the caller then goes field by field.
    static constexpr int BATCH_OPS = 8;
    virtual bool hv_many(Ctx &c, const void *phi, const void *v0, void *out0, int n, void *spec, double *means,
                         double *dots) = 0;
    virtual bool precond_resid_many(Ctx &c, const void *u0, const void *v0, const double *alphas, void *w0, int n,
                                    void *spec) = 0;
    // out = IFFT(G FFT(in) / N^3) with the DC set to zero (the non-local part of ngrad_cam)
    virtual void gconv(Ctx &c, const void *in, void *out) = 0;
    // x <- IFFT(FFT(x + nln(x) - coef * (*s) * v) / N^3 / (1 + G)), DC = 0
    // (one force step of SPM.cpp:34-99 / 103-168; s a device scalar)
    virtual void force(Ctx &c, void *x, const void *v, const double *s, double coef, ForceSpec *fs = nullptr) = 0;
    // Spring_Pairing's first tangent d = Phi1 - Phi0: partial sums of d^2 into pt
    // (one per block), the <f, d> sums of both replicas (OpFvOut, 7 per block)
    // into pf; returns the block count
    virtual long spm_tangent(Ctx &c, const void *p0, const void *p1, void *d, double *pt, double *pf) = 0;
    // hiosd_lobpcg_sieh's step of x (HISD.cpp:215-268) with the reflected force part tmpp
    virtual void hisd_step(Ctx &c, void *x, const void *tmpp, double bta) = 0;
    // force, then scal[SC_GN2] (or *fs->gn2) = nrmp(ngrad_cam(x))^2 of the updated x (no gradient field)
    virtual void force_ngrad(Ctx &c, void *x, const void *v, const double *s, double coef,
                             ForceSpec *fs = nullptr) = 0;
    virtual void gf_init(Ctx &c, const void *phi) = 0;
    virtual void gf_init_member(Ctx &c, int m, const void *phi, double Tau, double Gamma) = 0;
    virtual void gf_read_member(Ctx &c, int m, void *phi) = 0;
    virtual void gf_step(Ctx &c) = 0;
};

GridIface *make_grid(int N, int prec, int shape);
// FFT pass shapes of N: *s = table of (row, j, i) threads per pencil; returns the count
int grid_shapes(int N, const int (**s)[3]);

// generic helpers (spm_capi.hip)
void launch_reduce(Ctx &c, const double *part, long nblk, int nval, unsigned maxmask, double scale, double *out);

} // namespace spm
