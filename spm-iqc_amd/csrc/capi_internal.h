// capi_internal.h -- the context handle behind the C ABI and the helpers the
// entry-point translation units share.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/spm_iqc.h"
#include "spm_ctx.h"

struct spm_ctx {
    spm::Ctx c;
    hipStream_t own = nullptr;
    double *h_pin = nullptr; // pinned host scratch for scalars read back
    size_t dev_bytes = 0;
    // solver workspace (LOBPCG / SPM fields), grown on demand
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // block-operation scratch (Gram partials, coefficients)
    double *bscr = nullptr;
    size_t bscr_bytes = 0;
    // spectra of the operators applied to several fields in one launch, grown on demand
    void *specws = nullptr;
    size_t specws_bytes = 0;
    // gradient-flow steps replayed from a graph captured on `own` (valid while
    // the history buffer it writes stays where it was at capture)
    hipGraphExec_t gf_exec = nullptr;
    const void *gf_exec_hist = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr; // ordering with the caller's stream
};

// Enqueue n gradient-flow iterations: whole graphs of GF_GRAPH_STEPS on the
// context's own stream (ordered after and before the caller's stream by
// events), the remainder as single launches.
constexpr int GF_GRAPH_STEPS = 16;
int spm_gf_enqueue(spm_ctx *h, long n);

int spm_fail(int code, const std::string &msg);
// host tables of a context: twiddles exp(-2 pi i e / N), e < N, in FP64 and FP32
// (interleaved re, im), and k^2 of the table entry |n|^2 = n2 (get_ksquare_local,
// Initialization.cpp:132-148, with k_i = kscale n_i / L)
void spm_host_twiddles(int N, std::vector<double> &twd, std::vector<float> &twf);
double spm_host_k2(double kscale, double L, int n2);
// the checks spm_create makes of its parameters (no device work); who names the caller
int spm_check_params(const spm_params *p, const char *who);
// context of nmem batch members (1: a single field) with hist_cap iterations of history each
int spm_ctx_create(const spm_params *p, int nmem, long hist_cap, spm_ctx **out);
void spm_ctx_free(spm_ctx *h);
// control block of member m: iteration 0, the given limits (max_iter <= 0: stopped)
void spm_ctl_init(spm_ctx *h, int m, int max_iter, double res_tol, double gn_tol);
// shape recorded by the tuner for (p->N, p->precision, nmem), tuning on a miss
int spm_resolve_shape(const spm_params *p, int nmem, int *shape);
// two-pass step variant recorded by the tuner for (p->N, p->precision, nmem), tuning on a miss
int spm_resolve_step(const spm_params *p, int nmem, int *variant);
int spm_dev_alloc(spm_ctx *h, void **p, size_t bytes);
int spm_launch_status(const char *what);
int spm_read_scalars(spm_ctx *h, const double *dev, int n, double *out);

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return spm_fail(SPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

#define CHECK_CTX(h)                                                                          \
    if (!(h)) return spm_fail(SPM_EINVAL, "null context");                                    \
    spm::Ctx &c = (h)->c;                                                                     \
    (void)c

#define RCHK(expr)                                                                            \
    do {                                                                                      \
        int _rc = (expr);                                                                     \
        if (_rc) return _rc;                                                                  \
    } while (0)
