// capi_internal.h -- the context handle behind the C ABI and the helpers the
// entry-point translation units share.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

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
};

int spm_fail(int code, const std::string &msg);
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
