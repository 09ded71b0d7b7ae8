// noalloc.hip -- spm_ngrad_noalloc: ngrad_cam (LPsrc.cpp:119-173) with no
// device allocation, for callers that cannot allocate (code under graph
// capture, tight serving loops).  It runs the kernels spm_ngrad runs, with the
// tables, the spectrum work array, the partial sums and the scalars of the
// call carved out of `scratch`, one static array per GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "capi_internal.h"

using namespace spm;

// The work memory of spm_ngrad_noalloc.  Every GPU the library runs on holds
// its own copy until the process ends; every call on that GPU reuses it, so
// calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[SPM_NOALLOC_SCRATCH_BYTES];

namespace {

struct Region { // one call's carve of the scratch array (byte offsets)
    size_t tw, gtab, scal, part, W, end;
};

struct Device {
    unsigned char *ptr = nullptr; // device address of this GPU's `scratch`
    // what the table regions hold: (N, precision, L, kscale) of the last upload
    int N = 0, prec = -1;
    double L = 0, kscale = 0;
};

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
Device g_dev[MAX_DEVICES];
std::map<std::tuple<int, int, int>, GridIface *> g_grids; // (N, precision, shape) -> kernels

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

// Geometry of a context with these parameters (as spm_ctx_create sets it up)
// and the carve of the scratch array it needs.
void layout(Ctx &c, const spm_params *p, GridIface *g, Region &r)
{
    c.N = p->N;
    c.prec = p->precision;
    c.esz = c.prec == SPM_FP64 ? 8 : 4;
    c.NH = c.N / 2 + 1;
    const int align = c.prec == SPM_FP64 ? 8 : 16;
    c.NHP = (c.NH + align - 1) / align * align;
    c.R = (long)c.N * c.N * c.N;
    c.C = (long)c.N * c.N * c.NHP;
    c.L = p->L;
    c.kscale = p->kscale;
    c.Tau = p->Tau;
    c.Gamma = p->Gamma;
    c.dt = p->dt;
    c.shape = p->shape;
    c.gf2 = p->step_passes == 2 || (p->step_passes == 0 && (c.N == 64 || c.N == 128 || c.N == 256));
    c.nmem = 1;
    c.g = g;
    g->geometry(c.nb_rows, c.nb_j, c.nb_i, c.NHP);
    c.ntab = 3 * (c.N / 2) * (c.N / 2) + 1;
    const long mb = std::max(std::max(c.nb_rows, c.nb_j), std::max(c.nb_i, 1024L));
    c.part_slot = std::max(8 * mb, g->two_pass_partials());
    c.nslots = 2; // ngrad uses slots 0 and 1
    r.tw = 0;
    r.gtab = up256(r.tw + 2 * c.N * c.esz);
    r.scal = up256(r.gtab + sizeof(double) * c.ntab);
    r.part = up256(r.scal + sizeof(double) * SC_COUNT);
    r.W = up256(r.part + sizeof(double) * c.part_slot * c.nslots);
    r.end = up256(r.W + (size_t)c.C * 2 * c.esz);
}

} // namespace

extern "C" int spm_ngrad_noalloc(const spm_params *p, const void *phi, void *grad, double *nrm, void *hip_stream)
{
    if (!p || !phi || !grad) return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: null argument");
    if (phi == grad) return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: phi and grad must differ");
    RCHK(spm_check_params(p, "spm_ngrad_noalloc"));
    if (p->shape < 0) // the tuner would allocate a context to time the shapes
        return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: shape -1 (tuned) is not available; pass shape >= 0");
    const int (*shapes)[3] = nullptr;
    if (p->shape >= grid_shapes(p->N, &shapes))
        return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: no FFT pass shape " + std::to_string(p->shape) + " for N " +
                                        std::to_string(p->N));

    std::lock_guard<std::mutex> lock(g_mu);
    // kernels of (N, precision, shape): host objects, made once per process
    GridIface *&g = g_grids[std::make_tuple(p->N, p->precision, p->shape)];
    if (!g) g = make_grid(p->N, p->precision, p->shape);
    if (!g) return spm_fail(SPM_ENOTSUP, "spm_ngrad_noalloc: no kernels for this N");
    Ctx c;
    Region r;
    layout(c, p, g, r);
    if (r.end > SPM_NOALLOC_SCRATCH_BYTES)
        return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: N " + std::to_string(p->N) + " in this precision needs " +
                                        std::to_string(r.end) + " bytes of scratch, more than the " +
                                        std::to_string((size_t)SPM_NOALLOC_SCRATCH_BYTES) + " it has");

    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return spm_fail(SPM_EINVAL, "spm_ngrad_noalloc: device index out of range");
    Device &d = g_dev[dev];
    if (!d.ptr) { // once per device, with the device current: this GPU's copy of the array
        void *a = nullptr;
        HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        d.ptr = (unsigned char *)a;
    }
    hipStream_t st = (hipStream_t)hip_stream;
    c.stream = st;
    c.tw = d.ptr + r.tw;
    c.Gtab = (double *)(d.ptr + r.gtab);
    c.scal = (double *)(d.ptr + r.scal);
    c.part = (double *)(d.ptr + r.part);
    c.W = d.ptr + r.W;
    if (d.N != c.N || d.prec != c.prec || d.L != c.L || d.kscale != c.kscale) {
        // the tables spm_create builds for these parameters, copied in (synchronously,
        // so that the host copies can go); the record is set only once they are there
        std::vector<double> twd, Gt(c.ntab);
        std::vector<float> twf;
        spm_host_twiddles(c.N, twd, twf);
        for (int n2 = 0; n2 < c.ntab; ++n2) Gt[n2] = spm_get_GK(spm_host_k2(c.kscale, c.L, n2));
        d.N = 0;
        HIPCHK(hipMemcpyAsync(c.tw, c.prec == SPM_FP64 ? (const void *)twd.data() : (const void *)twf.data(),
                              2 * c.N * c.esz, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c.Gtab, Gt.data(), sizeof(double) * c.ntab, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        d.N = c.N;
        d.prec = c.prec;
        d.L = c.L;
        d.kscale = c.kscale;
    }
    g->ngrad(c, phi, grad);
    RCHK(spm_launch_status("spm_ngrad_noalloc"));
    if (nrm) {
        double g2 = 0.0;
        HIPCHK(hipMemcpyAsync(&g2, c.scal + SC_GN2, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *nrm = std::sqrt(g2);
    }
    return SPM_OK;
}
