// batch.hip -- batches of independent gradient flows with members loaded and
// retired while the others run (continuous batching), steps replayed from a
// HIP graph.
//
// A batch is a context of `capacity` members: member m owns the m-th state
// spectrum U and work spectrum W, the m-th share of each partial-sum slot,
// control block ctl[m], (Tau, Gamma) at prm[2m] and history hist + 3 m hist_cap.
// One step is the four passes of the single-field step (get_min_u_F,
// LPsrc.cpp:52-111), each launched once over all members (blockIdx.y);
// members whose ctl[m].stop is set return at once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "capi_internal.h"

using namespace spm;

namespace {
constexpr int GRAPH_STEPS = 16; // iterations per captured graph

__global__ void k_set_prm(double *prm, int m, double Tau, double Gamma)
{
    prm[2 * m] = Tau;
    prm[2 * m + 1] = Gamma;
}
} // namespace

struct spm_batch {
    spm_ctx *h = nullptr;
    int cap = 0;
    hipGraphExec_t gexec = nullptr;
    Ctl *h_ctl = nullptr; // pinned copies of the control blocks
};

static int batch_capture(spm_batch *b)
{
    Ctx &c = b->h->c;
    hipGraph_t g = nullptr;
    // relaxed: other threads of the process may be allocating or launching meanwhile
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
    for (int s = 0; s < GRAPH_STEPS; ++s) c.g->gf_step(c);
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (e != hipSuccess) return spm_fail(SPM_EHIP, std::string("spm_batch_steps: capture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        b->gexec = nullptr;
        return spm_fail(SPM_EHIP, std::string("spm_batch_steps: instantiate: ") + hipGetErrorString(e2));
    }
    return SPM_OK;
}

static int read_ctl(spm_batch *b)
{
    Ctx &c = b->h->c;
    HIPCHK(hipMemcpyAsync(b->h_ctl, c.ctl, sizeof(Ctl) * b->cap, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return SPM_OK;
}

#define CHECK_BATCH(b)                                                                        \
    if (!(b) || !(b)->h) return spm_fail(SPM_EINVAL, "null batch");                           \
    Ctx &c = (b)->h->c;                                                                       \
    (void)c

#define CHECK_SLOT(b, slot, what)                                                             \
    if ((slot) < 0 || (slot) >= (b)->cap)                                                     \
    return spm_fail(SPM_EINVAL, std::string(what) + ": slot " + std::to_string(slot) + " out of range")

extern "C" {

int spm_batch_create(const spm_params *p, int capacity, int hist_cap, spm_batch **out)
{
    if (!p || !out) return spm_fail(SPM_EINVAL, "spm_batch_create: null argument");
    *out = nullptr;
    if (capacity < 1 || capacity > 65535) return spm_fail(SPM_EINVAL, "spm_batch_create: capacity in 1..65535");
    if (hist_cap < 1) return spm_fail(SPM_EINVAL, "spm_batch_create: hist_cap >= 1");
    spm_batch *b = new spm_batch();
    int rc = spm_ctx_create(p, capacity, hist_cap, &b->h);
    if (rc) {
        delete b;
        return rc;
    }
    b->cap = capacity;
    if (hipHostMalloc((void **)&b->h_ctl, sizeof(Ctl) * capacity, hipHostMallocDefault) != hipSuccess) {
        spm_ctx_free(b->h);
        delete b;
        return spm_fail(SPM_ENOMEM, "spm_batch_create: pinned control blocks");
    }
    // the step graph is captured once, here: every member starts stopped, and
    // the graph reads members' state through their control blocks
    if ((rc = batch_capture(b))) {
        spm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return SPM_OK;
}

int spm_batch_destroy(spm_batch *b)
{
    if (!b) return SPM_OK;
    if (b->h) (void)hipStreamSynchronize(b->h->c.stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    spm_ctx_free(b->h);
    delete b;
    return SPM_OK;
}

int spm_batch_stream(const spm_batch *b, void **s)
{
    if (!b || !b->h || !s) return spm_fail(SPM_EINVAL, "spm_batch_stream: null argument");
    *s = (void *)b->h->c.stream;
    return SPM_OK;
}

int spm_batch_info(const spm_batch *b, int *capacity, size_t *dev_bytes)
{
    if (!b || !b->h) return spm_fail(SPM_EINVAL, "null batch");
    if (capacity) *capacity = b->cap;
    if (dev_bytes) *dev_bytes = b->h->dev_bytes;
    return SPM_OK;
}

int spm_batch_load(spm_batch *b, int slot, const void *phi, double Tau, double Gamma, int max_iter, double res_tol,
                   double gn_tol)
{
    CHECK_BATCH(b);
    CHECK_SLOT(b, slot, "spm_batch_load");
    if (!phi) return spm_fail(SPM_EINVAL, "spm_batch_load: null field");
    if (max_iter < 0 || max_iter > c.hist_cap)
        return spm_fail(SPM_EINVAL, "spm_batch_load: max_iter must be in 0.." + std::to_string(c.hist_cap));
    if (!std::isfinite(Tau) || !std::isfinite(Gamma)) return spm_fail(SPM_EINVAL, "spm_batch_load: Tau, Gamma");
    hipLaunchKernelGGL(k_set_prm, dim3(1), dim3(1), 0, c.stream, c.prm, slot, Tau, Gamma);
    spm_ctl_init(b->h, slot, max_iter, res_tol, gn_tol);
    c.g->gf_init_member(c, slot, phi, Tau, Gamma);
    return spm_launch_status("spm_batch_load");
}

int spm_batch_steps(spm_batch *b, int nsteps)
{
    CHECK_BATCH(b);
    if (nsteps < 0) return spm_fail(SPM_EINVAL, "spm_batch_steps: nsteps >= 0");
    int s = 0;
    if (nsteps >= GRAPH_STEPS) {
        for (; s + GRAPH_STEPS <= nsteps; s += GRAPH_STEPS) HIPCHK(hipGraphLaunch(b->gexec, c.stream));
    }
    for (; s < nsteps; ++s) c.g->gf_step(c);
    return spm_launch_status("spm_batch_steps");
}

int spm_batch_poll(spm_batch *b, int *iters, int *stopped)
{
    CHECK_BATCH(b);
    RCHK(read_ctl(b));
    int running = 0;
    for (int m = 0; m < b->cap; ++m) {
        if (iters) iters[m] = b->h_ctl[m].iter;
        if (stopped) stopped[m] = b->h_ctl[m].stop;
        running += b->h_ctl[m].stop ? 0 : 1;
    }
    return running;
}

int spm_batch_read(spm_batch *b, int slot, void *phi, double *hist, int *iters)
{
    CHECK_BATCH(b);
    CHECK_SLOT(b, slot, "spm_batch_read");
    RCHK(read_ctl(b));
    const Ctl st = b->h_ctl[slot];
    if (!st.stop) return spm_fail(SPM_EINVAL, "spm_batch_read: member " + std::to_string(slot) + " is running");
    if (phi) c.g->gf_read_member(c, slot, phi);
    RCHK(spm_launch_status("spm_batch_read"));
    if (hist && st.iter > 0)
        HIPCHK(hipMemcpyAsync(hist, c.hist + 3L * c.hist_cap * slot, sizeof(double) * 3 * st.iter,
                              hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (iters) *iters = st.iter;
    return SPM_OK;
}

int spm_batch_release(spm_batch *b, int slot)
{
    CHECK_BATCH(b);
    CHECK_SLOT(b, slot, "spm_batch_release");
    spm_ctl_init(b->h, slot, 0, 0.0, 0.0);
    return spm_launch_status("spm_batch_release");
}

} // extern "C"
