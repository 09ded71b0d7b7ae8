// tune.hip -- choice by measurement of the FFT pass shapes (threads per pencil
// of the row, j and i passes of the four-pass step and the operators) and of
// the two-pass step's variant, with a cache of the results on disk: the
// counterpart of the reference's FFTW_PATIENT planning and wisdom files
// (My_mpi_fft_set_up, My_FFT.cpp:45-97).
//
// Cache: one line per measurement, "arch N precision members choice us_per_step
// table", appended to $SPM_TUNE_CACHE (default $HOME/.cache/spm-iqc/shapes.txt);
// the last line for a key wins.  Members are bucketed to the next power of two.
// "table" is a signature of what was chosen from: of N's shape table in this
// build for shapes (a line recorded with another table -- a rebuild that added,
// dropped or reordered shapes -- never matches, and the shape is measured
// again); "sj" and the variant count for two-pass step variants, whose lines
// carry the arch as "<arch>/sj".
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace spm;

namespace {
std::mutex g_mu;
std::map<std::string, int> g_table;
bool g_loaded = false;

std::string cache_path()
{
    if (const char *e = std::getenv("SPM_TUNE_CACHE")) return e;
    const char *home = std::getenv("HOME");
    return std::string(home ? home : ".") + "/.cache/spm-iqc/shapes.txt";
}

std::string arch()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t k = a.find(':');
    return k == std::string::npos ? a : a.substr(0, k);
}

int bucket(int nmem)
{
    int b = 1;
    while (b < nmem) b *= 2;
    return b;
}

// FNV-1a of N's shape table (the threads per pencil of every shape, in order)
std::string table_sig(int N)
{
    const int (*s)[3] = nullptr;
    const int n = grid_shapes(N, &s);
    unsigned long long h = 1469598103934665603ull;
    auto mix = [&](int v) {
        h ^= (unsigned)v;
        h *= 1099511628211ull;
    };
    mix(n);
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) mix(s[k][a]);
    char buf[24];
    std::snprintf(buf, sizeof buf, "%016llx", h);
    return buf;
}

std::string key(const std::string &ar, int N, int prec, int nmem, const std::string &sig)
{
    return ar + " " + std::to_string(N) + " " + std::to_string(prec) + " " + std::to_string(bucket(nmem)) + " " + sig;
}

void load_cache() // g_mu held
{
    if (g_loaded) return;
    g_loaded = true;
    FILE *f = std::fopen(cache_path().c_str(), "r");
    if (!f) return;
    char line[256], ar[64], sig[32];
    int N, prec, nm, shape;
    double us;
    while (std::fgets(line, sizeof line, f)) // lines without a table signature are skipped
        if (std::sscanf(line, "%63s %d %d %d %d %lf %31s", ar, &N, &prec, &nm, &shape, &us, sig) == 7)
            g_table[key(ar, N, prec, nm, sig)] = shape;
    std::fclose(f);
}

std::string step_sig() { return "sj" + std::to_string(SPM_STEP_VARIANTS); }
std::string step_arch() { return arch() + "/sj"; }

void record(const std::string &ar, int N, int prec, int nmem, int shape, double us, const std::string &sig) // g_mu held
{
    g_table[key(ar, N, prec, nmem, sig)] = shape;
    const std::string path = cache_path();
    for (size_t k = path.find('/', 1); k != std::string::npos; k = path.find('/', k + 1))
        (void)mkdir(path.substr(0, k).c_str(), 0755); // mkdir -p of the parent; fopen reports failure
    if (FILE *f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, "%s %d %d %d %d %.3f %s\n", ar.c_str(), N, prec, bucket(nmem), shape, us, sig.c_str());
        std::fclose(f);
    }
}

// U <- pseudo-random spectrum of amplitude amp (a hash of the index), so the
// timed steps move realistic data (all-zero fields run measurably faster).
template <typename T> __global__ void k_fill_noise(C2<T> *p, long n, double amp)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double a = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        const double b = (double)((z << 21) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        p[i] = C2<T>{(T)(amp * a), (T)(amp * b)};
    }
}

// Time `steps` iterations of a batch of nmem fields with the parameters q,
// replayed from one captured graph (after one untimed replay), starting from
// a noise spectrum whose field has r.m.s. about 0.1.
int time_steps(const spm_params &q, int nmem, int steps, double *us)
{
    spm_ctx *h = nullptr;
    RCHK(spm_ctx_create(&q, nmem, 2L * steps + 2, &h));
    Ctx &c = h->c;
    int rc = SPM_OK;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto fail = [&](const char *what, hipError_t e) {
        rc = spm_fail(SPM_EHIP, std::string("spm_tune: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMemsetAsync(c.W, 0, (size_t)c.C * 2 * c.esz * nmem, c.stream)) != hipSuccess) fail("memset", e);
    if (!rc) {
        const double amp = 0.2 / std::sqrt(2.0 * (double)c.R); // field r.m.s. ~ 0.1
        if (c.prec == SPM_FP64)
            hipLaunchKernelGGL(k_fill_noise<double>, dim3(2048), dim3(256), 0, c.stream, (C2<double> *)c.U,
                               c.C * nmem, amp);
        else
            hipLaunchKernelGGL(k_fill_noise<float>, dim3(2048), dim3(256), 0, c.stream, (C2<float> *)c.U,
                               c.C * nmem, amp);
        if ((e = hipGetLastError()) != hipSuccess) fail("noise", e);
    }
    for (int m = 0; !rc && m < nmem; ++m) spm_ctl_init(h, m, 2 * steps + 2, -1.0, 0.0); // never stops early
    if (!rc && (e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
        fail("capture", e);
    if (!rc) {
        for (int s = 0; s < steps; ++s) c.g->gf_step(c);
        if ((e = hipStreamEndCapture(c.stream, &g)) != hipSuccess) fail("capture", e);
    }
    if (!rc && (e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0)) != hipSuccess) fail("instantiate", e);
    if (!rc && ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess))
        fail("event", e);
    if (!rc && (e = hipGraphLaunch(ge, c.stream)) != hipSuccess) fail("launch", e);
    if (!rc && (e = hipEventRecord(e0, c.stream)) != hipSuccess) fail("event", e);
    if (!rc && (e = hipGraphLaunch(ge, c.stream)) != hipSuccess) fail("launch", e);
    if (!rc && (e = hipEventRecord(e1, c.stream)) != hipSuccess) fail("event", e);
    if (!rc && (e = hipEventSynchronize(e1)) != hipSuccess) fail("sync", e);
    float ms = 0.f;
    if (!rc && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) fail("elapsed", e);
    if (!rc) *us = 1000.0 * ms / steps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    spm_ctx_free(h);
    return rc;
}

// the shape sets the passes of the four-pass step, which are the ones the
// other operators (LOBPCG, SPM forces) run; the two-pass step has its own
int time_shape(const spm_params *p, int nmem, int steps, int v, double *us)
{
    spm_params q = *p;
    q.shape = v;
    q.step_passes = 4;
    q.step_variant = 0;
    return time_steps(q, nmem, steps, us);
}

int time_step_variant(const spm_params *p, int nmem, int steps, int v, double *us)
{
    spm_params q = *p;
    q.shape = p->shape < 0 ? 0 : p->shape; // the two-pass step does not depend on it
    q.step_passes = 2;
    q.step_variant = SPM_STEP_EXPLICIT | v;
    return time_steps(q, nmem, steps, us);
}
} // namespace

int spm_resolve_shape(const spm_params *p, int nmem, int *shape)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_cache();
        auto it = g_table.find(key(arch(), p->N, p->precision, nmem, table_sig(p->N)));
        const int (*s)[3] = nullptr;
        if (it != g_table.end() && it->second >= 0 && it->second < grid_shapes(p->N, &s)) {
            *shape = it->second;
            return SPM_OK;
        }
    }
    return spm_tune(p, nmem, 20, shape, nullptr);
}

int spm_resolve_step(const spm_params *p, int nmem, int *variant)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_cache();
        auto it = g_table.find(key(step_arch(), p->N, p->precision, nmem, step_sig()));
        if (it != g_table.end() && it->second >= 0 && it->second < SPM_STEP_VARIANTS &&
            (it->second & SPM_STEP_U_MASK) != SPM_STEP_U_MASK) {
            *variant = it->second;
            return SPM_OK;
        }
    }
    return spm_tune_step(p, nmem, 20, variant, nullptr);
}

extern "C" {

int spm_shapes(int N, int *count, int *tpp, int max_shapes)
{
    if (!count) return spm_fail(SPM_EINVAL, "spm_shapes: null count");
    const int (*s)[3] = nullptr;
    const int n = grid_shapes(N, &s);
    if (n <= 0) return spm_fail(SPM_ENOTSUP, "spm_shapes: unsupported N " + std::to_string(N));
    *count = n;
    if (tpp)
        for (int k = 0; k < n && k < max_shapes; ++k)
            for (int a = 0; a < 3; ++a) tpp[3 * k + a] = s[k][a];
    return SPM_OK;
}

int spm_tune(const spm_params *p, int nmem, int steps, int *best, double *us_per_step)
{
    if (!p || !best) return spm_fail(SPM_EINVAL, "spm_tune: null argument");
    if (nmem < 1 || steps < 1) return spm_fail(SPM_EINVAL, "spm_tune: nmem >= 1, steps >= 1");
    const int (*s)[3] = nullptr;
    const int n = grid_shapes(p->N, &s);
    if (n <= 0) return spm_fail(SPM_ENOTSUP, "spm_tune: unsupported N " + std::to_string(p->N));
    // two sweeps over the candidates, the faster time of each kept: the first
    // candidates of a single sweep would run while the clocks still ramp up
    std::vector<double> best_us(n, 0.0);
    for (int sweep = 0; sweep < 2; ++sweep)
        for (int v = 0; v < n; ++v) {
            double us = 0.0;
            const int rc = time_shape(p, nmem, steps, v, &us);
            if (rc) return spm_fail(rc, "spm_tune: shape " + std::to_string(v) + ": " + spm_last_error());
            if (sweep == 0 || us < best_us[v]) best_us[v] = us;
        }
    int bi = 0;
    for (int v = 0; v < n; ++v) {
        if (us_per_step) us_per_step[v] = best_us[v];
        if (best_us[v] < best_us[bi]) bi = v;
    }
    const double bt = best_us[bi];
    *best = bi;
    std::lock_guard<std::mutex> lk(g_mu);
    load_cache();
    record(arch(), p->N, p->precision, nmem, bi, bt, table_sig(p->N));
    return SPM_OK;
}

int spm_tune_step(const spm_params *p, int nmem, int steps, int *best, double *us_per_variant)
{
    if (!p || !best) return spm_fail(SPM_EINVAL, "spm_tune_step: null argument");
    if (nmem < 1 || steps < 1) return spm_fail(SPM_EINVAL, "spm_tune_step: nmem >= 1, steps >= 1");
    if (p->N != 64 && p->N != 128 && p->N != 256)
        return spm_fail(SPM_ENOTSUP, "spm_tune_step: the two-pass step exists for N = 64, 128 and 256 only");
    std::vector<double> best_us(SPM_STEP_VARIANTS, 0.0);
    for (int sweep = 0; sweep < 2; ++sweep) // as spm_tune: two sweeps, the faster time kept
        for (int v = 0; v < SPM_STEP_VARIANTS; ++v) {
            double us = 0.0;
            const int rc = time_step_variant(p, nmem, steps, v, &us);
            if (rc) return spm_fail(rc, "spm_tune_step: variant " + std::to_string(v) + ": " + spm_last_error());
            if (sweep == 0 || us < best_us[v]) best_us[v] = us;
        }
    int bi = 0;
    for (int v = 0; v < SPM_STEP_VARIANTS; ++v) {
        if (us_per_variant) us_per_variant[v] = best_us[v];
        if (best_us[v] < best_us[bi]) bi = v;
    }
    const double bt = best_us[bi];
    *best = bi;
    std::lock_guard<std::mutex> lk(g_mu);
    load_cache();
    record(step_arch(), p->N, p->precision, nmem, bi, bt, step_sig());
    return SPM_OK;
}

} // extern "C"
