// lobpcg.hip -- the reference's LOBPCG variant for the lowest Hessian
// eigenpairs (hiosd_lobpcg_initialization / hiosd_lobpcg_inip,
// LOBPCG.cpp:10-487) on device-resident blocks of fields.
//
// Per iteration, as the reference: U = H V; alpha_i = <V_i, U_i>;
// W = precond(U - V alpha); W orthonormalised against V and in order with
// deflation at norm 1e-8; Y = H W; Vp orthonormalised against [V, W] and in
// order with deflation; Up = H Vp; Rayleigh-Ritz on the kept basis
// UU = [V, W, Vp] with the symmetrised Gram UU^T YY / N^3; V <- lowest k Ritz
// vectors, Vp <- old V; mean removal; V orthonormalised in order.
// The sequential Gram-Schmidt sweeps become block projections and CholQR2
// on FP64 Gram matrices computed on the matrix cores; every inner product
// keeps the reference's 1/N^3 weight.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "block_ops.h"

using namespace spm;

namespace spm {

// Cyclic Jacobi eigensolver for the symmetric m x m Rayleigh-Ritz matrix
// (replaces Eigen's EigenSolver, LOBPCG.cpp:362-382).  A row-major, destroyed;
// w ascending (sort_eig, BasicOperators.cpp:313-334); Z columns = vectors.
void sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &Z)
{
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) {
                tot += A[p * n + q] * A[p * n + q];
                if (p != q) off += A[p * n + q] * A[p * n + q];
            }
        if (off == 0.0 || off <= 1e-60 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < n; ++r) {
                    const double arp = A[r * n + p], arq = A[r * n + q];
                    A[r * n + p] = cs * arp - sn * arq;
                    A[r * n + q] = sn * arp + cs * arq;
                }
                for (int r = 0; r < n; ++r) {
                    const double apr = A[p * n + r], aqr = A[q * n + r];
                    A[p * n + r] = cs * apr - sn * aqr;
                    A[q * n + r] = sn * apr + cs * aqr;
                }
                A[p * n + q] = A[q * n + p] = 0.0;
                for (int r = 0; r < n; ++r) {
                    const double vrp = V[r * n + p], vrq = V[r * n + q];
                    V[r * n + p] = cs * vrp - sn * vrq;
                    V[r * n + q] = sn * vrp + cs * vrq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
    w.assign(n, 0.0);
    Z.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        w[i] = A[idx[i] * n + idx[i]];
        for (int r = 0; r < n; ++r) Z[r * n + i] = V[r * n + idx[i]];
    }
}

static int ensure_ws(spm_ctx *h, size_t bytes)
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

static inline void *fld(void *base, long i, const Ctx &c) { return (char *)base + (size_t)i * c.R * c.esz; }

// hiosd_lobpcg_initialization's start (LOBPCG.cpp:31-102): the caller's random
// block, each field mean-removed, then orthonormalised in order.
int lobpcg_init(spm_ctx *h, const void *start, void *V, int k)
{
    Ctx &c = h->c;
    std::vector<void *> Vp(k);
    std::vector<const void *> Vc(k);
    for (int i = 0; i < k; ++i) {
        Vp[i] = fld(V, i, c);
        Vc[i] = Vp[i];
        if (start != V) RCHK(field_copy(h, Vp[i], (const char *)start + (size_t)i * c.R * c.esz));
    }
    std::vector<double> mean(k), eye(k * k, 0.0), shift(k);
    RCHK(block_means(h, Vc.data(), k, mean.data()));
    for (int i = 0; i < k; ++i) {
        eye[i * k + i] = 1.0;
        shift[i] = -mean[i];
    }
    RCHK(block_combine(h, Vc.data(), k, Vp.data(), k, eye.data(), shift.data()));
    std::vector<char> keep(k);
    return block_ortho(h, Vp.data(), k, nullptr, 0, -1.0, (bool *)keep.data());
}

// The block state of a LOBPCG iteration: V (the caller's fields or, after the
// first Ritz update, workspace fields), U = H V, W, Y = H W, Vp, Up = H Vp, and a
// spare block T; keep flags of V, W, Vp.
struct LobBlock {
    std::vector<void *> V, U, W, Y, Vp, Up, T;
    std::vector<const void *> V0; // fixed orthonormal fields the sweep stays orthogonal to (GOSD)
    std::vector<char> keepbuf;
    bool *keepV() { return (bool *)keepbuf.data(); }
    bool *keepW() { return (bool *)keepbuf.data() + V.size(); }
    bool *keepP() { return (bool *)keepbuf.data() + 2 * V.size(); }
};

static std::vector<const void *> cv(const std::vector<void *> &a) { return std::vector<const void *>(a.begin(), a.end()); }

// fields a[0..n) one after another (R values apart): a block the batched
// operators can take as members
static bool contiguous(const std::vector<void *> &a, int n, const Ctx &c)
{
    for (int i = 1; i < n; ++i)
        if ((char *)a[i] != (char *)a[0] + (size_t)i * c.R * c.esz) return false;
    return true;
}

// Fields per launch of the batched operators: BATCH_OPS where a field's
// spectrum is at most 64 MB (N <= 128), where the operator passes are short and
// their launches cost a large share of an iteration; one at 256^3, where the
// passes are long and a batch gained nothing (LOBPCG iteration, block 8, two
// passes: 64^3 1.04 -> 0.50 ms FP64, 128^3 2.35 -> 1.75 ms; 256^3 12.78 against
// 12.93 ms FP64 and 8.95 against 8.57 ms FP32).
static int batch_chunk(const Ctx &c)
{
    const double spec = (double)c.C * 2 * c.esz;
    return spec <= 64e6 ? (int)GridIface::BATCH_OPS : 1;
}

static int ensure_specws(spm_ctx *h, int n)
{
    const size_t bytes = (size_t)n * h->c.C * 2 * h->c.esz;
    if (h->specws_bytes >= bytes) return SPM_OK;
    if (h->specws) {
        HIPCHK(hipStreamSynchronize(h->c.stream));
        HIPCHK(hipFree(h->specws));
        h->dev_bytes -= h->specws_bytes;
        h->specws = nullptr;
        h->specws_bytes = 0;
    }
    RCHK(spm_dev_alloc(h, &h->specws, bytes));
    h->specws_bytes = bytes;
    return SPM_OK;
}

// out_i = H in_i (hv_cam, LOBPCG.cpp:113-166) for the fields with keep[i] (keep
// NULL: all), the others' outputs left alone; with dots, dots[i] = <in_i, out_i>
// (device).  Contiguous blocks with every field kept go through the batched
// operators, batch_chunk fields per launch of each pass (the same kernels and
// sums per field as field by field, so the same values); anything else field by
// field.
static int hv_block(spm_ctx *h, const void *phi, const std::vector<void *> &in, const std::vector<void *> &out,
                    const bool *keep, double *dots)
{
    Ctx &c = h->c;
    const int n = (int)in.size();
    bool all = true;
    for (int i = 0; keep && i < n; ++i) all = all && keep[i];
    const int chunk = std::min(n, batch_chunk(c));
    if (all && chunk > 1 && contiguous(in, n, c) && contiguous(out, n, c)) {
        double *means = block_small(h);
        if (!means) return spm_fail(SPM_ENOMEM, "lobpcg: block scratch");
        means += 16;
        RCHK(ensure_specws(h, chunk));
        bool done = true;
        for (int i0 = 0; i0 < n && done; i0 += chunk)
            done = c.g->hv_many(c, phi, in[i0], out[i0], std::min(chunk, n - i0), h->specws, means,
                                dots ? dots + i0 : nullptr);
        if (done) return spm_launch_status("lobpcg hv");
        // no batched form (then nothing was launched): field by field below
    }
    for (int i = 0; i < n; ++i) {
        if (keep && !keep[i]) continue;
        if (dots) c.g->hv_dot(c, phi, in[i], out[i], dots + i);
        else c.g->hv(c, phi, in[i], out[i]);
    }
    return spm_launch_status("lobpcg hv");
}

// V = the k fields at Vout, U..T the first 6k fields of the workspace (zeroed)
static int lob_block(spm_ctx *h, void *Vout, int k, size_t extra_fields, LobBlock &b)
{
    Ctx &c = h->c;
    const size_t fb = (size_t)c.R * c.esz;
    RCHK(ensure_ws(h, (6 * (size_t)k + extra_fields) * fb));
    b.V.resize(k);
    b.U.resize(k);
    b.W.resize(k);
    b.Y.resize(k);
    b.Vp.resize(k);
    b.Up.resize(k);
    b.T.resize(k);
    for (int i = 0; i < k; ++i) {
        b.V[i] = fld(Vout, i, c);
        b.U[i] = fld(h->ws, 0 * k + i, c);
        b.W[i] = fld(h->ws, 1 * k + i, c);
        b.Y[i] = fld(h->ws, 2 * k + i, c);
        b.Vp[i] = fld(h->ws, 3 * k + i, c);
        b.Up[i] = fld(h->ws, 4 * k + i, c);
        b.T[i] = fld(h->ws, 5 * k + i, c);
    }
    b.keepbuf.assign(3 * k, 0);
    HIPCHK(hipMemsetAsync(h->ws, 0, (6 * (size_t)k + extra_fields) * fb, c.stream));
    return SPM_OK;
}

// ---------------------------------------------------------------- block sweep
// Small dense matrices, row-major.
using Mat = std::vector<double>;
// C (r x c) = A^T (A: n x r) B (n x c)
static Mat mtm(const Mat &A, const Mat &B, int n, int r, int c)
{
    Mat C((size_t)r * c, 0.0);
    for (int l = 0; l < n; ++l)
        for (int i = 0; i < r; ++i) {
            const double a = A[(size_t)l * r + i];
            if (a != 0.0)
                for (int j = 0; j < c; ++j) C[(size_t)i * c + j] += a * B[(size_t)l * c + j];
        }
    return C;
}
// C (r x c) = A (r x n) B (n x c)
static Mat mm(const Mat &A, const Mat &B, int r, int n, int c)
{
    Mat C((size_t)r * c, 0.0);
    for (int i = 0; i < r; ++i)
        for (int l = 0; l < n; ++l) {
            const double a = A[(size_t)i * n + l];
            if (a != 0.0)
                for (int j = 0; j < c; ++j) C[(size_t)i * c + j] += a * B[(size_t)l * c + j];
        }
    return C;
}
// the (i0.., j0..) block (r x c) of the n-column matrix G
static Mat sub(const Mat &G, int n, int i0, int j0, int r, int c)
{
    Mat S((size_t)r * c);
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < c; ++j) S[(size_t)i * c + j] = G[(size_t)(i0 + i) * n + j0 + j];
    return S;
}

// Cholesky with deflation of an orthogonal complement's Gram S = G - P^T P
// formed from Grams (G: the fields' own, P: their components along an
// orthonormal basis), where the subtraction loses digits.  The keep/drop
// decisions and R agree with those of the explicitly projected fields only
// where each remaining squared norm d_i is far above the rounding of the
// subtraction, about 1e-13 of the field's own squared norm g_i: d_i >= 1e-6
// g_i for a kept field, d_i + 1e3 (1e-13 g_i) < tol^2 for a dropped one.
static bool chol_from_grams(int n, const Mat &S, const double *g, double tol, bool *keep, Mat &Rinv)
{
    std::vector<double> piv(n);
    Rinv.assign((size_t)n * n, 0.0);
    chol_inv(n, S.data(), tol, true, keep, Rinv.data(), piv.data());
    for (int i = 0; i < n; ++i) {
        const double noise = 1e-13 * g[i];
        if (keep[i] ? !(piv[i] >= 1e-6 * g[i]) : !(piv[i] + 1e3 * noise < tol * tol)) return false;
    }
    return true;
}

// Rayleigh-Ritz of the symmetric pencil (A, B), B the measured Gram of the
// basis: B = R^T R, the eigenpairs (w, y) of R^-T A R^-1, Z = R^-1 y (B-orthonormal
// columns, ascending w).  Returns false when B is not positive definite.
static bool ritz_pencil(int m, const Mat &A, const Mat &B, std::vector<double> &w, Mat &Z)
{
    Mat Rinv;
    std::vector<char> keepb(m);
    Rinv.assign((size_t)m * m, 0.0);
    std::vector<double> piv(m);
    chol_inv(m, B.data(), -1.0, false, (bool *)keepb.data(), Rinv.data(), piv.data());
    for (int i = 0; i < m; ++i)
        if (!keepb[i] || !(piv[i] > 1e-8 * B[(size_t)i * m + i])) return false;
    Mat C = mm(mtm(Rinv, A, m, m, m), Rinv, m, m, m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < i; ++j) C[(size_t)i * m + j] = C[(size_t)j * m + i] = 0.5 * (C[(size_t)i * m + j] + C[(size_t)j * m + i]);
    Mat Y;
    sym_eig(m, C, w, Y);
    Z = mm(Rinv, Y, m, m, m);
    return true;
}

static int lobpcg_sweep_ref(spm_ctx *h, const void *phi, LobBlock &b, double *ritz, bool from_w = false);
static int lobpcg_rayleigh_ritz(spm_ctx *h, const void *phi, LobBlock &b, double *ritz);

// One iteration of hiosd_lobpcg_inip (LOBPCG.cpp:279-449) at phi, without fixed
// fields V0, in the fewest passes over the block:
//   U = H V and alpha_i = <V_i, U_i> from the same passes (hv_dot);
//   W = precond(U - alpha V), the residual formed in the preconditioner's input pass;
//   one Gram of [V, W, Vp]; from it the orthonormalisation of W against V and of
//   Vp against [V, W] (in order, with the reference's deflation at norm 1e-8) as
//   coefficient matrices, applied in one combine that writes both blocks.  Where
//   the complement's Gram cancels too much to decide (Vp close to span V, which
//   is where the iteration converges) Vp is projected explicitly and its Gram
//   measured again; where W's does, the sweep takes the reference-order path;
//   Y = H W, Up = H Vp on the kept fields;
//   one pass over [X | YY], X = [V, W, Vp] (kept), YY = [U, Y, Up]: A = X^T YY,
//   B = X^T X and the means of X; the Ritz pencil (A, B); the lowest k Ritz
//   vectors, their means removed and orthonormalised in order, in one combine
//   whose coefficients come from B and the means.
// Each step is the reference's in exact arithmetic; the Ritz values use the
// measured B rather than assuming the basis orthonormal.
static int lobpcg_sweep(spm_ctx *h, const void *phi, LobBlock &b, double *ritz)
{
    if (!b.V0.empty()) return lobpcg_sweep_ref(h, phi, b, ritz);
    Ctx &c = h->c;
    const int k = (int)b.V.size();
    std::vector<void *> &V = b.V, &U = b.U, &W = b.W, &Y = b.Y, &Vp = b.Vp, &Up = b.Up, &T = b.T;
    bool *keepW = b.keepW(), *keepP = b.keepP();
    // U = H V, alpha (:281-284)
    double *adev = block_small(h);
    if (!adev) return spm_fail(SPM_ENOMEM, "lobpcg: block scratch");
    RCHK(hv_block(h, phi, V, U, nullptr, adev));
    // W = precond(U - alpha V) (:284-300): batched with alpha on the device, or field by field
    const int chunk = std::min(k, batch_chunk(c));
    bool batched = chunk > 1 && contiguous(V, k, c) && contiguous(U, k, c) && contiguous(W, k, c);
    if (batched) {
        RCHK(ensure_specws(h, chunk));
        for (int i0 = 0; i0 < k && batched; i0 += chunk)
            batched = c.g->precond_resid_many(c, U[i0], V[i0], adev + i0, W[i0], std::min(chunk, k - i0),
                                              h->specws);
    }
    if (!batched) {
        std::vector<double> alpha(k);
        HIPCHK(hipMemcpyAsync(h->h_pin, adev, sizeof(double) * k, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        std::memcpy(alpha.data(), h->h_pin, sizeof(double) * k);
        for (int i = 0; i < k; ++i) c.g->precond_resid(c, U[i], V[i], alpha[i], W[i]);
    }
    RCHK(spm_launch_status("lobpcg precond"));
    // G = [V, W, Vp]^T [V, W, Vp]
    const int n3 = 3 * k;
    std::vector<const void *> X3(n3);
    for (int i = 0; i < k; ++i) {
        X3[i] = V[i];
        X3[k + i] = W[i];
        X3[2 * k + i] = Vp[i];
    }
    Mat G((size_t)n3 * n3);
    RCHK(block_gram(h, X3.data(), n3, X3.data(), n3, G.data()));
    // W against V, then in order (:301-318): W' = W - V PW, Gram WW - PW^T PW
    const Mat PW = sub(G, n3, 0, k, k, k), Gww = sub(G, n3, k, k, k, k);
    Mat SW = Gww, RW;
    {
        const Mat PtP = mtm(PW, PW, k, k, k);
        for (size_t e = 0; e < SW.size(); ++e) SW[e] -= PtP[e];
    }
    std::vector<double> gw(k), gp(k);
    for (int i = 0; i < k; ++i) {
        gw[i] = Gww[(size_t)i * k + i];
        gp[i] = G[(size_t)(2 * k + i) * n3 + 2 * k + i];
    }
    if (!chol_from_grams(k, SW, gw.data(), 1e-8, keepW, RW)) return lobpcg_sweep_ref(h, phi, b, ritz, true);
    // W^ = [V, W] CW, CW = [-PW RW^-1; RW^-1]
    Mat CW((size_t)2 * k * k, 0.0);
    {
        const Mat PR = mm(PW, RW, k, k, k);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                CW[(size_t)i * k + j] = -PR[(size_t)i * k + j];
                CW[(size_t)(k + i) * k + j] = RW[(size_t)i * k + j];
            }
    }
    // Vp against V and W^, then in order (:319-331): PV = V^T Vp, PH = W^^T Vp = CW^T [V, W]^T Vp
    const Mat PV = sub(G, n3, 0, 2 * k, k, k);
    const Mat PH = mtm(CW, sub(G, n3, 0, 2 * k, 2 * k, k), 2 * k, k, k);
    Mat SP = sub(G, n3, 2 * k, 2 * k, k, k), RP;
    {
        const Mat A1 = mtm(PV, PV, k, k, k), A2 = mtm(PH, PH, k, k, k);
        for (size_t e = 0; e < SP.size(); ++e) SP[e] -= A1[e] + A2[e];
    }
    const bool fastP = chol_from_grams(k, SP, gp.data(), 1e-8, keepP, RP);
    // Vp' = Vp - V PV - W^ PH = [V, W, Vp] D, D = [-PV - CW_V PH; -CW_W PH; I]
    Mat D((size_t)n3 * k, 0.0);
    {
        const Mat CH = mm(CW, PH, 2 * k, k, k);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                D[(size_t)i * k + j] = -PV[(size_t)i * k + j] - CH[(size_t)i * k + j];
                D[(size_t)(k + i) * k + j] = -CH[(size_t)(k + i) * k + j];
                D[(size_t)(2 * k + i) * k + j] = i == j ? 1.0 : 0.0;
            }
    }
    const Mat DP = fastP ? mm(D, RP, n3, k, k) : D;
    {   // one combine: [V, W, Vp] -> [W^, Vp^ (or Vp')]
        Mat coef((size_t)n3 * 2 * k, 0.0);
        for (int a = 0; a < n3; ++a)
            for (int j = 0; j < k; ++j) {
                if (a < 2 * k) coef[(size_t)a * 2 * k + j] = CW[(size_t)a * k + j];
                coef[(size_t)a * 2 * k + k + j] = DP[(size_t)a * k + j];
            }
        std::vector<void *> out(2 * k);
        for (int i = 0; i < k; ++i) {
            out[i] = W[i];
            out[k + i] = Vp[i];
        }
        RCHK(block_combine(h, X3.data(), n3, out.data(), 2 * k, coef.data(), nullptr));
    }
    if (!fastP) {
        // Vp' measured against [V, W^ (kept)]: c = M^-1 [V, W^]^T Vp', Gram Vp'^T Vp' - b^T c
        Mat G2((size_t)n3 * n3);
        RCHK(block_gram(h, X3.data(), n3, X3.data(), n3, G2.data()));
        std::vector<int> bi; // basis rows of G2: V and the kept W^
        for (int i = 0; i < k; ++i) bi.push_back(i);
        for (int i = 0; i < k; ++i)
            if (keepW[i]) bi.push_back(k + i);
        const int nb = (int)bi.size();
        Mat M((size_t)nb * nb), Bv((size_t)nb * k);
        for (int r = 0; r < nb; ++r) {
            for (int q = 0; q < nb; ++q) M[(size_t)r * nb + q] = G2[(size_t)bi[r] * n3 + bi[q]];
            for (int j = 0; j < k; ++j) Bv[(size_t)r * k + j] = G2[(size_t)bi[r] * n3 + 2 * k + j];
        }
        Mat MR((size_t)nb * nb);
        bool orth = true;
        {
            std::vector<char> kb(nb);
            std::vector<double> piv(nb);
            chol_inv(nb, M.data(), -1.0, false, (bool *)kb.data(), MR.data(), piv.data());
            for (int r = 0; r < nb; ++r) orth = orth && kb[r] && piv[r] > 0.5 * M[(size_t)r * nb + r];
        }
        if (!orth) { // [V, W^] far from orthonormal: Vp' against it in the reference's order
            std::vector<const void *> Bf;
            for (int r = 0; r < nb; ++r) Bf.push_back(X3[bi[r]]);
            RCHK(block_ortho(h, Vp.data(), k, Bf.data(), nb, 1e-8, keepP));
            return lobpcg_rayleigh_ritz(h, phi, b, ritz);
        }
        const Mat Cc = mm(MR, mtm(MR, Bv, nb, nb, k), nb, nb, k); // M^-1 Bv
        Mat S2 = sub(G2, n3, 2 * k, 2 * k, k, k);
        {
            const Mat BtC = mtm(Bv, Cc, nb, k, k);
            for (size_t e = 0; e < S2.size(); ++e) S2[e] -= BtC[e];
        }
        Mat R2((size_t)k * k);
        chol_inv(k, S2.data(), 1e-8, true, keepP, R2.data(), nullptr);
        // Vp^ = (Vp' - [V, W^] c) R2^-1
        std::vector<const void *> in;
        for (int r = 0; r < nb; ++r) in.push_back(X3[bi[r]]);
        for (int i = 0; i < k; ++i) in.push_back(Vp[i]);
        const int nin = nb + k;
        Mat coef((size_t)nin * k, 0.0);
        const Mat CR = mm(Cc, R2, nb, k, k);
        for (int r = 0; r < nb; ++r)
            for (int j = 0; j < k; ++j) coef[(size_t)r * k + j] = -CR[(size_t)r * k + j];
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) coef[(size_t)(nb + i) * k + j] = R2[(size_t)i * k + j];
        RCHK(block_combine(h, in.data(), nin, Vp.data(), k, coef.data(), nullptr));
    }
    return lobpcg_rayleigh_ritz(h, phi, b, ritz);
}

// Y = H W, Up = H Vp on the kept fields, then the Rayleigh-Ritz step, the mean
// removal and the orthonormalisation of the new V (LOBPCG.cpp:332-449) from one
// pass over [X | YY] and one combine.
static int lobpcg_rayleigh_ritz(spm_ctx *h, const void *phi, LobBlock &b, double *ritz)
{
    Ctx &c = h->c;
    const int k = (int)b.V.size();
    std::vector<void *> &V = b.V, &U = b.U, &W = b.W, &Y = b.Y, &Vp = b.Vp, &Up = b.Up, &T = b.T;
    bool *keepV = b.keepV(), *keepW = b.keepW(), *keepP = b.keepP();
    RCHK(hv_block(h, phi, W, Y, keepW, nullptr));
    RCHK(hv_block(h, phi, Vp, Up, keepP, nullptr));
    std::vector<const void *> X, YY;
    for (int i = 0; i < k; ++i) {
        X.push_back(V[i]);
        YY.push_back(U[i]);
    }
    for (int i = 0; i < k; ++i)
        if (keepW[i]) {
            X.push_back(W[i]);
            YY.push_back(Y[i]);
        }
    for (int i = 0; i < k; ++i)
        if (keepP[i]) {
            X.push_back(Vp[i]);
            YY.push_back(Up[i]);
        }
    const int m = (int)X.size();
    Mat A((size_t)m * m), B((size_t)m * m), Z;
    std::vector<double> mu(m), w;
    RCHK(block_gram_xy(h, X.data(), m, YY.data(), m, A.data(), B.data(), mu.data()));
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < i; ++j) {
            A[(size_t)i * m + j] = A[(size_t)j * m + i] = 0.5 * (A[(size_t)i * m + j] + A[(size_t)j * m + i]);
            B[(size_t)i * m + j] = B[(size_t)j * m + i] = 0.5 * (B[(size_t)i * m + j] + B[(size_t)j * m + i]);
        }
    if (!ritz_pencil(m, A, B, w, Z)) { // B singular: the reference's matrix, the basis taken as orthonormal
        Mat I((size_t)m * m, 0.0);
        for (int i = 0; i < m; ++i) I[(size_t)i * m + i] = 1.0;
        B = I;
        Mat Ps = A;
        sym_eig(m, Ps, w, Z);
    }
    // T = X Zk - mean, orthonormalised in order: Gram Zk^T B Zk - nu nu^T, nu = Zk^T mu
    Mat Zk((size_t)m * k);
    for (int a = 0; a < m; ++a)
        for (int i = 0; i < k; ++i) Zk[(size_t)a * k + i] = Z[(size_t)a * m + i];
    std::vector<double> nu(k, 0.0);
    for (int i = 0; i < k; ++i)
        for (int a = 0; a < m; ++a) nu[i] += Zk[(size_t)a * k + i] * mu[a];
    Mat ST = mtm(Zk, mm(B, Zk, m, m, k), m, k, k);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) ST[(size_t)i * k + j] -= nu[i] * nu[j];
    Mat RT((size_t)k * k);
    std::vector<double> piv(k);
    chol_inv(k, ST.data(), -1.0, false, keepV, RT.data(), piv.data());
    bool fastT = true;
    for (int i = 0; i < k; ++i) fastT = fastT && keepV[i] && piv[i] >= 1e-6;
    if (fastT) {
        const Mat CT = mm(Zk, RT, m, k, k);
        std::vector<double> shift(k, 0.0);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) shift[j] -= nu[i] * RT[(size_t)i * k + j];
        RCHK(block_combine(h, X.data(), m, T.data(), k, CT.data(), shift.data()));
    } else { // a Ritz vector close to a constant: mean removal and orthonormalisation in turn
        RCHK(block_combine(h, X.data(), m, T.data(), k, Zk.data(), nullptr));
        std::vector<double> mean(k), eye(k * k, 0.0), shift(k);
        RCHK(block_means(h, cv(T).data(), k, mean.data()));
        for (int i = 0; i < k; ++i) {
            eye[i * k + i] = 1.0;
            shift[i] = -mean[i];
        }
        RCHK(block_combine(h, cv(T).data(), k, T.data(), k, eye.data(), shift.data()));
        RCHK(block_ortho(h, T.data(), k, nullptr, 0, -1.0, keepV));
    }
    // Vp <- V, Up <- U (:391-400); V <- T
    std::vector<void *> oldVp = Vp, oldUp = Up;
    Vp = V;
    Up = U;
    V = T;
    T = oldVp;
    U = oldUp;
    if (ritz)
        for (int i = 0; i < k; ++i) ritz[i] = w[i];
    return SPM_OK;
}

// The sweep in the reference's order of operations (two projection rounds and
// CholQR2 per block, the Rayleigh-Ritz matrix assuming an orthonormal basis):
// with fixed fields V0 (GOSD), or from_w: W's complement Gram did not decide,
// U, alpha and W = precond(U - alpha V) already formed.
static int lobpcg_sweep_ref(spm_ctx *h, const void *phi, LobBlock &b, double *ritz, bool from_w)
{
    Ctx &c = h->c;
    const int k = (int)b.V.size();
    std::vector<void *> &V = b.V, &U = b.U, &W = b.W, &Y = b.Y, &Vp = b.Vp, &Up = b.Up, &T = b.T;
    bool *keepV = b.keepV(), *keepW = b.keepW(), *keepP = b.keepP();
    if (!from_w) {
        RCHK(hv_block(h, phi, V, U, nullptr, nullptr));
        std::vector<double> VU(k * k);
        RCHK(block_gram(h, cv(V).data(), k, cv(U).data(), k, VU.data()));
        {   // W = U - alpha V  (FuncsLinear2Cplx, :284)
            std::vector<const void *> in(2 * k);
            std::vector<double> coef(2 * k * k, 0.0);
            for (int i = 0; i < k; ++i) {
                in[i] = U[i];
                in[k + i] = V[i];
                coef[i * k + i] = 1.0;
                coef[(k + i) * k + i] = -VU[i * k + i];
            }
            RCHK(block_combine(h, in.data(), 2 * k, W.data(), k, coef.data(), nullptr));
        }
        for (int i = 0; i < k; ++i) c.g->precond(c, W[i]);
        RCHK(spm_launch_status("lobpcg precond"));
    }
    const int k0 = (int)b.V0.size();
    if (k0 == 0) {
        RCHK(block_ortho(h, W.data(), k, cv(V).data(), k, 1e-8, keepW));
    } else { // against V, then V0 (GOSD.cpp:205-207): V need not be orthogonal to V0 yet
        RCHK(block_project(h, W.data(), k, cv(V).data(), k));
        RCHK(block_ortho(h, W.data(), k, b.V0.data(), k0, 1e-8, keepW));
    }
    RCHK(hv_block(h, phi, W, Y, keepW, nullptr));
    for (int i = 0; i < k; ++i)
        if (!keepW[i]) RCHK(field_zero(h, Y[i]));
    if (k0 == 0) {
        std::vector<const void *> B(2 * k);
        for (int i = 0; i < k; ++i) {
            B[i] = V[i];
            B[k + i] = W[i];
        }
        RCHK(block_ortho(h, Vp.data(), k, B.data(), 2 * k, 1e-8, keepP));
    } else { // against V, W, then V0 in turn (GOSD.cpp:224-228)
        RCHK(block_project(h, Vp.data(), k, cv(V).data(), k));
        RCHK(block_project(h, Vp.data(), k, cv(W).data(), k));
        RCHK(block_ortho(h, Vp.data(), k, b.V0.data(), k0, 1e-8, keepP));
    }
    RCHK(hv_block(h, phi, Vp, Up, keepP, nullptr));
    for (int i = 0; i < k; ++i)
        if (!keepP[i]) RCHK(field_zero(h, Up[i]));
    // Rayleigh-Ritz on the kept basis (get_Merge / get_mat, :332-362)
    std::vector<const void *> UU, YY;
    for (int i = 0; i < k; ++i) {
        UU.push_back(V[i]);
        YY.push_back(U[i]);
    }
    for (int i = 0; i < k; ++i)
        if (keepW[i]) {
            UU.push_back(W[i]);
            YY.push_back(Y[i]);
        }
    for (int i = 0; i < k; ++i)
        if (keepP[i]) {
            UU.push_back(Vp[i]);
            YY.push_back(Up[i]);
        }
    const int m = (int)UU.size();
    std::vector<double> Pn(m * m), Ps(m * m), Z, vals;
    RCHK(block_gram(h, UU.data(), m, YY.data(), m, Pn.data()));
    for (int a = 0; a < m; ++a)
        for (int bb = 0; bb < m; ++bb) Ps[a * m + bb] = (Pn[bb * m + a] + Pn[a * m + bb]) * 0.5;
    sym_eig(m, Ps, vals, Z);
    // T = eta1^T UU (:402-403)
    std::vector<double> coef(m * k);
    for (int a = 0; a < m; ++a)
        for (int i = 0; i < k; ++i) coef[a * k + i] = Z[a * m + i];
    RCHK(block_combine(h, UU.data(), m, T.data(), k, coef.data(), nullptr));
    // Vp <- V, Up <- U (:391-400); V <- T
    std::vector<void *> oldVp = Vp, oldUp = Up;
    Vp = V;
    Up = U;
    V = T;
    T = oldVp;
    U = oldUp;
    // mean removal (:429-441), then orthonormalise in order (:443-449)
    std::vector<double> mean(k), eye(k * k, 0.0), shift(k);
    RCHK(block_means(h, cv(V).data(), k, mean.data()));
    for (int i = 0; i < k; ++i) {
        eye[i * k + i] = 1.0;
        shift[i] = -mean[i];
    }
    RCHK(block_combine(h, cv(V).data(), k, V.data(), k, eye.data(), shift.data()));
    // in order, each against the earlier ones and V0 (GOSD.cpp:344-350); the earlier
    // ones are orthogonal to V0, so this is V0's projection, then CholQR
    RCHK(block_ortho(h, V.data(), k, k0 ? b.V0.data() : nullptr, k0, -1.0, keepV));
    if (ritz)
        for (int i = 0; i < k; ++i) ritz[i] = vals[i];
    return SPM_OK;
}

// V back to the caller's fields at Vout when the last Ritz update left it in the workspace
static int lob_copy_out(spm_ctx *h, LobBlock &b, void *Vout)
{
    Ctx &c = h->c;
    for (int i = 0; i < (int)b.V.size(); ++i)
        if (b.V[i] != fld(Vout, i, c)) RCHK(field_copy(h, fld(Vout, i, c), b.V[i]));
    return SPM_OK;
}

// hiosd_lobpcg_inip (LOBPCG.cpp:207-487)
int lobpcg_run(spm_ctx *h, const void *phi, void *Vout, int k, int maxiter, double *ritz)
{
    Ctx &c = h->c;
    LobBlock b;
    RCHK(lob_block(h, Vout, k, 0, b));
    RCHK(block_ortho(h, b.V.data(), k, nullptr, 0, -1.0, b.keepV())); // :262-267
    for (int iter = 0; iter < maxiter; ++iter) RCHK(lobpcg_sweep(h, phi, b, ritz));
    RCHK(lob_copy_out(h, b, Vout));
    HIPCHK(hipStreamSynchronize(c.stream));
    return spm_launch_status("spm_lobpcg");
}

// hiosd_lobpcg_sieh (HISD.cpp:86-485): high-index saddle dynamics of x with the
// k unstable directions V, one LOBPCG sweep per step at the new x.
//   f = ngrad_cam(x); tmpp = 2 sum_i <f, V_i> V_i; g = f - tmpp;
//   bta = |<Dx, Dg>| / <Dg, Dg> (Barzilai-Borwein, Dg = g - g_prev, Dx = x - x_prev),
//         clamped to [betau dt, betat dt];
//   x <- the semi-implicit step with force f - tmpp (SOpHisdStep: real part of the spectrum);
//   then the sweep, f = ngrad_cam(x), stop when nrmp(f) < epsf or after maxiter steps.
// The first Dx is f - 2 tmpp + dt (the reference adds the constant dt, :185).
// With k0 > 0 fixed fields V0 this is hiosd_lobpcg_sieh_V0 (GOSD.cpp:123-531):
// the sweep keeps W, Vp and V orthogonal to V0 as well.
int hisd_run(spm_ctx *h, void *x, void *Vout, int k, const void *V0, int k0, double dt, double betat, double betau,
             double epsf, int maxiter, int *iters, double *res_out)
{
    Ctx &c = h->c;
    LobBlock b;
    RCHK(lob_block(h, Vout, k, 7, b));
    for (int i = 0; i < k0; ++i) b.V0.push_back(fld((void *)V0, i, c));
    void *f = fld(h->ws, 6 * k + 0, c), *gp = fld(h->ws, 6 * k + 2, c); // field 6k + 1: unused
    void *Dx = fld(h->ws, 6 * k + 3, c), *Dg = fld(h->ws, 6 * k + 4, c), *tmpp = fld(h->ws, 6 * k + 5, c);
    void *xp = fld(h->ws, 6 * k + 6, c);
    RCHK(block_ortho(h, b.V.data(), k, nullptr, 0, -1.0, b.keepV())); // :151-156
    // tmpp = s sum_i <f, V_i> V_i  (PV_Cplx, BasicOperators.cpp:360-371)
    auto pv = [&](double s) -> int {
        std::vector<double> a(k);
        const void *fv[1] = {f};
        RCHK(block_gram(h, cv(b.V).data(), k, fv, 1, a.data()));
        for (double &ai : a) ai *= s;
        void *out[1] = {tmpp};
        return block_combine(h, cv(b.V).data(), k, out, 1, a.data(), nullptr);
    };
    // out = a1 in1 + a2 in2 (+ shift)
    auto lin2 = [&](void *out, const void *in1, double a1, const void *in2, double a2, double shift) -> int {
        const void *in[2] = {in1, in2};
        void *o[1] = {out};
        const double coef[2] = {a1, a2};
        return block_combine(h, in, 2, o, 1, coef, shift != 0.0 ? &shift : nullptr);
    };
    c.g->ngrad(c, x, f);
    RCHK(pv(1.0));
    RCHK(lin2(Dx, f, 1.0, tmpp, -2.0, dt));
    double res = 0.0;
    int iter = 0;
    while (iter < maxiter) {
        RCHK(pv(2.0));
        {   // g = f - tmpp, Dg = g - gp, gp <- g in one pass (:199-204); each output sums
            // its inputs in the order f, tmpp, gp, so Dg rounds as (f - tmpp) - gp
            const void *in[3] = {f, tmpp, gp};
            void *o[2] = {Dg, gp};
            const double coef[6] = {1.0, 1.0, -1.0, -1.0, -1.0, 0.0};
            RCHK(block_combine(h, in, 3, o, 2, coef, nullptr));
        }
        double dd[2];
        {
            const void *X[2] = {Dx, Dg}, *Yv[1] = {Dg};
            RCHK(block_gram(h, X, 2, Yv, 1, dd));
        }
        double bta = std::fabs(dd[0] / dd[1]);
        bta = std::min(bta, betat * dt);
        bta = std::max(bta, betau * dt);
        RCHK(field_copy(h, xp, x));
        c.g->hisd_step(c, x, tmpp, bta);
        RCHK(lin2(Dx, x, 1.0, xp, -1.0, 0.0));
        RCHK(lobpcg_sweep(h, x, b, nullptr));
        c.g->ngrad(c, x, f);
        double g2;
        RCHK(spm_read_scalars(h, c.scal + SC_GN2, 1, &g2));
        res = std::sqrt(g2);
        iter = iter + 1;
        if (res < epsf) break;
    }
    RCHK(lob_copy_out(h, b, Vout));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (iters) *iters = iter;
    if (res_out) *res_out = res;
    return spm_launch_status("spm_hisd");
}

} // namespace spm

extern "C" {

int spm_lobpcg_init(spm_ctx *h, const void *start, void *V, int k)
{
    CHECK_CTX(h);
    if (!start || !V) return spm_fail(SPM_EINVAL, "spm_lobpcg_init: null field");
    if (k < 1 || k > 8) return spm_fail(SPM_EINVAL, "spm_lobpcg_init: block size k must be 1..8");
    return lobpcg_init(h, start, V, k);
}

int spm_lobpcg(spm_ctx *h, const void *phi, void *V, int k, int maxiter, double *ritz)
{
    CHECK_CTX(h);
    if (!phi || !V) return spm_fail(SPM_EINVAL, "spm_lobpcg: null field");
    if (k < 1 || k > 8) return spm_fail(SPM_EINVAL, "spm_lobpcg: block size k must be 1..8");
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_lobpcg: maxiter >= 0");
    return lobpcg_run(h, phi, V, k, maxiter, ritz);
}

int spm_hisd(spm_ctx *h, void *x, void *V, int k, double dt, double betat, double betau, double epsf, int maxiter,
             int *iters, double *res)
{
    CHECK_CTX(h);
    if (!x || !V || x == V) return spm_fail(SPM_EINVAL, "spm_hisd: bad fields");
    if (k < 1 || k > 8) return spm_fail(SPM_EINVAL, "spm_hisd: index k must be 1..8");
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_hisd: maxiter >= 0");
    if (!(dt > 0) || !(betau > 0) || !(betat >= betau))
        return spm_fail(SPM_EINVAL, "spm_hisd: need dt > 0 and 0 < betau <= betat");
    return hisd_run(h, x, V, k, nullptr, 0, dt, betat, betau, epsf, maxiter, iters, res);
}

int spm_gosd(spm_ctx *h, void *x, void *V, int k, const void *V0, int k0, double dt, double betat, double betau,
             double epsf, int maxiter, int *iters, double *res)
{
    CHECK_CTX(h);
    if (!x || !V || x == V) return spm_fail(SPM_EINVAL, "spm_gosd: bad fields");
    if (k < 1 || k > 8) return spm_fail(SPM_EINVAL, "spm_gosd: index k must be 1..8");
    if (k0 < 0 || 2 * k + k0 > 16) return spm_fail(SPM_EINVAL, "spm_gosd: need k0 >= 0 and 2 k + k0 <= 16");
    if (k0 > 0) {
        if (!V0) return spm_fail(SPM_EINVAL, "spm_gosd: null V0");
        // V0 is read throughout: it may not overlap x or V
        const char *a = (const char *)V0, *ae = a + (size_t)k0 * c.R * c.esz;
        const char *xv = (const char *)x, *xe = xv + (size_t)c.R * c.esz;
        const char *v = (const char *)V, *ve = v + (size_t)k * c.R * c.esz;
        if ((a < xe && xv < ae) || (a < ve && v < ae)) return spm_fail(SPM_EINVAL, "spm_gosd: V0 overlaps x or V");
    }
    if (maxiter < 0) return spm_fail(SPM_EINVAL, "spm_gosd: maxiter >= 0");
    if (!(dt > 0) || !(betau > 0) || !(betat >= betau))
        return spm_fail(SPM_EINVAL, "spm_gosd: need dt > 0 and 0 < betau <= betat");
    return hisd_run(h, x, V, k, k0 ? V0 : nullptr, k0, dt, betat, betau, epsf, maxiter, iters, res);
}

int spm_gram(spm_ctx *h, const void *X, int ma, const void *Y, int mb, double *out)
{
    CHECK_CTX(h);
    if (!X || !Y || !out) return spm_fail(SPM_EINVAL, "spm_gram: null argument");
    if (ma < 1 || mb < 1 || ma > BLK_MAX || mb > BLK_MAX) return spm_fail(SPM_EINVAL, "spm_gram: 1..32 fields");
    std::vector<const void *> xp(ma), yp(mb);
    for (int a = 0; a < ma; ++a) xp[a] = fld((void *)X, a, c);
    for (int b = 0; b < mb; ++b) yp[b] = fld((void *)Y, b, c);
    return block_gram(h, xp.data(), ma, yp.data(), mb, out);
}

int spm_rayleighq(spm_ctx *h, const void *phi, const void *v, double *out)
{
    CHECK_CTX(h);
    if (!phi || !v || !out) return spm_fail(SPM_EINVAL, "spm_rayleighq: null argument");
    RCHK(ensure_ws(h, (size_t)c.R * c.esz));
    c.g->hv(c, phi, v, h->ws);
    const void *X[1] = {v};
    const void *Y[2] = {h->ws, v};
    double s[2];
    RCHK(block_gram(h, X, 1, Y, 2, s));
    *out = s[0] / s[1];
    return SPM_OK;
}

} // extern "C"
