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

// hiosd_lobpcg_inip (LOBPCG.cpp:207-487)
int lobpcg_run(spm_ctx *h, const void *phi, void *Vout, int k, int maxiter, double *ritz)
{
    Ctx &c = h->c;
    const size_t fb = (size_t)c.R * c.esz;
    RCHK(ensure_ws(h, 6 * (size_t)k * fb));
    // buffers: V (caller), U, W, Y, Vp, Up, T in the workspace
    std::vector<void *> V(k), U(k), W(k), Y(k), Vp(k), Up(k), T(k);
    for (int i = 0; i < k; ++i) {
        V[i] = fld(Vout, i, c);
        U[i] = fld(h->ws, 0 * k + i, c);
        W[i] = fld(h->ws, 1 * k + i, c);
        Y[i] = fld(h->ws, 2 * k + i, c);
        Vp[i] = fld(h->ws, 3 * k + i, c);
        Up[i] = fld(h->ws, 4 * k + i, c);
        T[i] = fld(h->ws, 5 * k + i, c);
    }
    HIPCHK(hipMemsetAsync(h->ws, 0, 6 * (size_t)k * fb, c.stream));
    auto cv = [](const std::vector<void *> &a) { return std::vector<const void *>(a.begin(), a.end()); };
    std::vector<char> keepbuf(3 * k);
    bool *keepV = (bool *)keepbuf.data(), *keepW = keepV + k, *keepP = keepW + k;
    RCHK(block_ortho(h, V.data(), k, nullptr, 0, -1.0, keepV)); // :262-267
    std::vector<double> vals;
    for (int iter = 0; iter < maxiter; ++iter) {
        for (int i = 0; i < k; ++i) c.g->hv(c, phi, V[i], U[i]);
        RCHK(spm_launch_status("lobpcg hv"));
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
        RCHK(block_ortho(h, W.data(), k, cv(V).data(), k, 1e-8, keepW));
        for (int i = 0; i < k; ++i) {
            if (keepW[i]) c.g->hv(c, phi, W[i], Y[i]);
            else RCHK(field_zero(h, Y[i]));
        }
        {
            std::vector<const void *> B(2 * k);
            for (int i = 0; i < k; ++i) {
                B[i] = V[i];
                B[k + i] = W[i];
            }
            RCHK(block_ortho(h, Vp.data(), k, B.data(), 2 * k, 1e-8, keepP));
        }
        for (int i = 0; i < k; ++i) {
            if (keepP[i]) c.g->hv(c, phi, Vp[i], Up[i]);
            else RCHK(field_zero(h, Up[i]));
        }
        RCHK(spm_launch_status("lobpcg hv"));
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
        std::vector<double> Pn(m * m), Ps(m * m), Z;
        RCHK(block_gram(h, UU.data(), m, YY.data(), m, Pn.data()));
        for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) Ps[a * m + b] = (Pn[b * m + a] + Pn[a * m + b]) * 0.5;
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
        RCHK(block_ortho(h, V.data(), k, nullptr, 0, -1.0, keepV));
        if (ritz)
            for (int i = 0; i < k; ++i) ritz[i] = vals[i];
    }
    for (int i = 0; i < k; ++i)
        if (V[i] != fld(Vout, i, c)) RCHK(field_copy(h, fld(Vout, i, c), V[i]));
    HIPCHK(hipStreamSynchronize(c.stream));
    return spm_launch_status("spm_lobpcg");
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
