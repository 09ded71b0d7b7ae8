// block_ops.h -- operations on small blocks of fields (host side): the Gram
// matrix X^T Y on the matrix cores, linear combinations, means.  Used by the
// LOBPCG and SPM drivers.
#pragma once
#include "capi_internal.h"

namespace spm {

constexpr int BLK_MAX = 32; // fields per block operand

struct FieldPtrs {
    const void *p[BLK_MAX];
};
struct FieldPtrsW {
    void *p[BLK_MAX];
};

// out[a*mb + b] = inpfp(X_a, Y_b) = (1/R) sum_r X_a[r] Y_b[r]  (FP64 accumulation;
// BasicOperators.cpp:132-143 for every pair), host array, synchronises.
int block_gram(spm_ctx *h, const void *const *X, int ma, const void *const *Y, int mb, double *out);
// One pass over a Rayleigh-Ritz basis X (mx <= 24 fields) and Y (mx + my <= 48):
// A[a*my + b] = inpfp(X_a, Y_b), B[a*mx + b] = inpfp(X_a, X_b), mean[a] = mean of X_a
// (host arrays), synchronises.
int block_gram_xy(spm_ctx *h, const void *const *X, int mx, const void *const *Y, int my, double *A, double *B,
                  double *mean);
// out_i = sum_a coef[a*nout + i] in_a + shift[i] (shift may be NULL); in-place allowed
// (in one launch, every input read once, for up to 16 outputs of up to 24 inputs)
int block_combine(spm_ctx *h, const void *const *in, int nin, void *const *out, int nout, const double *coef,
                  const double *shift);
// out[a] = mean of X_a (host), synchronises
int block_means(spm_ctx *h, const void *const *X, int m, double *out);
// copy field src -> dst
int field_copy(spm_ctx *h, void *dst, const void *src);
int field_zero(spm_ctx *h, void *dst);
// 64 device doubles of the block scratch that no block operation uses
double *block_small(spm_ctx *h);

// Cholesky factor R of the symmetric n x n S (row-major) with the reference's
// sequential deflation: column i is dropped (keep[i] = false, row and column
// of Rinv zero) when its remaining norm is <= tol (deflate) or not positive.
// Rinv = R^-1 on the kept set, so that Q R^-1 has orthonormal kept columns;
// piv[i] (may be NULL) = the remaining squared norm of column i.
void chol_inv(int n, const double *S, double tol, bool deflate, bool *keep, double *Rinv, double *piv);

// Q <- Q - B (B^T Q), two rounds (B orthonormal: Q ends orthogonal to B)
int block_project(spm_ctx *h, void *const *Q, int nq, const void *const *B, int nb);

// Orthonormalise the fields Q[0..nq) in place, in order, as the reference's
// modified Gram-Schmidt does (Orthoi_v2V / Orthoi_Vi + nrmp + scaling,
// BasicOperators.cpp:229-255): first against the orthonormal fields
// B[0..nb) (which may include zero fields), then each against the earlier
// ones.  Field i is deflated (set to zero, keep[i] = false) when its norm
// after projection is <= tol (tol < 0: never).  Carried out as two rounds of
// classical projection and Cholesky QR on FP64 Gram matrices (CholQR2).
int block_ortho(spm_ctx *h, void *const *Q, int nq, const void *const *B, int nb, double tol, bool *keep);

} // namespace spm
