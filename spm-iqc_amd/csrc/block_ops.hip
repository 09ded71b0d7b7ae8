// block_ops.hip -- Gram matrices on the matrix cores, linear combinations and
// means of small blocks of fields, and the CholQR2 orthonormalisation built
// on them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "block_ops.h"

namespace spm {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int GRAM_BLOCKS = 2048; // MFMA Gram blocks of 4 waves (8 per SIMD); fixed, so the summation order is fixed
constexpr int GRAM_BLOCKS_SMALL = 1024; // streaming Gram: 4 waves per SIMD by its registers
constexpr int GRAM_TILE = 32;     // the Gram is computed as a padded 32 x 32 block

// Each wave accumulates C = X^T Y over its 16-point chunks with
// v_mfma_f64_16x16x4f64 (A: lane l holds A[l%16][l/16]; B: B[l/16][l%16];
// D: lane l holds D[l/16 + 4v][l%16], checked on gfx950).  The K index of a chunk is permuted so
// that lane l reads 4 consecutive points r0 + 4(l/16) .. +3 of its field in
// one 32-byte (FP64) or 16-byte (FP32) load; the sum over K is unchanged.
// SAME (Y is X, TA = TB): A and B of a tile pair take the same K mapping, so
// lane l's B values of tile t are its A values of tile t and Y is not read.
template <typename T, int TA, int TB, bool SAME>
__global__ void __launch_bounds__(256) k_gram(FieldPtrs X, int ma, FieldPtrs Y, int mb, long R,
                                             double *__restrict__ part)
{
    static_assert(!SAME || TA == TB, "k_gram: a block against itself has square tiles");
    __shared__ double sh[4][TA * TB * 4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    d4 acc[TA][TB];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) acc[ta][tb] = d4{0.0, 0.0, 0.0, 0.0};
    const T *xa[TA], *yb[TB];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) xa[ta] = (ta * 16 + row < ma) ? (const T *)X.p[ta * 16 + row] : nullptr;
#pragma unroll
    for (int tb = 0; tb < TB; ++tb)
        yb[tb] = (!SAME && tb * 16 + row < mb) ? (const T *)Y.p[tb * 16 + row] : nullptr;
    const long nchunk = R / 16;
    // lane data of chunk q: 4 consecutive points r0 + 4 (l / 16) .. +3 of its row's field,
    // in the field precision (FP32 fields: half the registers in flight), widened to
    // FP64 (exactly) where an MFMA takes it
    auto load = [&](long q, T (&a)[TA][4], T (&b)[TB][4]) {
        const long r0 = q * 16 + 4 * kq;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
            if (xa[ta]) {
                if constexpr (sizeof(T) == 8) {
                    const double2 u = *(const double2 *)(xa[ta] + r0), w = *(const double2 *)(xa[ta] + r0 + 2);
                    a[ta][0] = u.x; a[ta][1] = u.y; a[ta][2] = w.x; a[ta][3] = w.y;
                } else {
                    const float4 u = *(const float4 *)(xa[ta] + r0);
                    a[ta][0] = u.x; a[ta][1] = u.y; a[ta][2] = u.z; a[ta][3] = u.w;
                }
            } else {
                a[ta][0] = a[ta][1] = a[ta][2] = a[ta][3] = (T)0;
            }
        }
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
            if constexpr (SAME) continue;
            if (yb[tb]) {
                if constexpr (sizeof(T) == 8) {
                    const double2 u = *(const double2 *)(yb[tb] + r0), w = *(const double2 *)(yb[tb] + r0 + 2);
                    b[tb][0] = u.x; b[tb][1] = u.y; b[tb][2] = w.x; b[tb][3] = w.y;
                } else {
                    const float4 u = *(const float4 *)(yb[tb] + r0);
                    b[tb][0] = u.x; b[tb][1] = u.y; b[tb][2] = u.z; b[tb][3] = u.w;
                }
            } else {
                b[tb][0] = b[tb][1] = b[tb][2] = b[tb][3] = (T)0;
            }
        }
    };
    // one chunk ahead: the next chunk's loads are in flight during this chunk's MFMAs,
    // except for the FP64 4-tile Grams, which hold one more wave per SIMD without the
    // second set
    constexpr bool PF = !(sizeof(T) == 8 && TA * TB >= 4);
    const long step = (long)gridDim.x * 4;
    long q = (long)blockIdx.x * 4 + wid;
    T a[TA][4], b[TB][4];
    if (PF && q < nchunk) load(q, a, b);
    for (; q < nchunk; q += step) {
        T an[TA][4], bn[TB][4];
        if constexpr (PF) {
            if (q + step < nchunk) load(q + step, an, bn);
        } else {
            load(q, a, b);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int ta = 0; ta < TA; ++ta)
#pragma unroll
                for (int tb = 0; tb < TB; ++tb)
                    acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        (double)a[ta][s], (double)(SAME ? a[tb][s] : b[tb][s]), acc[ta][tb], 0, 0, 0);
        if constexpr (PF) {
#pragma unroll
            for (int ta = 0; ta < TA; ++ta)
#pragma unroll
                for (int s = 0; s < 4; ++s) a[ta][s] = an[ta][s];
#pragma unroll
            for (int tb = 0; tb < TB; ++tb)
#pragma unroll
                for (int s = 0; s < 4; ++s) b[tb][s] = bn[tb][s];
        }
    }
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
#pragma unroll
            for (int v = 0; v < 4; ++v) sh[wid][(ta * TB + tb) * 4 + v][lane] = acc[ta][tb][v];
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int ta = 0; ta < TA; ++ta)
#pragma unroll
            for (int tb = 0; tb < TB; ++tb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int e = (ta * TB + tb) * 4 + v;
                    const double s = ((sh[0][e][lane] + sh[1][e][lane]) + sh[2][e][lane]) + sh[3][e][lane];
                    const int ai = ta * 16 + kq + 4 * v, bj = tb * 16 + row;
                    part[(long)blockIdx.x * GRAM_TILE * GRAM_TILE + ai * GRAM_TILE + bj] = s;
                }
    }
}

// The Gram of blocks of at most SMALL_M fields, where an MFMA tile would keep
// only ma + mb of its 16 rows loading (a latency-bound stream).  Each thread
// sums 4 consecutive points per field and iteration into ma x mb FP64
// accumulators; the wave sums by a fixed butterfly, the 4 waves in a fixed
// order, and the block writes its partials in k_gram's layout (entry a*32 + b),
// so k_gram_reduce finishes both kernels.  SAME: Y is X (a block against
// itself), read once.
constexpr int SMALL_M = 4;
template <typename T, bool SAME>
__global__ void __launch_bounds__(256) k_gram_small(FieldPtrs X, int ma, FieldPtrs Y, int mb, long R,
                                                   double *__restrict__ part)
{
    __shared__ double sh[4][SMALL_M * SMALL_M];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[SMALL_M][SMALL_M];
#pragma unroll
    for (int a = 0; a < SMALL_M; ++a)
#pragma unroll
        for (int b = 0; b < SMALL_M; ++b) acc[a][b] = 0.0;
    auto load4 = [&](const void *f, long r, double (&v)[4]) {
        if constexpr (sizeof(T) == 8) {
            const double2 u = *((const double2 *)f + r / 2), w = *((const double2 *)f + r / 2 + 1);
            v[0] = u.x; v[1] = u.y; v[2] = w.x; v[3] = w.y;
        } else {
            const float4 u = *((const float4 *)f + r / 4);
            v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
        }
    };
    const long stride = (long)gridDim.x * blockDim.x * 4;
#pragma unroll 1 // 102 VGPRs, 4 waves per SIMD (unrolled twice: 174, 2 waves, FP32 Grams slower)
    for (long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4; r < R; r += stride) {
        double x[SMALL_M][4], y[SMALL_M][4];
#pragma unroll
        for (int a = 0; a < SMALL_M; ++a)
            if (a < ma) load4(X.p[a], r, x[a]);
#pragma unroll
        for (int b = 0; b < SMALL_M; ++b)
            if (!SAME && b < mb) load4(Y.p[b], r, y[b]);
#pragma unroll
        for (int a = 0; a < SMALL_M; ++a)
#pragma unroll
            for (int b = 0; b < SMALL_M; ++b)
                if (a < ma && b < mb) {
                    const double(&yb)[4] = SAME ? x[b] : y[b];
                    acc[a][b] += ((x[a][0] * yb[0] + x[a][1] * yb[1]) + x[a][2] * yb[2]) + x[a][3] * yb[3];
                }
    }
#pragma unroll
    for (int a = 0; a < SMALL_M; ++a)
#pragma unroll
        for (int b = 0; b < SMALL_M; ++b) {
            if (a >= ma || b >= mb) continue;
            double v = acc[a][b];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) sh[wid][a * SMALL_M + b] = v;
        }
    __syncthreads();
    if (threadIdx.x < SMALL_M * SMALL_M) {
        const int a = threadIdx.x / SMALL_M, b = threadIdx.x % SMALL_M;
        if (a < ma && b < mb)
            part[(long)blockIdx.x * GRAM_TILE * GRAM_TILE + a * GRAM_TILE + b] =
                ((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) + sh[3][threadIdx.x];
    }
}

// out[e] = scale * sum over blocks of part[b][e], for the 64 entries of this
// block: thread group g (16 groups of 64) sums the blocks b = g (mod 16) in
// order, then the 16 group sums combine by a fixed pairwise tree, so the
// result does not depend on timing.
__global__ void __launch_bounds__(1024) k_gram_reduce(const double *__restrict__ part, int nblk, long epb,
                                                     double scale, double *__restrict__ out)
{
    __shared__ double sh[16][64];
    const int el = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + el;
    double s = 0.0;
#pragma unroll 8
    for (int b = g; b < nblk; b += 16) s += part[(long)b * epb + e];
    sh[g][el] = s;
    __syncthreads();
#pragma unroll
    for (int w = 8; w > 0; w >>= 1) {
        if (g < w) sh[g][el] += sh[g + w][el];
        __syncthreads();
    }
    if (g == 0) out[e] = sh[0][el] * scale;
}

// The Gram of a Rayleigh-Ritz basis X (mx <= 16 TA fields) against [X, Y] (Y:
// my fields, mx + my <= 16 TB), and the means of [X, Y] (ONES: row mx of the
// row operand is the constant 1), in one pass that reads each field once.
// Column c < mx is X_c, which is row c of the same tile's row operand (tile
// rows and tile columns line up in the lanes), so it is not loaded again.
// The MFMA layout and the fixed summation order are those of k_gram; the
// partials of block b are the TA 16 x TB 16 entries at part + b (TA 16)(TB 16).
template <typename T, int TA, int TB>
__global__ void __launch_bounds__(256) k_gram_xy(FieldPtrs X, int mx, FieldPtrs Y, int my, int ones, long R,
                                                double *__restrict__ part)
{
    static_assert(TB >= TA, "the columns start with X");
    constexpr int CW = TB * 16;
    __shared__ double sh[4][TA * TB * 4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    d4 acc[TA][TB];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) acc[ta][tb] = d4{0.0, 0.0, 0.0, 0.0};
    const T *xa[TA], *yb[TB];
    bool one[TA], reuse[TB];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) {
        const int r = ta * 16 + row;
        xa[ta] = r < mx ? (const T *)X.p[r] : nullptr;
        one[ta] = ones && r == mx;
    }
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
        const int cc = tb * 16 + row;
        reuse[tb] = tb < TA && cc < mx;
        yb[tb] = (cc >= mx && cc - mx < my) ? (const T *)Y.p[cc - mx] : nullptr;
    }
    // Loads keep the field precision (FP32 fields: half the registers in flight);
    // each value is widened to FP64 (exactly) where an MFMA takes it.
    auto ld4 = [&](const T *f, long r0, T (&v)[4]) {
        if constexpr (sizeof(T) == 8) {
            const double2 u = *(const double2 *)(f + r0), w = *(const double2 *)(f + r0 + 2);
            v[0] = u.x; v[1] = u.y; v[2] = w.x; v[3] = w.y;
        } else {
            const float4 u = *(const float4 *)(f + r0);
            v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
        }
    };
    const long nchunk = R / 16;
    auto load = [&](long q, T (&a)[TA][4], T (&b)[TB][4]) {
        const long r0 = q * 16 + 4 * kq;
#pragma unroll
        for (int ta = 0; ta < TA; ++ta) {
            if (xa[ta]) ld4(xa[ta], r0, a[ta]);
            else a[ta][0] = a[ta][1] = a[ta][2] = a[ta][3] = one[ta] ? (T)1 : (T)0;
        }
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
            if (yb[tb]) ld4(yb[tb], r0, b[tb]);
            else b[tb][0] = b[tb][1] = b[tb][2] = b[tb][3] = (T)0;
        }
    };
    const long step = (long)gridDim.x * 4;
    long q = (long)blockIdx.x * 4 + wid;
    // the next chunk's loads in flight during this chunk's MFMAs, except for the
    // FP64 6-tile Gram, which holds a third wave per SIMD without the second set
    constexpr bool PF = !(sizeof(T) == 8 && TA * TB >= 6);
    T a[TA][4], b[TB][4];
    if (PF && q < nchunk) load(q, a, b);
    for (; q < nchunk; q += step) {
        T an[TA][4], bn[TB][4];
        if constexpr (PF) {
            if (q + step < nchunk) load(q + step, an, bn);
        } else {
            load(q, a, b);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int tb = 0; tb < TB; ++tb) {
                const double bv = (double)((tb < TA && reuse[tb]) ? a[tb < TA ? tb : 0][s] : b[tb][s]);
#pragma unroll
                for (int ta = 0; ta < TA; ++ta)
                    acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[ta][s], bv, acc[ta][tb], 0, 0, 0);
            }
        if constexpr (PF) {
#pragma unroll
            for (int ta = 0; ta < TA; ++ta)
#pragma unroll
                for (int s = 0; s < 4; ++s) a[ta][s] = an[ta][s];
#pragma unroll
            for (int tb = 0; tb < TB; ++tb)
#pragma unroll
                for (int s = 0; s < 4; ++s) b[tb][s] = bn[tb][s];
        }
    }
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
#pragma unroll
            for (int v = 0; v < 4; ++v) sh[wid][(ta * TB + tb) * 4 + v][lane] = acc[ta][tb][v];
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int ta = 0; ta < TA; ++ta)
#pragma unroll
            for (int tb = 0; tb < TB; ++tb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int e = (ta * TB + tb) * 4 + v;
                    const double sum = ((sh[0][e][lane] + sh[1][e][lane]) + sh[2][e][lane]) + sh[3][e][lane];
                    const int ai = ta * 16 + kq + 4 * v, bj = tb * 16 + row;
                    part[(long)blockIdx.x * (TA * 16) * CW + ai * CW + bj] = sum;
                }
    }
}

// k_gram_xy for X and Y of at most XYS fields each (the sweeps of HiSD with one
// or two directions), where an MFMA tile would keep only a few of its rows
// loading: a streaming VALU pass, four points per thread, every field read
// once.  Partials in k_gram_xy<T, 1, 1>'s layout (entry a * 16 + b of the
// 16 x 16 tile: rows X and the ones row mx, columns [X, Y]), so the same
// reduction finishes it.
constexpr int XYS = 3;
template <typename T>
__global__ void __launch_bounds__(256) k_gram_xy_small(FieldPtrs X, int mx, FieldPtrs Y, int my, long R,
                                                      double *__restrict__ part)
{
    constexpr int NE = XYS * (2 * XYS) + 2 * XYS; // X^T [X Y] and 1^T [X Y]
    __shared__ double sh[4][NE];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.0;
    auto load4 = [&](const void *f, long r, double (&v)[4]) {
        if constexpr (sizeof(T) == 8) {
            const double2 u = *((const double2 *)f + r / 2), w = *((const double2 *)f + r / 2 + 1);
            v[0] = u.x; v[1] = u.y; v[2] = w.x; v[3] = w.y;
        } else {
            const float4 u = *((const float4 *)f + r / 4);
            v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
        }
    };
    const long stride = (long)gridDim.x * blockDim.x * 4;
#pragma unroll 1
    for (long r = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4; r < R; r += stride) {
        double c[2 * XYS][4]; // columns [X, Y]
#pragma unroll
        for (int b = 0; b < 2 * XYS; ++b) {
            const bool isx = b < XYS;
            const int bi = isx ? b : b - XYS;
            if (bi < (isx ? mx : my)) load4(isx ? X.p[bi] : Y.p[bi], r, c[b]);
            else c[b][0] = c[b][1] = c[b][2] = c[b][3] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < XYS; ++a)
#pragma unroll
            for (int b = 0; b < 2 * XYS; ++b)
                acc[a * 2 * XYS + b] += ((c[a][0] * c[b][0] + c[a][1] * c[b][1]) + c[a][2] * c[b][2]) + c[a][3] * c[b][3];
#pragma unroll
        for (int b = 0; b < 2 * XYS; ++b) acc[XYS * 2 * XYS + b] += ((c[b][0] + c[b][1]) + c[b][2]) + c[b][3];
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        double v = acc[e];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) sh[wid][e] = v;
    }
    __syncthreads();
    // the 16 x 16 tile of this block: row a < mx (X_a) or a = mx (ones), column b < mx (X_b) or mx + b (Y_b)
    double *pb = part + (long)blockIdx.x * 256;
    for (int t = threadIdx.x; t < 256; t += blockDim.x) {
        const int a = t / 16, b = t % 16;
        double v = 0.0;
        const int ra = a < mx ? a : (a == mx ? XYS : -1);
        const int cb = b < mx ? b : (b - mx < my ? XYS + (b - mx) : -1);
        if (ra >= 0 && cb >= 0) {
            const int e = ra * 2 * XYS + cb;
            v = ((sh[0][e] + sh[1][e]) + sh[2][e]) + sh[3][e];
        }
        pb[t] = v;
    }
}

constexpr int COMBINE_NO = 8; // outputs per k_combine launch

// Operands of one k_combine launch, passed by value (read with uniform scalar
// loads): no upload, so a combine queues behind earlier work without draining
// the stream.  Two sizes: up to 32 inputs and 8 outputs (2.4 KB of kernel
// arguments), and up to 24 inputs and 16 outputs (3.5 KB).
template <int NI_, int NO_> struct CombineArgsT {
    static constexpr int NI = NI_, NO = NO_;
    const void *in[NI];
    void *out[NO];
    double coef[NI * NO]; // coef[a * NO + o]
    double shift[NO];
};
using CombineArgs = CombineArgsT<BLK_MAX, COMBINE_NO>;
using CombineArgsWide = CombineArgsT<24, 16>;

// out_o = sum_a coef[a][o] in_a + shift[o], o < NO: PT consecutive points per
// thread (4, or 2 for more than 8 outputs, which keeps the accumulators at 64
// registers), all inputs read once per point (so an output may alias an input).
template <typename T, int PT> struct PtVec;
template <> struct PtVec<double, 4> {
    static __device__ __forceinline__ void ld(const void *p, long i, double (&x)[4])
    {
        const double2 u = ((const double2 *)p)[2 * i], w = ((const double2 *)p)[2 * i + 1];
        x[0] = u.x; x[1] = u.y; x[2] = w.x; x[3] = w.y;
    }
    static __device__ __forceinline__ void st(void *p, long i, const double (&x)[4])
    {
        ((double2 *)p)[2 * i] = double2{x[0], x[1]};
        ((double2 *)p)[2 * i + 1] = double2{x[2], x[3]};
    }
};
template <> struct PtVec<double, 2> {
    static __device__ __forceinline__ void ld(const void *p, long i, double (&x)[2])
    {
        const double2 u = ((const double2 *)p)[i];
        x[0] = u.x; x[1] = u.y;
    }
    static __device__ __forceinline__ void st(void *p, long i, const double (&x)[2])
    {
        ((double2 *)p)[i] = double2{x[0], x[1]};
    }
};
template <> struct PtVec<float, 4> {
    static __device__ __forceinline__ void ld(const void *p, long i, double (&x)[4])
    {
        const float4 u = ((const float4 *)p)[i];
        x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
    }
    static __device__ __forceinline__ void st(void *p, long i, const double (&x)[4])
    {
        ((float4 *)p)[i] = float4{(float)x[0], (float)x[1], (float)x[2], (float)x[3]};
    }
};
template <> struct PtVec<float, 2> {
    static __device__ __forceinline__ void ld(const void *p, long i, double (&x)[2])
    {
        const float2 u = ((const float2 *)p)[i];
        x[0] = u.x; x[1] = u.y;
    }
    static __device__ __forceinline__ void st(void *p, long i, const double (&x)[2])
    {
        ((float2 *)p)[i] = float2{(float)x[0], (float)x[1]};
    }
};

template <typename T, int NO, int PT, class Args>
__global__ void __launch_bounds__(256) k_combine(const Args args, int nin, long RP)
{
    static_assert(NO <= Args::NO, "outputs");
    for (long rp = (long)blockIdx.x * blockDim.x + threadIdx.x; rp < RP; rp += (long)gridDim.x * blockDim.x) {
        double acc[NO][PT];
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
            for (int k = 0; k < PT; ++k) acc[o][k] = args.shift[o];
        for (int a = 0; a < nin; ++a) {
            double x[PT];
            PtVec<T, PT>::ld(args.in[a], rp, x);
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const double cf = args.coef[a * Args::NO + o];
#pragma unroll
                for (int k = 0; k < PT; ++k) acc[o][k] += cf * x[k];
            }
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) PtVec<T, PT>::st(args.out[o], rp, acc[o]);
    }
}

constexpr int SUM_BLOCKS = 256;

template <typename T>
__global__ void __launch_bounds__(256) k_sums(FieldPtrs X, long R, double *__restrict__ part)
{
    __shared__ double red[16];
    const T *x = (const T *)X.p[blockIdx.y];
    double s = 0.0;
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (long)gridDim.x * blockDim.x)
        s += (double)x[r];
    double v[1] = {s};
    block_reduce<1, 0u>(v, red);
    if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

__global__ void k_sums_reduce(const double *__restrict__ part, int nblk, int m, double scale,
                              double *__restrict__ out)
{
    const int a = threadIdx.x;
    if (a >= m) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(long)a * nblk + b];
    out[a] = s * scale;
}

// ---------------------------------------------------------------- host side

constexpr int XY_ROWS = 32, XY_COLS = 48; // k_gram_xy: at most 2 x 3 tiles
constexpr size_t GRAM_PART_EPB = (size_t)XY_ROWS * XY_COLS; // partials per block, k_gram and k_gram_xy

static int ensure_bscr(spm_ctx *h)
{
    const size_t need = sizeof(double) * ((size_t)GRAM_BLOCKS * GRAM_PART_EPB + 4 * 4096);
    if (h->bscr_bytes >= need) return SPM_OK;
    if (h->bscr) {
        HIPCHK(hipStreamSynchronize(h->c.stream));
        HIPCHK(hipFree(h->bscr));
        h->bscr = nullptr;
    }
    RCHK(spm_dev_alloc(h, (void **)&h->bscr, need));
    h->bscr_bytes = need;
    return SPM_OK;
}
// scratch regions inside bscr
static double *gram_part(spm_ctx *h) { return h->bscr; }
static double *small_dev(spm_ctx *h) { return h->bscr + (size_t)GRAM_BLOCKS * GRAM_PART_EPB; }
double *block_small(spm_ctx *h) { return ensure_bscr(h) == SPM_OK ? small_dev(h) + 2048 : nullptr; }

int block_gram(spm_ctx *h, const void *const *X, int ma, const void *const *Y, int mb, double *out)
{
    Ctx &c = h->c;
    if (ma < 1 || mb < 1 || ma > GRAM_TILE || mb > GRAM_TILE) return spm_fail(SPM_EINVAL, "block_gram: 1..32 fields");
    RCHK(ensure_bscr(h));
    FieldPtrs xp{}, yp{};
    for (int a = 0; a < ma; ++a) xp.p[a] = X[a];
    for (int b = 0; b < mb; ++b) yp.p[b] = Y[b];
    const int ta = ma > 16 ? 2 : 1, tb = mb > 16 ? 2 : 1;
    const bool small = ma <= SMALL_M && mb <= SMALL_M;
    const int nblk = (int)std::min<long>(small ? GRAM_BLOCKS_SMALL : GRAM_BLOCKS, std::max<long>(1, c.R / 64));
    double *part = gram_part(h);
#define GRAM_LAUNCH(T, A, B, S)                                                                                \
    hipLaunchKernelGGL((k_gram<T, A, B, S>), dim3(nblk), dim3(256), 0, c.stream, xp, ma, yp, mb, c.R, part)
#define GRAM_SMALL(T, SAME)                                                                                    \
    hipLaunchKernelGGL((k_gram_small<T, SAME>), dim3(nblk), dim3(256), 0, c.stream, xp, ma, yp, mb, c.R, part)
    // R = N^3 with 4 | N: a multiple of 64, so every thread's 4-point group is whole
    const bool same = ma == mb && std::equal(X, X + ma, Y);
    if (small) {
        if (c.prec == SPM_FP64) {
            if (same) GRAM_SMALL(double, true);
            else GRAM_SMALL(double, false);
        } else {
            if (same) GRAM_SMALL(float, true);
            else GRAM_SMALL(float, false);
        }
    } else if (c.prec == SPM_FP64) {
        if (ta == 1 && tb == 1) {
            if (same) GRAM_LAUNCH(double, 1, 1, true);
            else GRAM_LAUNCH(double, 1, 1, false);
        } else if (ta == 1) GRAM_LAUNCH(double, 1, 2, false);
        else if (tb == 1) GRAM_LAUNCH(double, 2, 1, false);
        else if (same) GRAM_LAUNCH(double, 2, 2, true);
        else GRAM_LAUNCH(double, 2, 2, false);
    } else {
        if (ta == 1 && tb == 1) {
            if (same) GRAM_LAUNCH(float, 1, 1, true);
            else GRAM_LAUNCH(float, 1, 1, false);
        } else if (ta == 1) GRAM_LAUNCH(float, 1, 2, false);
        else if (tb == 1) GRAM_LAUNCH(float, 2, 1, false);
        else if (same) GRAM_LAUNCH(float, 2, 2, true);
        else GRAM_LAUNCH(float, 2, 2, false);
    }
#undef GRAM_LAUNCH
#undef GRAM_SMALL
    double *dout = small_dev(h);
    // rows 0 .. ma-1 of the 32-wide partials: (ma * 32 + 63) / 64 groups of 64 entries
    hipLaunchKernelGGL(k_gram_reduce, dim3((ma * GRAM_TILE + 63) / 64), dim3(1024), 0, c.stream, part, nblk,
                       (long)GRAM_TILE * GRAM_TILE, 1.0 / (double)c.R, dout);
    RCHK(spm_launch_status("block_gram"));
    std::vector<double> full(GRAM_TILE * GRAM_TILE);
    HIPCHK(hipMemcpyAsync(full.data(), dout, sizeof(double) * ma * GRAM_TILE, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    for (int a = 0; a < ma; ++a)
        for (int b = 0; b < mb; ++b) out[a * mb + b] = full[a * GRAM_TILE + b];
    return SPM_OK;
}

int block_gram_xy(spm_ctx *h, const void *const *X, int mx, const void *const *Y, int my, double *A, double *B,
                  double *mean)
{
    Ctx &c = h->c;
    if (mx < 1 || my < 0 || mx + 1 > XY_ROWS || mx + my > XY_COLS) return spm_fail(SPM_EINVAL, "block_gram_xy: sizes");
    RCHK(ensure_bscr(h));
    FieldPtrs xp{}, yp{};
    for (int a = 0; a < mx; ++a) xp.p[a] = X[a];
    for (int b = 0; b < my; ++b) yp.p[b] = Y[b];
    const int ta = (mx + 1 + 15) / 16, tb = std::max(ta, (mx + my + 15) / 16);
    const bool small = mx <= XYS && my <= XYS;
    const int nblk = (int)std::min<long>(small ? GRAM_BLOCKS_SMALL : GRAM_BLOCKS, std::max<long>(1, c.R / 64));
    double *part = gram_part(h);
#define XY_LAUNCH(T, A_, B_)                                                                                   \
    hipLaunchKernelGGL((k_gram_xy<T, A_, B_>), dim3(nblk), dim3(256), 0, c.stream, xp, mx, yp, my, 1, c.R, part)
#define XY_BY_TILES(T)                                                                                         \
    if (small) hipLaunchKernelGGL(k_gram_xy_small<T>, dim3(nblk), dim3(256), 0, c.stream, xp, mx, yp, my, c.R, part); \
    else if (ta == 1 && tb == 1) XY_LAUNCH(T, 1, 1);                                                           \
    else if (ta == 1 && tb == 2) XY_LAUNCH(T, 1, 2);                                                           \
    else if (ta == 1) XY_LAUNCH(T, 1, 3);                                                                      \
    else if (tb == 2) XY_LAUNCH(T, 2, 2);                                                                      \
    else XY_LAUNCH(T, 2, 3);
    if (c.prec == SPM_FP64) {
        XY_BY_TILES(double)
    } else {
        XY_BY_TILES(float)
    }
#undef XY_BY_TILES
#undef XY_LAUNCH
    const int cw = tb * 16;
    const long used = (long)(mx + 1) * cw;
    double *dout = small_dev(h);
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((used + 63) / 64)), dim3(1024), 0, c.stream, part, nblk,
                       (long)ta * 16 * cw, 1.0 / (double)c.R, dout);
    RCHK(spm_launch_status("block_gram_xy"));
    std::vector<double> full(used);
    HIPCHK(hipMemcpyAsync(full.data(), dout, sizeof(double) * used, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    for (int a = 0; a < mx; ++a) {
        for (int b = 0; b < mx; ++b) B[a * mx + b] = full[a * cw + b];
        for (int b = 0; b < my; ++b) A[a * my + b] = full[a * cw + mx + b];
    }
    for (int b = 0; b < mx; ++b) mean[b] = full[mx * cw + b];
    return SPM_OK;
}

template <class Args>
static Args combine_args(const void *const *in, int nin, void *const *out, int no, int i0, int nout,
                         const double *coef, const double *shift)
{
    Args ca{};
    for (int a = 0; a < nin; ++a) ca.in[a] = in[a];
    for (int o = 0; o < no; ++o) {
        ca.out[o] = out[i0 + o];
        ca.shift[o] = shift ? shift[i0 + o] : 0.0;
        for (int a = 0; a < nin; ++a) ca.coef[a * Args::NO + o] = coef[a * nout + i0 + o];
    }
    return ca;
}

int block_combine(spm_ctx *h, const void *const *in, int nin, void *const *out, int nout, const double *coef,
                  const double *shift)
{
    Ctx &c = h->c;
    if (nin < 1 || nout < 1 || nin > BLK_MAX || nout > BLK_MAX) return spm_fail(SPM_EINVAL, "block_combine");
    // R = N^3, N >= 4: a multiple of 4
    // one point group per thread, no grid-stride loop: loads and stores share one
    // counter on gfx9, so a loop's next loads would wait for its stores (FP32
    // LOBPCG iteration at 256^3 7.67 -> 7.50 ms)
    auto nblocks = [&](long rp) { return (int)((rp + 255) / 256); };
    if (nout > COMBINE_NO && nout <= CombineArgsWide::NO && nin <= CombineArgsWide::NI) {
        // one launch: every input read once
        const CombineArgsWide ca = combine_args<CombineArgsWide>(in, nin, out, nout, 0, nout, coef, shift);
        const long RP = c.R / 2;
#define WIDE_LAUNCH(T, K)                                                                                      \
    hipLaunchKernelGGL((k_combine<T, K, 2, CombineArgsWide>), dim3(nblocks(RP)), dim3(256), 0, c.stream, ca, nin, RP)
#define WIDE_BY_NO(T)                                                                                          \
    switch (nout) {                                                                                            \
    case 9: WIDE_LAUNCH(T, 9); break;                                                                          \
    case 10: WIDE_LAUNCH(T, 10); break;                                                                        \
    case 11: WIDE_LAUNCH(T, 11); break;                                                                        \
    case 12: WIDE_LAUNCH(T, 12); break;                                                                        \
    case 13: WIDE_LAUNCH(T, 13); break;                                                                        \
    case 14: WIDE_LAUNCH(T, 14); break;                                                                        \
    case 15: WIDE_LAUNCH(T, 15); break;                                                                        \
    default: WIDE_LAUNCH(T, 16); break;                                                                        \
    }
        if (c.prec == SPM_FP64) {
            WIDE_BY_NO(double)
        } else {
            WIDE_BY_NO(float)
        }
#undef WIDE_BY_NO
#undef WIDE_LAUNCH
        return spm_launch_status("block_combine");
    }
    constexpr int NO = COMBINE_NO;
    if (nout > NO) // later launches would read inputs that earlier ones overwrote
        for (int a = 0; a < nin; ++a)
            for (int i = 0; i < nout; ++i)
                if (in[a] == out[i]) return spm_fail(SPM_EINVAL, "block_combine: in place with more than 8 outputs");
    const long R4 = c.R / 4;
    const int nb = nblocks(R4);
    for (int i0 = 0; i0 < nout; i0 += NO) {
        const int no = std::min(NO, nout - i0);
        const CombineArgs ca = combine_args<CombineArgs>(in, nin, out, no, i0, nout, coef, shift);
#define COMBINE_LAUNCH(T, K)                                                                                   \
    hipLaunchKernelGGL((k_combine<T, K, 4, CombineArgs>), dim3(nb), dim3(256), 0, c.stream, ca, nin, R4)
#define COMBINE_BY_NO(T)                                                                                       \
    switch (no) {                                                                                              \
    case 1: COMBINE_LAUNCH(T, 1); break;                                                                       \
    case 2: COMBINE_LAUNCH(T, 2); break;                                                                       \
    case 3: COMBINE_LAUNCH(T, 3); break;                                                                       \
    case 4: COMBINE_LAUNCH(T, 4); break;                                                                       \
    case 5: COMBINE_LAUNCH(T, 5); break;                                                                       \
    case 6: COMBINE_LAUNCH(T, 6); break;                                                                       \
    case 7: COMBINE_LAUNCH(T, 7); break;                                                                       \
    default: COMBINE_LAUNCH(T, 8); break;                                                                      \
    }
        if (c.prec == SPM_FP64) {
            COMBINE_BY_NO(double)
        } else {
            COMBINE_BY_NO(float)
        }
#undef COMBINE_BY_NO
#undef COMBINE_LAUNCH
    }
    return spm_launch_status("block_combine");
}

int block_means(spm_ctx *h, const void *const *X, int m, double *out)
{
    Ctx &c = h->c;
    if (m < 1 || m > BLK_MAX) return spm_fail(SPM_EINVAL, "block_means");
    RCHK(ensure_bscr(h));
    FieldPtrs xp{};
    for (int a = 0; a < m; ++a) xp.p[a] = X[a];
    double *part = gram_part(h);
    if (c.prec == SPM_FP64)
        hipLaunchKernelGGL(k_sums<double>, dim3(SUM_BLOCKS, m), dim3(256), 0, c.stream, xp, c.R, part);
    else
        hipLaunchKernelGGL(k_sums<float>, dim3(SUM_BLOCKS, m), dim3(256), 0, c.stream, xp, c.R, part);
    double *dout = small_dev(h);
    hipLaunchKernelGGL(k_sums_reduce, dim3(1), dim3(64), 0, c.stream, part, SUM_BLOCKS, m, 1.0 / (double)c.R, dout);
    RCHK(spm_launch_status("block_means"));
    HIPCHK(hipMemcpyAsync(h->h_pin, dout, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    std::memcpy(out, h->h_pin, sizeof(double) * m);
    return SPM_OK;
}

int field_copy(spm_ctx *h, void *dst, const void *src)
{
    if (dst == src) return SPM_OK;
    HIPCHK(hipMemcpyAsync(dst, src, (size_t)h->c.R * h->c.esz, hipMemcpyDeviceToDevice, h->c.stream));
    return SPM_OK;
}

int field_zero(spm_ctx *h, void *dst)
{
    HIPCHK(hipMemsetAsync(dst, 0, (size_t)h->c.R * h->c.esz, h->c.stream));
    return SPM_OK;
}

// Cholesky factor of the Gram S (n x n, row-major) with the reference's
// sequential deflation: column i is dropped when its remaining norm is <= tol.
// Returns Rinv (upper triangular, dropped columns zero) so that Q R^{-1} has
// orthonormal kept columns.
void chol_inv(int n, const double *S, double tol, bool deflate, bool *keep, double *Rinv, double *piv)
{
    std::vector<double> Rm(n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        keep[i] = true;
        double d = S[i * n + i];
        for (int l = 0; l < i; ++l) d -= Rm[l * n + i] * Rm[l * n + i];
        if (piv) piv[i] = d;
        if (deflate ? !(d > tol * tol) : !(d > 0.0)) {
            keep[i] = false;
            continue;
        }
        const double rii = std::sqrt(d);
        Rm[i * n + i] = rii;
        for (int j = i + 1; j < n; ++j) {
            double s = S[i * n + j];
            for (int l = 0; l < i; ++l) s -= Rm[l * n + i] * Rm[l * n + j];
            Rm[i * n + j] = s / rii;
        }
    }
    // Rinv: solve R X = I on the kept index set (back substitution per column)
    std::fill(Rinv, Rinv + n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        if (!keep[j]) continue;
        for (int i = j; i >= 0; --i) {
            if (!keep[i]) continue;
            double s = (i == j) ? 1.0 : 0.0;
            for (int l = i + 1; l <= j; ++l)
                if (keep[l]) s -= Rm[i * n + l] * Rinv[l * n + j];
            Rinv[i * n + j] = s / Rm[i * n + i];
        }
    }
}

int block_project(spm_ctx *h, void *const *Q, int nq, const void *const *B, int nb)
{
    if (nq < 1 || nb < 1) return SPM_OK;
    if (nq > 16 || nb > 16) return spm_fail(SPM_EINVAL, "block_project: at most 16 + 16 fields");
    std::vector<const void *> qin(Q, Q + nq);
    // two rounds of Q <- Q - B (B^T Q)
    {
        for (int round = 0; round < 2; ++round) {
            std::vector<double> P(nb * nq);
            RCHK(block_gram(h, B, nb, qin.data(), nq, P.data()));
            // out_i = Q_i - sum_j P[j][i] B_j : inputs [Q_0..Q_{nq-1}, B_0..B_{nb-1}]
            std::vector<const void *> in(nq + nb);
            std::vector<double> coef((nq + nb) * nq, 0.0);
            for (int i = 0; i < nq; ++i) {
                in[i] = Q[i];
                coef[i * nq + i] = 1.0;
            }
            for (int j = 0; j < nb; ++j) {
                in[nq + j] = B[j];
                for (int i = 0; i < nq; ++i) coef[(nq + j) * nq + i] = -P[j * nq + i];
            }
            RCHK(block_combine(h, in.data(), nq + nb, Q, nq, coef.data(), nullptr));
        }
    }
    return SPM_OK;
}

int block_ortho(spm_ctx *h, void *const *Q, int nq, const void *const *B, int nb, double tol, bool *keep)
{
    if (nq < 1) return SPM_OK;
    if (nq > 16 || nb > 16) return spm_fail(SPM_EINVAL, "block_ortho: at most 16 + 16 fields");
    std::vector<const void *> qin(Q, Q + nq);
    RCHK(block_project(h, Q, nq, B, nb));
    // CholQR with the reference's deflation, then a second CholQR on the kept fields
    std::vector<double> S(nq * nq), Rinv(nq * nq);
    RCHK(block_gram(h, qin.data(), nq, qin.data(), nq, S.data()));
    chol_inv(nq, S.data(), tol, tol >= 0.0, keep, Rinv.data(), nullptr);
    RCHK(block_combine(h, qin.data(), nq, Q, nq, Rinv.data(), nullptr));
    std::vector<int> kept;
    for (int i = 0; i < nq; ++i)
        if (keep[i]) kept.push_back(i);
    const int nk = (int)kept.size();
    if (nk > 0) {
        std::vector<const void *> kin(nk);
        std::vector<void *> kout(nk);
        for (int a = 0; a < nk; ++a) {
            kin[a] = Q[kept[a]];
            kout[a] = Q[kept[a]];
        }
        std::vector<double> S2(nk * nk), R2(nk * nk);
        bool k2[BLK_MAX];
        RCHK(block_gram(h, kin.data(), nk, kin.data(), nk, S2.data()));
        chol_inv(nk, S2.data(), -1.0, false, k2, R2.data(), nullptr);
        RCHK(block_combine(h, kin.data(), nk, kout.data(), nk, R2.data(), nullptr));
    }
    return SPM_OK;
}

} // namespace spm
