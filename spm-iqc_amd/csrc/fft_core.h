// fft_core.h -- device building blocks of the 3-D FFT passes (gfx950, HIP).
//
// A 1-D transform of length N is done by a group of TPP threads that each hold
// E = N/TPP complex values in registers.  Thread t always holds the elements
// t + TPP*m (m = 0..E-1) of its pencil before and after a transform: that is
// the layout in which the first Stockham stage reads and the last one writes,
// so pencils are loaded from and stored to HBM directly from registers, and
// LDS is used only to exchange values between stages.
//
// Stages follow the self-sorting Stockham formulation (Govindaraju et al.,
// "High performance discrete Fourier transforms on graphics processors",
// SC'08): stage s with radix R and Ns = product of the earlier radices takes
// butterfly j = t + TPP*p from elements j + q*N/R, twiddles element q by
// exp(-+2 pi i q (j mod Ns) / (Ns R)), and writes it to
// (j / Ns) * Ns * R + (j mod Ns) + q * Ns.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace spm {

template <typename T> struct C2 {
    T x, y;
};

template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by DIR * i  (DIR = -1 forward, +1 inverse)
template <int DIR, typename T> __device__ __forceinline__ C2<T> mul_i(C2<T> a)
{
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// ---------------------------------------------------------------- constants
// exp(-2 pi i e / R) for the radices used inside the small DFTs.
#define SPM_C1 0.92387953251128675613  /* cos(pi/8) */
#define SPM_S1 0.38268343236508977173  /* sin(pi/8) */
#define SPM_R2 0.70710678118654752440  /* sqrt(2)/2 */
#define SPM_H3 0.86602540378443864676  /* sqrt(3)/2 */

template <typename T, int R> __device__ __forceinline__ constexpr T kcos(int e)
{
    static_assert(16 % R == 0 || 12 % R == 0, "radix");
    if constexpr (16 % R == 0) {
        constexpr double c[16] = {1, SPM_C1, SPM_R2, SPM_S1, 0, -SPM_S1, -SPM_R2, -SPM_C1,
                                       -1, -SPM_C1, -SPM_R2, -SPM_S1, 0, SPM_S1, SPM_R2, SPM_C1};
        return (T)c[(e * (16 / R)) % 16];
    } else {
        constexpr double c[12] = {1, SPM_H3, 0.5, 0, -0.5, -SPM_H3, -1, -SPM_H3, -0.5, 0, 0.5, SPM_H3};
        return (T)c[(e * (12 / R)) % 12];
    }
}
template <typename T, int R> __device__ __forceinline__ constexpr T ksin(int e)
{
    if constexpr (16 % R == 0) {
        constexpr double s[16] = {0, SPM_S1, SPM_R2, SPM_C1, 1, SPM_C1, SPM_R2, SPM_S1,
                                       0, -SPM_S1, -SPM_R2, -SPM_C1, -1, -SPM_C1, -SPM_R2, -SPM_S1};
        return (T)s[(e * (16 / R)) % 16];
    } else {
        constexpr double s[12] = {0, 0.5, SPM_H3, 1, SPM_H3, 0.5, 0, -0.5, -SPM_H3, -1, -SPM_H3, -0.5};
        return (T)s[(e * (12 / R)) % 12];
    }
}

// ---------------------------------------------------------------- small DFTs
// In-place DFT of R values in registers, X[k] = sum_n x[n] exp(DIR 2 pi i n k / R).
template <int R, int DIR, typename T> __device__ __forceinline__ void dft(C2<T> *a);

template <int DIR, typename T> __device__ __forceinline__ void dft2(C2<T> *a)
{
    const C2<T> u = a[0], v = a[1];
    a[0] = cadd(u, v);
    a[1] = csub(u, v);
}

template <int DIR, typename T> __device__ __forceinline__ void dft3(C2<T> *a)
{
    const C2<T> t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
    const C2<T> m = {a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y};
    const C2<T> s = cscale(mul_i<DIR>(t2), (T)SPM_H3);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, s);
    a[2] = csub(m, s);
}

template <int DIR, typename T> __device__ __forceinline__ void dft4(C2<T> *a)
{
    const C2<T> s0 = cadd(a[0], a[2]), d0 = csub(a[0], a[2]);
    const C2<T> s1 = cadd(a[1], a[3]), d1 = mul_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(s0, s1);
    a[1] = cadd(d0, d1);
    a[2] = csub(s0, s1);
    a[3] = csub(d0, d1);
}

// R = A * B by one Cooley-Tukey split: n = B n1 + n2, k = k1 + A k2.
template <int A, int B, int DIR, typename T> __device__ __forceinline__ void dft_split(C2<T> *a)
{
    constexpr int R = A * B;
    C2<T> y[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = a[B * n1 + n2];
        dft<A, DIR>(y[n2]);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            const int e = (n2 * k1) % R;
            if (e != 0) {
                // constants after unrolling; +-i and -1 are taken exactly, without
                // the products by zero a general complex product would keep
                const T cr = kcos<T, R>(e), si = DIR < 0 ? -ksin<T, R>(e) : ksin<T, R>(e);
                if (cr == (T)0) y[n2][k1] = si > (T)0 ? mul_i<+1>(y[n2][k1]) : mul_i<-1>(y[n2][k1]);
                else if (si == (T)0) y[n2][k1] = C2<T>{-y[n2][k1].x, -y[n2][k1].y};
                else y[n2][k1] = cmul(y[n2][k1], C2<T>{cr, si});
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        C2<T> b[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) b[n2] = y[n2][k1];
        dft<B, DIR>(b);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) a[k1 + A * k2] = b[k2];
    }
}

template <int R, int DIR, typename T> __device__ __forceinline__ void dft(C2<T> *a)
{
    if constexpr (R == 1) {
    } else if constexpr (R == 2) dft2<DIR>(a);
    else if constexpr (R == 3) dft3<DIR>(a);
    else if constexpr (R == 4) dft4<DIR>(a);
    else if constexpr (R == 6) dft_split<3, 2, DIR>(a);
    else if constexpr (R == 8) dft_split<4, 2, DIR>(a);
    else if constexpr (R == 12) dft_split<4, 3, DIR>(a);
    else if constexpr (R == 16) dft_split<4, 4, DIR>(a);
    else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------- plans
// A length-N transform by TPP threads per pencil, E = N / TPP values each.
// Stage radices are chosen greedily from {16, 12, 8, 6, 4, 3, 2}: each must
// divide E (so a thread does E / R butterflies) and the length left.
__host__ __device__ constexpr int pick_radix(int E, int rem)
{
    const int cand[7] = {16, 12, 8, 6, 4, 3, 2};
    for (int i = 0; i < 7; ++i)
        if (E % cand[i] == 0 && rem % cand[i] == 0) return cand[i];
    return 0;
}

template <int N, int TPP_> struct FftPlan {
    static constexpr int TPP = TPP_, E = N / TPP_;
    static constexpr int R0 = pick_radix(E, N);
    static constexpr int R1 = N / R0 > 1 ? pick_radix(E, N / R0) : 1;
    static constexpr int R2 = N / (R0 * R1) > 1 ? pick_radix(E, N / (R0 * R1)) : 1;
    static constexpr int R3 = N / (R0 * R1 * R2) > 1 ? pick_radix(E, N / (R0 * R1 * R2)) : 1;
    static constexpr int NST = (R0 > 1) + (R1 > 1) + (R2 > 1) + (R3 > 1);
    static_assert(N % TPP_ == 0 && R0 * R1 * R2 * R3 == N, "unsupported (N, threads per pencil)");
};

// Default threads per pencil for each supported N.
template <int N> struct DefTPP;
#define SPM_DEFTPP(N_, T_)                                                                                 \
    template <> struct DefTPP<N_> {                                                                        \
        static constexpr int value = T_;                                                                   \
    };
SPM_DEFTPP(4, 1)
SPM_DEFTPP(8, 2)
SPM_DEFTPP(12, 2)
SPM_DEFTPP(16, 4)
SPM_DEFTPP(24, 2)
SPM_DEFTPP(32, 4)
SPM_DEFTPP(48, 4)
SPM_DEFTPP(64, 4)
SPM_DEFTPP(96, 8)
SPM_DEFTPP(128, 8)
SPM_DEFTPP(192, 16)
SPM_DEFTPP(256, 16)
#undef SPM_DEFTPP

// LDS slot of twiddle e in a stage's table: one spare slot after every 8, so
// the strided reads of the later stages (entries 4 q jm apart across lanes)
// fall in different banks.
__device__ __forceinline__ constexpr int twpad(int e) { return e + (e >> 3); }
template <int N> constexpr int twiddle_slots() { return N + N / 8; }

// LDS index of pencil element e.  PAD inserts one slot after every 16
// elements: the first stage writes element t*R + q from thread t, which
// without padding puts every thread of a pencil in the same bank.
template <bool PAD> __device__ __forceinline__ constexpr int lpad(int e) { return PAD ? e + (e >> 4) : e; }
template <int N, bool PAD> constexpr int pencil_slots() { return PAD ? N + N / 16 : N; }

// One Stockham stage.  v: the thread's E registers (cyclic layout on input).
// lds: this pencil's exchange area, element e at lds[lpad(e) * LS].
// tw: exp(-2 pi i e / N), e = 0..N-1.
template <typename T, int N, int TPP, int DIR, int R, int NS, bool LAST, int LS, bool PAD>
__device__ __forceinline__ void stockham_stage(C2<T> (&v)[N / TPP], C2<T> *lds, const C2<T> *__restrict__ tw,
                                               int t)
{
    constexpr int E = N / TPP, NB = E / R;
    static_assert(E % R == 0, "radix must divide the per-thread count");
    if constexpr (!LAST) __syncthreads(); // earlier readers of this LDS area are done
#pragma unroll
    for (int p = 0; p < NB; ++p) {
        C2<T> a[R];
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = v[p + q * NB];
        const int j = t + TPP * p;
        const int jm = (NS > 1) ? j % NS : 0;
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                C2<T> w = tw[twpad(jm * q * (N / (NS * R)))];
                if constexpr (DIR > 0) w.y = -w.y;
                a[q] = cmul(a[q], w);
            }
        }
        dft<R, DIR>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) v[p + q * NB] = a[q];
        } else {
            const int base = (j / NS) * NS * R + jm;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lpad<PAD>(base + q * NS) * LS] = a[q];
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = lds[lpad<PAD>(t + TPP * m) * LS];
    }
}

// The same stage with the exchange through a real-valued LDS area (half the
// bytes, so twice the blocks fit on a CU): real parts first, then imaginary
// parts, which wait in registers meanwhile.
template <typename T, int N, int TPP, int DIR, int R, int NS, bool LAST, int LS, bool PAD>
__device__ __forceinline__ void stockham_stage_split(C2<T> (&v)[N / TPP], T *lds, const C2<T> *__restrict__ tw,
                                                     int t)
{
    constexpr int E = N / TPP, NB = E / R;
    static_assert(E % R == 0, "radix must divide the per-thread count");
    if constexpr (!LAST) __syncthreads(); // earlier readers of this LDS area are done
    T hold[LAST ? 1 : E];
#pragma unroll
    for (int p = 0; p < NB; ++p) {
        C2<T> a[R];
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = v[p + q * NB];
        const int j = t + TPP * p;
        const int jm = (NS > 1) ? j % NS : 0;
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                C2<T> w = tw[twpad(jm * q * (N / (NS * R)))];
                if constexpr (DIR > 0) w.y = -w.y;
                a[q] = cmul(a[q], w);
            }
        }
        dft<R, DIR>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) v[p + q * NB] = a[q];
        } else {
            const int base = (j / NS) * NS * R + jm;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                lds[lpad<PAD>(base + q * NS) * LS] = a[q].x;
                hold[p * R + q] = a[q].y;
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m].x = lds[lpad<PAD>(t + TPP * m) * LS];
        __syncthreads();
#pragma unroll
        for (int p = 0; p < NB; ++p) {
            const int j = t + TPP * p;
            const int jm = (NS > 1) ? j % NS : 0;
            const int base = (j / NS) * NS * R + jm;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[lpad<PAD>(base + q * NS) * LS] = hold[p * R + q];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m].y = lds[lpad<PAD>(t + TPP * m) * LS];
    }
}

template <typename T, int N, int TPP, int DIR, int LS, bool PAD = false>
__device__ __forceinline__ void fft_pencil_split(C2<T> (&v)[N / TPP], T *lds, const C2<T> *__restrict__ tw, int t)
{
    using P = FftPlan<N, TPP>;
    if constexpr (P::NST == 1) {
        stockham_stage_split<T, N, TPP, DIR, P::R0, 1, true, LS, PAD>(v, lds, tw, t);
    } else if constexpr (P::NST == 2) {
        stockham_stage_split<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R1, P::R0, true, LS, PAD>(v, lds, tw, t);
    } else if constexpr (P::NST == 3) {
        stockham_stage_split<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R1, P::R0, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R2, P::R0 * P::R1, true, LS, PAD>(v, lds, tw, t);
    } else {
        stockham_stage_split<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R1, P::R0, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R2, P::R0 * P::R1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage_split<T, N, TPP, DIR, P::R3, P::R0 * P::R1 * P::R2, true, LS, PAD>(v, lds, tw, t);
    }
}

// Full length-N transform of the pencil held in v (cyclic in, cyclic out).
// Unnormalised in both directions.
template <typename T, int N, int TPP, int DIR, int LS, bool PAD>
__device__ __forceinline__ void fft_pencil(C2<T> (&v)[N / TPP], C2<T> *lds, const C2<T> *__restrict__ tw, int t)
{
    using P = FftPlan<N, TPP>;
    if constexpr (P::NST == 1) {
        stockham_stage<T, N, TPP, DIR, P::R0, 1, true, LS, PAD>(v, lds, tw, t);
    } else if constexpr (P::NST == 2) {
        stockham_stage<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R1, P::R0, true, LS, PAD>(v, lds, tw, t);
    } else if constexpr (P::NST == 3) {
        stockham_stage<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R1, P::R0, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R2, P::R0 * P::R1, true, LS, PAD>(v, lds, tw, t);
    } else {
        stockham_stage<T, N, TPP, DIR, P::R0, 1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R1, P::R0, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R2, P::R0 * P::R1, false, LS, PAD>(v, lds, tw, t);
        stockham_stage<T, N, TPP, DIR, P::R3, P::R0 * P::R1 * P::R2, true, LS, PAD>(v, lds, tw, t);
    }
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Block-wide sums of NV values in a fixed order (deterministic).  Result valid
// in thread 0.  red: LDS scratch of at least NV * (blockDim / 64) doubles.
// MAXMASK: bit q set => value q is reduced with max instead of sum.
template <int NV, unsigned MAXMASK = 0>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *red)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = ((MAXMASK >> q) & 1u) ? wave_max(v[q]) : wave_sum(v[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q * 16 + wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = red[q * 16];
            for (int w = 1; w < nw; ++w) s = ((MAXMASK >> q) & 1u) ? fmax(s, red[q * 16 + w]) : s + red[q * 16 + w];
            v[q] = s;
        }
    }
}

} // namespace spm
