// gf_split.h -- the gradient-flow step of a 64^3, 128^3 or 256^3 grid in two passes
// over the work array instead of four (gfx950, HIP).
//
// The length-N transform along j is split by one Cooley-Tukey step,
// j = 16 h + l (h < DH = N/16, l < 16):
//   stage A: DH-point DFTs over h for each l (rows 16 apart in one i plane),
//            then the twiddle exp(-2 pi i l k1 / N); digit k1 is stored at h;
//   stage B: 16-point DFTs over l for each k1 (16 adjacent rows); digit k2 is
//            stored at l, so position (h, l) holds frequency J = h + DH l.
// Stage A rides in the row (kz) pass and stage B in the i pass, so one
// iteration of get_min_u_F (LPsrc.cpp:52-111) is
//   pass 1 (rows):    A^-1, c2r, f(phi) and the density sum, r2c, A     W in/out
//   pass 2 (columns): i FFT, B, update with U and its sums, B^-1, i FFT^-1
//                                                          W in/out, U in/out
// (B and the i transform commute)
// six array passes per step instead of the ten of the axis-by-axis passes.
// The same passes make the transform round trips of the other operators.
//
// W of one field is stored as [i][h][ko][l][ki], kz = KT ko + ki, with
// KT = 2 (FP64) or 4 (FP32).  A pass-1 block reads, per (h, ko), the 64-byte
// half of a 128-byte line (2 l x KT ki); a pass-2 block reads, per i, every
// other 16 bytes of 512 (16 l x KT/2 ki), four i per wave instruction.  U is
// chunked per pass-2 block as [l][i][ka].  The two blocks sharing W lines
// are launched 8 block indices apart, i.e. on the same XCD at the same time,
// so each line comes from HBM once.  Entries with kz > N/2 are padding.
#pragma once
#include "grid_ops.h"

namespace spm {

template <typename T, int N> struct SJ {
    static_assert(N == 64 || N == 128 || N == 256, "two-pass gradient-flow step: N = 64, 128 or 256");
    static constexpr int NH = N / 2 + 1;                // valid kz
    static constexpr int DL = 16, DH = N / DL;          // ranges of the digits h and l: j = DL h + l
    static constexpr int KT = sizeof(T) == 8 ? 2 : 4;   // kz per contiguous group
    static constexpr int KO = (NH + KT - 1) / KT;       // kz groups
    static constexpr long HSTR = (long)KO * DL * KT;    // stride of h
    static constexpr long ISTR = DH * HSTR;             // stride of i
    static constexpr long CB = (long)N * ISTR;          // complex entries per field (W and U)
    __host__ __device__ static constexpr long wof(int i, int h, int ko, int l, int ki)
    {
        return (long)i * ISTR + (long)h * HSTR + ((long)ko * DL + l) * KT + ki;
    }
    // pass 1: S1 = 2 l of one i (DH S1 rows, DH S1 / 2 pairs of RTPP threads);
    // block b and b + 8 take l pairs 2 lp and 2 lp + 1 (one line)
    static constexpr int S1 = 2, RTPP = 16;
    static constexpr int NP = DH * S1 / 2;
    static constexpr int R_THREADS = NP * RTPP;
    static constexpr long R_BLOCKS = (long)N * (DL / S1);
    __device__ static void rblock(int b, int &i, int &lo)
    {
        const int x = b % 8, s = (b / 8) % 2, q = b / 16;
        const int P = q * 8 + x; // (i, lp), lp < DL / (2 S1)
        i = P / (DL / (2 * S1));
        lo = 2 * (P % (DL / (2 * S1))) + s;
    }
    // pass 2: (h, ko, group kg of KB ki), NC = DL KB columns (l, ka); the NSIB =
    // KT / KB blocks of one (h, ko) share its W lines and are 8 block indices
    // apart.  CTPP threads per i pencil, DL values each (one register array
    // serves the line role, DL values of l, and the column role, DL values of i)
    static constexpr int KB = KT / 2;
    static constexpr int NSIB = KT / KB;
    static constexpr int NC = DL * KB;
    static constexpr int CTPP = N / DL;
    static constexpr int C_THREADS = NC * CTPP;
    static constexpr int NPG = (DH * KO + 7) / 8 * 8; // (h, ko) pairs, padded to whole XCD rounds
    static constexpr long C_BLOCKS = (long)NSIB * NPG;
    // returns false for the padding pairs (no work; zero partial sums)
    __device__ static bool cblock(int b, int &h, int &ko, int &kg)
    {
        const int x = b % 8, s = (b / 8) % NSIB, q = b / (8 * NSIB);
        const int P = q * 8 + x; // (h, ko)
        h = P < DH * KO ? P / KO : 0;
        ko = P % KO;
        kg = s;
        return P < DH * KO;
    }
    // a pass-2 block with no valid column (kz of all its columns past N/2)
    __device__ static bool cblock_empty(bool real_blk, int ko, int kg) { return !real_blk || ko * KT + KB * kg >= NH; }
    // U: the state of one pass-2 block is one contiguous chunk [l][i][ka]
    // (digit k2 at l); r = i KB + ka is the pass-2 thread of the line role
    __host__ __device__ static constexpr long uof(int h, int ko, int kg, int l, int r)
    {
        return ((((long)h * KO + ko) * NSIB + kg) * DL + l) * ((long)N * KB) + r;
    }
    static_assert(R_THREADS == S1 * (N / 2), "pass 1: one (l, kz < N/2) column per thread");
    static_assert(C_THREADS == N * KB, "pass 2: one (i, ka) line per thread");
    static_assert((N * DL / (2 * S1)) % 8 == 0, "pass-1 block map");
};

// ----------------------------------------------------------------------
// Block-0 prologues closing an iteration (partial sums are complete at the
// kernel boundary).  E2 partials come from pass 1 of the previous iteration,
// the E1 / ||grad F||^2 / residual partials from pass 2 of this one.
// ----------------------------------------------------------------------

// Pass 2, block 0: sum of the density partials of pass 1 into e2[m].
struct ProSJDensity {
    static constexpr bool NONE = false;
    const double *pC;
    long nC;
    double *e2;
    __device__ __forceinline__ ProSJDensity bind(int m) const
    {
        ProSJDensity o = *this;
        o.pC = pC + (long)m * nC;
        o.e2 = e2 + m;
        return o;
    }
    __device__ int nblk() const { return 1; }
    __device__ void operator()(double *red) const
    {
        double v[1] = {0.0};
        for (long b = threadIdx.x; b < nC; b += blockDim.x) v[0] += pC[b];
        block_reduce<1, 0u>(v, red);
        if (threadIdx.x == 0) *e2 = v[0];
        __syncthreads();
    }
};

// Pass 2, block 0: one scalar per field of a round trip, out = scale * sum of
// the pass-1 partials (the mean that the output pass subtracts), in place of a
// reduction launch between the passes.
struct ProSum {
    static constexpr bool NONE = false;
    const double *p; // member m: partials at p + m n, scalar into out[m]
    long n;
    double scale;
    double *out;
    __device__ __forceinline__ ProSum bind(int m) const
    {
        ProSum o = *this;
        o.p = p + (long)m * n;
        o.out = out + m;
        return o;
    }
    __device__ int nblk() const { return 1; }
    __device__ void operator()(double *red) const
    {
        double v[1] = {0.0};
        for (long b = threadIdx.x; b < n; b += blockDim.x) v[0] += p[b];
        block_reduce<1, 0u>(v, red);
        if (threadIdx.x == 0) *out = v[0] * scale;
        __syncthreads();
    }
};

// Pass 1, block 0: E = E1 + E2, ||grad F||, residual into hist[3*iter ..],
// iter += 1, and the stop test of LPsrc.cpp:52.
struct ProSJFinalize {
    static constexpr bool NONE = false;
    const double *pA; // pass-2 partials: 3 values per block
    long nA;
    const double *e2;
    double R;
    Ctl *ctl;
    double *hist;
    long hstride;
    __device__ __forceinline__ ProSJFinalize bind(int m) const
    {
        ProSJFinalize o = *this;
        o.pA = pA + (long)m * 3 * nA;
        o.e2 = e2 + m;
        o.ctl = ctl + m;
        o.hist = hist + (long)m * hstride;
        return o;
    }
    __device__ int nblk() const { return 1; }
    __device__ void operator()(double *red) const
    {
        double v[3] = {0.0, 0.0, 0.0};
        for (long b = threadIdx.x; b < nA; b += blockDim.x) {
            v[0] += pA[b];
            v[1] += pA[nA + b];
            v[2] = fmax(v[2], pA[2 * nA + b]);
        }
        block_reduce<3, 4u>(v, red);
        if (threadIdx.x == 0) {
            const double E = 0.5 * v[0] + *e2 / R, gn = sqrt(v[1]), res = v[2];
            const int it = ctl->iter;
            hist[3L * it + 0] = E;
            hist[3L * it + 1] = gn;
            hist[3L * it + 2] = res;
            ctl->iter = it + 1;
            if (!(res > ctl->res_tol) || (ctl->gn_tol > 0 && gn <= ctl->gn_tol) || it + 1 >= ctl->max_iter)
                ctl->stop = 1;
        }
        __syncthreads();
    }
};

// ----------------------------------------------------------------------
// Pass 1: one i plane, the rows j = 16 h + l for all h and S1 consecutive l.
//   IN_SPEC:  input W: A^-1, then c2r of each row
//   !IN_SPEC: input real field (natural layout [i][j][k])
//   OUT_SPEC: r2c of op(x) per row, then A, into W; otherwise op only
// Rows pair up as (l even, l odd) of one h: the two real rows of a pair are
// the real and imaginary parts of one complex length-N FFT.  The column role
// of a thread (the 16 h of one (l, kz < N/2)) does stage A; the row role
// (TPP threads per pair) does the row FFTs; the row buffer in LDS converts.
// ----------------------------------------------------------------------
// At most 168 registers (3 waves per SIMD): FP64 needs 194 otherwise; the two it
// then spills cost less than the third block per CU gains (2 % of the step).
template <typename T, int N, bool IN_SPEC, bool OUT_SPEC, class Op, class Pro = ProNone>
__global__ void __launch_bounds__((SJ<T, N>::R_THREADS), 3)
    k_sj_rows(const void *in, void *out, const C2<T> *__restrict__ tw, Op op0, double *__restrict__ part,
              const Ctl *ctl, Pro pro0, long in_bs, long out_bs)
{
    using S = SJ<T, N>;
    constexpr int TPP = S::RTPP, E = N / TPP, NP = S::NP, S1 = S::S1, DH = S::DH, DL = S::DL, KT = S::KT;
    constexpr int NR = 2 * NP;
    constexpr int RS = N / 2 + 8;                      // row stride of the row buffer (reals)
    // Pencil exchange stride (reals), = 16 mod 32: the 4 pencils of a wave then
    // sit 16 banks apart, and the first stage's writes (element 16 t + q at
    // padded slot 17 t + q) of the 64 lanes fall in distinct banks.
    constexpr int RSTR = pencil_slots<N, true>() + (48 - pencil_slots<N, true>() % 32) % 32;
    constexpr int NL = NR * RS > NP * RSTR ? NR * RS : NP * RSTR;
    __shared__ __align__(16) T lds[NL];
    __shared__ T nyq[NR];            // kz = N/2 entry of each row (real)
    __shared__ C2<T> nyin[DH * S1];  // kz = N/2 column input, before A^-1
    __shared__ C2<T> stw[twiddle_slots<N>()];
    // W^(l h) of the block's S1 values of l (stage A and its inverse): a thread reads
    // its row at fixed offsets, with no per-twiddle index arithmetic (l h < N)
    constexpr int LTS = DH + 1;
    __shared__ C2<T> ltw[S1 * LTS];
    __shared__ double red[(Op::NRED > 0 ? Op::NRED : 1) * 16 + (Pro::NONE ? 0 : 64)];
    const int mb = blockIdx.y;
    if (ctl && ctl[mb].stop) return;
    const Op op = bind_member(op0, mb);
    if constexpr (!Pro::NONE) {
        const Pro pro = bind_member(pro0, mb);
        if ((int)blockIdx.x < pro.nblk()) pro(red);
    }
    load_twiddles<T, N>(stw, tw);
    int i, lo;
    S::rblock(blockIdx.x, i, lo);
    const int tid = threadIdx.x;
    static_assert((DL - 1) * (DH - 1) < N, "l h indexes the table directly");
    if (tid < S1 * DH) ltw[(tid / DH) * LTS + tid % DH] = tw[(S1 * lo + tid / DH) * (tid % DH)];
    // column role: the 16 h values of (l = S1 lo + li, kz)
    // FP64: a wave spans both l of 32 kz; FP32: one l and 64 kz, so that its row-buffer
    // accesses of 4-byte values fall in distinct banks (3.5 % of the FP32 step)
    constexpr bool LMAJ = sizeof(T) == 4;
    const int li = LMAJ ? tid / (N / 2) : (tid / KT) % S1;
    const int kzc = LMAJ ? tid % (N / 2) : (tid / (KT * S1)) * KT + tid % KT;
    const int ki = kzc % KT, ko = kzc / KT;
    const int kz = ko * KT + ki, l = S1 * lo + li;
    // row role: pair pp = rows ra, ra + 1 (row r = h S1 + li)
    const int t = tid % TPP, pp = tid / TPP;
    const int ra = 2 * pp, rb = ra + 1;
    T *my = lds + pp * RSTR;
    C2<T> v[E];
    C2<T> c[DH];
    C2<T> aux[!IN_SPEC && has_aux<Op>::value ? E : 1]; // the op's other field, loaded with the input
    double acc[Op::NRED > 0 ? Op::NRED : 1];
#pragma unroll
    for (int q = 0; q < (Op::NRED > 0 ? Op::NRED : 1); ++q) acc[q] = 0.0;

    if constexpr (IN_SPEC) {
        const C2<T> *W = (const C2<T> *)in + mb * in_bs;
#pragma unroll
        for (int h = 0; h < DH; ++h) c[h] = W[S::wof(i, h, ko, l, ki)];
        // the kz = N/2 column input goes straight to LDS (no register across the A^-1 below)
        if (tid < DH * S1) nyin[tid] = W[S::wof(i, tid / S1, (N / 2) / KT, S1 * lo + tid % S1, (N / 2) % KT)];
        __syncthreads(); // twiddles, nyin
        // A^-1: conjugate twiddle, inverse 16-point DFT over h
#pragma unroll
        for (int h = 1; h < DH; ++h) {
            C2<T> w = ltw[li * LTS + h];
            w.y = -w.y;
            c[h] = cmul(c[h], w);
        }
        dft<DH, +1>(c);
        // to the row buffer: real parts, then imaginary parts
        T hold[DH];
#pragma unroll
        for (int h = 0; h < DH; ++h) {
            lds[(h * S1 + li) * RS + kz] = c[h].x;
            hold[h] = c[h].y;
        }
        __syncthreads();
        // the kz = N/2 columns: only their real parts enter the c2r
        if constexpr (DH >= 8) { // one output per lane,
            // x[h] = Re sum_k conj(W^(k (DL h + ln))) y[k], W = exp(-2 pi i / N)
            if (tid < DH * S1) {
                const int cn = tid / DH, hn = tid % DH, ln = S1 * lo + cn;
                T sx = 0;
#pragma unroll
                for (int k = 0; k < DH; ++k) {
                    const C2<T> w = stw[twpad((k * (DL * hn + ln)) & (N - 1))], y = nyin[k * S1 + cn];
                    sx += w.x * y.x + w.y * y.y;
                }
                nyq[hn * S1 + cn] = sx;
            }
        } else if (tid < S1) { // a 4-point transform per column (N = 64: one wave per block)
#pragma unroll
            for (int h = 0; h < DH; ++h) {
                c[h] = nyin[h * S1 + tid];
                if (h > 0) {
                    C2<T> w = ltw[tid * LTS + h];
                    w.y = -w.y;
                    c[h] = cmul(c[h], w);
                }
            }
            dft<DH, +1>(c);
#pragma unroll
            for (int h = 0; h < DH; ++h) nyq[h * S1 + tid] = c[h].x;
        }
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPP * m;
            const int q = p <= N / 2 ? p : N - p;
            v[m].x = q < N / 2 ? lds[ra * RS + q] : (T)0;
            v[m].y = q < N / 2 ? lds[rb * RS + q] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < DH; ++h) lds[(h * S1 + li) * RS + kz] = hold[h];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPP * m;
            const int q = p <= N / 2 ? p : N - p;
            T ar = v[m].x, br = v[m].y, ai = 0, bi = 0;
            if (q == N / 2) {
                ar = nyq[ra];
                br = nyq[rb];
            } else if (q > 0) { // c2r reads only the real part of DC and Nyquist
                ai = lds[ra * RS + q];
                bi = lds[rb * RS + q];
            }
            // Z = Xa + i Xb on 0..N/2, conj(Xa) + i conj(Xb) above
            v[m] = p <= N / 2 ? C2<T>{ar - bi, ai + br} : C2<T>{ar + bi, br - ai};
        }
        fft_pencil_split<T, N, TPP, +1, 1, true>(v, my, stw, t);
    } else {
        const T *X = (const T *)in + mb * in_bs;
        const int ja = DL * (ra / S1) + S1 * lo + ra % S1, jb = ja + 1;
        const T *xa = X + ((long)i * N + ja) * N, *xb = X + ((long)i * N + jb) * N;
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = C2<T>{xa[t + TPP * m], xb[t + TPP * m]};
        if constexpr (has_aux<Op>::value) {
            const T *A = op.aux();
            const T *aa = A + ((long)i * N + ja) * N, *ab = A + ((long)i * N + jb) * N;
#pragma unroll
            for (int m = 0; m < E; ++m) aux[m] = C2<T>{aa[t + TPP * m], ab[t + TPP * m]};
        }
        __syncthreads(); // twiddles
    }
    // The pointwise work sits in a basic block of its own (the condition always
    // holds): scheduled in one region with the FFT code around it, it took
    // twice the registers.
    {
        constexpr int CH = op_chunk<Op>::value > 0 && op_chunk<Op>::value < E ? op_chunk<Op>::value : E;
        static_assert(E % CH == 0, "whole chunks");
        const int ja = DL * (ra / S1) + S1 * lo + ra % S1, jb = ja + 1;
        const long ga = ((long)i * N + ja) * N, gb = ((long)i * N + jb) * N;
#pragma unroll
        for (int m0 = 0; m0 < E; m0 += CH) {
            if (blockIdx.x < gridDim.x + m0) {
#pragma unroll
                for (int m = m0; m < m0 + CH; ++m) {
                    const int p = t + TPP * m;
                    if constexpr (!IN_SPEC && has_aux<Op>::value) {
                        v[m].x = op(v[m].x, aux[m].x, ga + p, acc);
                        v[m].y = op(v[m].y, aux[m].y, gb + p, acc);
                    } else {
                        v[m].x = op(v[m].x, ga + p, acc);
                        v[m].y = op(v[m].y, gb + p, acc);
                    }
                }
            }
        }
    }
    if constexpr (OUT_SPEC) {
        fft_pencil_split<T, N, TPP, -1, 1, true>(v, my, stw, t);
        // The indices of this phase again, from copies of the thread and block
        // index that the compiler cannot match with the ones above: otherwise it
        // keeps the addresses computed at the start live to here, and spills them.
        int tid = threadIdx.x, bx = blockIdx.x;
        asm volatile("" : "+v"(tid));
        asm volatile("" : "+s"(bx));
        int i, lo;
        S::rblock(bx, i, lo);
        const int li = LMAJ ? tid / (N / 2) : (tid / KT) % S1;
        const int kzc = LMAJ ? tid % (N / 2) : (tid / (KT * S1)) * KT + tid % KT;
        const int ki = kzc % KT, ko = kzc / KT;
        const int kz = ko * KT + ki, l = S1 * lo + li;
        const int t = tid % TPP;
        T *my = lds + (tid / TPP) * RSTR;
        // The column role separates the two rows of a pair itself: pair pp = h
        // holds rows (h, l even) and (h, l odd), and with Z its spectrum,
        //   Xa[k] = (Z[k] + conj Z[N-k]) / 2,  Xb[k] = -i (Z[k] - conj Z[N-k]) / 2.
        // Z goes through LDS as real parts, then imaginary parts; the kz = N/2
        // entries (real) go to nyq.
        static_assert(S1 == 2, "pair h = rows (h, 2 lo), (h, 2 lo + 1)");
        // With Z = Zr + i Zi at kz and Zs + i Zj at N - kz: Xa = (Zr + Zs, Zi - Zj) / 2
        // and Xb = (Zi + Zj, Zs - Zr) / 2; the real-part term u waits for the second
        // exchange (sg = +1 for l even, -1 for l odd)
        const T sg = li == 0 ? (T)1 : (T)-1;
        T u[DH];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) my[t + TPP * m] = v[m].x;
        __syncthreads();
#pragma unroll
        for (int h = 0; h < DH; ++h) u[h] = fma(sg, lds[h * RSTR + kz], lds[h * RSTR + (N - kz) % N]);
        if (tid < DH) nyq[tid * S1] = lds[tid * RSTR + N / 2]; // Xa[N/2] = Re Z[N/2]
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) my[t + TPP * m] = v[m].y;
        __syncthreads();
        {
            const T hx = (T)0.5;
#pragma unroll
            for (int h = 0; h < DH; ++h) {
                const T zi = lds[h * RSTR + kz], zj = lds[h * RSTR + (N - kz) % N];
                const T w = fma(-sg, zj, zi); // l even: Zi - Zj; l odd: Zi + Zj
                c[h] = li == 0 ? C2<T>{hx * u[h], hx * w} : C2<T>{hx * w, hx * u[h]};
            }
        }
        if (tid >= DH && tid < 2 * DH) nyq[(tid - DH) * S1 + 1] = lds[(tid - DH) * RSTR + N / 2]; // Xb[N/2] = Im Z[N/2]
        __syncthreads();
        // A: 16-point DFT over h, then the twiddle; digit k1 goes to position h
        C2<T> *Wo = (C2<T> *)out + mb * out_bs;
        auto stage_a = [&](int lr) { // lr: the row of ltw, l - S1 lo
            dft<DH, -1>(c);
#pragma unroll
            for (int h = 1; h < DH; ++h) c[h] = cmul(c[h], ltw[lr * LTS + h]);
        };
        stage_a(li);
#pragma unroll
        for (int h = 0; h < DH; ++h) Wo[S::wof(i, h, ko, l, ki)] = c[h];
        // kz = N/2 columns (real after the r2c)
        if constexpr (DH >= 8) { // one output per lane, X[k1] = sum_h W^(k1 (DL h + ln)) x[h]
            if (tid < DH * S1) {
                const int cn = tid / DH, k1 = tid % DH, ln = S1 * lo + cn;
                C2<T> sx = {(T)0, (T)0};
#pragma unroll
                for (int h = 0; h < DH; ++h) {
                    const C2<T> w = stw[twpad((k1 * (DL * h + ln)) & (N - 1))];
                    const T x = nyq[h * S1 + cn];
                    sx.x += w.x * x;
                    sx.y += w.y * x;
                }
                Wo[S::wof(i, k1, (N / 2) / KT, ln, (N / 2) % KT)] = sx;
            }
        } else if (tid < S1) {
            const int ln = S1 * lo + tid;
#pragma unroll
            for (int h = 0; h < DH; ++h) c[h] = C2<T>{nyq[h * S1 + tid], (T)0};
            stage_a(tid);
#pragma unroll
            for (int h = 0; h < DH; ++h) Wo[S::wof(i, h, (N / 2) / KT, ln, (N / 2) % KT)] = c[h];
        }
    }
    write_partials<Op::NRED, Op::MAXM>(acc, red, part);
}

// ----------------------------------------------------------------------
// Pass 2: one (h, ko, kg): the columns (l, ka), ki = KB kg + ka, over all i.
//   FWD: forward FFTs along i, then B (16-point DFTs over l) on W
//   op:  the spectral operator at frequency (i, J = h + DH l, kz); PRELOAD
//        operators read their state at the U offset
//   INV: B^-1, then inverse FFTs along i, into W
// Without FWD the input is read from the U layout, without INV the output is
// written to it.  The i transforms run in the column role, whose W accesses
// cover 4 x 512 bytes per wave (64 B of each 128-B line, the other half being
// the sibling block's); B and the operator run in the line role, whose U
// accesses are contiguous.  (B and the i transform commute.)
// ----------------------------------------------------------------------
template <typename T, int N, bool FWD, bool INV, class SOp, class Pro = ProNone>
__global__ void __launch_bounds__((SJ<T, N>::C_THREADS))
    k_sj_cols(const C2<T> *src, C2<T> *dst, const C2<T> *__restrict__ tw, SOp op0, double *__restrict__ part,
              const Ctl *ctl, Pro pro0, long bs)
{
    using S = SJ<T, N>;
    constexpr int TPP = S::CTPP, E = N / TPP, KT = S::KT, KB = S::KB, NC = S::NC, DL = S::DL, DH = S::DH;
    constexpr int LS = NC + KB; // LDS stride of i: the line-role accesses then cover all banks
    constexpr long LSTR = (long)N * KB; // U stride of l
    __shared__ T lds[N * LS];
    __shared__ C2<T> stw[twiddle_slots<N>()];
    __shared__ double red[(SOp::NRED > 0 ? SOp::NRED : 1) * 16 + (Pro::NONE ? 0 : 64)];
    const int mb = blockIdx.y;
    if (ctl && ctl[mb].stop) return;
    const SOp op = bind_member(op0, mb);
    if constexpr (!Pro::NONE) {
        const Pro pro = bind_member(pro0, mb);
        if ((int)blockIdx.x < pro.nblk()) pro(red);
    }
    src += mb * bs;
    if (dst) dst += mb * bs;
    load_twiddles<T, N>(stw, tw);
    int h, ko, kg;
    const bool real_blk = S::cblock(blockIdx.x, h, ko, kg);
    double acc[SOp::NRED > 0 ? SOp::NRED : 1];
#pragma unroll
    for (int q = 0; q < (SOp::NRED > 0 ? SOp::NRED : 1); ++q) acc[q] = 0.0;
    if (S::cblock_empty(real_blk, ko, kg)) { // nothing to read or write: zero partial sums only
        write_partials<SOp::NRED, SOp::MAXM>(acc, red, part);
        return;
    }
    // line role: thread r = (ia, ka) holds the DL values l of (i = ia, ki = KB kg + ka)
    const int r = threadIdx.x, ka = r % KB, ia = r / KB;
    const int kia = KB * kg + ka;
    const bool va = real_blk && ko * KT + kia < S::NH;
    const long ub = S::uof(h, ko, kg, 0, r);
    // column role: thread (cb, tb) holds column cb = (l, ka) at i = tb + TPP m
    const int cb = threadIdx.x % NC, tb = threadIdx.x / NC;
    const int kzb = ko * KT + KB * kg + cb % KB;
    const bool vb = real_blk && kzb < S::NH;
    const long wbase = S::wof(0, h, ko, cb / KB, KB * kg + cb % KB);
    C2<T> v[E];
    static_assert(E == DL, "one register array serves both roles");
    C2<T> pre[SOp::PRELOAD ? DL : 1]; // state, loaded at the start: its latency overlaps the transforms
    if constexpr (FWD) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = vb ? src[wbase + (long)(tb + TPP * m) * S::ISTR] : C2<T>{0, 0};
        if constexpr (SOp::PRELOAD) {
            // the state's loads once the W loads have landed: they then overlap the
            // transforms instead of competing with W (FP64 step at 256^3 1.5 us faster)
            __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
            for (int q = 0; q < DL; ++q) pre[q] = va ? op.load_pre(ub + q * LSTR) : C2<T>{0, 0};
        }
        fft_pencil_split<T, N, TPP, -1, LS>(v, lds + cb, stw, tb);
        // column role -> line role: real parts, then imaginary parts
        T hold[E];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) {
            lds[(tb + TPP * m) * LS + cb] = v[m].x;
            hold[m] = v[m].y;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DL; ++q) v[q].x = lds[ia * LS + q * KB + ka];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) lds[(tb + TPP * m) * LS + cb] = hold[m];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DL; ++q) v[q].y = lds[ia * LS + q * KB + ka];
        dft<16, -1>(v);
    } else {
#pragma unroll
        for (int q = 0; q < DL; ++q) v[q] = va ? src[ub + q * LSTR] : C2<T>{0, 0};
    }
    // The update in basic blocks of 4 elements (the conditions always hold):
    // scheduled in one region, with its table gathers hoisted, it took far
    // more registers.
    if (va) {
        const int kz = ko * KT + kia;
#pragma unroll
        for (int q0 = 0; q0 < DL; q0 += 4) {
            if (blockIdx.x < gridDim.x + q0) {
                if constexpr (gathers_group<SOp>::value && N <= 128) { // the group's loads first, then its updates
                    // (at 256^3 the per-element order is faster)
                    C2<T> ug[4], gg[4];
                    int n2[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        n2[q] = SOp::n2of(ia, h + DH * (q0 + q), kz);
                        gg[q] = op.load_gd(n2[q]);
                        ug[q] = op.load_pre(ub + (q0 + q) * LSTR);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) op.update(v[q0 + q], ug[q], gg[q], n2[q], kz, acc);
                } else {
#pragma unroll
                    for (int q = q0; q < q0 + 4; ++q)
                        op(v[q], pre[SOp::PRELOAD ? q : 0], ub + q * LSTR, ia, h + DH * q, kz, acc);
                }
            }
        }
        if constexpr (defers_store<SOp>::value)
#pragma unroll
            for (int q = 0; q < DL; ++q) op.store_u(ub + q * LSTR, v[q]);
    }
    if constexpr (INV) {
        dft<16, +1>(v);
        // line role -> column role
        T hold[E];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DL; ++q) {
            lds[ia * LS + q * KB + ka] = v[q].x;
            hold[q] = v[q].y;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m].x = lds[(tb + TPP * m) * LS + cb];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < DL; ++q) lds[ia * LS + q * KB + ka] = hold[q];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m].y = lds[(tb + TPP * m) * LS + cb];
        fft_pencil_split<T, N, TPP, +1, LS>(v, lds + cb, stw, tb);
        if (vb)
#pragma unroll
            for (int m = 0; m < E; ++m) dst[wbase + (long)(tb + TPP * m) * S::ISTR] = v[m];
    } else {
        if (va)
#pragma unroll
            for (int q = 0; q < DL; ++q) dst[ub + q * LSTR] = v[q];
    }
    write_partials<SOp::NRED, SOp::MAXM>(acc, red, part);
}

} // namespace spm
