// grid_ops.h -- 3-D FFT passes over one N^3 grid and the pointwise work of
// the Landau free-energy path fused into them (gfx950, HIP).
//
// Storage on the device (x-major, matching the reference's index order
// idx = (i*N + j)*N + k, Initialization.cpp:109-120):
//   real field  T[N][N][N]
//   half spectrum  C2<T>[N][N][NHP], kz = 0..N/2 valid, NHP >= N/2+1 padded so
//   every row starts on a 128-byte boundary.  Entry (i, j, kz) holds the
//   reference's u_F_local[(i*N + j)*N + kz] (normalised by 1/N^3); the other
//   half follows from Hermitian symmetry.
//
// A 3-D transform is three passes, one per axis, each reading and writing
// the array once:
//   k axis  "row" pass: two real rows packed as one complex row, one length-N
//           complex FFT, separated into two half rows (and the reverse).
//   j axis  column pass over each i plane (pencils of stride NHP).
//   i axis  column pass over the whole array (pencils of stride N*NHP); the
//           pointwise spectral operator (pair kernel G(k), preconditioner,
//           semi-implicit update, energy sums) runs here between the forward
//           and the inverse transform, so the spectrum of a step never makes
//           an extra trip through HBM.
#pragma once
#include <utility>

#include "fft_core.h"

namespace spm {

// Device control block for iterative loops (captured in graphs: kernels read
// it instead of taking a step index as an argument).
struct Ctl {
    int iter;       // iterations completed
    int stop;       // nonzero: later step kernels return immediately
    int max_iter;
    int flags;
    double res_tol; // stop when res <= res_tol
    double gn_tol;  // stop when ||grad F|| <= gn_tol
};

template <typename T, int N, int TPP_ = DefTPP<N>::value> struct Geo {
    static constexpr int TPP = TPP_, E = N / TPP_;
    // column passes: B adjacent columns per block (row segments of 256 B in
    // FP64, 128 B in FP32; measured faster than 128 B / 64 B in FP64 and than
    // 32 columns in either precision)
    static constexpr int B0 = 16;
    static constexpr int B = (64 / TPP > B0) ? 64 / TPP : B0;
    static constexpr int COL_THREADS = B * TPP;
    // FP64 column passes exchange real and imaginary parts separately through
    // an LDS tile of N*B doubles: half the LDS per block, twice the blocks per CU
    static constexpr bool XSPLIT = true;
    using XT = typename std::conditional<XSPLIT, T, C2<T>>::type;
    // row passes: PP row pairs per block
    static constexpr int PP0 = (sizeof(T) == 8 ? 128 : 256) / TPP;
    static constexpr int PP = PP0 > 0 ? PP0 : 1;
    static constexpr int ROW_THREADS = PP * TPP;
    static constexpr int RSTR = pencil_slots<N, true>() + 1; // LDS stride between row pencils (complex)
};

__device__ __forceinline__ int kfold(int i, int N) { return i < N / 2 ? i : i - N; }

// Batched launches: blockIdx.y selects the member of a batch of independent
// fields stored one after another.  An operator with member-dependent state
// (parameters, state arrays, partial-sum slots) has bind(m), which returns
// its copy for member m; operators without one are shared by all members.
template <class Op, class = void> struct has_bind : std::false_type {};
template <class Op>
struct has_bind<Op, std::void_t<decltype(std::declval<const Op &>().bind(0))>> : std::true_type {};
template <class Op> __device__ __forceinline__ Op bind_member(const Op &op, int m)
{
    if constexpr (has_bind<Op>::value)
        return op.bind(m);
    else
        return op;
}

// A real-space operator that gathers from other fields may set CHUNK: the row
// passes then apply it in basic blocks of CHUNK values per thread (or of all the
// values a thread holds, when those are fewer), so that the
// scheduler cannot hoist every gather of the thread at once (registers).
template <class Op, class = void> struct op_chunk : std::integral_constant<int, 0> {};
template <class Op> struct op_chunk<Op, std::void_t<decltype(Op::CHUNK)>> : std::integral_constant<int, Op::CHUNK> {};

// An input-pass operator that reads one other field at the same index may
// declare it as aux() and take its value as an argument, op(x, a, g, acc): the
// two-pass row kernel then loads it together with its input, at the start, so
// the two loads overlap.  op(x, g, acc) gathers it itself (the axis-by-axis passes).
template <class Op, class = void> struct has_aux : std::false_type {};
template <class Op> struct has_aux<Op, std::void_t<decltype(std::declval<const Op &>().aux())>> : std::true_type {};

// A spectral operator that updates a state array in place may leave the store
// of the new state to the pass (DEFER_STORE: the new state is the value it
// returns), which stores it after the whole update: on gfx9 loads and stores
// share one counter, so a store between two elements' loads makes the second
// element wait for the store to complete.
template <class S, class = void> struct defers_store : std::false_type {};
template <class S> struct defers_store<S, std::void_t<decltype(S::DEFER_STORE)>> : std::bool_constant<S::DEFER_STORE> {};
// A spectral operator with a gathered state (GATHER_GROUP) offers n2of, load_gd,
// load_pre and update, so the two-pass column kernel can issue the loads of a
// group of elements together before it updates them.
template <class S, class = void> struct gathers_group : std::false_type {};
template <class S> struct gathers_group<S, std::void_t<decltype(S::GATHER_GROUP)>> : std::bool_constant<S::GATHER_GROUP> {};

// ======================================================================
// Real-space operators applied to each real value inside the row pass.
// op(x, g, acc): x = value at real-field index g; returns the value the pass
// transforms next (if it does), may write outputs and accumulate FP64 sums.
// ======================================================================
struct OpNone {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    template <typename T> __device__ __forceinline__ T operator()(T x, long, double *) const { return x; }
};

template <typename T> struct OpStore { // c2r output: out[g] = x
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    T *out;
    __device__ __forceinline__ T operator()(T x, long g, double *) const
    {
        out[g] = x;
        return x;
    }
};

// Local free-energy density and nonlinear force of get_Ene / get_min_u_F
// (LPsrc.cpp:11-21, 56-61).  Sums the density (E2 = its mean) and returns
// f = Tau*u + Gamma*u^2 - u^3.  Optionally stores u.
template <typename T> struct OpGradFlowRow {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    T *phi_out;
    const double *prm; // batched: (Tau, Gamma) of member m at prm[2m], prm[2m+1]
    long out_bs;       // batched: member stride of phi_out
    __device__ __forceinline__ OpGradFlowRow bind(int m) const
    {
        OpGradFlowRow o = *this;
        if (prm) {
            o.Tau = prm[2 * m];
            o.Gamma = prm[2 * m + 1];
        }
        if (phi_out) o.phi_out = phi_out + (long)m * out_bs;
        return o;
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const
    {
        if (phi_out) phi_out[g] = x;
        if constexpr (sizeof(T) == 4) { // FP32 fields: pointwise math in FP32, the sum in FP64
            const float u = x, u2 = u * u, u3 = u2 * u, tau = (float)Tau, gam = (float)Gamma;
            acc[0] += (double)(-0.5f * tau * u2 - gam * u3 * (1.0f / 3.0f) + 0.25f * u2 * u2);
            return tau * u + gam * u2 - u3;
        } else {
            const double u = (double)x, u2 = u * u, u3 = u2 * u, u4 = u3 * u;
            acc[0] += -0.5 * Tau * u2 - Gamma * u3 * (1.0 / 3.0) + 0.25 * u4;
            return (T)(Tau * u + Gamma * u2 - u3);
        }
    }
};

// Energy density only (get_Ene's nln, LPsrc.cpp:11-21); passes u through.
template <typename T> struct OpDensity {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    __device__ __forceinline__ T operator()(T x, long, double *acc) const
    {
        const double u = (double)x, u2 = u * u, u3 = u2 * u, u4 = u3 * u;
        acc[0] += -0.5 * Tau * u2 - Gamma * u3 * (1.0 / 3.0) + 0.25 * u4;
        return x;
    }
};

// ngrad_cam input pass: transform phi, sum f(phi) (its mean is the mean of
// the gradient before it is removed, LPsrc.cpp:147-167).
template <typename T> struct OpNgradIn {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    __device__ __forceinline__ T operator()(T x, long, double *acc) const
    {
        const double u = (double)x, u2 = u * u;
        acc[0] += Tau * u + Gamma * u2 - u2 * u;
        return x;
    }
};

// The force round trip's output pass fused with ngrad_cam's input pass
// (SPM.cpp:235-236): the c2r result x is stored, the sum of
// f(x) = Tau x + Gamma x^2 - x^3 accumulated (its mean), and x goes on to the
// forward transform.
template <typename T> struct OpStoreNgradIn {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    T *out;
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const
    {
        out[g] = x;
        const double u = (double)x, u2 = u * u;
        acc[0] += Tau * u + Gamma * u2 - u2 * u;
        return x;
    }
};

// Spring_Pairing's first tangent (get_Tv, SPM.cpp:15-32) as the input pass of
// its G convolution: x = Phi0 (the pass's input), Phi1 gathered; the
// unnormalised tangent d = Phi1 - Phi0 is stored and transformed, and the sum of
// d^2 accumulated (nrmp(d)^2).
template <typename T> struct OpTangentIn {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    const T *p1;
    T *d;
    __device__ __forceinline__ const T *aux() const { return p1; }
    __device__ __forceinline__ T operator()(T x, T b, long g, double *acc) const
    {
        const T di = b - x;
        d[g] = di;
        acc[0] += (double)di * (double)di;
        return di;
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const { return (*this)(x, p1[g], g, acc); }
};

// The output pass of that G convolution: with w = (G*d)(g), the sums for
// <f_r, d> of both replicas without forming f_r = ngrad_cam(Phi_r)
// (LPsrc.cpp:147-169; <G*x, d> = <x, G*d> for the even kernel G):
// per replica r the sums of x_r w, nl_r d and nl_r (nl = Tau x + Gamma x^2 - x^3),
// then the sum of d.  G*d itself is not stored.
template <typename T> struct OpFvOut {
    static constexpr int NRED = 7;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    const T *x0, *x1, *d;
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const
    {
        const double w = (double)x, vi = (double)d[g];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const double u = (double)(r == 0 ? x0 : x1)[g];
            const double nl = Tau * u + Gamma * u * u - u * u * u;
            acc[3 * r] += u * w;
            acc[3 * r + 1] += nl * vi;
            acc[3 * r + 2] += nl;
        }
        acc[6] += vi;
        return x;
    }
};

// ngrad_cam epilogue (LPsrc.cpp:147-169): x = ifft(G phi_hat, DC = 0);
// grad = -((x - Tau u - Gamma u^2 + u^3) - mean), mean = -mean(f).
// Sums grad^2 for nrmp; grad = nullptr keeps only the sum.
template <typename T> struct OpNgradOut {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    static constexpr int CHUNK = 8;
    double Tau, Gamma;
    const T *phi;
    const double *mean_f; // device scalar: mean of f(phi)
    T *grad;
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const
    {
        const double u = (double)phi[g], u2 = u * u, u3 = u2 * u;
        const double gr = (double)x - Tau * u - Gamma * u2 + u3;
        const double out = -(gr + *mean_f);
        if (grad) grad[g] = (T)out;
        acc[0] += out * out;
        return x;
    }
};

// hv_cam input pass: transform v, sum c(phi) v with c = -Tau - 2 Gamma phi + 3 phi^2.
template <typename T> struct OpHvIn {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    const T *phi;
    __device__ __forceinline__ const T *aux() const { return phi; }
    __device__ __forceinline__ T operator()(T x, T pg, long, double *acc) const
    {
        const double u = (double)pg;
        acc[0] += (-Tau - 2 * Gamma * u + 3 * u * u) * (double)x;
        return x;
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const { return (*this)(x, phi[g], g, acc); }
};

// hv_cam epilogue (LOBPCG.cpp:136-156): hv = x + c(phi) v - mean(c v), with x
// = ifft(G v_hat, DC = 0) (the DC of G v_hat is the mean of ifft(G v_hat)).
// DOT: also the sum of v hv (the stored hv), for <v, H v> (LOBPCG.cpp:281-283).
template <typename T, bool DOT = false> struct OpHvOut {
    static constexpr int NRED = DOT ? 1 : 0;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma;
    const T *phi, *v;
    const double *mean_cv;
    T *hv;
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const
    {
        const double u = (double)phi[g], vg = (double)v[g];
        const T y = (T)((double)x + (-Tau - 2 * Gamma * u + 3 * u * u) * vg - *mean_cv);
        hv[g] = y;
        if constexpr (DOT) acc[0] += vg * (double)y;
        return x;
    }
};

// The residual u - alpha v of one LOBPCG block member (LOBPCG.cpp:284), formed
// as the input of the preconditioner's first pass instead of in a field of its own.
template <typename T> struct OpResid {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    double alpha;
    const T *v;
    __device__ __forceinline__ const T *aux() const { return v; }
    __device__ __forceinline__ T operator()(T x, T vg, long, double *) const
    {
        return (T)((double)x - alpha * (double)vg);
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const { return (*this)(x, v[g], g, acc); }
};

// The same three operators over a block of fields processed as batch members
// (blockIdx.y = m): member m reads and writes the fields at m * fs elements
// past the first, and its own mean / alpha at index m of a device array.
template <typename T, bool DOT = false> struct OpHvOutM : OpHvOut<T, DOT> {
    long fs;
    const double *means;
    __device__ __forceinline__ OpHvOutM bind(int m) const
    {
        OpHvOutM o = *this;
        o.v = this->v + (long)m * fs;
        o.hv = this->hv + (long)m * fs;
        o.mean_cv = means + m;
        return o;
    }
};
template <typename T> struct OpResidM : OpResid<T> {
    long fs;
    const double *alphas;
    __device__ __forceinline__ OpResidM bind(int m) const
    {
        OpResidM o = *this;
        o.v = this->v + (long)m * fs;
        o.alpha = alphas[m];
        return o;
    }
};
template <typename T> struct OpStoreM : OpStore<T> {
    long fs;
    __device__ __forceinline__ OpStoreM bind(int m) const
    {
        OpStoreM o = *this;
        o.out = this->out + (long)m * fs;
        return o;
    }
};

// Natural_Forces / MinMax_Forces input (SPM.cpp:53-86 with beta = 1): the
// three transforms FFT(x) + FFT(nln - mean) - FFT(coef <f, v> v) of the
// reference are one transform of y = x + nln(x) - coef s v, s = <f, v> (a
// device scalar); the mean of nln only moves the DC, which is set to zero.
template <typename T> struct OpForceIn {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma, coef;
    const T *v;
    const double *s;
    __device__ __forceinline__ const T *aux() const { return v; }
    __device__ __forceinline__ T operator()(T x, T vg, long, double *) const
    {
        const double u = (double)x, u2 = u * u, u3 = u2 * u;
        return (T)(u + (Tau * u + Gamma * u2 - u3) - coef * (*s) * (double)vg);
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const { return (*this)(x, v[g], g, acc); }
};

// HiSD step input (HISD.cpp:215-250): y = x + bta (Tau x + Gamma x^2 - x^3) -
// bta tmpp, tmpp the reflected part of the force (its mean only moves the DC,
// which the step zeroes).
template <typename T> struct OpHisdIn {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    double Tau, Gamma, bta;
    const T *tmpp;
    __device__ __forceinline__ const T *aux() const { return tmpp; }
    __device__ __forceinline__ T operator()(T x, T tg, long, double *) const
    {
        const double u = (double)x, u2 = u * u, u3 = u2 * u;
        return (T)(u + bta * (Tau * u + Gamma * u2 - u3) - bta * (double)tg);
    }
    __device__ __forceinline__ T operator()(T x, long g, double *acc) const { return (*this)(x, tmpp[g], g, acc); }
};

// ======================================================================
// Spectral operators applied inside the i-axis column pass.
// op(v, s, i, j, kz, acc): v = (unnormalised forward) value at spectrum index
// s with frequency indices (i, j, kz); may modify v, touch other arrays at s.
// ======================================================================
struct SOpNone {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = false, FINALIZE = false;
    template <typename T>
    __device__ __forceinline__ void operator()(C2<T> &, C2<T>, long, int, int, int, double *) const
    {
    }
};

template <typename T> struct SOpScale { // r2c completion: v *= sc
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = false, FINALIZE = false;
    double sc;
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T>, long, int, int, int, double *) const
    {
        v.x = (T)((double)v.x * sc);
        v.y = (T)((double)v.y * sc);
    }
};

// v <- sc * tab[|n|^2] * v, DC set to zero.  tab = G (ngrad_cam, hv_cam) or
// 1/(G + |Tau| + 0.02) (precond_camnew).
template <typename T, int N> struct SOpMulTab {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = false, FINALIZE = false;
    double sc;
    const double *tab;
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T>, long, int i, int j, int kz, double *) const
    {
        const int a = kfold(i, N), b = kfold(j, N);
        const int n2 = a * a + b * b + kz * kz;
        const double f = n2 == 0 ? 0.0 : sc * tab[n2];
        v.x = (T)((double)v.x * f);
        v.y = (T)((double)v.y * f);
    }
};

// HiSD step (HISD.cpp:252-262): v <- sc Re(v) / (1 + bta G), DC = 0.  Only the
// real part of each coefficient is kept, as the reference does, so the new
// field is the even part of the transform's result.
template <typename T, int N> struct SOpHisdStep {
    static constexpr int NRED = 0;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = false, FINALIZE = false;
    double sc, bta;
    const double *G;
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T>, long, int i, int j, int kz, double *) const
    {
        const int a = kfold(i, N), b = kfold(j, N);
        const int n2 = a * a + b * b + kz * kz;
        const double f = n2 == 0 ? 0.0 : sc / (1.0 + bta * G[n2]);
        v.x = (T)((double)v.x * f);
        v.y = (T)0;
    }
};

// Hermitian weight of a half-spectrum entry in a sum over the full spectrum.
template <int N> __device__ __forceinline__ double hweight(int kz) { return (kz == 0 || kz == N / 2) ? 1.0 : 2.0; }

// Spring_Pairing's force passes (Natural_Forces / MinMax_Forces, SPM.cpp:83-91):
// v <- sc B v with B = 1/(1+G) (DC = 0), which is the normalised spectrum X of
// the updated replica.  The first replica's pass stores X at the spectrum index
// (keep); the second (CROSS) reads it back and sums, over k != 0 with the
// Hermitian weights, Q_pq = sum w G Re(X_p conj X_q) for (p, q) = (0, 0), (0, 1),
// (1, 1): the means <x_p, G*x_q> of the two updated replicas (Parseval).
// FP64: the stored spectrum is loaded at the start of the pass (PRELOAD), so
// its latency overlaps the forward transform.  (FP32 pass-2 blocks of 512
// threads would then fit once per CU; they gather it where it is used.)
template <typename T, int N, bool CROSS> struct SOpForceSpec {
    static constexpr int NRED = CROSS ? 3 : 0;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = CROSS && sizeof(T) == 8, FINALIZE = false;
    double sc;
    const double *Btab, *Gtab;
    C2<T> *keep;         // !CROSS
    const C2<T> *other;  // CROSS
    __device__ __forceinline__ C2<T> load_pre(long s) const { return other[s]; }
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T> o, long s, int i, int j, int kz, double *acc) const
    {
        const int a = kfold(i, N), b = kfold(j, N);
        const int n2 = a * a + b * b + kz * kz;
        const double f = n2 == 0 ? 0.0 : sc * Btab[n2];
        v.x = (T)((double)v.x * f);
        v.y = (T)((double)v.y * f);
        if constexpr (CROSS) {
            if constexpr (!PRELOAD) o = other[s];
            const double wg = n2 == 0 ? 0.0 : hweight<N>(kz) * Gtab[n2];
            const double ox = (double)o.x, oy = (double)o.y, vx = (double)v.x, vy = (double)v.y;
            acc[0] += wg * (ox * ox + oy * oy);
            acc[1] += wg * (ox * vx + oy * vy);
            acc[2] += wg * (vx * vx + vy * vy);
        } else {
            keep[s] = v;
        }
    }
};

// Non-local energy sum of get_Ene (LPsrc.cpp:27-35): acc += w G |sc v|^2.
template <typename T, int N> struct SOpEnergy {
    static constexpr int NRED = 1;
    static constexpr unsigned MAXM = 0;
    static constexpr bool PRELOAD = false, FINALIZE = false;
    double sc;
    const double *G;
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T>, long, int i, int j, int kz, double *acc) const
    {
        const int a = kfold(i, N), b = kfold(j, N);
        const double re = (double)v.x * sc, im = (double)v.y * sc;
        acc[0] += hweight<N>(kz) * G[a * a + b * b + kz * kz] * (re * re + im * im);
    }
};

// One semi-implicit gradient-flow update (LPsrc.cpp:63-84) on the state
// spectrum U: f_hat = sc * v; U' = (dt f_hat + U) / (1 + dt G), DC = 0.
// Sums, for the pre-step state: w G |U|^2 (2 E1), w |G U - f_hat|^2 over
// k != 0 (||grad F||^2 by Parseval), and max |Re(U' - U)| (the reference's
// residual).  v becomes U' (then inverse-transformed by the same pass).
// PRE: U is loaded at the start of the pass (PRELOAD) so its latency overlaps
// the forward transform; otherwise it is gathered where it is used (fewer
// registers held across the transform).  G and D = 1 / (1 + dt G) come from one
// table entry (one gather, and a multiply instead of two FP64 divides per entry).
template <typename T, int N, bool PRE = true, bool NT = false> struct SOpGradFlowStep {
    static constexpr int NRED = 3;
    static constexpr unsigned MAXM = 4u; // acc[2] is a max
    static constexpr bool PRELOAD = PRE, FINALIZE = false;
    static constexpr bool DEFER_STORE = true; // the pass stores the new U (= the value returned in v)
    static constexpr bool GATHER_GROUP = !PRE;  // the two-pass kernel (N <= 128) loads U and the table 4 elements at a time
    double sc, dt;
    const C2<T> *GD; // (G, 1 / (1 + dt G)) per |n|^2, in the field precision
    C2<T> *U;
    long ubs; // batched: member stride of U
    // U is read and written with non-temporal hints when U and W together
    // overflow the last-level cache (so W, read again by the next pass, keeps
    // it; when both fit, the hints only cost).  FP32 takes the choice as the
    // template constant NT (a run-time choice splits every element's loads into
    // branches of their own); FP64 as the run-time flag ntr (its preloading
    // pass at 256^3 needs 20 more registers and runs slower with the constant).
    bool ntr;
    __device__ __forceinline__ bool nt() const { return sizeof(T) == 4 ? NT : ntr; }
    __device__ __forceinline__ SOpGradFlowStep bind(int m) const
    {
        SOpGradFlowStep o = *this;
        o.U = U + (long)m * ubs;
        return o;
    }
    __device__ __forceinline__ C2<T> load_pre(long s) const
    {
        if (nt()) {
            typedef T vt2 __attribute__((ext_vector_type(2)));
            const vt2 w = __builtin_nontemporal_load((const vt2 *)(U + s));
            return C2<T>{w.x, w.y};
        }
        return U[s];
    }
    __device__ __forceinline__ void store_u(long s, C2<T> x) const
    {
        if (nt()) {
            typedef T vt2 __attribute__((ext_vector_type(2)));
            vt2 w;
            w.x = x.x;
            w.y = x.y;
            __builtin_nontemporal_store(w, (vt2 *)(U + s));
        } else {
            U[s] = x;
        }
    }
    __device__ __forceinline__ static int n2of(int i, int j, int kz)
    {
        const int a = kfold(i, N), b = kfold(j, N);
        return a * a + b * b + kz * kz;
    }
    __device__ __forceinline__ C2<T> load_gd(int n2) const // FP32: one 8-byte vector load per entry
    {
        if constexpr (sizeof(T) == 4) {
            typedef T vt2 __attribute__((ext_vector_type(2)));
            const vt2 w = *(const vt2 *)(GD + n2);
            return C2<T>{w.x, w.y};
        }
        return GD[n2];
    }
    __device__ __forceinline__ void operator()(C2<T> &v, C2<T> u, long s, int i, int j, int kz, double *acc) const
    {
        const int n2 = n2of(i, j, kz);
        const C2<T> gd = load_gd(n2); // gathered here: fetching all E entries up front costs occupancy
        if constexpr (!PRELOAD) u = load_pre(s);
        update(v, u, gd, n2, kz, acc);
    }
    // the update of one element from its state u and table entry gd (FP32 without
    // branches: the DC entry is selected, so the loads of neighbouring elements can
    // be issued together; in FP64 that cost registers at 256^3)
    __device__ __forceinline__ void update(C2<T> &v, C2<T> u, C2<T> gd, int n2, int kz, double *acc) const
    {
        const bool dc = n2 == 0;
        if constexpr (sizeof(T) == 4) { // FP32 fields: the update in FP32, the three sums in FP64
            // every rounding written out (fused where fmaf says, nothing else
            // contracted), so all step variants compute the same values
#pragma clang fp contract(off)
            const float g = gd.x, d = gd.y, w = (kz == 0 || kz == N / 2) ? 1.0f : 2.0f;
            const float fr = v.x * (float)sc, fi = v.y * (float)sc, ur = u.x, ui = u.y, dtf = (float)dt;
            const float nr = dc ? 0.0f : fmaf(dtf, fr, ur) * d, ni = dc ? 0.0f : fmaf(dtf, fi, ui) * d;
            const float dr = fmaf(g, ur, -fr), di = fmaf(g, ui, -fi);
            acc[1] += dc ? 0.0 : (double)(w * fmaf(dr, dr, di * di));
            acc[0] += (double)(w * g * fmaf(ur, ur, ui * ui));
            acc[2] = fmax(acc[2], (double)fabsf(nr - ur));
            v = C2<T>{nr, ni};
            return;
        }
        const double g = (double)gd.x, d = (double)gd.y, w = hweight<N>(kz);
        const double fr = (double)v.x * sc, fi = (double)v.y * sc;
        const double ur = (double)u.x, ui = (double)u.y;
        double nr = (dt * fr + ur) * d, ni = (dt * fi + ui) * d;
        if (dc) {
            nr = 0.0;
            ni = 0.0;
        } else {
            const double dr = g * ur - fr, di = g * ui - fi;
            acc[1] += w * (dr * dr + di * di);
        }
        acc[0] += w * g * (ur * ur + ui * ui);
        const C2<T> un = {(T)nr, (T)ni};
        acc[2] = fmax(acc[2], fabs((double)un.x - ur));
        v = un;
    }
};

// ======================================================================
// Block-0 prologues of a column pass: small serial work that must follow the
// previous kernel (its partial sums are complete and visible at the kernel
// boundary) and needs no launch of its own.
// ======================================================================
struct ProNone {
    static constexpr bool NONE = true;
    __device__ int nblk() const { return 0; }
    __device__ void operator()(double *) const {}
};

// The two halves of closing one gradient-flow iteration.  Partial sums:
// pA from the i pass (w G |U|^2, w |G U - f_hat|^2, max residual), pC from
// the row pass of the same field (density).
// First level (blocks b < kf of the next pass): block b reduces its share of
// pA and pC into p2[q * kf + b].
struct ProGradFlowReduce {
    static constexpr bool NONE = false;
    const double *pA;
    long nA;
    const double *pC;
    long nC;
    double *p2;
    int kf;
    __device__ __forceinline__ ProGradFlowReduce bind(int m) const
    {
        ProGradFlowReduce o = *this;
        o.pA = pA + (long)m * 3 * nA;
        o.pC = pC + (long)m * nC;
        o.p2 = p2 + (long)m * 4 * kf;
        return o;
    }
    __device__ int nblk() const { return kf; }
    __device__ void operator()(double *red) const
    {
        const long b = blockIdx.x;
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (long i = nA * b / kf + threadIdx.x; i < nA * (b + 1) / kf; i += blockDim.x) {
            v[0] += pA[i];
            v[1] += pA[nA + i];
            v[2] = fmax(v[2], pA[2 * nA + i]);
        }
        for (long i = nC * b / kf + threadIdx.x; i < nC * (b + 1) / kf; i += blockDim.x) v[3] += pC[i];
        block_reduce<4, 4u>(v, red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) p2[q * kf + b] = v[q];
        __syncthreads();
    }
};

// Second level (block 0 of the pass after that): E = E1 + E2, ||grad F||,
// residual into hist[3*iter ..], iter += 1, and the stop test of LPsrc.cpp:52.
struct ProGradFlowFinalize {
    static constexpr bool NONE = false;
    const double *p2;
    int kf;
    double R;
    Ctl *ctl;
    double *hist;
    long hstride; // batched: member stride of hist (doubles)
    __device__ __forceinline__ ProGradFlowFinalize bind(int m) const
    {
        ProGradFlowFinalize o = *this;
        o.p2 = p2 + (long)m * 4 * kf;
        o.ctl = ctl + m;
        o.hist = hist + (long)m * hstride;
        return o;
    }
    __device__ int nblk() const { return 1; }
    __device__ void operator()(double *red) const
    {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = threadIdx.x; b < kf; b += blockDim.x) {
            v[0] += p2[b];
            v[1] += p2[kf + b];
            v[2] = fmax(v[2], p2[2 * kf + b]);
            v[3] += p2[3 * kf + b];
        }
        block_reduce<4, 4u>(v, red);
        if (threadIdx.x == 0) {
            const double E = 0.5 * v[0] + v[3] / R, gn = sqrt(v[1]), res = v[2];
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

// ======================================================================
// Kernels
// ======================================================================

template <int NRED, unsigned MAXM>
__device__ __forceinline__ void write_partials(double (&acc)[NRED > 0 ? NRED : 1], double *red, double *part)
{
    if constexpr (NRED > 0) {
        block_reduce<NRED, MAXM>(acc, red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < NRED; ++q)
                part[((long)blockIdx.y * NRED + q) * gridDim.x + blockIdx.x] = acc[q];
    }
}

// Twiddle table exp(-2 pi i e / N) into LDS (padded, twpad).  The first Stockham stage uses no
// twiddles and ends in a barrier, so no barrier is needed here.
template <typename T, int N>
__device__ __forceinline__ void load_twiddles(C2<T> *stw, const C2<T> *__restrict__ tw)
{
    for (int e = threadIdx.x; e < N; e += blockDim.x) stw[twpad(e)] = tw[e];
}

// Row (k-axis) pass over npairs pairs of rows.
//   IN_SPEC:  input rows are half spectra (row stride NHP): c2r (unnormalised)
//   !IN_SPEC: input rows are real (row stride N)
//   OUT_SPEC: forward r2c (unnormalised) of op(x) into out (row stride NHP)
// Half-spectrum input is read straight into the pencil layout (each element
// by the threads that need it; the second read of an element hits the cache);
// real input rows are contiguous and staged through LDS with 16-byte loads.
// Stockham exchanges and the final separation of the two packed rows go
// through LDS as real parts, then imaginary parts (half the LDS of complex
// exchanges).
template <typename T, int N, int TPP_, bool IN_SPEC, bool OUT_SPEC, class Op, class Pro = ProNone>
__global__ void __launch_bounds__((Geo<T, N, TPP_>::ROW_THREADS))
    k_rows(const void *in, C2<T> *out, int NHP, long npairs, const C2<T> *__restrict__ tw, Op op0,
           double *__restrict__ part, const Ctl *ctl, Pro pro0, long in_bs, long out_bs)
{
    using G = Geo<T, N, TPP_>;
    constexpr int TPP = G::TPP, E = G::E, PP = G::PP, NH = N / 2 + 1;
    constexpr int XCH = PP * G::RSTR, STG = IN_SPEC ? 0 : 2 * PP * N; // in reals
    __shared__ __align__(16) T lds[XCH > STG ? XCH : STG];
    __shared__ C2<T> stw[twiddle_slots<N>()];
    __shared__ double red[(Op::NRED > 0 ? Op::NRED : 1) * 16 + (Pro::NONE ? 0 : 64)];
    const int mb = blockIdx.y; // batch member
    if (ctl && ctl[mb].stop) return;
    const Op op = bind_member(op0, mb);
    if constexpr (!Pro::NONE) {
        const Pro pro = bind_member(pro0, mb);
        if ((int)blockIdx.x < pro.nblk()) pro(red);
    }
    if constexpr (IN_SPEC)
        in = (const C2<T> *)in + mb * in_bs;
    else
        in = (const T *)in + mb * in_bs;
    if (out) out += mb * out_bs;
    load_twiddles<T, N>(stw, tw);
    const int t = threadIdx.x % TPP, pp = threadIdx.x / TPP;
    const long pair = (long)blockIdx.x * PP + pp;
    const bool active = pair < npairs;
    const long ra = 2 * pair, rb = ra + 1;
    const long row0 = 2L * PP * blockIdx.x;
    const long nrows = 2 * npairs;
    T *my = lds + pp * G::RSTR;
    C2<T> v[E];
    double acc[Op::NRED > 0 ? Op::NRED : 1];
#pragma unroll
    for (int q = 0; q < (Op::NRED > 0 ? Op::NRED : 1); ++q) acc[q] = 0.0;

    if constexpr (IN_SPEC) {
        const C2<T> *la = (const C2<T> *)in + ra * NHP, *lb = la + NHP;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPP * m;
            const int q = p <= N / 2 ? p : N - p;
            C2<T> a = active ? la[q] : C2<T>{0, 0}, b = active ? lb[q] : C2<T>{0, 0};
            if (q == 0 || q == N / 2) { // c2r reads only the real part of DC and Nyquist
                a.y = 0;
                b.y = 0;
            }
            // Z = Xa + i Xb on 0..N/2, conj(Xa) + i conj(Xb) above
            v[m] = p <= N / 2 ? C2<T>{a.x - b.y, a.y + b.x} : C2<T>{a.x + b.y, b.x - a.y};
        }
        fft_pencil_split<T, N, TPP, +1, 1, true>(v, my, stw, t);
    } else {
        // 16-byte chunks of the block's 2*PP contiguous real rows
        constexpr int VPC = 16 / sizeof(T);
        using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
        const T *X = (const T *)in;
        const V *src = (const V *)(X + row0 * N);
        V *dstl = (V *)lds;
        const long nval = (nrows - row0 < 2 * PP ? nrows - row0 : 2 * PP) * (long)N;
        for (int c = threadIdx.x; c < 2 * PP * N / VPC; c += blockDim.x)
            if ((long)c * VPC < nval) dstl[c] = src[c];
        __syncthreads();
        const T *lr = lds + (2 * pp) * N;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPP * m;
            v[m] = C2<T>{lr[p], lr[N + p]};
        }
    }
    if (active) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPP * m;
            v[m].x = op(v[m].x, ra * N + p, acc);
            v[m].y = op(v[m].y, rb * N + p, acc);
        }
    }
    if constexpr (OUT_SPEC) {
        fft_pencil_split<T, N, TPP, -1, 1, true>(v, my, stw, t);
        // separation of the two rows: needs Z[k] and Z[N - k] (another thread's)
        constexpr int EH = E / 2 + 1;
        T ax[EH], bx[EH];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) my[t + TPP * m] = v[m].x;
        __syncthreads();
#pragma unroll
        for (int m = 0; m < EH; ++m) {
            const int k = t + TPP * m;
            ax[m] = k <= N / 2 ? my[k] : (T)0;
            bx[m] = k <= N / 2 ? my[(N - k) % N] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) my[t + TPP * m] = v[m].y;
        __syncthreads();
        if (active) {
#pragma unroll
            for (int m = 0; m < EH; ++m) {
                const int k = t + TPP * m;
                if (k <= N / 2) {
                    const T ay = my[k], by = my[(N - k) % N];
                    // A = Z[k], B = Z[N-k]: Xa = (A + conj B)/2, Xb = -i (A - conj B)/2
                    const T hx = (T)0.5;
                    out[ra * NHP + k] = C2<T>{hx * (ax[m] + bx[m]), hx * (ay - by)};
                    out[rb * NHP + k] = C2<T>{hx * (ay + by), hx * (bx[m] - ax[m])};
                }
            }
        }
    }
    write_partials<Op::NRED, Op::MAXM>(acc, red, part);
}

// Complex rows (the reference's full c2c layout): length-N transforms of
// nrows contiguous rows, out = sc * DFT_DIR(in).
template <typename T, int N, int TPP_, int DIR>
__global__ void __launch_bounds__((Geo<T, N, TPP_>::ROW_THREADS))
    k_rows_c2c(const C2<T> *__restrict__ in, C2<T> *__restrict__ out, long nrows, double sc,
               const C2<T> *__restrict__ tw)
{
    using G = Geo<T, N, TPP_>;
    constexpr int TPP = G::TPP, E = G::E, PP = G::PP;
    __shared__ C2<T> lds[PP * G::RSTR];
    __shared__ C2<T> stw[twiddle_slots<N>()];
    load_twiddles<T, N>(stw, tw);
    const int t = threadIdx.x % TPP, pp = threadIdx.x / TPP;
    const long row = (long)blockIdx.x * PP + pp;
    const bool active = row < nrows;
    C2<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = active ? in[row * N + t + TPP * m] : C2<T>{0, 0};
    fft_pencil<T, N, TPP, DIR, 1, true>(v, lds + pp * G::RSTR, stw, t);
    if (active)
#pragma unroll
        for (int m = 0; m < E; ++m) out[row * N + t + TPP * m] = cscale(v[m], (T)sc);
}

// Column pass, pencils of length N along the j or i axis.
// Element e of column col of outer block o: src[o*ostride + e*pstride + col].
// Valid columns: col < ncols and (col mod NHP) < NH.  iaxis: the pencil runs
// along i (frequencies i = e, j = col / NHP, kz = col mod NHP); otherwise
// along j (i = o, j = e, kz = col).
template <typename T, int N, int TPP_, bool FWD, bool INV, bool STORE, class SOp, class Pro = ProNone>
__global__ void __launch_bounds__((Geo<T, N, TPP_>::COL_THREADS))
    k_cols(const C2<T> *src, C2<T> *dst, long pstride, long ncols, long ostride, int NH, int NHP,
           int iaxis, const C2<T> *__restrict__ tw, SOp op0, double *__restrict__ part,
           const Ctl *ctl, Pro pro0, long bs)
{
    using G = Geo<T, N, TPP_>;
    constexpr int TPP = G::TPP, E = G::E, B = G::B;
    __shared__ typename G::XT lds[N * B];
    __shared__ C2<T> stw[twiddle_slots<N>()];
    __shared__ double red[(SOp::NRED > 0 ? SOp::NRED : 1) * 16 + (Pro::NONE ? 0 : 64)];
    const int mb = blockIdx.y; // batch member
    if (ctl && ctl[mb].stop) return;
    const SOp op = bind_member(op0, mb);
    if constexpr (!Pro::NONE) {
        const Pro pro = bind_member(pro0, mb);
        if ((int)blockIdx.x < pro.nblk()) pro(red);
    }
    src += mb * bs;
    if (dst) dst += mb * bs;
    load_twiddles<T, N>(stw, tw);
    const int c = threadIdx.x % B, t = threadIdx.x / B;
    const long tpo = (ncols + B - 1) / B;
    const long o = blockIdx.x / tpo;
    const long col = (blockIdx.x % tpo) * B + c;
    const int kz = (int)(col % NHP);
    const bool active = col < ncols && kz < NH;
    const long off = o * ostride + col;
    C2<T> v[E];
    C2<T> pre[SOp::PRELOAD ? E : 1];
    double acc[SOp::NRED > 0 ? SOp::NRED : 1];
#pragma unroll
    for (int q = 0; q < (SOp::NRED > 0 ? SOp::NRED : 1); ++q) acc[q] = 0.0;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = active ? src[off + (long)(t + TPP * m) * pstride] : C2<T>{0, 0};
    if constexpr (SOp::PRELOAD) {
#pragma unroll
        for (int m = 0; m < E; ++m) pre[m] = active ? op.load_pre(off + (long)(t + TPP * m) * pstride) : C2<T>{0, 0};
    }
    if constexpr (FWD) {
        if constexpr (G::XSPLIT)
            fft_pencil_split<T, N, TPP, -1, B>(v, lds + c, stw, t);
        else
            fft_pencil<T, N, TPP, -1, B, false>(v, lds + c, stw, t);
    }
    if (active) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int e = t + TPP * m;
            const int i = iaxis ? e : (int)o;
            const int j = iaxis ? (int)(col / NHP) : e;
            op(v[m], pre[SOp::PRELOAD ? m : 0], off + (long)e * pstride, i, j, kz, acc);
        }
        if constexpr (defers_store<SOp>::value)
#pragma unroll
            for (int m = 0; m < E; ++m) op.store_u(off + (long)(t + TPP * m) * pstride, v[m]);
    }
    if constexpr (INV) {
        if constexpr (G::XSPLIT)
            fft_pencil_split<T, N, TPP, +1, B>(v, lds + c, stw, t);
        else
            fft_pencil<T, N, TPP, +1, B, false>(v, lds + c, stw, t);
    }
    if constexpr (STORE) {
        if (active)
#pragma unroll
            for (int m = 0; m < E; ++m) dst[off + (long)(t + TPP * m) * pstride] = v[m];
    }
    write_partials<SOp::NRED, SOp::MAXM>(acc, red, part);
}

} // namespace spm
