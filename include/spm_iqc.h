/*
 * spm_iqc.h -- C ABI of the MI355X (gfx950) implementation of the
 * pseudo-spectral Landau free-energy path of JKGroupCUDA/SPM-IQC.
 *
 * The reference has no FFI: its interface is the set of free C++ functions in
 * its include/ headers, operating on process globals (Data.h:4-28).  This header is the
 * plain-C boundary those functions map onto, one entry point per reference
 * function (cited beside each).  The C++ layer with the reference's own
 * signatures (spm_compat.h) is built on top of it.
 *
 * Conventions
 *   - All field pointers are DEVICE pointers.  A real field is N^3 values of
 *     the context precision (double for SPM_FP64, float for SPM_FP32) in the
 *     reference's x-major order idx = (i*N + j)*N + k.  A half spectrum is
 *     N*N*nhp complex values (interleaved re, im) indexed (i*N + j)*nhp + kz,
 *     kz = 0..N/2, normalised by 1/N^3 like the reference's u_F_local.
 *   - Every call is enqueued on the context's HIP stream.  Calls that return a
 *     host scalar (double *) synchronise that stream.
 *   - Return value: 0 on success, a negative SPM_E* code otherwise;
 *     spm_last_error() gives a message.  (The reference has no error
 *     reporting: a failure there crashes the process.)
 *   - A context is not thread-safe; use one context per thread / request.
 */
#ifndef SPM_IQC_H
#define SPM_IQC_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SPM_FP64 0
#define SPM_FP32 1

#define SPM_OK 0
#define SPM_EINVAL (-1)   /* bad argument (size, pointer, parameter) */
#define SPM_ENOMEM (-2)   /* device allocation failed */
#define SPM_EHIP (-3)     /* HIP runtime error */
#define SPM_ENOTSUP (-4)  /* unsupported grid size or precision */

typedef struct spm_ctx spm_ctx;

/* Model and grid parameters: LPsystParameter (Initialization.cpp:7-42).
 * kscale: reciprocal metric k_i = kscale * n_i / L; the reference hard-codes
 * 1 (ProjMatrix = 1/L, Initialization.cpp:35-38). */
typedef struct {
    int N;          /* grid points per side: 4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256 */
    int precision;  /* SPM_FP64 or SPM_FP32 */
    double L;       /* box length (26) */
    double kscale;  /* 1.0 = the reference's metric; 2*pi = 2 pi n / L */
    double Tau;     /* -0.0061 */
    double Gamma;   /* 0.30 */
    double dt;      /* gradient-flow step (0.3) */
    int shape;      /* FFT pass shape: 0 = built-in default, k = shape k of this N
                       (spm_shapes), -1 = the tuned one (spm_tune; measured on
                       first use and kept in the tuning cache) */
    int step_passes; /* array passes per gradient-flow step (and per transform
                       round trip of the operators): 0 = default (2 at N = 64,
                       128 and 256, 4 otherwise), 4 = one pass per axis (any
                       N), 2 = the j transform split between the row and i
                       passes (N = 64, 128 and 256 only) */
    int step_variant; /* variant of the two-pass gradient-flow step: 0 = the
                       built-in choice for N and precision, -1 = the tuned one
                       (spm_tune_step; measured on first use and kept in the
                       tuning cache), or SPM_STEP_EXPLICIT | SPM_STEP_* flags.
                       Every variant computes the same values. */
} spm_params;

/* Two-pass step variants (spm_params.step_variant). */
#define SPM_STEP_EXPLICIT 8  /* the flags below are the variant */
#define SPM_STEP_GATHER_U 1  /* pass 2 gathers the state spectrum where it is used
                                (otherwise: loaded at the start of the pass) */
#define SPM_STEP_U_CACHED 2  /* state loads and stores without non-temporal hints */
#define SPM_STEP_U_STREAM 4  /* ... with them (neither: with them when the state and
                                the work spectrum exceed 200 MB together) */
#define SPM_STEP_U_MASK 6
#define SPM_STEP_VARIANTS 6  /* flag sets 0..5 (U_CACHED | U_STREAM is not one) */

typedef struct {
    int N, nh, nhp, precision;
    long long R;          /* N^3 */
    long long C;          /* N*N*nhp half-spectrum entries */
    size_t real_bytes;    /* bytes of one real field */
    size_t spec_bytes;    /* bytes of one half spectrum */
    size_t ctx_bytes;     /* device bytes owned by the context */
} spm_layout;

/* --- context: LPsystParameter + memAllocation_local + get_ksquare_local +
 *     get_epmckt2_local + get_ikt2_local + My_mpi_fft_set_up
 *     (Initialization.cpp:7-174, My_FFT.cpp:41-98); spm_destroy replaces
 *     Memfree_local_kspace / Memfree_local_data (Memfree.cpp:23-60). */
int spm_create(const spm_params *p, spm_ctx **out);
int spm_destroy(spm_ctx *c);
/* Stream for all later calls (NULL = the legacy default stream).  A new
 * context uses a non-blocking stream of its own. */
int spm_set_stream(spm_ctx *c, void *hip_stream);
int spm_get_layout(const spm_ctx *c, spm_layout *out);
const char *spm_last_error(void);

/* get_GK (Initialization.cpp:150-166), host function */
double spm_get_GK(double k2);

/* --- My_FFT (My_FFT.cpp:4-21): c2c on the reference's full complex layout
 *     (N^3 interleaved doubles/floats); forward normalised by 1/N^3, inverse
 *     unnormalised.  in may equal out. */
int spm_my_fft(spm_ctx *c, const void *x_R, void *x_F);
int spm_my_ifft(spm_ctx *c, const void *x_F, void *x_R);

/* Real-field transforms on the half-spectrum layout (the forms the path
 * uses internally): S = FFT(x)/N^3; x = IFFT(S), S preserved. */
int spm_r2c(spm_ctx *c, const void *x, void *S);
int spm_c2r(spm_ctx *c, const void *S, void *x);

/* --- LPsrc */
/* get_Ene (LPsrc.cpp:8-38) of the real field phi (u_F = FFT(phi)/N^3). */
int spm_energy(spm_ctx *c, const void *phi, double *E);
/* get_Ene (LPsrc.cpp:8-38) as the reference evaluates it from its two carried
 * states: the density term from the real field phi (u_R_local), the non-local
 * term 1/2 sum_k G |u_F_k|^2 from u_F, a full complex spectrum of N^3 values of
 * the context precision (the reference's u_F_local layout, normalised by
 * 1/N^3), which need not be the transform of phi. */
int spm_energy_uf(spm_ctx *c, const void *phi, const void *u_F, double *E);
/* ngrad_cam (LPsrc.cpp:119-173): grad = -dF/dphi; *nrm (may be NULL)
 * receives nrmp(grad) (BasicOperators.cpp:158-164). */
int spm_ngrad(spm_ctx *c, const void *phi, void *grad, double *nrm);
/* get_min_u_F (LPsrc.cpp:40-117) on phi, in place.  Iterates while
 * res > res_tol and ||grad F|| > gn_tol (gn_tol <= 0 disables it) and fewer
 * than max_iter iterations (the reference: res_tol 1e-16, max_iter 100000).
 * hist (host, may be NULL) receives 3 doubles per iteration: E and
 * ||grad F|| of the pre-step field and the residual max|Re(u_F' - u_F)|.
 * check_every: iterations queued between host convergence checks (>= 1).
 * *iters receives the iteration count. */
int spm_gradflow(spm_ctx *c, void *phi, int max_iter, double res_tol, double gn_tol, int check_every,
                 double *hist, int *iters);

/* The same loop in three asynchronous parts, for callers that drive the
 * iterations themselves (timing, overlap): begin = set-up of get_min_u_F's
 * state from phi; steps = enqueue nsteps iterations (those past the stop
 * condition do nothing); end = write the final field to phi (if not NULL),
 * synchronise, return the history as spm_gradflow does.
 * Iterations are replayed 16 at a time from a HIP graph on the context's own
 * stream, ordered after earlier work and before later work of the context's
 * stream by events (spm_gradflow does the same). */
int spm_gf_begin(spm_ctx *c, const void *phi, int max_iter, double res_tol, double gn_tol);
int spm_gf_steps(spm_ctx *c, int nsteps);
int spm_gf_end(spm_ctx *c, void *phi, double *hist, int *iters);

/* --- LOBPCG */
/* hv_cam (LOBPCG.cpp:113-166): hv = Hessian(phi) v minus its mean. */
int spm_hv(spm_ctx *c, const void *phi, const void *v, void *hv);
/* precond_camnew (LOBPCG.cpp:168-199), in place. */
int spm_precond(spm_ctx *c, void *w);

/* hiosd_lobpcg_initialization (LOBPCG.cpp:10-109): V (k fields, contiguous)
 * <- the random start block (k fields; the reference draws it from Eigen's
 * MatrixXd::Random) with each field's mean removed, orthonormalised in order
 * under inpfp.  start may equal V.  1 <= k <= 8. */
int spm_lobpcg_init(spm_ctx *c, const void *start, void *V, int k);
/* hiosd_lobpcg_inip (LOBPCG.cpp:207-487): maxiter iterations on V (k fields,
 * in place) for the Hessian at phi.  ritz (host, k doubles, may be NULL)
 * receives the lowest k Ritz values of the last Rayleigh-Ritz step (the
 * reference only prints them, LOBPCG.cpp:406-410).  1 <= k <= 8. */
int spm_lobpcg(spm_ctx *c, const void *phi, void *V, int k, int maxiter, double *ritz);
/* The block Gram of the Rayleigh-Ritz step (Pn = UUm * YYm^T / N^3,
 * LOBPCG.cpp:336-358): out[a*mb + b] = inpfp(X_a, Y_b) for the contiguous
 * blocks X (ma fields) and Y (mb fields), 1 <= ma, mb <= 32; FP64
 * accumulation on the matrix cores; out is a host array. */
int spm_gram(spm_ctx *c, const void *X, int ma, const void *Y, int mb, double *out);
/* rayleighq (BasicOperators.cpp:174-185): <v, H v> / <v, v>. */
int spm_rayleighq(spm_ctx *c, const void *phi, const void *v, double *out);

/* hiosd_lobpcg_sieh (HISD.cpp:86-485): high-index saddle dynamics for an
 * index-k saddle, in place on x and on its k unstable directions V (k
 * contiguous fields).  Each step is the semi-implicit step of x under the force
 * with its components along V reflected, with a Barzilai-Borwein step size
 * clamped to [betau dt, betat dt] and only the real part of the new spectrum
 * kept (as the reference does), then one LOBPCG sweep for V at the new x.  Stops
 * when nrmp(ngrad(x)) < epsf or after maxiter steps; *iters and *res (may be
 * NULL) receive the step count and the last residual.  The reference's driver
 * values are dt 0.8, betat 8, betau 0.02, epsf 1e-7, maxiter 10000
 * (HISD.cpp:16-22).  1 <= k <= 8. */
int spm_hisd(spm_ctx *c, void *x, void *V, int k, double dt, double betat, double betau, double epsf, int maxiter,
             int *iters, double *res);
/* hiosd_lobpcg_sieh_V0 (GOSD.cpp:123-531): spm_hisd with the k0 fixed
 * orthonormal fields V0 (the null space, k0 contiguous fields, read only, not
 * overlapping x or V): the embedded LOBPCG keeps its search directions W, Vp
 * and the updated V orthogonal to V0 (GOSD.cpp:206, 227, 347).  k0 = 0 is
 * spm_hisd.  The reference's drivers use k0 = 3, dt 0.8, betat 8, betau 1.3
 * (0.4 for critical_point_v), epsf 1e-9 (1e-7), maxiter 40000 (20000)
 * (GOSD.cpp:13-120).  1 <= k <= 8, 2 k + k0 <= 16. */
int spm_gosd(spm_ctx *c, void *x, void *V, int k, const void *V0, int k0, double dt, double betat, double betau,
             double epsf, int maxiter, int *iters, double *res);

/* --- SPM */
/* get_Tv (SPM.cpp:15-32): v = (phi1 - phi0) / nrmp(phi1 - phi0). */
int spm_get_tv(spm_ctx *c, const void *phi0, const void *phi1, void *v);
/* Natural_Forces (SPM.cpp:34-99) and MinMax_Forces (SPM.cpp:103-168; not
 * declared in the reference's SPM.h): one semi-implicit force step on x in
 * place along the unit tangent v. */
int spm_natural_force(spm_ctx *c, void *x, const void *v);
int spm_minmax_force(spm_ctx *c, void *x, const void *v);
/* Spring_Pairing (SPM.cpp:170-263) on the replicas phi0, phi1 in place:
 * iterates while res > tol and fewer than maxiter iterations (the reference
 * driver: 30000, 1e-9, main_SPM.cpp:122).  *iters, *res receive the count
 * and the last residual max_i nrmp(ngrad(phi_i)); res_hist (host, may be
 * NULL, maxiter doubles) the residual of every iteration. */
int spm_spring_pairing(spm_ctx *c, void *phi0, void *phi1, int maxiter, double tol, int *iters, double *res,
                       double *res_hist);

/* --- BasicOperators: inpfp / nrmp (BasicOperators.cpp:132-172), FP64 sums. */
int spm_inpfp(spm_ctx *c, const void *p, const void *q, double *out);
int spm_nrmp(spm_ctx *c, const void *p, double *out);
/* FuncsLinear2Cplx (BasicOperators.cpp:17-26) on real fields: r = a1 F1 + a2 F2 */
int spm_lin2(spm_ctx *c, void *r, double a1, const void *F1, double a2, const void *F2);

/* --- No device allocation.  No counterpart in the reference: this entry point
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 * ngrad_cam (LPsrc.cpp:119-173) as spm_ngrad computes it (the same kernels),
 * without a context: the model, grid and precision are p's (shape >= 0: the
 * tuned shape -1 is rejected), the stream is hip_stream (NULL = the legacy
 * default stream).  Every work buffer of the call (twiddles, the G table, the
 * spectrum work array, partial sums, scalars) lives in one static device array
 * of SPM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; sizes whose
 * work does not fit (N >= 192) are rejected with SPM_EINVAL before any device
 * work.  All calls on one GPU share that array and keep its last contents, so
 * calls that overlap on one GPU (on different streams or threads) race: order
 * them.  The first call on a device, and a call whose (N, precision, L, kscale)
 * differ from the previous call's there, copy in the tables and synchronise
 * hip_stream.  *nrm (may be NULL) receives nrmp(grad); the call then
 * synchronises hip_stream. */
#define SPM_NOALLOC_SCRATCH_BYTES (24u * 1024u * 1024u)
int spm_ngrad_noalloc(const spm_params *p, const void *phi, void *grad, double *nrm, void *hip_stream);

/* --- compat helpers: real part of a full complex field <-> real field */
int spm_cplx_to_real(spm_ctx *c, const void *z, void *x);
int spm_real_to_cplx(spm_ctx *c, const void *x, void *z);

/* --- Batches of independent gradient flows (continuous batching)
 *
 * No counterpart in the reference, which relaxes one field per process
 * (main_SPM.cpp:66-72, get_min_u_F LPsrc.cpp:40-117).  A batch holds
 * `capacity` members of one grid size, precision, L, kscale and dt; each
 * member is one get_min_u_F run with its own Tau, Gamma, limits and history
 * (hist_cap iterations).  spm_batch_steps advances every running member: each
 * pass of a step is one launch over all members, and the steps are replayed
 * from a HIP graph.  Members can be loaded, read and released between calls
 * while the others keep running.  Members compute exactly what spm_gradflow
 * computes for the same field and parameters (the same kernels).
 * A batch runs on a non-blocking stream of its own (spm_batch_stream). */
typedef struct spm_batch spm_batch;
int spm_batch_create(const spm_params *p, int capacity, int hist_cap, spm_batch **out);
int spm_batch_destroy(spm_batch *b);
int spm_batch_stream(const spm_batch *b, void **hip_stream);
/* capacity, device bytes held (state, partial sums, histories) */
int spm_batch_info(const spm_batch *b, int *capacity, size_t *dev_bytes);
/* Member `slot` starts get_min_u_F from phi (device, read on the batch stream:
 * finish writing it before the call and keep it until the next poll or read)
 * with its own Tau, Gamma and limits as spm_gradflow (max_iter <= hist_cap). */
int spm_batch_load(spm_batch *b, int slot, const void *phi, double Tau, double Gamma, int max_iter,
                   double res_tol, double gn_tol);
/* Enqueue nsteps iterations of every running member (asynchronous). */
int spm_batch_steps(spm_batch *b, int nsteps);
/* Synchronise; iters and stopped (host, capacity entries each, either may be
 * NULL) receive each member's iteration count and stop flag.  Returns the
 * number of running members (>= 0) or an error code. */
int spm_batch_poll(spm_batch *b, int *iters, int *stopped);
/* Synchronise; phi (device, may be NULL) <- the member's field; hist (host,
 * may be NULL, 3 * *iters doubles) <- its history as spm_gradflow's. */
int spm_batch_read(spm_batch *b, int slot, void *phi, double *hist, int *iters);
/* Stop the member (its slot can be loaded again). */
int spm_batch_release(spm_batch *b, int slot);

/* --- FFT pass shapes and the tuner (the counterpart of My_mpi_fft_set_up's
 * FFTW_PATIENT planning with wisdom files, My_FFT.cpp:45-97).
 * A shape is the number of threads per pencil of the row (k axis), j-axis and
 * i-axis passes; it fixes the radices of each pencil FFT and the block tiles. */
/* Shapes available for N: count, and (if tpp != NULL) 3 ints per shape. */
int spm_shapes(int N, int *count, int *tpp, int max_shapes);
/* Time `steps` gradient-flow iterations of a batch of nmem members for every
 * shape of p->N in p->precision (two sweeps over the shapes, the faster time
 * of each kept); *best <- the fastest; us_per_step (host, may be NULL, one per
 * shape) <- the times.  The result is recorded in the
 * tuning cache (the file named by SPM_TUNE_CACHE, default
 * $HOME/.cache/spm-iqc/shapes.txt), which spm_create and spm_batch_create
 * read for shape -1. */
int spm_tune(const spm_params *p, int nmem, int steps, int *best, double *us_per_step);
/* The same for the variants of the two-pass step (p->N = 64, 128 or 256):
 * every flag set 0 .. SPM_STEP_VARIANTS-1 timed on `steps` graph-replayed
 * iterations of a batch of nmem members at p->shape; *best <- the fastest flag
 * set (without SPM_STEP_EXPLICIT); us_per_variant (may be NULL,
 * SPM_STEP_VARIANTS doubles) <- the times.  Recorded in the tuning cache,
 * which spm_create and spm_batch_create read for step_variant -1. */
int spm_tune_step(const spm_params *p, int nmem, int steps, int *best, double *us_per_variant);

#ifdef __cplusplus
}
#endif
#endif
