/*
 * asd.h -- C ABI of libasd, the MI355X (gfx950) implementation of the two hot
 * stages of hwchen2017/ASD-kernel-fusion's kernel_fusion path:
 *
 *   1. RK4 integration of the Landau-Lifshitz-Gilbert equation on a periodic
 *      L x L lattice  (reference: kernel_fusion/llg_sqt_fusion.c:30-145, :350-361)
 *   2. the dynamical structure factor S(q,t) as a spin-spin correlation GEMM
 *      whose epilogue carries the phase weighting and the reduction
 *      (reference: kernel_fusion/gemm_kernel.h:355-422, llg_sqt_fusion.c:235-442)
 *
 * Plain C: pointers, sizes, status codes.  No torch, no HIP types in the
 * signatures (streams are passed as void*).  Every function returns ASD_OK (0)
 * or a negative status; asd_last_error() gives a message (thread-local).
 *
 * Conventions shared with the reference:
 *   site index i = row * L + col (llg_sqt_fusion.c:51); spins are
 *   [rep][site][xyz] doubles (:55); neighbours row+1, row-1, col-1, col+1
 *   (:181-184); q = 2 pi (mx, my) / L with mx pairing with the row offset
 *   (:292-293, :322).
 */
#ifndef ASD_H
#define ASD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status --- */
enum {
    ASD_OK = 0,
    ASD_EINVAL = -1,   /* bad argument (sizes, pointers, unsupported combination) */
    ASD_ENOMEM = -2,   /* device or host allocation failed                         */
    ASD_EHIP = -3,     /* a HIP runtime call failed                                */
    ASD_ENOTSUP = -4,  /* valid request this build does not support               */
    ASD_ENOTFINITE = -5/* NaN/Inf or |S| far from spin_length detected            */
};
const char *asd_last_error(void);
const char *asd_version(void);

/* --------------------------------------------------------- parameters --- */
/* Replaces the reference globals J1 and dt (llg_sqt_fusion.c:15, :17) and adds
 * easy-axis anisotropy, a field along z and Gilbert damping.  H_eff = J sum_nn S_j + (2 D S_z + B_z) z,
 * dS/dt = [H x S + (alpha/spin_length) S x (H x S)] * (1/(1 + alpha^2)) (reciprocal
 * formed once), which is the reference's h x S (llg_sqt_fusion.c:82) exactly when
 * D = B_z = alpha = 0. */
typedef struct asd_params {
    double J;           /* nearest-neighbour exchange (reference J1 = 1.0)          */
    double D;           /* easy-axis anisotropy                                       */
    double Bz;          /* field along z                                              */
    double alpha;       /* Gilbert damping                                            */
    double dt;          /* RK4 step                                                   */
    double spin_length; /* renormalisation length: 1.0 (baseline/, gemm/), 1.5 (kernel_fusion/) */
} asd_params;

typedef enum asd_precision {
    ASD_FP64 = 0,   /* v_mfma_f64_16x16x4f64, FP64 accumulate                         */
    ASD_FP32 = 1,   /* v_mfma_f32_16x16x4f32, FP32 accumulate                         */
    ASD_BF16X3 = 2, /* a = a_hi + a_lo (bf16); a_hi b_hi + a_hi b_lo + a_lo b_hi, FP32 acc */
    ASD_FP16X2 = 3, /* a = a_hi + 2^-11 a_lo (fp16); hi and lo products in two FP32 accs */
    ASD_FP16F8 = 4  /* the same split, with the cross products a_hi b_lo + a_lo b_hi formed
                     * from fp8 (e4m3) roundings of the four parts on the fp8 MFMA            */
} asd_precision;

/* --------------------------------------------------- host-side helpers --- */
/* Marsaglia (1972) uniform spins on S^2 into host memory out[nsites][3].  Site i
 * draws from its own counter-based stream: u = U(h(seed,i,2a)), v = U(h(seed,i,2a+1)),
 * a = 0,1,... until 0 < u^2+v^2 < 1; h = splitmix64 finaliser of a linear key.
 * Replaces the reference's time-seeded integer-division init (llg_sqt_fusion.c:195-216). */
int asd_init_spins(uint64_t seed, int nsites, double *out);

/* lattice momenta of the Gamma-X-M-Gamma path: 3L/2 pairs (mx,my) into q[2*(3L/2)]:
 * (k,0), (L/2,k), (L/2-k,L/2-k) for k = 0..L/2-1.  L even. Returns the count. */
int asd_qpath_gxmg(int L, int *q);

/* ------------------------------------------------------ device context --- */
typedef struct asd_ctx asd_ctx;
/* Binds to HIP device `device`; owns workspace (ping-pong state, packed operands,
 * displacement bins).  One context per host thread / stream. */
int asd_create(asd_ctx **ctx, int device);
int asd_destroy(asd_ctx *ctx);
/* stream (hipStream_t) all work of this context is ordered on; NULL = default */
int asd_set_stream(asd_ctx *ctx, void *stream);
int asd_synchronize(asd_ctx *ctx);

/* device memory for C hosts that do not include HIP headers (stream-ordered on
 * the context's stream; asd_memcpy kind: 0 host->device, 1 device->host,
 * 2 device->device; it returns after the copy has completed) */
int asd_malloc(asd_ctx *ctx, size_t bytes, void **dptr);
int asd_free(asd_ctx *ctx, void *dptr);
int asd_memcpy(asd_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* ----------------------------------------------- stage 1: RK4-LLG (FP64) --- */
/* Native, device-resident replacement for the reference time loop's stage 1
 * (llg_sqt_fusion.c:348 runge_kutta + :350-361 renormalisation), fused into one
 * kernel per step (all four stages per LDS tile, radius-4 halo), with the
 * snapshot writer fused in.  d_spins: device [nrep][L*L][3], in/out.
 * d_snaps: device [nsnap][nrep][L*L][3] or NULL; snapshot k = state after
 * k*snap_every steps (k = 0 is the input state).  The step loop is replayed from
 * a captured hipGraph when use_graph != 0.  Returns after the work has completed:
 * ASD_ENOTFINITE (message naming the replica and the first step) when a step produced
 * an unnormalised state c with |c|^2 NaN, infinite or outside [2^-900, 2^900], the range
 * where the step is defined (the reference's own NaN initial state,
 * llg_sqt_fusion.c:204-214, is caught at step 0); the states are then unusable. */
int asd_rk4_run(asd_ctx *ctx, const asd_params *p, int L, int nrep, double *d_spins,
                long nsteps, int snap_every, int nsnap, double *d_snaps, int use_graph);

/* asd_rk4_run without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and results (bit for bit) as
 * asd_rk4_run with use_graph = 0; the second state buffer is a fixed 16 MiB scratch array
 * in the library (one per GPU, kept until the process ends), so nrep * L * L * 3 must be
 * at most 2^21 doubles (ASD_EINVAL otherwise).  It only enqueues work on the context's
 * stream (no hipMalloc, no synchronisation, no numeric guard read-back) and returns
 * before the work completes.  Calls that overlap on one GPU share the scratch and race. */
int asd_rk4_run_noalloc(asd_ctx *ctx, const asd_params *p, int L, int nrep, double *d_spins,
                        long nsteps, int snap_every, int nsnap, double *d_snaps);

/* RK4 kernel for asd_rk4_run: 0 = automatic (LDS tiles sized by occupancy, two fused
 * steps per launch on the 8 x 8, 16 x 16 / 1024-thread and 32 x 32 paths; the streaming
 * strips once the one-step strips give at least 4 waves per CU, with two fused steps per
 * launch in the tolerance mode), 8, 16 or 32 = LDS tiles of that edge, 64 =
 * register-streaming column strips (one 64-lane wave per 56 columns of a band of rows;
 * lattices with L < 64 use 16 x 16 tiles), one step per launch for every forced choice.
 * Results do not depend on the choice (bit for bit, in either arithmetic mode). */
int asd_rk4_set_tile(asd_ctx *ctx, int tile);

/* RK4 arithmetic for every RK4 entry point of this context.  exact = 1 (default): the
 * reference's operation order (llg_sqt_fusion.c:30-145, :350-361) with correctly rounded
 * quotients, so every value equals the FP64 CPU restatement's (IEEE ==).  exact = 0, the
 * tolerance mode: fused multiply-adds in the field and cross products, the stage inputs
 * S + f h and the partial sums S + f dt/6 + f dt/3 + ... as single FMAs, and the
 * renormalisation by a Newton-refined reciprocal square root (|S| = spin_length to ~1e-16);
 * fewer instructions per site (the HBM-bound sweep is issue-bound) at a trajectory
 * difference of a few 1e-15 after 100 steps (1.5e-15 to 3.2e-15 measured on every kernel
 * path).  No counterpart in the reference. */
int asd_rk4_set_exact(asd_ctx *ctx, int exact);

/* ---------------------------------------- stage 2: S(q,tau) correlation --- */
/* Native replacement for the reference's S(q,t) stage (llg_sqt_fusion.c:370-442
 * + fused_dgemm), with time origins in the role of the reference's replicas:
 *   G_ij(tau) = (1/T0) sum_{t0<T0} S_i(t0+tau) . S_j(t0)        (GEMM, K = 3 T0)
 *   S(q,tau)  = (1/Ns) sum_ij [G_ij - fbar_i(tau).zbar_j] e^{-iq.(r_i - r_j)}
 * The GEMM runs on MFMA in `prec`; its epilogue reduces G into displacement
 * bins g_tau(dr,dc) = sum_i G_{i,i-d} (exact for lattice momenta) in FP64; a
 * second kernel evaluates the bins at every q with an exact L-entry twiddle table
 * (Kahan-compensated FP64) and subtracts the separable disconnected term
 * Fbar_q(tau).conj Fbar_q(0) computed in FP64.
 * d_snaps: device [nsnap][L*L][3] FP64, nsnap >= T0 + nlags - 1.
 * h_q: host (mx,my) pairs.  out: [nlags][nq][2] (re, im) doubles, host memory
 * (out_on_device = 0) or device memory (1). */
int asd_sqt(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags,
            const int *h_q, int nq, asd_precision prec, double *out, int out_on_device);

/* The reference's unfused S(q,t) data flow (gemm/llg_sqt_mkl.c:369-423) on the GPU,
 * FP64: per lag G = (1/T0) A_tau B^T by rocBLAS dgemm into an explicit Ns x Ns
 * buffer, C = G - fbar zbar^T, then sum_ij C_ij e^{-iq.(r_i - r_j)} for every q as a
 * second dgemm against the per-site cos/sin table plus an FP64 Kahan sum.  Same
 * arguments, output layout and results (to rounding) as asd_sqt with ASD_FP64; it
 * materialises Ns^2 doubles (2.1 GB at L = 128) and exists to measure what the fused
 * path saves. */
int asd_sqt_unfused(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                    int nq, double *out, int out_on_device);

/* Brute-force S(q,t) for L <= 16 (baseline/llg_sqt_baseline.c:340-418 with its
 * disconnected term applied): per (tau, q) the double site sum and the time-origin sum
 * evaluated directly, phases from FP64 sincos (no lattice tables).  An independent
 * device-side check of the fused path; same arguments and output layout as asd_sqt. */
int asd_sqt_bruteforce(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                       int nq, double *out, int out_on_device);

/* Same contraction for the reference's replica form (llg_sqt_fusion.c:235-268,
 * :370-398): A = nrep replicas at time t, B = the same replicas at time 0, one lag.
 * d_now, d_zero: device [nrep][L*L][3].  out: [nq][2] host doubles. */
int asd_sqt_replicas(asd_ctx *ctx, int L, int nrep, const double *d_now,
                     const double *d_zero, const int *h_q, int nq, asd_precision prec,
                     double *out);

/* ----------------------------------------------------------- profiling --- */
/* Per-stage timing with HIP events recorded on the context's stream (off by
 * default).  asd_stage_times() waits for the last recorded events and returns
 * milliseconds of the most recent call of each stage:
 *   t[0] operand packing, t[1] correlation GEMM + binning epilogue (one launch
 *   covering every lag), t[2] per-snapshot Fourier sums, t[3] finishing (bins ->
 *   S(q,tau), disconnected term), t[4] RK4 step loop (asd_rk4_run). */
int asd_set_timing(asd_ctx *ctx, int on);
int asd_stage_times(asd_ctx *ctx, double t[5]);

/* ------------------------------------------------------------ batching --- */
/* Independent trajectories of uneven L in ONE persistent launch per step
 * (work list of (trajectory, tile) items), step loop captured as a hipGraph.
 * Ls[ntraj], params[ntraj], d_spins[ntraj] (device [L*L][3] each),
 * d_snaps[ntraj] (device [nsnap][L*L][3] each, or NULL array).  The kernel variant (tile
 * shape, threads per tile, one or two fused steps per launch) is the one forced with
 * asd_batch_set_variant, else the one the autotuner recorded for this batch composition
 * (the multiset of L) in the current RK4 arithmetic mode (asd_rk4_set_exact), else
 * asd_autotune_batch runs first (a few ms on scratch states).
 * Every variant gives the same bits.  Returns after the work has completed, with the
 * numeric guard of asd_rk4_run (ASD_ENOTFINITE names the trajectory and the first step).
 * The counterpart of the reference's fixed blocking (gemm_kernel.h:293-328) is the tuned
 * variant. */
int asd_batch_rk4_run(asd_ctx *ctx, int ntraj, const int *Ls, const asd_params *params,
                      double *const *d_spins, long nsteps, int snap_every, int nsnap,
                      double *const *d_snaps);

/* Times every batch kernel variant for this composition (20 steps on scratch states of the
 * same sizes and parameters, best of two runs after a warm-up, in the current RK4
 * arithmetic mode) and records the fastest for the multiset Ls and that mode; best (may be NULL) receives its index, times_ms (may be NULL,
 * asd_batch_num_variants() entries) the times. */
int asd_autotune_batch(asd_ctx *ctx, int ntraj, const int *Ls, const asd_params *params, int *best,
                       double *times_ms);
int asd_batch_num_variants(void);
const char *asd_batch_variant_name(int variant);
/* force a variant for asd_batch_rk4_run (-1 = automatic) */
int asd_batch_set_variant(asd_ctx *ctx, int variant);
/* the variant asd_batch_rk4_run would take for this composition now: the forced one, the
 * recorded one, or -1 (none recorded: the next run autotunes first) */
int asd_batch_variant_for(asd_ctx *ctx, int ntraj, const int *Ls);

/* -------------------------------------------------------- autotuning --- */
/* Times every correlation-GEMM tile configuration for (L, K = 3 T0, prec) on
 * this device (8 lags of synthetic snapshots, best of 2 timed runs each) and
 * records the fastest in the context; later asd_sqt calls with that (L, K) use it
 * unless a configuration is forced.  best_cfg (may be NULL) receives its index,
 * times_ms (may be NULL, asd_sqt_num_configs(prec) entries) the GEMM times. */
int asd_autotune_sqt(asd_ctx *ctx, int L, int T0, asd_precision prec, int *best_cfg, double *times_ms);
/* number of tile configurations for a precision, and a printable name */
int asd_sqt_num_configs(asd_precision prec);
const char *asd_sqt_config_name(asd_precision prec, int cfg);
/* force a configuration for a precision (-1 = automatic: the one the autotuner recorded
 * for (L, K), else the precision's default, the direct-to-LDS ring kernel the autotuner
 * picks on most lattices: FP64 / FP32 / FP16F8 config 3, BF16X3 / FP16X2 config 5;
 * FP16F8 runs on the direct-to-LDS kernels only; asd_sqt_num_configs gives each count) */
int asd_sqt_set_config(asd_ctx *ctx, asd_precision prec, int cfg);

/* ------------------------------------------------------------ adjoint --- */
/* Loss = sum over (lag, q) of |S(q,tau) - S_target(q,tau)|^2 for the trajectory
 * started at d_S0 (device [L*L][3]) with parameters p, nsteps RK4 steps,
 * snapshots every snap_every steps (nsnap = T0 + nlags - 1 of them), and its
 * gradient with respect to theta = (J, D, B_z, alpha), by a hand-written discrete
 * adjoint of RK4-LLG (incl. renormalisation) and of the correlation, with
 * forward states checkpointed every `checkpoint_every` steps and recomputed in
 * reverse; one forward pass serves the loss and the gradient.  h_target: host
 * [nlags][nq][2].  FP64 throughout. */
int asd_sqt_loss_grad(asd_ctx *ctx, const asd_params *p, int L, const double *d_S0,
                      long nsteps, int snap_every, int T0, int nlags, const int *h_q,
                      int nq, const double *h_target, int checkpoint_every,
                      double *loss, double grad[4]);

/* Vector-Jacobian product of the same forward map (RK4 from d_S0, snapshots,
 * FP64 S(q,tau)): grad[p] = sum_{lag,q} (g_re dRe S/dtheta_p + g_im dIm S/dtheta_p)
 * for upstream h_g [nlags][nq][2] (host).  h_sqt (host, may be NULL) receives
 * S(q,tau).  With h_g == NULL only the forward runs (grad may be NULL); with h_g and
 * no h_sqt the correlation GEMM is skipped (the reverse sweep needs only the Fourier
 * sums of the snapshots).  checkpoint_every is rounded up to a multiple of snap_every. */
int asd_sqt_vjp(asd_ctx *ctx, const asd_params *p, int L, const double *d_S0, long nsteps,
                int snap_every, int T0, int nlags, const int *h_q, int nq, const double *h_g,
                int checkpoint_every, double grad[4], double *h_sqt);

/* ---------------------------------- reference-compatible entry points --- */
/* Host-pointer entries with the reference's argument meaning; they copy to the
 * device and back on every call (for parity and drop-in use, not speed).
 *
 * asd_euler replaces `void euler(double *spin0, double *derivative)`
 * (llg_sqt_fusion.c:30): derivative = f(spin0) for nrep lattices; the reference
 * globals N, Nfile, J1 become L, nrep, p->J. */
int asd_euler(const asd_params *p, int L, int nrep, const double *spin0, double *derivative);
/* replaces `void runge_kutta(double *spin, double *cur_input, double *k1, double *k2,
 * double *k3, double *k4)` (llg_sqt_fusion.c:98): cur_input = unnormalised
 * S^{n+1}, k_s = dt f(stage_s); spin unchanged. */
int asd_runge_kutta(const asd_params *p, int L, int nrep, const double *spin, double *cur_input,
                    double *k1, double *k2, double *k3, double *k4);
/* replaces `void fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda,
 * double *b, int *ldb, double *c, int *ldc, double *fuse_C, double *sum)`
 * (gemm_kernel.h:355), same arguments, returning a status: c (column-major, the
 * blocked 4-real/4-imaginary layout, ldc >= 2*ceil4(m)) += (alpha a b) o fuse_C and
 * sum[0], sum[1] += sum of the updated c.  Correct for every K (the reference
 * re-sums per 240-wide K block), every m (the reference drops m % 4 rows) and
 * every n (it drops the last column of odd n).  sum: the reference keeps one
 * pair per OpenMP thread and the caller adds them up; here the whole total is
 * added to sum[0..1] and sum[2..] are untouched, so the caller's loop over
 * pairs gives the same result. */
int asd_fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda, double *b,
                    int *ldb, double *c, int *ldc, double *fuse_C, double *sum);

#ifdef __cplusplus
}
#endif
#endif /* ASD_H */
