// asd_internal.h -- shared host/device definitions of libasd (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "asd.h"

namespace asd {

// ----------------------------------------------------------------- errors
int set_error(int code, const char *fmt, ...);
#define ASD_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::asd::set_error(ASD_EHIP, "%s failed: %s (%s:%d)", #call,          \
                                    hipGetErrorString(e_), __FILE__, __LINE__);        \
    } while (0)
#define ASD_TRY(expr)                                                                  \
    do {                                                                               \
        int r_ = (expr);                                                               \
        if (r_ != ASD_OK) return r_;                                                   \
    } while (0)
#define ASD_CHECK(cond, ...)                                                           \
    do {                                                                               \
        if (!(cond)) return ::asd::set_error(ASD_EINVAL, __VA_ARGS__);                 \
    } while (0)

// a device buffer that grows on demand and is reused across calls; freed with its owner
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t want);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// LLG parameters as the kernels take them (by value)
struct LLGParams {
    double J, D, Bz, alpha, dt, len;
    double a_over_len;  // alpha / spin_length
    double den;         // 1 + alpha^2
    double inv_den;     // 1 / (1 + alpha^2), formed once
    // the divisors of the k/6 and k/3 terms as kernel arguments: held in SGPRs they can be
    // operands of a VOP3 FMA, where a literal cannot (gfx9 encodings)
    double six, three, three_halves;
    double half_dt;     // dt / 2 (exact)
    // the tolerance mode's constants: 2 D, and the RK4 weights dt/6 and dt/3 of the increments
    double two_D, w6, w3;
};
LLGParams make_llg(const asd_params &p);

// Numeric guard of the RK4 kernels.  A site whose unnormalised state c has |c|^2 outside
// [2^-900, 2^900] (NaN and infinities included) is outside the range where the step is
// defined (the renormalisation's reciprocal square root needs it; after a valid step
// |S| = spin_length to rounding).  The kernels then record (step << 20 | trajectory) by
// atomicMin into *first, so the host learns the first such step.  The step of a launch is
// `step`, plus `every` times the device counter *ctr when ctr is set (graph replay, where
// the launch's arguments are fixed at capture).  first == NULL: no check.
struct Guard {
    unsigned long long *first = nullptr;
    const int *ctr = nullptr;
    long step = 0;
    int every = 0;
};
constexpr unsigned long long kGuardNone = ~0ull;

// one trajectory of a batch: state ping-pong a <-> b, snapshots [nsnap][L*L][3]
struct BatchTraj {
    double *a, *b, *snaps;
    LLGParams p;
    int L;
};
struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    int chunk = 0, nitems = 0, variant = -1;
    const void *trajs = nullptr, *items = nullptr, *guard = nullptr;
    uint64_t key = 0;
};

struct RK4Graph {
    hipGraphExec_t exec = nullptr;
    // key
    int L = 0, nrep = 0, every = 0, chunk = 0, tile = 0, fuse = 1, exact = 1;
    const void *a = nullptr, *b = nullptr, *snaps = nullptr, *guard = nullptr;
    LLGParams p{};
    long snap_stride = 0;
};

}  // namespace asd

struct asd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    // RK4 ping-pong state, the device-side snapshot counter and the numeric guard word
    asd::DevBuf state_b, counter, guard;
    asd::RK4Graph graph;
    // S(q,t) workspace
    asd::DevBuf opA, opB, opA2, opB2, bins, fq, twiddle, qdev, outdev, fix;
    // what twiddle and qdev hold (lattice side, q list, buffers), so repeated calls skip the upload
    int phases_L = 0;
    std::vector<int> phases_q;
    const void *phases_tw = nullptr, *phases_qd = nullptr;
    // compat (host-pointer) staging
    asd::DevBuf c0, c1, c2, c3, c4, c5, c6;
    // batching: work list, trajectory table, second state buffer
    asd::DevBuf work, tabs, batch_b;
    asd::BatchGraph bgraph;
    // batch kernel variant: forced (>= 0) or -1 (autotuned per batch composition, the sorted Ls)
    int batch_variant = -1;
    std::map<std::vector<int>, int> batch_tuned;
    // autotuner choice per precision (-1 = automatic); tuned per (L, K)
    int forced_cfg[5] = {-1, -1, -1, -1, -1};
    std::map<std::pair<int, int>, int> tuned[5];
    // RK4 tile edge: 0 = automatic; 8, 16 or 32 (LDS tiles) or 64 (register streaming)
    int rk4_tile = 0;
    // RK4 arithmetic: 1 = the reference's operation order (IEEE-equal values), 0 = the tolerance mode
    int rk4_exact = 1;
    // 1: S(q,tau) always takes the per-q path (the adjoint needs its F_q(t) sums)
    int sqt_generic_only = 0;
    // stage timing (asd_set_timing): events around pack, gemm, fourier, finish, rk4
    int timing = 0;
    hipEvent_t ev[8] = {};
    // adjoint workspace
    asd::DevBuf adj0, adj1, adj2, adj3, adj4, adj5;
    // unfused / brute-force S(q,tau) paths (sqt_ref.hip): rocBLAS handle and buffers
    void *blas = nullptr;
    asd::DevBuf ref0, ref1, ref2, ref3, ref4, ref5;
};

namespace asd {
// record stage event k on the context stream when timing is on
int mark(asd_ctx *ctx, int k);
// internal entry points shared between translation units
// one launch of batch kernel variant `variant` doing `steps` (1, or 2 for a two-step variant);
// it writes its input states as snapshot slot *snap_counter * snap_mul + snap_add (the counter
// may be NULL: slot snap_add), none where snap_add < 0, and bumps *snap_counter afterwards when
// bump_counter is set
int rk4_batch_launch(asd_ctx *ctx, int variant, int steps, const BatchTraj *d_trajs, const int2 *d_items,
                     int nitems, int parity, const int *snap_counter, int snap_mul, int snap_add,
                     int bump_counter, const Guard &guard = Guard());
int rk4_batch_num_variants();
const char *rk4_batch_variant_name(int v);
int rk4_batch_variant_tile(int v);
int rk4_batch_variant_steps(int v);
// one launch of `steps` RK4 steps (1, or 2 where rk4_fusable_steps allows it); the
// snapshot, if any, is of the input state
int rk4_launch_step(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *in,
                    double *out, double *snap_out, const int *snap_counter, long snap_stride,
                    int bump_counter, int steps = 1, const Guard &guard = Guard());
// steps one rk4_launch_step may fuse for this lattice: 2 where the automatic choice is an
// LDS tile path whose single step is latency-bound (8 x 8, 32 x 32 and 16 x 16 tiles with
// 1024 threads), else 1
int rk4_fusable_steps(const asd_ctx *ctx, int L, int nrep);
// tiny kernels shared by the drivers
__global__ void set_u64_kernel(unsigned long long *p, unsigned long long v);
__global__ void bump_counter_kernel(int *c);
int rk4_compat_stages(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *spin,
                      double *cur, double *k1, double *k2, double *k3, double *k4);
// one exact RK4 step's stage inputs Y2 = S + k1/2, Y3 = S + k2/2, Y4 = S + k3 and unnormalised sum
// c of a lattice, in one launch (the renormalised state goes to `scratch`, [L*L][3])
int rk4_stages_store(asd_ctx *ctx, const LLGParams &p, int L, const double *S, double *Y2, double *Y3, double *Y4,
                     double *c, double *scratch);
int rhs_launch(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *S, double *dS);
int sqt_num_configs(asd_precision prec);
const char *sqt_config_name(asd_precision prec, int cfg);
int sqt_autotune(asd_ctx *ctx, int L, int T0, asd_precision prec, int *best, double *times_ms);
int sqt_unfused_run(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                    int nq, double *d_out);
int sqt_bruteforce_run(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                       int nq, double *d_out);
void blas_destroy(asd_ctx *ctx);
int sqt_run(asd_ctx *ctx, int L, const double *d_snaps_a, const double *d_snaps_b, long ta_stride,
            int nsnap_a, int T0, int nlags, const int *h_q, int nq, asd_precision prec,
            double *d_out, int cfg);
// the Fourier sums F_q(t) of nt snapshots into ctx->fq ([t][q][xyz] complex, the layout sqt_run
// leaves there on its generic path), with ctx->twiddle and ctx->qdev set up as sqt_run does
int sqt_fourier(asd_ctx *ctx, int L, const double *d_snaps, long tstride, int nt, const int *h_q, int nq);
}  // namespace asd
