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
};
LLGParams make_llg(const asd_params &p);

// one trajectory of a batch: state ping-pong a <-> b, snapshots [nsnap][L*L][3]
struct BatchTraj {
    double *a, *b, *snaps;
    LLGParams p;
    int L;
};
struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    int chunk = 0, nitems = 0;
    const void *trajs = nullptr, *items = nullptr;
    uint64_t key = 0;
};

struct RK4Graph {
    hipGraphExec_t exec = nullptr;
    // key
    int L = 0, nrep = 0, every = 0, chunk = 0, tile = 0, fuse = 1;
    const void *a = nullptr, *b = nullptr, *snaps = nullptr;
    LLGParams p{};
    long snap_stride = 0;
};

}  // namespace asd

struct asd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    // RK4 ping-pong state and the device-side snapshot counter
    asd::DevBuf state_b, counter;
    asd::RK4Graph graph;
    // S(q,t) workspace
    asd::DevBuf opA, opB, opA2, opB2, bins, fq, twiddle, qdev, outdev, stats;
    // compat (host-pointer) staging
    asd::DevBuf c0, c1, c2, c3, c4, c5, c6;
    // batching: work list, trajectory table, second state buffer
    asd::DevBuf work, tabs, batch_b;
    asd::BatchGraph bgraph;
    // autotuner choice per precision (-1 = automatic); tuned per (L, K)
    int forced_cfg[4] = {-1, -1, -1, -1};
    std::map<std::pair<int, int>, int> tuned[4];
    // RK4 tile edge: 0 = automatic; 8, 16 or 32 (LDS tiles) or 64 (register streaming)
    int rk4_tile = 0;
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
int rk4_batch_launch(asd_ctx *ctx, const BatchTraj *d_trajs, const int2 *d_items, int nitems, int parity,
                     const int *snap_counter, int write_snap, int bump_counter);
// one launch of `steps` RK4 steps (1, or 2 where rk4_fusable_steps allows it); the
// snapshot, if any, is of the input state
int rk4_launch_step(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *in,
                    double *out, double *snap_out, const int *snap_counter, long snap_stride,
                    int bump_counter, int steps = 1);
// steps one rk4_launch_step may fuse for this lattice: 2 where the automatic choice is the
// small-lattice 8 x 8 tiles (one step is latency-bound there), else 1
int rk4_fusable_steps(const asd_ctx *ctx, int L, int nrep);
int rk4_compat_stages(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *spin,
                      double *cur, double *k1, double *k2, double *k3, double *k4);
int rk4_stages_store(asd_ctx *ctx, const LLGParams &p, int L, const double *S, double *Y2, double *Y3, double *Y4,
                     double *c, double *k1, double *k2, double *k3, double *k4);
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
}  // namespace asd
