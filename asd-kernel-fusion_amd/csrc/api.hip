// api.hip -- extern "C" entry points of libasd (declared in include/asd.h and
// include/asd_compat.h): context, RK4 driver, S(q,tau) driver, reference-compatible
// host-pointer entries.
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>

#include "asd_compat.h"
#include "asd_internal.h"

namespace asd {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int DevBuf::ensure(size_t want)
{
    if (want <= bytes) return ASD_OK;
    release();
    if (want == 0) return ASD_OK;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(ASD_ENOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    ptr = p;
    bytes = want;
    return ASD_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

static int check_params(const asd_params *p)
{
    ASD_CHECK(p, "params is NULL");
    ASD_CHECK(std::isfinite(p->J) && std::isfinite(p->D) && std::isfinite(p->Bz) &&
                  std::isfinite(p->alpha) && std::isfinite(p->dt) && std::isfinite(p->spin_length),
              "non-finite parameter");
    // the RK4 kernels' renormalisation assumes |S|^2 within 2^(+-900) of 1
    ASD_CHECK(p->spin_length >= 1e-100 && p->spin_length <= 1e100, "spin_length must be in [1e-100, 1e100] (got %g)",
              p->spin_length);
    ASD_CHECK(p->alpha >= 0, "alpha must be >= 0 (got %g)", p->alpha);
    return ASD_OK;
}

static int check_L(int L)
{
    ASD_CHECK(L >= 2 && L <= 1024, "L must be in [2, 1024] (got %d)", L);
    return ASD_OK;
}

// same hash as the counter-based stream documented in asd.h
static inline uint64_t hash64(uint64_t seed, uint64_t a, uint64_t b)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + a * 0xD1B54A32D192ED03ull + b * 0x8CB92BA72F3D8DD7ull +
                 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void set_int_kernel(int *p, int v) { *p = v; }

// longest step loop captured as one graph (longer snapshot intervals run eagerly)
constexpr long kMaxGraphSteps = 4096;
// snapshot intervals per captured RK4 chunk (up to; see rk4_drive)
constexpr int kMaxSnapsPerChunk = 16;
constexpr long kMaxChunkSteps = 512;

// ---------------------------------------------- scratch of asd_rk4_run_noalloc
// The allocation-free RK4 entry point keeps its second state buffer (the ping-pong
// partner of the caller's spins) in this file-scope array: 16 MiB per GPU, held for the
// life of the process.  Host code never uses the array's name as an address: asd_create
// looks up each device's copy once with hipGetSymbolAddress (that device current) and
// the entry point passes that pointer to the kernels.  Calls that overlap on one GPU share
// the array and race.
constexpr long kScratchDoubles = 1L << 21;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_lookup(int device)
{
    ASD_CHECK(device >= 0 && device < kMaxDevices, "device %d beyond the %d this build tracks", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return ASD_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(ASD_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                         hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(ptr);
    return ASD_OK;
}

// the guard word's verdict: ASD_OK, or ASD_ENOTFINITE naming the first step
static int guard_verdict(unsigned long long h, bool batch)
{
    if (h == kGuardNone) return ASD_OK;
    const long step = (long)(h >> 20);
    const int who = (int)(h & 0xFFFFF);
    return set_error(ASD_ENOTFINITE,
                     "RK4 state left the range of the step (|c|^2 NaN, infinite or outside [2^-900, 2^900]) "
                     "in %s %d, first in the launch starting at step %ld",
                     batch ? "trajectory" : "replica", who, step);
}

// the guard of a launch at offset k of a captured chunk of `chunk` steps after which one
// node bumps *ctr (the counter counts the chunks run): its step is chunk * *ctr + k
static Guard chunk_guard(unsigned long long *first, const int *ctr, int k, int chunk)
{
    Guard g;
    g.first = first;
    g.ctr = ctr;
    g.every = chunk;
    g.step = k;
    return g;
}
static Guard eager_guard(unsigned long long *first, long step)
{
    Guard g;
    g.first = first;
    g.step = step;
    return g;
}

int mark(asd_ctx *c, int k)
{
    if (!c->timing) return ASD_OK;
    ASD_HIP(hipEventRecord(c->ev[k], c->stream));
    return ASD_OK;
}

}  // namespace asd

using namespace asd;

extern "C" {

const char *asd_last_error(void) { return g_err; }
const char *asd_version(void) { return "asd-mi355x 0.1 (gfx950)"; }

int asd_init_spins(uint64_t seed, int nsites, double *out)
{
    ASD_CHECK(out && nsites >= 0, "bad arguments");
    for (int i = 0; i < nsites; i++) {
        double u, v, s;
        uint64_t ctr = 0;
        for (;;) {
            u = 2.0 * ((double)(hash64(seed, (uint64_t)i, ctr) >> 11) * 0x1.0p-53) - 1.0;
            v = 2.0 * ((double)(hash64(seed, (uint64_t)i, ctr + 1) >> 11) * 0x1.0p-53) - 1.0;
            ctr += 2;
            s = u * u + v * v;
            if (s < 1.0 && s > 0.0) break;
        }
        const double r = 2.0 * std::sqrt(1.0 - s);
        out[3L * i] = u * r;
        out[3L * i + 1] = v * r;
        out[3L * i + 2] = 1.0 - 2.0 * s;
    }
    return ASD_OK;
}

int asd_qpath_gxmg(int L, int *q)
{
    ASD_CHECK(L >= 2 && L % 2 == 0 && q, "L must be even");
    const int h = L / 2;
    int n = 0;
    for (int k = 0; k < h; k++) { q[2 * n] = k; q[2 * n + 1] = 0; n++; }
    for (int k = 0; k < h; k++) { q[2 * n] = h; q[2 * n + 1] = k; n++; }
    for (int k = 0; k < h; k++) { q[2 * n] = h - k; q[2 * n + 1] = h - k; n++; }
    return n;
}

int asd_create(asd_ctx **out, int device)
{
    ASD_CHECK(out, "ctx is NULL");
    int n = 0;
    ASD_HIP(hipGetDeviceCount(&n));
    ASD_CHECK(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    ASD_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    ASD_HIP(hipGetDeviceProperties(&prop, device));
    ASD_TRY(scratch_lookup(device));
    asd_ctx *c = new asd_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    *out = c;
    return ASD_OK;
}

int asd_destroy(asd_ctx *c)
{
    if (!c) return ASD_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    if (c->bgraph.exec) (void)hipGraphExecDestroy(c->bgraph.exec);
    for (hipEvent_t &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    blas_destroy(c);
    delete c;  // every DevBuf member frees its memory
    return ASD_OK;
}

int asd_set_stream(asd_ctx *c, void *stream)
{
    ASD_CHECK(c, "ctx is NULL");
    c->stream = static_cast<hipStream_t>(stream);
    if (c->graph.exec) {
        (void)hipGraphExecDestroy(c->graph.exec);
        c->graph = RK4Graph();
    }
    if (c->bgraph.exec) {
        (void)hipGraphExecDestroy(c->bgraph.exec);
        c->bgraph = BatchGraph();
    }
    return ASD_OK;
}

int asd_set_timing(asd_ctx *c, int on)
{
    ASD_CHECK(c, "ctx is NULL");
    if (on)
        for (hipEvent_t &e : c->ev)
            if (!e) ASD_HIP(hipEventCreate(&e));
    c->timing = on ? 1 : 0;
    return ASD_OK;
}

int asd_stage_times(asd_ctx *c, double t[5])
{
    ASD_CHECK(c && t, "NULL argument");
    ASD_CHECK(c->timing, "timing is off (asd_set_timing)");
    const int pairs[5][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {5, 6}};
    for (int k = 0; k < 5; k++) {
        float ms = 0.f;
        ASD_HIP(hipEventSynchronize(c->ev[pairs[k][1]]));
        hipError_t e = hipEventElapsedTime(&ms, c->ev[pairs[k][0]], c->ev[pairs[k][1]]);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ms = -1.f;  // stage not run since timing was enabled
        }
        t[k] = ms;
    }
    return ASD_OK;
}

int asd_malloc(asd_ctx *c, size_t bytes, void **dptr)
{
    ASD_CHECK(c && dptr, "NULL argument");
    ASD_HIP(hipSetDevice(c->device));
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(ASD_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return ASD_OK;
}

int asd_free(asd_ctx *c, void *dptr)
{
    ASD_CHECK(c, "ctx is NULL");
    ASD_HIP(hipStreamSynchronize(c->stream));
    if (dptr) ASD_HIP(hipFree(dptr));
    return ASD_OK;
}

int asd_memcpy(asd_ctx *c, void *dst, const void *src, size_t bytes, int kind)
{
    ASD_CHECK(c && dst && src, "NULL argument");
    ASD_CHECK(kind >= 0 && kind <= 2, "kind must be 0 (H2D), 1 (D2H) or 2 (D2D)");
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    ASD_HIP(hipMemcpyAsync(dst, src, bytes, k[kind], c->stream));
    ASD_HIP(hipStreamSynchronize(c->stream));
    return ASD_OK;
}

int asd_rk4_set_tile(asd_ctx *c, int tile)
{
    ASD_CHECK(c && (tile == 0 || tile == 8 || tile == 16 || tile == 32 || tile == 64),
              "tile must be 0 (automatic), 8, 16, 32 or 64");
    c->rk4_tile = tile;
    if (c->graph.exec) {  // captured launches use the previous tile
        (void)hipGraphExecDestroy(c->graph.exec);
        c->graph = RK4Graph();
    }
    return ASD_OK;
}

int asd_rk4_set_exact(asd_ctx *c, int exact)
{
    ASD_CHECK(c && (exact == 0 || exact == 1), "exact must be 0 or 1");
    c->rk4_exact = exact;
    return ASD_OK;  // graphs are keyed on the setting
}

int asd_synchronize(asd_ctx *c)
{
    ASD_CHECK(c, "ctx is NULL");
    ASD_HIP(hipStreamSynchronize(c->stream));
    return ASD_OK;
}

// --------------------------------------------------------------------- RK4

// The RK4 step loop shared by asd_rk4_run and asd_rk4_run_noalloc: `other` is the second
// state buffer, `counter` the snapshot counter (graph mode only), `first` the guard word
// (NULL: no check).  Everything is stream-ordered; nothing is allocated here.
static int rk4_drive(asd_ctx *c, const LLGParams &p, int L, int nrep, double *d_spins, double *other,
                     int *counter, unsigned long long *first, long nsteps, int snap_every, int nsnap,
                     double *d_snaps, bool use_graph)
{
    const long n = (long)nrep * L * L * 3;
    const long snap_stride = n;
    double *bufs[2] = {d_spins, other};
    hipStream_t st = c->stream;

    // snapshots due while stepping: those at s = k * every < nsteps; the one at
    // s == nsteps (if any) is the final state, copied at the end
    const long nsnap_in_loop = d_snaps ? std::min<long>(nsnap, (nsteps + snap_every - 1) / snap_every) : 0;
    long s = 0;
    int cur = 0;

    // launches of two fused steps where the kernel choice allows it and no snapshot falls
    // between the two steps of a launch (snapshots are of a launch's input state)
    const int fs = rk4_fusable_steps(c, L, nrep) == 2 && (!d_snaps || snap_every % 2 == 0) ? 2 : 1;

    // graph: one chunk = spc snapshot intervals, an even number of launches (the ping-pong
    // parity repeats).  The launch at offset k of a chunk with k % every == 0 writes snapshot
    // slot spc * *counter + k / every (its snapshot pointer is offset by k / every slots and
    // its slot stride is spc slots) and one node at the end of the chunk bumps the counter, so
    // the counter counts chunks.  spc is the largest count up to kMaxSnapsPerChunk (and
    // kMaxChunkSteps steps) that divides the snapshots due in the loop: one bump per chunk
    // instead of one per snapshot (a bump node costs about a step launch: 3 us at L = 128,
    // where a step is 2.9 us).
    auto chunk_of = [&](long m) {
        long ch = m * snap_every;
        if ((ch / fs) % 2) ch *= 2;
        return ch;
    };
    if (use_graph && d_snaps && snap_every >= 1 && nsnap_in_loop >= 2 && counter && chunk_of(1) <= kMaxGraphSteps) {
        long chunk = chunk_of(1);
        for (long m = std::min<long>(kMaxSnapsPerChunk, nsnap_in_loop); m > 1; m--) {
            const long ch = chunk_of(m);
            if (ch <= kMaxChunkSteps && nsnap_in_loop % (ch / snap_every) == 0 && nsteps >= ch) {
                chunk = ch;
                break;
            }
        }
        const long spc = chunk / snap_every;  // snapshot intervals per chunk
        const long nchunks = std::min<long>(nsteps / chunk, nsnap_in_loop / spc);
        if (nchunks >= 1) {
            RK4Graph &g = c->graph;
            const bool same = g.exec && g.L == L && g.nrep == nrep && g.every == snap_every && g.chunk == chunk &&
                              g.tile == c->rk4_tile && g.fuse == fs && g.a == bufs[0] && g.b == bufs[1] &&
                              g.snaps == d_snaps && std::memcmp(&g.p, &p, sizeof p) == 0 &&
                              g.snap_stride == snap_stride && g.guard == (const void *)first &&
                              g.exact == c->rk4_exact;
            if (!same) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                g = RK4Graph();
                hipStream_t cap;
                ASD_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                hipGraph_t graph;
                ASD_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                hipStream_t saved = c->stream;
                c->stream = cap;
                int r = ASD_OK;
                for (int k = 0; k < chunk && r == ASD_OK; k += fs) {
                    const bool snap_here = k % snap_every == 0;
                    const int j = k / fs;  // launch index: its input buffer alternates
                    const Guard gd = first ? chunk_guard(first, counter, k, (int)chunk) : Guard();
                    r = rk4_launch_step(c, p, L, nrep, bufs[j & 1], bufs[(j + 1) & 1],
                                        snap_here ? d_snaps + (k / snap_every) * snap_stride : nullptr, counter,
                                        spc * snap_stride, k + fs >= chunk ? 1 : 0, fs, gd);
                }
                c->stream = saved;
                hipError_t e1 = hipStreamEndCapture(cap, &graph);
                (void)hipStreamDestroy(cap);
                ASD_TRY(r);
                ASD_HIP(e1);
                hipError_t e2 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                ASD_HIP(e2);
                g.L = L; g.nrep = nrep; g.every = snap_every; g.chunk = (int)chunk; g.tile = c->rk4_tile; g.fuse = fs;
                g.a = bufs[0]; g.b = bufs[1]; g.snaps = d_snaps; g.p = p; g.snap_stride = snap_stride;
                g.guard = first;
                g.exact = c->rk4_exact;
            }
            hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, counter, 0);
            for (long ch = 0; ch < nchunks; ch++) ASD_HIP(hipGraphLaunch(g.exec, st));
            s = nchunks * chunk;
            cur = 0;  // an even number of launches per chunk: state back in bufs[0]
            // continue eagerly with explicit slots
        }
    }
    while (s < nsteps) {
        double *snap = nullptr;
        if (d_snaps && s % snap_every == 0 && s / snap_every < nsnap) snap = d_snaps + (s / snap_every) * snap_stride;
        const bool mid_snap = d_snaps && (s + 1) % snap_every == 0 && (s + 1) / snap_every < nsnap;
        const int k = fs == 2 && s + 2 <= nsteps && !mid_snap ? 2 : 1;
        ASD_TRY(rk4_launch_step(c, p, L, nrep, bufs[cur], bufs[cur ^ 1], snap, nullptr, snap_stride, 0, k,
                                first ? eager_guard(first, s) : Guard()));
        cur ^= 1;
        s += k;
    }
    if (d_snaps && nsteps % snap_every == 0 && nsteps / snap_every < nsnap)
        ASD_HIP(hipMemcpyAsync(d_snaps + (nsteps / snap_every) * snap_stride, bufs[cur], sizeof(double) * n,
                               hipMemcpyDeviceToDevice, st));
    if (cur == 1)
        ASD_HIP(hipMemcpyAsync(d_spins, bufs[1], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return ASD_OK;
}

static int check_rk4_args(asd_ctx *c, const asd_params *pp, int L, int nrep, double *d_spins, long nsteps,
                          int snap_every, int nsnap, double *d_snaps)
{
    ASD_CHECK(c && d_spins, "ctx/spins NULL");
    ASD_TRY(check_params(pp));
    ASD_TRY(check_L(L));
    ASD_CHECK(nrep >= 1 && nrep <= 65535, "nrep out of range");
    ASD_CHECK(nsteps >= 0, "nsteps < 0");
    ASD_CHECK(!d_snaps || (snap_every >= 1 && nsnap >= 0), "snap_every must be >= 1");
    return ASD_OK;
}

int asd_rk4_run(asd_ctx *c, const asd_params *pp, int L, int nrep, double *d_spins, long nsteps,
                int snap_every, int nsnap, double *d_snaps, int use_graph)
{
    ASD_TRY(check_rk4_args(c, pp, L, nrep, d_spins, nsteps, snap_every, nsnap, d_snaps));
    ASD_HIP(hipSetDevice(c->device));
    const LLGParams p = make_llg(*pp);
    const long n = (long)nrep * L * L * 3;
    ASD_TRY(c->state_b.ensure(sizeof(double) * n));
    ASD_TRY(c->counter.ensure(sizeof(int)));
    ASD_TRY(c->guard.ensure(sizeof(unsigned long long)));
    unsigned long long *first = c->guard.as<unsigned long long>();
    hipStream_t st = c->stream;
    ASD_TRY(mark(c, 5));
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, first, kGuardNone);
    ASD_TRY(rk4_drive(c, p, L, nrep, d_spins, c->state_b.as<double>(), c->counter.as<int>(), first, nsteps,
                      snap_every, nsnap, d_snaps, use_graph != 0));
    ASD_TRY(mark(c, 6));
    unsigned long long h = kGuardNone;
    ASD_HIP(hipMemcpyAsync(&h, first, sizeof h, hipMemcpyDeviceToHost, st));
    ASD_HIP(hipStreamSynchronize(st));
    return guard_verdict(h, false);
}

int asd_rk4_run_noalloc(asd_ctx *c, const asd_params *pp, int L, int nrep, double *d_spins, long nsteps,
                        int snap_every, int nsnap, double *d_snaps)
{
    ASD_TRY(check_rk4_args(c, pp, L, nrep, d_spins, nsteps, snap_every, nsnap, d_snaps));
    const long n = (long)nrep * L * L * 3;
    ASD_CHECK(n <= kScratchDoubles,
              "asd_rk4_run_noalloc: nrep * L * L * 3 = %ld doubles exceeds its fixed scratch of %ld doubles "
              "(16 MiB); use asd_rk4_run",
              n, kScratchDoubles);
    ASD_HIP(hipSetDevice(c->device));
    ASD_CHECK(c->device < kMaxDevices && g_scratch_ptr[c->device], "scratch of device %d not set up", c->device);
    return rk4_drive(c, make_llg(*pp), L, nrep, d_spins, g_scratch_ptr[c->device], nullptr, nullptr, nsteps,
                     snap_every, nsnap, d_snaps, false);
}

// ---------------------------------------------------------------- S(q,tau)

static int check_q(int L, const int *h_q, int nq)
{
    ASD_CHECK(h_q && nq >= 1 && nq <= 1 << 20, "bad q list");
    for (int i = 0; i < 2 * nq; i++) ASD_CHECK(h_q[i] >= 0 && h_q[i] < L, "q index out of [0, L)");
    return ASD_OK;
}

int asd_sqt(asd_ctx *c, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q, int nq,
            asd_precision prec, double *out, int out_on_device)
{
    ASD_CHECK(c && d_snaps && out, "NULL argument");
    ASD_TRY(check_L(L));
    ASD_CHECK(T0 >= 1 && nlags >= 1, "T0 and nlags must be >= 1");
    ASD_CHECK(prec >= ASD_FP64 && prec <= ASD_FP16F8, "unknown precision");
    ASD_TRY(check_q(L, h_q, nq));
    ASD_HIP(hipSetDevice(c->device));
    double *dout = out;
    if (!out_on_device) {
        ASD_TRY(c->outdev.ensure(sizeof(double) * 2 * (long)nlags * nq));
        dout = c->outdev.as<double>();
    }
    const int cfg = c->forced_cfg[prec];
    ASD_TRY(sqt_run(c, L, d_snaps, d_snaps, 3L * L * L, nsnap, T0, nlags, h_q, nq, prec, dout, cfg));
    if (!out_on_device) {
        ASD_HIP(hipMemcpyAsync(out, dout, sizeof(double) * 2 * (long)nlags * nq, hipMemcpyDeviceToHost, c->stream));
        ASD_HIP(hipStreamSynchronize(c->stream));
    }
    return ASD_OK;
}

// the unfused (library GEMM + explicit correlation matrix) and brute-force paths share
// asd_sqt's arguments, checks and output handling
static int sqt_other(asd_ctx *c, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q, int nq,
                     double *out, int out_on_device, bool brute)
{
    ASD_CHECK(c && d_snaps && out, "NULL argument");
    ASD_TRY(check_L(L));
    ASD_CHECK(T0 >= 1 && nlags >= 1, "T0 and nlags must be >= 1");
    ASD_TRY(check_q(L, h_q, nq));
    ASD_HIP(hipSetDevice(c->device));
    double *dout = out;
    if (!out_on_device) {
        ASD_TRY(c->outdev.ensure(sizeof(double) * 2 * (long)nlags * nq));
        dout = c->outdev.as<double>();
    }
    if (brute)
        ASD_TRY(sqt_bruteforce_run(c, L, d_snaps, nsnap, T0, nlags, h_q, nq, dout));
    else
        ASD_TRY(sqt_unfused_run(c, L, d_snaps, nsnap, T0, nlags, h_q, nq, dout));
    if (!out_on_device) {
        ASD_HIP(hipMemcpyAsync(out, dout, sizeof(double) * 2 * (long)nlags * nq, hipMemcpyDeviceToHost, c->stream));
        ASD_HIP(hipStreamSynchronize(c->stream));
    }
    return ASD_OK;
}

int asd_sqt_unfused(asd_ctx *c, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q, int nq,
                    double *out, int out_on_device)
{
    return sqt_other(c, L, d_snaps, nsnap, T0, nlags, h_q, nq, out, out_on_device, false);
}

int asd_sqt_bruteforce(asd_ctx *c, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                       int nq, double *out, int out_on_device)
{
    return sqt_other(c, L, d_snaps, nsnap, T0, nlags, h_q, nq, out, out_on_device, true);
}

int asd_sqt_replicas(asd_ctx *c, int L, int nrep, const double *d_now, const double *d_zero, const int *h_q,
                     int nq, asd_precision prec, double *out)
{
    ASD_CHECK(c && d_now && d_zero && out, "NULL argument");
    ASD_TRY(check_L(L));
    ASD_CHECK(nrep >= 1, "nrep must be >= 1");
    ASD_CHECK(prec >= ASD_FP64 && prec <= ASD_FP16F8, "unknown precision");
    ASD_TRY(check_q(L, h_q, nq));
    ASD_HIP(hipSetDevice(c->device));
    ASD_TRY(c->outdev.ensure(sizeof(double) * 2 * (long)nq));
    ASD_TRY(sqt_run(c, L, d_now, d_zero, 3L * L * L, nrep, nrep, 1, h_q, nq, prec, c->outdev.as<double>(),
                    c->forced_cfg[prec]));
    ASD_HIP(hipMemcpyAsync(out, c->outdev.ptr, sizeof(double) * 2 * nq, hipMemcpyDeviceToHost, c->stream));
    ASD_HIP(hipStreamSynchronize(c->stream));
    return ASD_OK;
}

int asd_sqt_num_configs(asd_precision prec)
{
    return (prec >= ASD_FP64 && prec <= ASD_FP16F8) ? sqt_num_configs(prec) : 0;
}
const char *asd_sqt_config_name(asd_precision prec, int cfg)
{
    return (prec >= ASD_FP64 && prec <= ASD_FP16F8) ? sqt_config_name(prec, cfg) : "invalid";
}
int asd_sqt_set_config(asd_ctx *c, asd_precision prec, int cfg)
{
    ASD_CHECK(c && prec >= ASD_FP64 && prec <= ASD_FP16F8, "bad arguments");
    ASD_CHECK(cfg >= -1 && cfg < asd_sqt_num_configs(prec), "config out of range");
    c->forced_cfg[prec] = cfg;
    return ASD_OK;
}
int asd_autotune_sqt(asd_ctx *c, int L, int T0, asd_precision prec, int *best, double *times_ms)
{
    ASD_CHECK(c, "ctx NULL");
    ASD_TRY(check_L(L));
    ASD_CHECK(T0 >= 1 && prec >= ASD_FP64 && prec <= ASD_FP16F8, "bad arguments");
    ASD_HIP(hipSetDevice(c->device));
    return sqt_autotune(c, L, T0, prec, best, times_ms);
}

// work list of (trajectory, tile) items for tiles of edge T, interleaved tile by tile so
// that every stretch of the list mixes lattice sizes (balanced persistent workgroups)
static std::vector<int2> batch_items(int ntraj, const int *Ls, int T)
{
    std::vector<int2> items;
    for (int t = 0; t < ntraj; t++) {
        const int nt = (Ls[t] + T - 1) / T;
        for (int k = 0; k < nt * nt; k++) items.push_back(make_int2(t, k));
    }
    std::stable_sort(items.begin(), items.end(), [](const int2 &x, const int2 &y) { return x.y < y.y; });
    return items;
}

// uploads the trajectory table and variant v's work list into the context's buffers
static int batch_tables(asd_ctx *c, const std::vector<BatchTraj> &tab, const int *Ls, int v, int *nitems)
{
    const int ntraj = (int)tab.size();
    const std::vector<int2> items = batch_items(ntraj, Ls, rk4_batch_variant_tile(v));
    *nitems = (int)items.size();
    ASD_TRY(c->tabs.ensure(sizeof(BatchTraj) * ntraj));
    ASD_TRY(c->work.ensure(sizeof(int2) * items.size()));
    ASD_HIP(hipMemcpyAsync(c->tabs.ptr, tab.data(), sizeof(BatchTraj) * ntraj, hipMemcpyHostToDevice, c->stream));
    ASD_HIP(hipMemcpyAsync(c->work.ptr, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice, c->stream));
    ASD_HIP(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
    return ASD_OK;
}

static int check_batch(int ntraj, const int *Ls, const asd_params *params, long *total)
{
    ASD_CHECK(ntraj >= 1 && ntraj <= (1 << 20) && Ls && params, "bad arguments");
    *total = 0;
    for (int t = 0; t < ntraj; t++) {
        ASD_TRY(check_L(Ls[t]));
        ASD_TRY(check_params(&params[t]));
        *total += 3L * Ls[t] * Ls[t];
    }
    return ASD_OK;
}

// a batch composition (the multiset of L) and the RK4 arithmetic mode it was timed in
static std::vector<int> batch_key(const asd_ctx *c, int ntraj, const int *Ls)
{
    std::vector<int> k(Ls, Ls + ntraj);
    std::sort(k.begin(), k.end());
    k.push_back(c->rk4_exact ? -1 : -2);
    return k;
}

// deterministic unit spins for the autotuner's timing runs (the values do not matter)
__global__ void fill_unit_spins_kernel(double *p, long nspins)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nspins; i += (long)gridDim.x * blockDim.x) {
        const double z = 2.0 * ((double)((i * 2654435761L) & 1023) + 0.5) / 1024.0 - 1.0, a = 0.61803398875 * (double)i;
        const double r = sqrt(1.0 - z * z);
        p[3 * i] = r * cos(a);
        p[3 * i + 1] = r * sin(a);
        p[3 * i + 2] = z;
    }
}

int asd_batch_num_variants(void) { return rk4_batch_num_variants(); }
const char *asd_batch_variant_name(int v) { return rk4_batch_variant_name(v); }

int asd_batch_set_variant(asd_ctx *c, int v)
{
    ASD_CHECK(c && v >= -1 && v < rk4_batch_num_variants(), "variant must be -1 (automatic) or in [0, %d)",
              rk4_batch_num_variants());
    c->batch_variant = v;
    return ASD_OK;
}

int asd_batch_variant_for(asd_ctx *c, int ntraj, const int *Ls)
{
    ASD_CHECK(c && ntraj >= 1 && Ls, "bad arguments");
    if (c->batch_variant >= 0) return c->batch_variant;
    auto it = c->batch_tuned.find(batch_key(c, ntraj, Ls));
    return it == c->batch_tuned.end() ? -1 : it->second;
}

int asd_autotune_batch(asd_ctx *c, int ntraj, const int *Ls, const asd_params *params, int *best, double *times_ms)
{
    ASD_CHECK(c, "ctx NULL");
    long total = 0;
    ASD_TRY(check_batch(ntraj, Ls, params, &total));
    ASD_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf a;
    ASD_TRY(a.ensure(sizeof(double) * total));
    ASD_TRY(c->batch_b.ensure(sizeof(double) * total));
    hipLaunchKernelGGL(fill_unit_spins_kernel, dim3(1024), dim3(256), 0, st, a.as<double>(), total / 3);
    std::vector<BatchTraj> tab(ntraj);
    for (long t = 0, off = 0; t < ntraj; off += 3L * Ls[t] * Ls[t], t++) {
        tab[t].a = a.as<double>() + off;
        tab[t].b = c->batch_b.as<double>() + off;
        tab[t].snaps = nullptr;
        tab[t].p = make_llg(params[t]);
        tab[t].L = Ls[t];
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ASD_HIP(hipEventCreate(&e0));
    hipError_t ee = hipEventCreate(&e1);
    if (ee != hipSuccess) {
        (void)hipEventDestroy(e0);
        ASD_HIP(ee);
    }
    // 20 steps per timing run (10 launches of the two-step variants), best of two runs
    constexpr int kSteps = 20;
    int rc = ASD_OK, bi = 0;
    double best_ms = 1e300;
    for (int v = 0; v < rk4_batch_num_variants() && rc == ASD_OK; v++) {
        int nitems = 0;
        rc = batch_tables(c, tab, Ls, v, &nitems);
        const int fs = rk4_batch_variant_steps(v);
        double ms = 1e300;
        for (int rep = 0; rep < 3 && rc == ASD_OK; rep++) {
            if (hipEventRecord(e0, st) != hipSuccess) rc = set_error(ASD_EHIP, "hipEventRecord failed");
            for (int k = 0; k < kSteps && rc == ASD_OK; k += fs)
                rc = rk4_batch_launch(c, v, fs, c->tabs.as<BatchTraj>(), c->work.as<int2>(), nitems, (k / fs) & 1,
                                      nullptr, 0, -1, 0);
            float t = 0.f;
            if (rc == ASD_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                rc = set_error(ASD_EHIP, "batch autotune timing failed");
            if (rep > 0) ms = std::min(ms, (double)t);  // the first run warms up
        }
        if (times_ms) times_ms[v] = ms;
        if (ms < 0.99 * best_ms) {  // near-ties keep the earlier variant
            best_ms = ms;
            bi = v;
        }
    }
    (void)hipStreamSynchronize(st);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != ASD_OK) return rc;
    c->batch_tuned[batch_key(c, ntraj, Ls)] = bi;
    if (best) *best = bi;
    return ASD_OK;
}

int asd_batch_rk4_run(asd_ctx *c, int ntraj, const int *Ls, const asd_params *params, double *const *d_spins,
                      long nsteps, int snap_every, int nsnap, double *const *d_snaps)
{
    ASD_CHECK(c && d_spins, "bad arguments");
    ASD_CHECK(nsteps >= 0, "nsteps < 0");
    ASD_CHECK(!d_snaps || (snap_every >= 1 && nsnap >= 0), "snap_every must be >= 1");
    long total = 0;
    ASD_TRY(check_batch(ntraj, Ls, params, &total));
    for (int t = 0; t < ntraj; t++) ASD_CHECK(d_spins[t], "d_spins[%d] is NULL", t);
    ASD_HIP(hipSetDevice(c->device));
    // the kernel variant: forced, recorded for this batch composition, or autotuned now
    int v = c->batch_variant;
    if (v < 0) {
        auto it = c->batch_tuned.find(batch_key(c, ntraj, Ls));
        if (it == c->batch_tuned.end()) {
            ASD_TRY(asd_autotune_batch(c, ntraj, Ls, params, &v, nullptr));
        } else {
            v = it->second;
        }
    }
    std::vector<BatchTraj> tab(ntraj);
    ASD_TRY(c->batch_b.ensure(sizeof(double) * total));
    ASD_TRY(c->counter.ensure(sizeof(int)));
    for (long t = 0, off = 0; t < ntraj; off += 3L * Ls[t] * Ls[t], t++) {
        tab[t].a = d_spins[t];
        tab[t].b = c->batch_b.as<double>() + off;
        tab[t].snaps = d_snaps ? d_snaps[t] : nullptr;
        tab[t].p = make_llg(params[t]);
        tab[t].L = Ls[t];
    }
    int nitems = 0;
    ASD_TRY(batch_tables(c, tab, Ls, v, &nitems));
    hipStream_t st = c->stream;
    const BatchTraj *dt = c->tabs.as<BatchTraj>();
    const int2 *dw = c->work.as<int2>();
    int *ctr = c->counter.as<int>();
    ASD_TRY(mark(c, 5));
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, ctr, 0);
    ASD_TRY(c->guard.ensure(sizeof(unsigned long long)));
    unsigned long long *first = c->guard.as<unsigned long long>();
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, first, kGuardNone);

    const int every = d_snaps ? snap_every : 10;
    // two fused steps per launch where the variant has them and no snapshot falls between them
    const int fs = rk4_batch_variant_steps(v) == 2 && every % 2 == 0 ? 2 : 1;
    const long nsnap_in_loop = d_snaps ? std::min<long>(nsnap, (nsteps + every - 1) / every) : 0;
    // a graph chunk is spc intervals of `every` steps, an even number of launches (the
    // ping-pong parity repeats), with one counter bump at its end (the counter counts
    // chunks); the launch at offset k with k % every == 0 writes snapshot slot
    // spc * *counter + k / every.  spc as in rk4_drive: the largest count up to
    // kMaxSnapsPerChunk (and kMaxChunkSteps steps) that divides the intervals to run.
    auto chunk_of = [&](long m) {
        long ch = m * every;
        if ((ch / fs) % 2) ch *= 2;
        return ch;
    };
    const long nint = d_snaps ? nsnap_in_loop : nsteps / every;
    long chunk = chunk_of(1);
    for (long m = std::min<long>(kMaxSnapsPerChunk, nint); m > 1; m--) {
        const long ch = chunk_of(m);
        if (ch <= kMaxChunkSteps && nint % (ch / every) == 0 && nsteps >= ch) {
            chunk = ch;
            break;
        }
    }
    const long spc = chunk / every;
    long nchunks = nsteps / chunk;
    if (d_snaps) nchunks = std::min<long>(nchunks, nsnap_in_loop / spc);
    if (chunk > kMaxGraphSteps) nchunks = 0;
    long s = 0;
    if (nchunks >= 1) {
        BatchGraph &g = c->bgraph;
        const uint64_t key = ((uint64_t)chunk << 40) ^ ((uint64_t)every << 32) ^ ((uint64_t)(d_snaps != nullptr) << 31) ^
                             ((uint64_t)c->rk4_exact << 30) ^ ((uint64_t)fs << 28) ^ (uint64_t)nitems;
        if (!(g.exec && g.key == key && g.variant == v && g.trajs == (const void *)dt && g.items == (const void *)dw &&
              g.guard == (const void *)first)) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = BatchGraph();
            hipStream_t cap;
            ASD_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            hipGraph_t graph;
            ASD_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            hipStream_t saved = c->stream;
            c->stream = cap;
            int r = ASD_OK;
            for (int k = 0; k < chunk && r == ASD_OK; k += fs) {
                const int slot = d_snaps && k % every == 0 ? k / every : -1;
                r = rk4_batch_launch(c, v, fs, dt, dw, nitems, (k / fs) & 1, ctr, (int)spc, slot, k + fs >= chunk,
                                     chunk_guard(first, ctr, k, (int)chunk));
            }
            c->stream = saved;
            hipError_t e1 = hipStreamEndCapture(cap, &graph);
            (void)hipStreamDestroy(cap);
            ASD_TRY(r);
            ASD_HIP(e1);
            hipError_t e2 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            ASD_HIP(e2);
            g.key = key;
            g.variant = v;
            g.trajs = dt;
            g.items = dw;
            g.guard = first;
            g.chunk = (int)chunk;
            g.nitems = nitems;
        }
        for (long ch = 0; ch < nchunks; ch++) ASD_HIP(hipGraphLaunch(g.exec, st));
        s = nchunks * chunk;
    }
    // remainder, eagerly, with explicit snapshot slots; the state is back in the a buffers
    // after the graph's even number of launches
    int cur = 0;
    while (s < nsteps) {
        const int slot = d_snaps && s % every == 0 && s / every < nsnap ? (int)(s / every) : -1;
        const bool mid_snap = d_snaps && (s + 1) % every == 0 && (s + 1) / every < nsnap;
        const int k = fs == 2 && s + 2 <= nsteps && !mid_snap ? 2 : 1;
        ASD_TRY(rk4_batch_launch(c, v, k, dt, dw, nitems, cur, nullptr, 0, slot, 0, eager_guard(first, s)));
        cur ^= 1;
        s += k;
    }
    for (int t = 0; t < ntraj; t++) {
        const size_t bytes = sizeof(double) * 3 * (size_t)Ls[t] * Ls[t];
        const double *fin = cur ? tab[t].b : tab[t].a;
        if (d_snaps && d_snaps[t] && nsteps % every == 0 && nsteps / every < nsnap)
            ASD_HIP(hipMemcpyAsync(d_snaps[t] + (nsteps / every) * 3L * Ls[t] * Ls[t], fin, bytes,
                                   hipMemcpyDeviceToDevice, st));
        if (cur) ASD_HIP(hipMemcpyAsync(d_spins[t], tab[t].b, bytes, hipMemcpyDeviceToDevice, st));
    }
    ASD_TRY(mark(c, 6));
    unsigned long long h = kGuardNone;
    ASD_HIP(hipMemcpyAsync(&h, first, sizeof h, hipMemcpyDeviceToHost, st));
    ASD_HIP(hipStreamSynchronize(st));
    return guard_verdict(h, true);
}

// ------------------------------------------------ reference-compatible entries

// one lazily created context per host thread for the host-pointer entries
static asd_ctx *compat_ctx(int *err)
{
    static thread_local asd_ctx *ctx = nullptr;
    if (!ctx) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        *err = asd_create(&ctx, dev);
        if (*err != ASD_OK) ctx = nullptr;
    }
    return ctx;
}

int asd_euler(const asd_params *p, int L, int nrep, const double *spin0, double *derivative)
{
    ASD_CHECK(spin0 && derivative, "NULL argument");
    ASD_TRY(check_params(p));
    ASD_TRY(check_L(L));
    ASD_CHECK(nrep >= 1, "nrep must be >= 1");
    int err = ASD_OK;
    asd_ctx *c = compat_ctx(&err);
    if (!c) return err;
    const size_t bytes = sizeof(double) * 3 * (size_t)nrep * L * L;
    ASD_TRY(c->c0.ensure(bytes));
    ASD_TRY(c->c1.ensure(bytes));
    ASD_HIP(hipMemcpyAsync(c->c0.ptr, spin0, bytes, hipMemcpyHostToDevice, c->stream));
    ASD_TRY(rhs_launch(c, make_llg(*p), L, nrep, c->c0.as<double>(), c->c1.as<double>()));
    ASD_HIP(hipMemcpyAsync(derivative, c->c1.ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    ASD_HIP(hipStreamSynchronize(c->stream));
    return ASD_OK;
}

int asd_runge_kutta(const asd_params *p, int L, int nrep, const double *spin, double *cur_input, double *k1,
                    double *k2, double *k3, double *k4)
{
    ASD_CHECK(spin && cur_input && k1 && k2 && k3 && k4, "NULL argument");
    ASD_TRY(check_params(p));
    ASD_TRY(check_L(L));
    ASD_CHECK(nrep >= 1, "nrep must be >= 1");
    int err = ASD_OK;
    asd_ctx *c = compat_ctx(&err);
    if (!c) return err;
    const size_t bytes = sizeof(double) * 3 * (size_t)nrep * L * L;
    DevBuf *b[6] = {&c->c0, &c->c1, &c->c2, &c->c3, &c->c4, &c->c5};
    for (DevBuf *x : b) ASD_TRY(x->ensure(bytes));
    ASD_HIP(hipMemcpyAsync(c->c0.ptr, spin, bytes, hipMemcpyHostToDevice, c->stream));
    ASD_TRY(rk4_compat_stages(c, make_llg(*p), L, nrep, c->c0.as<double>(), c->c1.as<double>(), c->c2.as<double>(),
                              c->c3.as<double>(), c->c4.as<double>(), c->c5.as<double>()));
    double *outs[5] = {cur_input, k1, k2, k3, k4};
    for (int i = 0; i < 5; i++)
        ASD_HIP(hipMemcpyAsync(outs[i], b[i + 1]->ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    ASD_HIP(hipStreamSynchronize(c->stream));
    return ASD_OK;
}

}  // extern "C"

// fused_dgemm compat lives in compat.hip; the globals-bound drop-ins below.
namespace asd {
struct CompatGlobals {
    int *N = nullptr, *Nfile = nullptr, *num_neighbor = nullptr;
    double *dt = nullptr, *J1 = nullptr;
    int (*nn)[2] = nullptr;
};
static CompatGlobals g_compat;

[[noreturn]] static void compat_die(const char *fn)
{
    fprintf(stderr, "libasd %s: %s\n", fn, asd_last_error());
    abort();
}

static int compat_params(asd_params *p)
{
    ASD_CHECK(g_compat.N && g_compat.Nfile && g_compat.dt && g_compat.J1 && g_compat.nn,
              "asd_compat_bind() was not called");
    static const int ref_nn[4][2] = {{1, 0}, {-1, 0}, {0, -1}, {0, 1}};
    ASD_CHECK(*g_compat.num_neighbor == 4, "num_neighbor must be 4");
    for (int j = 0; j < 4; j++)
        ASD_CHECK(g_compat.nn[j][0] == ref_nn[j][0] && g_compat.nn[j][1] == ref_nn[j][1],
                  "nn[][] must be the reference table {(1,0),(-1,0),(0,-1),(0,1)}");
    p->J = *g_compat.J1;
    p->D = 0.0;
    p->Bz = 0.0;
    p->alpha = 0.0;
    p->dt = *g_compat.dt;
    p->spin_length = 1.0;
    return ASD_OK;
}
}  // namespace asd

extern "C" {

int asd_compat_bind(int *N, int *Nfile, double *dt, double *J1, int (*nn)[2], int *num_neighbor)
{
    ASD_CHECK(N && Nfile && dt && J1 && nn && num_neighbor, "NULL global");
    g_compat.N = N;
    g_compat.Nfile = Nfile;
    g_compat.dt = dt;
    g_compat.J1 = J1;
    g_compat.nn = nn;
    g_compat.num_neighbor = num_neighbor;
    asd_params p;
    return compat_params(&p);
}

void euler(double *spin0, double *derivative)
{
    asd_params p;
    if (compat_params(&p) != ASD_OK || asd_euler(&p, *g_compat.N, *g_compat.Nfile, spin0, derivative) != ASD_OK)
        compat_die("euler");
}

void runge_kutta(double *spin, double *cur_input, double *k1, double *k2, double *k3, double *k4)
{
    asd_params p;
    if (compat_params(&p) != ASD_OK ||
        asd_runge_kutta(&p, *g_compat.N, *g_compat.Nfile, spin, cur_input, k1, k2, k3, k4) != ASD_OK)
        compat_die("runge_kutta");
}

void fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda, double *b, int *ldb, double *c,
                 int *ldc, double *fuse_C, double *sum)
{
    if (asd_fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_C, sum) != ASD_OK) compat_die("fused_dgemm");
}

}  // extern "C"
