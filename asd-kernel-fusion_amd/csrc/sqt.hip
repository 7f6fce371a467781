// sqt.hip -- S(q,tau) as a spin-spin correlation GEMM on MFMA with a fused
// displacement-binning epilogue, plus the FP64 finishing kernels.
//
// Reference data flow (kernel_fusion/llg_sqt_fusion.c, time origins in the role of
// the reference's replicas):
//   B operand  spin_zero[k + j*K] = S_j^c(t0)        :235-245   (k = 3 t0 + c)
//   A operand  spint[i + k*Ns]   = S_i^c(t0 + tau)   :370-377
//   averages   zbar_j, fbar_i                        :249-268, :380-398
//   prefill    C = -(fbar_i . zbar_j) Q_ij           :402-426
//   fused_dgemm: C += (alpha A B) o Q, sum += sum(C)  :429, gemm_kernel.h:355-422
//   sqt = sum / N^2                                  :432-442
// The reference materialises Q and C (2 Ns^2 doubles each: 4.3 GB apiece at L=128).
// Here nothing Ns^2-sized exists: because Q_ij depends only on r_i - r_j (mod L for
// lattice momenta), the epilogue reduces each G tile into displacement bins
//   g_tau(dr, dc) = sum_{i - j = (dr, dc) mod L} G_ij(tau)
// (LDS f64 atomics, then one global f64 atomic per bin per tile), and a finishing
// kernel evaluates sum_d g(d) e^{-iq.d} for every requested q from an exact L-entry
// twiddle table with Kahan-compensated FP64 sums.  The disconnected part
// sum_ij fbar_i.zbar_j Q_ij factorises into Fbar_q(tau).conj(Fbar_q(0)) and is
// computed in FP64 from per-snapshot Fourier sums F_q(t) = sum_i S_i(t) e^{-iq.r_i}.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

#include "asd_internal.h"

namespace asd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
// 16-byte staging vector: a native clang vector (HIP's uint4 is a struct with union
// members, which arrays of do not get promoted to registers and end up in scratch)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 128;  // tile edge of the FP64 / FP32 kernels; the split kernels' edges (128, 256) divide 2 TILE

// ------------------------------------------------------- fixed-point bins
// The displacement bins are sums of many products from many workgroups.  Floating-point
// atomics would make them depend on the order the adds land in, so the bins are integers
// instead: a product value v enters as the integer round(v 2^s), which integer atomics
// add exactly, so the sums (and S(q,tau)) repeat bit for bit.  The integer is formed by
// the magic-number rounding: for |x| < 2^51, t = x + 1.5 2^52 lies in [2^52, 2^53), where
// the double spacing is 1, so t = round(x) + 1.5 2^52 and bits(t) - bits(1.5 2^52) =
// round(x).  The kernels add bits(t) itself; each bin of a lag receives exactly Ns
// products (one j per i at every displacement), so the decoder subtracts Ns bits(1.5 2^52)
// (mod 2^64).  FP64 keeps a second word for the residual x - round(x) at a finer scale,
// which makes its bins the exactly rounded sums; the other precisions (errors >= 1e-7)
// keep one word.  The scale s comes from the largest snapshot component M, noted by the
// packing kernels: |G_ij| <= K M^2 (K = 3 T0), so s is chosen to keep every |x| below
// 2^50 and every bin below 2^62.
constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr unsigned long long kMagicBits = 0x4338000000000000ull;

struct FixScale {
    unsigned long long vmax_bits;  // bits of max |snapshot component| (non-negative doubles order as integers)
    double m0, m1;                 // 2^s0 (first word) and 2^s1 (second word, relative to the first)
    double d0, d1;                 // 2^-s0 and 2^-(s0 + s1) for the decoder
    // the split precisions pack v ps (a power of two putting max |v| in [64, 128), inside the
    // normal range of fp16 and e4m3 with room below: |v| of any size keeps their relative
    // precision) and convert their products with ms = 2^s0 / ps^2, so the bin words equal those of
    // the unscaled products
    double ps, ms;
    int bad;                       // a snapshot component is NaN or infinite
};

// max |v| over snapshots [0, nt) (components [t * tstride, t * tstride + n3)) into *vmax:
// grid-stride over components (x) and snapshots (y), one atomic per block.  The bits of
// non-negative doubles order as integers (NaN above Inf above every finite value).
__global__ __launch_bounds__(256) void vmax_kernel(const double *__restrict__ snaps, long tstride, int nt, long n3,
                                                   unsigned long long *__restrict__ vmax)
{
    __shared__ unsigned long long red[4];
    unsigned long long b = 0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (int t = blockIdx.y; t < nt; t += gridDim.y)
        for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n3; e += stride) {
            const unsigned long long x = (unsigned long long)__double_as_longlong(fabs(snaps[t * tstride + e]));
            b = x > b ? x : b;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = (unsigned long long)__shfl_xor((long long)b, o, 64);
        b = x > b ? x : b;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) b = red[w] > b ? red[w] : b;
        atomicMax(vmax, b);
    }
}

__global__ void fix_reset_kernel(FixScale *f)
{
    f->vmax_bits = 0;
    f->bad = 0;
}

// scales from the snapshot maximum (one thread); K = GEMM depth, Ns = sites
__global__ void fix_scale_kernel(FixScale *f, int K, int Ns)
{
    const double M = __longlong_as_double((long long)f->vmax_bits);
    int clog_ns = 0;
    while ((1L << clog_ns) < (long)Ns) clog_ns++;
    int e_el = 0;  // |G_ij| <= K M^2 <= 2^e_el
    int ep = 0;    // pack scale exponent of the split precisions
    if (!(M <= 1.7976931348623157e308)) {
        f->bad = 1;  // NaN or Inf
    } else if (M > 0.0) {
        int eM = 0, eK = 0;
        (void)frexp(M, &eM);  // M < 2^eM <= 2 M
        (void)frexp((double)K, &eK);
        e_el = 2 * eM + eK;
        ep = max(-480, min(480, 7 - eM));
    }
    int s0 = min(50, 62 - clog_ns) - e_el;
    s0 = max(-1000, min(1000, s0));
    const int s1 = min(50, 62 - clog_ns);  // residuals are <= 1/2
    f->m0 = ldexp(1.0, s0);
    f->m1 = ldexp(1.0, s1);
    f->d0 = ldexp(1.0, -s0);
    f->d1 = ldexp(1.0, -s0 - s1);
    f->ps = ldexp(1.0, ep);
    f->ms = ldexp(1.0, max(-1000, min(1000, s0 - 2 * ep)));
}

// one product value into the bin words of LDS (word w of bin b at sb[w * nbins + b])
template <int NW>
__device__ __forceinline__ void fix_add(unsigned long long *sb, int nbins, int b, double v, double m0, double m1)
{
    const double t = __builtin_fma(v, m0, kMagic);
    atomicAdd(&sb[b], (unsigned long long)__double_as_longlong(t));
    if constexpr (NW == 2) {
        const double r = __builtin_fma(v, m0, kMagic - t);  // v 2^s0 - round(v 2^s0), exact
        atomicAdd(&sb[nbins + b], (unsigned long long)__double_as_longlong(__builtin_fma(r, m1, kMagic)));
    }
}

// bins[w][lag][d] (integer words) -> g[lag][d] (doubles); NaN where a snapshot was not finite
__global__ void fix_decode_kernel(const unsigned long long *__restrict__ w, long n, int nw, int Ns,
                                  const FixScale *__restrict__ f, double *__restrict__ g)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const unsigned long long off = (unsigned long long)Ns * kMagicBits;
    double v = (double)(long long)(w[e] - off) * f->d0;
    if (nw == 2) v = __builtin_fma((double)(long long)(w[n + e] - off), f->d1, v);
    g[e] = f->bad ? __builtin_nan("") : v;
}

// ------------------------------------------------------------------ packing

// snapshots [t][i][c] -> k-major operand out[k][ld], k = 3 t + c; zero where t >= nt
// or i >= Ns (padding rows/columns), so padded products vanish exactly.
template <class T>
__global__ void pack_kmajor_kernel(const double *__restrict__ snaps, long tstride, int nt, int Ns,
                                   T *__restrict__ out, long ld)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= ld) return;
    const int t = k / 3, c = k - 3 * t;
    double v = (t < nt && i < Ns) ? snaps[(long)t * tstride + 3L * i + c] : 0.0;
    out[(long)k * ld + i] = (T)v;
}

// snapshots -> hi/lo split operands, k = 3 t + c, in row groups of 16 rows and k-blocks of
// 8 halves (16 B): element (i, k) at ((i / 16) nkb + k / 8) 128 + (i % 16) 8 + k % 8 halves,
// so the 16 rows of one k-block are 256 contiguous bytes and a 16-row, 32-deep chunk (four
// consecutive k-blocks) is one contiguous KiB from any k-block.
// v is the snapshot component times the pack scale fix->ps (a power of two, exact).
// BF16X3: hi = bf16(v), lo = bf16(v - hi).  FP16X2: hi = f16(v), lo = f16((v - hi) * 2^11).
// FP16F8: hi as FP16X2, the lo half holds the fp8 cross-product bytes (below).
// `nshift` shifted copies are written, copy r holding value(i, k + r) at (i, k), so a
// window starting at any k offset o is read from copy o % 8 at the k-block (o - o % 8) / 8.
// nkb: k-blocks per row group (rows >= Ns are zeros).  grid: (row groups of 32 rows,
// k-blocks); thread t: row 32 blockIdx.x + t / 8, k % 8 = t % 8.
// e4m3 (OCP FP8: 4 exponent bits, bias 7, 3 mantissa bits) of a float with |x| <= 448, rounded to
// nearest even: subnormals (|x| < 2^-6) are multiples of 2^-9, whose count rint(|x| 2^9) is also
// the code (8 lands on the smallest normal, code 0x08); normals keep the top 3 mantissa bits
// after a round-to-nearest-even add on the float's bits, a carry moving into the exponent.
__device__ __forceinline__ uint32_t to_e4m3(float x)
{
    const uint32_t sgn = (__float_as_uint(x) >> 24) & 0x80u;
    const float ax = __builtin_fabsf(x);
    if (ax < 0x1p-6f) return sgn | (uint32_t)__builtin_rintf(ax * 512.0f);
    uint32_t a = __float_as_uint(ax);
    a += 0x7FFFFu + ((a >> 20) & 1u);
    return sgn | ((a >> 20) - (120u << 3));
}

// FP16F8 writes, in place of the lo half, the cross-product bytes of the element: A (bside 0)
// e4m3(hi) in the low byte and e4m3(lo 2^11) in the high byte, B (bside 1) the other way round,
// so that A's hi byte meets B's lo byte and A's lo byte B's hi byte in the fp8 MFMA.
template <int PREC>
__global__ __launch_bounds__(256) void pack_split_kernel(const double *__restrict__ snaps, long tstride, int nt,
                                                         int Ns, uint16_t *__restrict__ hi,
                                                         uint16_t *__restrict__ lo, long nkb, long copy_stride,
                                                         int nshift, int bside, const FixScale *__restrict__ fix)
{
    const double ps = fix->ps;
    const int i = blockIdx.x * 32 + threadIdx.x / 8, e = threadIdx.x % 8;
    const long kb = blockIdx.y;
    const long o = ((long)(i / 16) * nkb + kb) * 128 + (i % 16) * 8 + e;
    for (int r = 0; r < nshift; r++) {
        const long kk = kb * 8 + e + r;
        const long t = kk / 3;
        const int c = (int)(kk - 3 * t);
        const float v = (t < nt && i < Ns) ? (float)(snaps[t * tstride + 3L * i + c] * ps) : 0.0f;
        uint16_t h, l;
        if (PREC == ASD_BF16X3) {
            __hip_bfloat16 bh = __float2bfloat16(v);
            const float rr = v - __bfloat162float(bh);
            __hip_bfloat16 bl = __float2bfloat16(rr);
            h = *reinterpret_cast<uint16_t *>(&bh);
            l = *reinterpret_cast<uint16_t *>(&bl);
        } else if (PREC == ASD_FP16X2) {
            __half fh = __float2half(v);
            const float rr = (v - __half2float(fh)) * 2048.0f;
            __half fl = __float2half(rr);
            h = *reinterpret_cast<uint16_t *>(&fh);
            l = *reinterpret_cast<uint16_t *>(&fl);
        } else {
            __half fh = __float2half(v);
            const float hv = __half2float(fh);
            const uint32_t xh = to_e4m3(hv), xl = to_e4m3((v - hv) * 2048.0f);
            h = *reinterpret_cast<uint16_t *>(&fh);
            l = (uint16_t)(bside ? (xl | xh << 8) : (xh | xl << 8));
        }
        hi[r * copy_stride + o] = h;
        lo[r * copy_stride + o] = l;
    }
}

// ------------------------------------------------------- work decomposition

struct Work {
    int lag, ti, tj;
};

// Work slots are processed in a grid-stride loop by a grid that is a multiple of 8: a
// persistent one (as many blocks as are resident at once) for the split-precision
// kernels, a capped one for the k-major kernels.  Blocks are dispatched round-robin
// over the 8 XCDs (block b -> XCD b % 8), so slot s always runs on XCD s % 8.  XCD x
// takes the x-th eighth of every lag's tiles and all XCDs walk the lags in step: the
// operands in flight are one lag's A window plus B, which the shared last-level cache
// holds.  Within an XCD, consecutive slots run together; tiles are ordered in groups
// of `group` tile-rows so that they share A and B panels in the XCD's L2.
__device__ __forceinline__ bool decode_work(long slot, int nti, int ntj, int nlags, int group, Work &w)
{
    const long per_lag = (long)nti * ntj, per_x = (per_lag + 7) / 8;
    const long u = slot / 8, lag = u / per_x;
    if (lag >= nlags) return false;
    const long r = (slot % 8) * per_x + (u - lag * per_x);
    if (r >= per_lag) return false;
    w.lag = (int)lag;
    const int span = group * ntj, g = (int)(r / span), first = g * group;
    const int gsz = min(nti - first, group), rr = (int)(r - (long)g * span);
    w.ti = first + rr % gsz;
    w.tj = rr / gsz;
    return true;
}

constexpr long kMaxGrid = 1L << 16;  // cap on the persistent grid (blocks)

// slots for nlags lags of per_lag tiles (each XCD gets ceil(per_lag / 8) per lag)
__host__ __device__ inline long num_slots(long per_lag, int nlags) { return 8 * ((per_lag + 7) / 8) * nlags; }

// --------------------------------------------------------- binning epilogue

struct BinGeom {
    int drmin, nbins;
};

__device__ __forceinline__ BinGeom bin_geom(int i0, int j0, int Ns, int L, int BM = TILE, int BN = TILE)
{
    const int il = min(i0 + BM, Ns) - 1, jl = min(j0 + BN, Ns) - 1;
    const int ri0 = i0 / L, ri1 = il / L, rj0 = j0 / L, rj1 = jl / L;
    BinGeom g;
    g.drmin = ri0 - rj1;
    g.nbins = ((ri1 - ri0) + (rj1 - rj0) + 1) * L;
    return g;
}

// the tile's LDS bin words (word w of bin b at sb[w * nbins + b]) into the lag's global
// bins (word w at gbins[w * wstride + d]); integer adds, so the order does not matter
template <int NW>
__device__ __forceinline__ void flush_bins(const unsigned long long *sb, const BinGeom &g, int L,
                                           unsigned long long *gbins, long wstride)
{
    for (int b = threadIdx.x; b < g.nbins; b += blockDim.x) {
        int dr = (g.drmin + b / L) % L;
        if (dr < 0) dr += L;
        const long d = (long)dr * L + b % L;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const unsigned long long v = sb[w * g.nbins + b];
            if (v != 0) atomicAdd(&gbins[w * wstride + d], v);
        }
    }
}

struct GemmArgs {
    const void *A;   // A operand base (lag 0)
    const void *B;   // B operand
    const void *A2;  // lo parts (split precisions)
    const void *B2;
    long ld;         // k-major: row stride in elements; split: k-blocks per row group
    long copy_stride;  // i-major A: elements between the 8 shifted copies
    int L, Ns, Kpad, nt, nlags;
    int K;           // k extent without the zero padding (3 T0)
    int group;       // tile-rows per group in the work order
    unsigned long long *bins;  // integer bin words [word][nlags][Ns]
    long wstride;    // nlags * Ns: words between the first and second word of a bin
    const FixScale *fix;
};

// s_waitcnt immediates (gfx9 encoding: vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[5:4]
// in [15:14]) that wait on one counter only
constexpr int wait_vmcnt(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
constexpr int kWaitLgkm0 = 15 | (7 << 4) | (3 << 14);

// ------------------------------------------------------------ FP64 / FP32
// k-major operands; v_mfma_{f64_16x16x4f64, f32_16x16x4f32}: lane l holds A[i=l%16][k=l/16],
// B[k=l/16][j=l%16].  The 4 outputs of lane l are D[4(l/16)+v][l%16] for f32 but
// D[l/16 + 4v][l%16] for f64 (measured on gfx950 with a lane-layout probe).
// 256 threads = 2 x 2 waves, each wave 64 x 64 = 4 x 4 MFMA blocks; BK = 16 per LDS stage,
// two stages (one barrier per stage).

template <class T> struct KMajorMfma;
template <> struct KMajorMfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return lane / 16 + 4 * v; }
};
template <> struct KMajorMfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

// DMA = 0: operand rows staged through registers, two LDS stages of 16 k rows.
// DMA = 1: rows moved by direct global -> LDS loads (buffer_load_dwordx4 ... lds: 16 B per lane,
// one load fills one 128-element row, FP64 with 64 lanes, FP32 with 32) through a ring of 4 stages
// of 8 k rows; each step
// reads the next stage's first fragments after its MFMAs, so the MFMAs after a barrier start
// on operands already in registers.
template <class T, int DMA = 0>
__global__ __launch_bounds__(256, 2) void sqt_gemm_kmajor_kernel(GemmArgs args)
{
    // rows padded by 16 elements: the four row groups of a fragment read (lane / 16) then fall on
    // disjoint banks for FP32 and pairwise disjoint for FP64 (FP32 0.7% faster, FP64 unchanged)
    constexpr int BK = 16, PAD = 16, LDS_LD = TILE + PAD;
    typedef typename KMajorMfma<T>::acc_t acc_t;
    __shared__ __attribute__((aligned(16))) unsigned char smem_raw[2 * 2 * BK * LDS_LD * sizeof(T) > 8 * 4096
                                                                       ? 2 * 2 * BK * LDS_LD * sizeof(T)
                                                                       : 8 * 4096];
    T(*sA)[BK][LDS_LD] = reinterpret_cast<T(*)[BK][LDS_LD]>(smem_raw);
    T(*sB)[BK][LDS_LD] = reinterpret_cast<T(*)[BK][LDS_LD]>(smem_raw + 2 * BK * LDS_LD * sizeof(T));

    const int nt_tiles = (args.Ns + TILE - 1) / TILE;
    const long nslots = num_slots((long)nt_tiles * nt_tiles, args.nlags);
    for (long slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    Work w;
    if (!decode_work(slot, nt_tiles, nt_tiles, args.nlags, args.group, w)) continue;
    __syncthreads();  // LDS of the previous work item fully consumed
    const int i0 = w.ti * TILE, j0 = w.tj * TILE;
    const long ld = args.ld;
    const T *Ap = static_cast<const T *>(args.A) + (long)(3 * w.lag) * ld + i0;
    const T *Bp = static_cast<const T *>(args.B) + j0;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave >> 1, wj = wave & 1;
    // global -> LDS mapping: row k = tid / 16, 8 consecutive columns
    constexpr int VEC = 8;
    const int lk = tid / 16, lc = (tid % 16) * VEC;

    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = acc_t{0, 0, 0, 0};

    if constexpr (DMA != 0) {
        // ring slot s, array a (0: A, 1: B), k row r at ((2 s + a) DK + r) LDS_LD elements; the
        // padded row stride keeps the fragment reads' bank pattern of the register-staged form
        constexpr int DK = 8, NS = 4, PER_STAGE = 4;  // k rows per stage, stages, loads per wave and stage
        static_assert(NS * 2 * DK * LDS_LD * sizeof(T) <= sizeof(smem_raw), "ring exceeds the LDS array");
        static_assert(DK == 2 * 4, "each of the 4 waves loads 2 A rows and 2 B rows per stage");
        constexpr int LPR = TILE * (int)sizeof(T) / 16;  // lanes that fill one row (16 B each)
        static_assert(LPR == 64 || LPR == 32, "a row is one load of 32 or 64 lanes");
        T *ring = reinterpret_cast<T *>(smem_raw);
        const int nkd = args.Kpad / DK;
        const int wv = __builtin_amdgcn_readfirstlane(wave);
        // one buffer resource per array and stage, based at the stage's first k row (64-bit
        // address arithmetic, scalar); a row's source is that base + r ld 8 B (SGPR) + 16 B per lane
        auto issue = [&](int kt, int st) {
            const long kb = (long)kt * DK * ld;
            const __amdgpu_buffer_rsrc_t rA =
                __builtin_amdgcn_make_buffer_rsrc((void *)(Ap + kb), 0, 0x7fffffff, 0x00020000);
            const __amdgpu_buffer_rsrc_t rB =
                __builtin_amdgcn_make_buffer_rsrc((void *)(Bp + kb), 0, 0x7fffffff, 0x00020000);
            if (LPR == 64 || lane < LPR) {  // FP32: the upper half-wave is masked off (it writes nothing)
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int r = 2 * wv + u;
                    const int so = (int)(r * ld * (long)sizeof(T));
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rA, (__attribute__((address_space(3))) void *)(ring + ((2 * st + 0) * DK + r) * LDS_LD), 16,
                        16 * lane, so, 0, 0);
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                        rB, (__attribute__((address_space(3))) void *)(ring + ((2 * st + 1) * DK + r) * LDS_LD), 16,
                        16 * lane, so, 0, 0);
                }
            }
        };
        auto frag = [&](int st, int kk, T(&fa)[4], T(&fb)[4]) {
            const int kr = kk * 4 + lane / 16;
            const T *sa = ring + ((2 * st + 0) * DK + kr) * LDS_LD + wi * 64 + lane % 16;
            const T *sb = ring + ((2 * st + 1) * DK + kr) * LDS_LD + wj * 64 + lane % 16;
#pragma unroll
            for (int m = 0; m < 4; m++) fa[m] = sa[m * 16];
#pragma unroll
            for (int n = 0; n < 4; n++) fb[n] = sb[n * 16];
        };
        auto mfmas = [&](const T(&fa)[4], const T(&fb)[4]) {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int n = 0; n < 4; n++) acc[m][n] = KMajorMfma<T>::mfma(fa[m], fb[n], acc[m][n]);
        };
        // prologue: stages 0 .. NS - 2 requested (a K shorter than that re-reads its last stage,
        // so the counts waited for stay exact), stage 0 landed everywhere, its first fragments read
#pragma unroll
        for (int st = 0; st < NS - 1; st++) issue(min(st, nkd - 1), st);
        __builtin_amdgcn_s_waitcnt(wait_vmcnt(PER_STAGE * (NS - 2)));
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        T pa[4], pb[4];
        frag(0, 0, pa, pb);
        // step k: wait for this wave's loads of stage k + 1 (stage k + 2's may stay in flight),
        // barrier (stage k + 1 is then in LDS for every wave, and every wave has finished reading
        // the slot of step k - 1), refill that slot with stage k + 3, then the MFMAs of the two
        // 4-deep steps of stage k (the first on the fragments read in step k - 1), and read stage
        // k + 1's first fragments.  Reads of slot k - 1 all fed MFMAs of step k - 1, so they are
        // complete by the barrier; only the prefetch of slot k + 1 is in flight across it.
        for (int k = 0; k < nkd; k++) {
            const int h = k & (NS - 1);
            if (k + 2 < nkd)
                __builtin_amdgcn_s_waitcnt(wait_vmcnt(PER_STAGE));
            else
                __builtin_amdgcn_s_waitcnt(wait_vmcnt(0));
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (k + NS - 1 < nkd) issue(k + NS - 1, (k + NS - 1) & (NS - 1));
            __builtin_amdgcn_s_setprio(1);
            T fa[4], fb[4];
            frag(h, 1, fa, fb);
            mfmas(pa, pb);
            mfmas(fa, fb);
            if (k + 1 < nkd) frag((h + 1) & (NS - 1), 0, pa, pb);
            __builtin_amdgcn_s_setprio(0);
            asm volatile("" ::: "memory");
        }
        // every load in flight (a short K's re-reads included) lands before the bins reuse the LDS
        __builtin_amdgcn_s_waitcnt(wait_vmcnt(0));
        __syncthreads();
    } else {
    T ra[VEC], rb[VEC];
    auto gload = [&](int kt) {
        const T *a = Ap + (long)(kt * BK + lk) * ld + lc;
        const T *b = Bp + (long)(kt * BK + lk) * ld + lc;
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            ra[v] = a[v];
            rb[v] = b[v];
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int v = 0; v < VEC; v++) {
            sA[buf][lk][lc + v] = ra[v];
            sB[buf][lk][lc + v] = rb[v];
        }
    };

    const int nk = args.Kpad / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload(kt + 1);
        // the stage's fragment reads and MFMAs at raised wave priority, its LDS stores and
        // barrier at the normal one: the SIMD's other wave (the other workgroup's) then gets
        // its stores and barrier issued in the gaps of this wave's MFMAs instead of
        // competing with them (FP64 at K = 1200: 950 -> 926 ms per 100 lags; priority
        // around each 4-deep step's MFMAs only, 998 ms, or over the stores too, 993 ms, lose)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int kk = 0; kk < BK / 4; kk++) {
            const int kr = kk * 4 + lane / 16;
            T fa[4], fb[4];
#pragma unroll
            for (int m = 0; m < 4; m++) fa[m] = sA[cur][kr][wi * 64 + m * 16 + lane % 16];
#pragma unroll
            for (int n = 0; n < 4; n++) fb[n] = sB[cur][kr][wj * 64 + n * 16 + lane % 16];
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int n = 0; n < 4; n++) acc[m][n] = KMajorMfma<T>::mfma(fa[m], fb[n], acc[m][n]);
        }
        __builtin_amdgcn_s_setprio(0);
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }
    }

    // ---- epilogue: displacement binning (FP64 keeps the residual word)
    constexpr int NW = std::is_same<T, double>::value ? 2 : 1;
    static_assert(sizeof(smem_raw) >= NW * 8 * (2 * TILE + 3 * 1024), "bins of a tile at L = 1024");
    const int L = args.L, Ns = args.Ns;
    const BinGeom g = bin_geom(i0, j0, Ns, L);
    const double m0 = args.fix->m0, m1 = args.fix->m1;
    unsigned long long *sb = reinterpret_cast<unsigned long long *>(smem_raw);
    // FP64: unwrapped bins where the LDS holds them (L <= 512; the split kernels' layout:
    // word (ri - rj - drmin) 2L + L + ci - cj, the flush adds the words of dc and dc - L):
    // a wave whose 64 rows and 64 columns each stay within one lattice row adds its products
    // at one per-lane base plus compile-time offsets, the 4 - |m - n| blocks of one
    // displacement diagonal d = m - n summed in registers first (28 words per lane and word
    // plane instead of 64).  FP64 at L = 128: 3852 -> 3814 ms (K = 300: 262 -> 255 ms); the
    // FP32 kernel measured 4.7% slower with it at K = 1200 (123 VGPRs, a different
    // occupancy) and keeps the wrapped bins.
    if (std::is_same<T, double>::value && 2L * NW * g.nbins * 8 <= (long)sizeof(smem_raw)) {
        const int UP = 2 * g.nbins;  // words per word plane
        for (int b = tid; b < NW * UP; b += 256) sb[b] = 0;
        __syncthreads();
        const int wi0 = i0 + wi * 64, wj0 = j0 + wj * 64, ciw = wi0 % L, cjw = wj0 % L;
        if (i0 + TILE <= Ns && j0 + TILE <= Ns && ciw + 64 <= L && cjw + 64 <= L) {
            // lane l holds rows 16 m + KMajorMfma<T>::row(l, v) and columns 16 n + l % 16 of the
            // wave's 64 x 64 block; row(l, v) = l / 16 + 4 v (FP64) or 4 (l / 16) + v (FP32)
            const int base = (wi0 / L - wj0 / L - g.drmin) * 2 * L + L + ciw - cjw - lane % 16 - 16 * 3;
            unsigned long long *sw = sb + base;
#pragma unroll
            for (int d = -3; d <= 3; d++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    unsigned long long w0 = 0, w1 = 0;
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const int n = m - d;
                        if (n < 0 || n > 3) continue;
                        const double x = (double)acc[m][n][v];
                        const double t = __builtin_fma(x, m0, kMagic);
                        w0 += (unsigned long long)__double_as_longlong(t);
                        if constexpr (NW == 2) {
                            const double r = __builtin_fma(x, m0, kMagic - t);  // exact residual
                            w1 += (unsigned long long)__double_as_longlong(__builtin_fma(r, m1, kMagic));
                        }
                    }
                    const int o = 16 * (d + 3) + KMajorMfma<T>::row(lane, v);
                    atomicAdd(&sw[o], w0);
                    if constexpr (NW == 2) atomicAdd(&sw[UP + o], w1);
                }
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = wi0 + m * 16 + KMajorMfma<T>::row(lane, v);
                    if (i >= Ns) continue;
                    const int ri = i / L, ci = i - ri * L;
#pragma unroll
                    for (int n = 0; n < 4; n++) {
                        const int j = wj0 + n * 16 + lane % 16;
                        if (j >= Ns) continue;
                        const int rj = j / L, cj = j - rj * L;
                        fix_add<NW>(sb, UP, (ri - rj - g.drmin) * 2 * L + L + ci - cj, (double)acc[m][n][v], m0, m1);
                    }
                }
        }
        __syncthreads();
        for (int b = tid; b < g.nbins; b += 256) {
            const int row = b / L, dc = b - row * L;
            int dr = (g.drmin + row) % L;
            if (dr < 0) dr += L;
            const long dst = (long)dr * L + dc;
#pragma unroll
            for (int wd = 0; wd < NW; wd++) {
                const unsigned long long v = sb[wd * UP + 2 * L * row + dc] + sb[wd * UP + 2 * L * row + L + dc];
                if (v != 0) atomicAdd(&args.bins[(long)w.lag * Ns + wd * args.wstride + dst], v);
            }
        }
        continue;
    }
    for (int b = tid; b < NW * g.nbins; b += 256) sb[b] = 0;
    int rj[4], cj[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int j = j0 + wj * 64 + n * 16 + lane % 16;
        rj[n] = j / L;
        cj[n] = j - rj[n] * L;
    }
    __syncthreads();
    if (std::is_same<T, double>::value && i0 + TILE <= Ns && j0 + TILE <= Ns && L >= 64) {
        // full tile: with i = ri L + ci and j = rj L + cj the bin (ri - rj - drmin) L + (ci - cj
        // mod L) is (i - j) - drmin L + (ci < cj ? L : 0).  A lane's rows are ib + off with
        // off < 64 <= L, so ci needs at most one wrap and no division (FP64 3.886 -> 3.832 s at
        // the L = 128 workload; the FP32 kernel measured 5% slower with it and keeps the divisions)
        const int ib = i0 + wi * 64, cib = ib % L;
        int jb[4];
#pragma unroll
        for (int n = 0; n < 4; n++) jb[n] = j0 + wj * 64 + n * 16 + lane % 16 + g.drmin * L;
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int off = m * 16 + KMajorMfma<T>::row(lane, v);
                int ci = cib + off;
                ci -= ci >= L ? L : 0;
#pragma unroll
                for (int n = 0; n < 4; n++)
                    fix_add<NW>(sb, g.nbins, ib + off - jb[n] + (ci < cj[n] ? L : 0), (double)acc[m][n][v], m0, m1);
            }
    } else {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = i0 + wi * 64 + m * 16 + KMajorMfma<T>::row(lane, v);
            if (i >= Ns) continue;
            const int ri = i / L, ci = i - ri * L;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const int j = j0 + wj * 64 + n * 16 + lane % 16;
                if (j >= Ns) continue;
                int dc = ci - cj[n];
                dc += dc < 0 ? L : 0;
                fix_add<NW>(sb, g.nbins, (ri - rj[n] - g.drmin) * L + dc, (double)acc[m][n][v], m0, m1);
            }
        }
    }
    __syncthreads();
    flush_bins<NW>(sb, g, L, args.bins + (long)w.lag * Ns, args.wstride);
    }
}

// ------------------------------------------------------------- BF16X3 / FP16X2
// i-major operands (k contiguous).  v_mfma_f32_32x32x16_{bf16,f16}: lane l holds
// A[i=l%32][k=8(l/32)+e], B[k=8(l/32)+e][j=l%32] (e = 0..7); the 16 outputs of lane l
// are D[8(v/4) + 4(l/32) + v%4][l%32] (checked with a lane-layout probe on gfx950).
// 256 threads = 2 x 2 waves, each wave WM x WN blocks of 32 x 32; BK = 32 per LDS stage.
//   BF16X3: acc = sum Ah Bh + Ah Bl + Al Bh                    G = acc
//   FP16X2: acc = sum Ah Bh; acc2 = sum Ah Bl + Al Bh (lo parts carry 2^11)
//           G = acc + 2^-11 acc2.  One accumulator with the hi products scaled by
//           2^11 instead was measured ~20x less accurate at K = 1200: the MFMA's
//           internal sum does not keep the small cross terms next to the large ones.
// The three product passes run outermost so that consecutive MFMAs write
// independent accumulators (back-to-back MFMAs on one accumulator stall on its
// latency).

template <int PREC> struct SplitMfma;
template <> struct SplitMfma<ASD_BF16X3> {
    typedef bf8 v_t;
    static __device__ __forceinline__ f16v mfma(v_t a, v_t b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    static constexpr bool dual = false;
};
template <> struct SplitMfma<ASD_FP16X2> {
    typedef h8 v_t;
    static __device__ __forceinline__ f16v mfma(v_t a, v_t b, f16v c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    static constexpr bool dual = true;
};
// FP16F8: the hi products as FP16X2; the cross products on v_mfma_f32_32x32x64_f8f6f4 with
// e4m3 operands (format 0 for A and B, no block scales).  Lane l supplies 32 bytes of row
// (A) or column (B) l % 32; byte e of lane l in A meets byte e of lane l' in B when l / 32 ==
// l' / 32 (checked with a lane-layout probe on gfx950), so any assignment of k slots to bytes
// that A and B share gives the same sum.
template <> struct SplitMfma<ASD_FP16F8> : SplitMfma<ASD_FP16X2> {
    static __device__ __forceinline__ f16v mfma8(v8i a, v8i b, f16v c)
    {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
    }
};

// Epilogue of the 32x32x16 split kernels: bins one wave's accumulators (rows
// i0 + r0 + 32 m + 8 (v / 4) + 4 (lane / 32) + v % 4, columns j0 + c0 + 32 n + lane % 32)
// into the tile's displacement bins in LDS (sb: one fixed-point word per bin, zeroed here;
// the caller has finished with that LDS) and adds them to the lag's global bins.  FP16X2
// keeps the cross products, scaled by 2^11, in acc2.
// The product value of accumulator element (m, n, v): FP16X2 folds acc + acc2 / 2048 in
// FP32 (one rounding at the accumulators' own precision), the same in every layout.
template <int WM, int WN, bool DUAL>
__device__ __forceinline__ double split_value(const f16v (&acc)[WM][WN],
                                              const f16v (&acc2)[DUAL ? WM : 1][DUAL ? WN : 1], int m, int n, int v)
{
    if constexpr (DUAL)
        return (double)__builtin_fmaf(acc2[m][n][v], 1.0f / 2048.0f, acc[m][n][v]);
    else
        return (double)acc[m][n][v];
}

// The element loop: one wave's accumulators into the tile's zeroed LDS bin words (geometry g).
template <int WM, int WN, bool DUAL>
__device__ __forceinline__ void bin_split_elems(const f16v (&acc)[WM][WN],
                                                const f16v (&acc2)[DUAL ? WM : 1][DUAL ? WN : 1],
                                                unsigned long long *sb, const BinGeom &g, int i0, int j0, int r0,
                                                int c0, int BM, int BN, int Ns, int L, double m0)
{
    const int lane = threadIdx.x & 63;
    if (i0 + BM <= Ns && j0 + BN <= Ns && L >= 32) {
        // full tile: with i = ri L + ci and j = rj L + cj the bin
        // (ri - rj - drmin) L + (ci - cj mod L) is (i - j) - drmin L + (ci < cj ? L : 0).
        // A lane's rows are ib + 32 m + off(v) with off(v) < 32 <= L, so ci needs at
        // most one wrap and no division per element.
        const int ib = i0 + r0 + 4 * (lane / 32);
        int jb[WN], cj[WN];
#pragma unroll
        for (int n = 0; n < WN; n++) {
            const int j = j0 + c0 + n * 32 + lane % 32;
            jb[n] = j + g.drmin * L;
            cj[n] = j % L;
        }
#pragma unroll
        for (int m = 0; m < WM; m++) {
            const int im = ib + m * 32, cim = im % L;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int off = 8 * (v / 4) + (v % 4);
                int ci = cim + off;
                ci -= ci >= L ? L : 0;
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    const double val = split_value<WM, WN, DUAL>(acc, acc2, m, n, v);
                    fix_add<1>(sb, g.nbins, im - jb[n] + (ci < cj[n] ? L : 0) + off, val, m0, 0.0);
                }
            }
        }
    } else {
        int rj[WN], cj[WN];
#pragma unroll
        for (int n = 0; n < WN; n++) {
            const int j = j0 + c0 + n * 32 + lane % 32;
            rj[n] = j / L;
            cj[n] = j - rj[n] * L;
        }
#pragma unroll
        for (int m = 0; m < WM; m++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = i0 + r0 + m * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (i >= Ns) continue;
                const int ri = i / L, ci = i - ri * L;
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    const int j = j0 + c0 + n * 32 + lane % 32;
                    if (j >= Ns) continue;
                    int dc = ci - cj[n];
                    dc += dc < 0 ? L : 0;
                    const double val = split_value<WM, WN, DUAL>(acc, acc2, m, n, v);
                    fix_add<1>(sb, g.nbins, (ri - rj[n] - g.drmin) * L + dc, val, m0, 0.0);
                }
            }
    }
}

template <int WM, int WN, bool DUAL>
__device__ __forceinline__ void bin_split_tile(const f16v (&acc)[WM][WN],
                                               const f16v (&acc2)[DUAL ? WM : 1][DUAL ? WN : 1],
                                               unsigned long long *sb, int i0, int j0, int r0, int c0, int BM,
                                               int BN, int Ns, int L, unsigned long long *gbins, double m0)
{
    const BinGeom g = bin_geom(i0, j0, Ns, L, BM, BN);
    for (int b = threadIdx.x; b < g.nbins; b += blockDim.x) sb[b] = 0;
    __syncthreads();
    bin_split_elems<WM, WN, DUAL>(acc, acc2, sb, g, i0, j0, r0, c0, BM, BN, Ns, L, m0);
    __syncthreads();
    flush_bins<1>(sb, g, L, gbins, 0);
}

// The same epilogue over an unwrapped LDS layout, for kernels whose LDS holds twice the
// tile's bins (at most 2 (BM + BN + 3 L) words, 55 KB at L = 1024): product (i, j) with
// lattice rows ri, rj and columns ci, cj goes to word (ri - rj - drmin) 2L + L + (ci - cj),
// with no reduction of ci - cj mod L; the flush adds the words of ci - cj = dc and dc - L,
// which are the one bin (ri - rj - drmin) L + dc of the wrapped layout.  A wave whose
// 32 WM rows and 32 WN columns each stay within one lattice row (every wave of the tiles
// at L = 64, 128, 256, ...) then has ri - rj fixed and ci - cj = const + 32 (m - n) +
// 8 (v / 4) + v % 4 - lane % 32 + 4 (lane / 32): its word addresses are one per-lane base
// plus compile-time offsets, and the blocks with equal m - n (same words) are summed in
// registers first.  Integer words, so the sums do not depend on the order.
template <int WM, int WN, bool DUAL>
__device__ __forceinline__ void bin_split_elems_unwrapped(const f16v (&acc)[WM][WN],
                                                          const f16v (&acc2)[DUAL ? WM : 1][DUAL ? WN : 1],
                                                          unsigned long long *su, const BinGeom &g, int i0, int j0,
                                                          int r0, int c0, int BM, int BN, int Ns, int L, double m0)
{
    const int lane = threadIdx.x & 63;
    const int wi0 = i0 + r0, wj0 = j0 + c0;  // the wave's first row and column
    const int ci0 = wi0 % L, cj0 = wj0 % L;
    if (i0 + BM <= Ns && j0 + BN <= Ns && ci0 + 32 * WM <= L && cj0 + 32 * WN <= L) {
        const int base = (wi0 / L - wj0 / L - g.drmin) * 2 * L + L + ci0 - cj0 + 4 * (lane / 32) - lane % 32 -
                         32 * (WN - 1);  // word of m - n = -(WN - 1), v = 0
        unsigned long long *sw = su + base;
#pragma unroll
        for (int d = -(WN - 1); d <= WM - 1; d++) {  // d = m - n
#pragma unroll
            for (int v = 0; v < 16; v++) {
                unsigned long long word = 0;
#pragma unroll
                for (int m = 0; m < WM; m++) {
                    const int n = m - d;
                    if (n < 0 || n >= WN) continue;
                    const double t = __builtin_fma(split_value<WM, WN, DUAL>(acc, acc2, m, n, v), m0, kMagic);
                    word += (unsigned long long)__double_as_longlong(t);
                }
                atomicAdd(&sw[32 * (d + WN - 1) + 8 * (v / 4) + (v % 4)], word);
            }
        }
    } else {
        int rj[WN], cj[WN];
#pragma unroll
        for (int n = 0; n < WN; n++) {
            const int j = wj0 + n * 32 + lane % 32;
            rj[n] = j / L;
            cj[n] = j - rj[n] * L;
        }
#pragma unroll
        for (int m = 0; m < WM; m++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = wi0 + m * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
                if (i >= Ns) continue;
                const int ri = i / L, ci = i - ri * L;
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    const int j = wj0 + n * 32 + lane % 32;
                    if (j >= Ns) continue;
                    const double t = __builtin_fma(split_value<WM, WN, DUAL>(acc, acc2, m, n, v), m0, kMagic);
                    atomicAdd(&su[(ri - rj[n] - g.drmin) * 2 * L + L + ci - cj[n]],
                              (unsigned long long)__double_as_longlong(t));
                }
            }
    }
}

template <int WM, int WN, bool DUAL>
__device__ __forceinline__ void bin_split_tile_unwrapped(const f16v (&acc)[WM][WN],
                                                         const f16v (&acc2)[DUAL ? WM : 1][DUAL ? WN : 1],
                                                         unsigned long long *su, int i0, int j0, int r0, int c0,
                                                         int BM, int BN, int Ns, int L, unsigned long long *gbins,
                                                         double m0)
{
    const BinGeom g = bin_geom(i0, j0, Ns, L, BM, BN);
    for (int b = threadIdx.x; b < 2 * g.nbins; b += blockDim.x) su[b] = 0;
    __syncthreads();
    bin_split_elems_unwrapped<WM, WN, DUAL>(acc, acc2, su, g, i0, j0, r0, c0, BM, BN, Ns, L, m0);
    __syncthreads();
    for (int b = threadIdx.x; b < g.nbins; b += blockDim.x) {
        const int row = b / L, dc = b - row * L;
        const unsigned long long v = su[2 * L * row + dc] + su[2 * L * row + L + dc];
        int dr = (g.drmin + row) % L;
        if (dr < 0) dr += L;
        if (v != 0) atomicAdd(&gbins[(long)dr * L + dc], v);
    }
}

template <int PREC, int WM, int WN, int WVM = 2, int WVN = 2>
__global__ __launch_bounds__(64 * WVM * WVN) void sqt_gemm_split_kernel(GemmArgs args)
{
    constexpr int BK = 32, LDS_LD = BK + 8;  // halves per LDS row (pad 16 B)
    constexpr int NT = 64 * WVM * WVN;       // WVM x WVN waves of WM x WN 32x32 blocks
    constexpr int BM = WVM * WM * 32, BN = WVN * WN * 32;
    typedef typename SplitMfma<PREC>::v_t v_t;
    // [buf][hi, lo][row][k]
    __shared__ __attribute__((aligned(16))) uint16_t sA[2][2][BM][LDS_LD];
    __shared__ __attribute__((aligned(16))) uint16_t sB[2][2][BN][LDS_LD];

    const int nti = (args.Ns + BM - 1) / BM, ntj = (args.Ns + BN - 1) / BN;
    const long nslots = num_slots((long)nti * ntj, args.nlags);
    for (long slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    Work w;
    if (!decode_work(slot, nti, ntj, args.nlags, args.group, w)) continue;
    __syncthreads();  // LDS of the previous work item fully consumed
    const int i0 = w.ti * BM, j0 = w.tj * BN;
    const long ld = args.ld;  // k-blocks per row group
    const long koff = 3L * w.lag, shift = koff & 7;
    const long acopy = shift * args.copy_stride + ((i0 / 16) * ld + (koff - shift) / 8) * 128;
    const uint16_t *Ah = static_cast<const uint16_t *>(args.A) + acopy;
    const uint16_t *Al = static_cast<const uint16_t *>(args.A2) + acopy;
    const uint16_t *Bh = static_cast<const uint16_t *>(args.B) + (j0 / 16) * ld * 128;
    const uint16_t *Bl = static_cast<const uint16_t *>(args.B2) + (j0 / 16) * ld * 128;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave / WVN, wj = wave % WVN;
    // global -> LDS: a row of one operand tile is 32 halves = 4 x 16 B (k-blocks); piece e of
    // an operand is (row 16 (e / 64) + e % 16, quarter (e / 16) % 4), so consecutive threads
    // read consecutive 16 B of a row group's contiguous KiB; thread t moves pieces t, t+NT, ...
    constexpr int PA = BM * 4 / NT, PB = BN * 4 / NT;  // pieces per thread per array
    static_assert(PA * NT == BM * 4 && PB * NT == BN * 4, "tile rows must split evenly over threads");

    constexpr bool DUAL = SplitMfma<PREC>::dual;
    f16v acc[WM][WN], acc2[DUAL ? WM : 1][DUAL ? WN : 1];
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[a][b][e] = 0.0f;
    if constexpr (DUAL) {
#pragma unroll
        for (int a = 0; a < WM; a++)
#pragma unroll
            for (int b = 0; b < WN; b++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc2[a][b][e] = 0.0f;
    }

    // two register stages: the loads for iteration kt+2 are issued while kt is
    // computed, so every global load has about two iterations to arrive.  The loop
    // is unrolled by two (h) so that every register-array index is a constant
    // (no lambdas capturing arrays: those push them to scratch).
    u32x4 ra[2][2][PA], rb[2][2][PB];  // [stage][hi, lo][piece]
#define ASD_SPLIT_GLOAD(KT, ST)                                                          \
    do {                                                                                 \
        const long kb_ = (long)(KT) * (BK / 8);                                          \
        _Pragma("unroll") for (int u = 0; u < PA; u++) {                                 \
            const int e = tid + NT * u, qq = (e >> 4) & 3;                                 \
            const long off = ((e >> 6) * ld + kb_ + qq) * 128 + (e & 15) * 8;           \
            ra[ST][0][u] = *reinterpret_cast<const u32x4 *>(Ah + off);                   \
            ra[ST][1][u] = *reinterpret_cast<const u32x4 *>(Al + off);                   \
        }                                                                                \
        _Pragma("unroll") for (int u = 0; u < PB; u++) {                                 \
            const int e = tid + NT * u, qq = (e >> 4) & 3;                                 \
            const long off = ((e >> 6) * ld + kb_ + qq) * 128 + (e & 15) * 8;           \
            rb[ST][0][u] = *reinterpret_cast<const u32x4 *>(Bh + off);                   \
            rb[ST][1][u] = *reinterpret_cast<const u32x4 *>(Bl + off);                   \
        }                                                                                \
    } while (0)
#define ASD_SPLIT_SSTORE(BUF, ST)                                                        \
    do {                                                                                 \
        _Pragma("unroll") for (int u = 0; u < PA; u++) {                                 \
            const int e = tid + NT * u, row = 16 * (e >> 6) + (e & 15), qq = (e >> 4) & 3; \
            *reinterpret_cast<u32x4 *>(&sA[BUF][0][row][qq * 8]) = ra[ST][0][u];         \
            *reinterpret_cast<u32x4 *>(&sA[BUF][1][row][qq * 8]) = ra[ST][1][u];         \
        }                                                                                \
        _Pragma("unroll") for (int u = 0; u < PB; u++) {                                 \
            const int e = tid + NT * u, row = 16 * (e >> 6) + (e & 15), qq = (e >> 4) & 3; \
            *reinterpret_cast<u32x4 *>(&sB[BUF][0][row][qq * 8]) = rb[ST][0][u];         \
            *reinterpret_cast<u32x4 *>(&sB[BUF][1][row][qq * 8]) = rb[ST][1][u];         \
        }                                                                                \
    } while (0)

    // Kpad is a multiple of 2 BK (the host pads with zeros), so nk is even and the
    // body below has no branch: the prefetch of the last two steps re-reads chunk
    // nk - 1 (in bounds, never stored to a buffer that is read again).  Loads that
    // are always issued let the compiler count outstanding loads exactly and wait
    // only for the ones each LDS store needs.
    const int nk = args.Kpad / BK;
    ASD_SPLIT_GLOAD(0, 0);
    ASD_SPLIT_GLOAD(1, 1);
    ASD_SPLIT_SSTORE(0, 0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt += 2) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int k = kt + h;
            {
                ASD_SPLIT_GLOAD(min(k + 2, nk - 1), h);
                // keep the prefetch ahead of this step's MFMAs (the scheduler would
                // otherwise sink it below the LDS stores to reuse their registers)
                __builtin_amdgcn_sched_barrier(0);
                // fragment reads and MFMAs at raised wave priority, the LDS stores and the
                // barrier at the normal one (8 waves of 32 x 64 at K = 1200: 233 -> 213 ms per
                // 100 lags; no change with 4 waves, one per SIMD)
                __builtin_amdgcn_s_setprio(1);
#pragma unroll
                for (int kk = 0; kk < BK / 16; kk++) {
                    const int kc = kk * 16 + 8 * (lane / 32);
                    v_t ah[WM], al[WM], bh[WN], bl[WN];
#pragma unroll
                    for (int m = 0; m < WM; m++) {
                        const int r = wi * WM * 32 + m * 32 + lane % 32;
                        ah[m] = *reinterpret_cast<const v_t *>(&sA[h][0][r][kc]);
                        al[m] = *reinterpret_cast<const v_t *>(&sA[h][1][r][kc]);
                    }
#pragma unroll
                    for (int n = 0; n < WN; n++) {
                        const int r = wj * WN * 32 + n * 32 + lane % 32;
                        bh[n] = *reinterpret_cast<const v_t *>(&sB[h][0][r][kc]);
                        bl[n] = *reinterpret_cast<const v_t *>(&sB[h][1][r][kc]);
                    }
#pragma unroll
                    for (int m = 0; m < WM; m++)
#pragma unroll
                        for (int n = 0; n < WN; n++) {
                            acc[m][n] = SplitMfma<PREC>::mfma(ah[m], bh[n], acc[m][n]);
                            if constexpr (DUAL) {
                                acc2[m][n] = SplitMfma<PREC>::mfma(ah[m], bl[n], acc2[m][n]);
                                acc2[m][n] = SplitMfma<PREC>::mfma(al[m], bh[n], acc2[m][n]);
                            } else {
                                acc[m][n] = SplitMfma<PREC>::mfma(ah[m], bl[n], acc[m][n]);
                                acc[m][n] = SplitMfma<PREC>::mfma(al[m], bh[n], acc[m][n]);
                            }
                        }
                }
                __builtin_amdgcn_s_setprio(0);
                ASD_SPLIT_SSTORE(h ^ 1, h ^ 1);
                __syncthreads();
            }
        }
    }
#undef ASD_SPLIT_GLOAD
#undef ASD_SPLIT_SSTORE

    bin_split_tile<WM, WN, DUAL>(acc, acc2, reinterpret_cast<unsigned long long *>(&sA[0][0][0][0]), i0, j0,
                                 wi * WM * 32, wj * WN * 32, BM, BN, args.Ns, args.L,
                                 args.bins + (long)w.lag * args.Ns, args.fix->ms);
    }
}

// The split-precision GEMM with operand tiles moved by direct global -> LDS loads
// (buffer_load_dwordx4 ... lds: 64 lanes x 16 B into 1 KiB of LDS at a wave-uniform base)
// through a ring of NS stages, so no staging registers and no LDS stores are needed
// and every chunk is requested NS - 1 steps before it is used.  A stage holds
// [A hi][A lo][B hi][B lo] in 1 KiB chunks of 16 rows x 32 halves, each stored as the
// operand is in HBM: k-block (16-byte quarter) q of row r at (16 q + r) 16 B.  A chunk's
// load then reads four contiguous 256-byte runs, and a fragment read (32 rows at one
// quarter) covers 512 contiguous bytes, every LDS bank evenly.
//
// Synchronisation per k step (one barrier): wait for this wave's loads of the step's
// stage (the loads of the NS - 2 later stages may stay in flight) and for its LDS
// reads of the previous step, barrier, then issue the stage that reuses the previous
// step's buffer, then compute.  The barrier is the bare s_barrier: the workgroup fence
// of __syncthreads would wait for every load in flight.  The loads are always issued
// (the last ones re-read chunk nk - 1) so the counts waited for are exact.
template <int PREC, int WM, int WN, int WVM, int WVN, int NS>
__global__ __launch_bounds__(64 * WVM * WVN) void sqt_gemm_dma_kernel(GemmArgs args)
{
    constexpr int BK = 32;                      // halves per row of a stage
    constexpr int NW = WVM * WVN;
    constexpr int BM = WVM * WM * 32, BN = WVN * WN * 32;
    constexpr int CA = BM / 16 / NW, CB = BN / 16 / NW;  // 1 KiB chunks per wave per array
    static_assert(CA * 16 * NW == BM && CB * 16 * NW == BN, "tile rows must split evenly over waves");
    constexpr int PER_STAGE = 2 * (CA + CB);    // loads per wave per stage
    static_assert(PER_STAGE * (NS - 2) <= 63, "vmcnt range");
    constexpr int OFF_AL = BM * BK, OFF_BH = 2 * BM * BK, OFF_BL = 2 * BM * BK + BN * BK;
    constexpr int STAGE = 2 * (BM + BN) * BK;   // halves
    typedef typename SplitMfma<PREC>::v_t v_t;
    __shared__ __attribute__((aligned(1024))) uint16_t lds[NS * STAGE];
    // the epilogue's unwrapped bins (2 (BM + BN + 3 L) words at most, L <= 1024) reuse the ring
    static_assert(NS * STAGE * 2 >= 2 * (BM + BN + 3 * 1024) * 8, "ring too small for the bins");

    const int nti = (args.Ns + BM - 1) / BM, ntj = (args.Ns + BN - 1) / BN;
    const long nslots = num_slots((long)nti * ntj, args.nlags);
    const int tid = threadIdx.x, lane = tid & 63;
    // wave-uniform by construction; readfirstlane lets the compiler keep it (and the
    // LDS bases of the loads) in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wi = wave / WVN, wj = wave % WVN;
    // load side: lane l of a chunk fills (quarter l / 16, row l % 16); read side: fragment
    // lane l reads row l % 32 of a 32-row block at quarter 2 kk + l / 32, at
    // (rows / 16) 512 + (16 quarter + row % 16) 8 halves from the block
    const int rrow = lane & 31, rhalf = lane >> 5;
    const int rd_off = (rrow >> 4) * 512 + (16 * rhalf + (rrow & 15)) * 8;

    for (long slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    Work w;
    if (!decode_work(slot, nti, ntj, args.nlags, args.group, w)) continue;
    __syncthreads();  // LDS (bins) of the previous work item fully consumed
    const int i0 = w.ti * BM, j0 = w.tj * BN;
    const long ld = args.ld;  // k-blocks per row group
    const long koff = 3L * w.lag, shift = koff & 7;
    const long acopy = shift * args.copy_stride + ((i0 / 16) * ld + (koff - shift) / 8) * 128;
    const uint16_t *Ah = static_cast<const uint16_t *>(args.A) + acopy;
    const uint16_t *Al = static_cast<const uint16_t *>(args.A2) + acopy;
    const uint16_t *Bh = static_cast<const uint16_t *>(args.B) + (j0 / 16) * ld * 128;
    const uint16_t *Bl = static_cast<const uint16_t *>(args.B2) + (j0 / 16) * ld * 128;
    // The loads are buffer loads into LDS (MUBUF), not global ones: the compiler then counts
    // them in vmcnt only, and can wait for single LDS fragment reads (behind a pending
    // global_load_lds it waits for every LDS read in flight).  Per k step, one resource per
    // operand array based at the step's first k-block of the tile's first row group (64-bit
    // address arithmetic, scalar); a chunk's source is that base + the chunk's row-group
    // offset (SGPR, below 16 ld 256 B) + 16 B per lane: one contiguous KiB, lane l filling
    // (quarter l / 16, row l % 16).
    const int lane_off = 16 * lane;
    const long step_bytes = 4 * 256;  // one 32-deep step: four k-blocks of a row group

    auto issue = [&](int kt, int st) {
        const long kb = (long)kt * step_bytes;
        auto rsrc = [&](const uint16_t *p) {
            return __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)p + kb), 0, 0x7fffffff, 0x00020000);
        };
        const __amdgpu_buffer_rsrc_t rAh = rsrc(Ah), rAl = rsrc(Al), rBh = rsrc(Bh), rBl = rsrc(Bl);
        uint16_t *base = lds + st * STAGE;
#pragma unroll
        for (int u = 0; u < CA; u++) {
            const int c = (wave * CA + u) * 512;
            const int so = (int)((wave * CA + u) * ld * 256);  // the chunk's row group
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rAh, (__attribute__((address_space(3))) void *)(base + c), 16,
                                                     lane_off, so, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rAl, (__attribute__((address_space(3))) void *)(base + OFF_AL + c),
                                                     16, lane_off, so, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < CB; u++) {
            const int c = (wave * CB + u) * 512;
            const int so = (int)((wave * CB + u) * ld * 256);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rBh, (__attribute__((address_space(3))) void *)(base + OFF_BH + c),
                                                     16, lane_off, so, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rBl, (__attribute__((address_space(3))) void *)(base + OFF_BL + c),
                                                     16, lane_off, so, 0, 0);
        }
    };

    constexpr bool DUAL = SplitMfma<PREC>::dual;
    f16v acc[WM][WN], acc2[DUAL ? WM : 1][DUAL ? WN : 1];
#pragma unroll
    for (int a = 0; a < WM; a++)
#pragma unroll
        for (int b = 0; b < WN; b++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[a][b][e] = 0.0f;
    if constexpr (DUAL) {
#pragma unroll
        for (int a = 0; a < WM; a++)
#pragma unroll
            for (int b = 0; b < WN; b++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc2[a][b][e] = 0.0f;
    }

    // k steps over the unpadded K: where the last step's second 16-deep half lies wholly
    // in the zero padding (K % 32 in 1..16), that step does its first half only.  The
    // products skipped are exact zeros, so the sums are unchanged.
    const int nk = (args.K + BK - 1) / BK;
    const bool half_tail = args.K % BK != 0 && args.K % BK <= 16;
    const int nk_full = half_tail ? nk - 1 : nk;

    // one k step in ring slot h: wait for this wave's loads of the slot (the loads of the
    // later stages in flight, up to NS - 2, may stay) and for its LDS reads of the previous
    // step, barrier, refill the slot the previous step used with chunk k + NS - 1, then the
    // MFMAs of KH halves.  The last NS - 2 steps have no chunk left to request and issue
    // nothing: they wait for exactly the stages still in flight (vmcnt(0) in the last step), so
    // the tile ends with no load outstanding instead of draining re-reads of its last chunk.
    auto step = [&](int k, int h, auto khc) {
        constexpr int KH = decltype(khc)::value;
        static_assert(NS - 2 <= 2, "the tail waits below cover up to two stages in flight");
        const int ahead = min(NS - 2, nk - 1 - k);  // later stages in flight
        if (ahead >= NS - 2)
            __builtin_amdgcn_s_waitcnt(wait_vmcnt(PER_STAGE * (NS - 2)));
        else if (ahead == 1)
            __builtin_amdgcn_s_waitcnt(wait_vmcnt(PER_STAGE));
        else
            __builtin_amdgcn_s_waitcnt(wait_vmcnt(0));
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (k + NS - 1 < nk) issue(k + NS - 1, (h + NS - 1) % NS);
        // fragment reads and MFMAs at raised wave priority, the waits, barrier and load
        // requests at the normal one (as in the k-major kernel; FP16X2 at K = 1200:
        // 320.5 -> 310.5 ms per 200 lags)
        __builtin_amdgcn_s_setprio(1);
        const uint16_t *sb = lds + h * STAGE;
        const uint16_t *sr = sb + rd_off;
        if constexpr (PREC == ASD_FP16F8) {
            // the hi halves of the step's 16-deep halves (the hh products on the fp16 MFMA), then
            // the cross bytes: lane l takes k-blocks l / 32 and 2 + l / 32 of its row, 32 bytes,
            // for one fp8 MFMA over the whole 32-deep stage.  In a half step (KH = 1) the second
            // half's cross bytes are read too: B is zero there, so those products vanish.
            v_t ah[KH][WM], bh[KH][WN];
            i32x4 ax[2][WM], bx[2][WN];
#pragma unroll
            for (int kk = 0; kk < KH; kk++) {
                const int qo = 2 * kk * 16 * 8;
#pragma unroll
                for (int m = 0; m < WM; m++)
                    ah[kk][m] = *reinterpret_cast<const v_t *>(sr + (wi * WM + m) * 1024 + qo);
#pragma unroll
                for (int n = 0; n < WN; n++)
                    bh[kk][n] = *reinterpret_cast<const v_t *>(sr + OFF_BH + (wj * WN + n) * 1024 + qo);
            }
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const int qo = 2 * kk * 16 * 8;
#pragma unroll
                for (int n = 0; n < WN; n++)
                    bx[kk][n] = *reinterpret_cast<const i32x4 *>(sr + OFF_BL + (wj * WN + n) * 1024 + qo);
#pragma unroll
                for (int m = 0; m < WM; m++)
                    ax[kk][m] = *reinterpret_cast<const i32x4 *>(sr + OFF_AL + (wi * WM + m) * 1024 + qo);
            }
#pragma unroll
            for (int kk = 0; kk < KH; kk++)
#pragma unroll
                for (int m = 0; m < WM; m++)
#pragma unroll
                    for (int n = 0; n < WN; n++)
                        acc[m][n] = SplitMfma<PREC>::mfma(ah[kk][m], bh[kk][n], acc[m][n]);
#pragma unroll
            for (int m = 0; m < WM; m++)
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    const v8i a8 = __builtin_shufflevector(ax[0][m], ax[1][m], 0, 1, 2, 3, 4, 5, 6, 7);
                    const v8i b8 = __builtin_shufflevector(bx[0][n], bx[1][n], 0, 1, 2, 3, 4, 5, 6, 7);
                    acc2[m][n] = SplitMfma<PREC>::mfma8(a8, b8, acc2[m][n]);
                }
            {
                constexpr int R0 = WM + WN, RT = (KH + 2) * (WM + WN), MT = (KH + 1) * WM * WN;
                constexpr int NI = MT < RT - R0 ? MT : RT - R0;
                __builtin_amdgcn_sched_group_barrier(0x100, R0, 0);
#pragma unroll
                for (int i = 0; i < NI; i++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
                if constexpr (RT - R0 > NI) __builtin_amdgcn_sched_group_barrier(0x100, RT - R0 - NI, 0);
                if constexpr (MT > NI) __builtin_amdgcn_sched_group_barrier(0x008, MT - NI, 0);
            }
        } else {
        // reads in the order the MFMAs use them: per 16-deep half A hi, B hi (the hh
        // products), B lo (hl), A lo (lh)
        v_t ah[KH][WM], al[KH][WM], bh[KH][WN], bl[KH][WN];
#pragma unroll
        for (int kk = 0; kk < KH; kk++) {
            const int qo = 2 * kk * 16 * 8;  // quarter 2 kk (+ l / 32 in rd_off)
#pragma unroll
            for (int m = 0; m < WM; m++)
                ah[kk][m] = *reinterpret_cast<const v_t *>(sr + (wi * WM + m) * 1024 + qo);
#pragma unroll
            for (int n = 0; n < WN; n++)
                bh[kk][n] = *reinterpret_cast<const v_t *>(sr + OFF_BH + (wj * WN + n) * 1024 + qo);
#pragma unroll
            for (int n = 0; n < WN; n++)
                bl[kk][n] = *reinterpret_cast<const v_t *>(sr + OFF_BL + (wj * WN + n) * 1024 + qo);
#pragma unroll
            for (int m = 0; m < WM; m++)
                al[kk][m] = *reinterpret_cast<const v_t *>(sr + OFF_AL + (wi * WM + m) * 1024 + qo);
        }
        // per half: every hh product, then every hl, then every lh (each accumulator still
        // takes hh, hl, lh of a half in that order), so the first MFMAs wait only for the
        // first A hi and B hi reads
#pragma unroll
        for (int kk = 0; kk < KH; kk++) {
#pragma unroll
            for (int m = 0; m < WM; m++)
#pragma unroll
                for (int n = 0; n < WN; n++) acc[m][n] = SplitMfma<PREC>::mfma(ah[kk][m], bh[kk][n], acc[m][n]);
#pragma unroll
            for (int m = 0; m < WM; m++)
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    if constexpr (DUAL)
                        acc2[m][n] = SplitMfma<PREC>::mfma(ah[kk][m], bl[kk][n], acc2[m][n]);
                    else
                        acc[m][n] = SplitMfma<PREC>::mfma(ah[kk][m], bl[kk][n], acc[m][n]);
                }
#pragma unroll
            for (int m = 0; m < WM; m++)
#pragma unroll
                for (int n = 0; n < WN; n++) {
                    if constexpr (DUAL)
                        acc2[m][n] = SplitMfma<PREC>::mfma(al[kk][m], bh[kk][n], acc2[m][n]);
                    else
                        acc[m][n] = SplitMfma<PREC>::mfma(al[kk][m], bh[kk][n], acc[m][n]);
                }
        }
        // issue order: the first WM + WN reads (A hi, B hi), then one MFMA and one read in
        // turn until every read is issued, then the remaining MFMAs; the LDS counter waits
        // for single reads (the loads into LDS are buffer loads)
        {
            constexpr int R0 = WM + WN, RT = 2 * KH * (WM + WN), MT = 3 * KH * WM * WN;
            static_assert(MT >= RT - R0, "interleave needs an MFMA per read");
            __builtin_amdgcn_sched_group_barrier(0x100, R0, 0);
#pragma unroll
            for (int i = 0; i < RT - R0; i++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
            if constexpr (MT > RT - R0) __builtin_amdgcn_sched_group_barrier(0x008, MT - (RT - R0), 0);
        }
        }
        __builtin_amdgcn_s_setprio(0);
        asm volatile("" ::: "memory");
    };

#pragma unroll
    for (int st = 0; st < NS - 1; st++) issue(min(st, nk - 1), st);
    for (int kt = 0; kt < nk_full; kt += NS) {
#pragma unroll
        for (int h = 0; h < NS; h++)
            if (kt + h < nk_full) step(kt + h, h, std::integral_constant<int, BK / 16>{});
    }
    if (half_tail) step(nk - 1, (nk - 1) % NS, std::integral_constant<int, 1>{});
    // every load in flight lands before the bins reuse the LDS
    __builtin_amdgcn_s_waitcnt(wait_vmcnt(0));
    __syncthreads();
    bin_split_tile_unwrapped<WM, WN, DUAL>(acc, acc2, reinterpret_cast<unsigned long long *>(lds), i0, j0,
                                           wi * WM * 32, wj * WN * 32, BM, BN, args.Ns, args.L,
                                           args.bins + (long)w.lag * args.Ns, args.fix->ms);
    }
}

// ------------------------------------------------------------- configurations
// Tile configurations the autotuner chooses from (per precision).
struct SqtCfg {
    int wm, wn, group;
    const char *name;
    int wvm = 2, wvn = 2;  // waves per workgroup along M and N
    int dma = 0;           // stages of the direct-to-LDS ring (0: register-staged loads)
};
static const SqtCfg kSplitCfgs[] = {
    {2, 2, 8, "128x128 tiles, 4 waves of 64x64, group 8"},
    {1, 2, 8, "128x128 tiles, 8 waves of 32x64, group 8", 4, 2},
    {2, 1, 8, "128x128 tiles, 8 waves of 64x32, group 8", 2, 4},
    {1, 2, 8, "128x128 tiles, 8 waves of 32x64, 4-stage LDS ring, group 8", 4, 2, 4},
    {2, 1, 8, "128x128 tiles, 8 waves of 64x32, 4-stage LDS ring, group 8", 2, 4, 4},
    // 32 tiles run at once per XCD: 4 tile-rows x 8 tile-columns share 4 A panels of 256
    // rows and 8 B panels of 128 rows (256 KB per k step, 83% L2 hits at best); group 8
    // (8 x 4) needs 320 KB (79%; 78% measured)
    {2, 2, 4, "256x128 tiles, 8 waves of 64x64, 3-stage LDS ring, group 4", 4, 2, 3},
    {2, 2, 16, "128x128 tiles, 4 waves of 64x64, group 16"},
    {2, 4, 8, "128x256 tiles, 4 waves of 64x128, group 8"},
    {4, 2, 8, "256x128 tiles, 4 waves of 128x64, group 8"},
    // fewer tile-rows per group: FP16F8 at K = 1200, 121-124 ms per 100 lags against 124-129 at group 4
    {2, 2, 2, "256x128 tiles, 8 waves of 64x64, 3-stage LDS ring, group 2", 4, 2, 3},
    {2, 2, 1, "256x128 tiles, 8 waves of 64x64, 3-stage LDS ring, group 1", 4, 2, 3},
};
constexpr int kNumSplit = sizeof(kSplitCfgs) / sizeof(kSplitCfgs[0]);
// The rows each split precision is offered, in its own configuration order.  BF16X3: all.
// FP16X2 keeps two accumulators, so the 64x128 / 128x64 register-staged wave tiles (rows 6-8)
// do not fit the register file.  FP16F8 runs on the direct-to-LDS kernels only.
static const int kFp16Rows[] = {0, 1, 2, 3, 4, 5, 9, 10};
static const int kF8Rows[] = {5, 3, 4, 9, 10};
constexpr int kNumFp16 = sizeof(kFp16Rows) / sizeof(kFp16Rows[0]);
constexpr int kNumF8 = sizeof(kF8Rows) / sizeof(kF8Rows[0]);
// row of kSplitCfgs for configuration cfg of a split precision
static int split_row(asd_precision prec, int cfg)
{
    return prec == ASD_FP16X2 ? kFp16Rows[cfg] : prec == ASD_FP16F8 ? kF8Rows[cfg] : cfg;
}
static const SqtCfg kKmajorCfgs[] = {
    {2, 2, 8, "128x128 tiles, 64x64 per wave, group 8"},
    {2, 2, 16, "128x128 tiles, 64x64 per wave, group 16"},
    {2, 2, 4, "128x128 tiles, 64x64 per wave, group 4"},
    {2, 2, 8, "128x128 tiles, 64x64 per wave, 4-stage direct-to-LDS ring, group 8", 2, 2, 4},
    {2, 2, 16, "128x128 tiles, 64x64 per wave, 4-stage direct-to-LDS ring, group 16", 2, 2, 4},
};
constexpr int kNumKmajor = sizeof(kKmajorCfgs) / sizeof(kKmajorCfgs[0]);

// the configuration an untuned (L, K) takes: the one the autotuner picks on most lattices (the
// direct-to-LDS rings)
static_assert(kNumKmajor > 3 && kNumFp16 > 5 && kNumSplit > 5 && kNumF8 > 3, "the default configurations exist for every precision");
static int default_config(asd_precision prec)
{
    return prec <= ASD_FP32 ? 3 : prec == ASD_FP16F8 ? 3 : 5;
}

int sqt_num_configs(asd_precision prec)
{
    return prec <= ASD_FP32 ? kNumKmajor : prec == ASD_FP16X2 ? kNumFp16 : prec == ASD_FP16F8 ? kNumF8 : kNumSplit;
}
const char *sqt_config_name(asd_precision prec, int cfg)
{
    if (cfg < 0 || cfg >= sqt_num_configs(prec)) return "invalid";
    return prec <= ASD_FP32 ? kKmajorCfgs[cfg].name : kSplitCfgs[split_row(prec, cfg)].name;
}

using GemmKernel = void (*)(GemmArgs);

// persistent grid: the blocks of `k` resident at once on the device (a multiple of 8,
// so that slot s stays on XCD s % 8), at most one per slot
static int resident_grid(GemmKernel k, int threads, long nslots, int device)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(device, reinterpret_cast<const void *>(k));
    auto it = cache.find(key);
    if (it == cache.end()) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, 0) != hipSuccess) per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 8;
        const long n = (long)std::max(per_cu, 1) * std::max(cus, 8) / 8 * 8;
        it = cache.emplace(key, (int)std::min(n, kMaxGrid)).first;
    }
    return (int)std::min<long>(it->second, nslots);
}

template <int PREC>
static void launch_split(const SqtCfg &c, GemmArgs ga, hipStream_t st, int device)
{
    const int BM = 32 * c.wvm * c.wm, BN = 32 * c.wvn * c.wn;
    const long per_lag = (long)((ga.Ns + BM - 1) / BM) * ((ga.Ns + BN - 1) / BN);
    const int threads = 64 * c.wvm * c.wvn;
    ga.group = c.group;
    GemmKernel k;
    if constexpr (PREC == ASD_FP16F8) {  // the direct-to-LDS kernels only
        if (c.dma == 4 && c.wvm == 4)
            k = sqt_gemm_dma_kernel<PREC, 1, 2, 4, 2, 4>;
        else if (c.dma == 4)
            k = sqt_gemm_dma_kernel<PREC, 2, 1, 2, 4, 4>;
        else
            k = sqt_gemm_dma_kernel<PREC, 2, 2, 4, 2, 3>;
    } else if (c.dma == 4 && c.wvm == 4)
        k = sqt_gemm_dma_kernel<PREC, 1, 2, 4, 2, 4>;
    else if (c.dma == 4)
        k = sqt_gemm_dma_kernel<PREC, 2, 1, 2, 4, 4>;
    else if (c.dma == 3)
        k = sqt_gemm_dma_kernel<PREC, 2, 2, 4, 2, 3>;
    else if (c.wvm == 4 && c.wvn == 2)
        k = sqt_gemm_split_kernel<PREC, 1, 2, 4, 2>;
    else if (c.wvm == 2 && c.wvn == 4)
        k = sqt_gemm_split_kernel<PREC, 2, 1, 2, 4>;
    else if constexpr (PREC == ASD_FP16X2)
        k = sqt_gemm_split_kernel<PREC, 2, 2>;
    else if (c.wm == 2 && c.wn == 2)
        k = sqt_gemm_split_kernel<PREC, 2, 2>;
    else if (c.wm == 2 && c.wn == 4)
        k = sqt_gemm_split_kernel<PREC, 2, 4>;
    else
        k = sqt_gemm_split_kernel<PREC, 4, 2>;
    const dim3 grid(resident_grid(k, threads, num_slots(per_lag, ga.nlags), device)), blk(threads);
    hipLaunchKernelGGL(k, grid, blk, 0, st, ga);
}

// ----------------------------------------------------------- FP64 finishing

struct KahanC {
    double re = 0, im = 0, cre = 0, cim = 0;
    __device__ __forceinline__ void add(double xr, double xi)
    {
        double y = xr - cre, t = re + y;
        cre = (t - re) - y;
        re = t;
        y = xi - cim;
        t = im + y;
        cim = (t - im) - y;
        im = t;
    }
};

// block-wide sum of NV doubles (256 threads)
template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV], double *red)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; k++) red[k * 256 + tid] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
#pragma unroll
            for (int k = 0; k < NV; k++) red[k * 256 + tid] += red[k * 256 + tid + s];
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = red[k * 256];
    __syncthreads();
}

// F[t][q][c] = sum_i S_i^c(t) e^{-i q.r_i}, (re, im); grid (nq, nt), 256 threads.
// phase index (mx row + my col) mod L into the exact table tw[k] = (cos, sin)(2 pi k / L).
__global__ __launch_bounds__(256) void fourier_kernel(const double *__restrict__ snaps, long tstride,
                                                      int L, const int *__restrict__ q,
                                                      const double2 *__restrict__ tw,
                                                      double *__restrict__ F, int nq)
{
    __shared__ double red[6 * 256];
    const int iq = blockIdx.x, t = blockIdx.y;
    const int mx = q[2 * iq], my = q[2 * iq + 1];
    const int Ns = L * L;
    const double *S = snaps + (long)t * tstride;
    KahanC acc[3];
    for (int i = threadIdx.x; i < Ns; i += 256) {
        const int r = i / L, c = i - r * L;
        const int k = (int)(((long)mx * r + (long)my * c) % L);
        const double2 e = tw[k];  // e^{-i phi} = cos - i sin
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double s = S[3L * i + d];
            acc[d].add(s * e.x, -s * e.y);
        }
    }
    double v[6] = {acc[0].re, acc[0].im, acc[1].re, acc[1].im, acc[2].re, acc[2].im};
    block_sum<6>(v, red);
    if (threadIdx.x < 6) F[((long)t * nq + iq) * 6 + threadIdx.x] = v[threadIdx.x];
}

// out[lag][q] = ( (1/T0) sum_d g_lag(d) e^{-iq.d}  -  Fbar_q(lag).conj Fbar_q(0) ) / Ns
__global__ __launch_bounds__(256) void finish_kernel(const double *__restrict__ bins, int L,
                                                     const int *__restrict__ q,
                                                     const double2 *__restrict__ tw,
                                                     const double *__restrict__ F,
                                                     const double *__restrict__ Fz, int nq, int T0,
                                                     double *__restrict__ out)
{
    __shared__ double red[12 * 256];
    const int iq = blockIdx.x, lag = blockIdx.y;
    const int mx = q[2 * iq], my = q[2 * iq + 1];
    const int Ns = L * L;
    const double *g = bins + (long)lag * Ns;
    KahanC s;
    for (int b = threadIdx.x; b < Ns; b += 256) {
        const int dr = b / L, dc = b - dr * L;
        const int k = (int)(((long)mx * dr + (long)my * dc) % L);
        const double2 e = tw[k];
        const double v = g[b];
        s.add(v * e.x, -v * e.y);
    }
    // Fbar(lag) and Fbar(0) = (1/T0) sum_t0 F(t0 + lag), F(t0)
    KahanC fb[3], zb[3];
    for (int t0 = threadIdx.x; t0 < T0; t0 += 256) {
        const double *a = F + ((long)(t0 + lag) * nq + iq) * 6;
        const double *z = Fz + ((long)t0 * nq + iq) * 6;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            fb[d].add(a[2 * d], a[2 * d + 1]);
            zb[d].add(z[2 * d], z[2 * d + 1]);
        }
    }
    double v[14] = {s.re, s.im};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v[2 + 2 * d] = fb[d].re; v[3 + 2 * d] = fb[d].im;
        v[8 + 2 * d] = zb[d].re; v[9 + 2 * d] = zb[d].im;
    }
    // two passes of 7 to stay within the reduction buffer
    double v0[7], v1[7];
#pragma unroll
    for (int k = 0; k < 7; k++) { v0[k] = v[k]; v1[k] = v[7 + k]; }
    block_sum<7>(v0, red);
    block_sum<7>(v1, red);
    if (threadIdx.x == 0) {
        const double invT = 1.0 / (double)T0;
        double dre = 0, dim = 0;
        const double f[6] = {v0[2], v0[3], v0[4], v0[5], v0[6], v1[0]};
        const double z[6] = {v1[1], v1[2], v1[3], v1[4], v1[5], v1[6]};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double ar = f[2 * d] * invT, ai = f[2 * d + 1] * invT;
            const double br = z[2 * d] * invT, bi = z[2 * d + 1] * invT;
            dre += ar * br + ai * bi;   // a * conj(b)
            dim += ai * br - ar * bi;
        }
        out[((long)lag * nq + iq) * 2] = (v0[0] * invT - dre) / (double)Ns;
        out[((long)lag * nq + iq) * 2 + 1] = (v0[1] * invT - dim) / (double)Ns;
    }
}

// deterministic unit-length pseudo-spins for timing runs (values do not matter)
__global__ void fill_pseudo_spins_kernel(double *p, long nspins)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nspins; i += (long)gridDim.x * blockDim.x) {
        const double a = 0.7548776662 * (double)i, z = 2.0 * (0.5698402910 * (double)i - floor(0.5698402910 * (double)i)) - 1.0;
        const double r = sqrt(fmax(0.0, 1.0 - z * z));
        p[3 * i] = r * cos(a);
        p[3 * i + 1] = r * sin(a);
        p[3 * i + 2] = z;
    }
}

// ------------------------------------------------ full momentum grid (separable)
// out[b][k][j] = sum_{x<L} in[b][x][j] e^{-2 pi i k x / L}, j < inner; `in` real
// (in_complex = 0) or complex (re, im interleaved); Kahan-compensated FP64.
// grid (ceil(L*inner/256), nbatch)
__global__ __launch_bounds__(256) void dft_strided_kernel(const double *__restrict__ in, int in_complex, long b0,
                                                          int L, int inner, const double2 *__restrict__ tw,
                                                          double *__restrict__ out)
{
    const long b = b0 + blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)L * inner) return;
    const int k = (int)(e / inner), j = (int)(e - (long)k * inner);
    const long base = b * (long)L * inner;
    KahanC s;
    for (int x = 0; x < L; x++) {
        const double2 w = tw[(int)(((long)k * x) % L)];  // e^{-i phi} = (w.x, -w.y)
        const long at = base + (long)x * inner + j;
        if (in_complex) {
            const double re = in[2 * at], im = in[2 * at + 1];
            s.add(re * w.x + im * w.y, im * w.x - re * w.y);
        } else {
            const double v = in[at];
            s.add(v * w.x, -v * w.y);
        }
    }
    out[2 * (base + e)] = s.re;
    out[2 * (base + e) + 1] = s.im;
}

// fbar[tau][i][c] = (1/T0) sum_{t0<T0} S_i^c(t0 + tau) as a Kahan-compensated sliding sum
__global__ void window_mean_kernel(const double *__restrict__ snaps, long tstride, long n, int T0, int nlags,
                                   double *__restrict__ fbar)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = 0.0, c = 0.0;
    auto add = [&](double x) {
        const double y = x - c, t = s + y;
        c = (t - s) - y;
        s = t;
    };
    for (int t0 = 0; t0 < T0; t0++) add(snaps[(long)t0 * tstride + e]);
    fbar[e] = s / (double)T0;
    for (int tau = 1; tau < nlags; tau++) {
        add(snaps[(long)(tau + T0 - 1) * tstride + e]);
        add(-snaps[(long)(tau - 1) * tstride + e]);
        fbar[(long)tau * n + e] = s / (double)T0;
    }
}

// out[lag][iq] = ( Sc[lag][q]/T0 - sum_c Fbar_c(lag) conj(Zbar_c) ) / Ns, q = (mx, my)
// Sc: [lag][mx][my] complex; Fb: [lag][mx][my][c] complex (DFT of the window means);
// Fz: the B side's lag-0 window-mean DFT (= Fb for the time-origin form)
__global__ void finish_full_kernel(const double *__restrict__ Sc, const double *__restrict__ Fb,
                                     const double *__restrict__ Fz, int L, const int *__restrict__ q, int nq,
                                     int nlags, int T0, double *__restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nq * nlags) return;
    const int lag = (int)(e / nq), iq = (int)(e - (long)lag * nq);
    const long Ns = (long)L * L;
    const long m = (long)q[2 * iq] * L + q[2 * iq + 1];
    const double *sc = Sc + 2 * ((long)lag * Ns + m);
    const double *f = Fb + 6 * ((long)lag * Ns + m), *z = Fz + 6 * m;
    double dre = 0, dim = 0;
    for (int c = 0; c < 3; c++) {
        dre += f[2 * c] * z[2 * c] + f[2 * c + 1] * z[2 * c + 1];
        dim += f[2 * c + 1] * z[2 * c] - f[2 * c] * z[2 * c + 1];
    }
    const double invT = 1.0 / (double)T0;
    out[2 * e] = (sc[0] * invT - dre) / (double)Ns;
    out[2 * e + 1] = (sc[1] * invT - dim) / (double)Ns;
}

__global__ void zero_kernel(double *p, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// --------------------------------------------------------------- host side

static long round_up(long a, long b) { return (a + b - 1) / b * b; }


// the exact twiddle table e^{2 pi i k / L} and the q list on the device (ctx->twiddle, ctx->qdev)
static int load_phases(asd_ctx *ctx, int L, const int *h_q, int nq)
{
    hipStream_t st = ctx->stream;
    ASD_TRY(ctx->twiddle.ensure(sizeof(double2) * L));
    ASD_TRY(ctx->qdev.ensure(sizeof(int) * 2 * nq));
    if (ctx->phases_L == L && ctx->phases_tw == ctx->twiddle.ptr && ctx->phases_qd == ctx->qdev.ptr &&
        ctx->phases_q.size() == (size_t)(2 * nq) && std::equal(h_q, h_q + 2 * nq, ctx->phases_q.begin()))
        return ASD_OK;  // already on the device
    std::vector<double2> tw(L);
    for (int k = 0; k < L; k++) {
        // exact phases: angle 2 pi k / L evaluated in long double
        const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
        tw[k] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    ASD_HIP(hipMemcpyAsync(ctx->twiddle.ptr, tw.data(), sizeof(double2) * L, hipMemcpyHostToDevice, st));
    ASD_HIP(hipMemcpyAsync(ctx->qdev.ptr, h_q, sizeof(int) * 2 * nq, hipMemcpyHostToDevice, st));
    ASD_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
    ctx->phases_L = L;
    ctx->phases_q.assign(h_q, h_q + 2 * nq);
    ctx->phases_tw = ctx->twiddle.ptr;
    ctx->phases_qd = ctx->qdev.ptr;
    return ASD_OK;
}

int sqt_fourier(asd_ctx *ctx, int L, const double *d_snaps, long tstride, int nt, const int *h_q, int nq)
{
    ASD_TRY(load_phases(ctx, L, h_q, nq));
    ASD_TRY(ctx->fq.ensure(sizeof(double) * 6 * (long)nt * nq));
    hipLaunchKernelGGL(fourier_kernel, dim3(nq, nt), dim3(256), 0, ctx->stream, d_snaps, tstride, L,
                       ctx->qdev.as<int>(), ctx->twiddle.as<double2>(), ctx->fq.as<double>(), nq);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

int sqt_run(asd_ctx *ctx, int L, const double *d_snaps_a, const double *d_snaps_b, long tstride,
            int nsnap_a, int T0, int nlags, const int *h_q, int nq, asd_precision prec, double *d_out,
            int cfg)
{
    if (cfg < 0) {
        auto it = ctx->tuned[prec].find(std::make_pair(L, 3 * T0));
        cfg = it != ctx->tuned[prec].end() ? it->second : default_config(prec);
    }
    ASD_CHECK(cfg < sqt_num_configs(prec), "configuration %d out of range", cfg);
    const int Ns = L * L;
    const int K = 3 * T0;
    const int nt_tiles = (Ns + TILE - 1) / TILE;
    const long ldi = round_up(Ns, 2 * TILE);  // padded site count (every tile shape divides it)
    const int nt_a = T0 + nlags - 1;      // snapshots the A windows touch
    ASD_CHECK(nsnap_a >= nt_a, "need nsnap >= T0 + nlags - 1 (%d < %d)", nsnap_a, nt_a);
    hipStream_t st = ctx->stream;

    ASD_TRY(load_phases(ctx, L, h_q, nq));
    // bins: fixed-point words [word][lag][d] (two words for FP64), then the decoded doubles
    const long nbins = (long)nlags * Ns;
    const int nw = prec == ASD_FP64 ? 2 : 1;
    ASD_TRY(ctx->bins.ensure(sizeof(double) * (nw + 1) * nbins));
    ASD_TRY(ctx->fix.ensure(sizeof(FixScale)));
    unsigned long long *words = ctx->bins.as<unsigned long long>();
    double *gbins = ctx->bins.as<double>() + nw * nbins;
    FixScale *fix = ctx->fix.as<FixScale>();
    ASD_TRY(mark(ctx, 0));
    hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, st, ctx->bins.as<double>(), nw * nbins);
    hipLaunchKernelGGL(fix_reset_kernel, dim3(1), dim3(1), 0, st, fix);
    {
        const long n3 = 3L * Ns;
        const unsigned g = (unsigned)std::min<long>(1024, (n3 + 255) / 256);
        hipLaunchKernelGGL(vmax_kernel, dim3(g, (unsigned)std::min(nt_a, 16)), dim3(256), 0, st, d_snaps_a, tstride,
                           nt_a, n3, &fix->vmax_bits);
        if (d_snaps_b != d_snaps_a)
            hipLaunchKernelGGL(vmax_kernel, dim3(g, (unsigned)std::min(T0, 16)), dim3(256), 0, st, d_snaps_b,
                               tstride, T0, n3, &fix->vmax_bits);
        hipLaunchKernelGGL(fix_scale_kernel, dim3(1), dim3(1), 0, st, fix, (int)K, Ns);
    }

    GemmArgs ga{};
    ga.L = L;
    ga.Ns = Ns;
    ga.nt = nt_tiles;
    ga.nlags = nlags;
    ga.K = (int)K;
    ga.bins = words;
    ga.wstride = nbins;
    ga.fix = fix;
    ga.group = kKmajorCfgs[prec <= ASD_FP32 ? cfg : 0].group;

    if (prec == ASD_FP64 || prec == ASD_FP32) {
        const size_t es = prec == ASD_FP64 ? 8 : 4;
        const long Kpad = round_up(K, 16);
        const long rowsA = 3L * (nlags - 1) + Kpad;
        ASD_TRY(ctx->opA.ensure(es * rowsA * ldi));
        ASD_TRY(ctx->opB.ensure(es * Kpad * ldi));
        const dim3 pb(256);
        if (prec == ASD_FP64) {
            hipLaunchKernelGGL(pack_kmajor_kernel<double>, dim3((unsigned)(ldi / 256 + 1), (unsigned)rowsA), pb, 0,
                               st, d_snaps_a, tstride, nt_a, Ns, ctx->opA.as<double>(), ldi);
            hipLaunchKernelGGL(pack_kmajor_kernel<double>, dim3((unsigned)(ldi / 256 + 1), (unsigned)Kpad), pb, 0,
                               st, d_snaps_b, tstride, T0, Ns, ctx->opB.as<double>(), ldi);
        } else {
            hipLaunchKernelGGL(pack_kmajor_kernel<float>, dim3((unsigned)(ldi / 256 + 1), (unsigned)rowsA), pb, 0,
                               st, d_snaps_a, tstride, nt_a, Ns, ctx->opA.as<float>(), ldi);
            hipLaunchKernelGGL(pack_kmajor_kernel<float>, dim3((unsigned)(ldi / 256 + 1), (unsigned)Kpad), pb, 0,
                               st, d_snaps_b, tstride, T0, Ns, ctx->opB.as<float>(), ldi);
        }
        ga.A = ctx->opA.ptr;
        ga.B = ctx->opB.ptr;
        ga.ld = ldi;
        ga.Kpad = (int)Kpad;
        ASD_TRY(mark(ctx, 1));
        const bool dma = kKmajorCfgs[cfg].dma != 0;
        GemmKernel k = prec == ASD_FP32 ? (dma ? sqt_gemm_kmajor_kernel<float, 1> : sqt_gemm_kmajor_kernel<float>)
                                        : (dma ? sqt_gemm_kmajor_kernel<double, 1> : sqt_gemm_kmajor_kernel<double>);
        // a capped grid of many more blocks than fit at once measured 2-4% faster than the
        // persistent grid for this kernel (FP32 / FP64 at L = 256)
        const int nblocks = (int)std::min(num_slots((long)nt_tiles * nt_tiles, nlags), kMaxGrid);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(256), 0, st, ga);
    } else {
        const long Kpad = round_up(K, 64);  // an even number of 32-wide k steps
        const long ldk = round_up(3L * (nlags - 1) + Kpad + 8, 8);  // k extent (halves) of the A copies
        const long nkb = ldk / 8;                                    // k-blocks of 16 B
        ASD_CHECK(nkb <= 65535, "too many lags for the split GEMM (%d)", nlags);
        // k-blocks per row group, odd so that row groups are not a power of two apart
        const long nkbs = nkb | 1;
        const long cs = nkbs * 8 * ldi;                              // one shifted copy
        ASD_TRY(ctx->opA.ensure(2 * 8 * cs));
        ASD_TRY(ctx->opA2.ensure(2 * 8 * cs));
        ASD_TRY(ctx->opB.ensure(2 * cs));
        ASD_TRY(ctx->opB2.ensure(2 * cs));
        const dim3 pb(256);
        auto packer = prec == ASD_BF16X3   ? pack_split_kernel<ASD_BF16X3>
                      : prec == ASD_FP16X2 ? pack_split_kernel<ASD_FP16X2>
                                           : pack_split_kernel<ASD_FP16F8>;
        hipLaunchKernelGGL(packer, dim3((unsigned)(ldi / 32), (unsigned)nkb), pb, 0, st, d_snaps_a, tstride, nt_a,
                           Ns, ctx->opA.as<uint16_t>(), ctx->opA2.as<uint16_t>(), nkbs, cs, 8, 0,
                           (const FixScale *)fix);
        hipLaunchKernelGGL(packer, dim3((unsigned)(ldi / 32), (unsigned)(Kpad / 8)), pb, 0, st, d_snaps_b, tstride,
                           T0, Ns, ctx->opB.as<uint16_t>(), ctx->opB2.as<uint16_t>(), nkbs, cs, 1, 1,
                           (const FixScale *)fix);
        ga.copy_stride = cs;
        ga.A = ctx->opA.ptr;
        ga.A2 = ctx->opA2.ptr;
        ga.B = ctx->opB.ptr;
        ga.B2 = ctx->opB2.ptr;
        ga.ld = nkbs;  // k-blocks per row group
        ga.Kpad = (int)Kpad;
        ASD_TRY(mark(ctx, 1));
        if (prec == ASD_BF16X3)
            launch_split<ASD_BF16X3>(kSplitCfgs[split_row(prec, cfg)], ga, st, ctx->device);
        else if (prec == ASD_FP16X2)
            launch_split<ASD_FP16X2>(kSplitCfgs[split_row(prec, cfg)], ga, st, ctx->device);
        else
            launch_split<ASD_FP16F8>(kSplitCfgs[split_row(prec, cfg)], ga, st, ctx->device);
    }
    ASD_HIP(hipGetLastError());
    ASD_TRY(mark(ctx, 2));
    hipLaunchKernelGGL(fix_decode_kernel, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, st,
                       (const unsigned long long *)words, nbins, nw, Ns, (const FixScale *)fix, gbins);

    const bool full = nq > 2 * L && !ctx->sqt_generic_only;  // many momenta: separable DFTs over the grid
    if (full) {
        const long n3 = 3L * Ns;
        const long wm = (long)nlags * n3;            // doubles of window means
        const bool two = d_snaps_b != d_snaps_a;     // replica form: B side has its own mean
        ASD_TRY(ctx->fq.ensure(sizeof(double) * (wm + n3 + 4 * (wm + n3) + 4L * nlags * Ns)));
        double *fbar = ctx->fq.as<double>();         // [tau][r][col][c], then the B mean
        double *pt = fbar + wm + n3;                 // complex [tau][r][my][c]
        double *Fb = pt + 2 * (wm + n3);             // complex [tau][mx][my][c], then the B side
        double *Sc = Fb + 2 * (wm + n3);             // complex [lag][mx][my]
        double *Ht = Sc + 2L * nlags * Ns;           // complex [lag][dr][my]
        const double2 *tw = ctx->twiddle.as<double2>();
        auto dft = [&](const double *in, int cplx, long nbatch, int inner, double *out) {
            for (long b0 = 0; b0 < nbatch; b0 += 65535) {
                const unsigned nb = (unsigned)std::min<long>(65535, nbatch - b0);
                hipLaunchKernelGGL(dft_strided_kernel, dim3((unsigned)(((long)L * inner + 255) / 256), nb), dim3(256),
                                   0, st, in, cplx, b0, L, inner, tw, out);
            }
        };
        ASD_TRY(mark(ctx, 2));
        hipLaunchKernelGGL(window_mean_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, d_snaps_a,
                           tstride, n3, T0, nlags, fbar);
        if (two)
            hipLaunchKernelGGL(window_mean_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, d_snaps_b,
                               tstride, n3, T0, 1, fbar + wm);
        const long nb = nlags + (two ? 1 : 0);       // window means are contiguous
        dft(fbar, 0, nb * L, 3, pt);                 // over col
        dft(pt, 1, nb, 3 * L, Fb);                   // over r
        ASD_TRY(mark(ctx, 3));
        dft(gbins, 0, (long)nlags * L, 1, Ht);  // bins over dc
        dft(Ht, 1, nlags, L, Sc);                                 // over dr
        const double *Fz = two ? Fb + 2 * wm : Fb;
        hipLaunchKernelGGL(finish_full_kernel, dim3((unsigned)(((long)nq * nlags + 255) / 256)), dim3(256), 0, st,
                           (const double *)Sc, (const double *)Fb, Fz, L, ctx->qdev.as<int>(), nq, nlags, T0, d_out);
        ASD_HIP(hipGetLastError());
        ASD_TRY(mark(ctx, 4));
        return ASD_OK;
    }

    // Fourier sums of every snapshot the windows touch (A side covers [0, nt_a);
    // B side is snapshots [0, T0) of d_snaps_b, which equals d_snaps_a in the
    // time-origin form; the replica form passes T0 = 1 window of its own)
    ASD_TRY(ctx->fq.ensure(sizeof(double) * 6 * (long)(nt_a + T0) * nq));
    double *Fa = ctx->fq.as<double>();
    hipLaunchKernelGGL(fourier_kernel, dim3(nq, nt_a), dim3(256), 0, st, d_snaps_a, tstride, L,
                       ctx->qdev.as<int>(), ctx->twiddle.as<double2>(), Fa, nq);
    const double *Fz = Fa;
    if (d_snaps_b != d_snaps_a) {
        double *Fb = Fa + 6L * nt_a * nq;
        hipLaunchKernelGGL(fourier_kernel, dim3(nq, T0), dim3(256), 0, st, d_snaps_b, tstride, L,
                           ctx->qdev.as<int>(), ctx->twiddle.as<double2>(), Fb, nq);
        Fz = Fb;
    }
    ASD_TRY(mark(ctx, 3));
    hipLaunchKernelGGL(finish_kernel, dim3(nq, nlags), dim3(256), 0, st, (const double *)gbins, L,
                       ctx->qdev.as<int>(), ctx->twiddle.as<double2>(), (const double *)Fa, Fz, nq, T0, d_out);
    ASD_HIP(hipGetLastError());
    ASD_TRY(mark(ctx, 4));
    return ASD_OK;
}

int sqt_autotune(asd_ctx *ctx, int L, int T0, asd_precision prec, int *best, double *times_ms)
{
    const int nl = 8;                       // lags per timing run
    const int nsnap = T0 + nl - 1;
    const long n = 3L * L * L * nsnap;
    DevBuf snaps, out;
    ASD_TRY(snaps.ensure(sizeof(double) * n));
    ASD_TRY(out.ensure(sizeof(double) * 2 * nl));
    hipLaunchKernelGGL(fill_pseudo_spins_kernel, dim3(1024), dim3(256), 0, ctx->stream, snaps.as<double>(), n / 3);
    const int q[2] = {0, 0};
    const int saved_timing = ctx->timing;
    for (hipEvent_t &e : ctx->ev)
        if (!e) ASD_HIP(hipEventCreate(&e));
    ctx->timing = 1;
    double best_ms = 1e300;
    int bi = 0;
    int rc = ASD_OK;
    for (int c = 0; c < sqt_num_configs(prec) && rc == ASD_OK; c++) {
        double ms = 1e300;
        for (int rep = 0; rep < 3 && rc == ASD_OK; rep++) {
            rc = sqt_run(ctx, L, snaps.as<double>(), snaps.as<double>(), 3L * L * L, nsnap, T0, nl, q, 1, prec,
                         out.as<double>(), c);
            if (rc != ASD_OK) break;
            float t = 0.f;
            if (hipEventSynchronize(ctx->ev[2]) != hipSuccess || hipEventElapsedTime(&t, ctx->ev[1], ctx->ev[2]) != hipSuccess) {
                rc = set_error(ASD_EHIP, "autotune timing failed");
                break;
            }
            if (rep > 0) ms = std::min(ms, (double)t);
        }
        if (times_ms) times_ms[c] = ms;
        // a later configuration must be more than 1% faster: near-ties keep the earlier one, so
        // the choice does not flip from run to run on timing noise
        if (ms < 0.99 * best_ms) {
            best_ms = ms;
            bi = c;
        }
    }
    ctx->timing = saved_timing;
    (void)hipStreamSynchronize(ctx->stream);
    snaps.release();
    out.release();
    if (rc != ASD_OK) return rc;
    ctx->tuned[prec][std::make_pair(L, 3 * T0)] = bi;
    if (best) *best = bi;
    return ASD_OK;
}

}  // namespace asd
