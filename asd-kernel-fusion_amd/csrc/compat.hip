// compat.hip -- the reference fused_dgemm contract on the GPU (FP64).
//
// Reference: kernel_fusion/gemm_kernel.h:355-422 (fused_dgemm), :39-88 (epilogue),
// :227-249 (alpha applied while packing A).  Contract restated:
//   g_ij   = sum_k (alpha a_ik) b_kj      a column-major (lda), b column-major (ldb)
//   c[p]  += g_ij q[p],  c[p+4] += g_ij q[p+4]   p = (i/4)*8 + i%4 + j*ldc
//   sum   += sum_ij (c[p], c[p+4])  of the updated c
// Unlike the reference this is correct for every K (no re-summing per K block), keeps
// the m % 4 tail rows and the last column of odd n.  Each workgroup computes a
// 64 x 64 tile on VALU FMAs (4 x 4 per thread, 16-deep LDS stages), applies the
// epilogue, and writes a Kahan-compensated partial sum; a second kernel reduces the
// partials in a fixed order, so the result is deterministic.
#include "asd_internal.h"

namespace asd {

struct Kahan2 {
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

__global__ __launch_bounds__(256) void fused_dgemm_kernel(int m, int n, int k, double alpha,
                                                          const double *__restrict__ a, long lda,
                                                          const double *__restrict__ b, long ldb,
                                                          double *__restrict__ c, long ldc,
                                                          const double *__restrict__ q,
                                                          double *__restrict__ partial)
{
    constexpr int T = 64, BK = 16;
    __shared__ double sA[BK][T + 1], sB[BK][T + 1];
    __shared__ double red[2][256];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = blockIdx.x * T, j0 = blockIdx.y * T;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < k; k0 += BK) {
        for (int e = threadIdx.x; e < BK * T; e += 256) {
            const int kk = e / T, r = e % T;
            const int i = i0 + r, j = j0 + r, kg = k0 + kk;
            sA[kk][r] = (i < m && kg < k) ? alpha * a[i + (long)kg * lda] : 0.0;
            sB[kk][r] = (j < n && kg < k) ? b[kg + (long)j * ldb] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk++) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                av[u] = sA[kk][tx + 16 * u];
                bv[u] = sB[kk][ty + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
        }
        __syncthreads();
    }
    Kahan2 s;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = i0 + tx + 16 * u, j = j0 + ty + 16 * v;
            if (i < m && j < n) {
                const long p = (long)(i / 4) * 8 + i % 4 + (long)j * ldc;
                const double g = acc[u][v];
                const double cr = c[p] + g * q[p];
                const double ci = c[p + 4] + g * q[p + 4];
                c[p] = cr;
                c[p + 4] = ci;
                s.add(cr, ci);
            }
        }
    red[0][threadIdx.x] = s.re;
    red[1][threadIdx.x] = s.im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            red[0][threadIdx.x] += red[0][threadIdx.x + st];
            red[1][threadIdx.x] += red[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long blk = (long)blockIdx.y * gridDim.x + blockIdx.x;
        partial[2 * blk] = red[0][0];
        partial[2 * blk + 1] = red[1][0];
    }
}

// deterministic Kahan reduction of the per-tile partials (one workgroup)
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *__restrict__ partial, long nblk,
                                                              double *__restrict__ out)
{
    __shared__ double red[2][256];
    Kahan2 s;
    for (long b = threadIdx.x; b < nblk; b += 256) s.add(partial[2 * b], partial[2 * b + 1]);
    red[0][threadIdx.x] = s.re;
    red[1][threadIdx.x] = s.im;
    __syncthreads();
    if (threadIdx.x == 0) {
        Kahan2 t;
        for (int i = 0; i < 256; i++) t.add(red[0][i], red[1][i]);
        out[0] = t.re;
        out[1] = t.im;
    }
}

}  // namespace asd

using namespace asd;

extern "C" int asd_fused_dgemm(int *m_, int *n_, int *k_, double *alpha_, double *a, int *lda_, double *b,
                               int *ldb_, double *c, int *ldc_, double *fuse_C, double *sum)
{
    ASD_CHECK(m_ && n_ && k_ && alpha_ && lda_ && ldb_ && ldc_ && sum, "NULL argument");
    const int m = *m_, n = *n_, k = *k_;
    const long lda = *lda_, ldb = *ldb_, ldc = *ldc_;
    ASD_CHECK(m >= 0 && n >= 0 && k >= 0, "negative dimension");
    if (m == 0 || n == 0) return ASD_OK;
    ASD_CHECK(c && fuse_C, "NULL c / fuse_C");
    ASD_CHECK(k == 0 || (a && b), "NULL a / b");
    ASD_CHECK(lda >= m && (k == 0 || ldb >= k), "lda < m or ldb < k");
    const long mpad = (m + 3) / 4 * 4;
    ASD_CHECK(ldc >= 2 * mpad, "ldc must be >= 2*ceil4(m) (blocked 4-re/4-im layout)");

    static thread_local asd_ctx *ctx = nullptr;
    if (!ctx) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        ASD_TRY(asd_create(&ctx, dev));
    }
    const size_t ba = sizeof(double) * lda * (size_t)std::max(k, 1);
    const size_t bb = sizeof(double) * ldb * (size_t)n;
    const size_t bc = sizeof(double) * ldc * (size_t)n;
    const dim3 grid((unsigned)((m + 63) / 64), (unsigned)((n + 63) / 64));
    const long nblk = (long)grid.x * grid.y;
    ASD_TRY(ctx->c0.ensure(ba));
    ASD_TRY(ctx->c1.ensure(k ? bb : 8));
    ASD_TRY(ctx->c2.ensure(bc));
    ASD_TRY(ctx->c3.ensure(bc));
    ASD_TRY(ctx->c4.ensure(sizeof(double) * (2 * nblk + 2)));
    hipStream_t st = ctx->stream;
    if (k) {
        ASD_HIP(hipMemcpyAsync(ctx->c0.ptr, a, ba, hipMemcpyHostToDevice, st));
        ASD_HIP(hipMemcpyAsync(ctx->c1.ptr, b, bb, hipMemcpyHostToDevice, st));
    }
    ASD_HIP(hipMemcpyAsync(ctx->c2.ptr, c, bc, hipMemcpyHostToDevice, st));
    ASD_HIP(hipMemcpyAsync(ctx->c3.ptr, fuse_C, bc, hipMemcpyHostToDevice, st));
    double *partial = ctx->c4.as<double>();
    hipLaunchKernelGGL(fused_dgemm_kernel, grid, dim3(256), 0, st, m, n, k, *alpha_, ctx->c0.as<double>(), lda,
                       ctx->c1.as<double>(), ldb, ctx->c2.as<double>(), ldc, ctx->c3.as<double>(), partial);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, st, (const double *)partial, nblk,
                       partial + 2 * nblk);
    ASD_HIP(hipGetLastError());
    double tot[2];
    ASD_HIP(hipMemcpyAsync(c, ctx->c2.ptr, bc, hipMemcpyDeviceToHost, st));
    ASD_HIP(hipMemcpyAsync(tot, partial + 2 * nblk, sizeof tot, hipMemcpyDeviceToHost, st));
    ASD_HIP(hipStreamSynchronize(st));
    sum[0] += tot[0];
    sum[1] += tot[1];
    return ASD_OK;
}
