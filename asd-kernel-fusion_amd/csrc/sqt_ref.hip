// sqt_ref.hip -- two S(q,tau) paths beside the fused one, for comparison and checking.
//
// asd_sqt_unfused: the reference's unfused data flow (gemm/llg_sqt_mkl.c:369-423) on
//   the GPU.  Per lag: G = (1/T0) A_tau B^T with rocBLAS dgemm into an explicit
//   Ns x Ns buffer (cblas_dgemm, :378), C = G - fbar zbar^T (:382-398), then the phase
//   reduction sum_ij C_ij e^{-iq.(r_i - r_j)} (:401-423, one q there).  Here the
//   reduction runs for all q as a second dgemm, C times the per-site [cos, sin](q.r_j)
//   table, and an FP64 Kahan sum over i.  It materialises Ns^2 doubles per lag (2.1 GB
//   at L = 128), which the fused path never does; it is kept to measure what fusing saves.
//
// asd_sqt_bruteforce: the reference's baseline loop (baseline/llg_sqt_baseline.c:340-418)
//   with its disconnected term applied (the baseline computes ssum before subtracting
//   fbar.zbar): per (tau, q) the double sum over sites and the sum over time origins,
//   with phases from FP64 sincos instead of the lattice twiddle table.  For L <= 16; an
//   independent device-side check of the fused epilogue maths.
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "asd_internal.h"

namespace asd {

namespace {

// X[k][i] = S_i^c(t), k = 3 t + c, for every snapshot t < nsnap (k-major, ld = Ns)
__global__ void ref_pack_kernel(const double *__restrict__ snaps, int Ns, long nk, double *__restrict__ X)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nk * Ns) return;
    const long k = e / Ns;
    const int i = (int)(e - k * Ns);
    const long t = k / 3;
    const int c = (int)(k - 3 * t);
    X[e] = snaps[t * 3L * Ns + 3L * i + c];
}

// window means, Kahan-compensated: fbar[i][c] = (1/T0) sum_{t0 < T0} S_i^c(t0 + tau)
__global__ void ref_window_mean_kernel(const double *__restrict__ snaps, int Ns, int T0, int tau,
                                       double *__restrict__ fbar)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3L * Ns) return;
    double s = 0.0, comp = 0.0;
    for (int t0 = 0; t0 < T0; t0++) {
        const double y = snaps[(long)(t0 + tau) * 3L * Ns + e] - comp, t = s + y;
        comp = (t - s) - y;
        s = t;
    }
    fbar[e] = s / (double)T0;
}

// C_ij = G_ij - fbar_i . zbar_j in place, G column-major (G[i + j Ns])
__global__ void ref_outer_subtract_kernel(double *__restrict__ G, int Ns, const double *__restrict__ fbar,
                                          const double *__restrict__ zbar)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)Ns * Ns) return;
    const long j = e / Ns;
    const int i = (int)(e - j * Ns);
    G[e] -= fbar[3 * i] * zbar[3 * j] + fbar[3 * i + 1] * zbar[3 * j + 1] + fbar[3 * i + 2] * zbar[3 * j + 2];
}

// S(q_m) = (1/Ns) sum_i (c_i - i s_i) (Wc_i + i Ws_i), W = C P with P[:, 2m] = cos(q_m.r),
// P[:, 2m+1] = sin(q_m.r): Re = sum c Wc + s Ws, Im = sum c Ws - s Wc.  One block per q.
__global__ void ref_phase_reduce_kernel(const double *__restrict__ W, const double *__restrict__ P, int Ns,
                                        double *__restrict__ out)
{
    const int m = blockIdx.x;
    const double *pc = P + 2L * m * Ns, *ps = pc + Ns, *wc = W + 2L * m * Ns, *ws = wc + Ns;
    double re = 0.0, rc = 0.0, im = 0.0, ic = 0.0;
    auto kadd = [](double &s, double &c, double x) {
        const double y = x - c, t = s + y;
        c = (t - s) - y;
        s = t;
    };
    for (int i = threadIdx.x; i < Ns; i += blockDim.x) {
        kadd(re, rc, pc[i] * wc[i] + ps[i] * ws[i]);
        kadd(im, ic, pc[i] * ws[i] - ps[i] * wc[i]);
    }
    __shared__ double sr[256], si[256];
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, ac = 0.0, b = 0.0, bc = 0.0;
        for (int k = 0; k < (int)blockDim.x; k++) {
            kadd(a, ac, sr[k]);
            kadd(b, bc, si[k]);
        }
        out[2 * m] = a / (double)Ns;
        out[2 * m + 1] = b / (double)Ns;
    }
}

// one block (256 threads) per (q, tau); thread i < Ns owns site i
__global__ __launch_bounds__(256) void ref_bruteforce_kernel(const double *__restrict__ snaps, int L, int T0,
                                                             const int *__restrict__ q, int nq,
                                                             double *__restrict__ out)
{
    const int iq = blockIdx.x, tau = blockIdx.y, Ns = L * L, i = threadIdx.x;
    const double qx = 2.0 * M_PI * q[2 * iq] / L, qy = 2.0 * M_PI * q[2 * iq + 1] / L;
    __shared__ double zbar[256][3];
    __shared__ double sr[256], si[256];
    const long ts = 3L * Ns;
    double fbar[3] = {0.0, 0.0, 0.0};
    if (i < Ns) {
        double z[3] = {0.0, 0.0, 0.0};
        for (int t0 = 0; t0 < T0; t0++)
            for (int c = 0; c < 3; c++) {
                z[c] += snaps[t0 * ts + 3 * i + c];
                fbar[c] += snaps[(t0 + tau) * ts + 3 * i + c];
            }
        for (int c = 0; c < 3; c++) {
            zbar[i][c] = z[c] / T0;
            fbar[c] /= T0;
        }
    }
    __syncthreads();
    double re = 0.0, im = 0.0;
    if (i < Ns) {
        const int ri = i / L, ci = i % L;
        for (int j = 0; j < Ns; j++) {
            double g = 0.0;
            for (int t0 = 0; t0 < T0; t0++)
                for (int c = 0; c < 3; c++) g += snaps[(t0 + tau) * ts + 3 * i + c] * snaps[t0 * ts + 3 * j + c];
            const double cval =
                g / T0 - (fbar[0] * zbar[j][0] + fbar[1] * zbar[j][1] + fbar[2] * zbar[j][2]);
            const int rj = j / L, cj = j % L;
            double s, co;
            sincos(qx * (ri - rj) + qy * (ci - cj), &s, &co);
            re += cval * co;  // e^{-i phi} = cos phi - i sin phi
            im -= cval * s;
        }
    }
    sr[i] = re;
    si[i] = im;
    __syncthreads();
    if (i == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 256; k++) {
            a += sr[k];
            b += si[k];
        }
        out[2 * ((long)tau * nq + iq)] = a / Ns;
        out[2 * ((long)tau * nq + iq) + 1] = b / Ns;
    }
}

#define ASD_BLAS(call)                                                                    \
    do {                                                                                  \
        rocblas_status s_ = (call);                                                       \
        if (s_ != rocblas_status_success)                                                 \
            return set_error(ASD_EHIP, "%s: rocblas status %d", #call, (int)s_);          \
    } while (0)

}  // namespace

int sqt_unfused_run(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                    int nq, double *d_out)
{
    const int Ns = L * L, K = 3 * T0;
    const long nk = 3L * (T0 + nlags - 1);
    ASD_CHECK(nsnap >= T0 + nlags - 1, "need nsnap >= T0 + nlags - 1 (%d < %d)", nsnap, T0 + nlags - 1);
    // the explicit correlation matrix and the per-site phase table (built on the host)
    // bound this path: Ns^2 and 2 nq Ns doubles
    ASD_CHECK(8.0 * Ns * Ns <= 64e9, "the unfused path stores Ns^2 doubles: L = %d is too large", L);
    ASD_CHECK(16.0 * nq * Ns <= 8e9, "the unfused path's phase table (2 nq Ns doubles) is too large: nq = %d", nq);
    hipStream_t st = ctx->stream;
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        ASD_BLAS(rocblas_create_handle(&h));
        ctx->blas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->blas);
    ASD_BLAS(rocblas_set_stream(h, st));
    DevBuf &X = ctx->ref0, &G = ctx->ref1, &P = ctx->ref2, &W = ctx->ref3, &mean = ctx->ref4;
    ASD_TRY(X.ensure(sizeof(double) * nk * Ns));
    ASD_TRY(G.ensure(sizeof(double) * (long)Ns * Ns));
    ASD_TRY(P.ensure(sizeof(double) * 2L * nq * Ns));
    ASD_TRY(W.ensure(sizeof(double) * 2L * nq * Ns));
    ASD_TRY(mean.ensure(sizeof(double) * 6L * Ns));
    // per-site phase table with the fused path's exact L-entry twiddles (angle 2 pi k / L
    // in long double): P[i + 2m Ns] = cos(q_m.r_i), P[i + (2m+1) Ns] = sin(q_m.r_i)
    {
        std::vector<double> tc(L), ts(L), hp(2L * nq * Ns);
        for (int k = 0; k < L; k++) {
            const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
            tc[k] = (double)cosl(ang);
            ts[k] = (double)sinl(ang);
        }
        for (int m = 0; m < nq; m++)
            for (int i = 0; i < Ns; i++) {
                const int idx = (int)(((long)h_q[2 * m] * (i / L) + (long)h_q[2 * m + 1] * (i % L)) % L);
                hp[2L * m * Ns + i] = tc[idx];
                hp[(2L * m + 1) * Ns + i] = ts[idx];
            }
        ASD_HIP(hipMemcpyAsync(P.ptr, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, st));
        ASD_HIP(hipStreamSynchronize(st));  // host vector goes out of scope
    }
    const long nx = nk * Ns;
    hipLaunchKernelGGL(ref_pack_kernel, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, d_snaps, Ns, nk,
                       X.as<double>());
    double *zbar = mean.as<double>() + 3L * Ns, *fbar = mean.as<double>();
    hipLaunchKernelGGL(ref_window_mean_kernel, dim3((unsigned)((3L * Ns + 255) / 256)), dim3(256), 0, st, d_snaps,
                       Ns, T0, 0, zbar);
    const double alpha = 1.0 / (double)T0, one = 1.0, zero = 0.0;
    for (int tau = 0; tau < nlags; tau++) {
        // G = (1/T0) A_tau B^T: A_tau = X rows 3 tau .. 3 tau + K - 1 (Ns x K, lda Ns), B = X rows 0 .. K - 1
        ASD_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, Ns, Ns, K, &alpha,
                               X.as<double>() + 3L * tau * Ns, Ns, X.as<double>(), Ns, &zero, G.as<double>(), Ns));
        hipLaunchKernelGGL(ref_window_mean_kernel, dim3((unsigned)((3L * Ns + 255) / 256)), dim3(256), 0, st,
                           d_snaps, Ns, T0, tau, fbar);
        hipLaunchKernelGGL(ref_outer_subtract_kernel, dim3((unsigned)(((long)Ns * Ns + 255) / 256)), dim3(256), 0,
                           st, G.as<double>(), Ns, (const double *)fbar, (const double *)zbar);
        // W = C P (Ns x 2 nq)
        ASD_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, Ns, 2 * nq, Ns, &one,
                               G.as<double>(), Ns, P.as<double>(), Ns, &zero, W.as<double>(), Ns));
        hipLaunchKernelGGL(ref_phase_reduce_kernel, dim3((unsigned)nq), dim3(256), 0, st, (const double *)W.ptr,
                           (const double *)P.ptr, Ns, d_out + 2L * tau * nq);
    }
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

int sqt_bruteforce_run(asd_ctx *ctx, int L, const double *d_snaps, int nsnap, int T0, int nlags, const int *h_q,
                       int nq, double *d_out)
{
    ASD_CHECK(L * L <= 256, "the brute-force path takes L <= 16 (got %d)", L);
    ASD_CHECK(nsnap >= T0 + nlags - 1, "need nsnap >= T0 + nlags - 1 (%d < %d)", nsnap, T0 + nlags - 1);
    ASD_TRY(ctx->ref5.ensure(sizeof(int) * 2L * nq));
    ASD_HIP(hipMemcpyAsync(ctx->ref5.ptr, h_q, sizeof(int) * 2L * nq, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(ref_bruteforce_kernel, dim3((unsigned)nq, (unsigned)nlags), dim3(256), 0, ctx->stream,
                       d_snaps, L, T0, (const int *)ctx->ref5.ptr, nq, d_out);
    ASD_HIP(hipStreamSynchronize(ctx->stream));  // h_q belongs to the caller
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

void blas_destroy(asd_ctx *ctx)
{
    if (ctx->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
    ctx->blas = nullptr;
}

}  // namespace asd
