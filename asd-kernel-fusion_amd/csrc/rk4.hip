// rk4.hip -- RK4 integration of the LLG equation on a periodic L x L lattice (FP64).
//
// Reference: kernel_fusion/llg_sqt_fusion.c
//   euler        :30-93   exchange field h = sum_nn J1 S_nb, f = h x S (v3d.h:48-53)
//   runge_kutta  :98-145  k_s = dt f(stage_s); cur = S + k1/6 + k2/3 + k3/3 + k4/6
//   renormalise  :350-361 S = len * cur / |cur|
// The reference makes ~14 full-array passes per step (9 OpenMP regions); here one
// kernel per step loads a (TR+8) x (TC+8) tile (radius-4 halo: each RK stage widens
// the dependency by one site) into LDS, runs all four stages there, renormalises
// and writes the TR x TC interior once: 24 B read + 24 B written per site-step.
//
// This file is compiled with -ffp-contract=off and keeps the reference's operation
// order (field accumulated neighbour by neighbour from 0.0, J multiplied per term,
// v3d_cross operand order, the k/6 + k/3 + k/3 + k/6 left-to-right sum, true
// divisions), so trajectories match the FP64 CPU restatement bit for bit.
#include "asd_internal.h"

namespace asd {

LLGParams make_llg(const asd_params &p)
{
    LLGParams q;
    q.J = p.J;
    q.D = p.D;
    q.Bz = p.Bz;
    q.alpha = p.alpha;
    q.dt = p.dt;
    q.len = p.spin_length;
    q.a_over_len = p.alpha / p.spin_length;
    q.den = 1.0 + p.alpha * p.alpha;
    q.inv_den = 1.0 / q.den;
    return q;
}

// f(S) at one site given its four neighbours in the reference order
// (row+1, row-1, col-1, col+1; llg_sqt_fusion.c:181-184).
__device__ __forceinline__ void llg_f(const LLGParams &p, double sx, double sy, double sz,
                                      const double *n0, const double *n1, const double *n2,
                                      const double *n3, double &fx, double &fy, double &fz)
{
    double hx = 0.0, hy = 0.0, hz = 0.0;
    hx += n0[0] * p.J; hy += n0[1] * p.J; hz += n0[2] * p.J;
    hx += n1[0] * p.J; hy += n1[1] * p.J; hz += n1[2] * p.J;
    hx += n2[0] * p.J; hy += n2[1] * p.J; hz += n2[2] * p.J;
    hx += n3[0] * p.J; hy += n3[1] * p.J; hz += n3[2] * p.J;
    hz += 2.0 * p.D * sz;
    hz += p.Bz;
    const double px = hy * sz - hz * sy;
    const double py = hz * sx - hx * sz;
    const double pz = hx * sy - hy * sx;
    const double qx = sy * pz - sz * py;
    const double qy = sz * px - sx * pz;
    const double qz = sx * py - sy * px;
    fx = (px + p.a_over_len * qx) * p.inv_den;
    fy = (py + p.a_over_len * qy) * p.inv_den;
    fz = (pz + p.a_over_len * qz) * p.inv_den;
}

__device__ __forceinline__ int wrap(int x, int L)
{
    x %= L;
    return x < 0 ? x + L : x;
}

// One full RK4 step (+ renormalisation) of one TR x TC output tile, one interior
// site per thread.  LDS holds the tile + halo twice, structure-of-arrays: S and the
// latest increment K.  The stage input Y_s = S + c k_{s-1} (c = 1/2, 1/2, 1) is formed
// when read -- the same expression the reference stores, so the same rounding -- which
// keeps LDS at two arrays (77 KB for 32 x 32: two workgroups per CU overlap each
// other's load latency and barriers).  Each thread keeps its sites' new increments in
// registers until every read of the previous K is done.
template <int TR, int TC> struct TileLDS {
    static constexpr int R = 4, RR = TR + 2 * R, RC = TC + 2 * R, NR = RR * RC;
    double S[3][NR], K[3][NR];
};

template <int TR, int TC>
__device__ __forceinline__ void rk4_tile(TileLDS<TR, TC> &sm, const LLGParams &p, int L, int tile_r, int tile_c,
                                         const double *__restrict__ in, double *__restrict__ out,
                                         double *__restrict__ snap)
{
    constexpr int R = 4, RR = TR + 2 * R, RC = TC + 2 * R, NR = RR * RC, NT = TR * TC;
    constexpr int MAXE = ((RR - 2) * (RC - 2) + NT - 1) / NT;  // region sites per thread, stage 1
    double(&sS)[3][NR] = sm.S;
    double(&sK)[3][NR] = sm.K;
    const int r0 = tile_r * TR - R, c0 = tile_c * TC - R;
    const int t = threadIdx.x;

    // tile coordinates lie in [-R, L + RR); when L >= RR one correction wraps them,
    // which avoids a runtime integer division per loaded site
    const bool near = L >= RR;
    for (int idx = t; idx < NR; idx += NT) {
        int gr = r0 + idx / RC, gc = c0 + idx % RC;
        if (near) {
            gr += gr < 0 ? L : (gr >= L ? -L : 0);
            gc += gc < 0 ? L : (gc >= L ? -L : 0);
        } else {
            gr = wrap(gr, L);
            gc = wrap(gc, L);
        }
        const double *s = in + 3 * ((long)gr * L + gc);
        sS[0][idx] = s[0]; sS[1][idx] = s[1]; sS[2][idx] = s[2];
    }
    // this thread's interior site
    const int ir = t / TC, ic = t % TC;
    const int grow = tile_r * TR + ir, gcol = tile_c * TC + ic;
    const bool owner = grow < L && gcol < L;
    const int own = (ir + R) * RC + (ic + R);
    __syncthreads();

    double ax, ay, az;  // cur = S + k1/6 + k2/3 + k3/3 + k4/6, accumulated left to right
    if (snap && owner) {
        double *o = snap + 3 * ((long)grow * L + gcol);
        o[0] = sS[0][own]; o[1] = sS[1][own]; o[2] = sS[2][own];
    }

    // unrolled: the region width w of each stage is a compile-time constant, so the
    // e / w, e % w index arithmetic needs no runtime division
#pragma unroll
    for (int stage = 1; stage <= 4; stage++) {
        // increments on the region within radius R - stage of the interior
        const int lo = stage, hr = RR - stage, hc = RC - stage;
        const int w = hc - lo, nsite = (hr - lo) * w;
        // Y_stage at LDS index i (llg_sqt_fusion.c:111, :121, :131)
        auto Y = [&](int d, int i) -> double {
            if (stage == 1) return sS[d][i];
            if (stage == 4) return sS[d][i] + sK[d][i];
            return sS[d][i] + sK[d][i] / 2.0;
        };
        double kr[MAXE][3];
#pragma unroll
        for (int j = 0; j < MAXE; j++) {
            const int e = t + j * NT;
            if (e < nsite) {
                const int idx = (lo + e / w) * RC + lo + e % w;
                double n[4][3];
                const int nb[4] = {idx + RC, idx - RC, idx - 1, idx + 1};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    n[q][0] = Y(0, nb[q]); n[q][1] = Y(1, nb[q]); n[q][2] = Y(2, nb[q]);
                }
                double fx, fy, fz;
                llg_f(p, Y(0, idx), Y(1, idx), Y(2, idx), n[0], n[1], n[2], n[3], fx, fy, fz);
                kr[j][0] = fx * p.dt; kr[j][1] = fy * p.dt; kr[j][2] = fz * p.dt;
            }
        }
        __syncthreads();  // every read of the previous increment is done
#pragma unroll
        for (int j = 0; j < MAXE; j++) {
            const int e = t + j * NT;
            if (e < nsite) {
                const int idx = (lo + e / w) * RC + lo + e % w;
                sK[0][idx] = kr[j][0]; sK[1][idx] = kr[j][1]; sK[2][idx] = kr[j][2];
            }
        }
        __syncthreads();
        const double kx = sK[0][own], ky = sK[1][own], kz = sK[2][own];
        if (stage == 1) {
            ax = sS[0][own] + kx / 6.0; ay = sS[1][own] + ky / 6.0; az = sS[2][own] + kz / 6.0;
        } else if (stage == 4) {
            ax = ax + kx / 6.0; ay = ay + ky / 6.0; az = az + kz / 6.0;
        } else {
            ax = ax + kx / 3.0; ay = ay + ky / 3.0; az = az + kz / 3.0;
        }
    }
    if (owner) {
        const double nrm = sqrt(ax * ax + ay * ay + az * az);
        double *o = out + 3 * ((long)grow * L + gcol);
        o[0] = p.len * ax / nrm;
        o[1] = p.len * ay / nrm;
        o[2] = p.len * az / nrm;
    }
}

// single lattice (nrep replicas along blockIdx.z); snapshot slot from *snap_counter
// when given (graph replay), else slot 0 of snap
template <int TR, int TC>
__global__ __launch_bounds__(TR *TC) void rk4_step_kernel(LLGParams p, int L,
                                                          const double *__restrict__ in,
                                                          double *__restrict__ out,
                                                          double *__restrict__ snap,
                                                          const int *__restrict__ snap_counter,
                                                          long snap_stride)
{
    __shared__ TileLDS<TR, TC> sm;
    const long base = (long)blockIdx.z * L * L * 3;
    double *sn = nullptr;
    if (snap) sn = snap + (snap_counter ? (long)*snap_counter : 0) * snap_stride + base;
    rk4_tile<TR, TC>(sm, p, L, blockIdx.y, blockIdx.x, in + base, out + base, sn);
}

// ------------------------------------------------------------------ batching
// Uneven-size independent trajectories in one persistent launch per step: the
// grid strides over a work list of (trajectory, tile) items; every trajectory
// has its own L, parameters and state/snapshot buffers (BatchTraj table).
template <int TR, int TC>
__global__ __launch_bounds__(TR *TC) void rk4_batch_kernel(const BatchTraj *__restrict__ trajs,
                                                           const int2 *__restrict__ items, int nitems,
                                                           int parity, const int *__restrict__ snap_counter,
                                                           int write_snap)
{
    __shared__ TileLDS<TR, TC> sm;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        const BatchTraj &tr = trajs[w.x];
        const int ntc = (tr.L + TC - 1) / TC;
        const double *in = parity ? tr.b : tr.a;
        double *out = parity ? tr.a : tr.b;
        double *sn = (write_snap && tr.snaps) ? tr.snaps + (long)*snap_counter * 3L * tr.L * tr.L : nullptr;
        __syncthreads();  // LDS of the previous item fully consumed
        rk4_tile<TR, TC>(sm, tr.p, tr.L, w.y / ntc, w.y % ntc, in, out, sn);
    }
}

__global__ void bump_counter_kernel(int *c) { *c += 1; }

// ------------------------------------------------- compat: euler / runge_kutta

// derivative = f(S), one site per thread (llg_sqt_fusion.c:30-93)
__global__ void rhs_kernel(LLGParams p, int L, long nsites_total, const double *__restrict__ S,
                           double *__restrict__ dS, double scale)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsites_total) return;
    const long ns = (long)L * L;
    const long rep = g / ns;
    const int i = (int)(g - rep * ns);
    const int row = i / L, col = i % L;
    const double *b = S + rep * ns * 3;
    const double *n0 = b + 3 * ((long)wrap(row + 1, L) * L + col);
    const double *n1 = b + 3 * ((long)wrap(row - 1, L) * L + col);
    const double *n2 = b + 3 * ((long)row * L + wrap(col - 1, L));
    const double *n3 = b + 3 * ((long)row * L + wrap(col + 1, L));
    const double *s = b + 3 * (long)i;
    double fx, fy, fz;
    llg_f(p, s[0], s[1], s[2], n0, n1, n2, n3, fx, fy, fz);
    double *o = dS + 3 * g;
    if (scale != 1.0) {  // k *= dt, as runge_kutta does in place
        fx *= scale; fy *= scale; fz *= scale;
    }
    o[0] = fx; o[1] = fy; o[2] = fz;
}

// cur = spin + k * c  (c = 1/2 as k/2.0, or 1)  -- llg_sqt_fusion.c:111, :121, :131
__global__ void axpy_stage_kernel(long n, const double *__restrict__ spin, const double *__restrict__ k,
                                  double *__restrict__ cur, int half)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    cur[j] = half ? spin[j] + k[j] / 2.0 : spin[j] + k[j];
}

// cur = spin + k1/6 + k2/3 + k3/3 + k4/6  -- llg_sqt_fusion.c:141
__global__ void rk4_combine_kernel(long n, const double *__restrict__ spin, const double *__restrict__ k1,
                                   const double *__restrict__ k2, const double *__restrict__ k3,
                                   const double *__restrict__ k4, double *__restrict__ cur)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    cur[j] = spin[j] + k1[j] / 6.0 + k2[j] / 3.0 + k3[j] / 3.0 + k4[j] / 6.0;
}

int rhs_launch(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *S, double *dS)
{
    const long n = (long)nrep * L * L;
    hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, L, n,
                       S, dS, 1.0);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

int rk4_compat_stages(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *spin,
                      double *cur, double *k1, double *k2, double *k3, double *k4)
{
    const long ns = (long)nrep * L * L, n = ns * 3;
    const dim3 gs((unsigned)((ns + 255) / 256)), ge((unsigned)((n + 255) / 256)), b(256);
    double *ks[4] = {k1, k2, k3, k4};
    // stage 1 reads spin directly (cur_input = spin, :101-103)
    for (int s = 0; s < 4; s++) {
        hipLaunchKernelGGL(rhs_kernel, gs, b, 0, ctx->stream, p, L, ns, s == 0 ? spin : (const double *)cur,
                           ks[s], p.dt);
        if (s < 3)
            hipLaunchKernelGGL(axpy_stage_kernel, ge, b, 0, ctx->stream, n, spin, (const double *)ks[s], cur,
                               s < 2 ? 1 : 0);
    }
    hipLaunchKernelGGL(rk4_combine_kernel, ge, b, 0, ctx->stream, n, spin, (const double *)k1,
                       (const double *)k2, (const double *)k3, (const double *)k4, cur);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

// Stage inputs of one RK4 step for the adjoint: Y2 = S + k1/2, Y3 = S + k2/2,
// Y4 = S + k3 and the unnormalised c = S + k1/6 + k2/3 + k3/3 + k4/6, with the
// same rounding steps as the fused kernel (so recomputed stages are bit-identical).
int rk4_stages_store(asd_ctx *ctx, const LLGParams &p, int L, const double *S, double *Y2, double *Y3, double *Y4,
                     double *c, double *k1, double *k2, double *k3, double *k4)
{
    const long ns = (long)L * L, n = ns * 3;
    const dim3 gs((unsigned)((ns + 255) / 256)), ge((unsigned)((n + 255) / 256)), b(256);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(rhs_kernel, gs, b, 0, st, p, L, ns, S, k1, p.dt);
    hipLaunchKernelGGL(axpy_stage_kernel, ge, b, 0, st, n, S, (const double *)k1, Y2, 1);
    hipLaunchKernelGGL(rhs_kernel, gs, b, 0, st, p, L, ns, (const double *)Y2, k2, p.dt);
    hipLaunchKernelGGL(axpy_stage_kernel, ge, b, 0, st, n, S, (const double *)k2, Y3, 1);
    hipLaunchKernelGGL(rhs_kernel, gs, b, 0, st, p, L, ns, (const double *)Y3, k3, p.dt);
    hipLaunchKernelGGL(axpy_stage_kernel, ge, b, 0, st, n, S, (const double *)k3, Y4, 0);
    hipLaunchKernelGGL(rhs_kernel, gs, b, 0, st, p, L, ns, (const double *)Y4, k4, p.dt);
    hipLaunchKernelGGL(rk4_combine_kernel, ge, b, 0, st, n, S, (const double *)k1, (const double *)k2,
                       (const double *)k3, (const double *)k4, c);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

// ------------------------------------------------------------- step launcher

int rk4_batch_launch(asd_ctx *ctx, const BatchTraj *d_trajs, const int2 *d_items, int nitems, int parity,
                     const int *snap_counter, int write_snap, int bump_counter)
{
    constexpr int TR = 16, TC = 16;
    const int grid = std::min(nitems, 4 * std::max(ctx->num_cu, 1));
    hipLaunchKernelGGL((rk4_batch_kernel<TR, TC>), dim3(grid), dim3(TR * TC), 0, ctx->stream, d_trajs, d_items,
                       nitems, parity, snap_counter, write_snap);
    if (bump_counter)
        hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(1), 0, ctx->stream, (int *)snap_counter);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

int rk4_launch_step(asd_ctx *ctx, const LLGParams &p, int L, int nrep, const double *in, double *out,
                    double *snap_out, const int *snap_counter, long snap_stride, int bump_counter)
{
    // 16 x 16 tiles (256 threads, 27.6 KB LDS, 5 workgroups per CU; 2.25 sites loaded
    // and 5.7 f-evaluations per output site) are the default: measured faster at every
    // size than 32 x 32 (1024 threads, 76.8 KB, 1.56 loads and 4.8 f-evaluations per
    // site), which the 128-register cap of 1024-thread workgroups limits to one
    // workgroup per CU.  32 x 32 stays selectable (asd_rk4_set_tile).
    const bool big = ctx->rk4_tile == 32;
    if (big) {
        constexpr int TR = 32, TC = 32;
        const dim3 grid((unsigned)((L + TC - 1) / TC), (unsigned)((L + TR - 1) / TR), (unsigned)nrep);
        hipLaunchKernelGGL((rk4_step_kernel<TR, TC>), grid, dim3(TR * TC), 0, ctx->stream, p, L, in, out,
                           snap_out, snap_counter, snap_stride);
    } else {
        constexpr int TR = 16, TC = 16;
        const dim3 grid((unsigned)((L + TC - 1) / TC), (unsigned)((L + TR - 1) / TR), (unsigned)nrep);
        hipLaunchKernelGGL((rk4_step_kernel<TR, TC>), grid, dim3(TR * TC), 0, ctx->stream, p, L, in, out,
                           snap_out, snap_counter, snap_stride);
    }
    if (bump_counter)
        hipLaunchKernelGGL(bump_counter_kernel, dim3(1), dim3(1), 0, ctx->stream, (int *)snap_counter);
    ASD_HIP(hipGetLastError());
    return ASD_OK;
}

}  // namespace asd
