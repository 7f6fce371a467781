// adjoint.hip -- hand-written discrete adjoint of RK4-LLG (with renormalisation)
// and of S(q,tau), giving vector-Jacobian products d<g, S>/d theta for
// theta = (J, D, B_z, alpha).
//
// Forward map per step (reference runge_kutta, kernel_fusion/llg_sqt_fusion.c:98-145,
// and renormalisation :350-361):
//   Y1 = S, Y2 = S + k1/2, Y3 = S + k2/2, Y4 = S + k3,  k_s = dt f(Y_s; theta)
//   c = S + k1/6 + k2/3 + k3/3 + k4/6,   S' = len c / |c|
// Reverse step (lambda' = dL/dS'):
//   lambda_c = len (lambda' - c^ (c^ . lambda')) / |c|
//   mu4 = lambda_c/6              -> lambda_Y4 = dt J_f(Y4)^T mu4
//   mu3 = lambda_c/3 + lambda_Y4  -> lambda_Y3 = dt J_f(Y3)^T mu3
//   mu2 = lambda_c/3 + lambda_Y3/2 -> lambda_Y2 = dt J_f(Y2)^T mu2
//   mu1 = lambda_c/6 + lambda_Y2/2 -> lambda_Y1 = dt J_f(S)^T mu1
//   lambda = lambda_c + lambda_Y1 + lambda_Y2 + lambda_Y3 + lambda_Y4
//   theta_bar += dt sum_s (df/dtheta)(Y_s)^T mu_s
// with f = (p + a S x p)/den, p = h x S, h = J sum_nn S_j + (2 D S_z + B_z) z,
// a = alpha/len, den = 1 + alpha^2.  For mu at site i, nu = (mu + a mu x S)/den:
//   dL/dh_i = S x nu                               (sent to the 4 neighbours times J)
//   dL/dS_i (local) = nu x h + (a/den) p x mu + 2 D (S x nu)_z z
//   dJ: (S x nu).sum_nn S ; dD: 2 (S x nu)_z S_z ; dB_z: (S x nu)_z ;
//   dalpha: mu . ((S x p)/(len den) - 2 alpha f / den)
// Stage inputs are recomputed from checkpoints with the forward kernels, so the
// recomputed trajectory is bit-identical to the forward one.
//
// S(q,tau) side (the correlation): with upstream g(q,tau) = dL/dRe S + i dL/dIm S and
// R = g/2, F_q(t) = sum_i S_i(t) e^{-iq.r_i}, Fbar_q(tau) = (1/T0) sum_t0 F_q(t0+tau):
//   Lambda_q(t) = 1/(Ns T0) sum_tau { [0<=t-tau<T0] R(tau) (F(t-tau) - Fbar(0))
//                                    + [t<T0] conj(R(tau)) (F(t+tau) - Fbar(tau)) }
//   dL/dS_i(t) = 2 Re sum_q conj(Lambda_q(t)) e^{-iq.r_i}
// i.e. the exact adjoint of the GEMM form G -> bins -> S(q,tau), written through the
// factorisation e^{-iq.(r_i - r_j)} = e^{-iq.r_i} e^{iq.r_j} that the forward
// epilogue's displacement bins evaluate.
#include <algorithm>
#include <utility>
#include <cmath>
#include <vector>

#include "asd_internal.h"

namespace asd {

__device__ __forceinline__ int wrapi(int x, int L)
{
    x %= L;
    return x < 0 ? x + L : x;
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// grad[k] = sum over blocks b (in order) of gacc[b][k]; one thread per parameter
__global__ void grad_reduce_kernel(const double *__restrict__ gacc, int nblk, double *__restrict__ grad)
{
    const int k = threadIdx.x;
    if (k >= 4) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += gacc[4L * b + k];
    grad[k] = s;
}

// The reverse sweep's per-segment work reads the segment index from the device (*sg), so one
// captured graph serves every full segment: dst = src[*sg * n ..] (the segment's checkpoint),
// dst = src (plain copy), *sg += delta (the next segment).
__global__ void copy_indexed_kernel(long n, double *__restrict__ dst, const double *__restrict__ src,
                                    const int *__restrict__ sg)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(long)*sg * n + i];
}
__global__ void copy_kernel(long n, double *__restrict__ dst, const double *__restrict__ src)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
__global__ void step_index_kernel(int *sg, int delta) { *sg += delta; }
__global__ void set_index_kernel(int *sg, int v) { *sg = v; }

// Fbar[tau][q][6] = (1/T0) sum_{t0<T0} F[t0+tau][q][6]
__global__ void fbar_kernel(const double *__restrict__ F, int nq, int T0, int nlags, double *__restrict__ Fbar)
{
    const int q = blockIdx.x, tau = blockIdx.y, k = threadIdx.x;
    if (k >= 6) return;
    double s = 0.0;
    for (int t0 = 0; t0 < T0; t0++) s += F[((long)(t0 + tau) * nq + q) * 6 + k];
    Fbar[((long)tau * nq + q) * 6 + k] = s / (double)T0;
}

// Lambda[t][q][6] (3 complex) from R = g/2; grid (nq, nt), 256 threads over tau
__global__ __launch_bounds__(256) void lambda_kernel(const double *__restrict__ g, const double *__restrict__ F,
                                                     const double *__restrict__ Fbar, int nq, int T0, int nlags,
                                                     int Ns, double *__restrict__ Lam)
{
    __shared__ double red[6][256];
    const int q = blockIdx.x, t = blockIdx.y;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double *F0 = Fbar + (long)q * 6;  // Fbar(0)
    for (int tau = threadIdx.x; tau < nlags; tau += 256) {
        const double rr = 0.5 * g[((long)tau * nq + q) * 2], ri = 0.5 * g[((long)tau * nq + q) * 2 + 1];
        const int a = t - tau;
        if (a >= 0 && a < T0) {
            const double *Fa = F + ((long)a * nq + q) * 6;
            for (int c = 0; c < 3; c++) {
                const double xr = Fa[2 * c] - F0[2 * c], xi = Fa[2 * c + 1] - F0[2 * c + 1];
                acc[2 * c] += rr * xr - ri * xi;
                acc[2 * c + 1] += rr * xi + ri * xr;
            }
        }
        if (t < T0) {
            const double *Fb = F + ((long)(t + tau) * nq + q) * 6;
            const double *Fbt = Fbar + ((long)tau * nq + q) * 6;
            for (int c = 0; c < 3; c++) {
                const double xr = Fb[2 * c] - Fbt[2 * c], xi = Fb[2 * c + 1] - Fbt[2 * c + 1];
                acc[2 * c] += rr * xr + ri * xi;        // conj(R) x
                acc[2 * c + 1] += rr * xi - ri * xr;
            }
        }
    }
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 6) Lam[((long)t * nq + q) * 6 + threadIdx.x] = red[threadIdx.x][0] / ((double)Ns * T0);
}

// dX[t][i][c] = 2 sum_q (Lr cos(q.r_i) - Li sin(q.r_i)); grid (ceil(Ns/256), nt)
__global__ void dx_kernel(const double *__restrict__ Lam, const int *__restrict__ q, const double2 *__restrict__ tw,
                          int nq, int L, double *__restrict__ dX)
{
    const int Ns = L * L;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
    if (i >= Ns) return;
    const int r = i / L, c = i - r * L;
    double a[3] = {0, 0, 0};
    for (int iq = 0; iq < nq; iq++) {
        const int k = (int)(((long)q[2 * iq] * r + (long)q[2 * iq + 1] * c) % L);
        const double2 e = tw[k];
        const double *l = Lam + ((long)t * nq + iq) * 6;
        for (int d = 0; d < 3; d++) a[d] += l[2 * d] * e.x - l[2 * d + 1] * e.y;
    }
    for (int d = 0; d < 3; d++) dX[((long)t * Ns + i) * 3 + d] = 2.0 * a[d];
}

// One reverse RK4 step in one launch, per T x T tile with a radius-5 halo: the renormalisation
// adjoint, then for the stages s = 4, 3, 2, 1 the local part of J_f^T (adj_jt1's formulas: dL/dh,
// the local dL/dS and the parameter partials) and its neighbour gather (adj_jt2's: lambda_Y =
// dt (gS + J sum_nn dL/dh)), with the stage inputs and dL/dh in LDS.  Each reverse stage widens
// the dependency by one site (the gather), and the first stage's field needs one more, so the
// tile's own sites (depth >= 5 in the loaded region) end with the exact lamS of the separate
// kernels' formulas.  Parameter partials are summed over the tile's own sites only, in a fixed
// order (thread, then a tree over the block), into the block's slot gacc[block][4].  The snapshot
// adjoint dX[*sg * mul + add] (add < 0: none) is added to the result, written to lam_out.
constexpr int kAdjTile = 8, kAdjHalo = 5, kAdjRR = kAdjTile + 2 * kAdjHalo, kAdjNR = kAdjRR * kAdjRR;
constexpr int kAdjThreads = 512;  // >= kAdjNR: one loaded site per thread
static_assert(kAdjThreads >= kAdjNR, "one site of the loaded region per thread");

__global__ __launch_bounds__(kAdjThreads) void adj_step_kernel(
    LLGParams p, int L, const double *__restrict__ S, const double *__restrict__ Y2, const double *__restrict__ Y3,
    const double *__restrict__ Y4, const double *__restrict__ c, const double *__restrict__ lam,
    double *__restrict__ lam_out, double *__restrict__ gacc, const double *__restrict__ dX,
    const int *__restrict__ sg, int mul, int add, int nslots)
{
    __shared__ double sY[3][kAdjNR], sG[3][kAdjNR];
    __shared__ double red[4][kAdjThreads];
    const int t = threadIdx.x;
    const int tr = blockIdx.y, tc = blockIdx.x;
    const bool has = t < kAdjNR;
    const int er = has ? t / kAdjRR : 0, ec = has ? t % kAdjRR : 0;
    const int depth = has ? min(min(er, kAdjRR - 1 - er), min(ec, kAdjRR - 1 - ec)) : -1;
    const int grow = wrapi(tr * kAdjTile - kAdjHalo + er, L), gcol = wrapi(tc * kAdjTile - kAdjHalo + ec, L);
    const long gi = (long)grow * L + gcol;
    const bool own = depth >= kAdjHalo && tr * kAdjTile + er - kAdjHalo < L && tc * kAdjTile + ec - kAdjHalo < L;
    // neighbours in the loaded region (row+1, row-1, col-1, col+1), valid where depth >= 1
    const int li = er * kAdjRR + ec;
    const int nb[4] = {li + kAdjRR, li - kAdjRR, li - 1, li + 1};

    // renormalisation adjoint: lamc = len (lam - u (u . lam)) / |c|, lamS = lamc
    double lamc[3] = {0, 0, 0}, lamS[3] = {0, 0, 0}, carry[3] = {0, 0, 0}, gS[3] = {0, 0, 0};
    if (has) {
        const double *cc = c + 3 * gi, *l = lam + 3 * gi;
        const double nrm = sqrt(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
        const double u[3] = {cc[0] / nrm, cc[1] / nrm, cc[2] / nrm};
        const double d = u[0] * l[0] + u[1] * l[1] + u[2] * l[2];
        for (int k = 0; k < 3; k++) {
            lamc[k] = p.len * (l[k] - u[k] * d) / nrm;
            lamS[k] = lamc[k];
        }
    }
    double part[4] = {0, 0, 0, 0};
    const double *Ys[4] = {S, Y2, Y3, Y4};
    const double ca[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double cb[4] = {0.5, 0.5, 1.0, 0.0};
    for (int sidx = 3; sidx >= 0; sidx--) {
        const int level = 3 - sidx;
        if (has) {
            const double *y = Ys[sidx] + 3 * gi;
            sY[0][li] = y[0]; sY[1][li] = y[1]; sY[2][li] = y[2];
        }
        __syncthreads();
        if (depth >= level + 1) {  // adj_jt1's formulas at this site
            const double s[3] = {sY[0][li], sY[1][li], sY[2][li]};
            double sn[3] = {0, 0, 0};
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 3; k++) sn[k] += sY[k][nb[j]];
            const double h[3] = {p.J * sn[0], p.J * sn[1], p.J * sn[2] + 2.0 * p.D * s[2] + p.Bz};
            double pp[3], sxp[3], f[3], mu[3], mxs[3], nu[3], ghv[3], t1[3], t2[3];
            cross3(h, s, pp);
            cross3(s, pp, sxp);
            for (int k = 0; k < 3; k++) f[k] = (pp[k] + p.a_over_len * sxp[k]) / p.den;
            for (int k = 0; k < 3; k++) mu[k] = ca[sidx] * lamc[k] + (sidx == 3 ? 0.0 : cb[sidx] * carry[k]);
            cross3(mu, s, mxs);
            for (int k = 0; k < 3; k++) nu[k] = (mu[k] + p.a_over_len * mxs[k]) / p.den;
            cross3(s, nu, ghv);
            cross3(nu, h, t1);
            cross3(pp, mu, t2);
            for (int k = 0; k < 3; k++) {
                sG[k][li] = ghv[k];
                gS[k] = t1[k] + (p.a_over_len / p.den) * t2[k];
            }
            gS[2] += 2.0 * p.D * ghv[2];
            if (own) {
                part[0] += p.dt * (ghv[0] * sn[0] + ghv[1] * sn[1] + ghv[2] * sn[2]);
                part[1] += p.dt * 2.0 * ghv[2] * s[2];
                part[2] += p.dt * ghv[2];
                double da = 0.0;
                for (int k = 0; k < 3; k++) da += mu[k] * (sxp[k] / (p.len * p.den) - 2.0 * p.alpha * f[k] / p.den);
                part[3] += p.dt * da;
            }
        }
        __syncthreads();
        if (depth >= level + 2) {  // adj_jt2's gather
            for (int k = 0; k < 3; k++) {
                double a = 0.0;
                for (int n = 0; n < 4; n++) a += sG[k][nb[n]];
                const double v = p.dt * (gS[k] + p.J * a);
                lamS[k] += v;
                carry[k] = v;
            }
        }
        __syncthreads();  // sY / sG are overwritten by the next stage
    }
    if (own) {
        const long slot = add >= 0 ? (long)*sg * mul + add : -1;
        double *o = lam_out + 3 * gi;
        for (int k = 0; k < 3; k++) {
            double v = lamS[k];
            if (slot >= 0 && slot < nslots) v += dX[slot * 3L * L * L + 3 * gi + k];
            o[k] = v;
        }
    }
    for (int k = 0; k < 4; k++) red[k][t] = part[k];
    __syncthreads();
    for (int st = kAdjThreads / 2; st > 0; st >>= 1) {
        if (t < st)
            for (int k = 0; k < 4; k++) red[k][t] += red[k][t + st];
        __syncthreads();
    }
    if (t < 4) gacc[4L * (blockIdx.y * gridDim.x + blockIdx.x) + t] += red[t][0];
}

static dim3 g1(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace asd

using namespace asd;

static int sqt_vjp_exact(asd_ctx *c, const asd_params *pp, int L, const double *d_S0, long nsteps, int every,
                         int T0, int nlags, const int *h_q, int nq, const double *h_g, int checkpoint_every,
                         double grad[4], double *h_sqt, const double *h_target = nullptr, double *loss = nullptr);

// The reverse sweep linearises the reference's operation order (rk4_stages_store), so the
// forward and recomputed trajectories take the exact RK4 arithmetic whatever the
// context's setting.
extern "C" int asd_sqt_vjp(asd_ctx *c, const asd_params *pp, int L, const double *d_S0, long nsteps, int every,
                           int T0, int nlags, const int *h_q, int nq, const double *h_g, int checkpoint_every,
                           double grad[4], double *h_sqt)
{
    ASD_CHECK(c, "NULL argument");
    const int saved = c->rk4_exact;
    c->rk4_exact = 1;
    const int r = sqt_vjp_exact(c, pp, L, d_S0, nsteps, every, T0, nlags, h_q, nq, h_g, checkpoint_every, grad, h_sqt);
    c->rk4_exact = saved;
    return r;
}

// h_target (with loss): the misfit L = sum |S - S_target|^2 is formed on the host from this
// forward pass and its gradient taken with g = 2 (S - S_target), in one forward pass.
static int sqt_vjp_exact(asd_ctx *c, const asd_params *pp, int L, const double *d_S0, long nsteps, int every,
                         int T0, int nlags, const int *h_q, int nq, const double *h_g, int checkpoint_every,
                         double grad[4], double *h_sqt, const double *h_target, double *loss)
{
    ASD_CHECK(c && pp && d_S0 && h_q && (grad || (!h_g && !h_target)) && (!h_target || loss), "NULL argument");
    ASD_CHECK(L >= 2 && L <= 1024, "L out of range");
    ASD_CHECK(every >= 1 && T0 >= 1 && nlags >= 1 && nq >= 1, "bad sizes");
    ASD_CHECK(checkpoint_every >= 1, "checkpoint_every must be >= 1");
    ASD_CHECK(pp->spin_length > 0 && pp->alpha >= 0, "bad parameters");
    for (int i = 0; i < 2 * nq; i++) ASD_CHECK(h_q[i] >= 0 && h_q[i] < L, "q index out of [0, L)");
    const int nsnap = T0 + nlags - 1;
    ASD_CHECK((long)(nsnap - 1) * every <= nsteps, "nsteps too small for T0 + nlags - 1 snapshots");
    ASD_HIP(hipSetDevice(c->device));
    // checkpoint interval: a multiple of the snapshot interval
    const int C = (checkpoint_every + every - 1) / every * every;
    const long ns = (long)L * L, n = 3 * ns;
    const long ncp = (nsteps + C - 1) / C + 1;
    hipStream_t st = c->stream;
    const LLGParams p = make_llg(*pp);

    // workspace: snapshots | checkpoints | segment states | 4 stage buffers | lamS | scratch | lam | dX |
    // per-block gradient partials | grad
    // the reverse step's tiles; gacc holds one parameter-partial slot per tile
    const dim3 tiles((unsigned)((L + kAdjTile - 1) / kAdjTile), (unsigned)((L + kAdjTile - 1) / kAdjTile));
    const int nblk = (int)(tiles.x * tiles.y);
    const long words = (long)nsnap * n + ncp * n + (long)C * n + 6 * n + n + (long)nsnap * n + 4L * nblk + 4 + 1;
    ASD_TRY(c->adj0.ensure(sizeof(double) * words));
    double *snaps = c->adj0.as<double>();
    double *cps = snaps + (long)nsnap * n;
    double *seg = cps + ncp * n;
    double *Y2 = seg + (long)C * n, *Y3 = Y2 + n, *Y4 = Y3 + n, *cc = Y4 + n;
    double *lamS = cc + n;      // dL/dS of one step; lam and lamS alternate within a segment
    double *scratch = lamS + n;  // the stage kernel's renormalised state (unused)
    double *lam = scratch + n;   // dL/dS at the start and end of every segment
    double *dX = lam + n;
    double *gacc = dX + (long)nsnap * n;
    double *dgrad = gacc + 4L * nblk;
    int *dsg = reinterpret_cast<int *>(dgrad + 4);  // the reverse sweep's segment index

    // ---- forward with checkpoints (every C steps) and snapshots (every `every` steps)
    ASD_TRY(c->adj1.ensure(sizeof(double) * n));
    double *S = c->adj1.as<double>();
    ASD_HIP(hipMemcpyAsync(S, d_S0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (nsteps == 0)  // the only snapshot is the initial state (nsnap == 1 here)
        ASD_HIP(hipMemcpyAsync(snaps, d_S0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    for (long a = 0, k = 0; a < nsteps || k == 0; a += C, k++) {
        ASD_HIP(hipMemcpyAsync(cps + k * n, S, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        if (a >= nsteps) break;
        const long len = std::min<long>(C, nsteps - a);
        const long s0 = a / every;  // first snapshot index of this segment
        const int nsn = (int)std::max<long>(0, std::min<long>(nsnap - s0, len / every + 1));
        ASD_TRY(asd_rk4_run(c, pp, L, 1, S, len, every, nsn, nsn > 0 ? snaps + s0 * n : nullptr, 0));
        if (a + len >= nsteps) break;
    }
    // ---- S(q,tau) in FP64, which leaves F_q(t) of all snapshots in c->fq; the reverse sweep
    // needs only F, so a gradient whose S(q,tau) nobody asked for skips the correlation GEMM
    const long m2 = 2L * nlags * nq;
    if (h_sqt || h_target || !h_g) {
        ASD_TRY(c->outdev.ensure(sizeof(double) * m2));
        c->sqt_generic_only = 1;
        const int rs = sqt_run(c, L, snaps, snaps, n, nsnap, T0, nlags, h_q, nq, ASD_FP64, c->outdev.as<double>(), -1);
        c->sqt_generic_only = 0;
        ASD_TRY(rs);
        if (h_sqt) ASD_HIP(hipMemcpyAsync(h_sqt, c->outdev.ptr, sizeof(double) * m2, hipMemcpyDeviceToHost, st));
    } else {
        ASD_TRY(sqt_fourier(c, L, snaps, n, nsnap, h_q, nq));
    }
    std::vector<double> gmis;
    if (h_target) {  // the misfit and its upstream gradient g = 2 (S - S_target)
        gmis.resize(m2);
        ASD_HIP(hipMemcpyAsync(gmis.data(), c->outdev.ptr, sizeof(double) * m2, hipMemcpyDeviceToHost, st));
        ASD_HIP(hipStreamSynchronize(st));
        double l = 0.0;
        for (long i = 0; i < m2; i++) {
            const double d = gmis[i] - h_target[i];
            l += d * d;
            gmis[i] = 2.0 * d;
        }
        *loss = l;
        h_g = gmis.data();
    }
    if (!h_g) {  // forward only
        ASD_HIP(hipStreamSynchronize(st));
        return ASD_OK;
    }
    // ---- adjoint of S(q,tau) -> dL/dX for every snapshot
    ASD_TRY(c->adj2.ensure(sizeof(double) * (2L * nlags * nq + 6L * nlags * nq + 6L * nsnap * nq)));
    double *gdev = c->adj2.as<double>();
    double *Fbar = gdev + 2L * nlags * nq;
    double *Lam = Fbar + 6L * nlags * nq;
    ASD_HIP(hipMemcpyAsync(gdev, h_g, sizeof(double) * 2 * (long)nlags * nq, hipMemcpyHostToDevice, st));
    const double *F = c->fq.as<double>();
    hipLaunchKernelGGL(fbar_kernel, dim3(nq, nlags), dim3(64), 0, st, F, nq, T0, nlags, Fbar);
    hipLaunchKernelGGL(lambda_kernel, dim3(nq, nsnap), dim3(256), 0, st, (const double *)gdev, F,
                       (const double *)Fbar, nq, T0, nlags, (int)ns, Lam);
    hipLaunchKernelGGL(dx_kernel, dim3((unsigned)((ns + 255) / 256), nsnap), dim3(256), 0, st,
                       (const double *)Lam, c->qdev.as<int>(), c->twiddle.as<double2>(), nq, L, dX);
    ASD_HIP(hipGetLastError());

    // ---- reverse sweep over checkpointed segments
    ASD_HIP(hipMemsetAsync(gacc, 0, sizeof(double) * 4L * nblk, st));
    ASD_HIP(hipMemsetAsync(lam, 0, sizeof(double) * n, st));
    if (nsteps % every == 0 && nsteps / every < nsnap)
        ASD_HIP(hipMemcpyAsync(lam, dX + (nsteps / every) * n, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    // One segment's work (the checkpoint at *dsg, the recomputed states, the reverse steps),
    // enqueued on c->stream; the same launches for every segment of `len` steps, since C is a
    // multiple of `every` (a snapshot falls on the same offsets in every segment).
    auto segment = [&](long len) -> int {
        hipStream_t s2 = c->stream;
        hipLaunchKernelGGL(copy_indexed_kernel, g1(n), dim3(256), 0, s2, n, seg, (const double *)cps,
                           (const int *)dsg);
        for (long m = 1; m < len; m++)
            ASD_TRY(rk4_launch_step(c, p, L, 1, seg + (m - 1) * n, seg + m * n, nullptr, nullptr, 0, 0));
        // lam and lamS alternate as the step's input and output; an odd segment ends with a copy,
        // so every segment starts and ends with dL/dS in lam
        double *lin = lam, *lout = lamS;
        for (long m = len - 1; m >= 0; m--) {
            const double *Sn = seg + m * n;
            ASD_TRY(rk4_stages_store(c, p, L, Sn, Y2, Y3, Y4, cc, scratch));
            // lin holds dL/dS_{n+1}; dL/dS_n goes to lout, with the snapshot adjoint of step
            // sg C + m added when that step is a snapshot step
            hipLaunchKernelGGL(adj_step_kernel, tiles, dim3(kAdjThreads), 0, s2, p, L, Sn, (const double *)Y2,
                               (const double *)Y3, (const double *)Y4, (const double *)cc, (const double *)lin, lout,
                               gacc, (const double *)dX, (const int *)dsg, C / every,
                               m % every == 0 ? (int)(m / every) : -1, nsnap);
            std::swap(lin, lout);
        }
        if (lin != lam) hipLaunchKernelGGL(copy_kernel, g1(n), dim3(256), 0, s2, n, lam, (const double *)lin);
        hipLaunchKernelGGL(step_index_kernel, dim3(1), dim3(1), 0, s2, dsg, -1);
        ASD_HIP(hipGetLastError());
        return ASD_OK;
    };
    const long nseg = (nsteps + C - 1) / C;
    hipLaunchKernelGGL(set_index_kernel, dim3(1), dim3(1), 0, st, dsg, (int)(nseg - 1));
    long nfull = nseg;
    if (nsteps % C) {  // the last segment is shorter: eagerly, first
        ASD_TRY(segment(nsteps % C));
        nfull--;
    }
    if (nfull >= 2) {
        // the full segments replay one captured graph (three small launches a step: the
        // recomputed state, the stage values, the reverse step; launch-bound)
        hipStream_t cap;
        ASD_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t graph;
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int r = ASD_OK;
        if (e == hipSuccess) {
            c->stream = cap;
            r = segment(C);
            c->stream = st;
            e = hipStreamEndCapture(cap, &graph);
        }
        if (e == hipSuccess) {  // the captured graph exists: instantiate it unless the capture failed
            if (r == ASD_OK) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
        }
        (void)hipStreamDestroy(cap);
        ASD_TRY(r);
        ASD_HIP(e);
        for (long sg = 0; sg < nfull && e == hipSuccess; sg++) e = hipGraphLaunch(exec, st);
        (void)hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        ASD_HIP(e);
    } else if (nfull == 1) {
        ASD_TRY(segment(C));
    }
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(1), dim3(64), 0, st, (const double *)gacc, nblk, dgrad);
    ASD_HIP(hipGetLastError());
    ASD_HIP(hipMemcpyAsync(grad, dgrad, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    ASD_HIP(hipStreamSynchronize(st));
    return ASD_OK;
}

extern "C" int asd_sqt_loss_grad(asd_ctx *c, const asd_params *p, int L, const double *d_S0, long nsteps,
                                 int snap_every, int T0, int nlags, const int *h_q, int nq, const double *h_target,
                                 int checkpoint_every, double *loss, double grad[4])
{
    ASD_CHECK(c && h_target && loss && grad, "NULL argument");
    // one forward pass (trajectory, checkpoints, S), the misfit and g = 2 (S - S_target) on
    // the host, then the reverse sweep over the same checkpoints
    const int saved = c->rk4_exact;
    c->rk4_exact = 1;
    const int r = sqt_vjp_exact(c, p, L, d_S0, nsteps, snap_every, T0, nlags, h_q, nq, nullptr, checkpoint_every,
                                grad, nullptr, h_target, loss);
    c->rk4_exact = saved;
    return r;
}
