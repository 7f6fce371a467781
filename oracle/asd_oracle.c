/*
 * asd_oracle.c -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY, see asd_oracle.h).
 *
 * Compiled with -ffp-contract=off so that every rounding step is the textbook
 * one written below; the operation order of the reference's euler/runge_kutta
 * is kept so that, at D = B_z = alpha = 0, results are bit-identical to the
 * reference built with the same flag (tests/test_oracle_vs_reference.py).
 */
#include "asd_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG --- */

uint64_t orc_hash64(uint64_t seed, uint64_t a, uint64_t b)
{
    /* linear key, then the splitmix64 finaliser (Steele, Lea, Flood 2014) */
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + a * 0xD1B54A32D192ED03ull +
                 b * 0x8CB92BA72F3D8DD7ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double orc_u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

void orc_init_spins(uint64_t seed, int nsites, double *S)
{
    /* Marsaglia (1972): (u,v) uniform in the unit disc, s = u^2+v^2,
     * S = (2u sqrt(1-s), 2v sqrt(1-s), 1-2s).  Site i draws from its own
     * counter stream (seed, i, 0), (seed, i, 1), ... so the draw is
     * order-independent.  Replaces the reference init (llg_sqt_fusion.c:195-216),
     * which divides integers and yields NaN. */
    for (int i = 0; i < nsites; i++) {
        double u, v, s;
        uint64_t ctr = 0;
        for (;;) {
            u = 2.0 * orc_u01(orc_hash64(seed, (uint64_t)i, ctr)) - 1.0;
            v = 2.0 * orc_u01(orc_hash64(seed, (uint64_t)i, ctr + 1)) - 1.0;
            ctr += 2;
            s = u * u + v * v;
            if (s < 1.0 && s > 0.0) break;
        }
        double r = 2.0 * sqrt(1.0 - s);
        S[3 * i + 0] = u * r;
        S[3 * i + 1] = v * r;
        S[3 * i + 2] = 1.0 - 2.0 * s;
    }
}

/* ------------------------------------------------------------- LLG RHS --- */

/* neighbour offsets (row, col) in the reference order, llg_sqt_fusion.c:181-184 */
static const int ORC_NB[4][2] = {{1, 0}, {-1, 0}, {0, -1}, {0, 1}};

/* f(S) at one site.  Follows euler (llg_sqt_fusion.c:49-89): exchange field
 * accumulated from 0.0 neighbour by neighbour as J*S_nb (:53, :74-76), then
 * p = h x S with v3d_cross's operand order (:82, v3d.h:48-53).  Extension:
 * h_z += 2 D S_z + B_z and Gilbert damping in Landau-Lifshitz form,
 * f = (p + (alpha/s) S x p) * (1 / (1 + alpha^2)), which is p exactly at alpha = 0
 * (the reciprocal is formed once; the GPU kernels use the same rounding). */
static void orc_site_rhs(const orc_params *p, int L, const double *S, long base, int i,
                         double *out)
{
    int row = i / L, col = i % L;
    const double *s = S + base + 3L * i;
    double hx = 0.0, hy = 0.0, hz = 0.0;
    for (int j = 0; j < 4; j++) {
        int nr = (row + ORC_NB[j][0] + L) % L;
        int nc = (col + ORC_NB[j][1] + L) % L;
        const double *t = S + base + 3L * (nr * L + nc);
        hx += (t[0] * p->J);
        hy += (t[1] * p->J);
        hz += (t[2] * p->J);
    }
    hz += 2.0 * p->D * s[2];
    hz += p->Bz;
    double px = hy * s[2] - hz * s[1];
    double py = hz * s[0] - hx * s[2];
    double pz = hx * s[1] - hy * s[0];
    double a = p->alpha / p->spin_length;
    double inv_den = 1.0 / (1.0 + p->alpha * p->alpha);
    double qx = s[1] * pz - s[2] * py;
    double qy = s[2] * px - s[0] * pz;
    double qz = s[0] * py - s[1] * px;
    out[0] = (px + a * qx) * inv_den;
    out[1] = (py + a * qy) * inv_den;
    out[2] = (pz + a * qz) * inv_den;
}

void orc_llg_rhs(const orc_params *p, int L, int nrep, const double *S, double *dS)
{
    long ns = (long)L * L;
    for (int r = 0; r < nrep; r++) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < (int)ns; i++)
            orc_site_rhs(p, L, S, r * ns * 3, i, dS + r * ns * 3 + 3L * i);
    }
}

/* runge_kutta (llg_sqt_fusion.c:98-145): k_s scaled by dt in place, stage inputs
 * spin + k1/2, spin + k2/2, spin + k3, and cur = spin + k1/6 + k2/3 + k3/3 + k4/6
 * summed left to right.  cur is not normalised. */
void orc_runge_kutta(const orc_params *p, int L, int nrep, const double *spin, double *cur,
                     double *k1, double *k2, double *k3, double *k4)
{
    long n = (long)nrep * L * L * 3;
    const double dt = p->dt;
    memcpy(cur, spin, n * sizeof(double));
    orc_llg_rhs(p, L, nrep, cur, k1);
    for (long j = 0; j < n; j++) {
        k1[j] *= dt;
        cur[j] = spin[j] + k1[j] / 2.0;
    }
    orc_llg_rhs(p, L, nrep, cur, k2);
    for (long j = 0; j < n; j++) {
        k2[j] *= dt;
        cur[j] = spin[j] + k2[j] / 2.0;
    }
    orc_llg_rhs(p, L, nrep, cur, k3);
    for (long j = 0; j < n; j++) {
        k3[j] *= dt;
        cur[j] = spin[j] + k3[j];
    }
    orc_llg_rhs(p, L, nrep, cur, k4);
    for (long j = 0; j < n; j++) {
        k4[j] *= dt;
        cur[j] = spin[j] + k1[j] / 6.0 + k2[j] / 3.0 + k3[j] / 3.0 + k4[j] / 6.0;
    }
}

/* renormalisation, llg_sqt_fusion.c:350-361 (len = 1.5 there; 1.0 in
 * baseline/llg_sqt_baseline.c:304-306 and gemm/llg_sqt_mkl.c:333-335) */
void orc_renormalise(long nspins, double len, const double *cur, double *S)
{
    for (long j = 0; j < nspins; j++) {
        double tx = cur[3 * j], ty = cur[3 * j + 1], tz = cur[3 * j + 2];
        double nrm = sqrt(tx * tx + ty * ty + tz * tz);
        S[3 * j] = len * tx / nrm;
        S[3 * j + 1] = len * ty / nrm;
        S[3 * j + 2] = len * tz / nrm;
    }
}

void orc_run(const orc_params *p, int L, int nrep, double *S, long nsteps, int snap_every,
             int nsnap, double *snaps)
{
    long n = (long)nrep * L * L * 3;
    double *cur = malloc(5 * n * sizeof(double));
    double *k1 = cur + n, *k2 = k1 + n, *k3 = k2 + n, *k4 = k3 + n;
    for (long s = 0; s <= nsteps; s++) {
        if (snaps && snap_every > 0 && s % snap_every == 0 && s / snap_every < nsnap)
            memcpy(snaps + (s / snap_every) * n, S, n * sizeof(double));
        if (s == nsteps) break;
        orc_runge_kutta(p, L, nrep, S, cur, k1, k2, k3, k4);
        orc_renormalise(n / 3, p->spin_length, cur, S);
    }
    free(cur);
}

double orc_energy(const orc_params *p, int L, const double *S)
{
    /* E = -J sum_<ij> S_i.S_j - D sum S_z^2 - B_z sum S_z (each bond once) */
    double e = 0.0;
    for (int r = 0; r < L; r++)
        for (int c = 0; c < L; c++) {
            const double *s = S + 3L * (r * L + c);
            const double *a = S + 3L * (((r + 1) % L) * L + c);
            const double *b = S + 3L * (r * L + (c + 1) % L);
            e -= p->J * (s[0] * a[0] + s[1] * a[1] + s[2] * a[2]);
            e -= p->J * (s[0] * b[0] + s[1] * b[1] + s[2] * b[2]);
            e -= p->D * s[2] * s[2] + p->Bz * s[2];
        }
    return e;
}

/* ----------------------------------------------------- fused_dgemm contract */

/* position of element (i, j) of a complex M x N matrix in the reference's
 * blocked layout: within each 4-row group 4 real then 4 imaginary values,
 * column stride ldc (llg_sqt_fusion.c:324-327, gemm_kernel.h:39-46). */
static long orc_blk(int i, int j, int ldc) { return (long)(i / 4) * 8 + i % 4 + (long)j * ldc; }

void orc_fused_dgemm(int m, int n, int k, double alpha, const double *a, int lda,
                     const double *b, int ldb, double *c, int ldc, const double *fuse_c,
                     double *sum)
{
    /* gemm_kernel.h:355-422: alpha is applied to A while packing (:403, :232-247),
     * the epilogue does C_re += g Q_re, C_im += g Q_im, stores C (STORE_MATRIX, :4)
     * and accumulates the updated C into sum (:59-88).  Restated for every K, M, N
     * (the reference double-counts for K > 240 and drops M%4 rows / N%2 columns). */
    double sre = 0.0, sim = 0.0;
#pragma omp parallel for reduction(+ : sre, sim) schedule(static)
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) {
            double g = 0.0;
            for (int kk = 0; kk < k; kk++)
                g += (alpha * a[i + (long)kk * lda]) * b[kk + (long)j * ldb];
            long pr = orc_blk(i, j, ldc), pi = pr + 4;
            c[pr] += g * fuse_c[pr];
            c[pi] += g * fuse_c[pi];
            sre += c[pr];
            sim += c[pi];
        }
    }
    sum[0] += sre;
    sum[1] += sim;
}

void orc_build_qmat(int L, int mx, int my, double *q)
{
    /* llg_sqt_fusion.c:290-333: Q_ij = cexp(-i (qx (r_i - r_j) + qy (c_i - c_j))),
     * qx = 2 pi vx / N pairs with the row offset, qy with the column offset. */
    int ns = L * L, ldq = 2 * ((ns + 3) / 4 * 4); /* = 2 Ns when Ns % 4 == 0 */
    double qx = 2.0 * M_PI * mx / (double)L;
    double qy = 2.0 * M_PI * my / (double)L;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ns; j++) {
        int rj = j / L, cj = j % L;
        for (int i = 0; i < ns; i++) {
            int ri = i / L, ci = i % L;
            double dffr = ri - rj, dffc = ci - cj;
            double complex t = cexp(CMPLX(0.0, -1.0) * (qx * dffr + qy * dffc));
            long pr = orc_blk(i, j, ldq);
            q[pr] = creal(t);
            q[pr + 4] = cimag(t);
        }
    }
}

/* ------------------------------------------------------------- S(q,tau) --- */

void orc_sqt_gemm(int L, const double *snaps, int T0, int nlags, int nq, const int *qlist,
                  double *out)
{
    /* Restates the reference S(q,t) data flow with time origins in the role of
     * replicas (llg_sqt_fusion.c:235-268 B operand and its average, :370-398
     * A operand and its average, :402-426 disconnected prefill, :429 fused GEMM,
     * :432-442 reduction and 1/N^2).  G is formed once per lag and reused for
     * every q, which is what the reference would compute per q. */
    const int ns = L * L, K = 3 * T0;
    const double alpha = 1.0 / (double)T0;
    double *B = malloc((size_t)K * ns * sizeof(double));
    double *A = malloc((size_t)K * ns * sizeof(double));
    double *zbar = malloc((size_t)ns * 3 * sizeof(double));
    double *fbar = malloc((size_t)ns * 3 * sizeof(double));
    double *G = malloc((size_t)ns * ns * sizeof(double));

    /* spin_zero: B[k + j*K], k = t0*3 + c  (:240-242) */
    for (int j = 0; j < ns; j++)
        for (int t0 = 0; t0 < T0; t0++)
            for (int c = 0; c < 3; c++)
                B[(t0 * 3 + c) + (long)j * K] = snaps[((long)t0 * ns + j) * 3 + c];
    /* zerot_avg (:249-268) */
    for (int j = 0; j < ns; j++)
        for (int c = 0; c < 3; c++) {
            double t = 0.0;
            for (int t0 = 0; t0 < T0; t0++) t += B[(t0 * 3 + c) + (long)j * K];
            zbar[3 * j + c] = t / (double)T0;
        }

    for (int lag = 0; lag < nlags; lag++) {
        /* spint_mkl: A[i + k*ns] (:374-376) and fint_avg (:380-398) */
        for (int t0 = 0; t0 < T0; t0++)
            for (int i = 0; i < ns; i++)
                for (int c = 0; c < 3; c++)
                    A[i + (long)(t0 * 3 + c) * ns] = snaps[((long)(t0 + lag) * ns + i) * 3 + c];
        for (int i = 0; i < ns; i++)
            for (int c = 0; c < 3; c++) {
                double t = 0.0;
                for (int t0 = 0; t0 < T0; t0++) t += A[i + (long)(t0 * 3 + c) * ns];
                fbar[3 * i + c] = t / (double)T0;
            }
        /* G = (alpha A) B */
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ns; j++)
            for (int i = 0; i < ns; i++) {
                double g = 0.0;
                for (int kk = 0; kk < K; kk++) g += (alpha * A[i + (long)kk * ns]) * B[kk + (long)j * K];
                G[(long)j * ns + i] = g;
            }
        for (int iq = 0; iq < nq; iq++) {
            double qx = 2.0 * M_PI * qlist[2 * iq] / (double)L;
            double qy = 2.0 * M_PI * qlist[2 * iq + 1] / (double)L;
            double sre = 0.0, sim = 0.0;
#pragma omp parallel for reduction(+ : sre, sim) schedule(static)
            for (int j = 0; j < ns; j++) {
                int rj = j / L, cj = j % L;
                double zx = zbar[3 * j], zy = zbar[3 * j + 1], zz = zbar[3 * j + 2];
                for (int i = 0; i < ns; i++) {
                    int ri = i / L, ci = i % L;
                    double dffr = ri - rj, dffc = ci - cj;
                    double complex e = cexp(CMPLX(0.0, -1.0) * (qx * dffr + qy * dffc));
                    double tx = fbar[3 * i], ty = fbar[3 * i + 1], tz = fbar[3 * i + 2];
                    double tmp = zx * tx + zy * ty + zz * tz;
                    tmp *= -1.0;
                    double cre = tmp * creal(e), cim = tmp * cimag(e);
                    double g = G[(long)j * ns + i];
                    cre += g * creal(e);
                    cim += g * cimag(e);
                    sre += cre;
                    sim += cim;
                }
            }
            out[2 * ((long)lag * nq + iq)] = sre / (double)ns;
            out[2 * ((long)lag * nq + iq) + 1] = sim / (double)ns;
        }
    }
    free(B); free(A); free(zbar); free(fbar); free(G);
}

/* F_q(t) = sum_i S_i(t) exp(-i q.r_i), for t in [0, nt) -> F[(t*nq + iq)*3 + c] */
static void orc_fourier(int L, const double *snaps, int nt, int nq, const int *qlist,
                        double complex *F)
{
    const int ns = L * L;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; t++)
        for (int iq = 0; iq < nq; iq++) {
            double qx = 2.0 * M_PI * qlist[2 * iq] / (double)L;
            double qy = 2.0 * M_PI * qlist[2 * iq + 1] / (double)L;
            double complex f0 = 0, f1 = 0, f2 = 0;
            for (int i = 0; i < ns; i++) {
                double complex e = cexp(CMPLX(0.0, -1.0) * (qx * (i / L) + qy * (i % L)));
                const double *s = snaps + ((long)t * ns + i) * 3;
                f0 += s[0] * e; f1 += s[1] * e; f2 += s[2] * e;
            }
            double complex *o = F + ((long)t * nq + iq) * 3;
            o[0] = f0; o[1] = f1; o[2] = f2;
        }
}

void orc_sqt_separable(int L, const double *snaps, int T0, int nlags, int nq,
                       const int *qlist, double *out)
{
    /* S(q,tau) = (1/Ns) [ (1/T0) sum_t0 F_q(t0+tau).conj F_q(t0)
     *                     - Fbar_q(tau).conj Fbar_q(0) ],  Fbar_q(tau) = (1/T0) sum_t0 F_q(t0+tau).
     * Exactly equal to the GEMM form because e^{-iq(r_i-r_j)} factorises. */
    const int ns = L * L, nt = T0 + nlags - 1;
    double complex *F = malloc((size_t)nt * nq * 3 * sizeof(double complex));
    orc_fourier(L, snaps, nt, nq, qlist, F);
    for (int lag = 0; lag < nlags; lag++)
        for (int iq = 0; iq < nq; iq++) {
            double complex corr = 0, fb[3] = {0, 0, 0}, zb[3] = {0, 0, 0};
            for (int t0 = 0; t0 < T0; t0++) {
                const double complex *a = F + ((long)(t0 + lag) * nq + iq) * 3;
                const double complex *b = F + ((long)t0 * nq + iq) * 3;
                for (int c = 0; c < 3; c++) {
                    corr += a[c] * conj(b[c]);
                    fb[c] += a[c];
                    zb[c] += b[c];
                }
            }
            corr /= (double)T0;
            double complex disc = 0;
            for (int c = 0; c < 3; c++) disc += (fb[c] / (double)T0) * conj(zb[c] / (double)T0);
            double complex s = (corr - disc) / (double)ns;
            out[2 * ((long)lag * nq + iq)] = creal(s);
            out[2 * ((long)lag * nq + iq) + 1] = cimag(s);
        }
    free(F);
}

/* ---------------------------------------------------- tangent-linear model */

/* f and its directional derivative at one site for the 4 parameter directions.
 * dS holds 4 tangent fields [ip][rep-less site][3]; df likewise. */
static void orc_site_rhs_tlm(const orc_params *p, int L, const double *S, const double *dS,
                             long ns, int i, double *f, double *df)
{
    int row = i / L, col = i % L;
    const double *s = S + 3L * i;
    /* f itself with orc_site_rhs's rounding steps, so the linearisation runs along
     * exactly the trajectory orc_run produces (in a chaotic run a 1-ulp difference
     * in the base trajectory grows into a visible change of the derivative) */
    double hx = 0.0, hy = 0.0, hz = 0.0, sx = 0, sy = 0, sz = 0; /* h and sum of neighbours */
    for (int j = 0; j < 4; j++) {
        int nr = (row + ORC_NB[j][0] + L) % L, nc = (col + ORC_NB[j][1] + L) % L;
        const double *t = S + 3L * (nr * L + nc);
        hx += (t[0] * p->J); hy += (t[1] * p->J); hz += (t[2] * p->J);
        sx += t[0]; sy += t[1]; sz += t[2];
    }
    hz += 2.0 * p->D * s[2];
    hz += p->Bz;
    double px = hy * s[2] - hz * s[1], py = hz * s[0] - hx * s[2], pz = hx * s[1] - hy * s[0];
    double a = p->alpha / p->spin_length, inv_den = 1.0 / (1.0 + p->alpha * p->alpha);
    double qx = s[1] * pz - s[2] * py, qy = s[2] * px - s[0] * pz, qz = s[0] * py - s[1] * px;
    f[0] = (px + a * qx) * inv_den; f[1] = (py + a * qy) * inv_den; f[2] = (pz + a * qz) * inv_den;
    for (int ip = 0; ip < 4; ip++) {
        const double *d = dS + ip * ns * 3;
        const double *ds = d + 3L * i;
        double nx = 0, ny = 0, nz = 0;
        for (int j = 0; j < 4; j++) {
            int nr = (row + ORC_NB[j][0] + L) % L, nc = (col + ORC_NB[j][1] + L) % L;
            const double *t = d + 3L * (nr * L + nc);
            nx += t[0]; ny += t[1]; nz += t[2];
        }
        double dJ = ip == 0, dD = ip == 1, dB = ip == 2, dA = ip == 3;
        double dhx = dJ * sx + p->J * nx;
        double dhy = dJ * sy + p->J * ny;
        double dhz = dJ * sz + p->J * nz + 2.0 * dD * s[2] + 2.0 * p->D * ds[2] + dB;
        double dpx = dhy * s[2] - dhz * s[1] + hy * ds[2] - hz * ds[1];
        double dpy = dhz * s[0] - dhx * s[2] + hz * ds[0] - hx * ds[2];
        double dpz = dhx * s[1] - dhy * s[0] + hx * ds[1] - hy * ds[0];
        double dqx = ds[1] * pz - ds[2] * py + s[1] * dpz - s[2] * dpy;
        double dqy = ds[2] * px - ds[0] * pz + s[2] * dpx - s[0] * dpz;
        double dqz = ds[0] * py - ds[1] * px + s[0] * dpy - s[1] * dpx;
        double da = dA / p->spin_length, dden = 2.0 * p->alpha * dA;
        double *o = df + ip * ns * 3 + 3L * i;
        /* d[(p + a q) / den] = (dp + da q + a dq) / den - f dden / den */
        o[0] = (dpx + da * qx + a * dqx) * inv_den - f[0] * dden * inv_den;
        o[1] = (dpy + da * qy + a * dqy) * inv_den - f[1] * dden * inv_den;
        o[2] = (dpz + da * qz + a * dqz) * inv_den - f[2] * dden * inv_den;
    }
}

void orc_tlm_run(const orc_params *p, int L, const double *S0, long nsteps, int snap_every,
                 int nsnap, double *snaps, double *dsnaps)
{
    const long ns = (long)L * L, n = ns * 3;
    double *S = malloc(n * sizeof(double)), *dS = calloc(4 * n, sizeof(double));
    double *Y = malloc(n * sizeof(double)), *dY = malloc(4 * n * sizeof(double));
    double *k = malloc(4 * n * sizeof(double)), *dk = malloc(16 * n * sizeof(double));
    double *f = malloc(n * sizeof(double)), *df = malloc(4 * n * sizeof(double));
    memcpy(S, S0, n * sizeof(double));
    static const double cst[4] = {0.5, 0.5, 1.0, 0.0};
    for (long st = 0; st <= nsteps; st++) {
        if (st % snap_every == 0 && st / snap_every < nsnap) {
            long si = st / snap_every;
            memcpy(snaps + si * n, S, n * sizeof(double));
            for (int ip = 0; ip < 4; ip++)
                memcpy(dsnaps + ((long)ip * nsnap + si) * n, dS + ip * n, n * sizeof(double));
        }
        if (st == nsteps) break;
        memcpy(Y, S, n * sizeof(double));
        memcpy(dY, dS, 4 * n * sizeof(double));
        for (int stg = 0; stg < 4; stg++) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < (int)ns; i++) orc_site_rhs_tlm(p, L, Y, dY, ns, i, f + 3L * i, df);
            for (long j = 0; j < n; j++) k[stg * n + j] = p->dt * f[j];
            for (long j = 0; j < 4 * n; j++) dk[stg * 4 * n + j] = p->dt * df[j];
            if (stg < 3) {
                for (long j = 0; j < n; j++) Y[j] = S[j] + cst[stg] * k[stg * n + j];
                for (int ip = 0; ip < 4; ip++)
                    for (long j = 0; j < n; j++)
                        dY[ip * n + j] = dS[ip * n + j] + cst[stg] * dk[(stg * 4 + ip) * n + j];
            }
        }
        /* c = S + k1/6 + k2/3 + k3/3 + k4/6, then S = len c/|c| */
        for (long s = 0; s < ns; s++) {
            double c[3], dc[4][3];
            for (int d = 0; d < 3; d++) {
                long j = 3 * s + d;
                c[d] = S[j] + k[j] / 6.0 + k[n + j] / 3.0 + k[2 * n + j] / 3.0 + k[3 * n + j] / 6.0;
                for (int ip = 0; ip < 4; ip++)
                    dc[ip][d] = dS[ip * n + j] + dk[(0 * 4 + ip) * n + j] / 6.0 +
                                dk[(1 * 4 + ip) * n + j] / 3.0 + dk[(2 * 4 + ip) * n + j] / 3.0 +
                                dk[(3 * 4 + ip) * n + j] / 6.0;
            }
            double nrm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (int ip = 0; ip < 4; ip++) {
                double cd = c[0] * dc[ip][0] + c[1] * dc[ip][1] + c[2] * dc[ip][2];
                for (int d = 0; d < 3; d++)
                    dS[ip * n + 3 * s + d] =
                        p->spin_length * (dc[ip][d] / nrm - c[d] * cd / (nrm * nrm * nrm));
            }
            for (int d = 0; d < 3; d++) S[3 * s + d] = p->spin_length * c[d] / nrm;
        }
    }
    free(S); free(dS); free(Y); free(dY); free(k); free(dk); free(f); free(df);
}

void orc_tlm_sqt(int L, const double *snaps, const double *dsnaps, int T0, int nlags, int nq,
                 const int *qlist, double *dout)
{
    /* derivative of the separable form (identical to the GEMM form's derivative,
     * since the two are the same bilinear function of the snapshots) */
    const int ns = L * L, nt = T0 + nlags - 1;
    double complex *F = malloc((size_t)nt * nq * 3 * sizeof(double complex));
    double complex *dF = malloc((size_t)nt * nq * 3 * sizeof(double complex));
    orc_fourier(L, snaps, nt, nq, qlist, F);
    orc_fourier(L, dsnaps, nt, nq, qlist, dF);
    for (int lag = 0; lag < nlags; lag++)
        for (int iq = 0; iq < nq; iq++) {
            double complex dcorr = 0, fb[3] = {0}, zb[3] = {0}, dfb[3] = {0}, dzb[3] = {0};
            for (int t0 = 0; t0 < T0; t0++) {
                const double complex *a = F + ((long)(t0 + lag) * nq + iq) * 3;
                const double complex *b = F + ((long)t0 * nq + iq) * 3;
                const double complex *da = dF + ((long)(t0 + lag) * nq + iq) * 3;
                const double complex *db = dF + ((long)t0 * nq + iq) * 3;
                for (int c = 0; c < 3; c++) {
                    dcorr += da[c] * conj(b[c]) + a[c] * conj(db[c]);
                    fb[c] += a[c]; zb[c] += b[c]; dfb[c] += da[c]; dzb[c] += db[c];
                }
            }
            dcorr /= (double)T0;
            double complex ddisc = 0;
            for (int c = 0; c < 3; c++)
                ddisc += (dfb[c] * conj(zb[c]) + fb[c] * conj(dzb[c])) / ((double)T0 * T0);
            double complex s = (dcorr - ddisc) / (double)ns;
            dout[2 * ((long)lag * nq + iq)] = creal(s);
            dout[2 * ((long)lag * nq + iq) + 1] = cimag(s);
        }
    free(F); free(dF);
}
