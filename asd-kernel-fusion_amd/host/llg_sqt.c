/*
 * llg_sqt.c -- C host driver for libasd (no HIP headers: only include/asd.h).
 *
 * Mode "reference" (default) mirrors the driver of kernel_fusion/llg_sqt_fusion.c
 * (main, :149-477): the same getopt flags -n L -m MCsweep -f Nfile -e dt -t trd
 * -x vx -y vy (vx = vy = L/10 unless -Q, as :178-179 overwrite them; -t is read and
 * ignored as there), Nfile replicas integrated together, 60 iterations of
 * { RK4 step + renormalisation ; S(q,t) at one q over the replicas }, the two stages
 * timed for iterations 20..59 and the line "L tlle tsqt" (seconds per iteration)
 * appended to the output file (:451-458).  Deviations, each a fix of a reference
 * defect: spins are Marsaglia-uniform from a seed (-s) instead of the NaN init of
 * :195-216; the spin length is 1.0 unless -l says otherwise (the reference uses
 * 1.5, :356-358); S(q,t) values are written (-O) instead of being discarded.
 *
 * Mode "pipeline" (-P) runs the time-origin form on one trajectory: RK4 for
 * -S steps with a snapshot every -E steps, then S(q,tau) for -G lags over -T
 * origins on the Gamma-X-M-Gamma path (or the full grid with -F) in precision -p,
 * and writes S(q,tau) as raw little-endian doubles [lag][q][re,im] to -O plus a
 * JSON header next to it (<file>.json).  -W file writes the snapshots the same way
 * (doubles [snapshot][site][x,y,z], site = row * L + col, plus <file>.json); -R file
 * reads such a file instead of integrating (its size must match -n, -S and -E), so
 * S(q,tau) can be recomputed from stored trajectories, on another host or precision.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "asd.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

#define CHECK(call)                                                                       \
    do {                                                                                  \
        int rc_ = (call);                                                                 \
        if (rc_ != ASD_OK) {                                                              \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, asd_last_error());        \
            exit(1);                                                                      \
        }                                                                                 \
    } while (0)

static int parse_prec(const char *s)
{
    if (!strcmp(s, "fp64")) return ASD_FP64;
    if (!strcmp(s, "fp32")) return ASD_FP32;
    if (!strcmp(s, "bf16x3")) return ASD_BF16X3;
    if (!strcmp(s, "fp16x2")) return ASD_FP16X2;
    if (!strcmp(s, "fp16f8")) return ASD_FP16F8;
    fprintf(stderr, "unknown precision %s\n", s);
    exit(1);
}

static const char *prec_name(int p)
{
    static const char *n[] = {"fp64", "fp32", "bf16x3", "fp16x2", "fp16f8"};
    return n[p];
}

int main(int argc, char **argv)
{
    /* reference defaults, llg_sqt_fusion.c:13-20 */
    int N = 100, MCsweep = 500, Nfile = 50, trd = 1, vx = 5, vy = 5;
    double dt = 0.02, J1 = 1.0;
    /* extensions */
    int keep_q = 0, pipeline = 0, full_grid = 0, prec = ASD_FP64;
    long steps = 8000;
    int every = 10, T0 = 400, nlags = 400;
    double D = 0.0, Bz = 0.0, alpha = 0.0, len = 1.0;
    uint64_t seed = 0;
    const char *out_timing = "mi355x_fusion_sqt.txt", *out_sqt = NULL, *out_snaps = NULL, *in_snaps = NULL;
    int ch;
    while ((ch = getopt(argc, argv, "n:m:f:e:t:x:y:QPFS:E:T:G:p:J:D:B:a:l:s:o:O:W:R:h")) != -1) {
        switch (ch) {
        case 'n': N = atoi(optarg); break;
        case 'm': MCsweep = atoi(optarg); break;
        case 'f': Nfile = atoi(optarg); break;
        case 'e': dt = atof(optarg); break;
        case 't': trd = atoi(optarg); break;
        case 'x': vx = atoi(optarg); break;
        case 'y': vy = atoi(optarg); break;
        case 'Q': keep_q = 1; break;
        case 'P': pipeline = 1; break;
        case 'F': full_grid = 1; break;
        case 'S': steps = atol(optarg); break;
        case 'E': every = atoi(optarg); break;
        case 'T': T0 = atoi(optarg); break;
        case 'G': nlags = atoi(optarg); break;
        case 'p': prec = parse_prec(optarg); break;
        case 'J': J1 = atof(optarg); break;
        case 'D': D = atof(optarg); break;
        case 'B': Bz = atof(optarg); break;
        case 'a': alpha = atof(optarg); break;
        case 'l': len = atof(optarg); break;
        case 's': seed = strtoull(optarg, NULL, 10); break;
        case 'o': out_timing = optarg; break;
        case 'O': out_sqt = optarg; break;
        case 'W': out_snaps = optarg; break;
        case 'R': in_snaps = optarg; break;
        default:
            fprintf(stderr,
                    "usage: %s [-n L -m MCsweep -f Nfile -e dt -t trd -x vx -y vy -Q]  (reference mode)\n"
                    "       %s -P [-n L -S steps -E every -T T0 -G lags -p fp64|fp32|bf16x3|fp16x2|fp16f8 -F\n"
                    "              -W snapshot_file | -R snapshot_file]\n"
                    "       common: -J J -D D -B Bz -a alpha -l spin_length -s seed -o timing_file -O sqt_file\n",
                    argv[0], argv[0]);
            return 2;
        }
    }
    (void)trd;
    if (!keep_q) { vx = N / 10; vy = N / 10; } /* llg_sqt_fusion.c:178-179 */
    asd_params p = {J1, D, Bz, alpha, dt, len};
    asd_ctx *ctx;
    CHECK(asd_create(&ctx, 0));
    const long ns = (long)N * N;

    if (!pipeline) {
        const int iters = 60, nq = 1;
        if (MCsweep < iters) MCsweep = iters; /* sqt[num] for num < 60 (:272, :340) */
        const size_t bytes = sizeof(double) * 3 * ns * Nfile;
        double *h = malloc(bytes);
        for (int r = 0; r < Nfile; r++) CHECK(asd_init_spins(seed + r, (int)ns, h + 3 * ns * r));
        if (len != 1.0)
            for (size_t i = 0; i < 3 * ns * Nfile; i++) h[i] *= len;
        double *d_spins, *d_zero;
        CHECK(asd_malloc(ctx, bytes, (void **)&d_spins));
        CHECK(asd_malloc(ctx, bytes, (void **)&d_zero));
        CHECK(asd_memcpy(ctx, d_spins, h, bytes, 0));
        CHECK(asd_memcpy(ctx, d_zero, h, bytes, 0));
        int q[2] = {vx % N, vy % N};
        double *sqt = calloc(2 * (size_t)MCsweep, sizeof(double));
        double tlle = 0, tsqt = 0;
        printf("N: %d, Nfile: %d\n", N, Nfile);
        for (int num = 0; num < iters; num++) {
            if (num % 5 == 0) printf("Step: %d\n", num);
            double t1 = now_s();
            CHECK(asd_rk4_run(ctx, &p, N, Nfile, d_spins, 1, 1, 0, NULL, 0));
            CHECK(asd_synchronize(ctx));
            double t2 = now_s();
            if (num >= 20) tlle += t2 - t1;
            t1 = now_s();
            CHECK(asd_sqt_replicas(ctx, N, Nfile, d_spins, d_zero, q, nq, (asd_precision)prec, sqt + 2 * num));
            t2 = now_s();
            if (num >= 20) tsqt += t2 - t1;
        }
        tlle /= 40.0;
        tsqt /= 40.0;
        FILE *fp = fopen(out_timing, "a");
        if (!fp) { perror(out_timing); return 1; }
        fprintf(fp, "%d %lf %lf\n", N, tlle, tsqt);
        fclose(fp);
        if (out_sqt) {
            FILE *fo = fopen(out_sqt, "w");
            if (!fo) { perror(out_sqt); return 1; }
            for (int num = 0; num < iters; num++) fprintf(fo, "%d %.17g %.17g\n", num, sqt[2 * num], sqt[2 * num + 1]);
            fclose(fo);
        }
        free(sqt);
        free(h);
        CHECK(asd_free(ctx, d_spins));
        CHECK(asd_free(ctx, d_zero));
    } else {
        const int nsnap = (int)(steps / every);
        if (nsnap < T0 + nlags - 1) {
            fprintf(stderr, "need steps/every >= T0 + lags - 1\n");
            return 1;
        }
        int nq = full_grid ? (int)ns : 3 * N / 2;
        int *q = malloc(sizeof(int) * 2 * nq);
        if (full_grid) {
            for (int a = 0; a < N; a++)
                for (int b = 0; b < N; b++) { q[2 * (a * N + b)] = a; q[2 * (a * N + b) + 1] = b; }
        } else if (asd_qpath_gxmg(N, q) != nq) {
            fprintf(stderr, "Gamma-X-M-Gamma needs even L\n");
            return 1;
        }
        double *h = malloc(sizeof(double) * 3 * ns);
        CHECK(asd_init_spins(seed, (int)ns, h));
        double *d_spins, *d_snaps;
        CHECK(asd_malloc(ctx, sizeof(double) * 3 * ns, (void **)&d_spins));
        CHECK(asd_malloc(ctx, sizeof(double) * 3 * ns * nsnap, (void **)&d_snaps));
        CHECK(asd_memcpy(ctx, d_spins, h, sizeof(double) * 3 * ns, 0));
        double *out = malloc(sizeof(double) * 2 * nq * (size_t)nlags);
        const size_t snap_doubles = 3 * (size_t)ns * nsnap;
        double t0 = now_s();
        if (in_snaps) {
            /* stored snapshots instead of the integration */
            FILE *fi = fopen(in_snaps, "rb");
            if (!fi) { perror(in_snaps); return 1; }
            double *hs = malloc(sizeof(double) * snap_doubles);
            const size_t got = fread(hs, sizeof(double), snap_doubles, fi);
            const int extra = fgetc(fi) != EOF;
            fclose(fi);
            if (got != snap_doubles || extra) {
                fprintf(stderr, "%s: expected %zu doubles (L=%d, %d snapshots)\n", in_snaps, snap_doubles, N, nsnap);
                return 1;
            }
            CHECK(asd_memcpy(ctx, d_snaps, hs, sizeof(double) * snap_doubles, 0));
            free(hs);
        } else {
            CHECK(asd_rk4_run(ctx, &p, N, 1, d_spins, steps, every, nsnap, d_snaps, 1));
        }
        CHECK(asd_synchronize(ctx));
        double t1 = now_s();
        if (out_snaps) {
            double *hs = malloc(sizeof(double) * snap_doubles);
            CHECK(asd_memcpy(ctx, hs, d_snaps, sizeof(double) * snap_doubles, 1));
            FILE *fo = fopen(out_snaps, "wb");
            if (!fo) { perror(out_snaps); return 1; }
            if (fwrite(hs, sizeof(double), snap_doubles, fo) != snap_doubles) { perror(out_snaps); return 1; }
            fclose(fo);
            free(hs);
            char hdr[4096];
            snprintf(hdr, sizeof hdr, "%s.json", out_snaps);
            FILE *fj = fopen(hdr, "w");
            if (!fj) { perror(hdr); return 1; }
            fprintf(fj,
                    "{\"format\": \"float64 little-endian [snapshot][site][x,y,z], site = row * L + col\", "
                    "\"L\": %d, \"snapshots\": %d, \"snap_every\": %d, \"steps\": %ld, \"dt\": %.17g, "
                    "\"J\": %.17g, \"D\": %.17g, \"Bz\": %.17g, \"alpha\": %.17g, \"spin_length\": %.17g, "
                    "\"seed\": %llu, \"source\": \"%s\"}\n",
                    N, nsnap, every, steps, dt, J1, D, Bz, alpha, len, (unsigned long long)seed,
                    in_snaps ? "read from file" : "asd_rk4_run");
            fclose(fj);
        }
        const double t1w = now_s();
        CHECK(asd_sqt(ctx, N, d_snaps, nsnap, T0, nlags, q, nq, (asd_precision)prec, out, 0));
        double t2 = now_s();
        if (in_snaps)
            printf("snapshots read %.4f s   ", t1 - t0);
        else
            printf("rk4 %.4f s (%.3e spin-steps/s)   ", t1 - t0, steps * (double)ns / (t1 - t0));
        printf("S(q,tau) %.4f s (%.2f TFLOP/s GEMM-equivalent, %s)\n", t2 - t1w,
               nlags * 2.0 * ns * ns * 3.0 * T0 / (t2 - t1w) / 1e12, prec_name(prec));
        if (out_sqt) {
            FILE *fo = fopen(out_sqt, "wb");
            if (!fo) { perror(out_sqt); return 1; }
            fwrite(out, sizeof(double), 2 * (size_t)nq * nlags, fo);
            fclose(fo);
            char hdr[4096];
            snprintf(hdr, sizeof hdr, "%s.json", out_sqt);
            FILE *fj = fopen(hdr, "w");
            if (!fj) { perror(hdr); return 1; }
            fprintf(fj,
                    "{\"format\": \"float64 little-endian [lag][q][re,im]\", \"L\": %d, \"lags\": %d, \"T0\": %d, "
                    "\"nq\": %d, \"qset\": \"%s\", \"precision\": \"%s\", \"steps\": %ld, \"snap_every\": %d, "
                    "\"dt\": %.17g, \"J\": %.17g, \"D\": %.17g, \"Bz\": %.17g, \"alpha\": %.17g, "
                    "\"spin_length\": %.17g, \"seed\": %llu, \"q\": [",
                    N, nlags, T0, nq, full_grid ? "full" : "gamma-x-m-gamma", prec_name(prec), steps, every, dt, J1,
                    D, Bz, alpha, len, (unsigned long long)seed);
            for (int i = 0; i < nq; i++) fprintf(fj, "%s[%d,%d]", i ? "," : "", q[2 * i], q[2 * i + 1]);
            fprintf(fj, "]}\n");
            fclose(fj);
        }
        free(out);
        free(h);
        free(q);
        CHECK(asd_free(ctx, d_spins));
        CHECK(asd_free(ctx, d_snaps));
    }
    CHECK(asd_destroy(ctx));
    return 0;
}
