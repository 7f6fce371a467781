/* Test instrumentation for the drop-in build of the reference driver (test infrastructure).
 *
 * oracle/Makefile links this file into a second drop-in build of
 * kernel_fusion/llg_sqt_fusion.c (the edit of INTEGRATION.md section 1, linked to libasd):
 *
 *   -Drand=asd_spy_rand                         the driver's random source, see below
 *   -Wl,--wrap=runge_kutta -Wl,--wrap=fused_dgemm
 *
 * so that the driver's own calls (llg_sqt_fusion.c:348 runge_kutta, :429 fused_dgemm) go
 * through the wrappers here, which call libasd's functions (__real_*) and write what went
 * in and came out to files under $ASD_SPY_DIR.  tests/test_gpu_dropin.py then checks those
 * values against the CPU oracle and the reference's own functions.
 *
 * Why the random source is replaced: the driver draws each spin component as
 * rand()/RAND_MAX, an integer division (llg_sqt_fusion.c:204-206), which is 0 for every
 * draw below RAND_MAX, so its spins are 0/0 = NaN and nothing it computes can be checked.
 * asd_spy_rand returns -RAND_MAX, 0 or RAND_MAX from a fixed-seed generator, so the same
 * expression gives components in {-1, 0, 1} (never all three zero): 26 directions, a
 * deterministic, non-trivial initial state that the driver normalises itself.
 *
 * Files (raw native doubles / ints):
 *   rk_first.bin   spin (in), cur, k1, k2, k3, k4 (out) of the first runge_kutta call
 *   rk_last.bin    the same for the most recent call (the 60th when the run ends)
 *   rk_calls.txt   number of runge_kutta calls
 *   fd_first.bin   int m, n, k, lda, ldb, ldc; double alpha; a; b; c (in); fuse_C;
 *                  c (out); sum[0], sum[1] added by the call
 *   fd_sums.bin    per call: the (re, im) the call added to sum[0], sum[1]
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>

extern int N, Nfile;

void __real_runge_kutta(double *spin, double *cur, double *k1, double *k2, double *k3, double *k4);
void __real_fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda, double *b, int *ldb,
                        double *c, int *ldc, double *fuse_C, double *sum);

/* -RAND_MAX, 0 or RAND_MAX; the third draw of a triple is nonzero when the first two
 * were zero.  64-bit LCG (Knuth's MMIX constants), top bits used. */
static unsigned long long spy_state = 0x2545F4914F6CDD1DULL;
static int spy_pos, spy_zeros;
int asd_spy_rand(void)
{
    spy_state = spy_state * 6364136223846793005ULL + 1442695040888963407ULL;
    int v = (int)((spy_state >> 33) % 3) - 1; /* -1, 0, 1 */
    if (spy_pos == 2 && spy_zeros == 2 && v == 0) v = (spy_state >> 40) & 1 ? 1 : -1;
    spy_zeros = (spy_pos == 0 ? 0 : spy_zeros) + (v == 0);
    spy_pos = (spy_pos + 1) % 3;
    return v * RAND_MAX;
}

static FILE *spy_open(const char *name, const char *mode)
{
    const char *dir = getenv("ASD_SPY_DIR");
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir ? dir : ".", name);
    FILE *f = fopen(path, mode);
    if (!f) {
        fprintf(stderr, "dropin_spy: cannot open %s\n", path);
        abort();
    }
    return f;
}

static void spy_put(FILE *f, const void *p, size_t size, size_t n)
{
    if (n && fwrite(p, size, n, f) != n) {
        fprintf(stderr, "dropin_spy: short write\n");
        abort();
    }
}

static long rk_calls;
void __wrap_runge_kutta(double *spin, double *cur, double *k1, double *k2, double *k3, double *k4)
{
    const size_t n = (size_t)Nfile * N * N * 3;
    double *in = malloc(n * sizeof(double));
    if (!in) abort();
    for (size_t i = 0; i < n; i++) in[i] = spin[i];
    __real_runge_kutta(spin, cur, k1, k2, k3, k4);
    FILE *f = spy_open(rk_calls == 0 ? "rk_first.bin" : "rk_last.bin", "wb");
    spy_put(f, in, sizeof(double), n);
    spy_put(f, cur, sizeof(double), n);
    spy_put(f, k1, sizeof(double), n);
    spy_put(f, k2, sizeof(double), n);
    spy_put(f, k3, sizeof(double), n);
    spy_put(f, k4, sizeof(double), n);
    fclose(f);
    free(in);
    rk_calls++;
    f = spy_open("rk_calls.txt", "w");
    fprintf(f, "%ld\n", rk_calls);
    fclose(f);
}

static long fd_calls;
void __wrap_fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda, double *b, int *ldb,
                        double *c, int *ldc, double *fuse_C, double *sum)
{
    const double s0 = sum[0], s1 = sum[1];
    FILE *f = NULL;
    const size_t nc = (size_t)*ldc * *n;
    if (fd_calls == 0) {
        f = spy_open("fd_first.bin", "wb");
        const int dims[6] = {*m, *n, *k, *lda, *ldb, *ldc};
        spy_put(f, dims, sizeof(int), 6);
        spy_put(f, alpha, sizeof(double), 1);
        spy_put(f, a, sizeof(double), (size_t)*lda * *k);
        spy_put(f, b, sizeof(double), (size_t)*ldb * *n);
        spy_put(f, c, sizeof(double), nc);
        spy_put(f, fuse_C, sizeof(double), nc);
    }
    __real_fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_C, sum);
    const double add[2] = {sum[0] - s0, sum[1] - s1};
    if (f) {
        spy_put(f, c, sizeof(double), nc);
        spy_put(f, add, sizeof(double), 2);
        fclose(f);
    }
    f = spy_open("fd_sums.bin", fd_calls == 0 ? "wb" : "ab");
    spy_put(f, add, sizeof(double), 2);
    fclose(f);
    fd_calls++;
}
