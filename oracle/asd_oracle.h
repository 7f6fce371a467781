/*
 * asd_oracle.h -- CPU FP64 oracle for the RK4-LLG + S(q,t) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The HIP library (libasd) never links, loads or calls anything here.
 *
 * It restates, in plain C, the algorithm of hwchen2017/ASD-kernel-fusion:
 *   kernel_fusion/llg_sqt_fusion.c   (euler, runge_kutta, renormalisation,
 *                                     operand layouts, phase matrix, S(q,t))
 *   kernel_fusion/gemm_kernel.h      (fused_dgemm contract)
 *   kernel_fusion/v3d.h              (cross-product operand order)
 *   gemm/llg_sqt_mkl.c:369-423       (the unfused form of the same S(q,t))
 * extended with the D / B_z / alpha terms, time-origin averaging and
 * multi-q output that the GPU path supports.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_llg_rhs, orc_runge_kutta  -- bit-level against the reference's own
 *       euler/runge_kutta built from source into oracle/_ref (D=B=alpha=0),
 *       exact cone spin-wave solution incl. D and B_z (4th-order convergence),
 *       ferromagnet fixed point, M / M_z / energy conservation, alpha>0 decay.
 *   orc_fused_dgemm               -- against the reference fused_dgemm built
 *       into oracle/_ref (K<=240, the range where the reference is correct)
 *       and a numpy DGEMM + epilogue for any K.
 *   orc_sqt_gemm                  -- against orc_sqt_separable (exact identity)
 *       and numpy FFT; sum rule over the full q grid.
 *   orc_init_spins                -- |S|=1, first/second moments of the
 *       uniform sphere distribution, determinism.
 *   orc_tlm_*                     -- against central finite differences of
 *       orc_run / orc_sqt_gemm.
 */
#ifndef ASD_ORACLE_H
#define ASD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    double J;           /* NN exchange (reference J1, llg_sqt_fusion.c:15)        */
    double D;           /* easy-axis anisotropy: H_z += 2 D S_z                    */
    double Bz;          /* field along z:         H_z += B_z                       */
    double alpha;       /* Gilbert damping                                          */
    double dt;          /* RK4 step (reference dt, llg_sqt_fusion.c:17)            */
    double spin_length; /* renormalisation length (kernel_fusion 1.5, others 1.0)  */
} orc_params;

/* counter-based 64-bit hash: splitmix64 finaliser of a linear key */
uint64_t orc_hash64(uint64_t seed, uint64_t a, uint64_t b);
/* Marsaglia (1972) uniform points on S^2, one independent counter stream per site */
void orc_init_spins(uint64_t seed, int nsites, double *S);

/* LLG right-hand side for nrep lattices, layout [rep][site][xyz] */
void orc_llg_rhs(const orc_params *p, int L, int nrep, const double *S, double *dS);
/* reference runge_kutta contract: cur = unnormalised S^{n+1}, k_s = dt * f(stage_s) */
void orc_runge_kutta(const orc_params *p, int L, int nrep, const double *spin,
                     double *cur, double *k1, double *k2, double *k3, double *k4);
/* S <- len * cur / |cur| per site */
void orc_renormalise(long nspins, double len, const double *cur, double *S);
/* nsteps RK4+renorm steps; snapshot k (< nsnap) = state after k*snap_every steps */
void orc_run(const orc_params *p, int L, int nrep, double *S, long nsteps,
             int snap_every, int nsnap, double *snaps);
double orc_energy(const orc_params *p, int L, const double *S);

/* reference fused_dgemm contract (gemm_kernel.h:355): C += (alpha A B) o Q in the
 * blocked 4-re/4-im layout, sum[0..1] += sum of the updated C.  Correct for all K. */
void orc_fused_dgemm(int m, int n, int k, double alpha, const double *a, int lda,
                     const double *b, int ldb, double *c, int ldc, const double *fuse_c,
                     double *sum);
/* phase matrix Q in the reference layout (llg_sqt_fusion.c:308-333); column stride
 * 2*ceil4(Ns) doubles (= 2 Ns, the reference's ldc, when Ns % 4 == 0) */
void orc_build_qmat(int L, int mx, int my, double *q);

/* S(q,tau) from snapshots [nsnap][Ns][3]; origins t0 = 0..T0-1, lags 0..nlags-1,
 * q list as (mx,my) pairs (q = 2 pi (mx,my)/L, mx pairs with row);
 * out[(lag*nq + iq)*2 + {0,1}] = (re, im).  GEMM form, as the reference. */
void orc_sqt_gemm(int L, const double *snaps, int T0, int nlags, int nq, const int *qlist,
                  double *out);
/* separable Fourier form of the same quantity */
void orc_sqt_separable(int L, const double *snaps, int T0, int nlags, int nq,
                       const int *qlist, double *out);

/* tangent-linear model of orc_run w.r.t. theta = (J, D, Bz, alpha):
 * dsnaps[ip][nsnap][Ns][3] = d snaps / d theta_ip (S0 fixed). */
void orc_tlm_run(const orc_params *p, int L, const double *S0, long nsteps, int snap_every,
                 int nsnap, double *snaps, double *dsnaps);
/* directional derivative of orc_sqt_gemm along dsnaps */
void orc_tlm_sqt(int L, const double *snaps, const double *dsnaps, int T0, int nlags, int nq,
                 const int *qlist, double *dout);

#ifdef __cplusplus
}
#endif
#endif
