/*
 * asd_compat.h -- literal drop-in for the three functions the reference driver
 * kernel_fusion/llg_sqt_fusion.c calls on its hot path, with the reference's
 * exact names and signatures.  A reference-style driver keeps its own globals
 * (N, Nfile, dt, J1, nn, num_neighbor; llg_sqt_fusion.c:13-21), deletes its
 * definitions of euler/runge_kutta and `#include "gemm_kernel.h"`, includes
 * this header, calls asd_compat_bind() once after setting nn[][], and links
 * libasd.  The functions then read the bound globals at every call, as the
 * reference's do.  They run on the GPU, copying host arrays in and out.
 *
 * Errors: the reference has no error channel (void functions); these print
 * asd_last_error() to stderr and abort() on invalid input or a HIP failure.
 */
#ifndef ASD_COMPAT_H
#define ASD_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Binds the caller's globals.  nn must hold the reference's neighbour table
 * {(1,0),(-1,0),(0,-1),(0,1)} (llg_sqt_fusion.c:181-184) and *num_neighbor == 4.
 * spin_length applies to nothing here (runge_kutta does not renormalise). */
int asd_compat_bind(int *N, int *Nfile, double *dt, double *J1, int (*nn)[2], int *num_neighbor);

/* llg_sqt_fusion.c:30 */
void euler(double *spin0, double *derivative);
/* llg_sqt_fusion.c:98 */
void runge_kutta(double *spin, double *cur_input, double *k1, double *k2, double *k3, double *k4);
/* gemm_kernel.h:355 (see asd_fused_dgemm in asd.h for the exact contract) */
void fused_dgemm(int *m, int *n, int *k, double *alpha, double *a, int *lda, double *b, int *ldb,
                 double *c, int *ldc, double *fuse_C, double *sum);

#ifdef __cplusplus
}
#endif
#endif /* ASD_COMPAT_H */
