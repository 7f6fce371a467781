"""Writes tests/golden/reference_vectors.npz: input/output vectors of the
reference's own kernel_fusion functions (euler, runge_kutta, fused_dgemm),
compiled from the reference sources into oracle/_ref by oracle/Makefile.

Run in a checkout where oracle/_ref was built (needs the reference sources at
build time):  python tests/golden/gen_reference_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def main():
    ref = oracle.RefFusion("parity")
    out = {}
    # euler / runge_kutta: L = 6, Nfile = 2, J1 = 1.0, dt = 0.02 (reference globals)
    L, nf, dt = 6, 2, 0.02
    ref.set_globals(L, nf, dt, 1.0)
    S = np.stack([oracle.init_spins(50 + r, L * L) for r in range(nf)])
    out["rk_L"], out["rk_nfile"], out["rk_dt"] = L, nf, dt
    out["rk_spin"] = S
    out["rk_euler"] = ref.euler(S)
    cur, k1, k2, k3, k4 = ref.runge_kutta(S)
    out["rk_cur"], out["rk_k1"], out["rk_k4"] = cur, k1, k4
    # fused_dgemm: m = n = 36 (L = 6), K = 30 (Nfile = 10), alpha = 1/10, q = (1, 1)
    Lq, K = 6, 30
    ns = Lq * Lq
    rng = np.random.default_rng(123)
    A = rng.standard_normal(ns * K)
    B = rng.standard_normal(K * ns)
    Q = oracle.build_qmat(Lq, 1, 1)
    C0 = rng.standard_normal(Q.size)
    c = C0.copy()
    sums = np.zeros(4096)
    ref.fused_dgemm(ns, ns, K, 0.1, A, ns, B, K, c, 2 * ns, Q.copy(), sums)
    out.update(fd_L=Lq, fd_K=K, fd_alpha=0.1, fd_A=A, fd_B=B, fd_Q=Q, fd_C0=C0, fd_C=c,
               fd_sum=np.array([sums[0::2].sum(), sums[1::2].sum()]))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
