"""Pin the oracle's S(q,tau) and fused-GEMM contract by identities.

* GEMM form (restating llg_sqt_fusion.c:235-442 with time origins as replicas)
  == separable Fourier form (exact factorisation of e^{-iq.(r_i - r_j)}).
* separable form == an independent numpy FFT computation.
* sum rule over the full q grid at lag 0.
* orc_fused_dgemm == numpy DGEMM + epilogue for every K, M % 4 != 0, odd N
  (the reference is only right for K <= 240, M % 4 == 0, even N).
* tangent-linear model == central finite differences.
"""
import numpy as np
import pytest


def full_grid(L):
    return [(mx, my) for mx in range(L) for my in range(L)]


def gxmg(L):
    """Gamma-X-M-Gamma, 3L/2 lattice momenta (mx pairs with row):
    (k,0) k<L/2; (L/2,k) k<L/2; (L/2-k, L/2-k) k<L/2 -- each segment starts at
    its first corner and stops before the next."""
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


def trajectory(orc, L, nsnap, every=5, seed=1, **kw):
    p = orc.params(dt=kw.pop("dt", 0.01), **kw)
    _, snaps = orc.run(p, L, orc.init_spins(seed, L * L), (nsnap - 1) * every, every, nsnap)
    return snaps


def sqt_numpy_fft(L, snaps, T0, nlags):
    """Independent restatement: F = fft2 over (row, col); S on the full grid."""
    F = np.fft.fft2(snaps.reshape(len(snaps), L, L, 3), axes=(1, 2))   # e^{-2pi i (mx r + my c)/L}
    out = np.empty((nlags, L, L), complex)
    for lag in range(nlags):
        a, b = F[lag:lag + T0], F[:T0]
        corr = (a * b.conj()).sum(axis=-1).mean(axis=0)
        disc = (a.mean(axis=0) * b.mean(axis=0).conj()).sum(axis=-1)
        out[lag] = (corr - disc) / (L * L)
    return out.reshape(nlags, L * L)


@pytest.mark.parametrize("L,T0,nlags", [(4, 6, 3), (6, 8, 4), (5, 4, 5)])
def test_gemm_form_equals_separable(orc, L, T0, nlags):
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
    q = full_grid(L)
    g = orc.sqt_gemm(L, snaps, T0, nlags, q)
    s = orc.sqt_separable(L, snaps, T0, nlags, q)
    assert np.abs(g - s).max() < 1e-12 * np.abs(s).max()


def test_separable_equals_numpy_fft(orc):
    L, T0, nlags = 8, 10, 6
    snaps = trajectory(orc, L, T0 + nlags - 1)
    s = orc.sqt_separable(L, snaps, T0, nlags, full_grid(L))
    f = sqt_numpy_fft(L, snaps, T0, nlags)
    assert np.abs(s - f).max() < 1e-12 * np.abs(f).max()


def test_sum_rule_lag0(orc):
    """sum over all q of S(q,0) = sum_i [ <|S_i|^2> - |zbar_i|^2 ] = Ns - sum |zbar|^2."""
    L, T0 = 6, 12
    snaps = trajectory(orc, L, T0)
    s0 = orc.sqt_gemm(L, snaps, T0, 1, full_grid(L))[0]
    zbar = snaps[:T0].mean(axis=0)
    assert abs(s0.sum() - (L * L - (zbar ** 2).sum())) < 1e-11 * L * L
    assert np.abs(s0.imag).max() < 1e-13        # lag 0: G symmetric -> real


def test_gxmg_path_layout(orc):
    q = gxmg(8)
    assert len(q) == 12 and q[0] == (0, 0) and q[4] == (4, 0) and q[8] == (4, 4) and q[-1] == (1, 1)


def _blk(i, j, ldc):
    return (i // 4) * 8 + i % 4 + j * ldc


@pytest.mark.parametrize("m,n,k", [(16, 16, 300), (10, 7, 5), (13, 12, 481), (8, 9, 1)])
def test_fused_dgemm_contract_numpy(orc, m, n, k):
    rng = np.random.default_rng(m * n + k)
    lda, ldb = m + 3, k + 2
    ldc = 2 * ((m + 3) // 4) * 4
    a = rng.standard_normal(lda * k)
    b = rng.standard_normal(ldb * n)
    c0 = rng.standard_normal(ldc * n)
    q = rng.standard_normal(ldc * n)
    alpha = 0.37
    A = a.reshape(k, lda).T[:m]           # column-major m x k
    Bm = b.reshape(n, ldb).T[:k]          # column-major k x n
    G = alpha * A @ Bm
    c_exp = c0.copy()
    tot = 0j
    for j in range(n):
        for i in range(m):
            pr = _blk(i, j, ldc)
            c_exp[pr] += G[i, j] * q[pr]
            c_exp[pr + 4] += G[i, j] * q[pr + 4]
            tot += complex(c_exp[pr], c_exp[pr + 4])
    c = c0.copy()
    s = np.array([1.0, -2.0])
    orc.fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, q, s)
    assert np.abs(c - c_exp).max() < 1e-12 * np.abs(c_exp).max()
    assert abs(complex(s[0] - 1.0, s[1] + 2.0) - tot) < 1e-11 * max(1, abs(tot))


def test_qmat_phase_convention(orc):
    """Q_ij = exp(-i(qx (r_i - r_j) + qy (c_i - c_j))), qx with the row offset."""
    L, mx, my = 6, 1, 2
    Q = orc.build_qmat(L, mx, my)
    ns = L * L
    for (i, j) in [(0, 0), (7, 3), (35, 1), (3, 20)]:
        ri, ci, rj, cj = i // L, i % L, j // L, j % L
        e = np.exp(-1j * 2 * np.pi * (mx * (ri - rj) + my * (ci - cj)) / L)
        pr = _blk(i, j, 2 * ns)
        assert abs(complex(Q[pr], Q[pr + 4]) - e) < 1e-14


@pytest.mark.parametrize("th", [(1.0, 0.1, 0.1, 0.01), (1.1, 0.12, 0.08, 0.0), (-0.7, 0.05, 0.2, 0.2)])
def test_tlm_matches_finite_differences(orc, th):
    L, every, nsnap, T0, nlags = 4, 5, 9, 5, 5
    S0 = orc.init_spins(3, L * L)
    J, D, Bz, al = th
    p = orc.params(J=J, D=D, Bz=Bz, alpha=al, dt=0.01)
    snaps, dsnaps = orc.tlm_run(p, L, S0, (nsnap - 1) * every, every, nsnap)
    _, s_chk = orc.run(p, L, S0, (nsnap - 1) * every, every, nsnap)
    # the tangent model linearises along orc_run's own trajectory, bit for bit: over
    # thousands of chaotic steps a 1-ulp offset in the base trajectory grows to ~1e-4
    # in the derivative
    assert np.array_equal(snaps, s_chk)
    q = gxmg(L)
    h = 1e-6
    for ip in range(4):
        tp, tm = list(th), list(th)
        tp[ip] += h
        tm[ip] -= h
        _, sp = orc.run(orc.params(*tp, dt=0.01), L, S0, (nsnap - 1) * every, every, nsnap)
        _, sm = orc.run(orc.params(*tm, dt=0.01), L, S0, (nsnap - 1) * every, every, nsnap)
        fd = (sp - sm) / (2 * h)
        assert np.abs(dsnaps[ip] - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
        dsq = orc.tlm_sqt(L, snaps, dsnaps[ip], T0, nlags, q)
        fdq = (orc.sqt_gemm(L, sp, T0, nlags, q) - orc.sqt_gemm(L, sm, T0, nlags, q)) / (2 * h)
        assert np.abs(dsq - fdq).max() < 1e-6 * max(1e-3, np.abs(fdq).max())
