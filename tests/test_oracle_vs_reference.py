"""Pin the CPU oracle against the reference's own kernel_fusion functions.

The reference (kernel_fusion/llg_sqt_fusion.c + gemm_kernel.h) is compiled from
its sources into oracle/_ref by oracle/Makefile and driven through its globals
(llg_sqt_fusion.c:13-21) and functions: euler (:30), runge_kutta (:98) and
fused_dgemm (gemm_kernel.h:355).  The reference has no tests or golden vectors
of its own, so these comparisons are the parity anchor for the restatement.
"""
import numpy as np
import pytest


def _spins(orc, L, nrep, seed):
    return np.stack([orc.init_spins(seed + r, L * L) for r in range(nrep)])


@pytest.mark.parametrize("L,nrep", [(4, 1), (6, 3), (7, 2), (8, 5), (16, 1)])
def test_euler_bit_exact(orc, ref_fusion, L, nrep):
    ref_fusion.set_globals(L, nrep, 0.01, 1.0)
    S = _spins(orc, L, nrep, 11)
    assert np.array_equal(ref_fusion.euler(S), orc.llg_rhs(orc.params(dt=0.01), L, S))


@pytest.mark.parametrize("J1", [1.0, -0.7, 2.5])
def test_euler_exchange_constant(orc, ref_fusion, J1):
    L = 6
    ref_fusion.set_globals(L, 2, 0.01, J1)
    S = _spins(orc, L, 2, 3)
    assert np.array_equal(ref_fusion.euler(S), orc.llg_rhs(orc.params(J=J1, dt=0.01), L, S))


@pytest.mark.parametrize("L,nrep,dt", [(4, 1, 0.01), (8, 3, 0.02), (9, 2, 0.05)])
def test_runge_kutta_bit_exact(orc, ref_fusion, L, nrep, dt):
    ref_fusion.set_globals(L, nrep, dt, 1.0)
    S = _spins(orc, L, nrep, 5)
    for a, b in zip(ref_fusion.runge_kutta(S), orc.runge_kutta(orc.params(dt=dt), L, S)):
        assert np.array_equal(a, b)


def test_trajectory_with_reference_renormalisation(orc, ref_fusion):
    """100 steps of reference runge_kutta + the 1.5-length renormalisation of
    llg_sqt_fusion.c:350-361 against orc_run(spin_length=1.5)."""
    L, nrep, dt = 8, 2, 0.01
    ref_fusion.set_globals(L, nrep, dt, 1.0)
    S = _spins(orc, L, nrep, 21)
    S_ref = S.copy()
    for _ in range(100):
        cur = ref_fusion.runge_kutta(S_ref)[0]
        S_ref = orc.renormalise(cur, 1.5)
    S_orc, _ = orc.run(orc.params(dt=dt, spin_length=1.5), L, S, 100)
    assert np.array_equal(S_ref, S_orc)
    assert np.allclose(np.linalg.norm(S_orc, axis=-1), 1.5, atol=1e-14, rtol=0)


def _fused_inputs(orc, L, K, seed):
    ns = L * L
    rng = np.random.default_rng(seed)
    A = rng.standard_normal(ns * K)
    B = rng.standard_normal(K * ns)
    Q = orc.build_qmat(L, max(L // 10, 1), max(L // 10, 1))
    Cp = rng.standard_normal(2 * ns * ns)
    return A, B, Q, Cp


@pytest.mark.parametrize("L,K", [(4, 3), (8, 9), (8, 150), (12, 240), (10, 60)])
def test_fused_dgemm_matches_reference(orc, ref_fusion, L, K):
    """K <= 240 (the reference's correct range), M = N = L^2 with M % 4 == 0."""
    ns = L * L
    A, B, Q, Cp = _fused_inputs(orc, L, K, L + K)
    c_ref, s_ref = Cp.copy(), np.zeros(4096)
    ref_fusion.fused_dgemm(ns, ns, K, 1.0 / 7, A, ns, B, K, c_ref, 2 * ns, Q.copy(), s_ref)
    c_orc, s_orc = Cp.copy(), np.zeros(2)
    orc.fused_dgemm(ns, ns, K, 1.0 / 7, A, ns, B, K, c_orc, 2 * ns, Q, s_orc)
    scale = np.abs(c_orc).max()
    assert np.abs(c_ref - c_orc).max() <= 1e-13 * scale
    tot = np.array([s_ref[0::2].sum(), s_ref[1::2].sum()])
    assert np.allclose(tot, s_orc, rtol=1e-11, atol=1e-11 * np.abs(c_orc).sum())


def test_reference_double_counts_above_k240(orc, ref_fusion):
    """Documented deviation: for K > 240 the reference sums the partially updated
    C once per 240-wide K block (gemm_kernel.h:378, :406).  The oracle (and the
    GPU path) sum once.  Model the reference's sum from the oracle's pieces."""
    L, K = 4, 300
    ns = L * L
    A, B, Q, Cp = _fused_inputs(orc, L, K, 1)
    c_ref, s_ref = Cp.copy(), np.zeros(4096)
    ref_fusion.fused_dgemm(ns, ns, K, 1.0, A, ns, B, K, c_ref, 2 * ns, Q.copy(), s_ref)
    # block 1: k in [0,240), block 2: [240,300): C after each block, summed
    c1, s1 = Cp.copy(), np.zeros(2)
    orc.fused_dgemm(ns, ns, 240, 1.0, A, ns, B, K, c1, 2 * ns, Q, s1)
    c2, s2 = c1.copy(), np.zeros(2)
    orc.fused_dgemm(ns, ns, 60, 1.0, A[240 * ns:].copy(), ns, B.reshape(ns, K)[:, 240:].copy().ravel(), 60,
                    c2, 2 * ns, Q, s2)
    modelled = s1 + s2
    tot = np.array([s_ref[0::2].sum(), s_ref[1::2].sum()])
    assert np.allclose(tot, modelled, rtol=1e-10, atol=1e-9)
    # and the final C is still right (the store is cumulative)
    c_full, s_full = Cp.copy(), np.zeros(2)
    orc.fused_dgemm(ns, ns, K, 1.0, A, ns, B, K, c_full, 2 * ns, Q, s_full)
    assert np.abs(c_ref - c_full).max() <= 1e-12 * np.abs(c_full).max()
    assert not np.allclose(tot, s_full, rtol=1e-6)


def test_sqt_gemm_form_matches_reference_kernel(orc, ref_fusion):
    """orc_sqt_gemm's composition (operand layouts, averages, disconnected
    prefill, 1/N^2) against the reference's fused kernel driven with the same
    operands built the way llg_sqt_fusion.c:235-268, :370-426 builds them."""
    L, T0, lag = 6, 20, 3
    ns, K = L * L, 3 * T0
    p = orc.params(dt=0.02)
    _, snaps = orc.run(p, L, orc.init_spins(4, ns), (T0 + lag) * 2, 2, T0 + lag)
    mx, my = 1, 2
    B = np.empty(K * ns)
    A = np.empty(ns * K)
    for t0 in range(T0):
        for c in range(3):
            B[(t0 * 3 + c) + np.arange(ns) * K] = snaps[t0, :, c]
            A[np.arange(ns) + (t0 * 3 + c) * ns] = snaps[t0 + lag, :, c]
    zbar = snaps[:T0].mean(axis=0)
    fbar = snaps[lag:lag + T0].mean(axis=0)
    Q = orc.build_qmat(L, mx, my)
    Cp = np.empty_like(Q)
    idx = np.arange(ns)
    for j in range(ns):
        base = (idx // 4) * 8 + idx % 4 + j * 2 * ns
        tmp = -(fbar @ zbar[j])
        Cp[base] = tmp * Q[base]
        Cp[base + 4] = tmp * Q[base + 4]
    s_ref = np.zeros(4096)
    ref_fusion.fused_dgemm(ns, ns, K, 1.0 / T0, A, ns, B, K, Cp, 2 * ns, Q, s_ref)
    s_ref = complex(s_ref[0::2].sum(), s_ref[1::2].sum()) / ns
    s_orc = orc.sqt_gemm(L, snaps, T0, lag + 1, [(mx, my)])[lag, 0]
    assert abs(s_ref - s_orc) <= 1e-12 * max(1.0, abs(s_orc))
