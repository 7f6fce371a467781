"""Reference-compatible entry points on the GPU against the oracle and the
reference's own functions (oracle/_ref): euler (llg_sqt_fusion.c:30),
runge_kutta (:98), fused_dgemm (gemm_kernel.h:355)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import need_gpu, say


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep", [(6, 1), (8, 3), (17, 2), (64, 1)])
def test_euler_and_runge_kutta(orc, L, nrep):
    need_gpu()
    import asd_amd
    say(f"compat euler/runge_kutta L={L} nrep={nrep}")
    S = np.stack([orc.init_spins(3 + r, L * L) for r in range(nrep)])
    for kw in (dict(), dict(J=0.8, D=0.1, Bz=0.3, alpha=0.05)):
        e_g = asd_amd.euler(asd_amd.params(dt=0.02, **kw), L, S)
        e_o = orc.llg_rhs(orc.params(dt=0.02, **kw), L, S)
        assert np.abs(e_g - e_o).max() <= 1e-14
        r_g = asd_amd.runge_kutta(asd_amd.params(dt=0.02, **kw), L, S)
        r_o = orc.runge_kutta(orc.params(dt=0.02, **kw), L, S)
        for a, b in zip(r_g, r_o):
            assert np.abs(a - b).max() <= 1e-14
        say(f"  bit-exact euler={np.array_equal(e_g, e_o)} rk4={all(np.array_equal(a, b) for a, b in zip(r_g, r_o))}")


def _blocked_inputs(orc, L, K, seed):
    ns = L * L
    rng = np.random.default_rng(seed)
    A = rng.standard_normal(ns * K)
    B = rng.standard_normal(K * ns)
    Q = orc.build_qmat(L, max(L // 10, 1), max(L // 10, 1))
    Cp = rng.standard_normal(Q.size)
    return A, B, Q, Cp


@pytest.mark.gpu
@pytest.mark.parametrize("L,K", [(4, 3), (8, 150), (12, 240), (10, 300), (16, 1200)])
def test_fused_dgemm_matches_oracle(orc, L, K):
    need_gpu()
    import asd_amd
    say(f"compat fused_dgemm L={L} K={K}")
    ns = L * L
    A, B, Q, Cp = _blocked_inputs(orc, L, K, L * K)
    ldc = Q.size // ns
    c_g, s_g = Cp.copy(), np.zeros(256)
    asd_amd.fused_dgemm(ns, ns, K, 1.0 / 50, A, ns, B, K, c_g, ldc, Q, s_g)
    c_o, s_o = Cp.copy(), np.zeros(2)
    orc.fused_dgemm(ns, ns, K, 1.0 / 50, A, ns, B, K, c_o, ldc, Q, s_o)
    assert np.abs(c_g - c_o).max() <= 1e-13 * np.abs(c_o).max()
    assert abs(complex(*s_g[:2]) - complex(*s_o)) <= 1e-11 * np.abs(c_o).sum()
    assert not s_g[2:].any()
    if K <= 240 and orc.RefFusion.available():
        ref = orc.RefFusion("parity")
        c_r, s_r = Cp.copy(), np.zeros(4096)
        ref.fused_dgemm(ns, ns, K, 1.0 / 50, A, ns, B, K, c_r, ldc, Q.copy(), s_r)
        assert np.abs(c_g - c_r).max() <= 1e-13 * np.abs(c_r).max()
        assert abs(complex(s_r[0::2].sum(), s_r[1::2].sum()) - complex(*s_g[:2])) <= 1e-11 * np.abs(c_r).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,k", [(13, 7, 5), (130, 66, 33), (1, 1, 1), (5, 3, 0)])
def test_fused_dgemm_ragged(orc, m, n, k):
    """m % 4 != 0, odd n, k = 0: shapes the reference drops or mishandles."""
    need_gpu()
    import asd_amd
    rng = np.random.default_rng(m + n + k)
    lda, ldb, ldc = m + 1, k + 2, 2 * ((m + 3) // 4 * 4) + 8
    a = rng.standard_normal(lda * max(k, 1))
    b = rng.standard_normal(ldb * n)
    c0 = rng.standard_normal(ldc * n)
    q = rng.standard_normal(ldc * n)
    c_g, s_g = c0.copy(), np.array([0.5, 0.25])
    asd_amd.fused_dgemm(m, n, k, 0.3, a, lda, b, ldb, c_g, ldc, q, s_g)
    c_o, s_o = c0.copy(), np.array([0.5, 0.25])
    orc.fused_dgemm(m, n, k, 0.3, a, lda, b, ldb, c_o, ldc, q, s_o)
    assert np.abs(c_g - c_o).max() <= 1e-13 * np.abs(c_o).max()
    assert np.abs(s_g - s_o).max() <= 1e-11 * np.abs(c_o).sum()


@pytest.mark.gpu
def test_drop_in_symbols_read_bound_globals(orc):
    """The exact-name entries (asd_compat.h) read the caller's globals at each call."""
    need_gpu()
    import asd_amd
    lib = asd_amd.lib()
    N, Nfile, num = C.c_int(8), C.c_int(2), C.c_int(4)
    dt, J1 = C.c_double(0.02), C.c_double(1.0)
    nn = (C.c_int * 8)(1, 0, -1, 0, 0, -1, 0, 1)
    assert lib.asd_compat_bind(C.byref(N), C.byref(Nfile), C.byref(dt), C.byref(J1), C.cast(nn, C.c_void_p),
                               C.byref(num)) == 0
    S = np.stack([orc.init_spins(9 + r, 64) for r in range(2)])
    out = np.empty_like(S)
    lib.euler(S.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    assert np.abs(out - orc.llg_rhs(orc.params(dt=0.02), 8, S)).max() <= 1e-14
    J1.value = 1.7                                   # changed global is seen by the next call
    bufs = [np.empty_like(S) for _ in range(5)]
    lib.runge_kutta(S.ctypes.data_as(C.c_void_p), *(b.ctypes.data_as(C.c_void_p) for b in bufs))
    ref = orc.runge_kutta(orc.params(J=1.7, dt=0.02), 8, S)
    for a, b in zip(bufs, ref):
        assert np.abs(a - b).max() <= 1e-14
