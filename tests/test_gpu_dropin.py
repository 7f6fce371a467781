"""The reference's own driver (kernel_fusion/llg_sqt_fusion.c main) with the drop-in edit
of INTEGRATION.md section 1 -- its euler/runge_kutta definitions and gemm_kernel.h
removed, asd_compat.h included, asd_compat_bind() called -- linked to libasd by
oracle/Makefile (oracle/_ref/llg_sqt_fusion_dropin).  Its hot-path calls
runge_kutta (:348) and fused_dgemm (:429) then run on the GPU; the run must
finish and append its "L tlle tsqt" line (:455-456).  The reference's own init
makes the spins NaN (:204-206), so the values are checked on a second build of the
same edit with oracle/dropin_spy.c linked in (recorded calls, a fixed-seed random
source): test_reference_driver_values_through_libasd."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import need_gpu, say

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "llg_sqt_fusion_dropin")


@pytest.mark.gpu
def test_reference_driver_runs_on_libasd(tmp_path):
    need_gpu()
    if not os.path.exists(BIN):
        pytest.skip("drop-in build of the reference driver not present (needs the reference sources)")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([BIN, "-n", "20", "-f", "10"], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                       env=env)
    say(r.stdout[-300:])
    assert r.returncode == 0, r.stderr
    line = open(tmp_path / "amd_gcc_11.2_fusion_sqt.txt").read().split()
    assert line[0] == "20" and float(line[1]) > 0 and float(line[2]) > 0


SPY = os.path.join(os.path.dirname(BIN), "llg_sqt_fusion_dropin_spy")


def _read_rk(path, n):
    x = np.fromfile(path, dtype=np.float64)
    assert x.size == 6 * n
    return x.reshape(6, n)    # spin (in), cur, k1, k2, k3, k4


@pytest.mark.gpu
def test_reference_driver_values_through_libasd(tmp_path, orc):
    """The same drop-in build with oracle/dropin_spy.c linked in: the driver's own
    runge_kutta (:348) and fused_dgemm (:429) calls are recorded as they go to libasd, and
    its random source is a fixed-seed one (the reference's integer division makes every
    spin NaN, :204-206).  Checked against the oracle (itself pinned bit for bit to the
    reference's functions): the first and the 60th RK4 call bit for bit, along the
    driver's whole 60-step trajectory with its 1.5 renormalisation (:350-361), and the
    S(q,t) sum of every one of its 60 fused_dgemm calls, recomputed from the oracle's
    trajectory with the driver's operand layouts, averages and disconnected prefill
    (:235-268, :370-426)."""
    need_gpu()
    if not os.path.exists(SPY):
        pytest.skip("instrumented drop-in build not present (needs the reference sources)")
    N, nfile, dt, steps = 20, 4, 0.02, 60
    ns, K = N * N, 3 * nfile
    env = dict(os.environ, OMP_NUM_THREADS="4", ASD_SPY_DIR=str(tmp_path))
    r = subprocess.run([SPY, "-n", str(N), "-f", str(nfile)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120, env=env)
    say(r.stdout[-200:])
    assert r.returncode == 0, r.stderr
    assert int(open(tmp_path / "rk_calls.txt").read()) == steps
    n = nfile * ns * 3
    first = _read_rk(tmp_path / "rk_first.bin", n)
    last = _read_rk(tmp_path / "rk_last.bin", n)
    S0 = first[0].reshape(nfile, ns, 3)
    assert np.isfinite(S0).all()
    assert np.abs(np.linalg.norm(S0, axis=-1) - 1.0).max() <= 1e-15

    # RK4 along the driver's trajectory, bit for bit
    p = orc.params(J=1.0, dt=dt)
    S = S0.copy()
    traj = [S0]
    for step in range(steps):
        out = orc.runge_kutta(p, N, S)
        if step == 0:
            for a, b in zip(first[1:], out):
                assert np.array_equal(a, b.ravel())
        if step == steps - 1:
            assert np.array_equal(last[0], S.ravel())
            for a, b in zip(last[1:], out):
                assert np.array_equal(a, b.ravel())
        S = orc.renormalise(out[0], 1.5)
        traj.append(S)

    # first fused_dgemm call: inputs as the driver built them, output against the oracle
    raw = open(tmp_path / "fd_first.bin", "rb").read()
    dims = np.frombuffer(raw[:24], dtype=np.int32)
    m, nn, k, lda, ldb, ldc = (int(v) for v in dims)
    assert (m, nn, k, lda, ldb, ldc) == (ns, ns, K, ns, K, 2 * ns)
    x = np.frombuffer(raw[24:], dtype=np.float64)
    alpha = x[0]
    o = 1
    a = x[o:o + lda * k]; o += lda * k
    b = x[o:o + ldb * nn]; o += ldb * nn
    c_in = x[o:o + ldc * nn]; o += ldc * nn
    q = x[o:o + ldc * nn]; o += ldc * nn
    c_out = x[o:o + ldc * nn]; o += ldc * nn
    add = x[o:o + 2]
    assert o + 2 == x.size and alpha == 1.0 / nfile
    c_orc, s_orc = c_in.copy(), np.zeros(2)
    orc.fused_dgemm(m, nn, k, alpha, a.copy(), lda, b.copy(), ldb, c_orc, ldc, q.copy(), s_orc)
    assert np.abs(c_out - c_orc).max() <= 1e-13 * np.abs(c_orc).max()
    assert np.allclose(add, s_orc, rtol=1e-11, atol=1e-11 * np.abs(c_orc).sum())

    # every call's S(q,t) sum, from the oracle's trajectory and the driver's data flow
    vx = vy = N // 10
    Q = orc.build_qmat(N, vx, vy)
    assert np.array_equal(Q, q)
    sums = np.fromfile(tmp_path / "fd_sums.bin", dtype=np.float64).reshape(-1, 2)
    assert sums.shape == (steps, 2) and np.isfinite(sums).all()
    B = np.ascontiguousarray(S0.transpose(1, 0, 2).reshape(ns, K))           # spin_zero, ldb = K
    zbar = S0.sum(axis=0) / nfile                                            # zerot_avg
    rows = (np.arange(ns) // 4) * 8 + np.arange(ns) % 4
    for step in range(steps):
        St = traj[step + 1]
        A = np.ascontiguousarray(St.transpose(0, 2, 1).reshape(K, ns))      # spint_mkl, lda = Ns
        fbar = St.sum(axis=0) / nfile                                        # fint_avg
        tmp = -(np.outer(fbar[:, 0], zbar[:, 0]) + np.outer(fbar[:, 1], zbar[:, 1]) +
                np.outer(fbar[:, 2], zbar[:, 2]))                            # [i, j]
        c = np.zeros(ldc * nn)
        col = np.arange(nn) * ldc
        re = (rows[:, None] + col[None, :]).ravel()
        c[re] = tmp.ravel() * Q[re]
        c[re + 4] = tmp.ravel() * Q[re + 4]
        if step == 0:   # the restated data flow builds what the driver passed
            assert np.array_equal(A.ravel(), a) and np.array_equal(B.ravel(), b)
            assert np.abs(c - c_in).max() <= 1e-15 * np.abs(c_in).max()
        s = np.zeros(2)
        orc.fused_dgemm(ns, ns, K, 1.0 / nfile, A.ravel(), ns, B.ravel(), K, c, ldc, Q.copy(), s)
        assert np.allclose(sums[step], s, rtol=1e-11, atol=1e-11 * np.abs(c).sum()), step
