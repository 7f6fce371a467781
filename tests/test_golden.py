"""Frozen vectors of the reference's own functions (tests/golden/reference_vectors.npz,
made by tests/golden/gen_reference_vectors.py from kernel_fusion/llg_sqt_fusion.c's
euler (:30), runge_kutta (:98) and gemm_kernel.h's fused_dgemm (:355) built from
source).  They pin the oracle without the reference build, and the GPU path
through the C ABI."""
import os

import numpy as np
import pytest

from gpu_util import need_gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def test_oracle_euler_runge_kutta_match_golden(orc):
    L, dt = int(GOLD["rk_L"]), float(GOLD["rk_dt"])
    S = GOLD["rk_spin"]
    p = orc.params(dt=dt)
    assert np.array_equal(orc.llg_rhs(p, L, S), GOLD["rk_euler"])
    cur, k1, _, _, k4 = orc.runge_kutta(p, L, S)
    assert np.array_equal(cur, GOLD["rk_cur"])
    assert np.array_equal(k1, GOLD["rk_k1"]) and np.array_equal(k4, GOLD["rk_k4"])


def test_oracle_fused_dgemm_matches_golden(orc):
    L, K, al = int(GOLD["fd_L"]), int(GOLD["fd_K"]), float(GOLD["fd_alpha"])
    ns = L * L
    c = GOLD["fd_C0"].copy()
    s = np.zeros(2)
    orc.fused_dgemm(ns, ns, K, al, GOLD["fd_A"], ns, GOLD["fd_B"], K, c, 2 * ns, GOLD["fd_Q"], s)
    assert np.abs(c - GOLD["fd_C"]).max() <= 1e-13 * np.abs(c).max()
    assert np.abs(s - GOLD["fd_sum"]).max() <= 1e-11 * np.abs(c).sum()


@pytest.mark.gpu
def test_gpu_matches_golden():
    need_gpu()
    import asd_amd
    L, dt = int(GOLD["rk_L"]), float(GOLD["rk_dt"])
    S = GOLD["rk_spin"]
    p = asd_amd.params(dt=dt)
    assert np.array_equal(asd_amd.euler(p, L, S), GOLD["rk_euler"])
    cur, k1, _, _, k4 = asd_amd.runge_kutta(p, L, S)
    assert np.array_equal(cur, GOLD["rk_cur"]) and np.array_equal(k4, GOLD["rk_k4"])
    Lq, K, al = int(GOLD["fd_L"]), int(GOLD["fd_K"]), float(GOLD["fd_alpha"])
    ns = Lq * Lq
    c = GOLD["fd_C0"].copy()
    s = np.zeros(2)
    asd_amd.fused_dgemm(ns, ns, K, al, GOLD["fd_A"], ns, GOLD["fd_B"], K, c, 2 * ns, GOLD["fd_Q"], s)
    assert np.abs(c - GOLD["fd_C"]).max() <= 1e-13 * np.abs(c).max()
    assert np.abs(s - GOLD["fd_sum"]).max() <= 1e-11 * np.abs(c).sum()
