"""The reference's own driver (kernel_fusion/llg_sqt_fusion.c main) with the drop-in edit
of INTEGRATION.md section 1 -- its euler/runge_kutta definitions and gemm_kernel.h
removed, asd_compat.h included, asd_compat_bind() called -- linked to libasd by
oracle/Makefile (oracle/_ref/llg_sqt_fusion_dropin).  Its hot-path calls
runge_kutta (:348) and fused_dgemm (:429) then run on the GPU; the run must
finish and append its "L tlle tsqt" line (:455-456).  (The reference's own
init makes the spins NaN, :204-206, so only the integration is checked here.)"""
import os
import subprocess

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
