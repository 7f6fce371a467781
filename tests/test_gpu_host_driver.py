"""The C host driver (asd-kernel-fusion_amd/host/llg_sqt.c): reference mode appends the
"L tlle tsqt" line of llg_sqt_fusion.c:455-456 and writes S(q,t); pipeline mode writes
S(q,tau) as raw doubles plus a JSON header, equal to asd_sqt on the same trajectory."""
import json
import os
import subprocess

import numpy as np
import pytest

from gpu_util import need_gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "asd-kernel-fusion_amd",
                   "asd_llg_sqt")


@pytest.mark.gpu
def test_reference_mode(tmp_path):
    need_gpu()
    r = subprocess.run([BIN, "-n", "16", "-f", "6", "-o", "t.txt", "-O", "s.txt"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    L, tlle, tsqt = open(tmp_path / "t.txt").read().split()
    assert L == "16" and float(tlle) > 0 and float(tsqt) > 0
    s = np.loadtxt(tmp_path / "s.txt")
    assert s.shape == (60, 3) and np.isfinite(s).all()


@pytest.mark.gpu
def test_pipeline_mode_matches_api(gctx, tmp_path):
    torch = need_gpu()
    import asd_amd
    L, steps, every, T0, G = 16, 300, 10, 20, 8
    r = subprocess.run([BIN, "-P", "-n", str(L), "-S", str(steps), "-E", str(every), "-T", str(T0), "-G", str(G),
                        "-e", "0.01", "-D", "0.1", "-s", "4", "-p", "fp64", "-O", "out.bin"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    hdr = json.load(open(tmp_path / "out.bin.json"))
    got = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(hdr["lags"], hdr["nq"], 2)
    S = torch.tensor(asd_amd.init_spins(4, L * L), device="cuda")
    nsnap = steps // every
    snaps = torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda")
    gctx.rk4_run(asd_amd.params(D=0.1), L, S, steps, every, nsnap, snaps)
    ref = gctx.sqt(L, snaps, T0, G, np.array(hdr["q"], dtype=np.int32), "fp64")
    assert np.abs(got[..., 0] + 1j * got[..., 1] - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.gpu
def test_pipeline_snapshot_file_round_trip(orc, tmp_path):
    """-W writes the trajectory's snapshots (equal to the oracle's, bit for bit); -R reads
    them back instead of integrating and gives the same S(q,tau); a file of the wrong size
    is rejected."""
    need_gpu()
    L, steps, every, T0, G = 16, 200, 10, 12, 6
    common = ["-P", "-n", str(L), "-S", str(steps), "-E", str(every), "-T", str(T0), "-G", str(G), "-e", "0.01",
              "-D", "0.1", "-B", "0.05", "-s", "7", "-p", "fp64"]
    r = subprocess.run([BIN, *common, "-W", "snaps.bin", "-O", "a.bin"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    hdr = json.load(open(tmp_path / "snaps.bin.json"))
    nsnap = steps // every
    assert (hdr["L"], hdr["snapshots"], hdr["snap_every"]) == (L, nsnap, every)
    snaps = np.fromfile(tmp_path / "snaps.bin", dtype="<f8").reshape(nsnap, L * L, 3)
    S0 = orc.init_spins(7, L * L)
    _, snaps_ref = orc.run(orc.params(dt=0.01, D=0.1, Bz=0.05), L, S0[None], steps, every, nsnap)
    assert np.array_equal(snaps, snaps_ref.reshape(nsnap, L * L, 3))

    r = subprocess.run([BIN, *common, "-R", "snaps.bin", "-O", "b.bin"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a = np.fromfile(tmp_path / "a.bin", dtype="<f8")
    b = np.fromfile(tmp_path / "b.bin", dtype="<f8")
    assert a.shape == b.shape and np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    # the split precisions through the driver's -p flag, on the same snapshots (K = 36: a small
    # sample, bound 2e-4 as for the library's small-sample checks)
    for prec in ("fp16x2", "fp16f8"):
        r = subprocess.run([BIN, *common[:-1], prec, "-R", "snaps.bin", "-O", f"{prec}.bin"], cwd=tmp_path,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        c = np.fromfile(tmp_path / f"{prec}.bin", dtype="<f8")
        assert c.shape == a.shape and np.abs(c - a).max() <= 2e-4 * np.abs(a).max(), prec

    (tmp_path / "short.bin").write_bytes((tmp_path / "snaps.bin").read_bytes()[:-8])
    r = subprocess.run([BIN, *common, "-R", "short.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "expected" in r.stderr
