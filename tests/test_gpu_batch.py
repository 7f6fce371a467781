"""Batched uneven-size trajectories (asd_batch_rk4_run): one persistent launch per
step over a (trajectory, tile) work list, graph-captured step loop.  Each
trajectory must equal its unbatched run and the oracle bit for bit (same exact
tile arithmetic)."""
import numpy as np
import pytest

from gpu_util import need_gpu, say


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps", [200, 205])
def test_batch_equals_single_runs(gctx, orc, nsteps):
    torch = need_gpu()
    import asd_amd
    Ls = [32, 48, 64, 96, 128, 32, 48, 20, 128, 64]
    rng = np.random.default_rng(7)
    Ds = rng.uniform(0, 0.2, len(Ls))
    every, nsnap = 10, nsteps // 10 + 1
    plist = [asd_amd.params(J=1.0, D=float(d), dt=0.01) for d in Ds]
    S0 = [asd_amd.init_spins(100 + t, L * L) for t, L in enumerate(Ls)]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    snaps = [torch.full((nsnap, L * L, 3), np.nan, dtype=torch.float64, device="cuda") for L in Ls]
    say(f"batch of {len(Ls)} trajectories, {nsteps} steps")
    gctx.batch_rk4_run(Ls, plist, spins, nsteps, every, nsnap, snaps)
    gctx.synchronize()
    for t, L in enumerate(Ls):
        s1 = torch.tensor(S0[t], device="cuda")
        sn1 = torch.full((nsnap, L * L, 3), np.nan, dtype=torch.float64, device="cuda")
        gctx.rk4_run(plist[t], L, s1, nsteps, every, nsnap, sn1)
        gctx.synchronize()
        a, b = spins[t].cpu().numpy(), s1.cpu().numpy()
        assert np.array_equal(a, b), (t, L, np.abs(a - b).max())
        sa, sb = snaps[t].cpu().numpy(), sn1.cpu().numpy()
        expect_written = min(nsnap, nsteps // every + 1) if nsteps % every == 0 else min(nsnap, (nsteps - 1) // every + 1)
        assert not np.isnan(sa[:expect_written]).any()
        assert np.array_equal(sa[:expect_written], sb[:expect_written])
    # two trajectories against the oracle
    for t in (2, 7):
        L = Ls[t]
        ref, _ = orc.run(orc.params(J=1.0, D=float(Ds[t]), dt=0.01), L, S0[t], nsteps)
        assert np.abs(spins[t].cpu().numpy() - ref).max() <= 1e-10


@pytest.mark.gpu
def test_batch_without_snapshots_repeat(gctx, orc):
    """Graph cache reuse across calls with different parameters in the table."""
    torch = need_gpu()
    import asd_amd
    Ls = [16, 40, 64]
    S0 = [asd_amd.init_spins(3 + t, L * L) for t, L in enumerate(Ls)]
    for D in (0.0, 0.15):
        plist = [asd_amd.params(D=D, Bz=0.1 * t, dt=0.01) for t in range(3)]
        spins = [torch.tensor(s, device="cuda") for s in S0]
        gctx.batch_rk4_run(Ls, plist, spins, 60)
        gctx.synchronize()
        for t, L in enumerate(Ls):
            ref, _ = orc.run(orc.params(D=D, Bz=0.1 * t, dt=0.01), L, S0[t], 60)
            assert np.abs(spins[t].cpu().numpy() - ref).max() <= 1e-10


@pytest.mark.gpu
def test_batch_ragged_sizes_bit_exact(gctx, orc):
    """Odd and ragged lattices (some smaller than the 24 x 24 loaded tile) in one batch,
    undamped J = 1 trajectories (the specialised arithmetic) next to damped J != 1 ones:
    every trajectory and snapshot equals the oracle's bit for bit."""
    torch = need_gpu()
    import asd_amd
    Ls = [9, 17, 33, 47, 100, 20]
    kws = [dict(D=0.1), dict(J=0.7, D=0.05, Bz=0.1, alpha=0.05), dict(D=0.2, Bz=0.02),
           dict(J=1.2, alpha=0.2), dict(D=0.1), dict(J=-0.5, D=0.1, alpha=0.01)]
    nsteps, every, nsnap = 30, 10, 3
    S0 = [orc.init_spins(80 + t, L * L) for t, L in enumerate(Ls)]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    snaps = [torch.zeros((nsnap, L * L, 3), dtype=torch.float64, device="cuda") for L in Ls]
    gctx.batch_rk4_run(Ls, [asd_amd.params(dt=0.01, **kw) for kw in kws], spins, nsteps, every, nsnap, snaps)
    gctx.synchronize()
    for t, (L, kw) in enumerate(zip(Ls, kws)):
        ref, ref_snaps = orc.run(orc.params(dt=0.01, **kw), L, S0[t], nsteps, every, nsnap)
        assert np.array_equal(spins[t].cpu().numpy(), ref), (t, L, kw)
        assert np.array_equal(snaps[t].cpu().numpy(), ref_snaps), (t, L, kw)
