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


def _separable(snaps, L, T0, nlags, q):
    """FP64 separable form of S(q,tau) with numpy's FFT (qx with the row)."""
    nsnap = snaps.shape[0]
    F = np.fft.fft2(snaps.reshape(nsnap, L, L, 3), axes=(1, 2))
    qa = np.asarray(q)
    Fq = F[:, qa[:, 0], qa[:, 1], :]
    out = np.empty((nlags, len(q)), complex)
    for lag in range(nlags):
        a, b = Fq[lag:lag + T0], Fq[:T0]
        corr = np.einsum("tqc,tqc->q", a, b.conj()) / T0
        out[lag] = (corr - (a.mean(axis=0) * b.mean(axis=0).conj()).sum(axis=-1)) / (L * L)
    return out


@pytest.mark.gpu
def test_batching_workload_full_size(gctx):
    """The batching workload at its own size: 96 trajectories, L in {32, 48, 64, 96, 128}
    (32 / 24 / 20 / 12 / 8), D ~ U[0, 0.2], alpha = 0, seeds 100..195, 2000 FP64 RK4 steps with
    a snapshot every 10 in one persistent launch per step.  Every trajectory and snapshot
    equals its unbatched run bit for bit; then S(q,tau) for lags 0..99 over 100 origins on
    each lattice's Gamma-X-M-Gamma path in FP16X2 (autotuned per L) is within 1e-5 of the
    separable FP64 identity for all 96."""
    torch = need_gpu()
    import asd_amd
    Ls = [32] * 32 + [48] * 24 + [64] * 20 + [96] * 12 + [128] * 8
    Ds = np.random.default_rng(2024).uniform(0.0, 0.2, len(Ls))
    nsteps, every, T0, nl = 2000, 10, 100, 100
    nsnap = nsteps // every
    plist = [asd_amd.params(J=1.0, D=float(d), dt=0.01) for d in Ds]
    S0 = [asd_amd.init_spins(100 + t, L * L) for t, L in enumerate(Ls)]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    snaps = [torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda") for L in Ls]
    say(f"batch workload: {len(Ls)} trajectories, {nsteps} steps")
    gctx.batch_rk4_run(Ls, plist, spins, nsteps, every, nsnap, snaps)
    gctx.synchronize()
    for t, L in enumerate(Ls):
        s1 = torch.tensor(S0[t], device="cuda")
        sn1 = torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda")
        gctx.rk4_run(plist[t], L, s1, nsteps, every, nsnap, sn1)
        assert torch.equal(spins[t], s1), (t, L)
        assert torch.equal(snaps[t], sn1), (t, L)
    say("  batched trajectories equal their single runs")
    for prec in ("fp16x2", "fp16f8"):
        for L in sorted(set(Ls)):
            gctx.autotune(L, T0, prec)
    worst = {"fp16x2": 0.0, "fp16f8": 0.0}
    for t, L in enumerate(Ls):
        q = [(k, 0) for k in range(L // 2)] + [(L // 2, k) for k in range(L // 2)] + \
            [(L // 2 - k, L // 2 - k) for k in range(L // 2)]
        ref = _separable(snaps[t].cpu().numpy(), L, T0, nl, q)
        for prec in worst:
            got = gctx.sqt(L, snaps[t], T0, nl, q, prec)
            e = float((np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max())
            worst[prec] = max(worst[prec], e)
            assert e <= 1e-5, (t, L, prec, e)
    say(f"  S(q,tau) of all {len(Ls)} trajectories, worst max error: fp16x2 {worst['fp16x2']:.3e}, "
        f"fp16f8 {worst['fp16f8']:.3e}")


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps,every", [(45, 10), (40, 5), (31, 4)])
def test_every_batch_variant_bit_exact(gctx, orc, nsteps, every):
    """Each batch kernel variant forced in turn (tile edge 8 / 16 / 32, one or two fused steps
    per launch): every trajectory and snapshot equals the oracle bit for bit, with an odd
    trailing step (45), snapshot spacing that forbids fusion (every 5), and a short graph
    chunk (every 4)."""
    torch = need_gpu()
    import asd_amd
    Ls = [9, 17, 33, 47, 64, 20]
    kws = [dict(D=0.1), dict(J=0.7, D=0.05, Bz=0.1, alpha=0.05), dict(D=0.2, Bz=0.02),
           dict(J=1.2, alpha=0.2), dict(D=0.1), dict(J=-0.5, D=0.1, alpha=0.01)]
    nsnap = nsteps // every + 1
    S0 = [orc.init_spins(90 + t, L * L) for t, L in enumerate(Ls)]
    refs = [orc.run(orc.params(dt=0.01, **kw), L, S0[t], nsteps, every, nsnap) for t, (L, kw) in
            enumerate(zip(Ls, kws))]
    nv = asd_amd.lib().asd_batch_num_variants()
    try:
        for v in range(nv):
            gctx.batch_set_variant(v)
            spins = [torch.tensor(s, device="cuda") for s in S0]
            snaps = [torch.zeros((nsnap, L * L, 3), dtype=torch.float64, device="cuda") for L in Ls]
            gctx.batch_rk4_run(Ls, [asd_amd.params(dt=0.01, **kw) for kw in kws], spins, nsteps, every, nsnap,
                               snaps)
            for t in range(len(Ls)):
                assert np.array_equal(spins[t].cpu().numpy(), refs[t][0]), (v, t)
                assert np.array_equal(snaps[t].cpu().numpy(), refs[t][1]), (v, t)
    finally:
        gctx.batch_set_variant(-1)


@pytest.mark.gpu
def test_batch_autotuner_records_and_uses_its_choice(gctx, orc):
    torch = need_gpu()
    import asd_amd
    Ls = [32] * 6 + [48] * 4 + [96] * 2
    plist = [asd_amd.params(D=0.05 + 0.01 * t, dt=0.01) for t in range(len(Ls))]
    assert gctx.batch_variant_for(Ls) == -1
    best, times = gctx.autotune_batch(Ls, plist)
    say(f"batch autotune: best {best} ({list(times)[best]}), ms {[round(t, 4) for t in times.values()]}")
    assert all(np.isfinite(t) and t > 0 for t in times.values())
    vals = list(times.values())
    expect = 0                       # a later choice must be more than 1% faster than the best so far
    for i, t in enumerate(vals):
        if t < 0.99 * vals[expect]:
            expect = i
    assert best == expect and vals[best] <= 1.0102 * min(vals)
    assert gctx.batch_variant_for(Ls) == best
    assert gctx.batch_variant_for(list(reversed(Ls))) == best      # keyed on the multiset of L
    assert gctx.batch_variant_for([32, 48]) == -1
    # a run with the recorded choice equals the oracle bit for bit
    S0 = [orc.init_spins(300 + t, L * L) for t, L in enumerate(Ls)]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    gctx.batch_rk4_run(Ls, plist, spins, 60)
    for t, L in enumerate(Ls):
        ref, _ = orc.run(orc.params(D=0.05 + 0.01 * t, dt=0.01), L, S0[t], 60)
        assert np.array_equal(spins[t].cpu().numpy(), ref)
    # an untuned composition is tuned by its first run and recorded
    Ls2 = [40, 24]
    gctx.batch_rk4_run(Ls2, [asd_amd.params(dt=0.01)] * 2,
                       [torch.tensor(orc.init_spins(5, L * L), device="cuda") for L in Ls2], 4)
    assert gctx.batch_variant_for(Ls2) >= 0
