"""The correlation-GEMM autotuner (asd_autotune_sqt): it times every tile configuration of a
precision for (L, K = 3 T0), records the fastest in the context, and later asd_sqt calls on
that (L, K) without a forced configuration use it.  Because the bins are order-independent
integer sums, "uses it" is checked bit for bit: the automatic call equals the call with
that configuration forced, and differs from no other configuration by more than the
precision's rounding.  The result is also checked against the FP64 oracle."""
import numpy as np
import pytest

from gpu_util import err_stats, need_gpu, say


def gxmg(L):
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bound", [("fp16x2", 1e-5), ("fp16f8", 1e-5), ("fp64", 1e-10), ("bf16x3", 5e-4)])
def test_autotuned_choice_is_recorded_and_used(gctx, orc, prec, bound):
    torch = need_gpu()
    import asd_amd
    L, T0, nl, every = 64, 60, 8, 5
    nsnap = T0 + nl - 1
    S0 = orc.init_spins(11, L * L)
    s = torch.tensor(S0, device="cuda")
    snaps = torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda")
    kw = dict(J=1.0, D=0.1, dt=0.01)
    gctx.rk4_run(asd_amd.params(**kw), L, s, (nsnap - 1) * every, every, nsnap, snaps)
    q = gxmg(L)
    best, times = gctx.autotune(L, T0, prec)
    names = list(times)
    say(f"autotune {prec} L={L} T0={T0}: best {best} ({names[best]}), ms {[round(t, 3) for t in times.values()]}")
    assert 0 <= best < len(names)
    assert all(np.isfinite(t) and t > 0 for t in times.values())
    vals = list(times.values())
    expect = 0                       # a later choice must be more than 1% faster than the best so far
    for i, t in enumerate(vals):
        if t < 0.99 * vals[expect]:
            expect = i
    assert best == expect and vals[best] <= 1.0102 * min(vals)
    auto = gctx.sqt(L, snaps, T0, nl, q, prec)
    try:
        gctx.set_config(prec, best)
        forced = gctx.sqt(L, snaps, T0, nl, q, prec)
    finally:
        gctx.set_config(prec, -1)
    assert np.array_equal(auto, forced)
    _, ref_snaps = orc.run(orc.params(**kw), L, S0, (nsnap - 1) * every, every, nsnap)
    ref = orc.sqt_separable(L, ref_snaps, T0, nl, q)
    mx, rms = err_stats(auto, ref)
    say(f"  against the FP64 oracle: max {mx:.3e} rms {rms:.3e}")
    assert mx <= bound
    # another (L, K) is not affected by this choice: it still runs (its default configuration or its own)
    other = gctx.sqt(L, snaps, T0 - 1, nl, q, prec)
    assert np.all(np.isfinite(other))
