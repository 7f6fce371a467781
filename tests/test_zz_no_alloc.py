"""asd_rk4_run_noalloc: the RK4 driver without any device allocation (its second state
buffer is a fixed 16 MiB scratch array in the library).  Through the C ABI it must give
the same bits as asd_rk4_run (eager launches) on two calls in a row, the second with a
smaller input (the scratch keeps the first call's contents, which must not leak into the
second), reject inputs beyond its scratch, and leave the device's free memory unchanged."""
import numpy as np
import pytest

from gpu_util import need_gpu, say


def _run_pair(gctx, L, nrep, nsteps, every, nsnap, seed, **kw):
    torch = need_gpu()
    import asd_amd
    S0 = np.tile(asd_amd.init_spins(seed, L * L), (nrep, 1))
    p = asd_amd.params(dt=0.01, **kw)
    a = torch.tensor(S0, device="cuda")
    b = torch.tensor(S0, device="cuda")
    sa = torch.full((nsnap, nrep * L * L, 3), np.nan, dtype=torch.float64, device="cuda")
    sb = torch.full((nsnap, nrep * L * L, 3), np.nan, dtype=torch.float64, device="cuda")
    gctx.rk4_run(p, L, a, nsteps, every, nsnap, sa, use_graph=False)
    gctx.rk4_run_noalloc(p, L, b, nsteps, every, nsnap, sb)
    gctx.synchronize()
    return a.cpu().numpy(), b.cpu().numpy(), sa.cpu().numpy(), sb.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_equals_rk4_run_on_two_calls(gctx):
    # first call: L = 512, 2 replicas (1.57e6 doubles of the 2^21 scratch), 13 steps (odd:
    # the final state comes back from the scratch); second call smaller: L = 48, 5 steps
    for L, nrep, nsteps, every, nsnap, seed, kw in ((512, 2, 13, 4, 4, 21, dict(D=0.1)),
                                                    (48, 1, 5, 2, 3, 22, dict(J=0.8, D=0.05, Bz=0.1,
                                                                              alpha=0.02))):
        say(f"noalloc: L={L} nrep={nrep} steps={nsteps}")
        a, b, sa, sb = _run_pair(gctx, L, nrep, nsteps, every, nsnap, seed, **kw)
        assert np.array_equal(a, b)
        written = min(nsnap, nsteps // every + 1)
        assert not np.isnan(sb[:written]).any()
        assert np.array_equal(sa, sb, equal_nan=True)


@pytest.mark.gpu
def test_noalloc_rejects_inputs_beyond_its_scratch(gctx):
    torch = need_gpu()
    import asd_amd
    L = 1024   # 3 L^2 = 3.1e6 doubles > 2^21
    s = torch.zeros((L * L, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(asd_amd.AsdError) as ei:
        gctx.rk4_run_noalloc(asd_amd.params(), L, s, 1)
    assert ei.value.code == asd_amd.ASD_EINVAL and "scratch" in str(ei.value)


@pytest.mark.gpu
def test_noalloc_leaves_free_memory_unchanged(gctx):
    torch = need_gpu()
    import asd_amd
    L = 256
    p = asd_amd.params(D=0.1)
    s = torch.tensor(asd_amd.init_spins(5, L * L), device="cuda")
    gctx.rk4_run_noalloc(p, L, s, 3)          # first launches load the kernels
    gctx.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    gctx.rk4_run_noalloc(p, L, s, 7)
    gctx.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    say(f"noalloc: free memory {free0} -> {free1}")
    assert free1 == free0
