"""Numeric guard of the RK4 drivers (asd_rk4_run, asd_batch_rk4_run) and the argument checks
of the Python wrappers on device tensors.

The reference's own initial state is NaN (integer rand()/RAND_MAX, kernel_fusion/
llg_sqt_fusion.c:204-214) and its time loop carries the NaNs silently through RK4 and
S(q,t).  Here a step whose unnormalised state leaves the range where the step is defined
(|c|^2 NaN, infinite or outside [2^-900, 2^900]) makes the driver return ASD_ENOTFINITE
with a message naming the replica / trajectory and the first step.  The inputs below are
ordinary numbers (NaN components, a zero spin); every kernel runs to its normal end."""
import numpy as np
import pytest

from gpu_util import need_gpu, say


def _state(L, seed=4):
    import asd_amd
    return asd_amd.init_spins(seed, L * L)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["graph", "eager", "streaming", "streaming2"])
def test_nan_state_is_reported_at_step_zero(gctx, mode):
    """streaming2: the tolerance mode's two-step strips (automatic choice, 256 replicas)."""
    torch = need_gpu()
    import asd_amd
    L, nrep = {"streaming": (128, 2), "streaming2": (64, 256)}.get(mode, (64, 2))
    say(f"guard: NaN spin, {mode}, L={L}")
    S = np.tile(_state(L), (nrep, 1))
    S[L * L + 5] = np.nan                    # replica 1, site 5
    spins = torch.tensor(S, device="cuda")
    snaps = torch.zeros((6,) + spins.shape, dtype=torch.float64, device="cuda")
    if mode == "streaming":
        gctx.rk4_set_tile(64)
    if mode == "streaming2":
        gctx.rk4_set_exact(False)
    try:
        with pytest.raises(asd_amd.AsdError) as ei:
            if mode == "graph":
                gctx.rk4_run(asd_amd.params(D=0.1), L, spins, 50, 10, 6, snaps)
            else:
                gctx.rk4_run(asd_amd.params(D=0.1), L, spins, 7, use_graph=False)
    finally:
        gctx.rk4_set_tile(0)
        gctx.rk4_set_exact(True)
    assert ei.value.code == asd_amd.ASD_ENOTFINITE
    msg = str(ei.value)
    say(f"  {msg}")
    assert "replica 1" in msg and "step 0" in msg


@pytest.mark.gpu
def test_reference_nan_initial_state_is_reported(gctx):
    """The reference's all-NaN initial state (llg_sqt_fusion.c:204-214)."""
    torch = need_gpu()
    import asd_amd
    L = 32
    spins = torch.full((L * L, 3), float("nan"), dtype=torch.float64, device="cuda")
    with pytest.raises(asd_amd.AsdError) as ei:
        gctx.rk4_run(asd_amd.params(), L, spins, 20)
    assert ei.value.code == asd_amd.ASD_ENOTFINITE and "step 0" in str(ei.value)


@pytest.mark.gpu
def test_zero_spin_is_reported(gctx):
    """A zero spin has f = 0, so c = 0 and |c|^2 = 0 is below the range."""
    torch = need_gpu()
    import asd_amd
    L = 16
    S = _state(L)
    S[3] = 0.0
    with pytest.raises(asd_amd.AsdError) as ei:
        gctx.rk4_run(asd_amd.params(), L, torch.tensor(S, device="cuda"), 4, use_graph=False)
    assert ei.value.code == asd_amd.ASD_ENOTFINITE


@pytest.mark.gpu
def test_batch_reports_the_trajectory(gctx):
    torch = need_gpu()
    import asd_amd
    Ls = [16, 24, 32, 20]
    S0 = [_state(L, 10 + t) for t, L in enumerate(Ls)]
    S0[2][7] = [np.nan, 0.0, 1.0]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    snaps = [torch.zeros((5, L * L, 3), dtype=torch.float64, device="cuda") for L in Ls]
    plist = [asd_amd.params(D=0.05) for _ in Ls]
    say("guard: batch with a NaN spin in trajectory 2")
    for sn in (snaps, None):   # with snapshots (counter per interval) and without (per chunk)
        with pytest.raises(asd_amd.AsdError) as ei:
            gctx.batch_rk4_run(Ls, plist, spins, 40, 10, 5, sn)
        msg = str(ei.value)
        say(f"  {msg}")
        assert ei.value.code == asd_amd.ASD_ENOTFINITE and "trajectory 2" in msg and "step 0" in msg


@pytest.mark.gpu
def test_context_usable_after_a_reported_state(gctx, orc):
    """The guard word is reset per call: a finite run after a reported one succeeds and
    equals the oracle."""
    torch = need_gpu()
    import asd_amd
    L = 24
    bad = torch.full((L * L, 3), float("nan"), dtype=torch.float64, device="cuda")
    with pytest.raises(asd_amd.AsdError):
        gctx.rk4_run(asd_amd.params(), L, bad, 3)
    S0 = _state(L)
    s = torch.tensor(S0, device="cuda")
    gctx.rk4_run(asd_amd.params(D=0.1), L, s, 30)
    ref, _ = orc.run(orc.params(D=0.1), L, S0, 30)
    assert np.array_equal(s.cpu().numpy(), ref)


@pytest.mark.gpu
def test_wrappers_reject_host_and_missized_device_buffers(gctx):
    torch = need_gpu()
    import asd_amd
    L = 16
    spins = torch.tensor(_state(L), device="cuda")
    p = asd_amd.params()
    bad_snaps = [torch.zeros((3, L * L, 3), dtype=torch.float64),                  # on the host
                 torch.zeros((2, L * L, 3), dtype=torch.float64, device="cuda"),   # too small for nsnap = 3
                 torch.zeros((3, L * L, 3), dtype=torch.float32, device="cuda")]   # wrong dtype
    for fn in (gctx.rk4_run, gctx.rk4_run_noalloc):
        for sn in bad_snaps:
            with pytest.raises(ValueError):
                fn(p, L, spins, 20, 10, 3, sn)
        with pytest.raises(ValueError):
            fn(p, L + 1, spins, 20)                                             # size does not match L
    with pytest.raises(ValueError):
        gctx.sqt(L, torch.zeros((5, L * L, 3), dtype=torch.float64, device="cuda"), 2, 2, [(0, 0)],
                 out=torch.zeros((2, 1, 2), dtype=torch.float64))               # out on the host
    with pytest.raises(ValueError):
        gctx.sqt(L, torch.zeros((5, L * L, 3), dtype=torch.float64, device="cuda"), 2, 2, [(0, 0)],
                 out=torch.zeros((2, 2, 2), dtype=torch.float64, device="cuda"))  # out mis-sized
