"""Run-to-run reproducibility: the displacement bins of S(q,tau) are fixed-point integer sums
(order-independent), and the adjoint's parameter gradient is reduced in a fixed order, so
repeated calls on the same snapshots give the same bits in every precision.

The full-size case is the paper-scale plain workload: L = 128, J = 1, D = 0.1, seed 1, 8000
RK4 steps with a snapshot every 10, 400 lags over 400 origins on Gamma-X-M-Gamma."""
import numpy as np
import pytest

from gpu_util import need_gpu, say


def gxmg(L):
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


@pytest.mark.gpu
def test_plain_L128_sqt_repeats_bit_for_bit(gctx):
    torch = need_gpu()
    import asd_amd
    L, T0, nl, every, nsteps = 128, 400, 400, 10, 8000
    say(f"determinism: L={L}, {nsteps} steps, T0={T0}, lags={nl}")
    s = torch.tensor(asd_amd.init_spins(1, L * L), device="cuda")
    snaps = torch.empty((nsteps // every, L * L, 3), dtype=torch.float64, device="cuda")
    gctx.rk4_run(asd_amd.params(J=1.0, D=0.1), L, s, nsteps, every, nsteps // every, snaps)
    q = gxmg(L)
    for prec in ("fp16x2", "fp64"):
        a = gctx.sqt(L, snaps, T0, nl, q, prec)
        b = gctx.sqt(L, snaps, T0, nl, q, prec)
        say(f"  {prec}: max |a - b| = {np.abs(a - b).max():.3e}")
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x2", "fp16f8", "fp64"])
def test_every_tile_configuration_repeats(gctx, prec):
    """Each configuration of the precision twice on the same snapshots (L = 40: ragged tiles)."""
    torch = need_gpu()
    import asd_amd
    L, T0, nl, every = 40, 30, 12, 5
    nsnap = T0 + nl - 1
    s = torch.tensor(asd_amd.init_spins(9, L * L), device="cuda")
    snaps = torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda")
    gctx.rk4_run(asd_amd.params(D=0.1, Bz=0.05), L, s, (nsnap - 1) * every, every, nsnap, snaps)
    q = gxmg(L)
    n = asd_amd.lib().asd_sqt_num_configs(asd_amd.PRECISIONS[prec])
    try:
        for cfg in range(n):
            gctx.set_config(prec, cfg)
            a = gctx.sqt(L, snaps, T0, nl, q, prec)
            b = gctx.sqt(L, snaps, T0, nl, q, prec)
            assert np.array_equal(a, b), (prec, cfg)
    finally:
        gctx.set_config(prec, -1)


@pytest.mark.gpu
def test_adjoint_gradient_repeats(gctx, orc):
    torch = need_gpu()
    import asd_amd
    L, every, T0, nlags = 32, 10, 20, 10
    nsteps = (T0 + nlags - 2) * every
    S0 = torch.tensor(orc.init_spins(3, L * L), device="cuda")
    q = gxmg(L)
    g = np.random.default_rng(0).standard_normal((nlags, len(q), 2))
    p = asd_amd.params(1.0, 0.1, 0.1, 0.01, 0.01)
    _, g1 = gctx.sqt_vjp(p, L, S0, nsteps, every, T0, nlags, q, g=g)
    _, g2 = gctx.sqt_vjp(p, L, S0, nsteps, every, T0, nlags, q, g=g)
    say(f"  adjoint gradient {g1}")
    assert np.array_equal(g1, g2)


@pytest.mark.gpu
def test_vjp_with_zero_steps(gctx):
    """nsteps = 0 with one origin and one lag: the only snapshot is the initial state and
    S(q, 0) of a single origin is 0 (connected minus disconnected term, to rounding); the
    gradient is exactly 0 (no steps, and the adjoint source F(t0) - Fbar(0) vanishes)."""
    torch = need_gpu()
    import asd_amd
    L = 8
    S0 = torch.tensor(asd_amd.init_spins(2, L * L), device="cuda")
    q = gxmg(L)
    s, grad = gctx.sqt_vjp(asd_amd.params(), L, S0, 0, 1, 1, 1, q, g=np.ones((1, len(q), 2)))
    assert np.all(np.isfinite(s)) and np.abs(s).max() <= 1e-11
    assert np.all(grad == 0.0)
