"""Hand-written adjoint (asd_sqt_vjp / asd_sqt_loss_grad / torch wrapper) against
the oracle's FP64 tangent-linear model (tight) and central finite differences
(loose, the trajectory is chaotic)."""
import numpy as np
import pytest

from gpu_util import need_gpu, say


def gxmg(L):
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


@pytest.mark.gpu
@pytest.mark.parametrize("theta,ckpt", [((1.0, 0.1, 0.1, 0.01), 10), ((1.1, 0.12, 0.08, 0.0), 50),
                                        ((0.9, 0.05, 0.2, 0.1), 5)])
def test_vjp_matches_tangent_linear(gctx, orc, theta, ckpt):
    torch = need_gpu()
    import asd_amd
    L, every, T0, nlags = 8, 5, 6, 4
    nsnap = T0 + nlags - 1
    nsteps = (nsnap - 1) * every + 3            # trailing steps past the last snapshot
    q = gxmg(L)
    S0 = orc.init_spins(3, L * L)
    J, D, Bz, al = theta
    say(f"adjoint theta={theta} checkpoint={ckpt}")
    rng = np.random.default_rng(1)
    g = rng.standard_normal((nlags, len(q), 2))
    s_gpu, grad = gctx.sqt_vjp(asd_amd.params(J, D, Bz, al, 0.01), L, torch.tensor(S0, device="cuda"), nsteps,
                               every, T0, nlags, q, g=g, checkpoint_every=ckpt)
    snaps, dsnaps = orc.tlm_run(orc.params(J, D, Bz, al, 0.01), L, S0, nsteps, every, nsnap)
    s_ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    assert np.abs(s_gpu - s_ref).max() <= 1e-10 * np.abs(s_ref).max()
    tlm = np.array([np.sum(g[..., 0] * d.real + g[..., 1] * d.imag)
                    for d in (orc.tlm_sqt(L, snaps, dsnaps[i], T0, nlags, q) for i in range(4))])
    say(f"  adjoint {grad}\n  tangent {tlm}")
    assert np.abs(grad - tlm).max() <= 1e-9 * max(1.0, np.abs(tlm).max())


@pytest.mark.gpu
@pytest.mark.parametrize("L", [5, 12, 20])
def test_vjp_ragged_lattices_match_tangent_linear(gctx, orc, L):
    """Lattices that are not a multiple of the reverse step's 8 x 8 tiles (L = 12, 20: partial
    tiles at the far edges) or smaller than one tile with its halo (L = 5: the halo wraps
    around the lattice more than once), against the oracle's tangent-linear model."""
    torch = need_gpu()
    import asd_amd
    every, T0, nlags = 5, 6, 4
    nsnap = T0 + nlags - 1
    nsteps = (nsnap - 1) * every + 7
    q = [(k % L, (2 * k) % L) for k in range(5)]
    S0 = orc.init_spins(8 + L, L * L)
    theta = (0.9, 0.07, 0.12, 0.03)
    rng = np.random.default_rng(L)
    g = rng.standard_normal((nlags, len(q), 2))
    s_gpu, grad = gctx.sqt_vjp(asd_amd.params(*theta, 0.01), L, torch.tensor(S0, device="cuda"), nsteps, every, T0,
                               nlags, q, g=g, checkpoint_every=10)
    snaps, dsnaps = orc.tlm_run(orc.params(*theta, 0.01), L, S0, nsteps, every, nsnap)
    s_ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    assert np.abs(s_gpu - s_ref).max() <= 1e-10 * np.abs(s_ref).max()
    tlm = np.array([np.sum(g[..., 0] * d.real + g[..., 1] * d.imag)
                    for d in (orc.tlm_sqt(L, snaps, dsnaps[i], T0, nlags, q) for i in range(4))])
    say(f"  L={L}: adjoint {grad}\n  tangent {tlm}")
    assert np.abs(grad - tlm).max() <= 1e-9 * max(1.0, np.abs(tlm).max())


@pytest.mark.gpu
def test_loss_grad_against_finite_differences(gctx, orc):
    torch = need_gpu()
    import asd_amd
    L, every, T0, nlags = 16, 10, 20, 10
    nsteps = (T0 + nlags - 2) * every
    q = gxmg(L)
    S0 = torch.tensor(orc.init_spins(3, L * L), device="cuda")
    ts = (1.1, 0.12, 0.08, 0.01)
    target, _ = gctx.sqt_vjp(asd_amd.params(*ts, 0.01), L, S0, nsteps, every, T0, nlags, q)
    th = [1.0, 0.1, 0.1, 0.01]
    loss, grad = gctx.sqt_loss_grad(asd_amd.params(*th, 0.01), L, S0, nsteps, every, T0, nlags, q, target)
    h = 1e-6
    fd = []
    for i in range(4):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        sp, _ = gctx.sqt_vjp(asd_amd.params(*tp, 0.01), L, S0, nsteps, every, T0, nlags, q)
        sm, _ = gctx.sqt_vjp(asd_amd.params(*tm, 0.01), L, S0, nsteps, every, T0, nlags, q)
        fd.append((np.sum(np.abs(sp - target) ** 2) - np.sum(np.abs(sm - target) ** 2)) / (2 * h))
    fd = np.array(fd)
    say(f"  loss {loss:.6e}\n  adjoint {grad}\n  central FD {fd}")
    assert np.abs(grad - fd).max() <= 1e-4 * np.abs(fd).max()


@pytest.mark.gpu
@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.1, 0.01), (1.1, 0.0, 0.3, 0.2)])
def test_gradient_exactly_zero_at_a_fixed_point(gctx, orc, theta):
    """A uniform ferromagnet along z is a fixed point for every (J, D, B_z, alpha): the
    exchange, anisotropy and field terms are all parallel to S, so f and its derivatives with
    respect to the parameters vanish exactly (h x S has only products with a zero factor).
    The trajectory is static, S(q, tau) = 0 up to the rounding of the Fourier sums (the
    connected and disconnected terms cancel) and dL/dtheta = 0 exactly for any target,
    although the adjoint state, driven by S - S_target, is not zero.  The oracle's
    tangent-linear model agrees."""
    torch = need_gpu()
    import asd_amd
    L, every, T0, nlags = 8, 5, 6, 4
    nsnap = T0 + nlags - 1
    nsteps = (nsnap - 1) * every
    q = gxmg(L)
    S0 = np.zeros((L * L, 3))
    S0[:, 2] = 1.0
    rng = np.random.default_rng(4)
    target = rng.standard_normal((nlags, len(q))) + 1j * rng.standard_normal((nlags, len(q)))
    p = asd_amd.params(*theta, 0.01)
    loss, grad = gctx.sqt_loss_grad(p, L, torch.tensor(S0, device="cuda"), nsteps, every, T0, nlags, q, target,
                                    checkpoint_every=10)
    s, _ = gctx.sqt_vjp(p, L, torch.tensor(S0, device="cuda"), nsteps, every, T0, nlags, q)
    say(f"  fixed point theta={theta}: loss {loss:.6e}, grad {grad}")
    assert np.abs(s).max() <= 1e-13
    assert loss == pytest.approx(float(np.sum(np.abs(target) ** 2)), rel=1e-14)
    assert np.all(grad == 0.0)
    snaps, dsnaps = orc.tlm_run(orc.params(*theta, 0.01), L, S0, nsteps, every, nsnap)
    assert np.all(snaps == S0) and np.all(dsnaps == 0.0)


@pytest.mark.gpu
def test_torch_autograd_wrapper(gctx, orc):
    torch = need_gpu()
    import asd_amd
    from asd_amd.autograd import SqtConfig, sqt_of_theta
    L, every, T0, nlags = 8, 5, 6, 4
    nsteps = (T0 + nlags - 2) * every
    cfg = SqtConfig(ctx=gctx, L=L, nsteps=nsteps, every=every, T0=T0, nlags=nlags,
                    q=np.array(gxmg(L), dtype=np.int32), checkpoint_every=10)
    S0 = torch.tensor(orc.init_spins(5, L * L), device="cuda")
    theta = torch.tensor([1.0, 0.1, 0.1, 0.01], dtype=torch.float64, requires_grad=True)
    target = torch.zeros(nlags, len(cfg.q), 2, dtype=torch.float64)
    s = sqt_of_theta(theta, S0, cfg)
    loss = ((s - target) ** 2).sum()
    loss.backward()
    l2, g2 = gctx.sqt_loss_grad(asd_amd.params(1.0, 0.1, 0.1, 0.01, 0.01), L, S0, nsteps, every, T0, nlags, cfg.q,
                                np.zeros((nlags, len(cfg.q))), checkpoint_every=10)
    assert abs(loss.item() - l2) <= 1e-12 * abs(l2)
    assert np.abs(theta.grad.numpy() - g2).max() <= 1e-10 * np.abs(g2).max()


@pytest.mark.gpu
def test_differentiation_workload_full_size(gctx, orc):
    """The differentiation workload at its own size: L = 64, theta = (J, D, B_z, alpha) =
    (1.0, 0.1, 0.1, 0.01), target S(q,tau) from a forward run at theta* = (1.1, 0.12, 0.08,
    0.01), seed 3, 4000 RK4 steps at dt = 0.01, snapshots every 10, lags 0..199 over 200
    origins on Gamma-X-M-Gamma; loss = sum |S - S_target|^2; adjoint with checkpoints every 50
    steps.  Checked against the oracle's FP64 tangent-linear model over the same 4000 steps
    (the tight check: after 4000 chaotic steps |grad| is ~1e12 and central differences at
    h = 1e-6 no longer resolve a derivative)."""
    torch = need_gpu()
    import asd_amd
    L, nsteps, every, T0, nl = 64, 4000, 10, 200, 200
    q = gxmg(L)
    S0h = orc.init_spins(3, L * L)
    S0 = torch.tensor(S0h, device="cuda")
    ts, th = (1.1, 0.12, 0.08, 0.01), (1.0, 0.1, 0.1, 0.01)
    say(f"differentiation workload: L={L}, {nsteps} steps, T0={T0}, lags={nl}, checkpoint 50")
    target, _ = gctx.sqt_vjp(asd_amd.params(*ts, 0.01), L, S0, nsteps, every, T0, nl, q)
    loss, grad = gctx.sqt_loss_grad(asd_amd.params(*th, 0.01), L, S0, nsteps, every, T0, nl, q, target, 50)
    s_now, _ = gctx.sqt_vjp(asd_amd.params(*th, 0.01), L, S0, nsteps, every, T0, nl, q)
    snaps, dsnaps = orc.tlm_run(orc.params(*th, dt=0.01), L, S0h, nsteps, every, T0 + nl - 1)
    s_ref = orc.sqt_separable(L, snaps, T0, nl, q)    # the GEMM form's exact identity, O(Ns nq) per lag
    assert np.abs(s_now - s_ref).max() <= 1e-10 * np.abs(s_ref).max()
    R = s_now - target
    tlm = np.array([2 * np.sum((R.conj() * orc.tlm_sqt(L, snaps, dsnaps[i], T0, nl, q)).real) for i in range(4)])
    rel = float(np.abs(grad - tlm).max() / np.abs(tlm).max())
    say(f"  loss {loss:.6e}\n  adjoint {grad}\n  tangent {tlm}\n  max rel diff {rel:.3e}")
    assert rel <= 1e-9
