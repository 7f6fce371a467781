"""Pin the oracle's LLG right-hand side and RK4 by closed forms and invariants.

These fix what the reference's own functions cannot (the D, B_z and alpha
terms are absent from kernel_fusion/llg_sqt_fusion.c:30-93): the exact
precessing-cone spin wave with J, D and B_z, its 4th-order convergence, the
ferromagnetic fixed point, conservation of M (and M_z) and energy at alpha = 0,
and monotone energy decay at alpha > 0.
"""
import numpy as np
import pytest


def cone_wave(L, mx, my, theta, t, J=1.0, D=0.0, Bz=0.0):
    """S_i = (sin th cos phi_i, sin th sin phi_i, cos th), phi_i = q.r_i - w t,
    w = cos th (J (2cos qx + 2cos qy - 4) - 2D) - B_z for dS/dt = h x S."""
    qx, qy = 2 * np.pi * mx / L, 2 * np.pi * my / L
    w = np.cos(theta) * (J * (2 * np.cos(qx) + 2 * np.cos(qy) - 4) - 2 * D) - Bz
    r, c = np.divmod(np.arange(L * L), L)
    phi = qx * r + qy * c - w * t
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                     np.full(L * L, np.cos(theta))], axis=-1)


def test_init_unit_length_and_moments(orc):
    S = orc.init_spins(0, 200_000)
    assert np.abs(np.linalg.norm(S, axis=1) - 1).max() < 4e-16
    assert np.abs(S.mean(axis=0)).max() < 5e-3          # E[S] = 0
    cov = S.T @ S / len(S)
    assert np.abs(cov - np.eye(3) / 3).max() < 5e-3     # E[S S^T] = I/3
    assert np.array_equal(S[:10], orc.init_spins(0, 10))  # counter streams: prefix-stable
    assert not np.array_equal(orc.init_spins(1, 10), S[:10])


def test_init_z_uniform(orc):
    """Archimedes: S_z of a uniform point on S^2 is uniform on [-1, 1]."""
    z = orc.init_spins(3, 100_000)[:, 2]
    hist, _ = np.histogram(z, bins=10, range=(-1, 1))
    assert np.abs(hist / len(z) - 0.1).max() < 0.01


def test_ferromagnet_is_fixed_point(orc):
    L = 8
    S = np.tile(np.array([0.6, 0.0, 0.8]), (L * L, 1))
    f = orc.llg_rhs(orc.params(J=1.3), L, S)
    assert np.abs(f).max() < 1e-15
    # along z with anisotropy and field: still a fixed point
    Sz = np.tile(np.array([0.0, 0.0, 1.0]), (L * L, 1))
    assert np.abs(orc.llg_rhs(orc.params(D=0.3, Bz=0.2, alpha=0.1), L, Sz)).max() == 0.0


@pytest.mark.parametrize("J,D,Bz", [(1.0, 0.0, 0.0), (1.0, 0.1, 0.0), (0.7, 0.05, 0.2)])
def test_rhs_matches_cone_derivative(orc, J, D, Bz):
    L, mx, my, th = 8, 1, 2, 0.7
    p = orc.params(J=J, D=D, Bz=Bz)
    h = 1e-6
    num = (cone_wave(L, mx, my, th, h, J, D, Bz) - cone_wave(L, mx, my, th, -h, J, D, Bz)) / (2 * h)
    f = orc.llg_rhs(p, L, cone_wave(L, mx, my, th, 0.0, J, D, Bz))
    assert np.abs(f - num).max() < 1e-8


@pytest.mark.parametrize("J,D,Bz", [(1.0, 0.0, 0.0), (1.0, 0.12, 0.08)])
def test_rk4_fourth_order_on_cone(orc, J, D, Bz):
    L, mx, my, th, T = 8, 1, 3, 0.9, 2.0
    errs = []
    for dt in (0.04, 0.02):
        n = int(round(T / dt))
        S, _ = orc.run(orc.params(J=J, D=D, Bz=Bz, dt=dt), L, cone_wave(L, mx, my, th, 0, J, D, Bz), n)
        errs.append(np.abs(S - cone_wave(L, mx, my, th, T, J, D, Bz)).max())
    ratio = errs[0] / errs[1]
    assert 12 < ratio < 20, (errs, ratio)     # 2^4 = 16
    assert errs[1] < 1e-6


def test_magnetisation_conserved(orc):
    """sum_i h_i x S_i = 0 for pure exchange (pairs cancel), so M is conserved
    by every RK4 stage; the renormalisation moves |S| by O(dt^5) per step."""
    L = 12
    S0 = orc.init_spins(9, L * L)
    p = orc.params(dt=0.01)
    f = orc.llg_rhs(p, L, S0)
    assert np.abs(f.sum(axis=0)).max() < 1e-13
    drift = []
    for dt, n in ((0.01, 200), (0.005, 400)):
        S, _ = orc.run(orc.params(dt=dt), L, S0, n)
        drift.append(np.abs(S.sum(axis=0) - S0.sum(axis=0)).max())
    assert drift[0] < 1e-7
    assert drift[0] / drift[1] > 10          # O(dt^4) globally


def test_mz_and_energy_conserved_with_anisotropy(orc):
    L = 10
    S0 = orc.init_spins(2, L * L)
    p = orc.params(D=0.1, Bz=0.2, dt=0.005)
    S, _ = orc.run(p, L, S0, 400)
    assert abs(S[:, 2].sum() - S0[:, 2].sum()) < 1e-8     # O(dt^4) renormalisation drift
    e0, e1 = orc.energy(p, L, S0), orc.energy(p, L, S)
    assert abs(e1 - e0) < 1e-7 * abs(e0)


def test_energy_decays_with_damping(orc):
    L = 8
    S = orc.init_spins(5, L * L)
    p = orc.params(D=0.1, Bz=0.3, alpha=0.1, dt=0.01)
    es = [orc.energy(p, L, S)]
    for _ in range(20):
        S, _ = orc.run(p, L, S, 10)
        es.append(orc.energy(p, L, S))
    assert np.all(np.diff(es) < 0)


def test_damping_single_spin_relaxes_toward_field(orc):
    """Isolated spin in a field (J = 0): S_z relaxes to +1 at rate ~ alpha B/(1+alpha^2)."""
    L = 2
    S = np.tile(np.array([1.0, 0.0, 0.0]), (4, 1))
    p = orc.params(J=0.0, Bz=1.0, alpha=0.2, dt=0.01)
    S1, _ = orc.run(p, L, S, 1000)    # t = 10
    # exact: cos(theta(t)) = tanh(alpha B t / (1+alpha^2)) starting at theta = pi/2
    expect = np.tanh(0.2 * 1.0 * 10.0 / (1 + 0.04))
    assert abs(S1[0, 2] - expect) < 1e-8


def test_renormalise_preserves_direction(orc):
    rng = np.random.default_rng(0)
    cur = rng.standard_normal((100, 3))
    S = orc.renormalise(cur, 1.5)
    assert np.abs(np.linalg.norm(S, axis=1) - 1.5).max() < 1e-15
    assert np.abs(np.cross(S, cur)).max() < 1e-14
    assert np.all((S * cur).sum(axis=1) > 0)


def test_snapshot_convention(orc):
    """snapshot k is the state after k * snap_every steps (k = 0 is the input)."""
    L = 4
    S0 = orc.init_spins(1, L * L)
    p = orc.params(dt=0.01)
    _, snaps = orc.run(p, L, S0, 30, 10, 4)
    assert np.array_equal(snaps[0], S0)
    for k in (1, 2, 3):
        Sk, _ = orc.run(p, L, S0, 10 * k)
        assert np.array_equal(snaps[k], Sk)
