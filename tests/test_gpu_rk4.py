"""RK4-LLG on the GPU (asd_rk4_run) against the FP64 oracle.

Bar: trajectories within 1e-10 of the oracle over the first 100
steps and |S_i| = spin_length within 1e-12.  The kernel keeps the reference's
rounding steps, so the comparison is also reported as bit-exact or not.
"""
import numpy as np
import pytest

from gpu_util import need_gpu, say

CASES = [
    # L, nrep, params
    (8, 1, dict()),                                        # lattice smaller than tile + halo
    (16, 3, dict()),                                       # replicas (reference Nfile)
    (33, 1, dict(D=0.1, Bz=0.05)),                         # odd L, ragged tiles
    (64, 2, dict(D=0.1, Bz=0.1, alpha=0.01)),
    (128, 1, dict(D=0.1)),                                 # config-2 lattice
    (48, 1, dict(J=-0.7, D=0.05, Bz=0.2, alpha=0.2, spin_length=1.5)),
    # enough 32 x 32 tiles to take the large-tile kernel: one big lattice, and a ragged
    # L = 40 (32 + 8) in 200 replicas
    (768, 1, dict(D=0.1, alpha=0.05)),
    (40, 200, dict(D=0.1, Bz=0.1)),
    # 256 tiles of 16 x 16: the automatic choice takes 16 x 16 tiles with 1024 threads
    (64, 16, dict(D=0.1, Bz=0.05, alpha=0.01)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep,kw", CASES)
def test_rk4_trajectory_matches_oracle(gctx, orc, L, nrep, kw):
    torch = need_gpu()
    import asd_amd
    say(f"rk4 L={L} nrep={nrep} {kw}")
    S0 = np.stack([orc.init_spins(7 + r, L * L) for r in range(nrep)])
    length = kw.get("spin_length", 1.0)
    S0 = S0 * length
    p_o = orc.params(dt=0.01, **kw)
    p_g = asd_amd.params(dt=0.01, **kw)
    nsteps, every, nsnap = (100, 10, 11) if L * L * nrep < 200_000 else (20, 10, 3)
    S_ref, snaps_ref = orc.run(p_o, L, S0, nsteps, every, nsnap)
    spins = torch.tensor(S0, device="cuda")
    snaps = torch.zeros((nsnap,) + S0.shape, dtype=torch.float64, device="cuda")
    gctx.rk4_run(p_g, L, spins, nsteps, every, nsnap, snaps)
    gctx.synchronize()
    S_gpu = spins.cpu().numpy()
    err = np.abs(S_gpu - S_ref).max()
    say(f"  max |S_gpu - S_oracle| = {err:.3e} (bit-exact: {np.array_equal(S_gpu, S_ref)})")
    assert err <= 1e-10
    assert np.abs(snaps.cpu().numpy() - snaps_ref).max() <= 1e-10
    assert np.abs(np.linalg.norm(S_gpu, axis=-1) - length).max() <= 1e-12


@pytest.mark.gpu
def test_graph_replay_equals_eager(gctx, orc):
    torch = need_gpu()
    import asd_amd
    L, n, every, nsnap = 40, 250, 10, 26
    p = asd_amd.params(D=0.1, Bz=0.02, dt=0.01)
    S0 = torch.tensor(orc.init_spins(5, L * L), device="cuda")
    res = []
    for graph in (True, False, True):            # third run replays the cached graph
        s = S0.clone()
        snaps = torch.full((nsnap, L * L, 3), np.nan, dtype=torch.float64, device="cuda")
        gctx.rk4_run(p, L, s, n, every, nsnap, snaps, use_graph=graph)
        gctx.synchronize()
        res.append((s.cpu().numpy(), snaps.cpu().numpy()))
    for s, sn in res[1:]:
        assert np.array_equal(s, res[0][0])
        assert np.array_equal(sn, res[0][1])
    assert not np.isnan(res[0][1]).any()          # all 26 snapshots written (25 in loop + final)
    _, sref = orc.run(orc.params(D=0.1, Bz=0.02, dt=0.01), L, orc.init_spins(5, L * L), n, every, nsnap)
    assert np.abs(res[0][1] - sref).max() <= 1e-10


@pytest.mark.gpu
def test_full_size_invariants(gctx, orc):
    """Config-2 lattice (L=128), 8000 steps: |S| = 1, M_z conserved (alpha = 0),
    energy conserved to RK4 accuracy."""
    torch = need_gpu()
    import asd_amd
    L = 128
    p = asd_amd.params(D=0.1, dt=0.01)
    S0 = orc.init_spins(1, L * L)
    s = torch.tensor(S0, device="cuda")
    gctx.rk4_run(p, L, s, 8000, 10, 0, None)
    gctx.synchronize()
    S = s.cpu().numpy()
    assert np.abs(np.linalg.norm(S, axis=1) - 1).max() <= 1e-12
    po = orc.params(D=0.1, dt=0.01)
    e0, e1 = orc.energy(po, L, S0), orc.energy(po, L, S)
    say(f"  Mz drift {abs(S[:, 2].sum() - S0[:, 2].sum()):.3e}  energy rel drift {abs(e1 - e0) / abs(e0):.3e}")
    assert abs(S[:, 2].sum() - S0[:, 2].sum()) < 1e-6 * L * L
    assert abs(e1 - e0) < 1e-5 * abs(e0)


@pytest.mark.gpu
def test_tile_edge_does_not_change_results(gctx, orc):
    torch = need_gpu()
    import asd_amd
    L = 72
    S0 = torch.tensor(orc.init_spins(8, L * L), device="cuda")
    p = asd_amd.params(D=0.1, Bz=0.05, alpha=0.02, dt=0.01)
    out = []
    for tile in (16, 32, 0, 64, 8):
        gctx.rk4_set_tile(tile)
        s = S0.clone()
        gctx.rk4_run(p, L, s, 30, 10, 0, None, use_graph=False)
        gctx.synchronize()
        out.append(s.cpu().numpy())
    gctx.rk4_set_tile(0)
    assert all(np.array_equal(out[0], o) for o in out[1:])


STREAM_CASES = [
    # L, nrep, params: strips of 56 columns and bands of min(56, L - 8) rows (these small
    # batches; large ones take 104-row bands), ragged at both ends; L = 48 is below the
    # 64-column wave and takes the tile kernel
    (64, 2, dict(D=0.1, Bz=0.1, alpha=0.01)),
    (100, 1, dict(D=0.1, Bz=0.05)),
    (130, 3, dict(J=-0.7, D=0.05, Bz=0.2, alpha=0.2, spin_length=1.5)),
    (257, 1, dict(D=0.1, alpha=0.05)),
    (128, 1, dict(D=0.1)),
    (48, 1, dict(D=0.1)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep,kw", STREAM_CASES)
def test_streaming_kernel_bit_exact(gctx, orc, L, nrep, kw):
    """The register-streaming step (asd_rk4_set_tile 64) gives the oracle's trajectory and
    snapshots bit for bit."""
    torch = need_gpu()
    import asd_amd
    say(f"rk4 streaming L={L} nrep={nrep} {kw}")
    S0 = np.stack([orc.init_spins(11 + r, L * L) for r in range(nrep)]) * kw.get("spin_length", 1.0)
    nsteps, every, nsnap = 40, 10, 5
    S_ref, snaps_ref = orc.run(orc.params(dt=0.01, **kw), L, S0, nsteps, every, nsnap)
    spins = torch.tensor(S0, device="cuda")
    snaps = torch.zeros((nsnap,) + S0.shape, dtype=torch.float64, device="cuda")
    gctx.rk4_set_tile(64)
    try:
        gctx.rk4_run(asd_amd.params(dt=0.01, **kw), L, spins, nsteps, every, nsnap, snaps, use_graph=False)
        gctx.synchronize()
    finally:
        gctx.rk4_set_tile(0)
    S_gpu = spins.cpu().numpy()
    say(f"  max |S_gpu - S_oracle| = {np.abs(S_gpu - S_ref).max():.3e}")
    assert np.array_equal(S_gpu, S_ref)
    assert np.array_equal(snaps.cpu().numpy(), snaps_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep", [(100, 12), (128, 8)])
def test_automatic_tile_shapes_bit_exact(gctx, orc, L, nrep):
    """Replica batches whose automatic choice is 32 x 32 tiles (L = 100 x 12: 192 tiles of
    32 x 32, ragged: 100 = 3 x 32 + 4) and 16 x 16 tiles with four threads per site
    (L = 128 x 8) give the oracle's trajectory and snapshots bit for bit."""
    torch = need_gpu()
    import asd_amd
    kw = dict(J=0.9, D=0.1, Bz=0.05, alpha=0.03)
    say(f"rk4 automatic tiles L={L} nrep={nrep} {kw}")
    S0 = np.stack([orc.init_spins(21 + r, L * L) for r in range(nrep)])
    nsteps, every, nsnap = 30, 10, 3
    S_ref, snaps_ref = orc.run(orc.params(dt=0.01, **kw), L, S0, nsteps, every, nsnap)
    spins = torch.tensor(S0, device="cuda")
    snaps = torch.zeros((nsnap,) + S0.shape, dtype=torch.float64, device="cuda")
    gctx.rk4_run(asd_amd.params(dt=0.01, **kw), L, spins, nsteps, every, nsnap, snaps)
    gctx.synchronize()
    assert np.array_equal(spins.cpu().numpy(), S_ref)
    assert np.array_equal(snaps.cpu().numpy(), snaps_ref)


@pytest.mark.gpu
def test_automatic_choice_takes_streaming_for_many_replicas(gctx, orc):
    """At L = 128 x 180 replicas (1080 waves) the automatic choice runs the streaming
    strips; the result equals both forced kernels bit for bit."""
    torch = need_gpu()
    import asd_amd
    L, nrep = 128, 180
    S0 = torch.tensor(np.stack([orc.init_spins(40 + r % 7, L * L) for r in range(nrep)]), device="cuda")
    p = asd_amd.params(D=0.1, dt=0.01)
    out = []
    try:
        for tile in (0, 16, 64):
            gctx.rk4_set_tile(tile)
            s = S0.clone()
            gctx.rk4_run(p, L, s, 20, 10, 0, None, use_graph=False)
            gctx.synchronize()
            out.append(s.cpu().numpy())
    finally:
        gctx.rk4_set_tile(0)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


@pytest.mark.gpu
def test_spin_length_outside_the_kernels_range_is_rejected(gctx, orc):
    """The streaming kernel's renormalisation assumes |S|^2 within 2^(+-900) of 1; the
    parameter check keeps spin_length in [1e-100, 1e100] and reports anything else."""
    torch = need_gpu()
    import asd_amd
    L = 64
    s = torch.tensor(orc.init_spins(3, L * L), device="cuda")
    for bad in (1e101, 1e-101, 0.0, -1.0):
        with pytest.raises(asd_amd.AsdError, match="spin_length"):
            gctx.rk4_run(asd_amd.params(dt=0.01, spin_length=bad), L, s, 1, 1, 0, None)
    # the ends of the range run (J scaled so the precession per step stays O(dt))
    for ok in (1e-100, 1e100):
        s2 = s * ok
        gctx.rk4_run(asd_amd.params(J=1.0 / ok, dt=0.01, spin_length=ok), L, s2, 2, 1, 0, None)
        gctx.synchronize()
        assert np.allclose(torch.linalg.norm(s2, dim=1).cpu().numpy() / ok, 1.0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep,kw", [(64, 1, dict(D=0.1, Bz=0.05)), (40, 3, dict(J=-0.7, D=0.05, alpha=0.2)),
                                       (128, 1, dict(D=0.1)), (256, 1, dict(D=0.05, Bz=0.2)),
                                       (320, 1, dict(J=0.8, D=0.1, alpha=0.05)), (448, 1, dict(D=0.1, Bz=0.05))])
def test_fused_two_step_launches_equal_single_steps(gctx, orc, L, nrep, kw):
    """On the 8 x 8, 16 x 16 / 1024-thread (L = 256, 320) and 32 x 32 (L = 448) tile paths a
    launch fuses two RK4 steps (snapshot interval even); with an odd interval every launch
    is one step.
    Both give the same trajectory and the same snapshots at common times, bit for bit,
    graph-replayed and eager, and the oracle's."""
    torch = need_gpu()
    import asd_amd
    S0 = np.stack([orc.init_spins(21 + r, L * L) for r in range(nrep)])
    p = asd_amd.params(dt=0.01, **kw)
    n = 103                                                 # odd: the last launch is one step
    res = {}
    for every, graph in ((10, True), (10, False), (5, True), (5, False)):
        nsnap = n // every + 1
        s = torch.tensor(S0, device="cuda")
        snaps = torch.full((nsnap,) + S0.shape, np.nan, dtype=torch.float64, device="cuda")
        gctx.rk4_run(p, L, s, n, every, nsnap, snaps, use_graph=graph)
        gctx.synchronize()
        res[(every, graph)] = (s.cpu().numpy(), snaps.cpu().numpy())
    ref = res[(5, False)]
    for key, (s, sn) in res.items():
        assert np.array_equal(s, ref[0]), key
        step = key[0] // 5
        assert np.array_equal(sn, ref[1][::step][: len(sn)]), key
    S_or, _ = orc.run(orc.params(dt=0.01, **kw), L, S0, n, 0, 0)
    assert np.array_equal(ref[0], S_or)


@pytest.mark.gpu
def test_plain_L128_1000_steps_equal_oracle(gctx, orc):
    """The benchmark lattice (L = 128, J = 1, D = 0.1, seed 1; two-step fused launches,
    graph-replayed with snapshots every 10) over 1000 steps: trajectory and every snapshot
    equal to the oracle's."""
    torch = need_gpu()
    import asd_amd
    L, n, every = 128, 1000, 10
    S0 = orc.init_spins(1, L * L)
    s = torch.tensor(S0, device="cuda")
    snaps = torch.empty((n // every, L * L, 3), dtype=torch.float64, device="cuda")
    gctx.rk4_run(asd_amd.params(J=1.0, D=0.1, dt=0.01), L, s, n, every, n // every, snaps)
    gctx.synchronize()
    S_ref, sn_ref = orc.run(orc.params(J=1.0, D=0.1, dt=0.01), L, S0, n, every, n // every)
    assert np.array_equal(s.cpu().numpy(), S_ref)
    assert np.array_equal(snaps.cpu().numpy(), sn_ref)


@pytest.mark.gpu
def test_streaming_equals_tiles_on_the_large_batch(gctx, orc):
    """L = 1024 x 24 replicas (the automatic choice streams, in 104-row bands: the small
    batches of the other streaming tests take 56-row bands): 10 steps equal to the 16 x 16
    tile kernel's and to the oracle's on the first replica."""
    torch = need_gpu()
    import asd_amd
    L, nrep, n = 1024, 24, 10
    S0 = torch.tensor(np.stack([orc.init_spins(60 + r, L * L) for r in range(nrep)]), device="cuda")
    p = asd_amd.params(J=1.0, D=0.1, Bz=0.02, dt=0.01)
    out = []
    try:
        for tile in (0, 16):
            gctx.rk4_set_tile(tile)
            s = S0.clone()
            gctx.rk4_run(p, L, s, n, 10, 0, None, use_graph=False)
            gctx.synchronize()
            out.append(s.cpu().numpy())
    finally:
        gctx.rk4_set_tile(0)
    assert np.array_equal(out[0], out[1])
    S_ref, _ = orc.run(orc.params(J=1.0, D=0.1, Bz=0.02, dt=0.01), L, S0[0].cpu().numpy(), n, 0, 0)
    assert np.array_equal(out[0][0], S_ref)


@pytest.mark.gpu
def test_zero_and_one_step_and_unused_snapshot_slots(gctx, orc):
    """nsteps = 0 leaves the state and writes snapshot 0 only; nsteps = 1 with every = 1
    writes the initial and the stepped state; slots beyond the run stay untouched."""
    torch = need_gpu()
    import asd_amd
    L = 32
    S0 = orc.init_spins(4, L * L)
    p = asd_amd.params(D=0.1, dt=0.01)
    for n, every in ((0, 1), (1, 1), (25, 10)):
        s = torch.tensor(S0, device="cuda")
        snaps = torch.full((5, L * L, 3), np.nan, dtype=torch.float64, device="cuda")
        gctx.rk4_run(p, L, s, n, every, 5, snaps)
        gctx.synchronize()
        S_ref, sn_ref = orc.run(orc.params(D=0.1, dt=0.01), L, S0, n, every, n // every + 1)
        used = n // every + 1
        assert np.array_equal(s.cpu().numpy(), S_ref), (n, every)
        assert np.array_equal(snaps[:used].cpu().numpy(), sn_ref), (n, every)
        assert np.isnan(snaps[used:].cpu().numpy()).all(), (n, every)


SWEEP_L = [9, 17, 23, 31, 33, 47, 65, 100, 129]
SWEEP_PARAMS = [dict(D=0.1), dict(J=-0.8, D=0.05, Bz=0.2, alpha=0.1, spin_length=1.3)]


@pytest.mark.gpu
@pytest.mark.parametrize("L", SWEEP_L)
def test_every_tile_shape_bit_exact_sweep(gctx, orc, L):
    """Every kernel choice (automatic, forced 8 / 16 / 32 tiles, streaming strips from
    L = 64) on ragged and odd lattices, including ones smaller than a loaded tile (the
    wrapped-coordinate load), gives the oracle's trajectory and snapshots bit for bit,
    undamped with J = 1 and damped with J != 1, B_z and spin_length != 1."""
    torch = need_gpu()
    import asd_amd
    nsteps, every, nsnap = 24, 6, 4
    tiles = (0, 8, 16, 32) + ((64,) if L >= 64 else ())
    try:
        for k, kw in enumerate(SWEEP_PARAMS):
            S0 = orc.init_spins(60 + L + k, L * L)[None] * kw.get("spin_length", 1.0)
            S_ref, snaps_ref = orc.run(orc.params(dt=0.01, **kw), L, S0, nsteps, every, nsnap)
            for tile in tiles:
                say(f"rk4 sweep L={L} tile={tile} {kw}")
                gctx.rk4_set_tile(tile)
                spins = torch.tensor(S0, device="cuda")
                snaps = torch.zeros((nsnap,) + S0.shape, dtype=torch.float64, device="cuda")
                gctx.rk4_run(asd_amd.params(dt=0.01, **kw), L, spins, nsteps, every, nsnap, snaps)
                gctx.synchronize()
                assert np.array_equal(spins.cpu().numpy(), S_ref), (L, tile, kw)
                assert np.array_equal(snaps.cpu().numpy(), snaps_ref), (L, tile, kw)
    finally:
        gctx.rk4_set_tile(0)


# ---------------------------------------------------------------- tolerance mode
# asd_rk4_set_exact(ctx, 0): fused multiply-adds and a Newton-refined reciprocal square root
# instead of the reference's rounding steps.  Bar: within 1e-10 of the oracle over the first
# 100 steps and |S| = spin_length within 1e-12, on every kernel path (forced LDS tiles of
# each edge, the register-streaming strips, the automatic choice with fused two-step
# launches, the graph-replayed loop with snapshots) and on the batched kernel.
FAST_CASES = [
    (16, 3, 8, dict()),
    (33, 1, 16, dict(D=0.1, Bz=0.05)),
    (64, 2, 32, dict(D=0.1, Bz=0.1, alpha=0.01)),
    (48, 1, 0, dict(J=-0.7, D=0.05, Bz=0.2, alpha=0.2, spin_length=1.5)),
    (128, 1, 0, dict(D=0.1)),                         # automatic: fused 8 x 8 tiles
    (128, 4, 64, dict(D=0.1)),                        # streaming, plain (J = 1, alpha = 0, B_z = 0)
    (128, 2, 64, dict(J=0.9, D=0.05, Bz=0.1, alpha=0.05, spin_length=1.5)),
    (256, 1, 0, dict(D=0.05, Bz=0.2)),                # automatic: fused 16 x 16 / 1024 tiles
]


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep,tile,kw", FAST_CASES)
def test_tolerance_mode_within_bar(gctx, orc, L, nrep, tile, kw):
    torch = need_gpu()
    import asd_amd
    say(f"rk4 tolerance mode L={L} nrep={nrep} tile={tile} {kw}")
    length = kw.get("spin_length", 1.0)
    S0 = np.stack([orc.init_spins(31 + r, L * L) for r in range(nrep)]) * length
    nsteps, every, nsnap = 100, 10, 11
    S_ref, snaps_ref = orc.run(orc.params(dt=0.01, **kw), L, S0, nsteps, every, nsnap)
    spins = torch.tensor(S0, device="cuda")
    snaps = torch.zeros((nsnap,) + S0.shape, dtype=torch.float64, device="cuda")
    gctx.rk4_set_exact(False)
    gctx.rk4_set_tile(tile)
    try:
        gctx.rk4_run(asd_amd.params(dt=0.01, **kw), L, spins, nsteps, every, nsnap, snaps)
    finally:
        gctx.rk4_set_tile(0)
        gctx.rk4_set_exact(True)
    S = spins.cpu().numpy()
    err = float(np.abs(S - S_ref).max())
    err_snaps = float(np.abs(snaps.cpu().numpy() - snaps_ref).max())
    norm = float(np.abs(np.linalg.norm(S, axis=-1) - length).max())
    say(f"  max |S - S_oracle| = {err:.3e} (snapshots {err_snaps:.3e}), max ||S| - len| = {norm:.3e}")
    assert err <= 1e-10 and err_snaps <= 1e-10
    assert norm <= 1e-12
    assert not np.array_equal(S, S_ref) or L * L * nrep < 512   # it is a different rounding


@pytest.mark.gpu
def test_tolerance_mode_batch_within_bar(gctx, orc):
    torch = need_gpu()
    import asd_amd
    Ls = [9, 32, 48, 100]
    kws = [dict(D=0.1), dict(J=0.7, D=0.05, Bz=0.1, alpha=0.05), dict(D=0.2, Bz=0.02), dict(D=0.1)]
    S0 = [orc.init_spins(60 + t, L * L) for t, L in enumerate(Ls)]
    spins = [torch.tensor(s, device="cuda") for s in S0]
    gctx.rk4_set_exact(False)
    try:
        gctx.batch_rk4_run(Ls, [asd_amd.params(dt=0.01, **kw) for kw in kws], spins, 100)
    finally:
        gctx.rk4_set_exact(True)
    for t, (L, kw) in enumerate(zip(Ls, kws)):
        ref, _ = orc.run(orc.params(dt=0.01, **kw), L, S0[t], 100)
        S = spins[t].cpu().numpy()
        assert np.abs(S - ref).max() <= 1e-10, (t, L)
        assert np.abs(np.linalg.norm(S, axis=-1) - 1.0).max() <= 1e-12


# the automatic choice on many replicas takes the streaming strips; in the tolerance mode
# its launches fuse two steps (rk4_stream2_kernel), which must equal one-step streaming
# launches (asd_rk4_set_tile 64) bit for bit: every case below has at least four waves of
# one-step strips per CU on a 256-CU part.  Cases cover each arithmetic specialisation
# (J == 1, alpha == 0, the plain B_z == 0 / |S| == 1), ragged strips (L = 100: 48-column
# strips 3 x 48 = 144) and segments that end mid-column.
STREAM2_CASES = [
    (64, 256, dict(D=0.1)),                                            # plain
    (100, 320, dict(J=0.9, D=0.05, Bz=0.1, alpha=0.05, spin_length=1.5)),
    (64, 256, dict(D=0.1, alpha=0.05)),                                # J == 1, damped
    (64, 256, dict(J=0.8, D=0.1, Bz=0.02)),                            # alpha == 0
    (1024, 8, dict(D=0.1)),                                            # the large lattice
    (64, 1100, dict(D=0.1, Bz=0.05)),                                  # shares of several columns
]


@pytest.mark.gpu
@pytest.mark.parametrize("L,nrep,kw", STREAM2_CASES)
def test_two_step_streaming_equals_one_step_streaming(gctx, orc, L, nrep, kw):
    torch = need_gpu()
    import asd_amd
    say(f"two-step streaming L={L} nrep={nrep} {kw}")
    length = kw.get("spin_length", 1.0)
    S0 = np.stack([orc.init_spins(70 + r, L * L) for r in range(nrep)]) * length
    p = asd_amd.params(dt=0.01, **kw)
    n = 43 if L < 1024 else 31                     # odd: the last launch is one step
    res = {}
    # forced one-step strips (64) as the reference; the automatic choice (0) graph-replayed,
    # eager, and with an odd interval (one-step launches); the tolerance mode's LDS tiles (16)
    runs = [(64, 5, False), (64, 10, True), (0, 10, True), (0, 10, False), (0, 5, False), (16, 10, False)]
    gctx.rk4_set_exact(False)
    try:
        for tile, every, graph in runs:
            nsnap = n // every + 1
            s = torch.tensor(S0, device="cuda")
            snaps = torch.full((nsnap,) + S0.shape, np.nan, dtype=torch.float64, device="cuda")
            gctx.rk4_set_tile(tile)
            gctx.rk4_run(p, L, s, n, every, nsnap, snaps, use_graph=graph)
            gctx.synchronize()
            res[(tile, every, graph)] = (s.cpu().numpy(), snaps.cpu().numpy())
            del s, snaps
    finally:
        gctx.rk4_set_tile(0)
        gctx.rk4_set_exact(True)
    ref = res[(64, 5, False)]
    for key, (s, sn) in res.items():
        assert np.array_equal(s, ref[0]), key
        step = key[1] // 5
        assert np.array_equal(sn, ref[1][::step][: len(sn)]), key
    # the oracle on the first and last replica (100 steps: the tolerance bar)
    for r in ((0, nrep - 1) if L < 1024 else (0,)):
        m = 100 if L < 1024 else 20
        s = torch.tensor(S0[r], device="cuda")
        gctx.rk4_set_exact(False)
        try:
            gctx.rk4_run(p, L, s, m)
        finally:
            gctx.rk4_set_exact(True)
        S_or, _ = orc.run(orc.params(dt=0.01, **kw), L, S0[r], m)
        err = float(np.abs(s.cpu().numpy() - S_or).max())
        say(f"  replica {r}: {m} steps, max |S - S_oracle| = {err:.3e}")
        assert err <= 1e-10
    S = res[(0, 10, True)][0]
    assert np.abs(np.linalg.norm(S, axis=-1) - length).max() <= 1e-12
