"""S(q,tau) on the GPU (asd_sqt) against the FP64 oracle.

FP64 path: within 1e-10 (relative to max|S| per lag) of the oracle's GEMM form
(the reference data flow of llg_sqt_fusion.c:235-442 with time origins as
replicas) and of the separable Fourier form.  FP32 / BF16X3 / FP16X2: max and
RMS error against the GPU FP64 result are printed, with bounds stated in the
test.  The replica form (asd_sqt_replicas) is checked against the reference's
own fused_dgemm driven with the operands llg_sqt_fusion.c builds.
"""
import numpy as np
import pytest

from gpu_util import err_stats, need_gpu, say


def gxmg(L):
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


def full_grid(L):
    return [(a, b) for a in range(L) for b in range(L)]


def trajectory(orc, L, nsnap, every=5, seed=1, **kw):
    p = orc.params(dt=0.01, **kw)
    _, snaps = orc.run(p, L, orc.init_spins(seed, L * L), (nsnap - 1) * every, every, nsnap)
    return snaps


def max_rel(a, ref):
    return float(np.abs(a - ref).max() / np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags,qset", [
    (8, 10, 5, "full"), (12, 8, 6, "full"), (16, 20, 8, "gxmg"), (6, 30, 4, "full"), (24, 12, 5, "gxmg")])
def test_fp64_matches_oracle_gemm_form(gctx, orc, L, T0, nlags, qset):
    torch = need_gpu()
    say(f"sqt fp64 L={L} T0={T0} nlags={nlags} q={qset}")
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
    q = full_grid(L) if qset == "full" else gxmg(L)
    ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    got = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, nlags, q, "fp64")
    e = max_rel(got, ref)
    say(f"  max rel err vs oracle GEMM form {e:.3e}")
    assert e <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags", [(64, 50, 20), (128, 20, 10), (48, 30, 12), (32, 100, 100)])
def test_fp64_matches_separable(gctx, orc, L, T0, nlags):
    torch = need_gpu()
    say(f"sqt fp64 vs separable L={L} T0={T0} nlags={nlags}")
    snaps = trajectory(orc, L, T0 + nlags - 1, every=10)
    q = gxmg(L)
    ref = orc.sqt_separable(L, snaps, T0, nlags, q)
    got = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, nlags, q, "fp64")
    e = max_rel(got, ref)
    say(f"  max rel err vs separable {e:.3e}")
    assert e <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bound", [("fp32", 1e-5), ("bf16x3", 1e-5), ("fp16x2", 1e-5), ("fp16f8", 1e-5)])
def test_reduced_precision_error(gctx, orc, prec, bound):
    """Error of the MFMA variants against the GPU FP64 result on identical snapshots,
    per (q, tau) relative to max_q |S(q, tau)| (the floor)."""
    torch = need_gpu()
    L, T0, nlags = 64, 60, 20
    snaps = torch.tensor(trajectory(orc, L, T0 + nlags - 1, every=10, D=0.1), device="cuda")
    q = gxmg(L)
    ref = gctx.sqt(L, snaps, T0, nlags, q, "fp64")
    got = gctx.sqt(L, snaps, T0, nlags, q, prec)
    mx, rms = err_stats(got, ref)
    say(f"sqt {prec} L={L}: max rel err {mx:.3e}, rms {rms:.3e} (floor: max_q|S| per lag)")
    assert mx <= bound


@pytest.mark.gpu
def test_full_grid_sum_rule(gctx, orc):
    """sum over the full q grid of S(q,0) = Ns - sum_i |zbar_i|^2 (size-independent check)."""
    torch = need_gpu()
    L, T0 = 32, 40
    snaps = trajectory(orc, L, T0, every=10)
    s0 = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, 1, full_grid(L), "fp64")[0]
    zbar = snaps[:T0].mean(axis=0)
    assert abs(s0.sum() - (L * L - (zbar ** 2).sum())) <= 1e-9 * L * L
    assert np.abs(s0.imag).max() <= 1e-10


@pytest.mark.gpu
def test_replica_form_matches_reference_kernel(gctx, orc):
    """asd_sqt_replicas == reference fused_dgemm (or its restatement) fed the way
    llg_sqt_fusion.c:235-268, :370-426 feeds it: A = replicas now, B = replicas at 0."""
    torch = need_gpu()
    L, nrep = 8, 20
    ns, K = L * L, 3 * nrep
    zero = np.stack([orc.init_spins(40 + r, ns) for r in range(nrep)])
    now, _ = orc.run(orc.params(dt=0.01), L, zero, 30)
    mx = my = 1
    B = np.empty(K * ns)
    A = np.empty(ns * K)
    for r in range(nrep):
        for c in range(3):
            B[(r * 3 + c) + np.arange(ns) * K] = zero[r, :, c]
            A[np.arange(ns) + (r * 3 + c) * ns] = now[r, :, c]
    zbar, fbar = zero.mean(axis=0), now.mean(axis=0)
    Q = orc.build_qmat(L, mx, my)
    Cp = np.empty_like(Q)
    idx = np.arange(ns)
    for j in range(ns):
        base = (idx // 4) * 8 + idx % 4 + j * 2 * ns
        tmp = -(fbar @ zbar[j])
        Cp[base] = tmp * Q[base]
        Cp[base + 4] = tmp * Q[base + 4]
    s = np.zeros(4096)
    if orc.RefFusion.available():
        orc.RefFusion("parity").fused_dgemm(ns, ns, K, 1.0 / nrep, A, ns, B, K, Cp, 2 * ns, Q, s)
    else:
        orc.fused_dgemm(ns, ns, K, 1.0 / nrep, A, ns, B, K, Cp, 2 * ns, Q, s)
    expect = complex(s[0::2].sum(), s[1::2].sum()) / ns
    got = gctx.sqt_replicas(L, torch.tensor(now, device="cuda"), torch.tensor(zero, device="cuda"),
                            [(mx, my)], "fp64")[0]
    assert abs(got - expect) <= 1e-11 * max(1.0, abs(expect))


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags,prec", [(16, 12, 6, "fp64"), (20, 8, 5, "fp16x2")])
def test_full_grid_path_equals_per_q_path(gctx, orc, L, T0, nlags, prec):
    """nq > 2L takes the separable full-grid DFT path; the same momenta asked in
    chunks of at most 2L take the per-q path -- both must agree."""
    torch = need_gpu()
    snaps = torch.tensor(trajectory(orc, L, T0 + nlags - 1, D=0.1), device="cuda")
    q = full_grid(L)
    rng = np.random.default_rng(L)
    q = [q[i] for i in rng.permutation(len(q))]
    a = gctx.sqt(L, snaps, T0, nlags, q, prec)
    b = np.concatenate([gctx.sqt(L, snaps, T0, nlags, q[i:i + 2 * L], prec) for i in range(0, len(q), 2 * L)], axis=1)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.gpu
def test_full_grid_replica_form(gctx, orc):
    torch = need_gpu()
    L, nrep = 8, 6
    zero = np.stack([orc.init_spins(60 + r, L * L) for r in range(nrep)])
    now, _ = orc.run(orc.params(dt=0.01, D=0.1), L, zero, 20)
    q = full_grid(L)
    dn, dz = torch.tensor(now, device="cuda"), torch.tensor(zero, device="cuda")
    a = gctx.sqt_replicas(L, dn, dz, q, "fp64")
    b = np.concatenate([gctx.sqt_replicas(L, dn, dz, q[i:i + 2 * L], "fp64") for i in range(0, len(q), 2 * L)])
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp16x2", 1e-5), ("fp16f8", 2e-4)])
def test_large_work_list_all_lags_computed(gctx, orc, prec, tol):
    """L = 256 with 70 lags is 18.4M tile work items (4.7e9 work-items if launched one
    workgroup per tile, beyond the 32-bit dispatch size); every lag must match the
    separable oracle (FP16F8 at this K = 12 averages few fp8 roundings: the small-sample bound)."""
    torch = need_gpu()
    L, T0, nlags = 256, 4, 70
    snaps = np.stack([orc.init_spins(1000 + t, L * L) for t in range(T0 + nlags - 1)])
    q = [(0, 0), (1, 0), (3, 5), (128, 0), (128, 128), (17, 200), (255, 1), (64, 32)]
    ref = orc.sqt_separable(L, snaps, T0, nlags, q)
    got = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, nlags, q, prec)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    say(f"  {prec} L=256 per-lag max rel err: first {err[0]:.2e} last {err[-1]:.2e} max {err.max():.2e}")
    assert err.max() <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("L,prec,tol", [(600, "fp64", 1e-10), (1024, "fp64", 1e-10), (600, "fp32", 1e-5),
                                        (1024, "fp16x2", 1e-5), (600, "bf16x3", 2e-4)])
def test_large_lattices_match_separable(gctx, orc, L, prec, tol):
    """Lattices past L = 512, where the FP64 kernel's LDS no longer holds a tile's unwrapped
    bins (its wrapped bins take over), with L = 600 not a multiple of any tile edge (waves
    whose rows cross a lattice row) and L = 1024, the largest lattice; every lag against the
    separable oracle."""
    torch = need_gpu()
    T0, nlags = 4, 2
    say(f"sqt {prec} L={L} T0={T0} nlags={nlags}")
    snaps = np.stack([orc.init_spins(2000 + t, L * L) for t in range(T0 + nlags - 1)])
    q = [(0, 0), (1, 0), (3, 5), (L // 2, 0), (L // 2, L // 2), (17, L - 3), (L - 1, 1)]
    ref = orc.sqt_separable(L, snaps, T0, nlags, q)
    got = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, nlags, q, prec)
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    say(f"  per-lag max rel err {err}")
    assert err.max() <= tol


# Oracle bounds at these small samples (5 to 22 time origins), relative to
# max |S| over all (q, tau).  BF16X3 keeps about 17 significant bits per operand, so
# with few origins to average over its error reaches 3e-5 to 1.2e-4 (5 origins); a
# wrong fragment layout or a dropped tile gives errors of order 1.
SMALL_SAMPLE_TOL = {"fp64": 1e-10, "fp32": 1e-5, "bf16x3": 2e-4, "fp16x2": 1e-5, "fp16f8": 2e-4}


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3", "fp16x2", "fp16f8"])
def test_every_kernel_config_matches_oracle(gctx, orc, prec):
    """Every tile / wave layout the autotuner may pick gives the oracle's S(q,tau):
    lattices whose site count is not a multiple of any tile edge (400, 1296) and
    one smaller than a tile (144), so partial tiles and masked rows are covered.
    K = 3 T0 covers one whole 32-deep k step (18, 21, 27), a last step whose second
    16-deep half lies in the zero padding after zero, one and two whole steps (15, 48,
    66; the direct-to-LDS kernels do only its first half); the FP64 / FP32 ring of 8-deep
    stages sees 2, 4, 6 and 10 stages, the first fewer than the 3 it requests up front.
    Each output element is accumulated over k in the same order by every layout,
    so the layouts agree with each other up to the order of the FP64 bin sums."""
    torch = need_gpu()
    import asd_amd
    n = asd_amd.lib().asd_sqt_num_configs(asd_amd.PRECISIONS[prec])
    try:
        for L, T0, nlags in ((20, 9, 5), (36, 6, 4), (12, 7, 3), (20, 5, 4), (36, 16, 3), (12, 22, 3)):
            snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
            q = gxmg(L)
            ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
            d = torch.tensor(snaps, device="cuda")
            first = None
            for cfg in range(n):
                name = asd_amd.lib().asd_sqt_config_name(asd_amd.PRECISIONS[prec], cfg).decode()
                say(f"sqt {prec} L={L} config {cfg} ({name})")
                gctx.set_config(prec, cfg)
                got = gctx.sqt(L, d, T0, nlags, q, prec)
                e = max_rel(got, ref)
                first = got if first is None else first
                ec = max_rel(got, first)
                say(f"  max rel err vs oracle {e:.3e}, vs config 0 {ec:.3e}")
                assert e <= SMALL_SAMPLE_TOL[prec], (L, cfg, name, e)
                assert ec <= 1e-12, (L, cfg, name, ec)
    finally:
        gctx.set_config(prec, -1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_direct_to_lds_ring_equals_register_staged_at_full_depth(gctx, prec):
    """At the plain L = 128 workload's shape (K = 1200: 150 stages of the direct-to-LDS ring,
    16384 x 16384 per lag), the ring kernel (configuration 3, the default) and the
    register-staged kernel (configuration 0) give the same bits: the k order of every
    product is the same and the bins are integer sums."""
    torch = need_gpu()
    L, T0, nl = 128, 400, 3
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    snaps = torch.randn((T0 + nl - 1, L * L, 3), dtype=torch.float64, device="cuda", generator=g)
    snaps /= torch.linalg.norm(snaps, dim=-1, keepdim=True)
    q = gxmg(L)
    try:
        out = []
        for cfg in (0, 3):
            gctx.set_config(prec, cfg)
            out.append(gctx.sqt(L, snaps, T0, nl, q, prec))
            say(f"sqt {prec} L={L} K={3 * T0} config {cfg}: sum |S| {np.abs(out[-1]).sum():.15e}")
    finally:
        gctx.set_config(prec, -1)
    assert np.all(np.isfinite(out[0])) and np.abs(out[0]).max() > 0
    assert np.array_equal(out[0], out[1])


def e4m3(x):
    """Nearest OCP FP8 e4m3 value (bias 7, 3 mantissa bits; below 2^-6 the multiples of 2^-9),
    ties to even, for |x| <= 448: the rounding FP16F8 applies to its cross-product parts."""
    x = np.asarray(x, dtype=np.float64)
    ax = np.abs(x)
    e = np.maximum(np.floor(np.log2(np.where(ax > 0, ax, 1.0))), -6.0)
    step = 2.0 ** (e - 3)
    return np.round(x / step) * step


def test_e4m3_rounding_known_values():
    """The e4m3 reference rounding on values whose codes are fixed by the format: the largest
    finite 448, the smallest normal 2^-6 and subnormal 2^-9, ties to even at both ends of a
    binade, a carry into the next binade, and a subnormal tie."""
    cases = {448.0: 448.0, 1.0: 1.0, 1.0625: 1.0, 1.1875: 1.25, 1.9375: 2.0, 2.0 ** -6: 2.0 ** -6,
             2.0 ** -9: 2.0 ** -9, 1.5 * 2.0 ** -9: 2.0 ** -8, 2.5 * 2.0 ** -9: 2.0 ** -8,
             0.3: 0.3125, -0.3: -0.3125, 0.0: 0.0, 2.0 ** -11: 0.0}
    for x, want in cases.items():
        assert e4m3(x) == want, (x, e4m3(x), want)


def pack_scale(x):
    """The split precisions' pack scale: the power of two 2^(7 - e) with max |x| < 2^e <= 2 max |x|,
    which puts the largest component in [64, 128)."""
    return 2.0 ** (7 - np.frexp(np.abs(x).max())[1])


def fp16f8_parts(x, ps):
    """FP16F8's three operand arrays for the components x, in x's units: hi = f16(v), its e4m3
    rounding, and e4m3((v - hi) 2^11) 2^-11, with v = f32(x ps) as the pack forms it."""
    v = (x * ps).astype(np.float32)
    hi = v.astype(np.float16).astype(np.float32)
    return hi / ps, e4m3(hi) / ps, e4m3((v - hi) * np.float32(2048.0)) / (2048.0 * ps)


def sqt_connected(L, X, Y, T0, nlags, q):
    """(1 / (Ns T0)) sum_t0 sum_c F_X(t0 + tau) conj F_Y(t0): the connected part of S(q,tau)
    for the operand arrays X (A side) and Y (B side), with the oracle's phase convention
    e^{-2 pi i (qx (i / L) + qy (i % L)) / L} (orc_sqt_separable, separable form)."""
    ns = L * L
    r, c = np.divmod(np.arange(ns), L)
    qa = np.asarray(q)
    ph = np.exp(-2j * np.pi * (np.outer(qa[:, 0], r) + np.outer(qa[:, 1], c)) / L)  # [nq][Ns]
    FX = np.einsum("qi,tic->tqc", ph, X)
    FY = np.einsum("qi,tic->tqc", ph, Y)
    out = np.empty((nlags, len(q)), dtype=complex)
    for tau in range(nlags):
        out[tau] = np.einsum("tqc,tqc->q", FX[tau:tau + T0], FY[:T0].conj()) / (ns * T0)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags", [(20, 30, 4), (36, 25, 3)])
def test_fp16f8_matches_its_arithmetic_on_the_cpu(gctx, orc, L, T0, nlags):
    """FP16F8 against a CPU restatement of its own arithmetic: v = f32(x ps) with the pack scale,
    hi = f16(v), the cross parts e4m3(hi) and e4m3((v - hi) 2^11), products
    hi.hi + 2^-11 (hi8.lo8 + lo8.hi8), divided by ps^2.  The
    connected part is linear in the operand products, so the restatement is the oracle's FP64
    S(q,tau) with the connected part of the exact snapshots replaced by those of the three
    products (separable form).  The GPU sums in FP32 where this sums in FP64, so the two agree
    to FP32 accumulation (bound 1e-6 relative to max_q |S| per lag), several times closer than
    FP16F8 is to FP64; a wrong byte pairing or rounding would give errors of the fp8 size."""
    torch = need_gpu()
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
    q = gxmg(L)
    H, H8, R8 = fp16f8_parts(snaps, pack_scale(snaps))
    ref64 = orc.sqt_gemm(L, snaps, T0, nlags, q)
    emul = (ref64 - sqt_connected(L, snaps, snaps, T0, nlags, q) + sqt_connected(L, H, H, T0, nlags, q)
            + sqt_connected(L, H8, R8, T0, nlags, q) + sqt_connected(L, R8, H8, T0, nlags, q))
    got = gctx.sqt(L, torch.tensor(snaps, device="cuda"), T0, nlags, q, "fp16f8")
    e_emul, e_64 = max_rel(got, emul), max_rel(got, ref64)
    say(f"sqt fp16f8 L={L} K={3 * T0}: max rel err vs its CPU restatement {e_emul:.3e}, vs FP64 {e_64:.3e}")
    assert e_emul <= 1e-6 and e_emul <= 0.25 * e_64


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp16x2", "fp16f8", "bf16x3"])
def test_split_precisions_scale_with_the_spin_length(gctx, orc, prec):
    """The split precisions pack the snapshots scaled by a power of two that puts their largest
    component in [64, 128), and convert the products back exactly, so spins of any length keep
    the same relative error: by a power of two the result scales bit for bit (2^12 is beyond fp8's
    range of 448 and 2^-16 puts every component among fp16's subnormals, unscaled), and by 1000
    it stays within the precision's bound of the FP64 result."""
    torch = need_gpu()
    L, T0, nlags = 32, 100, 6
    snaps = torch.tensor(trajectory(orc, L, T0 + nlags - 1, D=0.1), device="cuda")
    q = gxmg(L)
    base = gctx.sqt(L, snaps, T0, nlags, q, prec)
    for sc in (2.0 ** 12, 2.0 ** -16):
        got = gctx.sqt(L, snaps * sc, T0, nlags, q, prec)
        assert np.array_equal(got, base * sc * sc), (prec, sc)
    big = snaps * 1000.0
    e = max_rel(gctx.sqt(L, big, T0, nlags, q, prec), gctx.sqt(L, big, T0, nlags, q, "fp64"))
    e1 = max_rel(base, gctx.sqt(L, snaps, T0, nlags, q, "fp64"))
    say(f"sqt {prec} L={L} K={3 * T0}: spins x1000 max rel err vs FP64 {e:.3e} (x1: {e1:.3e})")
    assert e <= (1.5e-4 if prec == "bf16x3" else 1e-5)


@pytest.mark.gpu
def test_fp16f8_replica_form_matches_its_arithmetic(gctx, orc):
    """The replica form (A = replicas now, B = replicas at 0: two different arrays) in FP16F8
    against the CPU restatement of its roundings, built on the GPU's FP64 replica result
    (which test_replica_form_matches_reference_kernel pins to the reference kernel)."""
    torch = need_gpu()
    L, nrep = 16, 40
    zero = np.stack([orc.init_spins(70 + r, L * L) for r in range(nrep)])
    now, _ = orc.run(orc.params(dt=0.01, D=0.1), L, zero, 25)
    q = gxmg(L)
    dn, dz = torch.tensor(now, device="cuda"), torch.tensor(zero, device="cuda")
    s64 = gctx.sqt_replicas(L, dn, dz, q, "fp64")
    got = gctx.sqt_replicas(L, dn, dz, q, "fp16f8")

    ps = pack_scale(np.concatenate([now.ravel(), zero.ravel()]))
    (Hn, H8n, R8n), (Hz, H8z, R8z) = fp16f8_parts(now, ps), fp16f8_parts(zero, ps)
    conn = lambda X, Y: sqt_connected(L, X, Y, nrep, 1, q)[0]
    emul = s64 - conn(now, zero) + conn(Hn, Hz) + conn(H8n, R8z) + conn(R8n, H8z)
    scale = np.abs(s64).max()
    e_emul, e_64 = np.abs(got - emul).max() / scale, np.abs(got - s64).max() / scale
    say(f"sqt_replicas fp16f8 L={L} K={3 * nrep}: max rel err vs its CPU restatement {e_emul:.3e}, "
        f"vs FP64 {e_64:.3e}")
    assert e_emul <= 1e-6 and e_emul <= 0.25 * e_64


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags,qset", [(16, 20, 8, "full"), (20, 12, 6, "gxmg"), (32, 10, 4, "gxmg")])
def test_unfused_path_matches_fused_and_oracle(gctx, orc, L, T0, nlags, qset):
    """asd_sqt_unfused (rocBLAS GEMM into an explicit Ns x Ns matrix, outer subtract,
    phase reduction as a second GEMM) gives the fused FP64 path's and the oracle's
    S(q,tau) to 1e-10."""
    torch = need_gpu()
    say(f"sqt unfused L={L} T0={T0} nlags={nlags} q={qset}")
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
    q = full_grid(L) if qset == "full" else gxmg(L)
    d = torch.tensor(snaps, device="cuda")
    ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    un = gctx.sqt_unfused(L, d, T0, nlags, q)
    fu = gctx.sqt(L, d, T0, nlags, q, "fp64")
    e_o, e_f = max_rel(un, ref), max_rel(un, fu)
    say(f"  max rel err vs oracle {e_o:.3e}, vs fused {e_f:.3e}")
    assert e_o <= 1e-10 and e_f <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags", [(6, 9, 4), (8, 10, 5), (16, 6, 3)])
def test_bruteforce_matches_oracle_and_fused(gctx, orc, L, T0, nlags):
    """The direct double sum with sincos phases (no lattice tables, no binning) agrees
    with the oracle and with the fused path on the full momentum grid."""
    torch = need_gpu()
    say(f"sqt brute force L={L} T0={T0} nlags={nlags}")
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05, alpha=0.02)
    q = full_grid(L)
    d = torch.tensor(snaps, device="cuda")
    ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    bf = gctx.sqt_bruteforce(L, d, T0, nlags, q)
    fu = gctx.sqt(L, d, T0, nlags, q, "fp64")
    e_o, e_f = max_rel(bf, ref), max_rel(bf, fu)
    say(f"  max rel err vs oracle {e_o:.3e}, vs fused {e_f:.3e}")
    assert e_o <= 1e-10 and e_f <= 1e-10


@pytest.mark.gpu
def test_unfused_path_rejects_sizes_it_cannot_hold(gctx):
    """The unfused path stores Ns^2 doubles and a 2 nq Ns phase table; beyond its bounds it
    returns an error before allocating anything."""
    torch = need_gpu()
    snaps = torch.zeros((5, 256 * 256, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="phase table"):
        gctx.sqt_unfused(256, snaps, 3, 2, full_grid(256))            # 2 nq Ns = 8.6e9 doubles
    big = torch.zeros((3, 1024 * 1024, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="Ns\\^2"):
        gctx.sqt_unfused(1024, big, 2, 1, [(0, 0)])                   # Ns^2 = 1.1e12 doubles


@pytest.mark.gpu
@pytest.mark.parametrize("L,T0,nlags", [(9, 12, 4), (33, 10, 3)])
def test_odd_lattice_full_grid(gctx, orc, L, T0, nlags):
    """Odd L (the displacement bins wrap by L, the twiddle table has L entries; L = 33 gives
    ragged tiles on both sides): FP64 against the oracle's GEMM form, FP16X2 within 1e-5."""
    torch = need_gpu()
    say(f"sqt odd L={L} T0={T0} nlags={nlags} full grid")
    snaps = trajectory(orc, L, T0 + nlags - 1, D=0.1, Bz=0.05)
    q = full_grid(L)
    ref = orc.sqt_gemm(L, snaps, T0, nlags, q)
    d = torch.tensor(snaps, device="cuda")
    e64 = max_rel(gctx.sqt(L, d, T0, nlags, q, "fp64"), ref)
    mx, rms = err_stats(gctx.sqt(L, d, T0, nlags, q, "fp16x2"), ref)
    say(f"  fp64 {e64:.3e}; fp16x2 max {mx:.3e} rms {rms:.3e}")
    assert e64 <= 1e-10
    assert mx <= 1e-5


@pytest.mark.gpu
def test_single_time_origin_gives_zero(gctx, orc):
    """With one time origin the connected term cancels exactly: the window average of
    F_q(tau) conj F_q(0) is the product of the window means, so S(q, tau) = 0 for every q and
    lag (FP64 to rounding; FP16X2 to its product error)."""
    torch = need_gpu()
    L, nlags = 16, 5
    snaps = torch.tensor(trajectory(orc, L, nlags, D=0.1), device="cuda")
    q = full_grid(L)
    s64 = gctx.sqt(L, snaps, 1, nlags, q, "fp64")
    s16 = gctx.sqt(L, snaps, 1, nlags, q, "fp16x2")
    say(f"  T0 = 1: max |S| fp64 {np.abs(s64).max():.2e}, fp16x2 {np.abs(s16).max():.2e}")
    assert np.abs(s64).max() <= 1e-13
    assert np.abs(s16).max() <= 1e-5


@pytest.mark.gpu
def test_invalid_sizes_are_rejected(gctx, orc):
    torch = need_gpu()
    import asd_amd
    L = 16
    snaps = torch.tensor(trajectory(orc, L, 6), device="cuda")
    q = full_grid(L)
    for T0, nlags in ((0, 3), (3, 0), (4, 4)):          # the last needs 7 snapshots, 6 given
        with pytest.raises(asd_amd.AsdError):
            gctx.sqt(L, snaps, T0, nlags, q, "fp64")
    with pytest.raises(asd_amd.AsdError):
        gctx.sqt(L, snaps, 3, 2, [(0, L)], "fp64")       # q index out of [0, L)
    with pytest.raises(ValueError):                      # snapshots of another lattice size
        gctx.sqt(L, snaps[:, : L * L // 2].contiguous(), 3, 2, q, "fp64")
    for prec in (-1, 5):                                 # no such precision
        with pytest.raises(asd_amd.AsdError):
            gctx.sqt(L, snaps, 3, 2, q, prec)
        with pytest.raises(asd_amd.AsdError):
            gctx.sqt_replicas(L, snaps[0], snaps[1], q, prec)
