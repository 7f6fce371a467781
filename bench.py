#!/usr/bin/env python3
"""Benchmark of the RK4-LLG + S(q,tau) hot path on one MI355X.

Workload (the paper-scale plain configuration): L = 128 (16384 sites), J = 1,
D = 0.1, B_z = 0, alpha = 0, spins Marsaglia-uniform with seed 1; 8000 RK4
steps at dt = 0.01 with a snapshot every 10 steps (800 snapshots); S(q,tau) for
lags 0..399 averaged over 400 time origins (GEMM K = 1200 per lag) on the
Gamma-X-M-Gamma path (192 lattice momenta).

One "step" of this benchmark is one pass of the whole hot path over that input:
the 8000-step RK4 run (snapshot writes included) followed by S(q,tau) in FP64, the
reference's own arithmetic (its fused_dgemm is an FP64 DGEMM).  `value` is the
S(q,tau) throughput in algorithmic GEMM flops (sum over lags of 2 Ns^2 K = 2.577e14
per pass) per second of S(q,tau) time, and `roofline` is the FP64 correlation GEMM
against the FP64 MFMA peak.  Beside it: RK4 spin-steps/s (and on an HBM-bound batch),
and the mixed-precision leg -- the same pass with the GEMM in split FP16F8 on MFMA
(fp16 hi products, fp8 cross products), its own roofline (3 MFMA products per useful
flop, costing 2 in fp16 MFMA work), its max / RMS error against the FP64 result on the
same snapshots and the end-to-end time ratio FP64 / mixed -- with the FP16X2 split
(3 fp16 products) reported the same way beside it.  Inputs are resident on the GPU
before timing.

The CPU baseline is the reference's own kernel_fusion code (runge_kutta and
fused_dgemm from oracle/_ref, built from its sources with its AVX2/FMA flags)
timed on this host's cores on a bounded sample of the same workload, with the
reference's thread rule min(128, omp_get_max_threads()).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "asd-kernel-fusion_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "RK4 spin-steps/s; S(q,t) TFLOP/s at L=128; mixed vs FP64 time @ err≤1e-5"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "fp16x2": 2500.0, "fp16f8": 2500.0}
# MFMA products per useful flop, and their cost in dense fp16 MFMA work (fp8 runs at twice the rate)
PRODUCTS_PER_FLOP = {"fp64": 1, "fp32": 1, "bf16x3": 3, "fp16x2": 3, "fp16f8": {"fp16": 1, "fp8": 2}}
FP16_WORK_PER_FLOP = {"fp64": 1, "fp32": 1, "bf16x3": 3, "fp16x2": 3, "fp16f8": 2}
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu_facts():
    """Host CPU model and counts (the CPU baseline's context)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def reference_threads():
    """The reference's rule, min(128, omp_get_max_threads()) (llg_sqt_fusion.c:221-224);
    omp_get_max_threads() is OMP_NUM_THREADS when set, else the CPUs this process may use."""
    env = os.environ.get("OMP_NUM_THREADS", "").split(",")[0].strip()
    if env.isdigit() and int(env) > 0:
        n = int(env)
    else:
        try:
            n = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            n = os.cpu_count() or 1
    return min(128, n)


def cpu_baseline(L, T0, nq, nlags, nsteps_total, threads):
    """Reference kernel_fusion (oracle/_ref, AVX2/FMA -O3) on `threads` host cores,
    bounded sample: 200 RK4 steps at L, and one fused_dgemm call at L, K = 3 T0
    (one q, as the reference evaluates one q per call) with the prefill glue."""
    import numpy as np

    import oracle
    if not oracle.RefFusion.available("perf"):
        return None
    os.environ["OMP_NUM_THREADS"] = str(threads)
    ref = oracle.RefFusion("perf")
    ns = L * L
    p = oracle.params(J=1.0, D=0.0, dt=0.01)
    ref.set_globals(L, 1, 0.01, 1.0)
    S = oracle.init_spins(1, ns).reshape(1, ns, 3)
    n_rk = 200
    ref.runge_kutta(S)                                  # warm
    t0 = time.perf_counter()
    for _ in range(n_rk):
        cur = ref.runge_kutta(S)[0]
        S = oracle.renormalise(cur, 1.0)
    t_rk = time.perf_counter() - t0
    rk_rate = n_rk * ns / t_rk

    K = 3 * T0
    rng = np.random.default_rng(0)
    A = rng.standard_normal(ns * K)
    B = rng.standard_normal(K * ns)
    t0 = time.perf_counter()
    Q = oracle.build_qmat(L, L // 10, L // 10)          # once per run in the reference (:308-333)
    t_q = time.perf_counter() - t0
    C = np.empty_like(Q)
    sums = np.zeros(4096)
    t0 = time.perf_counter()
    np.multiply(Q, -0.5, out=C)                        # disconnected prefill C = -(f.z) Q (:402-426)
    ref.fused_dgemm(ns, ns, K, 1.0 / T0, A, ns, B, K, C, 2 * ns, Q, sums)
    t_call = time.perf_counter() - t0
    gflop = 2.0 * ns * ns * K / 1e9
    del Q, C
    return {
        "rk4_spin_steps_per_s": rk_rate,
        "sqt_tflops": gflop / 1e3 / t_call,
        "sqt_s_per_q_lag": t_call,
        "projected_sqt_s": t_call * nq * nlags,
        "projected_rk4_s": nsteps_total * ns / rk_rate,
        "qmat_build_s": t_q,
    }


def gxmg(L):
    h = L // 2
    return [(k, 0) for k in range(h)] + [(h, k) for k in range(h)] + [(h - k, h - k) for k in range(h)]


def err_stats(a, ref):
    import numpy as np
    scale = np.abs(ref).max(axis=1, keepdims=True)
    e = np.abs(a - ref) / scale
    return {"max": float(e.max()), "rms": float(np.sqrt((e ** 2).mean()))}


def workload_precision256(args):
    """L=256, J=1 D=0.05 B_z=0.2 alpha=0 seed 2, 4000 FP64 RK4 steps (snapshot every 10),
    lags 0..199 over 200 origins, full L x L momentum grid; S(q,t) in FP64, FP32, BF16X3,
    FP16X2 with max / RMS error per (q, tau) against FP64 on identical snapshots."""
    import numpy as np
    import torch

    import asd_amd
    L, T0, nl, every, nsteps = 256, 200, 200, 10, 4000
    ns = L * L
    ctx = asd_amd.Context(0)
    ctx.set_timing(True)
    p = asd_amd.params(J=1.0, D=0.05, Bz=0.2, alpha=0.0, dt=0.01)
    S = torch.tensor(asd_amd.init_spins(2, ns), device="cuda")
    nsnap = nsteps // every
    snaps = torch.empty((nsnap, ns, 3), dtype=torch.float64, device="cuda")
    ctx.rk4_run(p, L, S, nsteps, every, nsnap, snaps)
    rk = ctx.stage_times()["rk4"]
    q = asd_amd.full_grid(L)
    flop = nl * 2.0 * ns * ns * 3 * T0
    res, outs = {}, {}
    for prec in ("fp64", "fp32", "bf16x3", "fp16x2", "fp16f8"):
        ctx.autotune(L, T0, prec)
        out = torch.empty((nl, len(q), 2), dtype=torch.float64, device="cuda")
        ctx.sqt(L, snaps, T0, nl, q, prec, out=out)
        torch.cuda.synchronize()
        st = ctx.stage_times()
        total = st["pack"] + st["gemm"] + st["fourier"] + st["finish"]
        o = out.cpu().numpy()
        outs[prec] = o[..., 0] + 1j * o[..., 1]
        res[prec] = {"sqt_s": round(total / 1e3, 3), "gemm_s": round(st["gemm"] / 1e3, 3),
                     "gemm_tflops": round(flop / st["gemm"] / 1e9, 2)}
        log(f"  {prec}: {res[prec]}")
    for prec in ("fp32", "bf16x3", "fp16x2", "fp16f8"):
        res[prec]["err_vs_fp64"] = err_stats(outs[prec], outs["fp64"])
    zbar = snaps[:T0].mean(dim=0).cpu().numpy()
    sum_rule = float(abs(outs["fp64"][0].sum() - (ns - (zbar ** 2).sum())) / ns)
    print(json.dumps({"workload": "precision L=256 full grid", "L": L, "T0": T0, "lags": nl, "nq": len(q),
                      "rk4_s": round(rk / 1e3, 4), "gemm_flop": flop, "per_precision": res,
                      "error_floor": "relative to max_q |S(q,tau)| of each lag",
                      "fp64_sum_rule_rel_residual": sum_rule}), flush=True)


def workload_batch96(args):
    """96 trajectories, L in {32,48,64,96,128} (32/24/20/12/8), D ~ U[0,0.2], alpha=0,
    seeds 100..195, 2000 FP64 RK4 steps (snapshot every 10) in one persistent launch per
    step; S(q,t) for lags 0..99 over 100 origins on each lattice's Gamma-X-M-Gamma path
    in FP16X2 with tile shapes autotuned per L."""
    import numpy as np
    import torch

    import asd_amd
    Ls = [32] * 32 + [48] * 24 + [64] * 20 + [96] * 12 + [128] * 8
    rng = np.random.default_rng(2024)
    Ds = rng.uniform(0.0, 0.2, len(Ls))
    nsteps, every, T0, nl = 2000, 10, 100, 100
    nsnap = nsteps // every
    ctx = asd_amd.Context(0)
    ctx.set_timing(True)
    plist = [asd_amd.params(J=1.0, D=float(d), dt=0.01) for d in Ds]
    S0 = [torch.tensor(asd_amd.init_spins(100 + t, L * L), device="cuda") for t, L in enumerate(Ls)]
    spins = [s.clone() for s in S0]
    snaps = [torch.empty((nsnap, L * L, 3), dtype=torch.float64, device="cuda") for L in Ls]
    total_sites = sum(L * L for L in Ls)
    vbest, vtimes = ctx.autotune_batch(Ls, plist)
    log(f"  batch variants (ms per 20 steps): {vtimes}")
    ctx.batch_rk4_run(Ls, plist, spins, nsteps, every, nsnap, snaps)     # warm (graph capture)
    for t in range(len(Ls)):
        spins[t].copy_(S0[t])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.batch_rk4_run(Ls, plist, spins, nsteps, every, nsnap, snaps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"  batched RK4 {t_batch:.3f} s")
    # the tolerance mode (fused multiply-adds; its own autotuned variant)
    ctx.rk4_set_exact(False)
    vbest_f, vtimes_f = ctx.autotune_batch(Ls, plist)
    fast = [s.clone() for s in S0]
    ctx.batch_rk4_run(Ls, plist, fast, nsteps, every, nsnap, snaps)      # warm (graph capture)
    for t in range(len(Ls)):
        fast[t].copy_(S0[t])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.batch_rk4_run(Ls, plist, fast, nsteps, every, nsnap, snaps)
    torch.cuda.synchronize()
    t_fast = time.perf_counter() - t0
    # the two modes over the first 100 steps (the trajectory bar; later the dynamics'
    # sensitivity to initial conditions amplifies ulp-level differences)
    fast = [s.clone() for s in S0]
    ctx.batch_rk4_run(Ls, plist, fast, 100)
    ctx.rk4_set_exact(True)
    exact = [s.clone() for s in S0]
    ctx.batch_rk4_run(Ls, plist, exact, 100)
    fast_diff = max(float((a - b).abs().max()) for a, b in zip(fast, exact))
    log(f"  batched RK4, tolerance mode {t_fast:.3f} s (max |dS| vs exact after 100 steps {fast_diff:.1e})")
    # unbatched: one trajectory at a time (for the batching gain)
    t0 = time.perf_counter()
    for t, L in enumerate(Ls):
        s = S0[t].clone()
        ctx.rk4_run(plist[t], L, s, nsteps, every, nsnap, snaps[t])
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    flop = sum(nl * 2.0 * (L * L) ** 2 * 3 * T0 for L in Ls)
    tuned, t_sqt = {}, {}
    for prec in ("fp16x2", "fp16f8"):
        tuned[prec] = {}
        for L in sorted(set(Ls)):
            best, times = ctx.autotune(L, T0, prec)
            tuned[prec][L] = list(times)[best]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t, L in enumerate(Ls):
            ctx.sqt(L, snaps[t], T0, nl, gxmg(L), prec)
        torch.cuda.synchronize()
        t_sqt[prec] = time.perf_counter() - t0
    print(json.dumps({"workload": "batch 96 trajectories", "trajectories": len(Ls), "sites_total": total_sites,
                      "rk4_steps": nsteps, "rk4_batched_s": round(t_batch, 4),
                      "rk4_spin_steps_per_s": total_sites * nsteps / t_batch,
                      "rk4_unbatched_s": round(t_single, 4), "batching_speedup": round(t_single / t_batch, 2),
                      "rk4_hbm_frac": round(48.0 * total_sites * nsteps / t_batch / (HBM_PEAK_GBS * 1e9), 4),
                      "batch_variant": list(vtimes)[vbest],
                      "batch_variant_ms_per_20_steps": {k: round(v, 4) for k, v in vtimes.items()},
                      "tolerance_mode": {"rk4_batched_s": round(t_fast, 4),
                                         "rk4_spin_steps_per_s": total_sites * nsteps / t_fast,
                                         "batch_variant": list(vtimes_f)[vbest_f],
                                         "max_abs_diff_vs_exact_100_steps": fast_diff},
                      "sqt_fp16x2_s": round(t_sqt["fp16x2"], 3),
                      "sqt_tflops": round(flop / t_sqt["fp16x2"] / 1e12, 2),
                      "sqt_fp16f8_s": round(t_sqt["fp16f8"], 3),
                      "sqt_fp16f8_tflops": round(flop / t_sqt["fp16f8"] / 1e12, 2),
                      "autotuned_cfg_per_L": tuned}), flush=True)


def workload_adjoint64(args):
    """L=64, theta=(J,D,B_z,alpha)=(1.0,0.1,0.1,0.01), target from theta*=(1.1,0.12,0.08,0.01),
    seed 3, 4000 RK4 steps, snapshot every 10, lags 0..199 over 200 origins, Gamma-X-M-Gamma;
    loss sum |S - S_target|^2; hand-written adjoint (checkpoint every 50 steps) against
    central FD (h=1e-6) and the FP64 CPU tangent-linear model."""
    import numpy as np
    import torch

    import asd_amd
    L, nsteps, every, T0, nl = 64, 4000, 10, 200, 200
    q = gxmg(L)
    ctx = asd_amd.Context(0)
    S0 = torch.tensor(asd_amd.init_spins(3, L * L), device="cuda")
    ts, th = (1.1, 0.12, 0.08, 0.01), [1.0, 0.1, 0.1, 0.01]
    target, _ = ctx.sqt_vjp(asd_amd.params(*ts, 0.01), L, S0, nsteps, every, T0, nl, q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, grad = ctx.sqt_loss_grad(asd_amd.params(*th, 0.01), L, S0, nsteps, every, T0, nl, q, target, 50)
    t_adj = time.perf_counter() - t0
    log(f"  adjoint: loss {loss:.6e} grad {grad} ({t_adj:.2f} s)")
    h = 1e-6
    fd = []
    t0 = time.perf_counter()
    for i in range(4):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        sp, _ = ctx.sqt_vjp(asd_amd.params(*tp, 0.01), L, S0, nsteps, every, T0, nl, q)
        sm, _ = ctx.sqt_vjp(asd_amd.params(*tm, 0.01), L, S0, nsteps, every, T0, nl, q)
        fd.append((np.sum(np.abs(sp - target) ** 2) - np.sum(np.abs(sm - target) ** 2)) / (2 * h))
    t_fd = time.perf_counter() - t0
    fd = np.array(fd)
    log(f"  central FD: {fd} ({t_fd:.2f} s)")
    # short horizon (400 steps, t = 4): the trajectory has not yet decorrelated, so
    # central differences are meaningful there
    Ts, nls, ns_short = 20, 20, 390
    tgt_s, _ = ctx.sqt_vjp(asd_amd.params(*ts, 0.01), L, S0, ns_short, every, Ts, nls, q)
    _, g_s = ctx.sqt_loss_grad(asd_amd.params(*th, 0.01), L, S0, ns_short, every, Ts, nls, q, tgt_s, 50)
    fd_s = []
    for i in range(4):
        tp, tm = list(th), list(th)
        tp[i] += h
        tm[i] -= h
        sp, _ = ctx.sqt_vjp(asd_amd.params(*tp, 0.01), L, S0, ns_short, every, Ts, nls, q)
        sm, _ = ctx.sqt_vjp(asd_amd.params(*tm, 0.01), L, S0, ns_short, every, Ts, nls, q)
        fd_s.append((np.sum(np.abs(sp - tgt_s) ** 2) - np.sum(np.abs(sm - tgt_s) ** 2)) / (2 * h))
    fd_s = np.array(fd_s)
    out = {"workload": "adjoint L=64", "loss": loss, "grad_adjoint": grad.tolist(), "grad_central_fd": fd.tolist(),
           "max_rel_diff_fd": float(np.abs(grad - fd).max() / np.abs(fd).max()),
           "adjoint_s": round(t_adj, 3), "fd_8_forward_s": round(t_fd, 3),
           "note": "after 4000 steps the loss is chaotic in theta at the 1e-6 scale (|grad| ~ 1e12), so "
                   "central FD at h = 1e-6 is not a derivative there; the tangent-linear check is the tight one",
           "short_horizon": {"steps": ns_short, "T0": Ts, "lags": nls, "grad_adjoint": g_s.tolist(),
                             "grad_central_fd": fd_s.tolist(),
                             "max_rel_diff_fd": float(np.abs(g_s - fd_s).max() / np.abs(fd_s).max())}}
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        s_now, _ = ctx.sqt_vjp(asd_amd.params(*th, 0.01), L, S0, nsteps, every, T0, nl, q)
        snaps, dsnaps = oracle.tlm_run(oracle.params(*th, dt=0.01), L, S0.cpu().numpy(), nsteps, every, T0 + nl - 1)
        R = s_now - target
        tlm = np.array([2 * np.sum((R.conj() * oracle.tlm_sqt(L, snaps, dsnaps[i], T0, nl, q)).real)
                        for i in range(4)])
        out.update(grad_tangent_linear_cpu=tlm.tolist(),
                   max_rel_diff_tlm=float(np.abs(grad - tlm).max() / np.abs(tlm).max()),
                   tlm_cpu_s=round(time.perf_counter() - t0, 2))
    print(json.dumps(out), flush=True)


def workload_rk4sweep(args):
    """RK4 throughput against the HBM roofline (48 B per spin-step algorithmic) from the
    latency-bound single L=128 lattice to replica batches whose state exceeds the caches."""
    import torch

    import asd_amd
    ctx = asd_amd.Context(0)
    ctx.set_timing(True)
    ctx.rk4_set_exact(not args.rk4_fast)
    p = asd_amd.params(J=1.0, D=0.1, dt=0.01)
    rows = []
    for L, nrep, steps in ((128, 1, 2000), (256, 1, 1000), (1024, 1, 200), (1024, 8, 100), (1024, 32, 50),
                           (1024, 64, 30)):
        n = L * L * nrep
        S = torch.tensor(asd_amd.init_spins(5, L * L), device="cuda").repeat(nrep, 1)
        row = {"L": L, "replicas": nrep, "sites": n, "state_MB": round(n * 24 / 1e6, 1)}
        for tile in (0, 8, 16, 32, 64):   # 0: the automatic choice
            ctx.rk4_set_tile(tile)
            ctx.rk4_run(p, L, S, steps)   # untimed: the clocks settle
            torch.cuda.synchronize()
            ctx.rk4_run(p, L, S, steps)
            ms = ctx.stage_times()["rk4"]
            rate = n * steps / (ms * 1e-3)
            row[f"tile{tile}"] = {"spin_steps_per_s": rate, "us_per_step": round(ms * 1e3 / steps, 2),
                                  "algorithmic_GBps": round(48 * rate / 1e9, 1),
                                  "hbm_frac": round(48 * rate / (HBM_PEAK_GBS * 1e9), 4)}
        ctx.rk4_set_tile(0)
        rows.append(row)
        log(f"  {rows[-1]}")
    print(json.dumps({"workload": "rk4 sweep", "rows": rows}), flush=True)


def workload_rk4large(args):
    """The RK4 step alone on the HBM-bound batch (L = 1024, 64 replicas, 1.6 GB of state;
    automatic kernel choice), for profiling its traffic."""
    import torch

    import asd_amd
    ctx = asd_amd.Context(0)
    ctx.set_timing(True)
    ctx.rk4_set_exact(not args.rk4_fast)
    p = asd_amd.params(J=1.0, D=0.1, dt=0.01)
    L, nrep, steps = 1024, 64, 30
    S = torch.tensor(asd_amd.init_spins(5, L * L), device="cuda").repeat(nrep, 1)
    ctx.rk4_run(p, L, S, steps)   # untimed: the clocks settle (the first run is ~10% slower)
    torch.cuda.synchronize()
    ctx.rk4_run(p, L, S, steps)
    ms = ctx.stage_times()["rk4"]
    rate = nrep * L * L * steps / (ms * 1e-3)
    res = {"workload": "rk4 large", "L": L, "replicas": nrep, "steps": steps,
           "rk4_arithmetic": "tolerance" if args.rk4_fast else "exact",
           "us_per_step": round(ms * 1e3 / steps, 2), "spin_steps_per_s": rate,
           "algorithmic_GBps": round(48.0 * rate / 1e9, 1), "hbm_frac": round(48.0 * rate / (HBM_PEAK_GBS * 1e9), 4)}
    if args.calib_copy:
        del S
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")      # 1 GiB
        y = torch.empty_like(x)
        for _ in range(2):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        res["calib_copy_bytes"] = {"read": x.numel() * 8, "write": y.numel() * 8}
    print(json.dumps(res), flush=True)


def workload_unfused128(args):
    """What fusing saves: the L=128 S(q,t) of the benchmark line (400 lags x 400 origins, 192 q)
    through the unfused FP64 path (library GEMM into an explicit 16384 x 16384 correlation matrix
    per lag, outer subtract, phase reduction as a second GEMM: the reference's
    gemm/llg_sqt_mkl.c data flow) beside the fused FP64 and FP16X2 paths, on the same snapshots."""
    import numpy as np
    import torch

    import asd_amd
    L, T0, nlags, every, nsteps = 128, 400, 400, 10, 8000
    ns, nsnap = L * L, nsteps // every
    q = asd_amd.qpath_gxmg(L)
    ctx = asd_amd.Context(0)
    S = torch.tensor(asd_amd.init_spins(1, ns), device="cuda")
    snaps = torch.empty((nsnap, ns, 3), dtype=torch.float64, device="cuda")
    ctx.rk4_run(asd_amd.params(J=1.0, D=0.1, dt=0.01), L, S, nsteps, every, nsnap, snaps)
    torch.cuda.synchronize()
    ctx.autotune(L, T0, "fp16x2")
    res, out = {}, {}
    for name, fn in (("fused_fp64", lambda: ctx.sqt(L, snaps, T0, nlags, q, "fp64")),
                     ("fused_fp16x2", lambda: ctx.sqt(L, snaps, T0, nlags, q, "fp16x2")),
                     ("unfused_fp64", lambda: ctx.sqt_unfused(L, snaps, T0, nlags, q))):
        fn()  # warm (workspaces, library handle)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out[name] = fn()
        torch.cuda.synchronize()
        res[name + "_s"] = round(time.perf_counter() - t0, 4)
        log(f"  {name}: {res[name + '_s']} s")
    scale = np.abs(out["fused_fp64"]).max(axis=1, keepdims=True)
    res["unfused_vs_fused_fp64_max_rel_err"] = float((np.abs(out["unfused_fp64"] - out["fused_fp64"]) / scale).max())
    res["fusion_gain_fp64"] = round(res["unfused_fp64_s"] / res["fused_fp64_s"], 3)
    res["fusion_gain_vs_fp16x2"] = round(res["unfused_fp64_s"] / res["fused_fp16x2_s"], 3)
    res["unfused_explicit_matrix_GB"] = round(ns * ns * 8 / 1e9, 2)
    print(json.dumps({"workload": "unfused vs fused S(q,t), L=128, 400 lags x 400 origins, 192 q", **res}), flush=True)


def profiled_traffic(cfg_name, precision, L, T0, nlags):
    """HBM bytes per GEMM launch (FETCH_SIZE + WRITE_SIZE, scaled by the plain-copy calibration) from the
    committed rocprofv3 summary, when that summary was taken with this configuration on this workload;
    None otherwise (the counters need a profiler run of their own)."""
    import glob
    pattern = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           f"sqt_gemm_*_{precision}_L{L}_hbm_traffic.md")
    for path in sorted(glob.glob(pattern)):
        try:
            with open(path) as f:
                kv = dict(re.findall(r"^- (\w+): (.+)$", f.read(), re.M))
        except OSError:
            continue
        if (kv.get("config"), kv.get("precision"), kv.get("workload")) == (cfg_name, precision,
                                                                           f"L={L} T0={T0} nlags={nlags}"):
            return int(kv["traffic_bytes_per_launch"])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3", "fp16x2", "fp16f8"],
                    help="GEMM precision of the timed (headline) pass; fp64 is the reference's arithmetic")
    ap.add_argument("--mixed", default="fp16f8", choices=["fp32", "bf16x3", "fp16x2", "fp16f8"],
                    help="split precision of the mixed-precision leg")
    ap.add_argument("--mixed-also", default="fp16x2", choices=["none", "fp32", "bf16x3", "fp16x2", "fp16f8"],
                    help="a second split precision reported beside the mixed-precision leg")
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--nsteps", type=int, default=8000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--T0", type=int, default=400)
    ap.add_argument("--nlags", type=int, default=400)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="plain128",
                    choices=["plain128", "precision256", "batch96", "adjoint64", "rk4sweep", "rk4large", "unfused128"],
                    help="plain128 is the benchmark line; the others report their own JSON")
    ap.add_argument("--no-rk4-large", action="store_true",
                    help="skip the RK4 measurement on the 1024 x 1024 x 64-replica batch")
    ap.add_argument("--rk4-fast", action="store_true",
                    help="rk4large / rk4sweep: the RK4 tolerance mode instead of the exact arithmetic")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 1 GiB-in / 1 GiB-out elementwise copy (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload != "plain128":
        {"precision256": workload_precision256, "batch96": workload_batch96, "adjoint64": workload_adjoint64,
         "rk4sweep": workload_rk4sweep, "rk4large": workload_rk4large,
         "unfused128": workload_unfused128}[args.workload](args)
        return

    import numpy as np
    import torch

    import asd_amd

    L, T0, nlags, every = args.L, args.T0, args.nlags, args.every
    ns = L * L
    nsnap = args.nsteps // every
    assert nsnap >= T0 + nlags - 1, "not enough snapshots for T0 + nlags - 1"
    q = asd_amd.qpath_gxmg(L)
    nq = len(q)
    gemm_flop = float(nlags) * 2.0 * ns * ns * 3 * T0
    prm = asd_amd.params(J=1.0, D=0.1, Bz=0.0, alpha=0.0, dt=0.01)

    ctx = asd_amd.Context(0)
    ctx.set_timing(True)
    S0 = torch.tensor(asd_amd.init_spins(1, ns), device="cuda")
    spins = torch.empty_like(S0)
    snaps = torch.empty((nsnap, ns, 3), dtype=torch.float64, device="cuda")
    out_main = torch.empty((nlags, nq, 2), dtype=torch.float64, device="cuda")

    def one_pass(prec, out):
        spins.copy_(S0)
        ctx.rk4_run(prm, L, spins, args.nsteps, every, nsnap, snaps)
        st_rk = ctx.stage_times()["rk4"]
        ctx.sqt(L, snaps, T0, nlags, q, prec, out=out)
        st = ctx.stage_times()
        st["rk4"] = st_rk
        return st

    log(f"bench: L={L} nsteps={args.nsteps} nsnap={nsnap} T0={T0} nlags={nlags} nq={nq} prec={args.precision}")
    mixed = None if args.no_mixed or args.mixed == args.precision else args.mixed
    also = None if (not mixed or args.mixed_also in ("none", args.precision, mixed)) else args.mixed_also
    tuned = {}
    for prec in [args.precision] + [m for m in (mixed, also) if m]:
        best, times = ctx.autotune(L, T0, prec)
        tuned[prec] = {"chosen": list(times)[best], "gemm_ms_8_lags": {k: round(v, 3) for k, v in times.items()}}
        log(f"  autotune {prec}: {tuned[prec]}")
    for w in range(args.warmup):
        one_pass(args.precision, out_main)
        torch.cuda.synchronize()
        log(f"  warmup {w} done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stages = []
    for k in range(args.steps):
        stages.append(one_pass(args.precision, out_main))
        log(f"  step {k}: {stages[-1]}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    avg = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    sqt_ms = avg["pack"] + avg["gemm"] + avg["fourier"] + avg["finish"]
    rk_ms = avg["rk4"]
    value = gemm_flop / (sqt_ms * 1e-3) / 1e12
    rk_rate = args.nsteps * ns / (rk_ms * 1e-3)
    gemm_tflops = gemm_flop / (avg["gemm"] * 1e-3) / 1e12
    split = args.precision in ("fp16x2", "bf16x3", "fp16f8")

    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"plain L={L} (Ns={ns}), J=1 D=0.1 B=0 alpha=0 seed 1, {args.nsteps} RK4 steps "
                               f"dt=0.01, snapshot every {every} ({nsnap}), {nlags} lags x {T0} origins, "
                               f"Gamma-X-M-Gamma ({nq} q)",
                   "L": L, "rk4_steps": args.nsteps, "snapshots": nsnap, "T0": T0, "lags": nlags, "nq": nq,
                   "gemm_precision": args.precision, "rk4_precision": "fp64"},
        "roofline": {"bound": "mfma", "achieved": round(gemm_tflops, 3), "peak": PEAK_TFLOPS[args.precision],
                     "unit": "TFLOP/s", "frac": round(gemm_tflops / PEAK_TFLOPS[args.precision], 4),
                     "traffic": profiled_traffic(tuned[args.precision]["chosen"], args.precision, L, T0, nlags),
                     "kernel": f"S(q,t) GEMM, {args.precision}, {tuned[args.precision]['chosen']} "
                               f"(all {nlags} lags in one launch)",
                     "mfma_products_per_flop": PRODUCTS_PER_FLOP[args.precision]},
        # the reference applies the phase in the GEMM epilogue: 4 flops per G element and q; here the
        # epilogue bins G by displacement (one fixed-point add per element) and the phases act on the L^2
        # real bins per lag (a multiply-add each for cos and sin: 4 flops per bin and q)
        "epilogue": {"reference_fused_flops": 4.0 * nq * float(ns) * ns * nlags,
                     "binning_adds": float(ns) * ns * nlags,
                     "bin_dft_flops": 4.0 * ns * nq * nlags,
                     "fourier_ms": round(avg["fourier"], 3), "finish_ms": round(avg["finish"], 3)},
        "stage_ms": {k: round(v, 3) for k, v in avg.items()},
        "autotune": tuned,
        "rk4_spin_steps_per_s": rk_rate,
        "rk4_hbm_frac": round(48.0 * rk_rate / (HBM_PEAK_GBS * 1e9), 5),
        "sqt_s": round(sqt_ms / 1e3, 4),
    }

    def mixed_leg(prec):
        """One timed end-to-end pass in split precision `prec` after a warm one: times, the GEMM's
        roofline and the error against the FP64 result of the timed passes (same snapshots)."""
        out_mix = torch.empty_like(out_main)
        one_pass(prec, out_mix)                    # warm the split path (workspace)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stm = one_pass(prec, out_mix)
        torch.cuda.synchronize()
        wall_m = time.perf_counter() - t0
        sqtm_ms = stm["pack"] + stm["gemm"] + stm["fourier"] + stm["finish"]
        gm_tf = gemm_flop / stm["gemm"] / 1e9
        a = out_mix.cpu().numpy()
        b = out_main.cpu().numpy()
        sa, sb = a[..., 0] + 1j * a[..., 1], b[..., 0] + 1j * b[..., 1]
        scale = np.abs(sb).max(axis=1, keepdims=True)
        e = np.abs(sa - sb) / scale
        leg = {
            "precision": prec, "sqt_s": round(sqtm_ms / 1e3, 4), "gemm_s": round(stm["gemm"] / 1e3, 4),
            "sqt_tflops": round(gemm_flop / sqtm_ms / 1e9, 3), "end_to_end_s": round(wall_m, 4),
            "roofline": {"bound": "mfma", "achieved": round(gm_tf, 3), "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                         "frac": round(gm_tf / PEAK_TFLOPS[prec], 4),
                         "mfma_products_per_flop": PRODUCTS_PER_FLOP[prec],
                         "mfma_work_frac": round(FP16_WORK_PER_FLOP[prec] * gm_tf / PEAK_TFLOPS[prec], 4),
                         "traffic": profiled_traffic(tuned[prec]["chosen"], prec, L, T0, nlags),
                         "kernel": f"S(q,t) GEMM, {prec}, {tuned[prec]['chosen']}"},
            "err_vs_fp64": {"max": float(e.max()), "rms": float(np.sqrt((e ** 2).mean())),
                            "floor": "relative to max_q |S(q,tau)| of each lag, same snapshots"}}
        if args.precision == "fp64":
            leg["vs_fp64"] = round((rk_ms + sqt_ms) / (stm["rk4"] + sqtm_ms), 3)
        log(f"  mixed pass {prec}: {stm}")
        return leg, stm

    legs = {}
    if mixed:
        res["mixed_precision"], legs[mixed] = mixed_leg(mixed)
        if "vs_fp64" in res["mixed_precision"]:
            res["mixed_vs_fp64"] = res["mixed_precision"].pop("vs_fp64")
    if also:
        res["mixed_precision_" + also], legs[also] = mixed_leg(also)

    # calibration: the vendor library's plain fp16 GEMM (torch.matmul -> hipBLASLt) at one lag's
    # shape with K padded to a multiple of 64 (the library's faster shape: 16384 x 16384 x 1200
    # runs ~15% below x 1216), fp16 output, no epilogue, against the split kernel's MFMA work
    # rate (three products per useful flop, over the K its MFMAs cover: K rounded up to the
    # MFMA depth of 16)
    split_prec = args.precision if args.precision in ("fp16x2", "bf16x3") else next(
        (m for m in (mixed, also) if m in ("fp16x2", "bf16x3")), None)
    if split_prec:
        split_tf = gemm_tflops if split_prec == args.precision else gemm_flop / legs[split_prec]["gemm"] / 1e9
        kpad = (3 * T0 + 63) // 64 * 64
        kexec = (3 * T0 + 15) // 16 * 16
        xa = torch.randn(ns, kpad, device="cuda", dtype=torch.float16)
        xb = torch.randn(kpad, ns, device="cuda", dtype=torch.float16)
        for _ in range(3):
            xc = xa @ xb
        torch.cuda.synchronize()
        nmm = 20
        t_mm = time.perf_counter()
        for _ in range(nmm):
            xc = xa @ xb
        torch.cuda.synchronize()
        lib_tf = 2.0 * ns * ns * kpad * nmm / (time.perf_counter() - t_mm) / 1e12
        ours = 3.0 * split_tf * kexec / (3 * T0)
        res["library_calibration"] = {"hipblaslt_fp16_gemm_tflops": round(lib_tf, 1), "shape": f"{ns}x{ns}x{kpad}",
                                      "split_kernel_mfma_tflops": round(ours, 1),
                                      "ratio": round(ours / lib_tf, 3)}
        del xa, xb, xc
        log(f"  library calibration: {res['library_calibration']}")

    # calibration of the FP64 headline: the vendor library's plain DGEMM (torch.matmul on float64)
    # at one lag's shape, FP64 output, no epilogue, against the fused FP64 kernel's rate
    if args.precision == "fp64":
        xa = torch.randn(ns, 3 * T0, device="cuda", dtype=torch.float64)
        xb = torch.randn(3 * T0, ns, device="cuda", dtype=torch.float64)
        for _ in range(2):
            xc = xa @ xb
        torch.cuda.synchronize()
        nmm = 10
        t_mm = time.perf_counter()
        for _ in range(nmm):
            xc = xa @ xb
        torch.cuda.synchronize()
        lib64 = 2.0 * ns * ns * 3 * T0 * nmm / (time.perf_counter() - t_mm) / 1e12
        res["library_calibration_fp64"] = {"library_dgemm_tflops": round(lib64, 2), "shape": f"{ns}x{ns}x{3 * T0}",
                                           "fused_kernel_tflops": round(gemm_tflops, 2),
                                           "ratio": round(gemm_tflops / lib64, 3)}
        del xa, xb, xc
        log(f"  FP64 library calibration: {res['library_calibration_fp64']}")

    # the same RK4 step on a lattice batch far larger than the caches (L = 1024, 64
    # replicas: 1.6 GB of state), the HBM-bound regime; the automatic kernel choice
    # takes the register-streaming strips here
    # (the tolerance mode -- fused multiply-adds, within 1e-10 of the exact path over 100 steps
    # -- as the row's figures, the exact mode beside it)
    if not args.no_rk4_large:
        Lb, nb, steps_b = 1024, 64, 30
        Sb0 = torch.tensor(asd_amd.init_spins(5, Lb * Lb), device="cuda").repeat(nb, 1)
        row = {}
        for mode in ("exact", "tolerance"):
            ctx.rk4_set_exact(mode == "exact")
            Sb = Sb0.clone()
            ctx.rk4_run(prm, Lb, Sb, steps_b)   # untimed: the clocks settle (the first run is ~10% slower)
            torch.cuda.synchronize()
            ctx.rk4_run(prm, Lb, Sb, steps_b)
            ms_b = ctx.stage_times()["rk4"]
            rate_b = nb * Lb * Lb * steps_b / (ms_b * 1e-3)
            row[mode] = {"us_per_step": round(ms_b * 1e3 / steps_b, 2), "spin_steps_per_s": rate_b,
                         "algorithmic_GBps": round(48.0 * rate_b / 1e9, 1),
                         "hbm_frac": round(48.0 * rate_b / (HBM_PEAK_GBS * 1e9), 4)}
            del Sb
        ctx.rk4_set_exact(True)
        res["rk4_bandwidth_bound"] = {"L": Lb, "replicas": nb, "sites": nb * Lb * Lb, "steps": steps_b,
                                      "rk4_arithmetic": "tolerance mode (asd_rk4_set_exact(ctx, 0))",
                                      **row["tolerance"], "exact_mode": row["exact"]}
        del Sb0
        log(f"  rk4 bandwidth-bound: {res['rk4_bandwidth_bound']}")

    if not args.no_cpu:
        threads = args.cpu_threads or reference_threads()
        try:
            cb = cpu_baseline(L, T0, nq, nlags, args.nsteps, threads)
        except Exception as ex:  # the baseline is reported, never required
            log(f"  cpu baseline failed: {ex!r}")
            cb = None
        if cb:
            res["cpu_baseline"] = {
                "value": round(cb["sqt_tflops"], 5), "unit": "TFLOP/s", "cores": threads, "kind": "reference",
                "threads_rule": "min(128, omp_get_max_threads()) as llg_sqt_fusion.c:221-224",
                "host": host_cpu_facts(),
                "sample": f"reference kernel_fusion fused_dgemm at L={L}, K={3 * T0} (one q, one lag, with the "
                          f"disconnected prefill; 1 call) and 200 reference runge_kutta steps at L={L}",
                "rk4_spin_steps_per_s": cb["rk4_spin_steps_per_s"],
                "projected_full_sqt_s": round(cb["projected_sqt_s"], 1),
                "projected_full_rk4_s": round(cb["projected_rk4_s"], 2)}
        else:
            res["cpu_baseline"] = None
    if args.calib_copy:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")      # 1 GiB
        y = torch.empty_like(x)
        for _ in range(2):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        res["calib_copy_bytes"] = {"read": x.numel() * 8, "write": y.numel() * 8}
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
