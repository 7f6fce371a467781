"""C-ABI checks that need no GPU: libasd.so loads, exports every function the
headers in include/ declare, and its host-only helpers (no device calls) agree
with the oracle bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("asd.h", "asd_compat.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([a-z_0-9]+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for n in ("euler", "runge_kutta", "fused_dgemm", "asd_fused_dgemm", "asd_rk4_run", "asd_sqt"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import asd_amd
    path = asd_amd.LIB_PATH
    assert os.path.exists(path), "libasd.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    lib = asd_amd.lib()          # dlopen + argtypes for every declared function
    for n in declared_functions():
        assert hasattr(lib, n)
    assert set(declared_functions()) <= set(asd_amd.exported_symbols())


def test_init_spins_matches_oracle(orc):
    import asd_amd
    for seed in (0, 1, 2, 3, 100, 195):
        assert np.array_equal(asd_amd.init_spins(seed, 4096), orc.init_spins(seed, 4096))


def test_gxmg_path():
    import asd_amd
    q = asd_amd.qpath_gxmg(128)
    assert q.shape == (192, 2)
    assert tuple(q[0]) == (0, 0) and tuple(q[64]) == (64, 0) and tuple(q[128]) == (64, 64)
    assert tuple(q[-1]) == (1, 1)


def test_error_path_without_device_work():
    import asd_amd
    with pytest.raises(asd_amd.AsdError):
        asd_amd.qpath_gxmg(7)      # odd L rejected on the host



def test_fused_dgemm_wrapper_rejects_bad_arrays():
    """The host-pointer fused_dgemm hands raw pointers to C, so the wrapper checks dtype,
    layout and sizes first (before the library is touched: no GPU needed)."""
    import asd_amd
    m = n = k = 8
    ldc = 2 * m
    ok = dict(a=np.zeros(m * k), b=np.zeros(k * n), c=np.zeros(ldc * n), fuse_c=np.zeros(ldc * n), sums=np.zeros(2))

    def call(**kw):
        a = {**ok, **kw}
        asd_amd.fused_dgemm(m, n, k, 1.0, a["a"], m, a["b"], k, a["c"], ldc, a["fuse_c"], a["sums"])

    bad = [dict(a=np.zeros(m * k, dtype=np.float32)),           # dtype
           dict(b=np.zeros((n, 2 * k))[:, ::2]),                # not contiguous
           dict(c=np.zeros(ldc * n - 1)),                       # too small
           dict(sums=np.zeros(1)),
           dict(fuse_c=list(range(ldc * n)))]                   # not an ndarray
    for kw in bad:
        with pytest.raises(ValueError):
            call(**kw)
    with pytest.raises(ValueError):   # ldc below 2 ceil4(m)
        asd_amd.fused_dgemm(m, n, k, 1.0, ok["a"], m, ok["b"], k, ok["c"], ldc - 2, ok["fuse_c"], ok["sums"])


def test_context_wrappers_reject_bad_tensors_before_any_call():
    """rk4_run / rk4_run_noalloc / sqt hand device pointers to C, so the wrappers check
    that every buffer is a contiguous float64 GPU tensor of the right size first (explicit
    ValueError, not assert: python -O keeps them).  A CPU or mis-sized tensor raises
    before the library is touched, so this runs without a GPU (no context is created)."""
    import torch

    import asd_amd
    ctx = object.__new__(asd_amd.Context)   # no device: the checks run before any C call
    ctx.h = None
    L = 4
    p = asd_amd.params()
    cpu = torch.zeros(L * L, 3, dtype=torch.float64)
    meta = torch.empty(L * L, 3, dtype=torch.float64, device="meta")
    for fn in (ctx.rk4_run, ctx.rk4_run_noalloc):
        with pytest.raises(ValueError):
            fn(p, L, cpu, 10)                                   # spins on the host
        with pytest.raises(ValueError):
            fn(p, L, cpu.float(), 10)                           # wrong dtype (checked first: on the host)
        with pytest.raises(ValueError):
            fn(p, L, meta, 10, 5, 3, cpu.repeat(3, 1))          # snaps: not a GPU tensor
    with pytest.raises(ValueError):
        ctx.sqt(L, cpu.reshape(1, L * L, 3), 1, 1, [(0, 0)])    # snapshots on the host
    with pytest.raises(ValueError):
        ctx.sqt_unfused(L, cpu.reshape(1, L * L, 3), 1, 1, [(0, 0)])
    ctx.h = None
