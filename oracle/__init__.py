"""CPU FP64 oracle for the RK4-LLG + S(q,t) path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package.  It is the checker, never the product: the HIP
library (libasd) does not link, load or call anything here.

Two things live here:

* ``liboracle.so`` -- a plain-C restatement of hwchen2017/ASD-kernel-fusion's
  algorithm (kernel_fusion/llg_sqt_fusion.c, kernel_fusion/gemm_kernel.h),
  wrapped below with numpy in/out.
* ``_ref/libref_fusion*.so`` -- the reference's own kernel_fusion path compiled
  from its sources by ``oracle/Makefile`` (main renamed), driven through its
  globals and functions by :class:`RefFusion`.  Absent when the reference
  sources were not available at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [("J", C.c_double), ("D", C.c_double), ("Bz", C.c_double),
                ("alpha", C.c_double), ("dt", C.c_double), ("spin_length", C.c_double)]


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_hash64.restype = C.c_uint64
        L.orc_hash64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_init_spins.argtypes = [C.c_uint64, C.c_int, _dp]
        L.orc_llg_rhs.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, _dp, _dp]
        L.orc_runge_kutta.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_renormalise.argtypes = [C.c_long, C.c_double, _dp, _dp]
        L.orc_run.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, _dp, C.c_long, C.c_int, C.c_int, C.c_void_p]
        L.orc_energy.restype = C.c_double
        L.orc_energy.argtypes = [C.POINTER(OrcParams), C.c_int, _dp]
        L.orc_fused_dgemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int,
                                      _dp, C.c_int, _dp, _dp]
        L.orc_build_qmat.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        for f in (L.orc_sqt_gemm, L.orc_sqt_separable):
            f.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip, _dp]
        L.orc_tlm_run.argtypes = [C.POINTER(OrcParams), C.c_int, _dp, C.c_long, C.c_int, C.c_int, _dp, _dp]
        L.orc_tlm_sqt.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, _ip, _dp]
        _LIB = L
    return _LIB


def params(J=1.0, D=0.0, Bz=0.0, alpha=0.0, dt=0.01, spin_length=1.0) -> OrcParams:
    return OrcParams(J, D, Bz, alpha, dt, spin_length)


def init_spins(seed: int, nsites: int) -> np.ndarray:
    out = np.empty((nsites, 3))
    lib().orc_init_spins(seed, nsites, out)
    return out


def llg_rhs(p: OrcParams, L: int, S: np.ndarray) -> np.ndarray:
    S = np.ascontiguousarray(S, dtype=np.float64)
    nrep = S.size // (3 * L * L)
    out = np.empty_like(S)
    lib().orc_llg_rhs(C.byref(p), L, nrep, S, out)
    return out


def runge_kutta(p: OrcParams, L: int, spin: np.ndarray):
    spin = np.ascontiguousarray(spin, dtype=np.float64)
    nrep = spin.size // (3 * L * L)
    cur, k1, k2, k3, k4 = (np.empty_like(spin) for _ in range(5))
    lib().orc_runge_kutta(C.byref(p), L, nrep, spin, cur, k1, k2, k3, k4)
    return cur, k1, k2, k3, k4


def renormalise(cur: np.ndarray, length: float = 1.0) -> np.ndarray:
    cur = np.ascontiguousarray(cur, dtype=np.float64)
    out = np.empty_like(cur)
    lib().orc_renormalise(cur.size // 3, length, cur, out)
    return out


def run(p: OrcParams, L: int, S0: np.ndarray, nsteps: int, snap_every: int = 0, nsnap: int = 0):
    """Returns (final state, snapshots[nsnap] or None)."""
    S = np.array(S0, dtype=np.float64, copy=True, order="C")
    nrep = S.size // (3 * L * L)
    snaps = np.zeros((nsnap,) + S.shape) if nsnap else None
    lib().orc_run(C.byref(p), L, nrep, S, nsteps, snap_every, nsnap,
                  snaps.ctypes.data_as(C.c_void_p) if nsnap else None)
    return S, snaps


def energy(p: OrcParams, L: int, S: np.ndarray) -> float:
    return lib().orc_energy(C.byref(p), L, np.ascontiguousarray(S, dtype=np.float64))


def fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_c, s):
    """In-place on c and s (numpy float64 arrays)."""
    lib().orc_fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_c, s)


def build_qmat(L: int, mx: int, my: int) -> np.ndarray:
    ns = L * L
    q = np.zeros(2 * ((ns + 3) // 4 * 4) * ns)     # column stride 2 * ceil4(Ns)
    lib().orc_build_qmat(L, mx, my, q)
    return q


def _qarr(qlist):
    return np.ascontiguousarray(np.asarray(qlist, dtype=np.int32).reshape(-1, 2))


def sqt_gemm(L, snaps, T0, nlags, qlist) -> np.ndarray:
    q = _qarr(qlist)
    out = np.empty((nlags, len(q), 2))
    lib().orc_sqt_gemm(L, np.ascontiguousarray(snaps, dtype=np.float64), T0, nlags, len(q), q, out)
    return out[..., 0] + 1j * out[..., 1]


def sqt_separable(L, snaps, T0, nlags, qlist) -> np.ndarray:
    q = _qarr(qlist)
    out = np.empty((nlags, len(q), 2))
    lib().orc_sqt_separable(L, np.ascontiguousarray(snaps, dtype=np.float64), T0, nlags, len(q), q, out)
    return out[..., 0] + 1j * out[..., 1]


def tlm_run(p, L, S0, nsteps, snap_every, nsnap):
    ns = L * L
    snaps = np.empty((nsnap, ns, 3))
    dsnaps = np.empty((4, nsnap, ns, 3))
    lib().orc_tlm_run(C.byref(p), L, np.ascontiguousarray(S0, dtype=np.float64), nsteps, snap_every, nsnap,
                      snaps, dsnaps)
    return snaps, dsnaps


def tlm_sqt(L, snaps, dsnaps_one, T0, nlags, qlist) -> np.ndarray:
    q = _qarr(qlist)
    out = np.empty((nlags, len(q), 2))
    lib().orc_tlm_sqt(L, np.ascontiguousarray(snaps), np.ascontiguousarray(dsnaps_one), T0, nlags, len(q), q, out)
    return out[..., 0] + 1j * out[..., 1]


# ----------------------------------------------------------------------------------
# the reference's own kernel_fusion path, built from source into oracle/_ref
# ----------------------------------------------------------------------------------

class RefFusion:
    """Drives kernel_fusion/llg_sqt_fusion.c's functions through its globals.

    ``euler``/``runge_kutta`` read the globals N, Nfile, dt, J1, nn and
    num_neighbor (llg_sqt_fusion.c:13-21); ``fused_dgemm`` takes everything by
    pointer (gemm_kernel.h:355).  ``variant`` "parity" is built without FMA
    contraction, "perf" with the reference's AVX2/FMA flags at -O3.
    """

    def __init__(self, variant: str = "parity"):
        name = {"parity": "libref_fusion.so", "perf": "libref_fusion_perf.so"}[variant]
        path = os.path.join(_HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        self.lib.euler.argtypes = [dp, dp]
        self.lib.runge_kutta.argtypes = [dp] * 6
        self.lib.fused_dgemm.argtypes = [ip, ip, ip, dp, dp, ip, dp, ip, dp, ip, dp, dp]

    @staticmethod
    def available(variant: str = "parity") -> bool:
        name = {"parity": "libref_fusion.so", "perf": "libref_fusion_perf.so"}[variant]
        return os.path.exists(os.path.join(_HERE, "_ref", name))

    def set_globals(self, L: int, nfile: int, dt: float, J1: float = 1.0):
        C.c_int.in_dll(self.lib, "N").value = L
        C.c_int.in_dll(self.lib, "Nfile").value = nfile
        C.c_double.in_dll(self.lib, "dt").value = dt
        C.c_double.in_dll(self.lib, "J1").value = J1
        C.c_int.in_dll(self.lib, "num_neighbor").value = 4
        nn = (C.c_int * 8).in_dll(self.lib, "nn")
        for i, v in enumerate((1, 0, -1, 0, 0, -1, 0, 1)):   # llg_sqt_fusion.c:181-184
            nn[i] = v

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def euler(self, spin0: np.ndarray) -> np.ndarray:
        spin0 = np.ascontiguousarray(spin0, dtype=np.float64)
        out = np.empty_like(spin0)
        self.lib.euler(self._p(spin0), self._p(out))
        return out

    def runge_kutta(self, spin: np.ndarray):
        spin = np.ascontiguousarray(spin, dtype=np.float64)
        cur, k1, k2, k3, k4 = (np.empty_like(spin) for _ in range(5))
        self.lib.runge_kutta(*(self._p(x) for x in (spin, cur, k1, k2, k3, k4)))
        return cur, k1, k2, k3, k4

    def fused_dgemm(self, m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_c, sums):
        ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
        self.lib.fused_dgemm(ci(m), ci(n), ci(k), C.byref(C.c_double(alpha)), self._p(a), ci(lda),
                             self._p(b), ci(ldb), self._p(c), ci(ldc), self._p(fuse_c), self._p(sums))
