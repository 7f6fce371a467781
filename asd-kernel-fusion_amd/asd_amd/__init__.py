"""asd_amd -- Python host side of libasd (MI355X / gfx950).

Thin ctypes layer over the C ABI in ``include/asd.h``.  Device buffers are torch
tensors on ``cuda:*`` (PyTorch-ROCm is only the allocator and stream provider
here); every compute call goes through ``libasd.so``.  There is no fallback: if
the library is missing or fails to load, every call raises.

Reference interface mirrored (hwchen2017/ASD-kernel-fusion, kernel_fusion/):
  ``euler``         llg_sqt_fusion.c:30   derivative of the spin field
  ``runge_kutta``   llg_sqt_fusion.c:98   one unnormalised RK4 step + increments
  ``fused_dgemm``   gemm_kernel.h:355     GEMM with the phase-weighted summing epilogue
and the native, device-resident pipeline that replaces the reference time loop
(llg_sqt_fusion.c:340-449): :func:`rk4_run`, :func:`sqt`.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libasd.so")

FP64, FP32, BF16X3, FP16X2, FP16F8 = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "fp16x2": FP16X2, "fp16f8": FP16F8}

ASD_OK = 0
ASD_EINVAL, ASD_ENOMEM, ASD_EHIP, ASD_ENOTSUP, ASD_ENOTFINITE = -1, -2, -3, -4, -5


class AsdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libasd error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """asd_params: J, D, B_z, alpha, dt, spin_length."""
    _fields_ = [("J", C.c_double), ("D", C.c_double), ("Bz", C.c_double),
                ("alpha", C.c_double), ("dt", C.c_double), ("spin_length", C.c_double)]


def params(J=1.0, D=0.0, Bz=0.0, alpha=0.0, dt=0.01, spin_length=1.0) -> Params:
    return Params(J, D, Bz, alpha, dt, spin_length)


_lib = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes)
_SIGS = {
    "asd_last_error": (C.c_char_p, []),
    "asd_version": (C.c_char_p, []),
    "asd_init_spins": (C.c_int, [C.c_uint64, C.c_int, _vp]),
    "asd_qpath_gxmg": (C.c_int, [C.c_int, _vp]),
    "asd_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "asd_destroy": (C.c_int, [_vp]),
    "asd_set_stream": (C.c_int, [_vp, _vp]),
    "asd_synchronize": (C.c_int, [_vp]),
    "asd_rk4_run": (C.c_int, [_vp, C.POINTER(Params), C.c_int, C.c_int, _vp, C.c_long, C.c_int, C.c_int, _vp,
                              C.c_int]),
    "asd_rk4_run_noalloc": (C.c_int, [_vp, C.POINTER(Params), C.c_int, C.c_int, _vp, C.c_long, C.c_int, C.c_int,
                                      _vp]),
    "asd_sqt": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_int]),
    "asd_sqt_replicas": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "asd_sqt_unfused": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int]),
    "asd_sqt_bruteforce": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int]),
    "asd_batch_rk4_run": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_long, C.c_int, C.c_int, _vp]),
    "asd_autotune_sqt": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _ip, _dp]),
    "asd_autotune_batch": (C.c_int, [_vp, C.c_int, _vp, _vp, _ip, _dp]),
    "asd_batch_num_variants": (C.c_int, []),
    "asd_batch_variant_name": (C.c_char_p, [C.c_int]),
    "asd_batch_set_variant": (C.c_int, [_vp, C.c_int]),
    "asd_batch_variant_for": (C.c_int, [_vp, C.c_int, _vp]),
    "asd_sqt_num_configs": (C.c_int, [C.c_int]),
    "asd_sqt_config_name": (C.c_char_p, [C.c_int, C.c_int]),
    "asd_sqt_set_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "asd_sqt_loss_grad": (C.c_int, [_vp, C.POINTER(Params), C.c_int, _vp, C.c_long, C.c_int, C.c_int, C.c_int,
                                    _vp, C.c_int, _vp, C.c_int, _dp, _dp]),
    "asd_malloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "asd_free": (C.c_int, [_vp, _vp]),
    "asd_memcpy": (C.c_int, [_vp, _vp, _vp, C.c_size_t, C.c_int]),
    "asd_sqt_vjp": (C.c_int, [_vp, C.POINTER(Params), C.c_int, _vp, C.c_long, C.c_int, C.c_int, C.c_int, _vp,
                              C.c_int, _vp, C.c_int, _dp, _vp]),
    "asd_rk4_set_tile": (C.c_int, [_vp, C.c_int]),
    "asd_rk4_set_exact": (C.c_int, [_vp, C.c_int]),
    "asd_set_timing": (C.c_int, [_vp, C.c_int]),
    "asd_stage_times": (C.c_int, [_vp, _dp]),
    "asd_euler": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int, _vp, _vp]),
    "asd_runge_kutta": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "asd_fused_dgemm": (C.c_int, [_ip, _ip, _ip, _dp, _vp, _ip, _vp, _ip, _vp, _ip, _vp, _vp]),
    "asd_compat_bind": (C.c_int, [_ip, _ip, _dp, _dp, _vp, _ip]),
    "euler": (None, [_vp, _vp]),
    "runge_kutta": (None, [_vp] * 6),
    "fused_dgemm": (None, [_ip, _ip, _ip, _dp, _vp, _ip, _vp, _ip, _vp, _ip, _vp, _vp]),
}


def lib():
    """Load libasd.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code: int):
    if code != ASD_OK:
        raise AsdError(code, lib().asd_last_error().decode())


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def init_spins(seed: int, nsites: int) -> np.ndarray:
    out = np.empty((nsites, 3))
    _check(lib().asd_init_spins(seed, nsites, _np_ptr(out)))
    return out


def qpath_gxmg(L: int) -> np.ndarray:
    q = np.empty((3 * L // 2, 2), dtype=np.int32)
    n = lib().asd_qpath_gxmg(L, _np_ptr(q))
    if n < 0:
        _check(n)
    return q


def _check_f64_cuda(x, name: str, numel: int = None):
    """Raise ValueError unless x is a contiguous float64 CUDA tensor (of `numel` elements)."""
    import torch
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float64 or not x.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 tensor on the GPU")
    if numel is not None and x.numel() != numel:
        raise ValueError(f"{name} holds {x.numel()} doubles, expected {numel}")


def full_grid(L: int) -> np.ndarray:
    mx, my = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    return np.ascontiguousarray(np.stack([mx.ravel(), my.ravel()], axis=1).astype(np.int32))


class Context:
    """An asd_ctx bound to one device; work is ordered on torch's current stream
    of that device at construction time (or ``stream``)."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self._torch = torch
        self.device = device
        h = C.c_void_p()
        _check(lib().asd_create(C.byref(h), device))
        self.h = h
        st = stream if stream is not None else torch.cuda.current_stream(device)
        self.stream = st
        _check(lib().asd_set_stream(self.h, C.c_void_p(st.cuda_stream)))

    def close(self):
        if self.h:
            lib().asd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().asd_synchronize(self.h))

    def rk4_set_tile(self, tile: int = 0):
        _check(lib().asd_rk4_set_tile(self.h, tile))

    def rk4_set_exact(self, exact: bool = True):
        """True: the reference's RK4 operation order (IEEE-equal values, the default);
        False: the tolerance mode (fused multiply-adds, fewer instructions per site)."""
        _check(lib().asd_rk4_set_exact(self.h, 1 if exact else 0))

    def set_timing(self, on: bool = True):
        _check(lib().asd_set_timing(self.h, 1 if on else 0))

    STAGES = ("pack", "gemm", "fourier", "finish", "rk4")

    def stage_times(self) -> dict:
        """ms of the last call of each stage, from HIP events on this context's stream."""
        t = (C.c_double * 5)()
        _check(lib().asd_stage_times(self.h, t))
        return dict(zip(self.STAGES, list(t)))

    # -------------------------------------------------------------- RK4
    def _rk4_args(self, L, spins, snap_every, nsnap, snaps):
        if L < 2:
            raise ValueError(f"L must be >= 2 (got {L})")
        _check_f64_cuda(spins, "spins")
        nrep = spins.numel() // (3 * L * L)
        if nrep < 1 or nrep * 3 * L * L != spins.numel():
            raise ValueError(f"spins must hold nrep x {L * L} x 3 doubles (got {spins.numel()})")
        sp = None
        if snaps is not None:
            if snap_every < 1 or nsnap < 0:
                raise ValueError("snap_every must be >= 1 and nsnap >= 0 with a snapshot buffer")
            _check_f64_cuda(snaps, "snaps", nsnap * spins.numel())
            sp = C.c_void_p(snaps.data_ptr())
        return nrep, sp

    def rk4_run(self, p: Params, L: int, spins, nsteps: int, snap_every: int = 0, nsnap: int = 0, snaps=None,
                use_graph: bool = True):
        """spins: cuda float64 tensor [nrep, L*L, 3] or [L*L, 3] (in/out);
        snaps: cuda float64 tensor [nsnap, ...spins.shape] or None.  Raises AsdError with
        code ASD_ENOTFINITE when a step leaves the range where it is defined (NaN states)."""
        nrep, sp = self._rk4_args(L, spins, snap_every, nsnap, snaps)
        _check(lib().asd_rk4_run(self.h, C.byref(p), L, nrep, C.c_void_p(spins.data_ptr()), nsteps, snap_every,
                                 nsnap, sp, 1 if use_graph else 0))

    def rk4_run_noalloc(self, p: Params, L: int, spins, nsteps: int, snap_every: int = 0, nsnap: int = 0,
                        snaps=None):
        """asd_rk4_run_noalloc: as rk4_run(use_graph=False) with no device allocation (fixed 16 MiB
        scratch: nrep * L * L * 3 <= 2^21); only enqueues the work (no numeric guard read-back)."""
        nrep, sp = self._rk4_args(L, spins, snap_every, nsnap, snaps)
        _check(lib().asd_rk4_run_noalloc(self.h, C.byref(p), L, nrep, C.c_void_p(spins.data_ptr()), nsteps,
                                         snap_every, nsnap, sp))

    def batch_rk4_run(self, Ls, plist, spins_list, nsteps: int, snap_every: int = 0, nsnap: int = 0,
                      snaps_list=None):
        """Uneven-size trajectories in one persistent launch per step.  spins_list[t]:
        cuda float64 [L_t*L_t, 3] (in/out); snaps_list[t]: [nsnap, L_t*L_t, 3] or None."""
        n = len(Ls)
        if len(plist) != n or len(spins_list) != n or (snaps_list is not None and len(snaps_list) != n):
            raise ValueError("one parameter set, state and snapshot buffer per trajectory")
        t = self._torch
        for i, (L, s) in enumerate(zip(Ls, spins_list)):
            if not (s.is_cuda and s.dtype == t.float64 and s.is_contiguous()) or s.numel() != 3 * L * L:
                raise ValueError(f"spins_list[{i}] must be a contiguous cuda float64 tensor of [{L * L}, 3]")
            if snaps_list is not None:
                x = snaps_list[i]
                if not (x.is_cuda and x.dtype == t.float64 and x.is_contiguous()) or x.numel() < nsnap * 3 * L * L:
                    raise ValueError(f"snaps_list[{i}] must hold [{nsnap}, {L * L}, 3] float64 on the GPU")
        Ls_c = (C.c_int * n)(*Ls)
        P = (Params * n)(*plist)
        sp = (C.c_void_p * n)(*[s.data_ptr() for s in spins_list])
        sn = None
        if snaps_list is not None:
            sn = (C.c_void_p * n)(*[s.data_ptr() for s in snaps_list])
        _check(lib().asd_batch_rk4_run(self.h, n, Ls_c, P, sp, nsteps, snap_every, nsnap, sn))

    def autotune_batch(self, Ls, plist):
        """Times every batch kernel variant for this composition; returns (best, {name: ms})
        and records the choice for later batch_rk4_run calls with the same multiset of L."""
        n = len(Ls)
        if len(plist) != n or n < 1:
            raise ValueError("one parameter set per trajectory")
        nv = lib().asd_batch_num_variants()
        times = (C.c_double * nv)()
        best = C.c_int()
        _check(lib().asd_autotune_batch(self.h, n, (C.c_int * n)(*Ls), (Params * n)(*plist), C.byref(best), times))
        names = [lib().asd_batch_variant_name(i).decode() for i in range(nv)]
        return best.value, dict(zip(names, list(times)))

    def batch_set_variant(self, variant: int = -1):
        _check(lib().asd_batch_set_variant(self.h, variant))

    def batch_variant_for(self, Ls) -> int:
        """The variant batch_rk4_run would take for this composition (-1: not tuned yet)."""
        r = lib().asd_batch_variant_for(self.h, len(Ls), (C.c_int * len(Ls))(*Ls))
        if r < -1:
            _check(r)
        return r

    # ---------------------------------------------------------- autotuning
    def autotune(self, L: int, T0: int, precision="fp16x2"):
        """Times every tile configuration for (L, K=3 T0); returns (best, {name: ms})."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        n = lib().asd_sqt_num_configs(prec)
        times = (C.c_double * n)()
        best = C.c_int()
        _check(lib().asd_autotune_sqt(self.h, L, T0, prec, C.byref(best), times))
        names = [lib().asd_sqt_config_name(prec, i).decode() for i in range(n)]
        return best.value, dict(zip(names, list(times)))

    def set_config(self, precision, cfg: int):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        _check(lib().asd_sqt_set_config(self.h, prec, cfg))

    def _check_state(self, S0, L: int):
        t = self._torch
        if not (S0.is_cuda and S0.dtype == t.float64 and S0.is_contiguous()) or S0.numel() != 3 * L * L:
            raise ValueError(f"S0 must be a contiguous cuda float64 tensor of [{L * L}, 3]")

    # ------------------------------------------------------------ adjoint
    def sqt_vjp(self, p: Params, L: int, S0, nsteps: int, every: int, T0: int, nlags: int, q, g=None,
                checkpoint_every: int = 50):
        """Returns (S(q,tau) complex [nlags, nq], grad[4] or None) for upstream g
        (float64 [nlags, nq, 2], dL/dRe S and dL/dIm S)."""
        self._check_state(S0, L)
        q = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(-1, 2))
        s = np.empty((nlags, len(q), 2))
        grad = np.zeros(4)
        gp = None
        if g is not None:
            g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(nlags, len(q), 2))
            gp = _np_ptr(g)
        _check(lib().asd_sqt_vjp(self.h, C.byref(p), L, C.c_void_p(S0.data_ptr()), nsteps, every, T0, nlags,
                                 _np_ptr(q), len(q), gp, checkpoint_every,
                                 grad.ctypes.data_as(_dp), _np_ptr(s)))
        return s[..., 0] + 1j * s[..., 1], (grad if g is not None else None)

    def sqt_loss_grad(self, p: Params, L: int, S0, nsteps: int, every: int, T0: int, nlags: int, q, target,
                      checkpoint_every: int = 50):
        self._check_state(S0, L)
        q = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(-1, 2))
        t = np.ascontiguousarray(np.stack([np.real(target), np.imag(target)], axis=-1), dtype=np.float64)
        if t.size != 2 * nlags * len(q):
            raise ValueError(f"target must be complex [{nlags}, {len(q)}] (got {np.shape(target)})")
        loss = C.c_double()
        grad = np.zeros(4)
        _check(lib().asd_sqt_loss_grad(self.h, C.byref(p), L, C.c_void_p(S0.data_ptr()), nsteps, every, T0, nlags,
                                       _np_ptr(q), len(q), _np_ptr(t), checkpoint_every, C.byref(loss),
                                       grad.ctypes.data_as(_dp)))
        return loss.value, grad

    # ------------------------------------------------------------ S(q,t)
    def sqt(self, L: int, snaps, T0: int, nlags: int, q, precision="fp64", out=None):
        """snaps: cuda float64 [nsnap, L*L, 3].  Returns complex128 numpy [nlags, nq]
        (or fills the cuda float64 tensor ``out`` [nlags, nq, 2] and returns it)."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        q = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(-1, 2))
        _check_f64_cuda(snaps, "snaps")
        nsnap = snaps.shape[0] if snaps.dim() > 0 else 0
        if snaps.numel() != nsnap * 3 * L * L:
            raise ValueError(f"snaps must be [nsnap, {L * L}, 3] (got {tuple(snaps.shape)})")
        if out is not None:
            _check_f64_cuda(out, "out", 2 * nlags * len(q))
            _check(lib().asd_sqt(self.h, L, C.c_void_p(snaps.data_ptr()), nsnap, T0, nlags, _np_ptr(q), len(q),
                                 prec, C.c_void_p(out.data_ptr()), 1))
            return out
        res = np.empty((nlags, len(q), 2))
        _check(lib().asd_sqt(self.h, L, C.c_void_p(snaps.data_ptr()), nsnap, T0, nlags, _np_ptr(q), len(q), prec,
                             _np_ptr(res), 0))
        return res[..., 0] + 1j * res[..., 1]

    def _sqt_other(self, fn, L, snaps, T0, nlags, q):
        q = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(-1, 2))
        _check_f64_cuda(snaps, "snaps")
        if snaps.numel() != snaps.shape[0] * 3 * L * L:
            raise ValueError(f"snaps must be [nsnap, {L * L}, 3] (got {tuple(snaps.shape)})")
        res = np.empty((nlags, len(q), 2))
        _check(fn(self.h, L, C.c_void_p(snaps.data_ptr()), snaps.shape[0], T0, nlags, _np_ptr(q), len(q),
                  _np_ptr(res), 0))
        return res[..., 0] + 1j * res[..., 1]

    def sqt_unfused(self, L: int, snaps, T0: int, nlags: int, q):
        """The unfused FP64 path (explicit correlation matrix, library GEMMs); as sqt()."""
        return self._sqt_other(lib().asd_sqt_unfused, L, snaps, T0, nlags, q)

    def sqt_bruteforce(self, L: int, snaps, T0: int, nlags: int, q):
        """Direct double sum over sites for L <= 16 (independent check); as sqt()."""
        return self._sqt_other(lib().asd_sqt_bruteforce, L, snaps, T0, nlags, q)

    def sqt_replicas(self, L: int, now, zero, q, precision="fp64"):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        q = np.ascontiguousarray(np.asarray(q, dtype=np.int32).reshape(-1, 2))
        nrep = now.numel() // (3 * L * L)
        for name, x in (("now", now), ("zero", zero)):
            if not (x.is_cuda and x.dtype == self._torch.float64 and x.is_contiguous()) or x.numel() != nrep * 3 * L * L \
                    or nrep < 1:
                raise ValueError(f"{name} must be a contiguous cuda float64 tensor of [nrep, {L * L}, 3]")
        res = np.empty((len(q), 2))
        _check(lib().asd_sqt_replicas(self.h, L, nrep, C.c_void_p(now.data_ptr()), C.c_void_p(zero.data_ptr()),
                                      _np_ptr(q), len(q), prec, _np_ptr(res)))
        return res[..., 0] + 1j * res[..., 1]


# --------------------------------------------------- reference-compatible (host)

def euler(p: Params, L: int, spin0: np.ndarray) -> np.ndarray:
    spin0 = np.ascontiguousarray(spin0, dtype=np.float64)
    out = np.empty_like(spin0)
    _check(lib().asd_euler(C.byref(p), L, spin0.size // (3 * L * L), _np_ptr(spin0), _np_ptr(out)))
    return out


def runge_kutta(p: Params, L: int, spin: np.ndarray):
    spin = np.ascontiguousarray(spin, dtype=np.float64)
    outs = [np.empty_like(spin) for _ in range(5)]
    _check(lib().asd_runge_kutta(C.byref(p), L, spin.size // (3 * L * L), _np_ptr(spin),
                                 *(_np_ptr(o) for o in outs)))
    return tuple(outs)


def fused_dgemm(m, n, k, alpha, a, lda, b, ldb, c, ldc, fuse_c, sums):
    """gemm_kernel.h:355 semantics; c and sums (numpy float64) updated in place.
    a: m x k column-major (lda >= m), b: k x n (ldb >= k), c and fuse_c: the blocked
    4-real/4-imaginary layout with ldc >= 2 ceil4(m) per column, sums: at least 2."""
    need = {"a": (a, lda * k), "b": (b, ldb * n), "c": (c, ldc * n), "fuse_c": (fuse_c, ldc * n),
            "sums": (sums, 2)}
    if m < 0 or n < 0 or k < 0 or lda < max(m, 1) or ldb < max(k, 1) or ldc < 2 * ((m + 3) // 4 * 4):
        raise ValueError(f"bad sizes m={m} n={n} k={k} lda={lda} ldb={ldb} ldc={ldc}")
    for name, (arr, size) in need.items():
        if not (isinstance(arr, np.ndarray) and arr.dtype == np.float64 and arr.flags.c_contiguous):
            raise ValueError(f"{name} must be a C-contiguous float64 numpy array")
        if arr.size < size:
            raise ValueError(f"{name} holds {arr.size} doubles, needs {size}")
    ci = lambda v: C.byref(C.c_int(v))  # noqa: E731
    _check(lib().asd_fused_dgemm(ci(m), ci(n), ci(k), C.byref(C.c_double(alpha)), _np_ptr(a), ci(lda),
                                 _np_ptr(b), ci(ldb), _np_ptr(c), ci(ldc), _np_ptr(fuse_c), _np_ptr(sums)))


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)
