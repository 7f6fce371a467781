"""Shared helpers for the GPU parity tests (they run through libasd's C ABI)."""
import sys

import numpy as np
import pytest


def need_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b, floor_frac=1e-3):
    """max |a-b| / max(|b|, floor_frac * max|b|) -- relative error with a floor."""
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), floor_frac * np.abs(b).max())
    return float((np.abs(np.asarray(a) - b) / scale).max())


def err_stats(a, ref):
    """(max, rms) of |a - ref| / max|ref| per lag, reduced over lags."""
    a, ref = np.asarray(a), np.asarray(ref)
    scale = np.abs(ref).max(axis=-1, keepdims=True)
    e = np.abs(a - ref) / scale
    return float(e.max()), float(np.sqrt((e ** 2).mean()))
