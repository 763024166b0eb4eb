"""Helpers for the GPU parity tests: device buffers are torch tensors, calls go through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import blaze_amd as bz  # noqa: E402


def have_gpu() -> bool:
    return torch.cuda.is_available()


def dev(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def empty(shape, dtype):
    tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
           np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64}[np.dtype(dtype)]
    return torch.empty(shape, dtype=tdt, device="cuda")


def full(shape, dtype, value):
    t = empty(shape, dtype)
    t.fill_(value)
    return t


def p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def host(t) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


_ctx = None


def ctx() -> bz.Context:
    global _ctx
    if _ctx is None:
        _ctx = bz.Context(stream=torch.cuda.current_stream().cuda_stream)
    return _ctx
