"""ctypes handle on the continuous-batching server of libblaze_gpu.so (bzg_server_* in
include/blaze_gpu.h).  Operands are device pointers (torch tensors serve as HBM buffers);
results stay in server-owned HBM until released.
"""
from __future__ import annotations

import ctypes as C

import blaze_amd as bz


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_inflight", C.c_int32), ("streams", C.c_int32),
                ("solve_batch_max", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("info", C.c_int32), ("iterations", C.c_int32),
                ("result", C.c_void_p), ("result_elems", C.c_int64), ("t_submit_us", C.c_double),
                ("t_admit_us", C.c_double), ("t_done_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "admitted", "completed", "released", "launches",
                                         "solve_batches", "max_solve_batch", "gemm_batches", "max_gemm_batch",
                                         "inflight", "bytes_in_use", "bytes_peak", "budget")]


KINDS = ("gemm", "spmv", "solve")


class Server:
    def __init__(self, budget_bytes: int, max_inflight: int = 32, streams: int = 4, solve_batch_max: int = 4096):
        cfg = ServerConfig(budget_bytes, max_inflight, streams, solve_batch_max)
        self.h = C.c_void_p()
        bz.check("bzg_server_create", bz.lib().bzg_server_create(C.byref(cfg), C.byref(self.h)))

    def close(self):
        if self.h:
            bz.check("bzg_server_destroy", bz.lib().bzg_server_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_csr(self, rowptr, colidx, val, n_cols: int) -> int:
        """Host numpy arrays (int32, int32, float64); returns the resident matrix handle."""
        h = C.c_int32()
        bz.check("bzg_server_add_csr", bz.lib().bzg_server_add_csr(
            self.h, len(rowptr) - 1, n_cols, rowptr.ctypes.data_as(C.c_void_p), colidx.ctypes.data_as(C.c_void_p),
            val.ctypes.data_as(C.c_void_p), C.byref(h)))
        return h.value

    def submit_gemm(self, m, n, k, A_ptr, lda, B_ptr, ldb) -> int:
        rid = C.c_int64()
        bz.check("bzg_submit_gemm", bz.lib().bzg_submit_gemm(self.h, m, n, k, C.c_void_p(A_ptr), lda,
                                                             C.c_void_p(B_ptr), ldb, C.byref(rid)))
        return rid.value

    def submit_spmv(self, csr: int, x_ptr) -> int:
        rid = C.c_int64()
        bz.check("bzg_submit_spmv", bz.lib().bzg_submit_spmv(self.h, csr, C.c_void_p(x_ptr), C.byref(rid)))
        return rid.value

    def submit_solve(self, n, nrhs, A_ptr, B_ptr) -> int:
        rid = C.c_int64()
        bz.check("bzg_submit_solve", bz.lib().bzg_submit_solve(self.h, n, nrhs, C.c_void_p(A_ptr), C.c_void_p(B_ptr),
                                                               C.byref(rid)))
        return rid.value

    def hold(self, on: bool) -> None:
        """Queue without admitting while on; lifting the hold admits the queued burst at once."""
        bz.check("bzg_server_hold", bz.lib().bzg_server_hold(self.h, 1 if on else 0))

    def poll(self, rid: int) -> bool:
        d = C.c_int32()
        bz.check("bzg_poll", bz.lib().bzg_poll(self.h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid: int) -> RequestResult:
        r = RequestResult()
        bz.check("bzg_wait", bz.lib().bzg_wait(self.h, rid, C.byref(r)))
        return r

    def release(self, rid: int) -> None:
        bz.check("bzg_release", bz.lib().bzg_release(self.h, rid))

    def stats(self) -> dict:
        s = ServerStats()
        bz.check("bzg_server_get_stats", bz.lib().bzg_server_get_stats(self.h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in ServerStats._fields_}


def result_tensor(res: RequestResult, shape):
    """Copy of a request's result (device) as a new torch tensor of the given shape."""
    import torch
    out = torch.empty(shape, dtype=torch.float64, device="cuda")
    nbytes = res.result_elems * 8
    assert out.numel() == res.result_elems
    ctx = _copy_ctx()
    ctx.call("bzg_memcpy", C.c_void_p(out.data_ptr()), C.c_void_p(res.result), nbytes, 2)
    return out


_cctx = None


def _copy_ctx():
    global _cctx
    if _cctx is None:
        _cctx = bz.Context()
    return _cctx
