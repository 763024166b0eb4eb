"""Python handle on the gfx950 multiplication & solve library (libblaze_gpu.so).

The product is the C ABI in include/blaze_gpu.h (and the C++ `blaze::gpu` front end in
include/blaze/gpu/); this module only loads the built library with ctypes so that tests and
the benchmark can call it with device pointers (torch tensors serve as device memory).
Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libblaze_gpu.so"
GEN_PATH = LIB_DIR / "libbzg_gen.so"

_lib = None
_gen = None

i64, i32, f64, f32, vp = C.c_int64, C.c_int32, C.c_double, C.c_float, C.c_void_p

# name -> (restype, argtypes); mirrors include/blaze_gpu.h
SIGNATURES = {
    "bzg_create": (i32, [C.POINTER(vp)]),
    "bzg_destroy": (i32, [vp]),
    "bzg_set_stream": (i32, [vp, vp]),
    "bzg_get_stream": (vp, [vp]),
    "bzg_own_stream": (vp, [vp]),
    "bzg_synchronize": (i32, [vp]),
    "bzg_status_string": (C.c_char_p, [i32]),
    "bzg_last_error": (C.c_char_p, []),
    "bzg_profile_enable": (i32, [vp, i32]),
    "bzg_profile_read": (i32, [vp, C.c_char_p, C.POINTER(f64), C.POINTER(i64)]),
    "bzg_malloc": (i32, [C.POINTER(vp), C.c_size_t]),
    "bzg_free": (i32, [vp]),
    "bzg_memcpy": (i32, [vp, vp, vp, C.c_size_t, C.c_int]),
    "bzg_memset_zero": (i32, [vp, vp, C.c_size_t]),
    "bzg_dgemm": (i32, [vp, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64]),
    "bzg_dgemm_grouped": (i32, [vp, i64, vp, vp]),
    "bzg_gemm_autotune": (i32, [vp, i32, i64, i64, i64, C.POINTER(i32), C.POINTER(f32)]),
    "bzg_gemm_tuning_clear": (i32, []),
    "bzg_sgemm": (i32, [vp, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64]),
    "bzg_igemm_ex": (i32, [vp, i32, i32, i64, i64, i64, i32, vp, i64, vp, i64, i32, vp, i64]),
    "bzg_igemm_digits": (i32, [vp, i32, i32, i64, i64, i64, i32, vp, i64, vp, i64, i32, vp, i64]),
    "bzg_igemv_t": (i32, [vp, i64, i64, i32, vp, i64, vp, i32, vp]),
    "bzg_igemm": (i32, [vp, i64, i64, i64, i32, vp, i64, vp, i64, i32, vp, i64]),
    "bzg_dtranspose": (i32, [vp, i64, i64, vp, i64, vp, i64]),
    "bzg_dgemm_ex": (i32, [vp, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64]),
    "bzg_sgemm_ex": (i32, [vp, i32, i32, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64]),
    "bzg_dgemv_t": (i32, [vp, i64, i64, f64, vp, i64, vp, f64, vp]),
    "bzg_sgemv_t": (i32, [vp, i64, i64, f32, vp, i64, vp, f32, vp]),
    "bzg_dgemv": (i32, [vp, i64, i64, f64, vp, i64, vp, f64, vp]),
    "bzg_sgemv": (i32, [vp, i64, i64, f32, vp, i64, vp, f32, vp]),
    "bzg_igemv": (i32, [vp, i64, i64, i32, vp, i64, vp, i32, vp]),
    "bzg_csr_plan_create": (i32, [vp, i64, i64, vp, C.POINTER(vp)]),
    "bzg_csr_plan_set_sum_start": (i32, [vp, i32]),
    "bzg_csc_to_csr_host": (i32, [i64, i64, vp, vp, vp, i32, vp, vp, vp]),
    "bzg_mm_read_csr": (i32, [C.c_char_p, vp]),
    "bzg_csr_host_free": (i32, [vp]),
    "bzg_csr_plan_destroy": (i32, [vp]),
    "bzg_csr_plan_tune": (i32, [vp, i32, i32, i32]),
    "bzg_csr_plan_describe": (i32, [vp, C.c_char_p, C.c_size_t]),
    "bzg_csr_plan_autotune": (i32, [vp, vp, i32, vp, vp, vp, vp, vp, C.POINTER(f32)]),
    "bzg_dcsrmv": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "bzg_scsrmv": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "bzg_icsrmv": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dcg": (i32, [vp, vp, i64, vp, vp, vp, vp, vp, i32, vp]),
    "bzg_dgesv_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dgesv_batched_noalloc": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dgetrf_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dgetrf": (i32, [vp, i64, vp, i64, vp, vp]),
    "bzg_daxpy": (i32, [vp, i64, C.c_double, vp, vp]),
    "bzg_saxpy": (i32, [vp, i64, C.c_float, vp, vp]),
    "bzg_iaxpy": (i32, [vp, i64, i32, vp, vp]),
    "bzg_dgetrs": (i32, [vp, i64, i64, vp, i64, vp, vp, i64]),
    "bzg_dgesv": (i32, [vp, i64, i64, vp, i64, vp, vp, i64, vp]),
    "bzg_csr_plan_pack_sell": (i32, [vp, vp, vp, vp, vp, i32]),
    "bzg_dsellmv": (i32, [vp, vp, vp, vp]),
    "bzg_ssellmv": (i32, [vp, vp, vp, vp]),
    "bzg_isellmv": (i32, [vp, vp, vp, vp]),
    "bzg_zgemm_ex": (i32, [vp, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64]),
    "bzg_cgemm_ex": (i32, [vp, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64]),
    "bzg_zgemv_ex": (i32, [vp, i32, i64, i64, vp, vp, i64, vp, vp, vp]),
    "bzg_cgemv_ex": (i32, [vp, i32, i64, i64, vp, vp, i64, vp, vp, vp]),
    "bzg_zgesv": (i32, [vp, i64, i64, vp, i64, vp, i64, vp]),
    "bzg_ztranspose": (i32, [vp, i64, i64, vp, i64, vp, i64]),
    "bzg_dgetri": (i32, [vp, i64, vp, i64, vp]),
    "bzg_ddet_lu": (i32, [vp, i64, vp, i64, vp, vp, vp]),
    "bzg_dgetrs_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_ddet_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dinv_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp]),
    "bzg_dsgesv_batched": (i32, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, f64, i32, vp]),
    "bzg_batched_order": (i32, [vp, i64, vp, vp]),
    "bzg_solver_autotune": (i32, [vp, i32, i64, vp, vp, vp, vp, vp, i64, vp, i64, vp, C.POINTER(i32), C.POINTER(f32)]),
    "bzg_dgemm_fp64acc": (i32, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, i32]),
    "bzg_dgemm_fp64acc_i8": (i32, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, i32]),
    "bzg_fp64acc_last_plan": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(f64)]),
    "bzg_fp64acc_slices": (i32, [i64, i32]),
    "bzg_fp64acc_operand_bits": (i32, [i64, i32, i32]),
    "bzg_server_create": (i32, [vp, C.POINTER(vp)]),
    "bzg_server_destroy": (i32, [vp]),
    "bzg_server_add_csr": (i32, [vp, i64, i64, vp, vp, vp, C.POINTER(i32)]),
    "bzg_submit_gemm": (i32, [vp, i64, i64, i64, vp, i64, vp, i64, C.POINTER(i64)]),
    "bzg_submit_spmv": (i32, [vp, i32, vp, C.POINTER(i64)]),
    "bzg_submit_solve": (i32, [vp, i32, i32, vp, vp, C.POINTER(i64)]),
    "bzg_server_hold": (i32, [vp, i32]),
    "bzg_poll": (i32, [vp, i64, C.POINTER(i32)]),
    "bzg_wait": (i32, [vp, i64, vp]),
    "bzg_release": (i32, [vp, i64]),
    "bzg_server_get_stats": (i32, [vp, vp]),
}


class BlazeGpuError(RuntimeError):
    def __init__(self, fn: str, status: int):
        msg = lib().bzg_status_string(status).decode()
        detail = lib().bzg_last_error().decode()
        super().__init__(f"{fn} failed with status {status} ({msg}): {detail}")
        self.status = status


def lib() -> C.CDLL:
    """The loaded libblaze_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; "
                              "g.build()'` (or python blaze_amd/build.py) first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def gen() -> C.CDLL:
    """The host-side synthetic input generators (libbzg_gen.so)."""
    global _gen
    if _gen is None:
        if not GEN_PATH.exists():
            raise ImportError(f"{GEN_PATH} is missing: build first")
        _gen = C.CDLL(str(GEN_PATH))
        _gen.bzg_gen_laplace3d_nnz.restype = i64
    return _gen


def check(fn: str, status: int) -> int:
    if status < 0:
        raise BlazeGpuError(fn, status)
    return status


class Context:
    """Owns a bzg_context bound to the current HIP device; optionally on a given stream."""

    def __init__(self, stream: int | None = None):
        h = vp()
        check("bzg_create", lib().bzg_create(C.byref(h)))
        self.handle = h
        if stream is not None:
            check("bzg_set_stream", lib().bzg_set_stream(h, vp(stream)))

    def call(self, name: str, *args) -> int:
        return check(name, getattr(lib(), name)(self.handle, *args))

    def profile(self, on: bool = True) -> None:
        check("bzg_profile_enable", lib().bzg_profile_enable(self.handle, int(on)))

    def profile_read(self, kernel: str):
        """(total device ms, launches) of the named kernel since the last read."""
        ms, n = f64(), i64()
        check("bzg_profile_read", lib().bzg_profile_read(self.handle, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def sync(self) -> None:
        check("bzg_synchronize", lib().bzg_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().bzg_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrHost(C.Structure):
    _fields_ = [("n_rows", i64), ("n_cols", i64), ("nnz", i64), ("rowptr", C.POINTER(i32)),
                ("colidx", C.POINTER(i32)), ("val", C.POINTER(f64))]


def read_matrix_market(path):
    """(rowptr, colidx, val, (n_rows, n_cols)) of a Matrix Market coordinate file, read by the
    library's bzg_mm_read_csr (host arrays copied into numpy, then freed)."""
    import numpy as np
    h = CsrHost()
    check("bzg_mm_read_csr", lib().bzg_mm_read_csr(str(path).encode(), C.byref(h)))
    shape = (int(h.n_rows), int(h.n_cols))
    try:
        rp = np.ctypeslib.as_array(h.rowptr, (h.n_rows + 1,)).copy()
        ci = np.ctypeslib.as_array(h.colidx, (max(h.nnz, 1),))[:h.nnz].copy()
        va = np.ctypeslib.as_array(h.val, (max(h.nnz, 1),))[:h.nnz].copy()
    finally:
        lib().bzg_csr_host_free(C.byref(h))  # clears h
    return rp, ci, va, shape
