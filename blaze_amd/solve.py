"""Batched solve helpers over the C ABI (packing of uneven-size systems into one pool).

Layout (include/blaze_gpu.h): A_i row-major n_i x n_i at A + a_off[i]; B_i/X_i row-major
n_i x nrhs_i at B + b_off[i]; ipiv_i at ipiv + p_off[i].  Device memory comes from torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import blaze_amd as bz


def pack(As, Bs):
    n = np.array([a.shape[0] for a in As], np.int32)
    r = np.array([b.shape[1] for b in Bs], np.int32)
    a_off = np.zeros(len(As), np.int64)
    b_off = np.zeros(len(As), np.int64)
    if len(As):
        a_off[1:] = np.cumsum((n.astype(np.int64) ** 2))[:-1]
        b_off[1:] = np.cumsum(n.astype(np.int64) * r)[:-1]
    Ap = np.concatenate([np.ascontiguousarray(a, np.float64).ravel() for a in As]) if As else np.zeros(1)
    Bp = np.concatenate([np.ascontiguousarray(b, np.float64).ravel() for b in Bs]) if Bs else np.zeros(1)
    return n, r, a_off, b_off, Ap, Bp


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _p(t):
    return C.c_void_p(t.data_ptr())


def order_longest_first(ctx: bz.Context, n: np.ndarray):
    import torch
    order = torch.empty(max(len(n), 1), dtype=torch.int32, device="cuda")
    ctx.call("bzg_batched_order", len(n), n.ctypes.data_as(C.c_void_p), _p(order))
    return order


def gesv_batched(ctx: bz.Context, As, Bs, with_ipiv: bool = False, lpt: bool = True):
    """Plain FP64 batched LU solve. Returns (Xs, infos[, LUs, ipivs])."""
    import torch
    n, r, a_off, b_off, Ap, Bp = pack(As, Bs)
    p_off = np.zeros(len(As), np.int64)
    if len(As):
        p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    dn, dr, da, db, dp = (_dev(torch, x) for x in (n, r, a_off, b_off, p_off))
    dA, dB = _dev(torch, Ap), _dev(torch, Bp)
    dpiv = torch.zeros(max(int(n.sum()), 1), dtype=torch.int32, device="cuda")
    dinfo = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    order = order_longest_first(ctx, n) if lpt else None
    ctx.call("bzg_dgesv_batched", len(As), _p(dn), _p(dr), _p(da), _p(db), _p(dp), _p(dA), _p(dB),
             _p(dpiv), _p(dinfo), _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    Xh, Ah, ph, infoh = dB.cpu().numpy(), dA.cpu().numpy(), dpiv.cpu().numpy(), dinfo.cpu().numpy()
    Xs = [Xh[b_off[i]:b_off[i] + n[i] * r[i]].reshape(n[i], r[i]) for i in range(len(As))]
    if not with_ipiv:
        return Xs, infoh[:len(As)]
    LUs = [Ah[a_off[i]:a_off[i] + n[i] * n[i]].reshape(n[i], n[i]) for i in range(len(As))]
    ipivs = [ph[p_off[i]:p_off[i] + n[i]] for i in range(len(As))]
    return Xs, infoh[:len(As)], LUs, ipivs


def dsgesv_batched(ctx: bz.Context, As, Bs, tol: float = 0.0, max_iter: int = 30, lpt: bool = True):
    """Mixed-precision batched solve. Returns (Xs, iters, infos)."""
    import torch
    n, r, a_off, b_off, Ap, Bp = pack(As, Bs)
    dn, dr, da, db = (_dev(torch, x) for x in (n, r, a_off, b_off))
    dA, dB = _dev(torch, Ap), _dev(torch, Bp)
    dX = torch.full_like(dB, float("nan"))
    dit = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    dinfo = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    order = order_longest_first(ctx, n) if lpt else None
    ctx.call("bzg_dsgesv_batched", len(As), _p(dn), _p(dr), _p(da), _p(db), _p(dA), _p(dB), _p(dX),
             _p(dit), _p(dinfo), tol, max_iter, _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    Xh = dX.cpu().numpy()
    Xs = [Xh[b_off[i]:b_off[i] + n[i] * r[i]].reshape(n[i], r[i]) for i in range(len(As))]
    return Xs, dit.cpu().numpy()[:len(As)], dinfo.cpu().numpy()[:len(As)]


def getrf_batched(ctx: bz.Context, As, lpt: bool = True):
    """Batched factorisation (dgetrf per system). Returns (LUs, ipivs, infos)."""
    import torch
    n = np.array([a.shape[0] for a in As], np.int32)
    a_off = np.zeros(len(As), np.int64); p_off = np.zeros(len(As), np.int64)
    if len(As):
        a_off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    Ap = np.concatenate([np.ascontiguousarray(a, np.float64).ravel() for a in As])
    dn, da, dp, dA = (_dev(torch, x) for x in (n, a_off, p_off, Ap))
    dpiv = torch.zeros(max(int(n.sum()), 1), dtype=torch.int32, device="cuda")
    dinfo = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    order = order_longest_first(ctx, n) if lpt else None
    ctx.call("bzg_dgetrf_batched", len(As), _p(dn), _p(da), _p(dp), _p(dA), _p(dpiv), _p(dinfo),
             _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    Ah, ph = dA.cpu().numpy(), dpiv.cpu().numpy()
    LUs = [Ah[a_off[i]:a_off[i] + n[i] * n[i]].reshape(n[i], n[i]) for i in range(len(As))]
    ipivs = [ph[p_off[i]:p_off[i] + n[i]] for i in range(len(As))]
    return LUs, ipivs, dinfo.cpu().numpy()[:len(As)]


def det_batched(ctx: bz.Context, As, lpt: bool = True):
    """Batched determinants (det() per system, DMatDetExpr.h:384-424, by the LU rule of detNxN).
    Returns (dets, infos); each A_i is factored in a device copy."""
    import torch
    n = np.array([a.shape[0] for a in As], np.int32)
    a_off = np.zeros(len(As), np.int64); p_off = np.zeros(len(As), np.int64)
    if len(As):
        a_off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    Ap = np.concatenate([np.ascontiguousarray(a, np.float64).ravel() for a in As] + [np.zeros(1)])
    dn, da, dp, dA = (_dev(torch, x) for x in (n, a_off, p_off, Ap))
    dpiv = torch.zeros(max(int(n.sum()), 1), dtype=torch.int32, device="cuda")
    dinfo = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    ddet = torch.full((max(len(As), 1),), float("nan"), dtype=torch.float64, device="cuda")
    order = order_longest_first(ctx, n) if lpt and len(As) else None
    ctx.call("bzg_ddet_batched", len(As), _p(dn), _p(da), _p(dp), _p(dA), _p(dpiv), _p(dinfo), _p(ddet),
             _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    return ddet.cpu().numpy()[:len(As)], dinfo.cpu().numpy()[:len(As)]


def inv_batched(ctx: bz.Context, As, lpt: bool = True):
    """Batched inverses (invert() by LU per system, Inversion.h:3337-3347). Returns (results,
    infos): the inverse where info == 0, the LU factors where info > 0 (singular)."""
    import torch
    n = np.array([a.shape[0] for a in As], np.int32)
    a_off = np.zeros(len(As), np.int64); p_off = np.zeros(len(As), np.int64)
    if len(As):
        a_off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    Ap = np.concatenate([np.ascontiguousarray(a, np.float64).ravel() for a in As] + [np.zeros(1)])
    dn, da, dp, dA = (_dev(torch, x) for x in (n, a_off, p_off, Ap))
    dpiv = torch.zeros(max(int(n.sum()), 1), dtype=torch.int32, device="cuda")
    dinfo = torch.full((max(len(As), 1),), -99, dtype=torch.int32, device="cuda")
    order = order_longest_first(ctx, n) if lpt and len(As) else None
    ctx.call("bzg_dinv_batched", len(As), _p(dn), _p(da), _p(dp), _p(dA), _p(dpiv), _p(dinfo),
             _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    Ah = dA.cpu().numpy()
    return [Ah[a_off[i]:a_off[i] + n[i] * n[i]].reshape(n[i], n[i]) for i in range(len(As))], \
        dinfo.cpu().numpy()[:len(As)]


def getrs_batched(ctx: bz.Context, LUs, ipivs, Bs, lpt: bool = True):
    """Batched solve with given factors (dgetrs per system). Returns (Xs, infos)."""
    import torch
    n, r, a_off, b_off, Ap, Bp = pack(LUs, Bs)
    p_off = np.zeros(len(LUs), np.int64)
    if len(LUs):
        p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    Pp = np.concatenate([np.asarray(p, np.int32) for p in ipivs])
    dn, dr, da, db, dp, dA, dB, dpiv = (_dev(torch, x) for x in (n, r, a_off, b_off, p_off, Ap, Bp, Pp))
    dinfo = torch.full((max(len(LUs), 1),), -99, dtype=torch.int32, device="cuda")
    order = order_longest_first(ctx, n) if lpt else None
    ctx.call("bzg_dgetrs_batched", len(LUs), _p(dn), _p(dr), _p(da), _p(db), _p(dp), _p(dA), _p(dpiv), _p(dB),
             _p(dinfo), _p(order) if order is not None else C.c_void_p(0))
    ctx.sync()
    Xh = dB.cpu().numpy()
    return [Xh[b_off[i]:b_off[i] + n[i] * r[i]].reshape(n[i], r[i]) for i in range(len(LUs))], \
        dinfo.cpu().numpy()[:len(LUs)]


def lu_rowmajor(ctx: bz.Context, A):
    """lu(A, L, U, P) with the reference's row-major convention (blaze/math/dense/LU.h:98-140,
    222-330): the column-major getrf sees A^T, so A = L * U * P with L lower triangular (its
    diagonal from the factorisation), U unit upper triangular and P the column permutation
    P(i, permut[i]) = 1.  Square A only.  Runs bzg_dtranspose + bzg_dgetrf_batched on the device."""
    import torch
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    assert A.shape == (n, n)
    dA = _dev(torch, A)
    dT = torch.empty_like(dA)
    ctx.call("bzg_dtranspose", n, n, _p(dA), n, _p(dT), n)
    LUs, ipivs, infos = getrf_batched_device(ctx, dT, n)
    R = LUs.T  # the factors of A^T, read back in A's row-major layout
    L = np.tril(R)
    U = np.triu(R, 1) + np.eye(n)
    permut = np.arange(n)
    for i, p in enumerate(ipivs - 1):
        if p != i:
            permut[[i, p]] = permut[[p, i]]
    P = np.zeros((n, n))
    P[np.arange(n), permut] = 1.0
    return L, U, P, int(infos)


def getrf_batched_device(ctx: bz.Context, dA, n: int):
    """One n x n system already on the device (factored in place). Returns (LU, ipiv, info)."""
    import torch
    dn = _dev(torch, np.array([n], np.int32))
    dz = _dev(torch, np.zeros(1, np.int64))
    dpiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    dinfo = torch.full((1,), -99, dtype=torch.int32, device="cuda")
    ctx.call("bzg_dgetrf_batched", 1, _p(dn), _p(dz), _p(dz), _p(dA), _p(dpiv), _p(dinfo), C.c_void_p(0))
    ctx.sync()
    return dA.cpu().numpy().reshape(n, n), dpiv.cpu().numpy(), dinfo.cpu().numpy()[0]


# ---- single system of any order (bzg_dgetrf / bzg_dgesv / bzg_dgetri / bzg_ddet_lu) ---------
def _pad(torch, A, lda):
    """Device copy of the row-major A with leading dimension lda (>= n columns)."""
    n, m = A.shape
    buf = np.zeros((n, lda))
    buf[:, :m] = A
    return _dev(torch, buf)


def getrf(ctx: bz.Context, A, lda: int | None = None):
    """(LU, ipiv 1-based, info) of the n x n A through bzg_dgetrf (factors of P A = L U)."""
    import torch
    n = A.shape[0]
    lda = lda or max(n, 1)
    dA = _pad(torch, np.ascontiguousarray(A, np.float64), lda)
    dp = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    di = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.call("bzg_dgetrf", n, _p(dA), lda, _p(dp), _p(di))
    torch.cuda.synchronize()
    return dA.cpu().numpy()[:, :n].copy(), dp.cpu().numpy()[:n].copy(), int(di.cpu()[0])


def gesv(ctx: bz.Context, A, B):
    """(X, LU, ipiv, info) of A X = B through bzg_dgesv (B: n x nrhs)."""
    import torch
    n, r = B.shape
    dA = _dev(torch, np.ascontiguousarray(A, np.float64))
    dB = _dev(torch, np.ascontiguousarray(B, np.float64))
    dp = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    di = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.call("bzg_dgesv", n, r, _p(dA), max(n, 1), _p(dp), _p(dB), max(r, 1), _p(di))
    torch.cuda.synchronize()
    return dB.cpu().numpy(), dA.cpu().numpy(), dp.cpu().numpy()[:n].copy(), int(di.cpu()[0])


def inv(ctx: bz.Context, A):
    """(A^-1, info) through bzg_dgetrf + bzg_dgetri (A^-1 only when info == 0)."""
    import torch
    n = A.shape[0]
    dA = _dev(torch, np.ascontiguousarray(A, np.float64))
    dp = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    di = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.call("bzg_dgetrf", n, _p(dA), max(n, 1), _p(dp), _p(di))
    torch.cuda.synchronize()
    info = int(di.cpu()[0])
    if info == 0:
        ctx.call("bzg_dgetri", n, _p(dA), max(n, 1), _p(dp))
        torch.cuda.synchronize()
    return dA.cpu().numpy(), info


def det(ctx: bz.Context, A):
    """(det A, info) through bzg_dgetrf + bzg_ddet_lu."""
    import torch
    n = A.shape[0]
    dA = _dev(torch, np.ascontiguousarray(A, np.float64))
    dp = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    di = torch.zeros(1, dtype=torch.int32, device="cuda")
    dd = torch.zeros(1, dtype=torch.float64, device="cuda")
    ctx.call("bzg_dgetrf", n, _p(dA), max(n, 1), _p(dp), _p(di))
    ctx.call("bzg_ddet_lu", n, _p(dA), max(n, 1), _p(dp), _p(di), _p(dd))
    torch.cuda.synchronize()
    return float(dd.cpu()[0]), int(di.cpu()[0])
