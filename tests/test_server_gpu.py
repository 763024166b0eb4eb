"""GPU tests of the continuous-batching server (bzg_server_* / bzg_submit_*).

* Every result is bit-identical alone (submitted, completed and released one at a time) and
  in a mix (all submitted at once from two client threads, solves of order <= 256 batched
  together, larger ones through the single-system FP64 LU, work spread over four streams) --
  the kernels compute each request independently.
* Accuracy per class: GEMM within Higham's bound of the FP64 product, SpMV rows of up to 32
  nonzeros bit-identical to the oracle's storage-order sum, solves to backward error <= 1e-12.
* Admission: results hold HBM budget until released; a request that does not fit waits, a
  request larger than the whole budget is refused with BZG_ERR_OUT_OF_MEMORY.
"""
import threading
import time

import numpy as np
import pytest

import oracle
from tests.gpu_util import have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _csr_laplace(nx):
    import ctypes as C
    import blaze_amd as bz
    g = bz.gen()
    nnz = g.bzg_gen_laplace3d_nnz(C.c_int64(nx))
    rp = np.zeros(nx ** 3 + 1, np.int32); ci = np.zeros(nnz, np.int32); va = np.zeros(nnz)
    g.bzg_gen_laplace3d(C.c_int64(nx), rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                        va.ctypes.data_as(C.c_void_p))
    return rp, ci, va


def _csr_powerlaw(rng, n, maxlen):
    lens = np.minimum(rng.zipf(1.5, n), maxlen).astype(np.int64)
    rp = np.zeros(n + 1, np.int32); rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, int(k), replace=False)) for k in lens]).astype(np.int32)
    va = rng.standard_normal(int(rp[-1]))
    return rp, ci, va


def _workload(torch, rng):
    mats = [_csr_laplace(12), _csr_powerlaw(rng, 3000, 600)]
    reqs = []
    for (m, n, k) in [(32, 32, 32), (100, 37, 250), (257, 129, 64), (512, 512, 512), (33, 1000, 7)]:
        A = torch.from_numpy(rng.uniform(-1, 1, (m, k))).cuda()
        B = torch.from_numpy(rng.uniform(-1, 1, (k, n))).cuda()
        reqs.append(("gemm", (m, n, k), A, B))
    for h, (rp, ci, va) in enumerate(mats):
        x = torch.from_numpy(rng.uniform(-1, 1, len(rp) - 1)).cuda()
        reqs.append(("spmv", h, x, None))
        reqs.append(("spmv", h, torch.from_numpy(rng.uniform(-1, 1, len(rp) - 1)).cuda(), None))
    for n in (8, 17, 64, 200, 256, 300, 512):
        r = int(rng.integers(1, 9))
        A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        B = rng.uniform(-1, 1, (n, r))
        reqs.append(("solve", (n, r), torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()))
    return mats, reqs


def _submit(srv, req):
    kind, shape, a, b = req
    if kind == "gemm":
        m, n, k = shape
        return srv.submit_gemm(m, n, k, a.data_ptr(), k, b.data_ptr(), n)
    if kind == "spmv":
        return srv.submit_spmv(shape, a.data_ptr())
    n, r = shape
    return srv.submit_solve(n, r, a.data_ptr(), b.data_ptr())


def _shape_of(req, mats):
    kind, shape, a, b = req
    if kind == "gemm":
        return (shape[0], shape[1])
    if kind == "spmv":
        return (len(mats[shape][0]) - 1,)
    return shape


def test_results_identical_alone_and_in_mix():
    import torch
    from blaze_amd.serve import Server, result_tensor
    rng = np.random.default_rng(70)
    mats, reqs = _workload(torch, rng)
    srv = Server(budget_bytes=8 << 30, max_inflight=16, streams=4)
    try:
        handles = [srv.add_csr(rp, ci, va, len(rp) - 1) for rp, ci, va in mats]
        assert handles == [0, 1]
        # alone: one request at a time
        alone = []
        for q in reqs:
            rid = _submit(srv, q)
            res = srv.wait(rid)
            assert res.status == 0
            alone.append((result_tensor(res, _shape_of(q, mats)).cpu().numpy(), res.info, res.iterations))
            srv.release(rid)
        # mix: everything at once from two client threads, in interleaved order
        ids = [None] * len(reqs)
        order = list(rng.permutation(len(reqs)))

        def client(part):
            for i in part:
                ids[i] = _submit(srv, reqs[i])

        th = [threading.Thread(target=client, args=(order[j::2],)) for j in range(2)]
        srv.hold(True)  # the whole burst is queued before anything is admitted
        for t in th:
            t.start()
        for t in th:
            t.join()
        srv.hold(False)
        mixed = []
        for i, q in enumerate(reqs):
            res = srv.wait(ids[i])
            assert res.status == 0
            mixed.append((result_tensor(res, _shape_of(q, mats)).cpu().numpy(), res.info, res.iterations))
        st = srv.stats()
        for i in range(len(reqs)):
            srv.release(ids[i])
        assert st["completed"] == 2 * len(reqs)
        for (a, ia, ta), (b, ib, tb), q in zip(alone, mixed, reqs):
            assert np.array_equal(a, b), q[0]
            assert ia == ib and ta == tb
        # accuracy per class
        for (got, info, it), (kind, shape, a, b) in zip(alone, reqs):
            if kind == "gemm":
                A, B = a.cpu().numpy(), b.cpu().numpy()
                bound = 2 * shape[2] * 2.0 ** -53 * (np.abs(A) @ np.abs(B))
                assert np.all(np.abs(got - A @ B) <= bound + 1e-300)
            elif kind == "spmv":
                rp, ci, va = mats[shape]
                ref = oracle.csrmv(rp, ci, va, a.cpu().numpy())
                short = np.diff(rp) <= 32
                assert np.array_equal(got[short], ref[short])
                assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
            else:
                # orders <= 256: FP32 LU + refinement (it >= 0); above: the FP64 LU alone (-1)
                assert info == 0 and (it >= 0 if shape[0] <= 256 else it == -1), (shape, it)
                be = oracle.backward_error(a.cpu().numpy(), b.cpu().numpy(), got)
                assert be <= 1e-12, (shape, be)
    finally:
        srv.close()
    # the held burst's solves of order <= 256 formed one batch (the alone pass ran them one by
    # one); the larger ones ran on their own
    n_batched = sum(1 for q in reqs if q[0] == "solve" and q[1][0] <= 256)
    assert n_batched >= 2 and st["max_solve_batch"] == n_batched


def test_budget_admission_and_release():
    import torch
    import blaze_amd as bz
    from blaze_amd.serve import Server
    n = 512  # each result: 512 x 512 doubles = 2 MiB
    A = torch.rand(n, n, dtype=torch.float64, device="cuda")
    srv = Server(budget_bytes=5 << 20, max_inflight=8, streams=2)
    try:
        r1 = srv.submit_gemm(n, n, n, A.data_ptr(), n, A.data_ptr(), n)
        r2 = srv.submit_gemm(n, n, n, A.data_ptr(), n, A.data_ptr(), n)
        assert srv.wait(r1).status == 0 and srv.wait(r2).status == 0
        r3 = srv.submit_gemm(n, n, n, A.data_ptr(), n, A.data_ptr(), n)
        time.sleep(0.2)
        assert not srv.poll(r3)              # 2 unreleased results hold 4 MiB of 5 MiB
        assert srv.stats()["bytes_in_use"] == 4 << 20
        srv.release(r1)
        assert srv.wait(r3).status == 0      # admitted once r1's bytes came back
        srv.release(r2)
        srv.release(r3)
        assert srv.stats()["bytes_in_use"] == 0
        with pytest.raises(bz.BlazeGpuError) as e:
            srv.submit_gemm(2048, 2048, 8, A.data_ptr(), 8, A.data_ptr(), 2048)  # 32 MiB result
        assert e.value.status == -4
    finally:
        srv.close()


def test_add_csr_failure_returns_budget():
    # a matrix the plan builder rejects (rowptr[0] != 0) frees its device copies and gives its
    # budget back: the same budget then still holds a valid matrix of almost its whole size
    import numpy as np
    import blaze_amd as bz
    from blaze_amd.serve import Server
    n = 20000
    rp = np.arange(n + 1, dtype=np.int32)                 # one entry per row: 4 (n+1) + 12 n bytes
    ci = np.arange(n, dtype=np.int32); va = np.ones(n)
    need = 4 * (n + 1) + 12 * n
    # all rows short: the server keeps the sliced-ELL copy (64-row slices of width 1 and the
    # slice offsets); both forms are reserved while the copy is made, then only the copy
    slices = (n + 63) // 64
    sell = slices * 64 * 12 + (slices + 1) * 8
    srv = Server(budget_bytes=need + sell + 1024, max_inflight=4, streams=1)
    try:
        bad = rp.copy(); bad[0] = 1
        for _ in range(3):
            with pytest.raises(bz.BlazeGpuError) as e:
                srv.add_csr(bad, ci, va, n)
            assert e.value.status == -1
            assert srv.stats()["bytes_in_use"] == 0
        h = srv.add_csr(rp, ci, va, n)
        assert h == 0 and srv.stats()["bytes_in_use"] == sell
    finally:
        srv.close()
