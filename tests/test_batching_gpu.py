"""GPU tests of the batching layer: grouped GEMM, the autotuners and HIP-graph capture.

* bzg_dgemm_grouped: every product of an uneven batch (empty, 1 x 1, odd leading dimensions,
  alpha / beta, k = 0, a rejected descriptor) is bit-identical to its own bzg_dgemm call, which
  is checked against the oracle's bound (Higham: 2 gamma_k |A||B|).
* Autotuners (GEMM tile shape, SpMV plan, solver grid): results are identical before and after
  tuning -- every configuration sums each element in the same order.
* HIP graphs: the grouped GEMM and the mixed-precision batched solve are captured once and
  replayed three times with new operand values written between replays (eager torch work in
  between); every replay equals the eager call on the same values.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

PROBLEM = np.dtype([("m", "<i8"), ("n", "<i8"), ("k", "<i8"), ("A", "<u8"), ("lda", "<i8"), ("B", "<u8"),
                    ("ldb", "<i8"), ("C", "<u8"), ("ldc", "<i8"), ("alpha", "<f8"), ("beta", "<f8")])


def _p(t):
    return C.c_void_p(t.data_ptr())


def _make_batch(torch, rng, shapes):
    ops = []
    for (m, n, k, pad, alpha, beta) in shapes:
        lda, ldb, ldc = max(k, 1) + pad, max(n, 1) + pad, max(n, 1) + pad
        A = torch.from_numpy(rng.uniform(-1, 1, (max(m, 1), lda))).cuda()
        B = torch.from_numpy(rng.uniform(-1, 1, (max(k, 1), ldb))).cuda()
        C0 = torch.from_numpy(rng.uniform(-1, 1, (max(m, 1), ldc))).cuda()
        ops.append(dict(m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, A=A, B=B, C0=C0, alpha=alpha, beta=beta))
    return ops


def _descs(torch, ops, Cs):
    d = np.zeros(len(ops), PROBLEM)
    for i, (o, Ct) in enumerate(zip(ops, Cs)):
        d[i] = (o["m"], o["n"], o["k"], o["A"].data_ptr(), o["lda"], o["B"].data_ptr(), o["ldb"], Ct.data_ptr(),
                o["ldc"], o["alpha"], o["beta"])
    return torch.from_numpy(d.view(np.uint8)).cuda()


SHAPES = [(0, 5, 5, 0, 1.0, 0.0), (1, 1, 1, 0, 1.0, 0.0), (5, 0, 3, 1, 1.0, 0.0), (7, 9, 0, 0, 1.0, 0.5),
          (65, 63, 130, 1, 1.0, 0.0), (200, 31, 17, 3, -2.0, 0.25), (128, 128, 128, 0, 1.0, 0.0),
          (333, 257, 64, 2, 0.5, 1.0), (17, 1000, 400, 0, 1.0, 0.0), (1000, 17, 33, 1, 1.0, -1.0)] + \
         [(int(m), int(n), int(k), 0, 1.0, 0.0) for m, n, k in np.random.default_rng(5).integers(1, 300, (30, 3))]


def test_grouped_gemm_matches_individual_calls():
    import torch
    rng = np.random.default_rng(80)
    ops = _make_batch(torch, rng, SHAPES)
    Cg = [o["C0"].clone() for o in ops]
    Ci = [o["C0"].clone() for o in ops]
    desc = _descs(torch, ops, Cg)
    # one rejected descriptor: ldc < n
    bad = np.zeros(1, PROBLEM)
    bad[0] = (4, 8, 4, ops[1]["A"].data_ptr(), 4, ops[1]["B"].data_ptr(), 8, Cg[1].data_ptr(), 2, 1.0, 0.0)
    desc = torch.cat([desc, torch.from_numpy(bad.view(np.uint8)).cuda()])
    status = torch.full((len(ops) + 1,), 7, dtype=torch.int32, device="cuda")
    ctx().call("bzg_dgemm_grouped", len(ops) + 1, _p(desc), _p(status))
    for o, Ct in zip(ops, Ci):
        ctx().call("bzg_dgemm", o["m"], o["n"], o["k"], o["alpha"], _p(o["A"]), o["lda"], _p(o["B"]), o["ldb"],
                   o["beta"], _p(Ct), o["ldc"])
    torch.cuda.synchronize()
    assert status.cpu().tolist() == [0] * len(ops) + [-1]
    for o, g, i in zip(ops, Cg, Ci):
        assert torch.equal(g, i), (o["m"], o["n"], o["k"])
        m, n, k = o["m"], o["n"], o["k"]
        if m and n:
            A = o["A"].cpu().numpy()[:m, :k]; B = o["B"].cpu().numpy()[:k, :n]
            C0 = o["C0"].cpu().numpy()[:m, :n]
            ref = o["alpha"] * (A @ B) + o["beta"] * C0
            bound = (2 * k + 2) * 2.0 ** -53 * (abs(o["alpha"]) * (np.abs(A) @ np.abs(B)) + abs(o["beta"]) * np.abs(C0))
            assert np.all(np.abs(g.cpu().numpy()[:m, :n] - ref) <= bound + 1e-300)


def test_gemm_autotune_changes_speed_not_results():
    import torch
    import blaze_amd as bz
    rng = np.random.default_rng(81)
    m, n, k = 1500, 700, 333
    A = torch.from_numpy(rng.uniform(-1, 1, (m, k))).cuda(); B = torch.from_numpy(rng.uniform(-1, 1, (k, n))).cuda()
    C1 = torch.empty(m, n, dtype=torch.float64, device="cuda"); C2 = torch.empty_like(C1)
    ctx().call("bzg_dgemm", m, n, k, 1.0, _p(A), k, _p(B), n, 0.0, _p(C1), n)
    shape, ms = C.c_int32(-1), C.c_float(0)
    ctx().call("bzg_gemm_autotune", 0, m, n, k, C.byref(shape), C.byref(ms))
    assert shape.value in (0, 1, 2) and ms.value > 0
    ctx().call("bzg_dgemm", m, n, k, 1.0, _p(A), k, _p(B), n, 0.0, _p(C2), n)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    bz.check("bzg_gemm_tuning_clear", bz.lib().bzg_gemm_tuning_clear())


def test_spmv_autotune_keeps_results():
    import torch
    import blaze_amd as bz
    from tests.test_gemv_spmv_gpu import Plan
    rng = np.random.default_rng(82)
    n = 20000
    lens = np.minimum(rng.zipf(1.5, n), 500)
    rp = np.zeros(n + 1, np.int32); rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, int(c), replace=False)) for c in lens]).astype(np.int32)
    va = rng.standard_normal(int(rp[-1])); x = rng.uniform(-1, 1, n)
    plan = Plan(rp, n)
    d = [torch.from_numpy(a).cuda() for a in (rp, ci, va, x)]
    y1 = torch.empty(n, dtype=torch.float64, device="cuda"); y2 = torch.empty_like(y1)
    ctx().call("bzg_dcsrmv", plan.h, *[_p(t) for t in d], _p(y1))
    ms = C.c_float(0)
    bz.check("bzg_csr_plan_autotune", bz.lib().bzg_csr_plan_autotune(ctx().handle, plan.h, 0, _p(d[0]), _p(d[1]),
                                                                     _p(d[2]), _p(d[3]), _p(y2), C.byref(ms)))
    ctx().call("bzg_dcsrmv", plan.h, *[_p(t) for t in d], _p(y2))
    torch.cuda.synchronize()
    assert ms.value > 0
    assert torch.equal(y1, y2), plan.describe()
    ref = oracle.csrmv(rp, ci, va, x)
    short = np.diff(rp) <= 32
    assert np.array_equal(y2.cpu().numpy()[short], ref[short])


def test_spmv_every_short_row_kernel_is_bit_identical():
    # LDS row blocks (2048 / 4096 nonzeros, cached / non-temporal loads) and lane groups: every short row is the
    # oracle's sequential sum bit for bit (SVecDVecInnerExpr.h:127-135), in fp64 and fp32, with
    # runs of short rows broken by long ones and empty rows in between
    import torch
    import blaze_amd as bz
    from tests.test_gemv_spmv_gpu import Plan
    rng = np.random.default_rng(83)
    n = 30000
    lens = rng.integers(0, 9, n)
    lens[rng.choice(n, 60, replace=False)] = rng.integers(33, 700, 60)
    rp = np.zeros(n + 1, np.int32); rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, int(c), replace=False)) for c in lens]).astype(np.int32)
    short = lens <= 32
    for dt, fn in ((np.float64, "bzg_dcsrmv"), (np.float32, "bzg_scsrmv")):
        va = rng.standard_normal(int(rp[-1])).astype(dt); x = rng.uniform(-1, 1, n).astype(dt)
        ref = oracle.csrmv(rp, ci, va, x)
        plan = Plan(rp, n)
        d = [torch.from_numpy(a).cuda() for a in (rp, ci, va, x)]
        first = None
        for cfg in ((0, 1, 0), (0, 2, 0), (0, 1, 1), (4, 1, 4), (1, 2, 8), (16, 1, 2)):
            bz.check("tune", bz.lib().bzg_csr_plan_tune(plan.h, *cfg))
            y = torch.full((n,), 7.0, dtype=d[2].dtype, device="cuda")
            ctx().call(fn, plan.h, *[_p(t) for t in d], _p(y))
            got = y.cpu().numpy()
            assert np.array_equal(got[short].view(np.uint8), ref[short].view(np.uint8)), (dt, cfg)
            if first is None:
                first = got
            assert np.array_equal(got.view(np.uint8), first.view(np.uint8)), (dt, cfg)


def _solve_batch(torch, rng, sizes, nrhs=2):
    from blaze_amd import solve
    As = [rng.uniform(-1, 1, (n, n)) + n * np.eye(n) for n in sizes]
    Bs = [rng.uniform(-1, 1, (n, nrhs)) for n in sizes]
    n, r, a_off, b_off, Ap, Bp = solve.pack(As, Bs)
    dev = [torch.from_numpy(v).cuda() for v in (n, r, a_off, b_off, Ap, Bp)]
    return As, Bs, dev


def test_solver_autotune_keeps_results():
    import torch
    rng = np.random.default_rng(83)
    sizes = list(rng.integers(4, 300, 300))
    As, Bs, (dn, dr, da, db, dA, dB) = _solve_batch(torch, rng, sizes)
    count = len(sizes)
    X1 = torch.empty_like(dB); X2 = torch.empty_like(dB)
    it = torch.empty(count, dtype=torch.int32, device="cuda"); inf = torch.empty_like(it)
    args = (count, _p(dn), _p(dr), _p(da), _p(db), _p(dA), _p(dB))
    ctx().call("bzg_dsgesv_batched", *args, _p(X1), _p(it), _p(inf), 0.0, 30, C.c_void_p(0))
    A0 = dA.clone(); B0 = dB.clone()
    for mixed in (1, 0):
        wg, ms = C.c_int32(0), C.c_float(0)
        ctx().call("bzg_solver_autotune", mixed, count, _p(dn), _p(dr), _p(da), _p(db), _p(dA), dA.numel(),
                   _p(dB), dB.numel(), C.c_void_p(0), C.byref(wg), C.byref(ms))
        assert wg.value in (1, 2, 3, 4, 6, 8) and ms.value > 0
    ctx().call("bzg_dsgesv_batched", *args, _p(X2), _p(it), _p(inf), 0.0, 30, C.c_void_p(0))
    torch.cuda.synchronize()
    assert torch.equal(dA, A0) and torch.equal(dB, B0)   # the tuner left the operands alone
    assert torch.equal(X1, X2)


def test_hip_graph_capture_and_replay():
    import torch
    import blaze_amd as bz
    rng = np.random.default_rng(84)
    s = torch.cuda.Stream()
    gctx = bz.Context(stream=s.cuda_stream)
    # grouped GEMM operands
    ops = _make_batch(torch, rng, [(m, n, k, 0, 1.0, 0.0) for m, n, k in rng.integers(1, 200, (12, 3))])
    Cg = [o["C0"].clone() for o in ops]
    desc = _descs(torch, ops, Cg)
    # batched mixed solve operands
    sizes = list(rng.integers(4, 300, 64))
    As, Bs, (dn, dr, da, db, dA, dB) = _solve_batch(torch, rng, sizes)
    count = len(sizes)
    X = torch.empty_like(dB); it = torch.empty(count, dtype=torch.int32, device="cuda"); inf = torch.empty_like(it)

    def work():
        gctx.call("bzg_dgemm_grouped", len(ops), _p(desc), C.c_void_p(0))
        gctx.call("bzg_dsgesv_batched", count, _p(dn), _p(dr), _p(da), _p(db), _p(dA), _p(dB), _p(X), _p(it),
                  _p(inf), 0.0, 30, C.c_void_p(0))

    with torch.cuda.stream(s):
        work()                     # eager warm-up: workspaces reach their size before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        work()
    torch.cuda.synchronize()
    for rep in range(3):
        # new operand values (eager torch work between replays)
        for o in ops:
            o["A"].uniform_(-1, 1); o["B"].uniform_(-1, 1)
        dB.uniform_(-1, 1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got_c = [c.clone() for c in Cg]; got_x = X.clone(); got_it = it.clone()
        with torch.cuda.stream(s):
            work()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(got_c, Cg)), rep
        assert torch.equal(got_x, X) and torch.equal(got_it, it), rep
        assert (inf == 0).all()
    gctx.close()
