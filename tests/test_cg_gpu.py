"""GPU parity: conjugate gradients on a resident CSR matrix (bzg_dcg), the loop of the
reference's CG benchmark (blazemark/src/blaze/CG.cpp:114-135) on its 5-point matrix.

* Against the oracle's restatement of the same loop (oracle.cg: the reference's sequential
  row sums for the products, numpy dot products): x within 1e-9 relative after 40 iterations
  (the two dot-product orders differ by rounding, which CG carries along but does not amplify
  at this conditioning), and r.r within 1e-6 relative.
* Convergence: with b random, ||b - A x|| <= 1e-9 ||b|| after enough iterations.
* Deterministic: two calls give bit-identical x; a HIP-graph capture replays bit-identically.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _p(t):
    return C.c_void_p(t.data_ptr())


class Problem:
    def __init__(self, torch, N, seed, zero_b):
        from tests.test_gemv_spmv_gpu import Plan
        self.rp, self.ci, self.va = oracle.laplace2d(N)
        self.n = N * N
        rng = np.random.default_rng(seed)
        self.b = np.zeros(self.n) if zero_b else rng.uniform(-1, 1, self.n)
        self.x0 = rng.uniform(-1, 1, self.n)
        self.plan = Plan(self.rp, self.n)
        self.d = [torch.from_numpy(a).cuda() for a in (self.rp, self.ci, self.va, self.b)]

    def run(self, torch, iterations, cx=None):
        x = torch.from_numpy(self.x0.copy()).cuda()
        rr = C.c_double(0)
        (cx or ctx()).call("bzg_dcg", self.plan.h, self.n, *[_p(t) for t in self.d], _p(x), iterations,
                           C.byref(rr))
        return x, rr.value


@pytest.mark.parametrize("N,zero_b", [(16, True), (48, False), (100, False)])
def test_cg_matches_oracle_loop(N, zero_b):
    import torch
    pb = Problem(torch, N, 90 + N, zero_b)
    x, rr = pb.run(torch, 40)
    xo, rro = oracle.cg(pb.rp, pb.ci, pb.va, pb.b, pb.x0, 40)
    got = x.cpu().numpy()
    assert np.abs(got - xo).max() <= 1e-9 * max(np.abs(xo).max(), 1e-300), N
    assert abs(rr - rro) <= 1e-6 * rro, (rr, rro)
    # r.r reported is the true residual of x (up to the recurrence's rounding)
    res = pb.b - oracle.csrmv(pb.rp, pb.ci, pb.va, got)
    assert abs(res @ res - rr) <= 1e-6 * max(rr, 1e-300) + 1e-24


def test_cg_converges_and_is_deterministic():
    import torch
    N = 64
    pb = Problem(torch, N, 7, False)
    x1, rr1 = pb.run(torch, 600)
    x2, rr2 = pb.run(torch, 600)
    assert torch.equal(x1, x2) and rr1 == rr2
    res = pb.b - oracle.csrmv(pb.rp, pb.ci, pb.va, x1.cpu().numpy())
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(pb.b)


def test_cg_graph_capture_replays_identically():
    import torch
    import blaze_amd as bz
    pb = Problem(torch, 40, 11, False)
    s = torch.cuda.Stream()
    gctx = bz.Context(stream=s.cuda_stream)
    x = torch.from_numpy(pb.x0.copy()).cuda()
    x0 = x.clone()

    def work():
        gctx.call("bzg_dcg", pb.plan.h, pb.n, *[_p(t) for t in pb.d], _p(x), 15, None)
    with torch.cuda.stream(s):
        work()                      # eager: workspace reaches its size
    torch.cuda.synchronize()
    eager = x.clone()
    g = torch.cuda.CUDAGraph()
    x.copy_(x0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        work()
    for rep in range(3):
        x.copy_(x0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, eager), rep
    gctx.close()


def test_cg_general_plan_path_matches_oracle():
    # the same loop on a plan retuned to lane groups (vector-CSR SpMV kernels instead of LDS
    # row blocks) meets the same bar against the oracle
    import torch
    import blaze_amd as bz
    pb = Problem(torch, 48, 91, False)
    bz.check("tune", bz.lib().bzg_csr_plan_tune(pb.plan.h, 4, 1, 4))
    assert "vector-csr" in pb.plan.describe()
    x, rr = pb.run(torch, 40)
    xo, rro = oracle.cg(pb.rp, pb.ci, pb.va, pb.b, pb.x0, 40)
    assert np.abs(x.cpu().numpy() - xo).max() <= 1e-9 * max(np.abs(xo).max(), 1e-300)
    assert abs(rr - rro) <= 1e-6 * rro


def test_cg_sliced_plan_is_bit_identical():
    # the plan's sliced-ELL copy serves the loop's SpMV once packed: same row sums, same x
    import torch
    import blaze_amd as bz
    pb = Problem(torch, 64, 7, False)
    x1, rr1 = pb.run(torch, 30)
    bz.check("pack", bz.lib().bzg_csr_plan_pack_sell(ctx().handle, pb.plan.h, _p(pb.d[0]), _p(pb.d[1]), _p(pb.d[2]), 0))
    x2, rr2 = pb.run(torch, 30)
    assert torch.equal(x1, x2) and rr1 == rr2
