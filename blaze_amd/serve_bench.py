"""Served mixed stream through the continuous-batching server (bzg_server_*).

Workload: Poisson arrivals at a swept rate; 50 % GEMM (M, N, K log-uniform in [32, 4096],
FP64), 30 % SpMV against 8 resident CSR matrices (7-point Laplacians and Zipf s = 1.5 power-law
matrices, n from 1e5 to 2e6), 20 % dense solves (n log-uniform in [8, 512], 1-8 right-hand
sides, 90 % A = R + n I, 10 % cond = 1e8).  Operands are resident in HBM (GEMM operands are
prefixes of two 4096 x 4096 random matrices, solve systems come from a pre-generated pool);
results are released as soon as the client has seen them.  Admission by a fixed HBM budget.
Reported per rate: throughput, p50/p99 latency (submit -> result complete on the device) per
class, achieved TFLOP/s (GEMM, solves) and GB/s (SpMV), worst error per class on a sample
(GEMM: relative to torch's FP64 matmul; SpMV: relative to torch's FP64 CSR product; solves:
normwise backward error), and the server's batching statistics.
"""
from __future__ import annotations

import ctypes as C
import sys
import threading
import time

import numpy as np

SEED = 9001


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


class Workload:
    def __init__(self, b, n_solve_pool=1500):
        import torch
        import blaze_amd as bz
        self.torch, self.b = torch, b
        rng = np.random.default_rng(SEED)
        self.rng = rng
        g = bz.gen()
        # 8 resident CSR matrices
        self.mats = []
        for nx in (47, 64, 100, 126):
            nnz = g.bzg_gen_laplace3d_nnz(C.c_int64(nx))
            rp = np.empty(nx ** 3 + 1, np.int32); ci = np.empty(nnz, np.int32); va = np.empty(nnz)
            g.bzg_gen_laplace3d(C.c_int64(nx), rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                va.ctypes.data_as(C.c_void_p))
            self.mats.append(("laplace", rp, ci, va))
        for n in (100_000, 300_000, 1_000_000, 2_000_000):
            rp = np.empty(n + 1, np.int32)
            g.bzg_gen_zipf_rowptr(C.c_uint64(SEED + n), C.c_int64(n), C.c_double(1.5), C.c_int32(2000),
                                  rp.ctypes.data_as(C.c_void_p))
            nnz = int(rp[-1])
            ci = np.empty(nnz, np.int32); va = np.empty(nnz)
            g.bzg_gen_zipf_fill(C.c_uint64(SEED + n), C.c_int64(n), C.c_int64(n), rp.ctypes.data_as(C.c_void_p),
                                ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p))
            self.mats.append(("zipf", rp, ci, va))
        self.x = [torch.from_numpy(rng.uniform(-1, 1, len(m[1]) - 1)).cuda() for m in self.mats]
        # GEMM operand pools
        self.Ag = torch.rand(4096 * 4096, dtype=torch.float64, device="cuda") * 2 - 1
        self.Bg = torch.rand(4096 * 4096, dtype=torch.float64, device="cuda") * 2 - 1
        # solve pool
        ns = np.exp(rng.uniform(np.log(8), np.log(512), n_solve_pool)).astype(np.int32)
        rs = rng.integers(1, 9, n_solve_pool).astype(np.int32)
        kind = (rng.random(n_solve_pool) < 0.10).astype(np.int8)
        a_off = np.zeros(n_solve_pool, np.int64); b_off = np.zeros(n_solve_pool, np.int64)
        a_off[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
        b_off[1:] = np.cumsum(ns.astype(np.int64) * rs)[:-1]
        A = np.empty(int(a_off[-1] + int(ns[-1]) ** 2)); B = np.empty(int(b_off[-1] + int(ns[-1]) * int(rs[-1])))
        g.bzg_gen_solve_batch(C.c_uint64(SEED), C.c_int64(n_solve_pool), ns.ctypes.data_as(C.c_void_p),
                              rs.ctypes.data_as(C.c_void_p), kind.ctypes.data_as(C.c_void_p),
                              a_off.ctypes.data_as(C.c_void_p), b_off.ctypes.data_as(C.c_void_p),
                              A.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p))
        self.sA, self.sB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        self.sys = list(zip(ns.tolist(), rs.tolist(), a_off.tolist(), b_off.tolist()))
        self._sparse = {}

    def make_requests(self, count):
        rng = self.rng
        u = rng.random(count)
        out = []
        for i in range(count):
            if u[i] < 0.5:
                m, n, k = np.exp(rng.uniform(np.log(32), np.log(4096), 3)).astype(int)
                out.append(("gemm", (int(m), int(n), int(k))))
            elif u[i] < 0.8:
                out.append(("spmv", int(rng.integers(0, len(self.mats)))))
            else:
                out.append(("solve", int(rng.integers(0, len(self.sys)))))
        return out

    def submit(self, srv, req):
        kind, p = req
        if kind == "gemm":
            m, n, k = p
            return srv.submit_gemm(m, n, k, self.Ag.data_ptr(), k, self.Bg.data_ptr(), n)
        if kind == "spmv":
            return srv.submit_spmv(p, self.x[p].data_ptr())
        n, r, ao, bo = self.sys[p]
        return srv.submit_solve(n, r, self.sA.data_ptr() + 8 * ao, self.sB.data_ptr() + 8 * bo)

    def work(self, req):
        """(flops, bytes) of one request."""
        kind, p = req
        if kind == "gemm":
            m, n, k = p
            return 2.0 * m * n * k, 8.0 * (m * k + k * n + m * n)
        if kind == "spmv":
            rp = self.mats[p][1]
            nnz, n = float(rp[-1]), len(rp) - 1
            return 2.0 * nnz, nnz * 12 + (n + 1) * 4 + 2.0 * n * 8
        n, r = self.sys[p][0], self.sys[p][1]
        return (2.0 / 3.0) * n ** 3 + 2.0 * n * n * r, 8.0 * n * n

    def check(self, req, res):
        """Error of one completed request (see module docstring)."""
        torch = self.torch
        from blaze_amd.serve import result_tensor
        kind, p = req
        if kind == "gemm":
            m, n, k = p
            got = result_tensor(res, (m, n))
            ref = self.Ag[: m * k].view(m, k) @ self.Bg[: k * n].view(k, n)
            return float((got - ref).norm() / ref.norm())
        if kind == "spmv":
            _, rp, ci, va = self.mats[p]
            got = result_tensor(res, (len(rp) - 1,))
            if p not in self._sparse:
                self._sparse[p] = torch.sparse_csr_tensor(
                    torch.from_numpy(rp.astype(np.int64)), torch.from_numpy(ci.astype(np.int64)),
                    torch.from_numpy(va), size=(len(rp) - 1, len(rp) - 1)).cuda()
            ref = self._sparse[p] @ self.x[p]
            return float((got - ref).norm() / ref.norm())
        n, r, ao, bo = self.sys[p]
        X = result_tensor(res, (n, r))
        A = self.sA[ao: ao + n * n].view(n, n)
        B = self.sB[bo: bo + n * r].view(n, r)
        R = B - A @ X
        den = A.abs().sum(1).max() * X.abs().max(0).values + B.abs().max(0).values
        return float((R.abs().max(0).values / den).max())


def run_rate(wl, srv, rate, count):
    """Poisson arrivals at `rate` req/s; returns the per-class summary."""
    reqs = wl.make_requests(count)
    gaps = wl.rng.exponential(1.0 / rate, count)
    ids = [None] * count
    results = [None] * count
    done = threading.Event()

    keep = {}

    def collector():
        for i in range(count):
            while ids[i] is None:
                time.sleep(20e-6)
            res = srv.wait(ids[i])
            results[i] = (res.t_submit_us, res.t_admit_us, res.t_done_us, res.status, res.info, res.iterations)
            if i % 97 == 0:
                keep[i] = res                     # accuracy sample: checked after the timed run
            else:
                srv.release(ids[i])
        done.set()

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(count):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        ids[i] = wl.submit(srv, reqs[i])
    done.wait()
    th.join()
    # device-side span: first submission to last completion (server clock)
    wall = (max(r[2] for r in results) - min(r[0] for r in results)) / 1e6
    for i, res in keep.items():
        results[i] = results[i] + (wl.check(reqs[i], res),)
        srv.release(ids[i])
    out = {"offered_req_s": rate, "requests": count, "span_s": wall, "throughput_req_s": count / wall}
    for cls in ("gemm", "spmv", "solve"):
        idx = [i for i in range(count) if reqs[i][0] == cls]
        if not idx:
            continue
        lat = np.array([(results[i][2] - results[i][0]) / 1e3 for i in idx])
        qwait = np.array([(results[i][1] - results[i][0]) / 1e3 for i in idx])
        fl = sum(wl.work(reqs[i])[0] for i in idx)
        by = sum(wl.work(reqs[i])[1] for i in idx)
        errs = [results[i][6] for i in idx if len(results[i]) > 6]
        d = {"count": len(idx), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "mean_queue_ms": float(qwait.mean()), "failed": int(sum(results[i][3] != 0 for i in idx)),
             "max_error_sampled": float(max(errs)) if errs else None, "sampled": len(errs)}
        if cls == "spmv":
            d["achieved_GB_s"] = by / wall / 1e9
        else:
            d["achieved_TFLOP_s"] = fl / wall / 1e12
        if cls == "solve":
            # orders <= 256: FP32 LU + refinement, a negative count = the FP64 fallback; orders
            # 257..512 run on the single-system FP64 LU by design (count -1)
            big = [i for i in idx if wl.sys[reqs[i][1]][0] > 256]
            d["fallback_fp64"] = int(sum(1 for i in idx if results[i][5] < 0 and wl.sys[reqs[i][1]][0] <= 256))
            d["fp64_lu_n_gt_256"] = len(big)
            if big:
                lat_big = np.array([(results[i][2] - results[i][0]) / 1e3 for i in big])
                d["p99_ms_n_le_256"] = float(np.percentile(
                    [(results[i][2] - results[i][0]) / 1e3 for i in idx if i not in set(big)] or [0.0], 99))
                d["p99_ms_n_gt_256"] = float(np.percentile(lat_big, 99))
            d["nonzero_info"] = int(sum(1 for i in idx if results[i][4] != 0))
        out[cls] = d
    return out


def run(b, rates=(1000, 4000, 16000), count=1500, budget_gb=200):
    from blaze_amd.serve import Server
    t0 = time.perf_counter()
    wl = Workload(b)
    _log(f"[serve] workload ready in {time.perf_counter() - t0:.1f}s")
    srv = Server(budget_bytes=budget_gb << 30, max_inflight=32, streams=4, solve_batch_max=4096)
    try:
        for _, rp, ci, va in wl.mats:
            srv.add_csr(rp, ci, va, len(rp) - 1)
        run_rate(wl, srv, 2000, 200)  # warm-up (workspaces, plans, code objects)
        out = {"workload": "Poisson mix 50% GEMM [32,4096]^3 fp64 / 30% SpMV on 8 resident CSR (n 1e5-2e6) / "
                           "20% solves n in [8,512] (FP32 LU + FP64 refinement)",
               "hbm_budget_GB": budget_gb, "streams": 4, "max_inflight": 32, "rates": []}
        for rate in rates:
            r = run_rate(wl, srv, rate, count)
            st = srv.stats()
            r["server"] = {k: st[k] for k in ("launches", "solve_batches", "max_solve_batch", "gemm_batches",
                                              "max_gemm_batch", "bytes_peak")}
            out["rates"].append(r)
            _log(f"[serve] rate {rate}/s: {r['throughput_req_s']:.0f} req/s; "
                 + "; ".join(f"{c} p50 {r[c]['p50_ms']:.2f} p99 {r[c]['p99_ms']:.2f} ms" for c in ("gemm", "spmv", "solve") if c in r))
        return out
    finally:
        srv.close()
