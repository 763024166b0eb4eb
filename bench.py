#!/usr/bin/env python3
"""Benchmark of the MI355X multiplication & solve path (one JSON line on stdout).

Headline step: one FP64-accurate GEMM C = A * B at M = N = K = 8192 (A, B uniform[-1,1)),
computed from residue slices on the matrix cores (bzg_dgemm_fp64acc[_i8]); value = whole-job
TFLOP/s counted as 2*M*N*K FP64-equivalent flops per step, inputs resident in HBM.

Beside it (fields of the same JSON line):
  native_fp64      the plain FP64 MFMA GEMM on the same inputs (the comparator) and the ratio;
  fp64acc_*        both slice variants; errors are relative Frobenius norms over 16 whole rows
                   of C against an extended-precision (x87 long double) product;
  fp64acc_8192_by_sigma   the same product for entries u 2^e, e = round(N(0, sigma)), sigma in
                   {0, 4, 10}: rate of the automatic moduli count, ratio to native, error;
  rel_frob_err_by_sigma_512x512x8192   full-matrix errors per sigma on a smaller M, N;
  fp64acc_16384x1024x16384   the rectangular case;
  plain_gemm_512   the plain FP64 GEMM at 512^3 on mt19937_64 inputs (the reference's
                   CPU-runnable case), rate and FP64 MFMA fraction;
  solve            100,000 systems, n ~ U{4..256}, nrhs ~ U{1..8}, 90% A = R + n I, 10%
                   Q1 diag(logspace(0,-8,n)) Q2^T with Haar-random Q (cond 1e8): plain FP64
                   batched LU vs FP32 LU + FP64 refinement, systems/s, p50/p99 batch latency,
                   worst backward error (target 1e-12), fallback count;
  spmv             CSR SpMV GB/s (algorithmic bytes) for the 128^3 7-point Laplacian (fp64,
                   fp32) and the n = 1e6 power-law matrix, rotating copies past the 256 MB MALL;
  gemv, cg, serve  dense GEMV, the CG loop, the continuous-batching server (p50/p99 per class);
  complex          complex<double> / complex<float> GEMM at 4096^3 (A B and A^H B) and
                   complex<double> GEMV at 16384^2, with Blaze's complex GEMM on the host;
  roofline         dominant kernel (the residue-slice GEMM) from HIP events on its stream;
  cpu_baseline     the reference's own GEMM (Blaze, compiled from its headers into oracle/_ref)
                   on the host cores at the headline size; every other leg carries its own
                   `cpu_baseline` (Blaze GEMM 512^3, SpMV, GEMV, solve() with LAPACK, complex
                   GEMM).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "FP64-accurate GEMM TFLOP/s vs native FP64; batched-solve p50/p99 at 1e-12"
SEED = 128753984
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "int8": 5000.0, "hbm": 8000.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def vp(t):
    return C.c_void_p(t.data_ptr())


class Bench:
    def __init__(self, args):
        import torch
        import blaze_amd as bz
        self.torch, self.bz, self.args = torch, bz, args
        torch.cuda.set_device(0)
        self.stream = torch.cuda.Stream()
        torch.cuda.set_stream(self.stream)
        self.ctx = bz.Context(stream=self.stream.cuda_stream)
        self.gen = bz.gen()

    # -- helpers ------------------------------------------------------------------------
    def sync(self):
        self.torch.cuda.synchronize()

    def timed(self, fn, steps, warmup):
        for _ in range(warmup):
            fn()
        self.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        self.sync()
        return (time.perf_counter() - t0) / steps

    def uniform(self, shape, stream_id):
        a = np.empty(shape, np.float64)
        self.gen.bzg_gen_uniform_ctr(C.c_uint64(SEED), C.c_uint64(stream_id), C.c_int64(a.size),
                                     a.ctypes.data_as(C.c_void_p))
        return a

    def graded(self, shape, stream_id, sigma):
        a = np.empty(shape, np.float64)
        self.gen.bzg_gen_graded(C.c_uint64(SEED), C.c_uint64(stream_id), C.c_int64(a.size),
                                C.c_double(sigma), a.ctypes.data_as(C.c_void_p))
        return a

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda")

    @staticmethod
    def ref_entries(A, B, rows, cols):
        """Extended-precision (long double) C[rows, cols]."""
        Al = A[rows].astype(np.longdouble)
        Bl = B[:, cols].T.astype(np.longdouble)
        return np.sum(Al * Bl, axis=1)

    @staticmethod
    def rel_frob(got, ref):
        ref = ref.astype(np.longdouble)
        return float(np.sqrt(np.sum((got.astype(np.longdouble) - ref) ** 2) / np.sum(ref ** 2)))

    # -- GEMM legs -------------------------------------------------------------------------
    def gemm_case(self, m, n, k, A, B, steps, warmup, variants, whole_rows=16):
        """Native FP64 and each split variant: rate, and the relative Frobenius error over
        `whole_rows` complete rows of C (every column) against an extended-precision product."""
        torch = self.torch
        dA, dB = self.dev(A), self.dev(B)
        dC = torch.empty((m, n), dtype=torch.float64, device="cuda")
        fl = 2.0 * m * n * k
        rows = np.sort(np.random.default_rng(7).choice(m, size=min(whole_rows, m), replace=False))
        ref = self.ref_rows(A, B, rows)
        drows = torch.from_numpy(rows).to("cuda")
        err = lambda: self.rel_frob(dC.index_select(0, drows).cpu().numpy(), ref)
        sample = f"{len(rows)} whole rows of C ({len(rows) * n} entries) against an x87 long-double product"
        out = {}
        t = self.timed(lambda: self.ctx.call("bzg_dgemm", m, n, k, 1.0, vp(dA), k, vp(dB), n, 0.0, vp(dC), n),
                       steps, warmup)
        out["native_fp64"] = {"tflops": fl / t / 1e12, "ms": t * 1e3, "rel_frob_err": err(),
                              "err_sample": sample}
        for name, fn in variants:
            t = self.timed(lambda: self.ctx.call(fn, m, n, k, vp(dA), k, vp(dB), n, vp(dC), n, 0),
                           steps, warmup)
            nm, bits, est = C.c_int32(), C.c_int32(), C.c_double()
            self.ctx.call("bzg_fp64acc_last_plan", C.byref(nm), C.byref(bits), C.byref(est))
            out[name] = {"tflops": fl / t / 1e12, "ms": t * 1e3, "rel_frob_err": err(), "err_sample": sample,
                         "vs_native": (fl / t / 1e12) / out["native_fp64"]["tflops"],
                         "moduli": nm.value, "operand_bits": bits.value, "estimated_err": est.value}
        del dA, dB, dC
        return out

    @staticmethod
    def ref_rows(A, B, rows):
        """Extended-precision (x87 long double, 64-bit significand) rows of A @ B, one thread per
        block of rows."""
        from concurrent.futures import ThreadPoolExecutor
        Bl = B.astype(np.longdouble)
        parts = [p for p in np.array_split(rows, min(16, len(rows))) if len(p)]
        with ThreadPoolExecutor(len(parts)) as ex:
            outs = list(ex.map(lambda r: A[r].astype(np.longdouble) @ Bl, parts))
        return np.concatenate(outs)

    @staticmethod
    def ref_full(A, B):
        """Extended-precision (x87 long double, 64-bit significand) A @ B, rows split over threads."""
        from concurrent.futures import ThreadPoolExecutor
        Bl = B.astype(np.longdouble)
        parts = np.array_split(np.arange(A.shape[0]), 16)
        with ThreadPoolExecutor(16) as ex:
            outs = list(ex.map(lambda r: A[r].astype(np.longdouble) @ Bl, parts))
        return np.concatenate(outs)

    def sigma_case(self, A, B, variants):
        """Full relative Frobenius error of native FP64 and each split variant (automatic count)."""
        torch = self.torch
        m, k = A.shape
        n = B.shape[1]
        ref = self.ref_full(A, B)
        den = np.sqrt(np.sum(ref ** 2))
        dA, dB = self.dev(A), self.dev(B)
        dC = torch.empty((m, n), dtype=torch.float64, device="cuda")
        self.ctx.call("bzg_dgemm", m, n, k, 1.0, vp(dA), k, vp(dB), n, 0.0, vp(dC), n)
        out = {"native_fp64": float(np.sqrt(np.sum((dC.cpu().numpy() - ref) ** 2)) / den)}
        for name, fn in variants:
            self.ctx.call(fn, m, n, k, vp(dA), k, vp(dB), n, vp(dC), n, 0)
            got = dC.cpu().numpy()
            nm, bits, est = C.c_int32(), C.c_int32(), C.c_double()
            self.ctx.call("bzg_fp64acc_last_plan", C.byref(nm), C.byref(bits), C.byref(est))
            out[name] = {"rel_frob_err": float(np.sqrt(np.sum((got - ref) ** 2)) / den), "moduli": nm.value,
                         "operand_bits": bits.value, "estimated_err": est.value}
        return out

    def roofline_fp64acc(self, m, n, k, A, B, fn, slices, reps=5):
        torch = self.torch
        dA, dB = self.dev(A), self.dev(B)
        dC = torch.empty((m, n), dtype=torch.float64, device="cuda")
        self.ctx.call(fn, m, n, k, vp(dA), k, vp(dB), n, vp(dC), n, 0)
        self.sync()
        self.ctx.profile(True)
        self.ctx.profile_read("fp64acc_gemm")
        for _ in range(reps):
            self.ctx.call(fn, m, n, k, vp(dA), k, vp(dB), n, vp(dC), n, 0)
        ms, launches = self.ctx.profile_read("fp64acc_gemm")
        tot_ms, _ = self.ctx.profile_read("fp64acc_total")
        self.ctx.profile(False)
        avg_s = ms / launches / 1e3
        ops = slices * 2.0 * m * n * k
        peak = PEAK["int8"] if fn.endswith("i8") else PEAK["bf16"]
        traffic = None
        tf = ROOT / "profiles" / "traffic.json"
        if tf.exists():
            traffic = json.loads(tf.read_text()).get(f"{fn}_{m}x{n}x{k}")
        return {"kernel": "residue_gemm_kernel (" + ("i8" if fn.endswith("i8") else "bf16") + ")",
                "bound": "mfma", "achieved": ops / avg_s / 1e12, "peak": peak,
                "unit": "TOP/s" if fn.endswith("i8") else "TFLOP/s",
                "frac": ops / avg_s / 1e12 / peak, "traffic": traffic,
                "avg_launch_ms": avg_s * 1e3, "launches": launches,
                "share_of_step": (ms / (tot_ms if tot_ms else ms)),
                "ops_per_launch": ops}

    # -- batched solves --------------------------------------------------------------------
    def solve_leg(self, count, steps, warmup):
        torch = self.torch
        rng = np.random.default_rng(SEED)
        n = rng.integers(4, 257, count).astype(np.int32)
        r = rng.integers(1, 9, count).astype(np.int32)
        kind = (rng.random(count) < 0.10).astype(np.int8)
        a_off = np.zeros(count, np.int64); b_off = np.zeros(count, np.int64)
        a_off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        b_off[1:] = np.cumsum(n.astype(np.int64) * r)[:-1]
        na = int(a_off[-1] + n[-1].astype(np.int64) ** 2); nb = int(b_off[-1] + n[-1] * r[-1])
        t0 = time.perf_counter()
        A = np.empty(na); B = np.empty(nb)
        self.gen.bzg_gen_solve_batch(C.c_uint64(SEED), C.c_int64(count), n.ctypes.data_as(C.c_void_p),
                                     r.ctypes.data_as(C.c_void_p), kind.ctypes.data_as(C.c_void_p),
                                     a_off.ctypes.data_as(C.c_void_p), b_off.ctypes.data_as(C.c_void_p),
                                     A.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p))
        log(f"[solve] generated {count} systems ({na * 8 / 1e9:.1f} GB) in {time.perf_counter() - t0:.1f}s")
        dn, dr, da, db = self.dev(n), self.dev(r), self.dev(a_off), self.dev(b_off)
        dA0, dB0 = self.dev(A), self.dev(B)
        self.solve_inputs = (n, r, a_off, b_off, A, B) if "cpu" in self.args.legs else None
        del A
        order = torch.empty(count, dtype=torch.int32, device="cuda")
        self.ctx.call("bzg_batched_order", count, n.ctypes.data_as(C.c_void_p), vp(order))
        flops = float(np.sum((2.0 / 3.0) * n.astype(np.float64) ** 3 - 0.5 * n.astype(np.float64) ** 2
                             + (5.0 / 6.0) * n + r * (2.0 * n.astype(np.float64) ** 2 - n)))
        out = {"systems": count, "flops_getrf_getrs": flops}
        # project autotuner: persistent-grid size (workgroups per CU) of each solver on this batch
        tuned = {}
        for mixed in (0, 1):
            wg, tms = C.c_int32(0), C.c_float(0)
            self.ctx.call("bzg_solver_autotune", mixed, count, vp(dn), vp(dr), vp(da), vp(db), vp(dA0), na,
                          vp(dB0), nb, vp(order), C.byref(wg), C.byref(tms))
            tuned["mixed" if mixed else "fp64"] = {"workgroups_per_cu": wg.value, "ms": tms.value}
        out["autotuned_grid"] = tuned
        log(f"[solve] autotuned grid: {tuned}")

        # plain FP64 (A, B overwritten: restore from the pristine copies before each run)
        dA, dB = torch.empty_like(dA0), torch.empty_like(dB0)
        dinfo = torch.empty(count, dtype=torch.int32, device="cuda")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        lat = []
        for it in range(warmup + steps):
            dA.copy_(dA0); dB.copy_(dB0)
            ev[0].record()
            self.ctx.call("bzg_dgesv_batched", count, vp(dn), vp(dr), vp(da), vp(db), C.c_void_p(0),
                          vp(dA), vp(dB), C.c_void_p(0), vp(dinfo), vp(order))
            ev[1].record()
            self.sync()
            if it >= warmup:
                lat.append(ev[0].elapsed_time(ev[1]) / 1e3)
        X64 = dB.clone()
        out["fp64_lu"] = self._lat_stats(lat, count, flops)
        out["fp64_lu"]["info_nonzero"] = int((dinfo != 0).sum().item())
        del dA

        # mixed precision, two stopping tests: the 1e-12 target (tol = 1e-12 / sqrt(256), so
        # ||r|| <= 1e-12 ||A|| ||x|| for every n <= 256) and LAPACK dsgesv's own (eps = 2^-53)
        dX = torch.empty_like(dB0)
        dit = torch.empty(count, dtype=torch.int32, device="cuda")
        for key, tol in (("mixed_fp32_lu_fp64_ir", 1e-12 / 16), ("mixed_lapack_eps", 0.0)):
            lat = []
            for it in range(warmup + steps):
                ev[0].record()
                self.ctx.call("bzg_dsgesv_batched", count, vp(dn), vp(dr), vp(da), vp(db), vp(dA0), vp(dB0),
                              vp(dX), vp(dit), vp(dinfo), tol, 30, vp(order))
                ev[1].record()
                self.sync()
                if it >= warmup:
                    lat.append(ev[0].elapsed_time(ev[1]) / 1e3)
            rec = self._lat_stats(lat, count, flops)
            its = dit.cpu().numpy()
            rec.update({
                "tol": tol, "fallback_fp64": int((its < 0).sum()), "mean_iterations": float(its[its >= 0].mean()),
                "max_iterations": int(its[its >= 0].max()) if (its >= 0).any() else None,
                "speedup_vs_fp64": rec["systems_per_s"] / out["fp64_lu"]["systems_per_s"]})
            # accuracy: normwise backward error on a sample, on the GPU in FP64 (grouped by n)
            rec["max_backward_error"] = self._backward_errors(n, r, a_off, b_off, dA0, dB0, dX)
            out[key] = rec
        out["speedup_mixed_vs_fp64"] = out["mixed_fp32_lu_fp64_ir"]["speedup_vs_fp64"]
        out["fp64_lu"]["max_backward_error"] = self._backward_errors(n, r, a_off, b_off, dA0, dB0, X64)
        # det() on every system (batched getrf + detNxN's rule) and inv() on the first 10^4
        # (getrf, then getrs against identity columns, eight per pass); A restored before each call
        p_off = np.zeros(count, np.int64); p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
        dp = self.dev(p_off)
        dpiv = torch.empty(int(n.sum()), dtype=torch.int32, device="cuda")
        ddet = torch.empty(count, dtype=torch.float64, device="cuda")
        dA = torch.empty_like(dA0)
        getrf_flops = float(np.sum((2.0 / 3.0) * n.astype(np.float64) ** 3 - 0.5 * n.astype(np.float64) ** 2
                                   + (5.0 / 6.0) * n))
        for key, cnt, fn in (("det_batched", count, "bzg_ddet_batched"), ("inv_batched_1e4", 10_000, "bzg_dinv_batched")):
            lat = []
            for it in range(warmup + max(2, steps // 2)):
                dA.copy_(dA0)
                ev[0].record()
                if fn == "bzg_ddet_batched":
                    self.ctx.call(fn, cnt, vp(dn), vp(da), vp(dp), vp(dA), vp(dpiv), vp(dinfo), vp(ddet), vp(order))
                else:
                    self.ctx.call(fn, cnt, vp(dn), vp(da), vp(dp), vp(dA), vp(dpiv), vp(dinfo), C.c_void_p(0))
                ev[1].record()
                self.sync()
                if it >= warmup:
                    lat.append(ev[0].elapsed_time(ev[1]) / 1e3)
            nn = n[:cnt].astype(np.float64)
            fl = getrf_flops if cnt == count else float(np.sum((2.0 / 3.0) * nn ** 3 - 0.5 * nn ** 2 + (5.0 / 6.0) * nn
                                                               + (4.0 / 3.0) * nn ** 3))
            rec = {"systems": cnt, "p50_ms": float(np.median(lat) * 1e3), "gflops": fl / float(np.median(lat)) / 1e9,
                   "flops": ("getrf" if cnt == count else "getrf + getri-equivalent (4/3 n^3)"),
                   "info_nonzero": int((dinfo[:cnt] != 0).sum().item())}
            if cnt == count:
                rec["nonfinite_det"] = int((~torch.isfinite(ddet)).sum().item())
            out[key] = rec
        del dA
        # both roofline fractions: getrf + getrs flops against the published FP64 (FP32 for the
        # mixed solver's factorisation) peak, and the algorithmic HBM bytes against 8 TB/s.
        # FP64 LU: A read and its factors written (16 n^2 B) plus B read and X written
        # (16 n nrhs B); mixed: A read in FP64, its FP32 copy written and read back (8 + 4 + 4
        # n^2 B), plus B read and X written.
        n2 = n.astype(np.float64) ** 2
        nr = n.astype(np.float64) * r
        bytes_fp64 = float(np.sum(16 * n2 + 16 * nr))
        bytes_mixed = float(np.sum(16 * n2 + 16 * nr))
        for key, peak, by in (("fp64_lu", PEAK["fp64"], bytes_fp64), ("mixed_fp32_lu_fp64_ir", PEAK["fp32"], bytes_mixed),
                              ("mixed_lapack_eps", PEAK["fp32"], bytes_mixed)):
            sec = out[key]["p50_ms"] / 1e3
            out[key]["roofline"] = {"compute_frac": flops / sec / 1e12 / peak, "compute_peak_tflops": peak,
                                    "hbm_frac": by / sec / 1e9 / PEAK["hbm"], "bytes": by}
        return out

    def _lat_stats(self, lat, count, flops):
        lat = np.array(lat)
        return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "systems_per_s": count / float(np.median(lat)), "gflops": flops / float(np.median(lat)) / 1e9,
                "runs": len(lat)}

    def _backward_errors(self, n, r, a_off, b_off, dA, dB, dX, sample=10_000):
        """Worst normwise backward error over a random sample of systems (FP64 on the GPU)."""
        torch = self.torch
        worst = 0.0
        pick = np.random.default_rng(3).choice(len(n), size=min(sample, len(n)), replace=False)
        for nn in np.unique(n[pick]):
            idx = pick[n[pick] == nn]
            for rr in np.unique(r[idx]):
                sel = idx[r[idx] == rr]
                ai = torch.from_numpy((a_off[sel][:, None] + np.arange(nn * nn)[None, :]).ravel()).to("cuda")
                bi = torch.from_numpy((b_off[sel][:, None] + np.arange(nn * rr)[None, :]).ravel()).to("cuda")
                Am = dA[ai].view(len(sel), nn, nn)
                Bm = dB[bi].view(len(sel), nn, rr)
                Xm = dX[bi].view(len(sel), nn, rr)
                R = Bm - torch.bmm(Am, Xm)
                an = Am.abs().sum(dim=2).amax(dim=1)
                be = R.abs().amax(dim=1) / (an[:, None] * Xm.abs().amax(dim=1) + Bm.abs().amax(dim=1))
                worst = max(worst, float(be.max().item()))
        return worst

    # -- SpMV --------------------------------------------------------------------------------
    def spmv_leg(self, steps):
        torch, bz = self.torch, self.bz
        out = {}
        g = self.gen
        nx = 128
        nr = nx ** 3
        nnz = g.bzg_gen_laplace3d_nnz(C.c_int64(nx))
        rp = np.zeros(nr + 1, np.int32); ci = np.zeros(nnz, np.int32); va = np.zeros(nnz)
        g.bzg_gen_laplace3d(C.c_int64(nx), rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                            va.ctypes.data_as(C.c_void_p))
        x = self.uniform(nr, 11)
        self.spmv_inputs = {"laplace128_fp64": (rp, ci, va, x, nr)}
        out["laplace128_fp64"] = self._spmv_run(rp, ci, va, x, nr, "bzg_dcsrmv", steps)
        out["laplace128_fp32"] = self._spmv_run(rp, ci, va.astype(np.float32), x.astype(np.float32), nr,
                                                "bzg_scsrmv", steps)
        # the same products on the plan's sliced-ELL copy (bzg_csr_plan_pack_sell, packed once per
        # matrix copy), bit-identical to the CSR kernels
        out["laplace128_fp64_sell"] = self._sell_run(rp, ci, va, x, nr, 0, steps)
        out["laplace128_fp32_sell"] = self._sell_run(rp, ci, va.astype(np.float32), x.astype(np.float32), nr, 1,
                                                     steps)
        n = 1_000_000
        rp = np.zeros(n + 1, np.int32)
        g.bzg_gen_zipf_rowptr(C.c_uint64(SEED), C.c_int64(n), C.c_double(1.5), C.c_int32(2000),
                              rp.ctypes.data_as(C.c_void_p))
        nnz = int(rp[-1])
        ci = np.zeros(nnz, np.int32); va = np.zeros(nnz)
        g.bzg_gen_zipf_fill(C.c_uint64(SEED), C.c_int64(n), C.c_int64(n), rp.ctypes.data_as(C.c_void_p),
                            ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p))
        x = self.uniform(n, 12)
        self.spmv_inputs["zipf1e6_s1.5_fp64"] = (rp, ci, va, x, n)
        out["zipf1e6_s1.5_fp64"] = self._spmv_run(rp, ci, va, x, n, "bzg_dcsrmv", steps)
        out["zipf1e6_s1.5_fp64"]["mean_row"] = nnz / n
        # y = trans(S) * x (TSMatDVecMultExpr) on the same matrix: the host pack gives the CSR of
        # trans(S) (rows of Poisson length ~34, columns ascending), its plan sums from +0
        trp = np.zeros(n + 1, np.int32); tci = np.zeros(nnz, np.int32); tva = np.zeros(nnz)
        t0 = time.perf_counter()
        bz.check("pack", bz.lib().bzg_csc_to_csr_host(n, n, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                                      va.ctypes.data_as(C.c_void_p), 8, trp.ctypes.data_as(C.c_void_p),
                                                      tci.ctypes.data_as(C.c_void_p), tva.ctypes.data_as(C.c_void_p)))
        pack_s = time.perf_counter() - t0
        out["zipf1e6_s1.5_trans_fp64"] = self._spmv_run(trp, tci, tva, x, n, "bzg_dcsrmv", steps, plus_zero=True)
        out["zipf1e6_s1.5_trans_fp64"]["host_pack_s"] = pack_s
        return out

    def _spmv_run(self, rp, ci, va, x, ncols, fn, steps, plus_zero=False):
        torch, bz = self.torch, self.bz
        n = len(rp) - 1
        nnz = int(rp[-1])
        esz = va.dtype.itemsize
        by = nnz * (esz + 4) + (n + 1) * 4 + 2 * n * esz
        copies = max(2, int(np.ceil(600e6 / by)))  # working set > 2x the 256 MB MALL
        mats = [[self.dev(a) for a in (rp, ci, va)] for _ in range(copies)]
        dx = self.dev(x)
        dy = torch.empty(n, dtype=dx.dtype, device="cuda")
        plan = C.c_void_p()
        bz.check("plan", bz.lib().bzg_csr_plan_create(self.ctx.handle, n, ncols, rp.ctypes.data_as(C.c_void_p),
                                                      C.byref(plan)))
        if plus_zero:
            bz.check("sum start", bz.lib().bzg_csr_plan_set_sum_start(plan, 1))
        it = [0]

        def one():
            mm = mats[it[0] % copies]; it[0] += 1
            self.ctx.call(fn, plan, vp(mm[0]), vp(mm[1]), vp(mm[2]), vp(dx), vp(dy))
        one(); self.sync()
        self.ctx.profile(True); self.ctx.profile_read("csrmv")
        t = self.timed(one, steps, 2)
        ms, launches = self.ctx.profile_read("csrmv")
        self.ctx.profile(False)
        # back-to-back device throughput: 8 rounds over the rotated copies captured in one HIP
        # graph and replayed (no host launch gap between calls), timed with events
        s = torch.cuda.Stream()
        gctx = bz.Context(stream=s.cuda_stream)
        calls = 8 * copies

        def many():
            for i in range(calls):
                mm = mats[i % copies]
                gctx.call(fn, plan, vp(mm[0]), vp(mm[1]), vp(mm[2]), vp(dx), vp(dy))
        with torch.cuda.stream(s):
            many()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            many()
        g.replay()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 5
        ev[0].record()
        for _ in range(reps):
            g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        gt = ev[0].elapsed_time(ev[1]) / 1e3 / (reps * calls)
        del g
        gctx.close()
        buf = C.create_string_buffer(256)
        bz.lib().bzg_csr_plan_describe(plan, buf, 256)
        bz.lib().bzg_csr_plan_destroy(plan)
        kt = ms / launches / 1e3
        return {"gbps": by / t / 1e9, "kernel_gbps": by / kt / 1e9,
                "graph_gbps": by / gt / 1e9, "frac_hbm": by / gt / 1e9 / PEAK["hbm"],
                "us": t * 1e6, "kernel_us": kt * 1e6, "graph_us": gt * 1e6, "bytes": by, "nnz": nnz,
                "copies_rotated": copies, "plan": buf.value.decode(),
                "timing": ("us/gbps: host-launched calls (ctypes); kernel_*: HIP events around each call; "
                           "graph_* and frac_hbm: calls replayed back to back from a HIP graph")}

    def _sell_run(self, rp, ci, va, x, ncols, dtype, steps):
        """y = S * x from sliced-ELL copies (one plan and copy per rotated matrix copy), calls
        replayed back to back from a HIP graph; `bytes` = the CSR product's algorithmic bytes (as
        the CSR legs), `sell_bytes` = what the sliced layout reads (padded slots, no row pointers)."""
        torch, bz = self.torch, self.bz
        n = len(rp) - 1
        nnz = int(rp[-1])
        esz = va.dtype.itemsize
        by = nnz * (esz + 4) + (n + 1) * 4 + 2 * n * esz
        lens = np.diff(rp)
        pad = np.zeros(((n + 63) // 64) * 64, lens.dtype)
        pad[:n] = lens
        slots = int(pad.reshape(-1, 64).max(axis=1).sum()) * 64
        sell_by = slots * (esz + 4) + 2 * n * esz
        copies = max(2, int(np.ceil(600e6 / by)))
        fn = ("bzg_dsellmv", "bzg_ssellmv", "bzg_isellmv")[dtype]
        plans = []
        for _ in range(copies):
            d = [self.dev(a) for a in (rp, ci, va)]
            pl = C.c_void_p()
            bz.check("plan", bz.lib().bzg_csr_plan_create(self.ctx.handle, n, ncols, rp.ctypes.data_as(C.c_void_p),
                                                          C.byref(pl)))
            bz.check("pack", bz.lib().bzg_csr_plan_pack_sell(self.ctx.handle, pl, vp(d[0]), vp(d[1]), vp(d[2]), dtype))
            del d  # the plan holds its own sliced copy
            plans.append(pl)
        dx = self.dev(x)
        dy = torch.empty(n, dtype=dx.dtype, device="cuda")
        self.ctx.call(fn, plans[0], vp(dx), vp(dy))
        self.sync()
        s = torch.cuda.Stream()
        gctx = bz.Context(stream=s.cuda_stream)
        calls = 8 * copies

        def many():
            for i in range(calls):
                gctx.call(fn, plans[i % copies], vp(dx), vp(dy))
        with torch.cuda.stream(s):
            many()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            many()
        g.replay()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 5
        ev[0].record()
        for _ in range(reps):
            g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        gt = ev[0].elapsed_time(ev[1]) / 1e3 / (reps * calls)
        del g
        gctx.close()
        for pl in plans:
            bz.lib().bzg_csr_plan_destroy(pl)
        return {"graph_us": gt * 1e6, "graph_gbps": by / gt / 1e9, "frac_hbm": by / gt / 1e9 / PEAK["hbm"],
                "bytes": by, "sell_bytes": sell_by, "sell_gbps": sell_by / gt / 1e9, "nnz": nnz,
                "copies_rotated": copies,
                "timing": "calls replayed back to back from a HIP graph over the rotated copies, timed with events"}

    # -- dense GEMV, y = A * x and y = trans(A) * x (DMatDVecMultExpr / TDMatDVecMultExpr) ------
    def gemv_leg(self, m=16384, n=16384, reps=5):
        """Row-major m x n A; calls replayed back to back from a HIP graph (two copies of A, so
        the 2 GiB FP64 working set is well beyond the 256 MB MALL), timed with events.  Bytes per
        call: m n + m + n elements (A once, x, y)."""
        torch, bz = self.torch, self.bz
        out = {}
        for dt, pre in ((torch.float64, "d"), (torch.float32, "s")):
            mats = [torch.empty(m, n, dtype=dt, device="cuda").uniform_(-1, 1) for _ in range(2)]
            xs = {False: torch.empty(n, dtype=dt, device="cuda").uniform_(-1, 1),
                  True: torch.empty(m, dtype=dt, device="cuda").uniform_(-1, 1)}
            ys = {False: torch.empty(m, dtype=dt, device="cuda"), True: torch.empty(n, dtype=dt, device="cuda")}
            by = (m * n + m + n) * mats[0].element_size()
            for tr in (False, True):
                fn = f"bzg_{pre}gemv" + ("_t" if tr else "")
                s = torch.cuda.Stream()
                gctx = bz.Context(stream=s.cuda_stream)
                calls = 16

                def many():
                    for i in range(calls):
                        gctx.call(fn, m, n, 1.0, vp(mats[i % 2]), n, vp(xs[tr]), 0.0, vp(ys[tr]))
                with torch.cuda.stream(s):
                    many()
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=s):
                    many()
                g.replay()
                torch.cuda.synchronize()
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record()
                for _ in range(reps):
                    g.replay()
                ev[1].record()
                torch.cuda.synchronize()
                us = ev[0].elapsed_time(ev[1]) * 1e3 / (reps * calls)
                del g
                gctx.close()
                key = f"{'trans_' if tr else ''}{pre}gemv_{m}x{n}"
                out[key] = {"us": us, "gbps": by / us / 1e3, "frac_hbm": by / us / 1e3 / PEAK["hbm"], "bytes": by}
            del mats
        out["timing"] = "calls replayed back to back from a HIP graph, events around the replays"
        return out

    # -- conjugate gradients (the reference's CG benchmark loop, blazemark CG.cpp) -------------
    def complex_leg(self, n=4096, mv=16384, reps=5):
        """complex<double> / complex<float> products at n^3 (8 n^3 real flops each; op pairs N.N and
        H.N) and complex<double> GEMV at mv x mv (16 bytes per element of A read once), timed with
        events on the library's stream; uniform[-1,1) real and imaginary parts."""
        torch, bz = self.torch, self.bz
        ctx = bz.Context(stream=torch.cuda.current_stream().cuda_stream)
        out = {}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

        def timed(f):
            f()
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(reps):
                f()
            ev[1].record()
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / reps / 1e3
        for dt, fn, peak in ((torch.float64, "bzg_zgemm_ex", PEAK["fp64"]), (torch.float32, "bzg_cgemm_ex", PEAK["fp32"])):
            A = torch.empty(2 * n * n, dtype=dt, device="cuda").uniform_(-1, 1)
            B = torch.empty(2 * n * n, dtype=dt, device="cuda").uniform_(-1, 1)
            Cm = torch.empty(2 * n * n, dtype=dt, device="cuda")
            one = np.array([1.0, 0.0], np.float64 if dt == torch.float64 else np.float32)
            zero = np.zeros(2, one.dtype)
            for opa, tag in ((0, "nn"), (2, "hn")):
                t = timed(lambda: ctx.call(fn, opa, 0, n, n, n, one.ctypes.data_as(C.c_void_p), vp(A), n, vp(B), n,
                                           zero.ctypes.data_as(C.c_void_p), vp(Cm), n))
                fl = 8.0 * n ** 3
                out[f"{fn[4]}gemm_{tag}_{n}"] = {"ms": t * 1e3, "tflops": fl / t / 1e12, "frac_peak": fl / t / 1e12 / peak}
            if dt == torch.float64:
                self.zgemm_inputs = (A.cpu().numpy(), B.cpu().numpy(), out[f"zgemm_nn_{n}"]["tflops"])
            del A, B, Cm
        A = torch.empty(2 * mv * mv, dtype=torch.float64, device="cuda").uniform_(-1, 1)
        x = torch.empty(2 * mv, dtype=torch.float64, device="cuda").uniform_(-1, 1)
        y = torch.empty(2 * mv, dtype=torch.float64, device="cuda")
        one = np.array([1.0, 0.0]); zero = np.zeros(2)
        by = 16 * (mv * mv + 2 * mv)
        for op, tag in ((0, "n"), (2, "h")):
            t = timed(lambda: ctx.call("bzg_zgemv_ex", op, mv, mv, one.ctypes.data_as(C.c_void_p), vp(A), mv, vp(x),
                                       zero.ctypes.data_as(C.c_void_p), vp(y)))
            out[f"zgemv_{tag}_{mv}"] = {"us": t * 1e6, "gbps": by / t / 1e9, "frac_hbm": by / t / 1e9 / PEAK["hbm"]}
        del A
        ctx.close()
        return out

    def cg_leg(self, N=2048, iterations=50, reps=5):
        """5-point matrix N^2 x N^2, b = 0, random start (as blazemark); the whole bzg_dcg call
        (SpMV + 3 streaming kernels per iteration) captured in a HIP graph and replayed."""
        torch, bz = self.torch, self.bz
        n = N * N
        # rows in the order top, left, centre, right, bottom (CG.cpp:86-103)
        i, j = np.divmod(np.arange(n, dtype=np.int64), N)
        cnt = 1 + (i > 0) + (j > 0) + (j < N - 1) + (i < N - 1)
        rp = np.zeros(n + 1, np.int64); rp[1:] = np.cumsum(cnt)
        ci = np.empty(int(rp[-1]), np.int32); va = np.empty(int(rp[-1]))
        pos = rp[:-1].copy()
        for cond, col, v in ((i > 0, (i - 1) * N + j, -1.0), (j > 0, i * N + j - 1, -1.0),
                             (np.ones(n, bool), i * N + j, 4.0), (j < N - 1, i * N + j + 1, -1.0),
                             (i < N - 1, (i + 1) * N + j, -1.0)):
            ci[pos[cond]] = col[cond]; va[pos[cond]] = v; pos[cond] += 1
        rp = rp.astype(np.int32)
        nnz = int(rp[-1])
        d = [self.dev(a) for a in (rp, ci, va)]
        b = torch.zeros(n, dtype=torch.float64, device="cuda")
        x0 = self.dev(self.uniform(n, 31))
        x = x0.clone()
        plan = C.c_void_p()
        bz.check("plan", bz.lib().bzg_csr_plan_create(self.ctx.handle, n, n, rp.ctypes.data_as(C.c_void_p),
                                                      C.byref(plan)))
        # every row holds <= 5 entries: the plan's sliced-ELL copy serves the SpMV inside the loop
        bz.check("pack", bz.lib().bzg_csr_plan_pack_sell(self.ctx.handle, plan, vp(d[0]), vp(d[1]), vp(d[2]), 0))
        s = torch.cuda.Stream()
        gctx = bz.Context(stream=s.cuda_stream)

        def call():
            gctx.call("bzg_dcg", plan, n, vp(d[0]), vp(d[1]), vp(d[2]), vp(b), vp(x), iterations, None)
        with torch.cuda.stream(s):
            call()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            call()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        x.copy_(x0); g.replay(); torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        t_call = ev[0].elapsed_time(ev[1]) / 1e3 / reps
        del g
        gctx.close()
        bz.lib().bzg_csr_plan_destroy(plan)
        spmv_bytes = nnz * 12 + (n + 1) * 4 + 2 * n * 8
        it_bytes = spmv_bytes + 88 * n  # + d.h (16 n), x/r update (48 n), d update (24 n)
        t_it = t_call / (iterations + 1)
        return {"N": N, "unknowns": n, "nnz": nnz, "iterations": iterations, "ms_per_call": t_call * 1e3,
                "us_per_iteration": t_it * 1e6, "bytes_per_iteration": it_bytes,
                "gbps": it_bytes / t_it / 1e9, "frac_hbm": it_bytes / t_it / 1e9 / PEAK["hbm"],
                "spmv": "sliced ELL (bzg_csr_plan_pack_sell)",
                "timing": "one bzg_dcg call captured in a HIP graph, replayed; per iteration = call / (iterations + 1)"}

    # -- plain GEMM at 512^3 (the reference's CPU-runnable GEMM case) --------------------------
    def gemm512_leg(self, nn=512, calls=200, reps=5):
        """C = A * B, M = N = K = 512, A then B from one mt19937_64 stream (u = (x >> 11) 2^-53 * 2
        - 1, row-major); calls replayed back to back from a HIP graph, timed with events.  Rate
        in 2 N^3 flops (and blazemark's 2 N^3 - N^2) against the FP64 MFMA peak."""
        torch, bz = self.torch, self.bz
        A = np.empty((nn, nn)); B = np.empty((nn, nn))
        self.gen.bzg_gen_uniform_mt(C.c_uint64(SEED), C.c_int64(A.size), A.ctypes.data_as(C.c_void_p),
                                    C.c_int64(B.size), B.ctypes.data_as(C.c_void_p))
        self.gemm512_inputs = (A, B)
        dA, dB = self.dev(A), self.dev(B)
        dC = torch.empty((nn, nn), dtype=torch.float64, device="cuda")
        s = torch.cuda.Stream()
        gctx = bz.Context(stream=s.cuda_stream)

        def many():
            for _ in range(calls):
                gctx.call("bzg_dgemm", nn, nn, nn, 1.0, vp(dA), nn, vp(dB), nn, 0.0, vp(dC), nn)
        with torch.cuda.stream(s):
            many()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            many()
        g.replay()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        us = ev[0].elapsed_time(ev[1]) * 1e3 / (reps * calls)
        del g
        gctx.close()
        # single host-launched calls (ctypes launch included), for comparison
        t1 = self.timed(lambda: self.ctx.call("bzg_dgemm", nn, nn, nn, 1.0, vp(dA), nn, vp(dB), nn, 0.0, vp(dC), nn),
                        20, 3)
        fl = 2.0 * nn ** 3
        return {"workload": f"C = A*B, DynamicMatrix<double,rowMajor> {nn}^3, A,B uniform[-1,1) from mt19937_64",
                "us": us, "tflops": fl / us / 1e6, "frac_fp64_mfma_peak": fl / us / 1e6 / PEAK["fp64"],
                "tflops_blazemark_convention": (2.0 * nn ** 3 - nn ** 2) / us / 1e6,
                "us_host_launched": t1 * 1e6,
                "timing": "calls replayed back to back from a HIP graph, events around the replays"}

    # -- CPU baseline -------------------------------------------------------------------------
    def cpu_cores(self):
        cores = os.cpu_count() or 1
        try:
            cores = min(cores, len(os.sched_getaffinity(0)))
        except AttributeError:
            pass
        return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))

    def cpu_baseline(self, rec):
        """The reference's own kernels (Blaze compiled from its headers into oracle/_ref; OpenMP
        over the host cores; blazemark protocol: one warm-up, then the best of a few timed
        assignments) on the operands the GPU legs used.  Returns the headline object (GEMM at the
        headline size) and fills each GPU leg's record with its own `cpu_baseline`."""
        import oracle
        cores = self.cpu_cores()
        if not oracle.have_ref():
            nn = 512
            A = self.uniform((nn, nn), 21); B = self.uniform((nn, nn), 22)
            t0 = time.perf_counter()
            oracle.gemm(A, B)
            t = time.perf_counter() - t0
            return {"value": 2.0 * nn ** 3 / t / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
                    "sample": f"oracle C restatement of Blaze's GEMM, {nn}^3, 1 core"}
        L = oracle.ref()
        L.ref_set_threads(C.c_int(cores))
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        proto = f"OpenMP {cores} threads, best of the timed runs after one warm-up (blazemark protocol)"
        # headline: C = A * B at the headline size, the GEMM leg's uniform[-1,1) operands
        A, B = getattr(self, "gemm_inputs", (None, None))
        nn = self.args.size
        if A is None:
            A = self.uniform((nn, nn), 1); B = self.uniform((nn, nn), 2)
        log(f"[cpu] Blaze C = A * B {nn}^3 on {cores} cores")
        t = L.ref_time_dgemm(C.c_int64(nn), C.c_int64(nn), C.c_int64(nn), ptr(A), ptr(B), C.c_int(2))
        head = {"value": 2.0 * nn ** 3 / t / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "reference",
                "sample": f"Blaze C = A * B, DynamicMatrix<double,rowMajor> {nn}^3, uniform[-1,1) (the GEMM leg's "
                          f"operands), {proto}, 2 timed runs; FP64 (the CPU has no split-precision path)",
                "ms": t * 1e3}
        if "value" in rec and rec.get("value"):
            head["gpu_over_cpu"] = rec["value"] / head["value"]
        self.gemm_inputs = None
        # config 1: 512^3 from mt19937_64
        if getattr(self, "gemm512_inputs", None) is not None and "plain_gemm_512" in rec:
            A5, B5 = self.gemm512_inputs
            t5 = L.ref_time_dgemm(C.c_int64(512), C.c_int64(512), C.c_int64(512), ptr(A5), ptr(B5), C.c_int(50))
            rec["plain_gemm_512"]["cpu_baseline"] = {
                "value": 2.0 * 512 ** 3 / t5 / 1e12, "unit": "TFLOP/s", "us": t5 * 1e6, "cores": cores,
                "kind": "reference", "sample": f"Blaze C = A * B 512^3, same mt19937_64 operands, {proto}, 50 timed runs"}
            rec["plain_gemm_512"]["gpu_over_cpu"] = rec["plain_gemm_512"]["tflops"] / rec["plain_gemm_512"]["cpu_baseline"]["value"]
        # SpMV: y = S * x on the SpMV leg's matrices (CompressedMatrix<double,rowMajor>)
        for key, inp in (getattr(self, "spmv_inputs", None) or {}).items():
            rp, ci, va, x, ncols = inp
            log(f"[cpu] Blaze y = S * x ({key}) on {cores} cores")
            ts = L.ref_time_dcsrmv(C.c_int64(len(rp) - 1), C.c_int64(ncols), ptr(rp), ptr(ci), ptr(va), ptr(x),
                                   C.c_int(10))
            leg = rec.get("spmv", {}).get(key)
            if isinstance(leg, dict):
                leg["cpu_baseline"] = {"value": leg["bytes"] / ts / 1e9, "unit": "GB/s", "us": ts * 1e6, "cores": cores,
                                       "kind": "reference",
                                       "sample": f"Blaze CompressedMatrix<double,rowMajor> y = S * x, same matrix and x "
                                                 f"(algorithmic bytes as the GPU leg), {proto}, 10 timed runs"}
                leg["gpu_over_cpu"] = leg["graph_gbps"] / leg["cpu_baseline"]["value"]
        self.spmv_inputs = None
        # GEMV: y = A * x, 16384^2 FP64
        gv = rec.get("gemv", {})
        if isinstance(gv, dict) and "dgemv_16384x16384" in gv:
            m = n = 16384
            Ag = self.uniform((m, n), 41); xg = self.uniform(n, 42)
            log(f"[cpu] Blaze y = A * x {m}x{n} on {cores} cores")
            tg = L.ref_time_dgemv(C.c_int64(m), C.c_int64(n), ptr(Ag), ptr(xg), C.c_int(5))
            by = gv["dgemv_16384x16384"]["bytes"]
            gv["dgemv_16384x16384"]["cpu_baseline"] = {
                "value": by / tg / 1e9, "unit": "GB/s", "us": tg * 1e6, "cores": cores, "kind": "reference",
                "sample": f"Blaze DynamicMatrix<double,rowMajor> {m}x{n} y = A * x, uniform[-1,1), {proto}, 5 timed runs"}
            gv["dgemv_16384x16384"]["gpu_over_cpu"] = gv["dgemv_16384x16384"]["gbps"] / (by / tg / 1e9)
            del Ag
        # complex<double> GEMM: Blaze's own complex kernel at 1024^3 (a bounded sample: the CPU
        # product at the GPU leg's 4096^3 would take minutes), the leading 1024 x 1024 blocks of
        # the GPU leg's operands
        cx = rec.get("complex", {})
        zin = getattr(self, "zgemm_inputs", None)
        if zin is not None and isinstance(cx, dict):
            Az, Bz, gpu_tf = zin
            nz = int(round((Az.size // 2) ** 0.5))
            s = 1024
            Azs = np.ascontiguousarray(Az.view(np.complex128).reshape(nz, nz)[:s, :s])
            Bzs = np.ascontiguousarray(Bz.view(np.complex128).reshape(nz, nz)[:s, :s])
            log(f"[cpu] Blaze complex<double> C = A * B {s}^3 on {cores} cores")
            tz = L.ref_time_zgemm(C.c_int64(s), C.c_int64(s), C.c_int64(s), ptr(Azs), ptr(Bzs), C.c_int(2))
            cx["cpu_baseline"] = {"value": 8.0 * s ** 3 / tz / 1e12, "unit": "TFLOP/s", "ms": tz * 1e3, "cores": cores,
                                  "kind": "reference",
                                  "sample": f"Blaze DynamicMatrix<complex<double>,rowMajor> C = A * B {s}^3 (the leading "
                                            f"blocks of the GPU leg's operands), {proto}, 2 timed runs"}
            cx["gpu_over_cpu"] = gpu_tf / cx["cpu_baseline"]["value"]
        self.zgemm_inputs = None
        # batched solves: the reference's solve(A, X, B) per system (closed forms n <= 6, LAPACK
        # dgesv_ above), OpenMP over the systems, LAPACK sequential
        sv = rec.get("solve", {})
        inp = getattr(self, "solve_inputs", None)
        if inp is not None and isinstance(sv, dict) and "fp64_lu" in sv:
            if oracle.have_ref_lapack():
                n, r, a_off, b_off, As, Bs = inp
                log(f"[cpu] Blaze solve() on {len(n)} systems, {cores} threads")
                _, ts, fails = oracle.ref_solve_batch(n, r, a_off, b_off, As, Bs, threads=cores, reps=1)
                sv["cpu_baseline"] = {"value": len(n) / ts, "unit": "systems/s", "ms": ts * 1e3, "cores": cores,
                                      "kind": "reference", "fails": fails,
                                      "sample": f"all {len(n)} systems of the GPU leg: Blaze solve(A, X, B) per system "
                                                "(closed forms n <= 6, dgesv_ from MKL's libmkl_rt, sequential, "
                                                f"n >= 7) under an OpenMP loop over the systems ({cores} threads), one "
                                                "warm-up pass then one timed pass",
                                      "gflops": sv["flops_getrf_getrs"] / ts / 1e9}
                for key in ("fp64_lu", "mixed_fp32_lu_fp64_ir"):
                    if key in sv:
                        sv[key]["gpu_over_cpu"] = sv[key]["systems_per_s"] / sv["cpu_baseline"]["value"]
            else:
                sv["cpu_baseline"] = {"error": "the reference's solve() needs a LAPACK; none was found when "
                                               "oracle/_ref was built"}
        self.solve_inputs = None
        return head


def gemm_leg(b, rec, args):
    """Headline: FP64-accurate GEMM vs native FP64 at args.size^3, error per sigma, rectangular case."""
    import blaze_amd as bz
    m = n = k = args.size
    log(f"[gemm] generating {m}x{k} inputs")
    A = b.uniform((m, k), 1); B = b.uniform((k, n), 2)
    variants = [("fp64acc_bf16", "bzg_dgemm_fp64acc"), ("fp64acc_int8", "bzg_dgemm_fp64acc_i8")]
    res = b.gemm_case(m, n, k, A, B, args.steps, args.warmup, variants)
    log(f"[gemm] {json.dumps(res)}")
    ok = {v: res[v]["rel_frob_err"] <= 1e-14 for v, _ in variants}
    if args.variant == "auto":
        cands = [v for v, _ in variants if ok[v]] or [v for v, _ in variants]
        head = max(cands, key=lambda v: res[v]["tflops"])
    else:
        head = "fp64acc_" + args.variant
    fn = dict(variants)[head]
    slices = res[head]["moduli"]  # chosen by the library from the operands (automatic count)
    rec["value"] = res[head]["tflops"]
    rec["ms_per_step"] = res[head]["ms"]
    rec["dtype"] = "int8" if head.endswith("int8") else "bf16"
    rec["config"] = {"workload": f"FP64-accurate GEMM C = A*B, M=N=K={m}, A,B uniform[-1,1) (sigma=0), "
                                 f"row-major DynamicMatrix<double>",
                     "variant": head, "slices": slices,
                     "slice_arith": ("INT8 x INT8 -> INT32 residues mod m<=255" if head.endswith("int8")
                                     else "BF16 x BF16 -> FP32 residues mod m<=511"),
                     "recombination": ("FP64 CRT, weights split in " + ("40" if head.endswith("int8") else "39")
                                       + "-bit chunks (exact partial sums), one final rounding"),
                     "moduli_choice": "automatic: smallest count with estimated rel. Frobenius error <= 8e-15 (3.2e-15 when single entries dominate rows/columns)",
                     "rel_frob_err": res[head]["rel_frob_err"], "err_sample": res[head]["err_sample"]}
    rec["native_fp64"] = res["native_fp64"]
    rec["fp64acc_vs_native"] = res[head]["vs_native"]
    rec["fp64acc_8192"] = {v: res[v] for v, _ in variants}
    rec["roofline"] = b.roofline_fp64acc(m, n, k, A, B, fn, slices)
    b.gemm_inputs = (A, B)  # the CPU baseline times the reference on the same operands
    # the same product for graded exponents u 2^e, e = round(N(0, sigma)): rate of the automatic
    # moduli count and its error over whole rows (INT8 slices, the headline variant)
    by_sigma = {}
    for sigma in (4, 10):
        log(f"[gemm] sigma={sigma}: generating {m}x{k} graded inputs")
        Ag = b.graded((m, k), 130 + sigma, sigma); Bg = b.graded((k, n), 140 + sigma, sigma)
        r_s = b.gemm_case(m, n, k, Ag, Bg, max(3, args.steps // 2), 1, [("fp64acc_int8", "bzg_dgemm_fp64acc_i8")])
        by_sigma[f"sigma{sigma}"] = r_s
        log(f"[gemm] sigma={sigma}: {json.dumps(r_s)}")
        del Ag, Bg
    by_sigma["sigma0"] = {"native_fp64": res["native_fp64"], "fp64acc_int8": res["fp64acc_int8"]}
    rec[f"fp64acc_{m}_by_sigma"] = by_sigma
    # error per exponent spread (smaller size keeps the run short; K is what matters)
    # error per exponent spread: the full relative Frobenius error of a 512 x 512 x k product
    # (for wide spreads a few entries carry most of ||C||: a sampled ratio would mislead)
    errs = {}
    ms_ = 512
    for sigma in (0, 4, 10):
        Ag = b.graded((ms_, k), 30 + sigma, sigma); Bg = b.graded((k, ms_), 40 + sigma, sigma)
        errs[f"sigma{sigma}"] = b.sigma_case(Ag, Bg, variants)
    rec[f"rel_frob_err_by_sigma_{ms_}x{ms_}x{k}"] = errs
    # the plain products in the other element types (same size, HIP events on the library's
    # stream): FP32 on MFMA; INT32 (wrap-around exact) as 10 INT8 digit products on the INT8
    # MFMA, where bzg_igemm sends products of this size
    torch = b.torch
    other = {}
    for name, fn32, dt, one in (("fp32", "bzg_sgemm", torch.float32, 1.0), ("int32", "bzg_igemm", torch.int32, 1)):
        if dt == torch.int32:
            dA = torch.randint(-2 ** 31, 2 ** 31, (m, k), dtype=torch.int64, device="cuda").to(dt)
            dB = torch.randint(-2 ** 31, 2 ** 31, (k, n), dtype=torch.int64, device="cuda").to(dt)
        else:
            dA = torch.from_numpy(A.astype(np.float32)).cuda(); dB = torch.from_numpy(B.astype(np.float32)).cuda()
        dC = torch.empty(m, n, dtype=dt, device="cuda")
        zero = 0 if dt == torch.int32 else 0.0
        call = lambda: b.ctx.call(fn32, m, n, k, one, vp(dA), k, vp(dB), n, zero, vp(dC), n)
        call(); b.sync()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        reps = 3
        ev[0].record()
        for _ in range(reps):
            call()
        ev[1].record(); b.sync()
        ms = ev[0].elapsed_time(ev[1]) / reps
        peak = PEAK["fp32"] if name == "fp32" else None
        other[name] = {"ms": ms, "tflops": 2.0 * m * n * k / ms / 1e9,
                       "frac_peak": (2.0 * m * n * k / ms / 1e9 / peak) if peak else None}
        if name == "int32":  # the INT8 work the 10 digit products issue, against the INT8 peak
            other[name]["int8_mfma_frac_peak"] = 10 * 2.0 * m * n * k / ms / 1e9 / PEAK["int8"]
        del dA, dB, dC
    rec["plain_gemm_other_types"] = other
    log(f"[gemm] other element types: {json.dumps(other)}")
    if m == 8192:
        mr, nr_, kr = 16384, 1024, 16384
        Ar = b.uniform((mr, kr), 3); Br = b.uniform((kr, nr_), 4)
        rec["fp64acc_16384x1024x16384"] = b.gemm_case(mr, nr_, kr, Ar, Br, max(3, args.steps // 2), 1, variants)
        del Ar, Br
    del A, B
    rec["plain_gemm_512"] = b.gemm512_leg()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--systems", type=int, default=100_000)
    ap.add_argument("--legs", default="gemm,solve,spmv,gemv,complex,cg,serve,cpu",
                    help="comma list of gemm, solve, spmv, gemv, complex, cg, serve, cpu, copy (1 GiB copy for "
                         "counter calibration)")
    ap.add_argument("--variant", default="auto", choices=["auto", "bf16", "int8"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(sharding across GPUs is not implemented)"}))
        sys.exit(2)
    legs = set(args.legs.split(","))
    b = Bench(args)
    import blaze_amd as bz
    m = n = k = args.size
    rec = {"metric": METRIC, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    if "gemm" in legs:
        gemm_leg(b, rec, args)
    if "solve" in legs:
        try:
            rec["solve"] = b.solve_leg(args.systems, max(3, args.steps // 2), 1)
            log(f"[solve] {json.dumps(rec['solve'])}")
            rec["batched_solve_p50_ms"] = rec["solve"]["mixed_fp32_lu_fp64_ir"]["p50_ms"]
            rec["batched_solve_p99_ms"] = rec["solve"]["mixed_fp32_lu_fp64_ir"]["p99_ms"]
        except Exception as e:  # report, never hide
            rec["solve"] = {"error": repr(e)}
    if "spmv" in legs:
        try:
            rec["spmv"] = b.spmv_leg(max(20, args.steps * 3))
            log(f"[spmv] {json.dumps(rec['spmv'])}")
        except Exception as e:
            rec["spmv"] = {"error": repr(e)}
    if "gemv" in legs:
        try:
            rec["gemv"] = b.gemv_leg()
            log(f"[gemv] {json.dumps(rec['gemv'])}")
        except Exception as e:
            rec["gemv"] = {"error": repr(e)}
    if "complex" in legs:
        try:
            rec["complex"] = b.complex_leg()
            log(f"[complex] {json.dumps(rec['complex'])}")
        except Exception as e:
            rec["complex"] = {"error": repr(e)}
    if "cg" in legs:
        try:
            rec["cg"] = b.cg_leg()
            log(f"[cg] {json.dumps(rec['cg'])}")
        except Exception as e:
            rec["cg"] = {"error": repr(e)}
    if "serve" in legs:
        try:
            from blaze_amd import serve_bench
            rec["serve"] = serve_bench.run(b)
        except ImportError:
            pass
        except Exception as e:
            rec["serve"] = {"error": repr(e)}
    if "cpu" in legs:
        rec["cpu_baseline"] = b.cpu_baseline(rec)
    if "copy" in legs:
        # counter calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE: one 1 GiB device copy
        src = b.torch.ones(1 << 27, dtype=b.torch.float64, device="cuda")
        dst = b.torch.empty_like(src)
        b.sync()
        dst.copy_(src)
        b.sync()
        rec["copy_bytes"] = 1 << 30
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
