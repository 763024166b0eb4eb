// Continuous-batching request server over the GEMM, SpMV and batched-solve entry points.
//
// No reference counterpart: the reference evaluates one expression per call on the calling
// thread.  This is the serving loop around the same operations (C = A * B,
// DMatDMatMultExpr.h:516; y = S * x, SMatDVecMultExpr.h:204; solve(A, x, b), LSE.h:6343).
//
// Design:
//   * Operands are device-resident (the caller's HBM buffers; CSR matrices are uploaded once
//     with bzg_server_add_csr and stay resident with their plans).  Results are written into
//     server-owned HBM and stay there until the client releases them.
//   * One dispatcher thread owns admission and completion.  A request is admitted when fewer
//     than max_inflight launches are on the device and its bytes (result + staging) fit in
//     the HBM budget; admission is FIFO (the head blocks the queue), except that every
//     admissible solve request waiting behind the head joins the head's solve batch.
//   * Admitted work is spread over `streams` HIP streams (default priority), one library
//     context per stream, so requests that arrive while others run overlap on the device.
//   * Solve requests admitted together run as ONE batched launch (bzg_dsgesv_batched: FP32 LU +
//     FP64 refinement, FP64 fallback), longest system first.  Small GEMM requests (up to 2^31
//     flops) waiting together likewise run as one grouped launch (bzg_dgemm_grouped).
//   * Completion is detected with HIP events (no stream synchronisation in the loop).
// Every kernel computes each request's result independently of what else is in flight (GEMM
// tiles depend only on the shape, SpMV on the plan, a solve on its own system), so a result
// is bit-identical alone and in any mix.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace bzg {
namespace {

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

size_t size_class(size_t bytes) {
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}

// Size-class cache of device or pinned host buffers (power-of-two classes); never returns
// memory to the runtime before the server is destroyed, so the loop never calls hipFree.
struct Pool {
  bool host = false;
  std::unordered_map<size_t, std::vector<void*>> free_;
  std::vector<void*> all_;
  bzg_status get(size_t bytes, void** p, size_t* cls) {
    *cls = size_class(bytes);
    auto& fl = free_[*cls];
    if (!fl.empty()) {
      *p = fl.back();
      fl.pop_back();
      return BZG_OK;
    }
    if (host) {
      BZG_HIP_CHECK(hipHostMalloc(p, *cls, hipHostMallocDefault));
    } else {
      BZG_HIP_CHECK(hipMalloc(p, *cls));
    }
    all_.push_back(*p);
    return BZG_OK;
  }
  void put(void* p, size_t cls) {
    if (p) free_[cls].push_back(p);
  }
  void release_all() {
    for (void* p : all_) {
      if (host)
        (void)hipHostFree(p);
      else
        (void)hipFree(p);
    }
    all_.clear();
    free_.clear();
  }
};

enum Kind { kGemm = 0, kSpmv = 1, kSolve = 2 };

struct Group;

struct Request {
  int64_t id = 0;
  int kind = kGemm;
  int64_t m = 0, n = 0, k = 0, lda = 0, ldb = 0;
  const double* A = nullptr;
  const double* B = nullptr;
  int32_t csr = -1;
  bzg_status status = BZG_OK;
  int32_t info = 0, iter = 0;
  double* out = nullptr;      // result in HBM (C, y or X)
  int64_t out_elems = 0;
  size_t out_cls = 0;         // pool class of `out` (0: part of the group's pool)
  Group* group = nullptr;
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool done = false;
  bool direct = false;        // solve of order > 256: alone, through the single-system FP64 LU
};

struct Csr {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int32_t* rowptr = nullptr;
  int32_t* colidx = nullptr;
  double* val = nullptr;
  bzg_csr_plan* plan = nullptr;
  bool sliced = false;        // the plan holds the matrix as sliced ELL; the CSR arrays are freed
  size_t bytes = 0;
};

// One launch on one stream: a GEMM, an SpMV, or a batch of solves.
struct Group {
  std::vector<Request*> reqs;
  int stream = 0;
  hipEvent_t ev = nullptr;
  // solve batch: device pool (meta + packed B + X) and its pinned host mirror of the meta
  void* dpool = nullptr;
  size_t dcls = 0;
  void* hpool = nullptr;
  size_t hcls = 0;
  size_t budget_bytes = 0;    // bytes charged to the budget for the pool (released with it)
  int live = 0;               // requests of the batch not yet released
  bool gemm_batch = false;    // a grouped GEMM launch (descriptor pool freed at completion)
  double est_us = 0;          // estimated device time (stream choice)
  int32_t* h_iter = nullptr;  // pinned copies of iter / info, valid once the event completed
  int32_t* h_info = nullptr;
};

// Copies each system's B (n_i x nrhs_i, row-major, anywhere in HBM) into the packed pool.
__global__ __launch_bounds__(256) void gather_rhs_kernel(int count, const int32_t* __restrict__ nv,
                                                         const int32_t* __restrict__ rv,
                                                         const int64_t* __restrict__ src,
                                                         const int64_t* __restrict__ b_off,
                                                         double* __restrict__ dst) {
  const int s = blockIdx.x;
  if (s >= count) return;
  const double* from = reinterpret_cast<const double*>(src[s]);
  const int len = nv[s] * rv[s];
  for (int e = threadIdx.x; e < len; e += blockDim.x) dst[b_off[s] + e] = from[e];
}

}  // namespace
}  // namespace bzg

struct bzg_server {
  bzg_server_config cfg{};
  int device = 0;
  std::vector<bzg_context*> ctx;
  std::vector<double> load;  // estimated device microseconds in flight per stream
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<bzg::Request*> incoming;  // submitted, not yet seen by the dispatcher (mu)
  std::deque<bzg::Request*> pending;   // dispatcher: waiting for admission
  std::vector<bzg::Group*> inflight;   // dispatcher: launched, not complete
  std::unordered_map<int64_t, bzg::Request*> reqs;  // every unreleased request (mu)
  std::vector<bzg::Csr> csrs;                        // resident matrices (mu)
  std::vector<hipEvent_t> events;                    // free events (dispatcher)
  bzg::Pool dev, host;                               // (mu)
  size_t used = 0, peak = 0;                         // budget bytes in use (mu)
  int64_t next_id = 1;
  bzg_server_stats st{};
  bool stop = false;
  bool hold = false;  // admissions paused by bzg_server_hold (mu)
  bzg_status fatal = BZG_OK;
  std::thread th;
};

namespace bzg {
namespace {

// Byte layout of a solve batch's device pool: metadata (n, nrhs, order, iter, info as int32;
// a_off, b_off, source pointers as int64), then the packed right-hand sides, then X.
struct SolveLayout {
  size_t off_a, off_b, off_src, meta, off_B, off_X, total;
  SolveLayout(int64_t cnt, int64_t elems) {
    const size_t i32b = (size_t)cnt * 4, i64b = (size_t)cnt * 8;
    off_a = (5 * i32b + 7) / 8 * 8;
    off_b = off_a + i64b;
    off_src = off_b + i64b;
    meta = off_src + i64b;
    off_B = (meta + 255) / 256 * 256;
    off_X = off_B + (size_t)elems * 8;
    total = off_X + (size_t)elems * 8;
  }
};

// Byte layout of a direct solve's device pool: a copy of A (factored in place), X, the pivots
// and info.
struct DirectLayout {
  size_t off_X, off_piv, off_info, total;
  DirectLayout(int64_t n, int64_t k) {
    off_X = ((size_t)n * n * 8 + 255) / 256 * 256;
    off_piv = off_X + ((size_t)n * k * 8 + 255) / 256 * 256;
    off_info = off_piv + ((size_t)n * 4 + 255) / 256 * 256;
    total = off_info + 256;
  }
};

// Budget bytes a request holds until released (a solve alone: its own batch pool).
size_t request_bytes(const bzg_server* s, const Request* r) {
  if (r->kind == kGemm) return size_class((size_t)r->m * r->n * sizeof(double));
  if (r->kind == kSpmv) return size_class((size_t)s->csrs[r->csr].n_rows * sizeof(double));
  if (r->direct) return size_class(DirectLayout(r->n, r->k).total);
  return size_class(SolveLayout(1, r->n * r->k).total);
}

// Rough device time of a request, only used to balance the streams: GEMM at 40 TFLOP/s,
// SpMV at 4 TB/s, a solve at 1 TFLOP/s, plus a launch cost.
double estimate_us(const bzg_server* s, const Request* r) {
  if (r->kind == kGemm) return 10.0 + 2.0 * r->m * r->n * r->k / 40e6;
  if (r->kind == kSpmv) {
    const Csr& M = s->csrs[r->csr];
    return 10.0 + (M.nnz * 12.0 + M.n_rows * 20.0) / 4e6;
  }
  return 20.0 + (2.0 / 3.0) * r->n * r->n * r->n / 1e6;
}

int pick_stream(const bzg_server* s) {
  int best = 0;
  for (int i = 1; i < (int)s->load.size(); ++i)
    if (s->load[i] < s->load[best]) best = i;
  return best;
}

hipEvent_t take_event(bzg_server* s) {
  if (!s->events.empty()) {
    hipEvent_t e = s->events.back();
    s->events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

// Launches one GEMM or SpMV request (dispatcher thread, mu held for the pool).
bzg_status launch_single(bzg_server* s, Request* r, Group* g) {
  bzg_context* c = s->ctx[g->stream];
  void* p = nullptr;
  bzg_status st = s->dev.get((size_t)r->out_elems * sizeof(double), &p, &r->out_cls);
  if (st) return st;
  r->out = static_cast<double*>(p);
  if (r->kind == kGemm) {
    st = bzg_dgemm(c, r->m, r->n, r->k, 1.0, r->A, r->lda, r->B, r->ldb, 0.0, r->out, r->n);
  } else {
    const Csr& M = s->csrs[r->csr];
    st = M.sliced ? bzg_dsellmv(c, M.plan, r->A, r->out) : bzg_dcsrmv(c, M.plan, M.rowptr, M.colidx, M.val, r->A, r->out);
  }
  return st;
}

// GEMMs up to this many flops are fused with the other waiting ones into one grouped launch
// (64 x 64 tiles); larger ones run alone with the tile shape chosen for their size.  Either
// way every element is summed in the same order, so the result does not depend on the choice.
constexpr double kGroupFlops = 2.0 * 1024 * 1024 * 1024;
bool groupable(const Request* r) { return r->kind == kGemm && 2.0 * r->m * r->n * r->k <= kGroupFlops; }

// Launches several GEMM requests as one bzg_dgemm_grouped call (descriptors staged through
// pinned memory into the group's device pool).
bzg_status launch_gemms(bzg_server* s, Group* g) {
  bzg_context* c = s->ctx[g->stream];
  const int cnt = (int)g->reqs.size();
  const size_t bytes = (size_t)cnt * sizeof(bzg_dgemm_problem);
  bzg_status st = s->dev.get(bytes, &g->dpool, &g->dcls);
  if (!st) st = s->host.get(bytes, &g->hpool, &g->hcls);
  if (st) return st;
  auto* h = static_cast<bzg_dgemm_problem*>(g->hpool);
  for (int i = 0; i < cnt; ++i) {
    Request* r = g->reqs[i];
    void* p = nullptr;
    st = s->dev.get((size_t)r->out_elems * sizeof(double), &p, &r->out_cls);
    if (st) return st;
    r->out = static_cast<double*>(p);
    h[i] = bzg_dgemm_problem{r->m, r->n, r->k, r->A, r->lda, r->B, r->ldb, r->out, r->n, 1.0, 0.0};
  }
  BZG_HIP_CHECK(hipMemcpyAsync(g->dpool, g->hpool, bytes, hipMemcpyHostToDevice, c->stream));
  return bzg_dgemm_grouped(c, cnt, static_cast<const bzg_dgemm_problem*>(g->dpool), nullptr);
}

// Launches a batch of solve requests as one bzg_dsgesv_batched call.
bzg_status launch_solves(bzg_server* s, Group* g) {
  bzg_context* c = s->ctx[g->stream];
  const int cnt = (int)g->reqs.size();
  // longest first (the persistent kernel takes systems in this order)
  std::vector<int> ord(cnt);
  for (int i = 0; i < cnt; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return g->reqs[a]->n > g->reqs[b]->n; });
  // meta layout (bytes): n[cnt] i32, nrhs[cnt] i32, order[cnt] i32, iter[cnt] i32, info[cnt] i32,
  // (pad to 8) a_off[cnt] i64, b_off[cnt] i64, src[cnt] i64, then packed B, then X
  int64_t tot = 0;
  std::vector<int64_t> boff(cnt);
  for (int i = 0; i < cnt; ++i) {
    boff[i] = tot;
    tot += g->reqs[i]->n * g->reqs[i]->k;
  }
  const SolveLayout L(cnt, tot);
  const size_t i32b = (size_t)cnt * 4;
  const size_t off_a = L.off_a, off_b = L.off_b, off_src = L.off_src, meta = L.meta;
  const size_t off_B = L.off_B, off_X = L.off_X;
  bzg_status st = s->dev.get(L.total, &g->dpool, &g->dcls);
  if (!st) st = s->host.get(meta, &g->hpool, &g->hcls);
  if (st) return st;
  char* h = static_cast<char*>(g->hpool);
  char* d = static_cast<char*>(g->dpool);
  int32_t* hn = reinterpret_cast<int32_t*>(h);
  int32_t* hr = hn + cnt;
  int32_t* ho = hr + cnt;
  g->h_iter = ho + cnt;
  g->h_info = g->h_iter + cnt;
  int64_t* ha = reinterpret_cast<int64_t*>(h + off_a);
  int64_t* hb = reinterpret_cast<int64_t*>(h + off_b);
  int64_t* hs = reinterpret_cast<int64_t*>(h + off_src);
  // A stays where the client keeps it: offsets relative to the lowest A address
  const double* abase = g->reqs[0]->A;
  for (auto* r : g->reqs) abase = std::min(abase, r->A);
  for (int i = 0; i < cnt; ++i) {
    const Request* r = g->reqs[i];
    hn[i] = (int32_t)r->n;
    hr[i] = (int32_t)r->k;
    ho[i] = ord[i];
    g->h_iter[i] = 0;
    g->h_info[i] = 0;
    ha[i] = (int64_t)(r->A - abase);
    hb[i] = boff[i];
    hs[i] = (int64_t)(uintptr_t)r->B;
  }
  BZG_HIP_CHECK(hipMemcpyAsync(d, h, meta, hipMemcpyHostToDevice, c->stream));
  const int32_t* dn = reinterpret_cast<const int32_t*>(d);
  const int32_t* dr = dn + cnt;
  const int32_t* dord = dr + cnt;
  int32_t* diter = const_cast<int32_t*>(dord + cnt);
  int32_t* dinfo = diter + cnt;
  const int64_t* da = reinterpret_cast<const int64_t*>(d + off_a);
  const int64_t* db = reinterpret_cast<const int64_t*>(d + off_b);
  const int64_t* dsrc = reinterpret_cast<const int64_t*>(d + off_src);
  double* Bp = reinterpret_cast<double*>(d + off_B);
  double* Xp = reinterpret_cast<double*>(d + off_X);
  hipLaunchKernelGGL(gather_rhs_kernel, dim3((unsigned)cnt), dim3(256), 0, c->stream, cnt, dn, dr, dsrc, db, Bp);
  BZG_LAUNCH_CHECK();
  st = bzg_dsgesv_batched(c, cnt, dn, dr, da, db, abase, Bp, Xp, diter, dinfo, 0.0, 30, dord);
  if (st) return st;
  // iter / info back to the pinned mirror (read when the group's event has completed)
  BZG_HIP_CHECK(hipMemcpyAsync(g->h_iter, diter, 2 * i32b, hipMemcpyDeviceToHost, c->stream));
  for (int i = 0; i < cnt; ++i) {
    g->reqs[i]->out = Xp + boff[i];
    g->reqs[i]->out_cls = 0;
  }
  g->live = cnt;
  return BZG_OK;
}

// A solve of order 257..512 on its own: the batched mixed solver's kernel for these orders
// takes one wavefront per system (about 10 ms for n = 512), the single-system FP64 LU
// (bzg_dgetrf + bzg_dgetrs: recursive, MFMA trailing updates) a few ms.  The result is FP64
// LU's (iterations = -1, as for a fallback); info comes back with the group's event.
bzg_status launch_direct(bzg_server* s, Group* g) {
  bzg_context* c = s->ctx[g->stream];
  Request* r = g->reqs[0];
  const DirectLayout L(r->n, r->k);
  bzg_status st = s->dev.get(L.total, &g->dpool, &g->dcls);
  if (!st) st = s->host.get(2 * sizeof(int32_t), &g->hpool, &g->hcls);
  if (st) return st;
  char* d = static_cast<char*>(g->dpool);
  double* Ap = reinterpret_cast<double*>(d);
  double* Xp = reinterpret_cast<double*>(d + L.off_X);
  int32_t* piv = reinterpret_cast<int32_t*>(d + L.off_piv);
  int32_t* dinfo = reinterpret_cast<int32_t*>(d + L.off_info);
  g->h_iter = static_cast<int32_t*>(g->hpool);
  g->h_info = g->h_iter + 1;
  g->h_iter[0] = -1;
  g->h_info[0] = 0;
  BZG_HIP_CHECK(hipMemcpyAsync(Ap, r->A, (size_t)r->n * r->n * 8, hipMemcpyDeviceToDevice, c->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(Xp, r->B, (size_t)r->n * r->k * 8, hipMemcpyDeviceToDevice, c->stream));
  if ((st = bzg_dgetrf(c, r->n, Ap, r->n, piv, dinfo))) return st;
  if ((st = bzg_dgetrs(c, r->n, r->k, Ap, r->n, piv, Xp, r->k))) return st;
  BZG_HIP_CHECK(hipMemcpyAsync(g->h_info, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  r->out = Xp;
  r->out_cls = 0;
  g->live = 1;
  return BZG_OK;
}

void release_group_pool(bzg_server* s, Group* g) {
  s->dev.put(g->dpool, g->dcls);
  s->host.put(g->hpool, g->hcls);
  s->used -= g->budget_bytes;
  delete g;
}

void dispatcher(bzg_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    bool progressed = false;
    bool held = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      while (!s->incoming.empty()) {
        s->pending.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
      if (s->stop && s->pending.empty() && s->inflight.empty()) break;
      held = s->hold && !s->stop;  // destroy admits and finishes everything
    }
    // ---- completions ----
    for (size_t gi = 0; gi < s->inflight.size();) {
      Group* g = s->inflight[gi];
      const hipError_t q = hipEventQuery(g->ev);
      if (q == hipErrorNotReady) {
        ++gi;
        continue;
      }
      const double t = now_us();
      {
        std::lock_guard<std::mutex> lk(s->mu);
        for (size_t i = 0; i < g->reqs.size(); ++i) {
          Request* r = g->reqs[i];
          if (q != hipSuccess) r->status = BZG_ERR_HIP;
          if (r->kind == kSolve) {
            r->iter = g->h_iter[i];
            r->info = g->h_info[i];
          }
          r->t_done = t;
          r->done = true;
          s->st.completed++;
        }
        s->load[g->stream] -= g->est_us;
        s->events.push_back(g->ev);
        if (g->reqs[0]->kind != kSolve) {
          for (auto* r : g->reqs) r->group = nullptr;
          s->dev.put(g->dpool, g->dcls);
          s->host.put(g->hpool, g->hcls);
          delete g;
        }
      }
      s->cv_done.notify_all();
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->inflight[gi] = s->inflight.back();
        s->inflight.pop_back();
        s->st.inflight = (int64_t)s->inflight.size();
      }
      progressed = true;
    }
    // ---- admissions ----
    while (!held && !s->pending.empty() && (int)s->inflight.size() < s->cfg.max_inflight) {
      Request* head = s->pending.front();
      std::unique_lock<std::mutex> lk(s->mu);
      const size_t need = request_bytes(s, head);
      if (s->used + need > s->cfg.hbm_budget_bytes) break;  // FIFO: wait for releases
      Group* g = new Group();
      g->stream = pick_stream(s);
      g->ev = take_event(s);
      if (!g->ev) {
        delete g;
        s->fatal = BZG_ERR_HIP;
        break;
      }
      const double t = now_us();
      bzg_status st = BZG_OK;
      if (groupable(head)) {
        // the head's GEMM batch: every admissible waiting small GEMM, in arrival order
        for (auto it = s->pending.begin(); it != s->pending.end();) {
          Request* r = *it;
          if (!groupable(r)) { ++it; continue; }
          const size_t b = request_bytes(s, r);
          if ((int)g->reqs.size() >= s->cfg.solve_batch_max || s->used + b > s->cfg.hbm_budget_bytes) break;
          s->used += b;
          g->reqs.push_back(r);
          r->group = g;
          r->t_admit = t;
          it = s->pending.erase(it);
        }
        g->gemm_batch = true;
        st = launch_gemms(s, g);
        s->st.gemm_batches++;
        s->st.max_gemm_batch = std::max<int64_t>(s->st.max_gemm_batch, (int64_t)g->reqs.size());
      } else if (head->kind != kSolve) {
        s->pending.pop_front();
        g->reqs.push_back(head);
        head->group = g;
        s->used += need;
        head->t_admit = t;
        st = launch_single(s, head, g);
      } else if (head->direct) {
        s->pending.pop_front();
        g->reqs.push_back(head);
        head->group = g;
        head->t_admit = t;
        g->budget_bytes = need;
        s->used += need;
        st = launch_direct(s, g);
        s->st.solve_batches++;
        s->st.max_solve_batch = std::max<int64_t>(s->st.max_solve_batch, 1);
      } else {
        // the head's batch: every admissible waiting solve of order <= 256, in arrival order
        // (the pool of the whole batch must fit; the head fits: checked above)
        size_t bytes = 0;
        int64_t elems = 0;
        for (auto it = s->pending.begin(); it != s->pending.end();) {
          Request* r = *it;
          if (r->kind != kSolve || r->direct) { ++it; continue; }
          const int64_t cnt = (int64_t)g->reqs.size() + 1;
          const size_t b = size_class(SolveLayout(cnt, elems + r->n * r->k).total);
          if (cnt > s->cfg.solve_batch_max || s->used + b > s->cfg.hbm_budget_bytes) break;
          bytes = b;
          elems += r->n * r->k;
          g->reqs.push_back(r);
          r->group = g;
          r->t_admit = t;
          it = s->pending.erase(it);
        }
        g->budget_bytes = bytes;
        s->used += bytes;
        st = launch_solves(s, g);
        s->st.solve_batches++;
        s->st.max_solve_batch = std::max<int64_t>(s->st.max_solve_batch, (int64_t)g->reqs.size());
      }
      s->peak = std::max(s->peak, s->used);
      s->st.admitted += (int64_t)g->reqs.size();
      s->st.launches++;
      if (st == BZG_OK && hipEventRecord(g->ev, s->ctx[g->stream]->stream) != hipSuccess) st = BZG_ERR_HIP;
      if (st != BZG_OK) {
        // not launched: complete the requests with the error, give the bytes back
        for (auto* r : g->reqs) {
          r->status = st;
          r->done = true;
          r->t_done = now_us();
          s->st.completed++;
        }
        s->events.push_back(g->ev);
        g->ev = nullptr;
        if (g->reqs[0]->kind == kSolve) {
          g->live = (int)g->reqs.size();
        } else {
          for (auto* r : g->reqs) r->group = nullptr;
          s->dev.put(g->dpool, g->dcls);
          s->host.put(g->hpool, g->hcls);
          delete g;
        }
        lk.unlock();
        s->cv_done.notify_all();
        continue;
      }
      for (auto* r : g->reqs) g->est_us += estimate_us(s, r);
      s->load[g->stream] += g->est_us;
      s->inflight.push_back(g);
      s->st.inflight = (int64_t)s->inflight.size();
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->incoming.empty() && !s->stop)
        s->cv_work.wait_for(lk, std::chrono::microseconds(s->inflight.empty() ? 2000 : 20));
    }
  }
}

}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_server_create(const bzg_server_config* cfg, bzg_server** out) {
  BZG_REQUIRE(cfg && out, BZG_ERR_INVALID_ARGUMENT, "server_create: null argument");
  BZG_REQUIRE(cfg->hbm_budget_bytes > 0 && cfg->max_inflight >= 1 && cfg->streams >= 1 && cfg->streams <= 16 &&
                  cfg->solve_batch_max >= 1,
              BZG_ERR_INVALID_ARGUMENT, "server_create: bad configuration");
  auto* s = new bzg_server();
  s->cfg = *cfg;
  s->host.host = true;
  BZG_HIP_CHECK(hipGetDevice(&s->device));
  for (int i = 0; i < cfg->streams; ++i) {
    bzg_context* c = nullptr;
    bzg_status st = bzg_create(&c);
    if (st) {
      for (auto* x : s->ctx) bzg_destroy(x);
      delete s;
      return st;
    }
    s->ctx.push_back(c);
    s->load.push_back(0.0);
  }
  s->st.budget = (int64_t)cfg->hbm_budget_bytes;
  s->th = std::thread(bzg::dispatcher, s);
  *out = s;
  return BZG_OK;
}

bzg_status bzg_server_destroy(bzg_server* s) {
  BZG_REQUIRE(s != nullptr, BZG_ERR_INVALID_ARGUMENT, "server_destroy: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->th.join();
  for (auto* c : s->ctx) bzg_synchronize(c);
  for (auto& kv : s->reqs) {
    bzg::Request* r = kv.second;
    if (r->group && r->kind == bzg::kSolve && --r->group->live == 0) bzg::release_group_pool(s, r->group);
    delete r;
  }
  for (auto& m : s->csrs) {
    bzg_csr_plan_destroy(m.plan);
    (void)hipFree(m.rowptr);
    (void)hipFree(m.colidx);
    (void)hipFree(m.val);
  }
  for (auto e : s->events) (void)hipEventDestroy(e);
  s->dev.release_all();
  s->host.release_all();
  for (auto* c : s->ctx) bzg_destroy(c);
  delete s;
  return BZG_OK;
}

bzg_status bzg_server_add_csr(bzg_server* s, int64_t n_rows, int64_t n_cols, const int32_t* rowptr,
                              const int32_t* colidx, const double* val, int32_t* handle) {
  BZG_REQUIRE(s && rowptr && handle && n_rows >= 0 && n_cols >= 0, BZG_ERR_INVALID_ARGUMENT,
              "server_add_csr: bad argument");
  const int64_t nnz = rowptr[n_rows];
  BZG_REQUIRE(nnz >= 0 && (nnz == 0 || (colidx && val)), BZG_ERR_INVALID_ARGUMENT, "server_add_csr: bad arrays");
  bzg::Csr M;
  M.n_rows = n_rows;
  M.n_cols = n_cols;
  M.nnz = nnz;
  M.bytes = (size_t)(n_rows + 1) * 4 + (size_t)nnz * 12;
  // rows of <= 32 nonzeros throughout: the matrix is kept as the plan's sliced-ELL copy (64-row
  // slices, padded to each slice's longest row), which the SpMV reads with coalesced loads;
  // both forms are reserved until the copy exists, then the CSR arrays and their bytes go
  size_t sell_bytes = 0;
  {
    int64_t slots = 0;
    bool short_rows = true;
    for (int64_t s0 = 0; s0 < n_rows && short_rows; s0 += 64) {
      int32_t w = 0;
      for (int64_t r = s0; r < std::min<int64_t>(n_rows, s0 + 64); ++r) w = std::max(w, rowptr[r + 1] - rowptr[r]);
      short_rows = w <= 32;
      slots += 64 * (int64_t)w;
    }
    if (short_rows && n_rows > 0) sell_bytes = (size_t)slots * 12 + (size_t)((n_rows + 63) / 64 + 1) * 8;
  }
  M.bytes += sell_bytes;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    BZG_REQUIRE(s->used + M.bytes <= s->cfg.hbm_budget_bytes, BZG_ERR_OUT_OF_MEMORY,
                "server_add_csr: matrix does not fit in the HBM budget");
    s->used += M.bytes;
    s->peak = std::max(s->peak, s->used);
  }
  bzg_context* c = s->ctx[0];
  // every failure below frees what was allocated and returns the budget reserved above
  auto fail = [&](bzg_status st, const std::string& what) {
    if (M.plan) bzg_csr_plan_destroy(M.plan);
    if (M.rowptr) hipFree(M.rowptr);
    if (M.colidx) hipFree(M.colidx);
    if (M.val) hipFree(M.val);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->used -= M.bytes;
    }
    if (!what.empty()) bzg::set_error("server_add_csr: " + what);
    return st;
  };
  hipError_t e = hipMalloc(&M.rowptr, (size_t)(n_rows + 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&M.colidx, std::max<size_t>((size_t)nnz * 4, 4));
  if (e == hipSuccess) e = hipMalloc(&M.val, std::max<size_t>((size_t)nnz * 8, 8));
  if (e != hipSuccess) return fail(BZG_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpyAsync(M.rowptr, rowptr, (size_t)(n_rows + 1) * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && nnz) e = hipMemcpyAsync(M.colidx, colidx, (size_t)nnz * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && nnz) e = hipMemcpyAsync(M.val, val, (size_t)nnz * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(BZG_ERR_HIP, std::string("copy: ") + hipGetErrorString(e));
  bzg_status st = bzg_csr_plan_create(c, n_rows, n_cols, rowptr, &M.plan);
  if (st) {
    M.plan = nullptr;
    return fail(st, "");  // the plan's own message stays in bzg_last_error
  }
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return fail(BZG_ERR_HIP, std::string("plan: ") + hipGetErrorString(e));
  if (sell_bytes) {
    st = bzg_csr_plan_pack_sell(c, M.plan, M.rowptr, M.colidx, M.val, 0);
    if (st) return fail(st, "");
    hipFree(M.rowptr);
    hipFree(M.colidx);
    hipFree(M.val);
    M.rowptr = nullptr;
    M.colidx = nullptr;
    M.val = nullptr;
    M.sliced = true;
    const size_t csr_bytes = M.bytes - sell_bytes;
    M.bytes = sell_bytes;
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= csr_bytes;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->csrs.push_back(M);
  *handle = (int32_t)s->csrs.size() - 1;
  return BZG_OK;
}

static bzg_status submit(bzg_server* s, bzg::Request* r, int64_t* id) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) {
      delete r;
      return s->fatal;
    }
    if (bzg::request_bytes(s, r) > s->cfg.hbm_budget_bytes) {
      delete r;
      bzg::set_error("server: request larger than the whole HBM budget");
      return BZG_ERR_OUT_OF_MEMORY;
    }
    r->id = s->next_id++;
    r->t_submit = bzg::now_us();
    s->reqs[r->id] = r;
    s->incoming.push_back(r);
    s->st.submitted++;
    *id = r->id;
  }
  s->cv_work.notify_one();
  return BZG_OK;
}

bzg_status bzg_submit_gemm(bzg_server* s, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                           const double* B, int64_t ldb, int64_t* id) {
  BZG_REQUIRE(s && id && m > 0 && n > 0 && k >= 0 && A && B && lda >= std::max<int64_t>(k, 1) &&
                  ldb >= n,
              BZG_ERR_INVALID_ARGUMENT, "submit_gemm: bad argument");
  auto* r = new bzg::Request();
  r->kind = bzg::kGemm;
  r->m = m; r->n = n; r->k = k; r->lda = lda; r->ldb = ldb; r->A = A; r->B = B;
  r->out_elems = m * n;
  return submit(s, r, id);
}

bzg_status bzg_submit_spmv(bzg_server* s, int32_t csr, const double* x, int64_t* id) {
  BZG_REQUIRE(s && id && x, BZG_ERR_INVALID_ARGUMENT, "submit_spmv: bad argument");
  int64_t rows = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    BZG_REQUIRE(csr >= 0 && csr < (int32_t)s->csrs.size(), BZG_ERR_INVALID_ARGUMENT, "submit_spmv: unknown matrix");
    rows = s->csrs[csr].n_rows;
  }
  BZG_REQUIRE(rows > 0, BZG_ERR_INVALID_ARGUMENT, "submit_spmv: empty matrix");
  auto* r = new bzg::Request();
  r->kind = bzg::kSpmv;
  r->csr = csr;
  r->A = x;
  r->out_elems = rows;
  return submit(s, r, id);
}

bzg_status bzg_submit_solve(bzg_server* s, int32_t n, int32_t nrhs, const double* A, const double* B,
                            int64_t* id) {
  BZG_REQUIRE(s && id && A && B, BZG_ERR_INVALID_ARGUMENT, "submit_solve: bad argument");
  BZG_REQUIRE(n >= 1 && n <= 512 && nrhs >= 1 && nrhs <= 8, BZG_ERR_UNSUPPORTED,
              "submit_solve: n must be 1..512 and nrhs 1..8");
  BZG_REQUIRE(((uintptr_t)A & 7) == 0 && ((uintptr_t)B & 7) == 0, BZG_ERR_INVALID_ARGUMENT,
              "submit_solve: misaligned operand");
  auto* r = new bzg::Request();
  r->kind = bzg::kSolve;
  r->n = n; r->k = nrhs; r->A = A; r->B = B;
  r->out_elems = (int64_t)n * nrhs;
  r->direct = n > 256;
  return submit(s, r, id);
}

static void fill_result(const bzg::Request* r, bzg_request_result* out) {
  out->status = r->status;
  out->kind = r->kind;
  out->info = r->info;
  out->iterations = r->iter;
  out->result = r->out;
  out->result_elems = r->out_elems;
  out->t_submit_us = r->t_submit;
  out->t_admit_us = r->t_admit;
  out->t_done_us = r->t_done;
}

bzg_status bzg_poll(bzg_server* s, int64_t id, int32_t* done) {
  BZG_REQUIRE(s && done, BZG_ERR_INVALID_ARGUMENT, "poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  BZG_REQUIRE(it != s->reqs.end(), BZG_ERR_INVALID_ARGUMENT, "poll: unknown request");
  *done = it->second->done ? 1 : 0;
  return BZG_OK;
}

bzg_status bzg_server_hold(bzg_server* s, int32_t hold) {
  BZG_REQUIRE(s != nullptr, BZG_ERR_INVALID_ARGUMENT, "server_hold: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->hold = hold != 0;
  }
  s->cv_work.notify_one();
  return BZG_OK;
}

bzg_status bzg_wait(bzg_server* s, int64_t id, bzg_request_result* out) {
  BZG_REQUIRE(s && out, BZG_ERR_INVALID_ARGUMENT, "wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  BZG_REQUIRE(it != s->reqs.end(), BZG_ERR_INVALID_ARGUMENT, "wait: unknown request");
  bzg::Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->done || s->fatal != BZG_OK; });
  if (!r->done) return s->fatal;
  fill_result(r, out);
  return BZG_OK;
}

bzg_status bzg_release(bzg_server* s, int64_t id) {
  BZG_REQUIRE(s != nullptr, BZG_ERR_INVALID_ARGUMENT, "release: null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    BZG_REQUIRE(it != s->reqs.end(), BZG_ERR_INVALID_ARGUMENT, "release: unknown request");
    bzg::Request* r = it->second;
    BZG_REQUIRE(r->done, BZG_ERR_BUSY, "release: request still running");
    if (r->kind == bzg::kSolve) {
      if (r->group && --r->group->live == 0) bzg::release_group_pool(s, r->group);
    } else {
      s->dev.put(r->out, r->out_cls);
      if (r->t_admit > 0) s->used -= bzg::request_bytes(s, r);
    }
    s->reqs.erase(it);
    s->st.released++;
    delete r;
  }
  s->cv_work.notify_one();
  return BZG_OK;
}

bzg_status bzg_server_get_stats(bzg_server* s, bzg_server_stats* out) {
  BZG_REQUIRE(s && out, BZG_ERR_INVALID_ARGUMENT, "server_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->bytes_in_use = (int64_t)s->used;
  out->bytes_peak = (int64_t)s->peak;
  return BZG_OK;
}

}  // extern "C"
