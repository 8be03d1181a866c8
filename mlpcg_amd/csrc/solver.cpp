// Native PCG driver (npcg / npcg_flex / npsd of cg_tests.py) over either the full grid
// with the matrix-free Laplacian or CSR over the compressed unknowns.
//
// One iteration is a fixed sequence of kernels on the caller's stream; the scalars
// (rho, beta, alpha, ||r||, the stop decision) live in a device SolveState updated by the
// one-block finalize kernels of the reductions.  The host runs up to kLookahead iterations ahead of the GPU:
// after queueing iteration k it waits only for the status of iteration k - kLookahead,
// copied on a side stream.  Once the device sets `done`, the kernels of already-queued
// iterations return immediately, so the iteration count and the result are exactly
// those of a loop that checks after every iteration (cg_tests.py:70, :150, :274).
#include <cmath>
#include <cstring>

#include "runtime.hpp"

namespace mlpcg {

constexpr int kLookahead = 2;

Solver::~Solver() {
  drop_graphs();
  if (work) (void)hipStreamSynchronize(work);
  if (work) (void)hipStreamDestroy(work);
  if (ev_in) (void)hipEventDestroy(ev_in);
  if (side) (void)hipStreamSynchronize(side);
  for (int i = 0; i < 4; ++i) {
    if (ev_iter[i]) (void)hipEventDestroy(ev_iter[i]);
    if (ev_stat[i]) (void)hipEventDestroy(ev_stat[i]);
  }
  if (side) (void)hipStreamDestroy(side);
  if (st_host) (void)hipHostFree(st_host);
  for (void* p : allocs) (void)hipFree(p);
}

static int salloc(Solver* S, void** out, size_t bytes, bool zero, hipStream_t st) {
  if (bytes == 0) bytes = 16;
  MLPCG_HIP(hipMalloc(out, bytes));
  S->allocs.push_back(*out);
  if (zero) MLPCG_HIP(hipMemsetAsync(*out, 0, bytes, st));
  return 0;
}

static int solver_common(Solver* S, hipStream_t st) {
  const size_t vb = (size_t)S->n * sizeof(float);
  for (int i = 0; i < 2; ++i) {
    if (salloc(S, (void**)&S->x[i], vb, true, st)) return -1;
    if (salloc(S, (void**)&S->r[i], vb, true, st)) return -1;
    if (salloc(S, (void**)&S->p[i], vb, true, st)) return -1;
    if (salloc(S, (void**)&S->aq_hist[i], vb, true, st)) return -1;
  }
  for (int i = 0; i < 3; ++i)
    if (salloc(S, (void**)&S->q_hist[i], vb, true, st)) return -1;
  if (salloc(S, (void**)&S->z, vb, true, st)) return -1;
  if (salloc(S, (void**)&S->w, vb, true, st)) return -1;
  if (salloc(S, (void**)&S->st, sizeof(SolveState), true, st)) return -1;
  if (salloc(S, (void**)&S->partials, reduction_workspace_bytes(), true, st)) return -1;
  MLPCG_HIP(hipHostMalloc((void**)&S->st_host, 4 * sizeof(SolveState), hipHostMallocDefault));
  memset(S->st_host, 0, 4 * sizeof(SolveState));
  MLPCG_HIP(hipStreamCreateWithFlags(&S->side, hipStreamNonBlocking));
  MLPCG_HIP(hipStreamCreateWithFlags(&S->work, hipStreamNonBlocking));
  MLPCG_HIP(hipEventCreateWithFlags(&S->ev_in, hipEventDisableTiming));
  for (int i = 0; i < 4; ++i) {
    MLPCG_HIP(hipEventCreateWithFlags(&S->ev_iter[i], hipEventDisableTiming));
    MLPCG_HIP(hipEventCreateWithFlags(&S->ev_stat[i], hipEventDisableTiming));
  }
  return 0;
}

int solver_create_flat(Solver** out, const int8_t* flags, int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(out && flags, "solver_create_flat: null argument");
  MLPCG_REQUIRE(n1 > 0 && n2 > 0 && n3 > 0, "solver_create_flat: bad grid");
  Solver* S = new Solver();
  S->layout = 0;
  S->n1 = n1;
  S->n2 = n2;
  S->n3 = n3;
  S->V = S->n = (long)n1 * n2 * n3;
  if (salloc(S, (void**)&S->code, (size_t)S->V, false, st) || solver_common(S, st) ||
      launch_flags_code(flags, S->code, n1, n2, n3, st)) {
    delete S;
    return -1;
  }
  *out = S;
  return 0;
}

int solver_create_csr(Solver** out, const int* rowptr, const int* col, const float* val, long n,
                      const int* fluid_idx, int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(out && rowptr && col && val && fluid_idx, "solver_create_csr: null argument");
  MLPCG_REQUIRE(n > 0, "solver_create_csr: empty system");
  Solver* S = new Solver();
  S->layout = 1;
  S->n1 = n1;
  S->n2 = n2;
  S->n3 = n3;
  S->V = (long)n1 * n2 * n3;
  S->n = n;
  S->rowptr = rowptr;
  S->col = col;
  S->val = val;
  S->fluid_idx = fluid_idx;
  if (solver_common(S, st) || salloc(S, (void**)&S->grid_in, (size_t)S->V * sizeof(float), true, st) ||
      salloc(S, (void**)&S->grid_out, (size_t)S->V * sizeof(float), true, st)) {
    delete S;
    return -1;
  }
  *out = S;
  return 0;
}

int Solver::solve(Precond* P, const float* b, const float* x0, float* x_out, int method, int max_it,
                  double tol, double atol, double* history_out, SolveInfo* info, hipStream_t st) {
  MLPCG_REQUIRE(b && x_out, "solve: null vector");
  MLPCG_REQUIRE(method >= kNpcg && method <= kNpsd, "solve: bad method %d", method);
  MLPCG_REQUIRE(max_it >= 1, "solve: max_it must be >= 1");
  if (P) {
    MLPCG_REQUIRE(P->dims[0][0] == n1 && P->dims[0][1] == n2 && P->dims[0][2] == n3,
                  "solve: preconditioner grid does not match the solver grid");
    MLPCG_REQUIRE(P->dtype != kF64, "solve: FP64 plans are for checking only; use f32/bf16/f16");
  }
  if (history_out && hist_cap < max_it + 1) {
    void* h = nullptr;
    if (salloc(this, &h, (size_t)(max_it + 1) * sizeof(double), false, st)) return -1;
    history = (double*)h;
    hist_cap = max_it + 1;
  }
  // Graph replay runs on the solver's own stream (the legacy default stream cannot be
  // captured); it is ordered after the caller's stream and the solve synchronises it.
  hipStream_t run = st;
  if (use_graphs) {
    MLPCG_HIP(hipEventRecord(ev_in, st));
    MLPCG_HIP(hipStreamWaitEvent(work, ev_in, 0));
    run = work;
    if (graph_P != (P ? P->gen : 0) || graph_b != b || graph_method != method) drop_graphs();
  }
  // initial state
  SolveState init;
  memset(&init, 0, sizeof(init));
  init.tol = tol;
  init.atol = atol;
  init.max_it = max_it;
  init.method = method;
  init.history = history_out ? history : nullptr;
  SolveState* hs = st_host;  // slot 0 carries the initial state up
  MLPCG_HIP(hipStreamSynchronize(side));
  *hs = init;
  MLPCG_HIP(hipMemcpyAsync(this->st, hs, sizeof(SolveState), hipMemcpyHostToDevice, run));
  const size_t vb = (size_t)n * sizeof(float);
  if (x0) MLPCG_HIP(hipMemcpyAsync(x[0], x0, vb, hipMemcpyDeviceToDevice, run));
  else MLPCG_HIP(hipMemsetAsync(x[0], 0, vb, run));
  MLPCG_HIP(hipMemsetAsync(p[0], 0, vb, run));
  MLPCG_HIP(hipMemsetAsync(p[1], 0, vb, run));
  const int* done = &this->st->done;

  FlatArgs fa0;
  memset(&fa0, 0, sizeof(fa0));
  fa0.code = code;
  fa0.b = b;
  fa0.st = this->st;
  fa0.partials = partials;
  fa0.n1 = n1;
  fa0.n2 = n2;
  fa0.n3 = n3;
  CsrArgs ca0;
  memset(&ca0, 0, sizeof(ca0));
  ca0.rowptr = rowptr;
  ca0.col = col;
  ca0.val = val;
  ca0.b = b;
  ca0.st = this->st;
  ca0.partials = partials;
  ca0.n = n;

  int rc = 0;
#define TRY(x)         \
  do {                 \
    rc = (x);          \
    if (rc) return rc; \
  } while (0)
  // r_0 = b - A x_0
  if (layout == 0) {
    FlatArgs fa = fa0;
    fa.x_old = x[0];
    fa.r = r[0];
    TRY(launch_resid_flat(fa, 1, run));
  } else {
    CsrArgs ca = ca0;
    ca.x_old = x[0];
    ca.r = r[0];
    TRY(launch_resid_csr(ca, 1, run));
  }

  // z = P(r); on the flat layout a fused last block also reduces z.dot0 and z.dot1
  // (*nbd partial blocks, 0 if not)
  auto precondition = [&](const float* rin, float* zout, hipStream_t s, const float* dot0 = nullptr,
                          const float* dot1 = nullptr, unsigned* nbd = nullptr) -> int {
    if (nbd) *nbd = 0;
    if (!P) {
      MLPCG_HIP(hipMemcpyAsync(zout, rin, vb, hipMemcpyDeviceToDevice, s));
      return 0;
    }
    const double* scale = &this->st->scale;
    if (layout == 0) return P->apply(rin, zout, scale, code, done, s, dot0, dot1, partials, nbd);
    int rc2 = launch_scatter(fluid_idx, rin, grid_in, n, done, s);
    if (rc2) return rc2;
    rc2 = P->apply(grid_in, grid_out, scale, nullptr, done, s);
    if (rc2) return rc2;
    return launch_gather(fluid_idx, grid_out, zout, n, done, s);
  };

  // One iteration (k counts from 1; cnt = directions created so far, npsd only).
  auto enqueue_iter = [&](int k, int cnt, hipStream_t s) -> int {
    const int cur = (k - 1) & 1, nxt = k & 1;
    const float* rk = r[cur];
    FlatArgs fa = fa0;
    CsrArgs ca = ca0;
    int rc2 = 0;
#define TRY2(x)          \
  do {                   \
    rc2 = (x);           \
    if (rc2) return rc2; \
  } while (0)
    if (method != kNpsd) {
      unsigned nbd = 0;
      TRY2(precondition(rk, z, s, rk, method == kNpcgFlex ? r[nxt] : nullptr, &nbd));
      if (nbd)
        TRY2(launch_finalize_rz(partials, nbd, this->st, s));
      else
        TRY2(launch_rz_dot(rk, z, r[nxt], n, this->st, partials, s));
      if (layout == 0) {
        fa.z = z;
        fa.p_old = p[cur];
        fa.p_new = p[nxt];
        fa.w = w;
        TRY2(launch_p_spmv_flat(fa, s));
        fa.x_old = x[cur];
        fa.x_new = x[nxt];
        fa.p = p[nxt];
        fa.r = r[nxt];
        TRY2(launch_resid_flat(fa, 0, s));
      } else {
        ca.z = z;
        ca.p_old = p[cur];
        ca.p_new = p[nxt];
        ca.w = w;
        TRY2(launch_p_spmv_csr(ca, s));
        ca.x_old = x[cur];
        ca.x_new = x[nxt];
        ca.p = p[nxt];
        ca.r = r[nxt];
        TRY2(launch_resid_csr(ca, 0, s));
      }
      return 0;
    }
    // npsd: q = P(r), projected against the kept directions N (newest) and O; direction
    // d lives in q_hist[d % 3], its A-product and alfa in slot d % 2.
    float* q = z;
    const int kept = cnt < 2 ? cnt : 2;
    const float* pN = kept >= 1 ? q_hist[(cnt - 1) % 3] : nullptr;
    const float* pO = kept >= 2 ? q_hist[(cnt - 2) % 3] : nullptr;
    const float* aN = kept >= 1 ? aq_hist[(cnt - 1) & 1] : nullptr;
    const float* aO = kept >= 2 ? aq_hist[cnt & 1] : nullptr;
    const double* alfaN = &this->st->alfa[(cnt - 1) & 1];
    const double* alfaO = kept >= 2 ? &this->st->alfa[cnt & 1] : nullptr;
    unsigned nbd = 0;
    TRY2(precondition(rk, q, s, aN, aO, &nbd));
    if (kept >= 1) {
      if (nbd)
        TRY2(launch_finalize_proj(partials, nbd, this->st, alfaN, alfaO, s));
      else
        TRY2(launch_npsd_proj(q, aN, aO, alfaN, alfaO, n, this->st, partials, s));
    }
    float* q_new = q_hist[cnt % 3];
    if (layout == 0) {
      fa.z = q;
      fa.p_old = pO;
      fa.p_new = q_new;
      fa.r = const_cast<float*>(rk);
      fa.w = aq_hist[cnt & 1];
      TRY2(launch_npsd_dir_flat(fa, pN, &this->st->alfa[cnt & 1], s));
      fa.x_old = x[cur];
      fa.x_new = x[nxt];
      fa.p = q_new;
      fa.r = r[nxt];
      TRY2(launch_resid_flat(fa, 0, s));
    } else {
      ca.z = q;
      ca.p_old = pO;
      ca.p_new = q_new;
      ca.r = const_cast<float*>(rk);
      ca.w = aq_hist[cnt & 1];
      TRY2(launch_npsd_dir_csr(ca, pN, &this->st->alfa[cnt & 1], s));
      ca.x_old = x[cur];
      ca.x_new = x[nxt];
      ca.p = q_new;
      ca.r = r[nxt];
      TRY2(launch_resid_csr(ca, 0, s));
    }
    return 0;
#undef TRY2
  };

  int k = 0;
  int npsd_count = 0;  // directions created (npsd)
  for (k = 1; k <= max_it; ++k) {
    // npcg / npcg_flex: iterations k and k+2 launch the same kernels with the same
    // arguments (buffer parity k & 1), so two captured graphs cover the solve
    const bool replay = use_graphs && method != kNpsd;
    if (replay) {
      const int par = k & 1;
      if (!graph_exec[par]) {
        hipGraph_t g = nullptr;
        MLPCG_HIP(hipStreamBeginCapture(run, hipStreamCaptureModeThreadLocal));
        const int rc3 = enqueue_iter(k, npsd_count, run);
        const hipError_t ce = hipStreamEndCapture(run, &g);
        if (rc3) {
          if (g) (void)hipGraphDestroy(g);
          return rc3;
        }
        MLPCG_HIP(ce);
        const hipError_t ie = hipGraphInstantiate(&graph_exec[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        MLPCG_HIP(ie);
        graph_P = P ? P->gen : 0;
        graph_b = b;
        graph_method = method;
      }
      MLPCG_HIP(hipGraphLaunch(graph_exec[par], run));
    } else {
      TRY(enqueue_iter(k, npsd_count, run));
    }
    if (method == kNpsd) ++npsd_count;
    // status of iteration k goes to the host on the side stream
    const int e = k & 3;
    MLPCG_HIP(hipEventRecord(ev_iter[e], run));
    MLPCG_HIP(hipStreamWaitEvent(side, ev_iter[e], 0));
    MLPCG_HIP(hipMemcpyAsync(&st_host[e], this->st, sizeof(SolveState), hipMemcpyDeviceToHost, side));
    MLPCG_HIP(hipEventRecord(ev_stat[e], side));
    if (k > kLookahead) {
      const int e2 = (k - kLookahead) & 3;
      MLPCG_HIP(hipEventSynchronize(ev_stat[e2]));
      if (st_host[e2].done) break;
    }
  }
#undef TRY
  // final state
  MLPCG_HIP(hipStreamSynchronize(run));
  MLPCG_HIP(hipStreamSynchronize(side));
  SolveState fin;
  MLPCG_HIP(hipMemcpy(&fin, this->st, sizeof(SolveState), hipMemcpyDeviceToHost));
  const int iters = fin.iter;
  // the count comes from device memory and sizes the copy into the caller's history
  // (max_it + 1 doubles): refuse anything outside [0, max_it] instead of trusting it
  MLPCG_REQUIRE(iters >= 0 && iters <= max_it, "solve: device iteration count %d outside [0, %d]", iters,
                max_it);
  MLPCG_HIP(hipMemcpyAsync(x_out, x[iters & 1], vb, hipMemcpyDeviceToDevice, run));
  if (history_out)
    MLPCG_HIP(hipMemcpy(history_out, history, (size_t)(iters + 1) * sizeof(double), hipMemcpyDeviceToHost));
  MLPCG_HIP(hipStreamSynchronize(run));
  if (info) {
    info->iters = iters;
    info->rel_res = fin.norm;
    info->norm_b = fin.norm_b;
    const double tolv = fmax(tol, atol / fin.norm_b);
    if (!std::isfinite(fin.norm)) info->status = 2;
    else if ((method == kNpsd) ? (fin.norm < tolv) : (fin.norm <= tolv)) info->status = 0;
    else info->status = 1;
  }
  return 0;
}

void Solver::drop_graphs() {
  for (int i = 0; i < 2; ++i) {
    if (graph_exec[i]) (void)hipGraphExecDestroy(graph_exec[i]);
    graph_exec[i] = nullptr;
  }
  graph_P = 0;
  graph_b = nullptr;
  graph_method = -1;
}

}  // namespace mlpcg
