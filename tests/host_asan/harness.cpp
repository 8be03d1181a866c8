// Drives the C ABI of libmlpcg_hip.so through the host-only build (fake_hip.cpp +
// fake_kernels.cpp, AddressSanitizer and LeakSanitizer on): plan and solver lifecycles in
// the orders the test suite and the Python wrappers use them, every solver method and
// layout, graph replay across plan replacement, every building-block entry point and the
// error paths.  Exits 0 when every check holds; ASan aborts on a bad access.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fake_hip.hpp"
#include "mlpcg.h"

namespace {

int g_fail = 0;
#define EXPECT(cond, ...)                           \
  do {                                              \
    if (!(cond)) {                                  \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);            \
      std::fprintf(stderr, "\n");                   \
      ++g_fail;                                     \
    }                                               \
  } while (0)
#define OK(call) EXPECT((call) == 0, "%s -> %s", #call, mlpcg_last_error())

void* dalloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) std::abort();
  std::memset(p, 0, bytes);
  return p;
}
void dfree(void* p) { (void)hipFree(p); }

size_t esz(int dt) { return dt == MLPCG_F64 ? 8 : (dt == MLPCG_F32 ? 4 : 2); }
size_t accsz(int dt) { return dt == MLPCG_F64 ? 8 : 4; }

// Scene: walls, a solid ball, air above 3/4 of the height, fluid elsewhere
// (flag codes of lib/read_data.py:13-16).
struct Scene {
  int n1, n2, n3;
  long V = 0, nf = 0;
  std::vector<int8_t> flags;
  std::vector<int> fluid, rowptr, col;
  std::vector<float> val;
  int8_t* d_flags = nullptr;
  int* d_fluid = nullptr;
  int* d_rowptr = nullptr;
  int* d_col = nullptr;
  float* d_val = nullptr;

  Scene(int a, int b, int c) : n1(a), n2(b), n3(c) {
    V = (long)n1 * n2 * n3;
    flags.assign(V, 2);
    std::vector<int> id(V, -1);
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n2; ++j)
        for (int k = 0; k < n3; ++k) {
          const long v = ((long)i * n2 + j) * n3 + k;
          const bool wall = i == 0 || j == 0 || k == 0 || i == n1 - 1 || j == n2 - 1 || k == n3 - 1;
          const double di = i - n1 / 2.0, dj = j - n2 / 2.0, dk = k - n3 / 3.0;
          const double r = n1 / 6.0;
          if (wall || di * di + dj * dj + dk * dk < r * r) flags[v] = 1;
          else if (k > 3 * n3 / 4) flags[v] = 3;
          if (flags[v] == 2) {
            id[v] = (int)fluid.size();
            fluid.push_back((int)v);
          }
        }
    nf = (long)fluid.size();
    rowptr.push_back(0);
    for (long f = 0; f < nf; ++f) {
      const long v = fluid[f];
      const int i = (int)(v / ((long)n2 * n3)), j = (int)((v / n3) % n2), k = (int)(v % n3);
      const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
      int diag = 0;
      std::vector<std::pair<int, float>> row;
      for (auto& q : nb) {
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= n1 || q[1] >= n2 || q[2] >= n3) continue;
        const long u = ((long)q[0] * n2 + q[1]) * n3 + q[2];
        if (flags[u] != 1) ++diag;
        if (flags[u] == 2) row.push_back({id[u], -1.f});
      }
      row.push_back({(int)f, (float)diag});
      for (auto& e : row) {
        col.push_back(e.first);
        val.push_back(e.second);
      }
      rowptr.push_back((int)col.size());
    }
    d_flags = (int8_t*)dalloc(V);
    std::memcpy(d_flags, flags.data(), V);
    d_fluid = (int*)dalloc(4 * nf);
    std::memcpy(d_fluid, fluid.data(), 4 * nf);
    d_rowptr = (int*)dalloc(4 * rowptr.size());
    std::memcpy(d_rowptr, rowptr.data(), 4 * rowptr.size());
    d_col = (int*)dalloc(4 * col.size());
    std::memcpy(d_col, col.data(), 4 * col.size());
    d_val = (float*)dalloc(4 * val.size());
    std::memcpy(d_val, val.data(), 4 * val.size());
  }
  ~Scene() {
    dfree(d_flags);
    dfree(d_fluid);
    dfree(d_rowptr);
    dfree(d_col);
    dfree(d_val);
  }
};

// Weights of a plan in state_dict order (device buffers of the accumulate type).
struct Weights {
  std::vector<void*> w, b;
  Weights(int kind, int n, int C, int dtype) {
    const int nblk = kind == MLPCG_MODEL_SMALL_SPD ? n + 1 : 2 * n + 1;
    const size_t a = accsz(dtype);
    for (int i = 0; i < nblk; ++i) {
      w.push_back(dalloc(27 * 27 * C * a));
      b.push_back(dalloc(27 * a));
    }
    for (int i = 0; i < 2 * n; ++i) {
      w.push_back(dalloc(27 * C * a));
      b.push_back(dalloc(a));
    }
  }
  ~Weights() {
    for (void* p : w) dfree(p);
    for (void* p : b) dfree(p);
  }
};

mlpcg_precond* make_plan(int kind, int n, int C, int g1, int g2, int g3, int dtype, int nearest, int mask,
                         int swap, const Weights& W, void** image_out, int fused_mode = -1, int tail = 1) {
  const long V = (long)g1 * g2 * g3;
  void* img = dalloc(C * V * accsz(dtype));
  mlpcg_precond* P = nullptr;
  const int rc = fused_mode < 0 ? mlpcg_precond_create(&P, kind, n, C, g1, g2, g3, dtype, nearest, mask, swap, img,
                                                       W.w.data(), W.b.data(), nullptr)
                                : mlpcg_precond_create_ex(&P, kind, n, C, g1, g2, g3, dtype, nearest, mask, swap,
                                                          fused_mode, tail, img, W.w.data(), W.b.data(), nullptr);
  EXPECT(rc == 0 && P, "precond_create kind %d n %d C %d dtype %d grid %dx%dx%d: %s", kind, n, C, dtype, g1, g2,
         g3, mlpcg_last_error());
  *image_out = img;
  return P;
}

// One case of the preconditioned-history test: plan, solver, one solve with a history,
// destroy (solver first or plan first).
void history_case(const Scene& sc, int csr, int method, bool plan_first) {
  Weights W(MLPCG_MODEL_SMALL_SM, 3, 1, MLPCG_F32);
  void* img = nullptr;
  mlpcg_precond* P = make_plan(MLPCG_MODEL_SMALL_SM, 3, 1, sc.n1, sc.n2, sc.n3, MLPCG_F32, 0, 0, 0, W, &img);
  mlpcg_solver* S = nullptr;
  if (csr)
    OK(mlpcg_solver_create_csr(&S, sc.d_rowptr, sc.d_col, sc.d_val, sc.nf, sc.d_fluid, sc.n1, sc.n2, sc.n3,
                               nullptr));
  else
    OK(mlpcg_solver_create_flags(&S, sc.d_flags, sc.n1, sc.n2, sc.n3, nullptr));
  const long n = csr ? sc.nf : sc.V;
  float* b = (float*)dalloc(4 * n);
  float* x = (float*)dalloc(4 * n);
  const int it = 8;
  std::vector<double> hist(it + 1, -1.0);
  mlpcg_solve_info info;
  OK(mlpcg_solver_solve(S, P, b, nullptr, x, method, it, 1e-12, 1e-12, hist.data(), &info, nullptr));
  EXPECT(info.iters == it, "history case csr %d method %d: %d iterations", csr, method, info.iters);
  for (int k = 0; k <= info.iters; ++k) EXPECT(hist[k] > 0.0, "history slot %d not written", k);
  if (plan_first) {
    mlpcg_precond_destroy(P);
    mlpcg_solver_destroy(S);
  } else {
    mlpcg_solver_destroy(S);
    mlpcg_precond_destroy(P);
  }
  dfree(b);
  dfree(x);
  dfree(img);
}

void solver_cases(const Scene& sc) {
  // the order of tests/test_gpu_solver.py::test_preconditioned_history_vs_oracle
  for (int csr = 0; csr < 2; ++csr)
    for (int method = 0; method < 3; ++method)
      for (int pf = 0; pf < 2; ++pf) history_case(sc, csr, method, pf != 0);

  // reuse of one solver: growing and shrinking max_it with and without history, x0,
  // no preconditioner, a plan of every storage type
  for (int csr = 0; csr < 2; ++csr) {
    mlpcg_solver* S = nullptr;
    if (csr)
      OK(mlpcg_solver_create_csr(&S, sc.d_rowptr, sc.d_col, sc.d_val, sc.nf, sc.d_fluid, sc.n1, sc.n2, sc.n3,
                                 nullptr));
    else
      OK(mlpcg_solver_create_flags(&S, sc.d_flags, sc.n1, sc.n2, sc.n3, nullptr));
    const long n = csr ? sc.nf : sc.V;
    float* b = (float*)dalloc(4 * n);
    float* x = (float*)dalloc(4 * n);
    float* x0 = (float*)dalloc(4 * n);
    for (int dtype : {MLPCG_F32, MLPCG_F16, MLPCG_BF16}) {
      Weights W(MLPCG_MODEL_SMALL_SM, 3, 1, dtype);
      void* img = nullptr;
      mlpcg_precond* P = make_plan(MLPCG_MODEL_SMALL_SM, 3, 1, sc.n1, sc.n2, sc.n3, dtype, 0, 0, 0, W, &img);
      for (int method = 0; method < 3; ++method)
        for (int max_it : {5, 40, 3, 1, 64}) {
          std::vector<double> hist(max_it + 1, -1.0);
          mlpcg_solve_info info;
          const bool with_hist = (max_it % 2) == 1;
          OK(mlpcg_solver_solve(S, (max_it == 3) ? nullptr : P, b, (max_it == 40) ? x0 : nullptr, x, method,
                                max_it, 1e-3, 1e-12, with_hist ? hist.data() : nullptr, &info, nullptr));
          EXPECT(info.iters >= 1 && info.iters <= max_it, "iters %d out of [1, %d]", info.iters, max_it);
        }
      mlpcg_precond_destroy(P);
      dfree(img);
    }
    mlpcg_solver_destroy(S);
    dfree(b);
    dfree(x);
    dfree(x0);
  }
}

// Graph replay with a plan that is destroyed and replaced between solves: the solver must
// not replay iterations captured with the freed plan.
void graph_cases(const Scene& sc) {
  for (int csr = 0; csr < 2; ++csr) {
    mlpcg_solver* S = nullptr;
    if (csr)
      OK(mlpcg_solver_create_csr(&S, sc.d_rowptr, sc.d_col, sc.d_val, sc.nf, sc.d_fluid, sc.n1, sc.n2, sc.n3,
                                 nullptr));
    else
      OK(mlpcg_solver_create_flags(&S, sc.d_flags, sc.n1, sc.n2, sc.n3, nullptr));
    OK(mlpcg_solver_set_graphs(S, 1));
    const long n = csr ? sc.nf : sc.V;
    float* b = (float*)dalloc(4 * n);
    float* b2 = (float*)dalloc(4 * n);
    float* x = (float*)dalloc(4 * n);
    for (int round = 0; round < 4; ++round) {
      Weights W(MLPCG_MODEL_SMALL_SM, 3, 1, MLPCG_F32);
      void* img = nullptr;
      mlpcg_precond* P = make_plan(MLPCG_MODEL_SMALL_SM, 3, 1, sc.n1, sc.n2, sc.n3, MLPCG_F32, 0, 0, 0, W, &img);
      // the first graphed solve with a new plan captures its iterations afresh, even when
      // the new plan happens to sit at the address of the one destroyed before it
      const long cap0 = fakehip::captures();
      for (int method : {MLPCG_NPCG_FLEX, MLPCG_NPCG, MLPCG_NPSD, MLPCG_NPCG_FLEX})
        for (float* rhs : {b, b, b2}) {
          mlpcg_solve_info info;
          OK(mlpcg_solver_solve(S, P, rhs, nullptr, x, method, 12, 1e-12, 1e-12, nullptr, &info, nullptr));
          if (method == MLPCG_NPCG_FLEX && rhs == b && fakehip::captures() == cap0)
            EXPECT(false, "graphed solve with a new plan replayed graphs captured before it (round %d)", round);
        }
      // replace the plan by a new one straight away (its Precond often reuses the freed
      // address when the allocator does not hold freed memory back)
      const long cap1 = fakehip::captures();
      mlpcg_precond_destroy(P);
      P = nullptr;
      OK(mlpcg_precond_create(&P, MLPCG_MODEL_SMALL_SM, 3, 1, sc.n1, sc.n2, sc.n3, MLPCG_F32, 0, 0, 0, img,
                              W.w.data(), W.b.data(), nullptr));
      mlpcg_solve_info info;
      OK(mlpcg_solver_solve(S, P, b2, nullptr, x, MLPCG_NPCG_FLEX, 12, 1e-12, 1e-12, nullptr, &info, nullptr));
      EXPECT(fakehip::captures() > cap1, "graphed solve with a replaced plan replayed stale graphs (round %d)",
             round);
      mlpcg_precond_destroy(P);
      dfree(img);
    }
    OK(mlpcg_solver_set_graphs(S, 0));
    mlpcg_solver_destroy(S);
    dfree(b);
    dfree(b2);
    dfree(x);
  }
}

void plan_cases() {
  struct G {
    int n, g1, g2, g3;
  };
  // 32^3 l=3 (test suite), 64^3 l=4 (fused level 0 for BF16, coarse tail), 16x16x64 l=2
  // (fused on every level with mode 2), 8^3 l=1
  const G grids[] = {{3, 32, 32, 32}, {4, 64, 64, 64}, {2, 16, 16, 64}, {1, 8, 8, 8}};
  for (int fused_mode : {1, 2, 0})
    for (int tail : {1, 0}) {
      for (const G& g : grids)
        for (int kind = 0; kind < 3; ++kind)
          for (int dtype : {MLPCG_F64, MLPCG_F32, MLPCG_F16, MLPCG_BF16})
            for (int C : {1, 2, 3}) {
              if (g.n == 4 && (C != 1 || fused_mode == 0)) continue;  // keep the run short
              const int mask = (kind == 0 && C >= 2) ? 1 : 0;
              for (int swap = 0; swap < (kind == 0 ? 2 : 1); ++swap) {
                Weights W(kind, g.n, C, dtype);
                void* img = nullptr;
                mlpcg_precond* P =
                    make_plan(kind, g.n, C, g.g1, g.g2, g.g3, dtype, swap, mask, swap, W, &img, fused_mode, tail);
                if (!P) continue;
                const int nfused = mlpcg_precond_fused_blocks(P);
                EXPECT(nfused >= 0 && (nfused == 0 || (dtype == MLPCG_BF16 && fused_mode > 0)),
                       "fused blocks %d in a plan of dtype %d, mode %d, C %d", nfused, dtype, fused_mode, C);
                EXPECT(!mlpcg_precond_coarse_tail(P) || (tail && kind == 0 && !swap),
                       "coarse tail on a plan that did not ask for it");
                const long V = (long)g.g1 * g.g2 * g.g3;
                const size_t vs = dtype == MLPCG_F64 ? 8 : 4;
                void* r = dalloc(V * vs);
                void* z = dalloc(V * vs);
                uint8_t* om = (uint8_t*)dalloc(V);
                double* sc = (double*)dalloc(8);
                OK(mlpcg_precond_apply(P, r, z, sc, om, nullptr));
                OK(mlpcg_precond_apply(P, r, z, nullptr, nullptr, nullptr));
                OK(mlpcg_precond_update(P, img, W.w.data(), W.b.data(), nullptr));
                OK(mlpcg_precond_apply(P, r, z, sc, nullptr, nullptr));
                EXPECT(mlpcg_precond_bytes(P) > 0, "plan bytes");
                // every non-transposed block alone, in the plan's storage and vector types
                const int nb = kind == MLPCG_MODEL_SMALL_SPD ? g.n + 1 : 2 * g.n + 1;
                for (int k = 0; k < nb; ++k) {
                  const int lvl = (k == nb - 1) ? g.n : (k % g.n);
                  const long Vl = V >> (3 * lvl);
                  void* xb = dalloc(Vl * 8);
                  void* yb = dalloc(Vl * 8);
                  const int vt = dtype == MLPCG_F64 ? MLPCG_F64 : MLPCG_F32;
                  OK(mlpcg_precond_apply_block(P, k, xb, dtype, yb, dtype, nullptr, nullptr));
                  OK(mlpcg_precond_apply_block(P, k, xb, vt, yb, dtype, sc, nullptr));
                  dfree(xb);
                  dfree(yb);
                }
                EXPECT(mlpcg_precond_apply_block(P, 99, r, dtype, z, dtype, nullptr, nullptr) != 0,
                       "apply_block accepted a block out of range");
                mlpcg_precond_destroy(P);
                dfree(r);
                dfree(z);
                dfree(om);
                dfree(sc);
                dfree(img);
              }
            }
    }
}

void block_cases() {
  const int g1 = 8, g2 = 8, g3 = 16;
  const long V = (long)g1 * g2 * g3;
  for (int dtype : {MLPCG_F64, MLPCG_F32, MLPCG_F16, MLPCG_BF16})
    for (int C : {1, 2, 3})
      for (int B : {1, 3}) {
        const size_t es = esz(dtype), as = accsz(dtype);
        void* img = dalloc(C * V * es);
        void* x = dalloc(B * V * es);
        void* y = dalloc(B * V * es);
        void* g = dalloc(B * V * es);
        void* gx = dalloc(B * V * es);
        void* W = dalloc(27 * 27 * C * as);
        void* bias = dalloc(27 * as);
        void* gW = dalloc(27 * 27 * C * as);
        void* gb = dalloc(27 * as);
        void* ws = dalloc(mlpcg_sm_block_workspace_size(C, B, g1, g2, g3, dtype));
        OK(mlpcg_sm_block_inference(img, x, W, bias, y, ws, C, B, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_block_forward(img, x, W, bias, y, ws, C, B, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_block_backward(g, img, x, W, bias, gx, gW, gb, ws, C, B, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_block_trans_inference(img, x, W, bias, y, ws, C, B, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_block_trans_forward(img, x, W, bias, y, ws, C, B, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_block_trans_backward(g, img, x, W, bias, gx, gW, gb, ws, C, B, g1, g2, g3, dtype, nullptr));
        void* lw = dalloc(27 * C * as);
        void* lb = dalloc(as);
        void* z = dalloc(as);
        void* ybar = dalloc(27 * C * as);
        void* lws = dalloc(mlpcg_sm_linear_workspace_size(C, g1, g2, g3, dtype));
        OK(mlpcg_sm_linear_inference(img, lw, lb, z, ybar, lws, C, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_linear_forward(img, lw, lb, z, ybar, lws, C, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_linear_backward(z, ybar, lw, lb, C, dtype, nullptr));
        OK(mlpcg_sm_linear_inference_noalloc(img, lw, lb, z, ybar, C, g1, g2, g3, dtype, nullptr));
        OK(mlpcg_sm_linear_inference_noalloc(img, lw, lb, z, ybar, C, g1 / 2, g2, g3 / 2, dtype, nullptr));
        EXPECT(mlpcg_sm_linear_inference_noalloc(img, lw, lb, z, ybar, 4, g1, g2, g3, dtype, nullptr) != 0,
               "noalloc accepted C = 4");
        for (void* p : {img, x, y, g, gx, W, bias, gW, gb, ws, lw, lb, z, ybar, lws}) dfree(p);
      }
  // streamed-stencil building blocks
  for (int dtype : {MLPCG_F32, MLPCG_BF16}) {
    const size_t es = esz(dtype);
    void* img = dalloc(V * 4);
    void* W = dalloc(27 * 27 * 4);
    void* bias = dalloc(27 * 4);
    void* K = dalloc(27 * V * es);
    OK(mlpcg_stencil_weights(img, MLPCG_F32, W, bias, K, dtype, 0, 1, g1, g2, g3, nullptr));
    OK(mlpcg_stencil_weights(img, MLPCG_F32, W, bias, K, dtype, 1, 1, g1, g2, g3, nullptr));
    void* x = dalloc(2 * V * 4);
    void* y = dalloc(2 * V * 4);
    void* prev = dalloc(2 * V * 4);
    void* c = dalloc(8);
    OK(mlpcg_sm_apply(K, dtype, x, dtype, y, dtype, prev, c, c, nullptr, nullptr, nullptr, 2, g1, g2, g3, nullptr));
    OK(mlpcg_avgpool2(x, y, 2, g1, g2, g3, dtype, nullptr));
    OK(mlpcg_upsample2(x, y, 2, g1, g2, g3, 0, 1.0, prev, c, c, dtype, nullptr));
    OK(mlpcg_upsample2_adjoint(x, y, 2, g1 / 2, g2 / 2, g3 / 2, 1, 0.125, dtype, nullptr));
    for (void* p : {img, W, bias, K, x, y, prev, c}) dfree(p);
  }
  // Laplacian and dot
  Scene sc(8, 8, 16);
  uint8_t* code = (uint8_t*)dalloc(sc.V);
  float* xv = (float*)dalloc(4 * sc.V);
  float* yv = (float*)dalloc(4 * sc.V);
  OK(mlpcg_flags_code(sc.d_flags, code, sc.n1, sc.n2, sc.n3, nullptr));
  OK(mlpcg_spmv_flags(code, xv, yv, sc.n1, sc.n2, sc.n3, nullptr));
  OK(mlpcg_spmv_csr(sc.d_rowptr, sc.d_col, sc.d_val, xv, yv, sc.nf, MLPCG_F32, nullptr));
  void* dws = dalloc(mlpcg_dot_workspace_size());
  double* out = (double*)dalloc(8);
  OK(mlpcg_dot(xv, yv, sc.V, out, dws, nullptr));
  // Lanczos vector kernels: a CSR FP64 product, dots, the three-term update, normalising
  {
    const long n = sc.nf;
    double* vd = (double*)dalloc(8 * sc.val.size());
    double* a = (double*)dalloc(8 * n);
    double* b = (double*)dalloc(8 * n);
    double* w = (double*)dalloc(8 * n);
    double* s = (double*)dalloc(4 * 8);
    void* fws = dalloc(mlpcg_dot_f64_workspace_size());
    OK(mlpcg_spmv_csr(sc.d_rowptr, sc.d_col, vd, a, w, n, MLPCG_F64, nullptr));
    OK(mlpcg_dot_f64(w, a, n, s, fws, nullptr));
    OK(mlpcg_lanczos_update(w, a, nullptr, s, nullptr, n, nullptr));
    OK(mlpcg_lanczos_update(w, a, b, s, s + 1, n, nullptr));
    OK(mlpcg_dot_f64(w, w, n, s + 2, fws, nullptr));
    OK(mlpcg_normalize_f64(b, w, s + 2, s + 3, n, nullptr));
    OK(mlpcg_normalize_f64(b, w, s + 2, nullptr, n, nullptr));
    EXPECT(mlpcg_dot_f64(w, a, n, nullptr, fws, nullptr) != 0, "dot_f64 accepted a null output");
    EXPECT(mlpcg_lanczos_update(w, a, b, s, nullptr, n, nullptr) != 0, "update accepted v0 without beta");
    for (void* p : {(void*)vd, (void*)a, (void*)b, (void*)w, (void*)s, fws}) dfree(p);
  }
  for (void* p : {(void*)code, (void*)xv, (void*)yv, dws, (void*)out}) dfree(p);
}

void error_cases(const Scene& sc) {
  mlpcg_precond* P = nullptr;
  mlpcg_solver* S = nullptr;
  Weights W(MLPCG_MODEL_SMALL_SM, 3, 1, MLPCG_F32);
  void* img = dalloc(sc.V * 4);
  // bad plan arguments: each returns -1 with a message and allocates nothing that outlives it
  EXPECT(mlpcg_precond_create(nullptr, 0, 3, 1, 32, 32, 32, MLPCG_F32, 0, 0, 0, img, W.w.data(), W.b.data(),
                              nullptr) != 0, "null handle accepted");
  EXPECT(mlpcg_precond_create(&P, 7, 3, 1, 32, 32, 32, MLPCG_F32, 0, 0, 0, img, W.w.data(), W.b.data(),
                              nullptr) != 0, "bad kind accepted");
  EXPECT(mlpcg_precond_create(&P, 0, 3, 4, 32, 32, 32, MLPCG_F32, 0, 0, 0, img, W.w.data(), W.b.data(),
                              nullptr) != 0, "C = 4 accepted");
  EXPECT(mlpcg_precond_create(&P, 0, 3, 1, 30, 32, 32, MLPCG_F32, 0, 0, 0, img, W.w.data(), W.b.data(),
                              nullptr) != 0, "indivisible grid accepted");
  EXPECT(mlpcg_precond_create(&P, 0, 3, 1, 32, 32, 32, 9, 0, 0, 0, img, W.w.data(), W.b.data(), nullptr) != 0,
         "bad dtype accepted");
  EXPECT(mlpcg_precond_create(&P, 0, 3, 1, 32, 32, 32, MLPCG_F32, 0, 1, 0, img, W.w.data(), W.b.data(),
                              nullptr) != 0, "mask with C = 1 accepted");
  EXPECT(mlpcg_precond_create(&P, 0, 3, 1, 32, 32, 32, MLPCG_F32, 0, 0, 0, nullptr, W.w.data(), W.b.data(),
                              nullptr) != 0, "null image accepted");
  EXPECT(mlpcg_precond_create_ex(&P, 0, 3, 1, 32, 32, 32, MLPCG_BF16, 0, 0, 0, 3, 1, img, W.w.data(), W.b.data(),
                                 nullptr) != 0, "fused mode 3 accepted");
  EXPECT(mlpcg_precond_fused_blocks(nullptr) == -1 && mlpcg_precond_coarse_tail(nullptr) == -1, "null plan query");
  EXPECT(std::strlen(mlpcg_last_error()) > 0, "no error message");
  EXPECT(P == nullptr, "a failed create wrote the handle");
  EXPECT(mlpcg_precond_update(nullptr, img, W.w.data(), W.b.data(), nullptr) != 0, "null plan accepted");
  EXPECT(mlpcg_precond_apply(nullptr, img, img, nullptr, nullptr, nullptr) != 0, "null plan accepted");
  EXPECT(mlpcg_solver_create_flags(nullptr, sc.d_flags, sc.n1, sc.n2, sc.n3, nullptr) != 0, "null handle");
  EXPECT(mlpcg_solver_create_flags(&S, nullptr, sc.n1, sc.n2, sc.n3, nullptr) != 0, "null flags");
  EXPECT(mlpcg_solver_create_flags(&S, sc.d_flags, 0, sc.n2, sc.n3, nullptr) != 0, "empty grid");
  EXPECT(mlpcg_solver_create_csr(&S, sc.d_rowptr, sc.d_col, sc.d_val, 0, sc.d_fluid, sc.n1, sc.n2, sc.n3,
                                 nullptr) != 0, "empty CSR system");
  EXPECT(S == nullptr, "a failed create wrote the handle");
  OK(mlpcg_solver_create_flags(&S, sc.d_flags, sc.n1, sc.n2, sc.n3, nullptr));
  float* b = (float*)dalloc(sc.V * 4);
  float* x = (float*)dalloc(sc.V * 4);
  mlpcg_solve_info info;
  EXPECT(mlpcg_solver_solve(S, nullptr, b, nullptr, x, 5, 10, 1e-6, 0, nullptr, &info, nullptr) != 0, "method");
  EXPECT(mlpcg_solver_solve(S, nullptr, b, nullptr, x, 0, 0, 1e-6, 0, nullptr, &info, nullptr) != 0, "max_it 0");
  EXPECT(mlpcg_solver_solve(S, nullptr, nullptr, nullptr, x, 0, 10, 1e-6, 0, nullptr, &info, nullptr) != 0,
         "null b");
  // a plan on another grid, and an FP64 plan
  Weights W2(MLPCG_MODEL_SMALL_SM, 2, 1, MLPCG_F32);
  void* img2 = nullptr;
  P = make_plan(MLPCG_MODEL_SMALL_SM, 2, 1, 16, 16, 16, MLPCG_F32, 0, 0, 0, W2, &img2);
  EXPECT(mlpcg_solver_solve(S, P, b, nullptr, x, 0, 10, 1e-6, 0, nullptr, &info, nullptr) != 0, "grid mismatch");
  mlpcg_precond_destroy(P);
  Weights W3(MLPCG_MODEL_SMALL_SM, 3, 1, MLPCG_F64);
  void* img3 = nullptr;
  P = make_plan(MLPCG_MODEL_SMALL_SM, 3, 1, sc.n1, sc.n2, sc.n3, MLPCG_F64, 0, 0, 0, W3, &img3);
  EXPECT(mlpcg_solver_solve(S, P, b, nullptr, x, 0, 10, 1e-6, 0, nullptr, &info, nullptr) != 0, "FP64 plan");
  mlpcg_precond_destroy(P);
  mlpcg_precond_destroy(nullptr);
  mlpcg_solver_destroy(nullptr);
  mlpcg_solver_destroy(S);
  for (void* p : {img, img2, img3, (void*)b, (void*)x}) dfree(p);
}

}  // namespace

int main() {
  {
    Scene sc(32, 32, 32);
    std::printf("scene 32^3: %ld fluid cells\n", sc.nf);
    solver_cases(sc);
    std::printf("solver cases done\n");
    graph_cases(sc);
    std::printf("graph cases done\n");
    error_cases(sc);
    std::printf("error cases done\n");
  }
  plan_cases();
  std::printf("plan cases done\n");
  block_cases();
  std::printf("block cases done\n");
  EXPECT(fakehip::live_device_allocations() == 0, "%ld device allocations leaked",
         fakehip::live_device_allocations());
  EXPECT(fakehip::live_streams() == 0, "%ld streams leaked", fakehip::live_streams());
  EXPECT(fakehip::live_events() == 0, "%ld events leaked", fakehip::live_events());
  EXPECT(fakehip::live_graphs() == 0, "%ld graphs leaked", fakehip::live_graphs());
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "ALL HOST CHECKS PASSED", g_fail);
  return g_fail ? 1 : 0;
}
