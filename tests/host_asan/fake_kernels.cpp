// Host-build stand-ins for the kernel launchers (the launch_* functions of the .hip
// sources), linked ahead of the real ones.  Each launcher enqueues (fake_hip.cpp) a
// closure that checks, with AddressSanitizer, every device buffer extent the real kernel
// reads or writes -- so a buffer the host code sized too small, or freed while a plan,
// a solver or a captured graph still uses it, is reported -- and emulates the solver
// state of the residual finalize (iteration count, stop flag, history slot), which is
// the only device state the host code reads back.  The sizing helpers
// (reduction_workspace_bytes, fused_blocks, mean_patch_workspace, ...) are the real ones.
#include <sanitizer/asan_interface.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../mlpcg_amd/csrc/kernels.hpp"
#include "../../mlpcg_amd/csrc/pcg.hpp"
#include "fake_hip.hpp"

namespace {

using namespace mlpcg;

// Reports (through ASan, with allocation / free stacks) the first bad byte of a region.
void check_region(const char* who, const void* p, size_t bytes) {
  if (p == nullptr || bytes == 0) return;
  void* bad = __asan_region_is_poisoned(const_cast<void*>(p), bytes);
  if (bad) {
    std::fprintf(stderr, "fake kernel %s: %zu-byte extent at %p reaches a bad byte at %p\n", who, bytes, p,
                 bad);
    volatile char c = *(volatile char*)bad;  // ASan reports this access
    (void)c;
    std::abort();
  }
}
#define RD(p, n) check_region(who, (p), (size_t)(n))
#define WR(p, n) check_region(who, (p), (size_t)(n))

size_t esz(int dt) { return dt == kF64 ? 8 : (dt == kF32 ? 4 : 2); }
size_t accsz(int dt) { return dt == kF64 ? 8 : 4; }
bool done_now(const int* done) { return done != nullptr && *done != 0; }

// block counts of the reduction kernels (pcg.hip: one item per thread, at most 16384
// blocks; flat kernels take runs of vz cells)
unsigned red_blocks(long items) {
  long nb = (items + 255) / 256;
  if (nb > 16384) nb = 16384;
  return (unsigned)(nb < 1 ? 1 : nb);
}
int flat_vz(int n3) { return (n3 % 8 == 0) ? 8 : (n3 % 4 == 0 ? 4 : (n3 % 2 == 0 ? 2 : 1)); }
unsigned flat_blocks(const FlatArgs& a) {
  return red_blocks((long)a.n1 * a.n2 * (a.n3 / flat_vz(a.n3)));
}

void partials_check(const char* who, const double* partials, int rows, unsigned nb) {
  if ((size_t)rows * nb * sizeof(double) > reduction_workspace_bytes()) {
    std::fprintf(stderr, "fake kernel %s: %d x %u partials exceed the reduction workspace\n", who, rows, nb);
    std::abort();
  }
  WR(partials, (size_t)rows * nb * sizeof(double));
}

// residual finalize: r = b - A x was recomputed; the emulated residual halves per iteration
void fin_resid(SolveState* st, int init) {
  const char* who = "finalize";
  WR(st, sizeof(SolveState));
  if (init) {
    st->iter = 0;
    st->done = 0;
    st->norm_b = 1.0;
    st->norm = 1.0;
    st->rr = 1.0;
    st->scale = 1.0;
  } else {
    if (st->done) return;
    st->iter += 1;
    st->norm = std::ldexp(1.0, -st->iter);
    st->rr = st->norm * st->norm;
    st->scale = 1.0 / st->norm;
    const double stop = std::fmax(st->tol, st->atol / st->norm_b);
    const bool cont = (st->method == kNpsd) ? !(st->norm < stop) : (st->norm > stop);
    st->done = (cont && st->iter < st->max_it) ? 0 : 1;
  }
  if (st->history && st->iter <= st->max_it) {
    WR(st->history + st->iter, sizeof(double));
    st->history[st->iter] = st->norm;
  }
}

}  // namespace

namespace mlpcg {

// ------------------------------------------------------------------ PCG side
int launch_flags_code(const int8_t* flags, uint8_t* code, int n1, int n2, int n3, hipStream_t st) {
  const long V = (long)n1 * n2 * n3;
  fakehip::enqueue(st, [=] {
    const char* who = "flags_code";
    RD(flags, V);
    WR(code, V);
    for (long v = 0; v < V; ++v) code[v] = flags[v] == kFlagFluid ? 7 : 0;
  });
  return 0;
}

int launch_resid_flat(const FlatArgs& a, int init, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "resid_flat";
    if (!init && a.st->done) return;
    const long V = (long)a.n1 * a.n2 * a.n3;
    RD(a.code, V);
    RD(a.b, 4 * V);
    RD(a.x_old, 4 * V);
    if (!init) {
      RD(a.p, 4 * V);
      WR(a.x_new, 4 * V);
    }
    WR(a.r, 4 * V);
    partials_check(who, a.partials, 2, flat_blocks(a));
    fin_resid(a.st, init);
  });
  return 0;
}

int launch_p_spmv_flat(const FlatArgs& a, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "p_spmv_flat";
    if (a.st->done) return;
    const long V = (long)a.n1 * a.n2 * a.n3;
    RD(a.code, V);
    RD(a.z, 4 * V);
    RD(a.p_old, 4 * V);
    WR(a.p_new, 4 * V);
    WR(a.w, 4 * V);
    partials_check(who, a.partials, 1, flat_blocks(a));
  });
  return 0;
}

int launch_npsd_dir_flat(const FlatArgs& a, const float* pN, double* alfa_out,
                         hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "npsd_dir_flat";
    if (a.st->done) return;
    const long V = (long)a.n1 * a.n2 * a.n3;
    RD(a.code, V);
    RD(a.z, 4 * V);
    RD(a.r, 4 * V);
    if (pN) {
      RD(pN, 4 * V);
      if (a.p_old) RD(a.p_old, 4 * V);
    }
    WR(a.p_new, 4 * V);
    WR(a.w, 4 * V);
    partials_check(who, a.partials, 3, flat_blocks(a));
    WR(alfa_out, sizeof(double));
    *alfa_out = 1.0;
  });
  return 0;
}

int launch_rz_dot(const float* r, const float* z, const float* r_old, long n, SolveState* s,
                  double* partials, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "rz_dot";
    if (s->done) return;
    RD(r, 4 * n);
    RD(z, 4 * n);
    if (r_old) RD(r_old, 4 * n);
    partials_check(who, partials, 2, red_blocks(n / 16 + 1));
  });
  return 0;
}

int launch_scatter(const int* idx, const float* src, float* grid, long n, const int* done, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "scatter";
    if (done_now(done)) return;
    RD(idx, 4 * n);
    RD(src, 4 * n);
    for (long i = 0; i < n; ++i) WR(grid + idx[i], 4);
  });
  return 0;
}

int launch_gather(const int* idx, const float* grid, float* dst, long n, const int* done, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "gather";
    if (done_now(done)) return;
    RD(idx, 4 * n);
    WR(dst, 4 * n);
    for (long i = 0; i < n; ++i) RD(grid + idx[i], 4);
  });
  return 0;
}

static void csr_check(const char* who, const CsrArgs& a) {
  RD(a.rowptr, 4 * (a.n + 1));
  const long nnz = a.rowptr[a.n];
  RD(a.col, 4 * nnz);
  RD(a.val, 4 * nnz);
}

int launch_p_spmv_csr(const CsrArgs& a, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "p_spmv_csr";
    if (a.st->done) return;
    csr_check(who, a);
    RD(a.z, 4 * a.n);
    RD(a.p_old, 4 * a.n);
    WR(a.p_new, 4 * a.n);
    WR(a.w, 4 * a.n);
    partials_check(who, a.partials, 1, red_blocks(a.n));
  });
  return 0;
}

int launch_resid_csr(const CsrArgs& a, int init, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "resid_csr";
    if (!init && a.st->done) return;
    csr_check(who, a);
    RD(a.b, 4 * a.n);
    RD(a.x_old, 4 * a.n);
    if (!init) {
      RD(a.p, 4 * a.n);
      WR(a.x_new, 4 * a.n);
    }
    WR(a.r, 4 * a.n);
    partials_check(who, a.partials, 2, red_blocks(a.n));
    fin_resid(a.st, init);
  });
  return 0;
}

int launch_npsd_dir_csr(const CsrArgs& a, const float* pN, double* alfa_out,
                        hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "npsd_dir_csr";
    if (a.st->done) return;
    csr_check(who, a);
    RD(a.z, 4 * a.n);
    RD(a.r, 4 * a.n);
    if (pN) {
      RD(pN, 4 * a.n);
      if (a.p_old) RD(a.p_old, 4 * a.n);
    }
    WR(a.p_new, 4 * a.n);
    WR(a.w, 4 * a.n);
    partials_check(who, a.partials, 3, red_blocks(a.n));
    WR(alfa_out, sizeof(double));
    *alfa_out = 1.0;
  });
  return 0;
}

int launch_finalize_rz(double* partials, unsigned nb, SolveState* s, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "finalize_rz";
    if (s->done) return;
    RD(partials, 2 * nb * sizeof(double));
  });
  return 0;
}

int launch_finalize_proj(double* partials, unsigned nb, SolveState* s, const double* alfaN,
                         const double* alfaO, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "finalize_proj";
    if (s->done) return;
    RD(partials, 2 * nb * sizeof(double));
    RD(alfaN, sizeof(double));
    if (alfaO) RD(alfaO, sizeof(double));
  });
  return 0;
}

int launch_npsd_proj(const float* q, const float* aN, const float* aO, const double* alfaN,
                     const double* alfaO, long n, SolveState* s, double* partials, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "npsd_proj";
    if (s->done) return;
    RD(q, 4 * n);
    RD(aN, 4 * n);
    if (aO) RD(aO, 4 * n);
    RD(alfaN, sizeof(double));
    if (alfaO) RD(alfaO, sizeof(double));
    partials_check(who, partials, 2, red_blocks(n / 16 + 1));
  });
  return 0;
}

int launch_spmv_csr(const int* rowptr, const int* col, const float* val, const float* x, float* y, long n,
                    hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "spmv_csr";
    RD(rowptr, 4 * (n + 1));
    RD(col, 4 * (long)rowptr[n]);
    RD(val, 4 * (long)rowptr[n]);
    for (long e = 0; e < rowptr[n]; ++e) RD(x + col[e], 4);
    WR(y, 4 * n);
  });
  return 0;
}

int launch_spmv_csr_f64(const int* rowptr, const int* col, const double* val, const double* x, double* y,
                        long n, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "spmv_csr_f64";
    RD(rowptr, 4 * (n + 1));
    RD(col, 4 * (long)rowptr[n]);
    RD(val, 8 * (long)rowptr[n]);
    for (long e = 0; e < rowptr[n]; ++e) RD(x + col[e], 8);
    WR(y, 8 * n);
  });
  return 0;
}

int launch_spmv_flat(const uint8_t* code, const float* x, float* y, int n1, int n2, int n3, hipStream_t st) {
  const long V = (long)n1 * n2 * n3;
  fakehip::enqueue(st, [=] {
    const char* who = "spmv_flat";
    RD(code, V);
    RD(x, 4 * V);
    WR(y, 4 * V);
  });
  return 0;
}

int launch_dot(const float* x, const float* y, long n, double* partials, double* out,
               hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "dot";
    RD(x, 4 * n);
    RD(y, 4 * n);
    partials_check(who, partials, 1, red_blocks(n));
    WR(out, sizeof(double));
  });
  return 0;
}

// ------------------------------------------------------------------ Lanczos (FP64)
int launch_dot_f64(const double* a, const double* b, long n, double* out, double* partials, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "dot_f64";
    RD(a, 8 * n);
    RD(b, 8 * n);
    long nb = (n / 8 + 255) / 256;
    nb = nb < 1 ? 1 : (nb > 8192 ? 8192 : nb);
    if ((size_t)nb * 8 > dot_f64_workspace_bytes()) std::abort();
    WR(partials, 8 * nb);
    WR(out, 8);
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    *out = s;
  });
  return 0;
}

int launch_lanczos_update(double* w, const double* v1, const double* v0, const double* alpha,
                          const double* beta, long n, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "lanczos_update";
    WR(w, 8 * n);
    RD(v1, 8 * n);
    RD(alpha, 8);
    if (v0) {
      RD(v0, 8 * n);
      RD(beta, 8);
    }
  });
  return 0;
}

int launch_normalize_f64(double* out, const double* w, const double* nrm2, double* norm_out, long n,
                         hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "normalize_f64";
    WR(out, 8 * n);
    RD(w, 8 * n);
    RD(nrm2, 8);
    if (norm_out) WR(norm_out, 8);
  });
  return 0;
}

// ------------------------------------------------------------------ preconditioner side
int launch_avgpool2(int dtype, const void* in, void* out, int B, int n1, int n2, int n3, const int* done,
                    hipStream_t st) {
  MLPCG_REQUIRE(n1 % 2 == 0 && n2 % 2 == 0 && n3 % 2 == 0, "avgpool2: odd grid");
  const size_t V = (size_t)n1 * n2 * n3, s = esz(dtype);
  fakehip::enqueue(st, [=] {
    const char* who = "avgpool2";
    if (done_now(done)) return;
    RD(in, B * V * s);
    WR(out, B * (V / 8) * s);
  });
  return 0;
}

int launch_sm_apply(int tk, int tx, int ty, const ApplyArgs& a, hipStream_t st) {
  const size_t V = (size_t)a.n1 * a.n2 * a.n3;
  fakehip::enqueue(st, [=] {
    const char* who = "sm_apply";
    if (done_now(a.done)) return;
    RD(a.K, 27 * V * esz(tk));
    RD(a.x, a.B * V * esz(tx));
    WR(a.y, a.B * V * esz(ty));
    if (a.prev) {
      RD(a.prev, a.B * V * esz(tk));
      RD(a.c0, accsz(tk));
      RD(a.c1, accsz(tk));
    }
    if (a.mask_a) RD(a.mask_a, V);
    if (a.mask_b) RD(a.mask_b, V);
    if (a.in_scale) RD(a.in_scale, 8);
  });
  return 0;
}

int launch_sm_apply_transposed(int tk, const ApplyArgs& a, hipStream_t st) {
  const size_t V = (size_t)a.n1 * a.n2 * a.n3;
  fakehip::enqueue(st, [=] {
    const char* who = "sm_apply_t";
    RD(a.K, 27 * V * esz(tk));
    RD(a.x, a.B * V * esz(tk));
    WR(a.y, a.B * V * esz(tk));
  });
  return 0;
}

int launch_sm_fused(int tw, int tx, int ty, int C, const FusedArgs& a, hipStream_t st) {
  MLPCG_REQUIRE(tw == kBF16 && fused_eligible(C, 1, a.n1, a.n2, a.n3), "sm_fused: not eligible");
  const size_t V = (size_t)a.n1 * a.n2 * a.n3;
  const unsigned nb = fused_blocks(C, a.n1, a.n2, a.n3);
  fakehip::enqueue(st, [=] {
    const char* who = "sm_fused";
    if (done_now(a.done)) return;
    RD(a.img, C * V * 2);
    RD(a.apack, fused_weights_bytes(C));
    RD(a.x, V * esz(tx));
    WR(a.y, V * esz(ty));
    if (a.pool_out) WR(a.pool_out, V / 8 * esz(ty));
    if (a.prev) {
      RD(a.prev, V * 2);
      RD(a.c0, 4);
      RD(a.c1, 4);
    }
    if (a.mask_a) RD(a.mask_a, V);
    if (a.mask_b) RD(a.mask_b, V);
    if (a.in_scale) RD(a.in_scale, 8);
    if (a.dot0) {
      RD(a.dot0, 4 * V);
      if (a.dot1) RD(a.dot1, 4 * V);
      partials_check(who, a.partials, 2, nb);
    }
  });
  return 0;
}

int launch_fused_kwrite(int C, const FusedArgs& a, void* K, hipStream_t st) {
  MLPCG_REQUIRE(fused_eligible(C, 1, a.n1, a.n2, a.n3), "fused_kwrite: not eligible");
  const size_t V = (size_t)a.n1 * a.n2 * a.n3;
  fakehip::enqueue(st, [=] {
    const char* who = "fused_kwrite";
    RD(a.img, C * V * 2);
    RD(a.apack, fused_weights_bytes(C));
    WR(K, 27 * V * 2);
  });
  return 0;
}

int launch_fused_sgrad(int C, const FusedArgs& a, void* work, void* dx, hipStream_t st) {
  MLPCG_REQUIRE(fused_eligible(C, 1, a.n1, a.n2, a.n3), "fused_sgrad: not eligible");
  const size_t V = (size_t)a.n1 * a.n2 * a.n3;
  const size_t wb = fused_sgrad_bytes(a.n1, a.n2, a.n3);
  fakehip::enqueue(st, [=] {
    const char* who = "fused_sgrad";
    RD(a.img, C * V * 2);
    RD(a.apack, fused_weights_bytes(C));
    RD(a.x, V * 2);
    WR(work, wb);
    WR(dx, V * 2);
  });
  return 0;
}

int launch_upsample2(int dtype, const UpArgs& a, hipStream_t st) {
  MLPCG_REQUIRE(a.n1 % 2 == 0 && a.n2 % 2 == 0 && a.n3 % 2 == 0, "upsample2: odd grid");
  const size_t V = (size_t)a.n1 * a.n2 * a.n3, s = esz(dtype);
  fakehip::enqueue(st, [=] {
    const char* who = "upsample2";
    if (done_now(a.done)) return;
    RD(a.in, a.B * (V / 8) * s);
    WR(a.out, a.B * V * s);
    if (a.prev) {
      RD(a.prev, a.B * V * s);
      RD(a.c0, accsz(dtype));
      RD(a.c1, accsz(dtype));
    }
  });
  return 0;
}

int launch_upsample2_adjoint(int dtype, const void* gin, void* gout, int B, int n1, int n2, int n3,
                             int nearest, double scale, hipStream_t st) {
  const size_t V = (size_t)n1 * n2 * n3, s = esz(dtype);
  fakehip::enqueue(st, [=] {
    const char* who = "upsample2_adjoint";
    RD(gin, B * 8 * V * s);
    WR(gout, B * V * s);
  });
  return 0;
}

int launch_coarse_tail(int dtype, const TailArgs& a, hipStream_t st) {
  MLPCG_REQUIRE(tail_eligible(dtype, a.f1, a.f2, a.f3), "coarse_tail: not eligible");
  const size_t vf = (size_t)a.f1 * a.f2 * a.f3, s = esz(dtype);
  fakehip::enqueue(st, [=] {
    const char* who = "coarse_tail";
    if (done_now(a.done)) return;
    RD(a.x_in, 8 * vf * s);
    RD(a.K_pre, 27 * vf * s);
    RD(a.K_post, 27 * vf * s);
    RD(a.K_l, 27 * (vf / 8) * s);
    RD(a.c0, 4);
    RD(a.c1, 4);
    WR(a.out, 8 * vf * s);
  });
  return 0;
}

int launch_stencil_weights(int dtype, const void* img, const void* W, const void* b, void* K, int C, int n1,
                           int n2, int n3, bool trans, hipStream_t st) {
  return launch_stencil_weights2(dtype, dtype, img, W, b, K, C, n1, n2, n3, trans, st);
}

int launch_stencil_weights2(int tin, int tk, const void* img, const void* W, const void* b, void* K, int C,
                            int n1, int n2, int n3, bool trans, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "stencil_weights: C");
  const size_t V = (size_t)n1 * n2 * n3;
  fakehip::enqueue(st, [=] {
    const char* who = "stencil_weights";
    RD(img, C * V * esz(tin));
    RD(W, 27 * 27 * C * accsz(tin));
    RD(b, 27 * accsz(tin));
    WR(K, 27 * V * esz(tk));
  });
  return 0;
}

int launch_sm_linear(int dtype, const void* img, const void* w, const void* b, void* z, void* ybar,
                     void* workspace, int C, int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "sm_linear: C");
  const size_t V = (size_t)n1 * n2 * n3;
  const size_t ws = mean_patch_workspace(dtype, C, n1, n2, n3);
  fakehip::enqueue(st, [=] {
    const char* who = "sm_linear";
    RD(img, C * V * esz(dtype));
    WR(workspace, ws);
    RD(w, 27 * C * accsz(dtype));
    RD(b, accsz(dtype));
    WR(z, accsz(dtype));
    WR(ybar, 27 * C * accsz(dtype));
    std::memset(z, 0, accsz(dtype));
  });
  return 0;
}

int launch_sm_linear_pair(int dtype, const void* img, const void* w0, const void* b0, void* z0,
                          const void* w1, const void* b1, void* z1, void* ybar, void* workspace, int C,
                          int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "sm_linear: C");
  MLPCG_REQUIRE(!w1 || (b1 && z1), "sm_linear: pair");
  const size_t V = (size_t)n1 * n2 * n3;
  const size_t ws = mean_patch_workspace(dtype, C, n1, n2, n3);
  fakehip::enqueue(st, [=] {
    const char* who = "sm_linear_pair";
    RD(img, C * V * esz(dtype));
    WR(workspace, ws);
    RD(w0, 27 * C * accsz(dtype));
    RD(b0, accsz(dtype));
    WR(z0, accsz(dtype));
    if (w1) {
      RD(w1, 27 * C * accsz(dtype));
      RD(b1, accsz(dtype));
      WR(z1, accsz(dtype));
      std::memset(z1, 0, accsz(dtype));
    }
    WR(ybar, 27 * C * accsz(dtype));
    std::memset(z0, 0, accsz(dtype));
  });
  return 0;
}

int launch_weight_grad(int dtype, const void* a, const void* c, const void* img, void* gW, void* gb,
                       void* workspace, int C, int B, int n1, int n2, int n3, hipStream_t st) {
  const size_t V = (size_t)n1 * n2 * n3;
  const size_t ws = weight_grad_workspace(dtype, C, B, n1, n2, n3);
  fakehip::enqueue(st, [=] {
    const char* who = "weight_grad";
    RD(a, B * V * esz(dtype));
    RD(c, B * V * esz(dtype));
    RD(img, C * V * esz(dtype));
    WR(workspace, ws);
    WR(gW, 27 * 27 * C * accsz(dtype));
    WR(gb, 27 * accsz(dtype));
  });
  return 0;
}

int launch_linear_backward(int dtype, const void* g, const void* ybar, void* gw, void* gb, int C,
                           hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "linear_backward";
    RD(g, accsz(dtype));
    RD(ybar, 27 * C * accsz(dtype));
    WR(gw, 27 * C * accsz(dtype));
    WR(gb, accsz(dtype));
  });
  return 0;
}

int launch_nonzero_mask(int acc, const void* ch, uint8_t* m, long V, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "nonzero_mask";
    RD(ch, V * accsz(acc));
    WR(m, V);
  });
  return 0;
}

int launch_square(int acc, const void* in, void* out, int n, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "square";
    RD(in, n * accsz(acc));
    WR(out, n * accsz(acc));
  });
  return 0;
}

int launch_cast(int tin, int tout, const void* in, void* out, long n, hipStream_t st) {
  fakehip::enqueue(st, [=] {
    const char* who = "cast";
    RD(in, n * esz(tin));
    WR(out, n * esz(tout));
  });
  return 0;
}

int launch_fused_pack(int tw, const void* W, const void* b, int C, void* wbuf, hipStream_t st, bool fold) {
  MLPCG_REQUIRE(tw == kBF16 && C >= 1 && C <= 3 && (!fold || C == 3), "fused_pack: not eligible");
  fakehip::enqueue(st, [=] {
    const char* who = "fused_pack";
    RD(W, 27 * 27 * C * 4);
    RD(b, 27 * 4);
    WR(wbuf, fused_weights_bytes(fold ? 2 : C));
  });
  return 0;
}

}  // namespace mlpcg
