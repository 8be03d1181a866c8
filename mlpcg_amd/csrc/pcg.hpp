// PCG-side device state and kernel argument blocks.
#pragma once

#include "common.hpp"

namespace mlpcg {

constexpr int8_t kFlagFluid = 2;  // lib/read_data.py:13-16: BOUNDARY 0, SOLID 1, FLUID 2, AIR 3

enum Method : int { kNpcg = 0, kNpcgFlex = 1, kNpsd = 2 };

// Solver scalars, resident in device memory for the whole solve.  Updated only by the
// one-block finalize kernels of the reductions.
struct SolveState {
  double rho, rho_old, pw, alpha, beta;
  double rr, norm_b, norm, scale;  // scale = 1 / ||r|| (input normalisation of P)
  double tol, atol;
  double proj[2];                  // npsd projection coefficients (newest, older direction)
  double cross;                    // npsd: newest direction . A (previous direction)
  double alfa[2];                  // npsd: q.Aq of the last two directions
  double* history;                 // optional device array [max_it + 1] of ||r||/||b||
  int iter, max_it, done, method;
};

struct FlatArgs {
  const uint8_t* code;
  const float* b;
  const float* x_old;
  float* x_new;
  float* r;
  const float* p;      // direction used by the residual update (x_new = x_old + alpha p)
  const float* z;
  const float* p_old;
  float* p_new;
  float* w;
  SolveState* st;
  double* partials;
  int n1, n2, n3;
};

struct CsrArgs {
  const int* rowptr;
  const int* col;
  const float* val;
  const float* b;
  const float* x_old;
  float* x_new;
  float* r;
  const float* p;
  const float* z;
  const float* p_old;
  float* p_new;
  float* w;
  SolveState* st;
  double* partials;
  long n;
};

size_t reduction_workspace_bytes();
// doubles of partial rows a reduction workspace holds (rows x blocks must not exceed it)
long reduction_capacity();
int launch_flags_code(const int8_t* flags, uint8_t* code, int n1, int n2, int n3, hipStream_t st);
int launch_resid_flat(const FlatArgs& a, int init, hipStream_t st);
int launch_p_spmv_flat(const FlatArgs& a, hipStream_t st);
int launch_npsd_dir_flat(const FlatArgs& a, const float* pN, double* alfa_out,
                         hipStream_t st);
int launch_rz_dot(const float* r, const float* z, const float* r_old, long n, SolveState* s,
                  double* partials, hipStream_t st);
int launch_scatter(const int* idx, const float* src, float* grid, long n, const int* done,
                   hipStream_t st);
int launch_gather(const int* idx, const float* grid, float* dst, long n, const int* done,
                  hipStream_t st);
int launch_p_spmv_csr(const CsrArgs& a, hipStream_t st);
int launch_resid_csr(const CsrArgs& a, int init, hipStream_t st);
int launch_npsd_dir_csr(const CsrArgs& a, const float* pN, double* alfa_out,
                        hipStream_t st);
// Finalizers for dot products computed elsewhere (the fused apply's epilogue):
// partials [2][nb] of z.r and z.r_old (-> rho, beta), or of q.AN and q.AO (-> proj).
int launch_finalize_rz(double* partials, unsigned nb, SolveState* s, hipStream_t st);
int launch_finalize_proj(double* partials, unsigned nb, SolveState* s, const double* alfaN,
                         const double* alfaO, hipStream_t st);
int launch_npsd_proj(const float* q, const float* aN, const float* aO, const double* alfaN,
                     const double* alfaO, long n, SolveState* s, double* partials, hipStream_t st);
int launch_spmv_csr(const int* rowptr, const int* col, const float* val, const float* x, float* y,
                    long n, hipStream_t st);
int launch_spmv_csr_f64(const int* rowptr, const int* col, const double* val, const double* x,
                        double* y, long n, hipStream_t st);
int launch_spmv_flat(const uint8_t* code, const float* x, float* y, int n1, int n2, int n3,
                     hipStream_t st);
// Training loss of one right-hand side on the full grid (M: cells with code != 0):
// r = b - M A M x (x FP32 or BF16) and out2 = [r.r, b.b] (FP64, two-kernel reduction over
// `partials`); g = ((float)*coef * A M r) * M in the activation type (FP32 or BF16).
int launch_loss_resid(int tx, const void* x, const float* b, const uint8_t* code, float* r, int n1, int n2,
                      int n3, double* partials, double* out2, hipStream_t st);
int launch_loss_grad(int tg, const float* r, const uint8_t* code, const double* coef, void* g, int n1, int n2,
                     int n3, hipStream_t st);
// Level combination y = c0 x0 + c1 x1 (FP32 or BF16 elements, coefficients of type ct
// rounded to the element type first) and its backward g0 = c0 g, g1 = c1 g (either may be
// null), out2 = [g.x0, g.x1] accumulated in FP64, stored as out_dt (FP64 or FP32).
int launch_comb(int dt, const void* x0, const void* x1, const void* c0, const void* c1, int ct, void* y, long n,
                hipStream_t st);
int launch_comb_grad(int dt, const void* g, const void* x0, const void* x1, const void* c0, const void* c1, int ct,
                     void* g0, void* g1, long n, double* partials, void* out2, int out_dt, hipStream_t st);
int launch_dot(const float* x, const float* y, long n, double* partials,
               double* out, hipStream_t st);

// Lanczos vector kernels (lanczos.hip), FP64
size_t dot_f64_workspace_bytes();
int launch_dot_f64(const double* a, const double* b, long n, double* out, double* partials, hipStream_t st);
int launch_lanczos_update(double* w, const double* v1, const double* v0, const double* alpha,
                          const double* beta, long n, hipStream_t st);
int launch_normalize_f64(double* out, const double* w, const double* nrm2, double* norm_out, long n,
                         hipStream_t st);

}  // namespace mlpcg
