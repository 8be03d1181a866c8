// PCG kernels: the fluid Laplacian applied matrix-free from a per-voxel code (full grid)
// or from CSR (compressed unknowns), fused with the vector updates and FP64 dot products
// of cg_tests.py (npcg :26-101, npcg_flex :103-183, npsd :189-284).
//
// Full-grid ("flat") vectors are zero outside the unknowns.  code[v] = 0 for a non-fluid
// cell, else 1 + (number of non-solid face neighbours) = 1 + A_vv; a neighbour outside
// the grid counts as solid.  Then for a fluid v:  (A x)(v) = A_vv x(v) - sum_nbr x(nbr),
// the sum running over all in-grid face neighbours (x vanishes on non-fluid cells).
//
// Every reduction is two kernels (block partials, then a one-block finalize that adds
// them in block order, see common.hpp); the finalize updates the solver scalars in
// SolveState on the device, so an iteration needs no host round trip.  Every kernel
// returns at once when state->done is set.
#include <initializer_list>

#include "common.hpp"
#include "pcg.hpp"

namespace mlpcg {

// Reduction kernels run up to kMaxRedBlocks blocks (one z-run per thread at 256^3, so
// enough bytes are in flight to cover HBM latency); their partials are [NR][blocks].
constexpr int kMaxRedBlocks = 16384;
constexpr int kDotRun = 16;  // floats per thread per vector in rz_dot

// ---------------------------------------------------------------------------
// code from flags
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void flags_code_kernel(const int8_t* __restrict__ flags,
                                                              uint8_t* __restrict__ code, int n1,
                                                              int n2, int n3) {
  const long V = (long)n1 * n2 * n3;
  const long v = (long)blockIdx.x * kThreads + threadIdx.x;
  if (v >= V) return;
  const int k = (int)(v % n3), j = (int)((v / n3) % n2), i = (int)(v / ((long)n2 * n3));
  if (flags[v] != kFlagFluid) {
    code[v] = 0;
    return;
  }
  const long s1 = (long)n2 * n3;
  int d = 0;
  if (i > 0 && flags[v - s1] >= kFlagFluid) ++d;
  if (i < n1 - 1 && flags[v + s1] >= kFlagFluid) ++d;
  if (j > 0 && flags[v - n3] >= kFlagFluid) ++d;
  if (j < n2 - 1 && flags[v + n3] >= kFlagFluid) ++d;
  if (k > 0 && flags[v - 1] >= kFlagFluid) ++d;
  if (k < n3 - 1 && flags[v + 1] >= kFlagFluid) ++d;
  code[v] = (uint8_t)(1 + d);
}

// ---------------------------------------------------------------------------
// Flat stencil helpers: a thread owns a z-run of VZ voxels.
// ---------------------------------------------------------------------------
template <int VZ>
struct Run {
  int i, j, k0;
  long v0;
};

template <int VZ>
__device__ __forceinline__ bool run_at(long cid, int n1, int n2, int n3, Run<VZ>& r) {
  const int n3c = n3 / VZ;
  const long nch = (long)n1 * n2 * n3c;
  if (cid >= nch) return false;
  r.k0 = (int)(cid % n3c) * VZ;
  const long t = cid / n3c;
  r.j = (int)(t % n2);
  r.i = (int)(t / n2);
  r.v0 = ((long)r.i * n2 + r.j) * n3 + r.k0;
  return true;
}

// Loads u = x + a*p (or just x when p == nullptr) on the 7-point neighbourhood of a run:
// c[0..VZ+1] is the centre row with one halo cell each side, s[4][VZ] the four side rows
// (i-1, i+1, j-1, j+1).  Out-of-grid values are 0.  u is one FP32 fused multiply-add (the
// storage precision of x, one rounding) so a residual computed from it is the residual of
// the stored iterate.
template <typename TA>
__device__ __forceinline__ TA upd(float x, float p, float a, bool has_p) {
  return has_p ? (TA)fmaf(a, p, x) : (TA)x;
}

template <int VZ, typename TA>
__device__ __forceinline__ void load_star(const float* __restrict__ x, const float* __restrict__ p,
                                          float a, const Run<VZ>& r, int n1, int n2, int n3,
                                          TA (&c)[VZ + 2], TA (&s)[4][VZ]) {
  const long s1 = (long)n2 * n3;
  const int lane = threadIdx.x & 63;
  float xv[VZ], pv[VZ];
  load_vec<VZ>(x + r.v0, xv);
  if (p) load_vec<VZ>(p + r.v0, pv);
#pragma unroll
  for (int t = 0; t < VZ; ++t) c[t + 1] = upd<TA>(xv[t], p ? pv[t] : 0.f, a, p != nullptr);
  TA lft = wave_from_prev(c[VZ]), rgt = wave_from_next(c[1]);
  if (lane == 0 && r.k0 > 0) lft = upd<TA>(x[r.v0 - 1], p ? p[r.v0 - 1] : 0.f, a, p != nullptr);
  if (lane == 63 && r.k0 + VZ < n3) rgt = upd<TA>(x[r.v0 + VZ], p ? p[r.v0 + VZ] : 0.f, a, p != nullptr);
  c[0] = r.k0 > 0 ? lft : TA(0);
  c[VZ + 1] = r.k0 + VZ < n3 ? rgt : TA(0);
  const long offs[4] = {-s1, s1, -(long)n3, (long)n3};
  const bool ok[4] = {r.i > 0, r.i < n1 - 1, r.j > 0, r.j < n2 - 1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (ok[q]) {
      load_vec<VZ>(x + r.v0 + offs[q], xv);
      if (p) load_vec<VZ>(p + r.v0 + offs[q], pv);
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[q][t] = upd<TA>(xv[t], p ? pv[t] : 0.f, a, p != nullptr);
    } else {
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[q][t] = TA(0);
    }
  }
}

template <int VZ, typename TA>
__device__ __forceinline__ TA lap_at(const TA (&c)[VZ + 2], const TA (&s)[4][VZ], int t, int code) {
  const TA nb = c[t] + c[t + 2] + s[0][t] + s[1][t] + s[2][t] + s[3][t];
  return (TA)(code - 1) * c[t + 1] - nb;
}

// r = b - A x_new,  x_new = x_old + alpha p (alpha from the state unless init).
// Reduces [r.r, b.b].
template <int VZ>
__global__ __launch_bounds__(kThreads) void resid_flat_kernel(FlatArgs a, int init) {
  SolveState* st = a.st;
  if (!init && st->done) return;
  const float alpha = init ? 0.f : (float)st->alpha;
  const float* pp = init ? nullptr : a.p;
  double red[2] = {0.0, 0.0};
  const long nch = (long)a.n1 * a.n2 * (a.n3 / VZ);
  for (long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    Run<VZ> r;
    run_at<VZ>(cid, a.n1, a.n2, a.n3, r);
    double c[VZ + 2], s[4][VZ];
    load_star<VZ, double>(a.x_old, pp, alpha, r, a.n1, a.n2, a.n3, c, s);
    float bv[VZ], code[VZ];
    // single-use streams and outputs bypass L2 (the x and p rows of the neighbourhood stay)
    load_vec_nt<VZ>(a.b + r.v0, bv);
    load_vec_nt<VZ>(a.code + r.v0, code);
    float rv[VZ], xn[VZ];
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      const int cd = (int)code[t];
      const double res = cd ? (double)bv[t] - lap_at<VZ, double>(c, s, t, cd) : 0.0;
      rv[t] = (float)res;
      xn[t] = cd ? (float)c[t + 1] : 0.f;
      red[0] += res * res;
      red[1] += (double)bv[t] * (double)bv[t];
    }
    store_vec_nt<VZ>(a.r + r.v0, rv);
    if (!init) store_vec_nt<VZ>(a.x_new + r.v0, xn);
  }
  block_partials<2>(red, a.partials);
}

// p_new = z + beta p_old; w = A p_new; reduces p_new . w -> alpha = rho / (p.w).
template <int VZ>
__global__ __launch_bounds__(kThreads) void p_spmv_flat_kernel(FlatArgs a) {
  SolveState* st = a.st;
  if (st->done) return;
  const float beta = (float)st->beta;
  double red[1] = {0.0};
  const long nch = (long)a.n1 * a.n2 * (a.n3 / VZ);
  for (long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    Run<VZ> r;
    run_at<VZ>(cid, a.n1, a.n2, a.n3, r);
    float c[VZ + 2], s[4][VZ];
    // p_new = z + beta * p_old, evaluated on the neighbourhood
    load_star<VZ, float>(a.z, a.p_old, beta, r, a.n1, a.n2, a.n3, c, s);
    float code[VZ];
    load_vec_nt<VZ>(a.code + r.v0, code);
    float wv[VZ], pn[VZ];
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      const int cd = (int)code[t];
      const float w = cd ? lap_at<VZ, float>(c, s, t, cd) : 0.f;
      pn[t] = cd ? c[t + 1] : 0.f;
      wv[t] = w;
      red[0] += (double)pn[t] * (double)w;
    }
    store_vec_nt<VZ>(a.p_new + r.v0, pn);
    store_vec_nt<VZ>(a.w + r.v0, wv);
  }
  block_partials<1>(red, a.partials);
}

// rho = r.z (and z.(r - r_old) for the flexible variant) -> beta.
__global__ __launch_bounds__(kThreads) void rz_dot_kernel(const float* __restrict__ r,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ r_old, long n,
                                                          SolveState* st, double* partials) {
  if (st->done) return;
  double red[2] = {0.0, 0.0};
  // each thread owns runs of kDotRun consecutive floats, all loads issued before use so
  // every wave keeps 3 x 4 KB in flight
  const long nr = n / kDotRun;
  const bool flex = (st->method == kNpcgFlex) && r_old != nullptr;
  for (long q = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; q < nr;
       q += (long)gridDim.x * kThreads) {
    float rv[kDotRun], zv[kDotRun], ov[kDotRun];
    load_vec<kDotRun>(r + kDotRun * q, rv);
    load_vec<kDotRun>(z + kDotRun * q, zv);
    if (flex) load_vec<kDotRun>(r_old + kDotRun * q, ov);
#pragma unroll
    for (int t = 0; t < kDotRun; ++t) red[0] += (double)rv[t] * (double)zv[t];
    if (flex) {
#pragma unroll
      for (int t = 0; t < kDotRun; ++t) red[1] += (double)zv[t] * ((double)rv[t] - (double)ov[t]);
    }
  }
  // tail (n not a multiple of kDotRun): first block only
  if (blockIdx.x == 0) {
    for (long e = nr * kDotRun + threadIdx.x; e < n; e += kThreads) {
      red[0] += (double)r[e] * (double)z[e];
      if (flex) red[1] += (double)z[e] * ((double)r[e] - (double)r_old[e]);
    }
  }
  block_partials<2>(red, partials);
}

// ---------------------------------------------------------------------------
// CSR (compressed unknowns)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void scatter_kernel(const int* __restrict__ idx,
                                                           const float* __restrict__ src,
                                                           float* __restrict__ grid, long n,
                                                           const int* done) {
  if (solve_done(done)) return;
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) grid[idx[i]] = src[i];
}

__global__ __launch_bounds__(kThreads) void gather_kernel(const int* __restrict__ idx,
                                                          const float* __restrict__ grid,
                                                          float* __restrict__ dst, long n,
                                                          const int* done) {
  if (solve_done(done)) return;
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = grid[idx[i]];
}

__global__ __launch_bounds__(kThreads) void p_spmv_csr_kernel(CsrArgs a) {
  SolveState* st = a.st;
  if (st->done) return;
  const float beta = (float)st->beta;
  double red[1] = {0.0};
  for (long row = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; row < a.n;
       row += (long)gridDim.x * kThreads) {
    const int e0 = a.rowptr[row], e1 = a.rowptr[row + 1];
    float w = 0.f;
    for (int e = e0; e < e1; ++e) {
      const int col = a.col[e];
      w += a.val[e] * (a.z[col] + beta * a.p_old[col]);
    }
    const float pn = a.z[row] + beta * a.p_old[row];
    a.p_new[row] = pn;
    a.w[row] = w;
    red[0] += (double)pn * (double)w;
  }
  block_partials<1>(red, a.partials);
}

__global__ __launch_bounds__(kThreads) void resid_csr_kernel(CsrArgs a, int init) {
  SolveState* st = a.st;
  if (!init && st->done) return;
  // the iterate update is the flat kernel's: one FP32 fused multiply-add
  const float alpha = init ? 0.f : (float)st->alpha;
  double red[2] = {0.0, 0.0};
  for (long row = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; row < a.n;
       row += (long)gridDim.x * kThreads) {
    const int e0 = a.rowptr[row], e1 = a.rowptr[row + 1];
    double ax = 0.0;
    for (int e = e0; e < e1; ++e) {
      const int col = a.col[e];
      const double xc = init ? (double)a.x_old[col] : (double)fmaf(alpha, a.p[col], a.x_old[col]);
      ax += (double)a.val[e] * xc;
    }
    const double bv = a.b[row];
    const double res = bv - ax;
    a.r[row] = (float)res;
    if (!init) a.x_new[row] = fmaf(alpha, a.p[row], a.x_old[row]);
    red[0] += res * res;
    red[1] += bv * bv;
  }
  block_partials<2>(red, a.partials);
}

// Plain y = A x (CSR, FP32) and y = A x (flat, FP32) for the drop-in operator.
__global__ __launch_bounds__(kThreads) void spmv_csr_kernel(const int* __restrict__ rowptr,
                                                            const int* __restrict__ col,
                                                            const float* __restrict__ val,
                                                            const float* __restrict__ x,
                                                            float* __restrict__ y, long n) {
  const long row = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (row >= n) return;
  float s = 0.f;
  for (int e = rowptr[row]; e < rowptr[row + 1]; ++e) s += val[e] * x[col[e]];
  y[row] = s;
}

__global__ __launch_bounds__(kThreads) void spmv_csr_f64_kernel(const int* __restrict__ rowptr,
                                                                const int* __restrict__ col,
                                                                const double* __restrict__ val,
                                                                const double* __restrict__ x,
                                                                double* __restrict__ y, long n) {
  const long row = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (row >= n) return;
  double s = 0.0;
  for (int e = rowptr[row]; e < rowptr[row + 1]; ++e) s += val[e] * x[col[e]];
  y[row] = s;
}

template <int VZ>
__global__ __launch_bounds__(kThreads) void spmv_flat_kernel(const uint8_t* __restrict__ code,
                                                             const float* __restrict__ x,
                                                             float* __restrict__ y, int n1, int n2,
                                                             int n3) {
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  Run<VZ> r;
  if (!run_at<VZ>(cid, n1, n2, n3, r)) return;
  float c[VZ + 2], s[4][VZ];
  load_star<VZ, float>(x, nullptr, 0.f, r, n1, n2, n3, c, s);
  float cd[VZ], out[VZ];
  load_vec<VZ>(code + r.v0, cd);
#pragma unroll
  for (int t = 0; t < VZ; ++t) out[t] = cd[t] != 0.f ? lap_at<VZ, float>(c, s, t, (int)cd[t]) : 0.f;
  store_vec<VZ>(y + r.v0, out);
}

// ---------------------------------------------------------------------------
// Training loss ||b - M A M x||^2 / ||b||^2 of one right-hand side (M: the fluid cells,
// code != 0) for a network output x on the full grid that need not vanish off the fluid:
// the residual with the partial sums of r.r and b.b, and the gradient c M A M r.  The same
// arithmetic as x.float() * fluid, the matrix-free A, b - A x and (c * A(r * fluid)) *
// fluid in torch, in one pass each.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(bf16 v) { return (float)v; }

// x * (code != 0) on the 7-point neighbourhood of a run (layout of load_star)
template <int VZ, typename TX>
__device__ __forceinline__ void load_masked_star(const TX* __restrict__ x, const uint8_t* __restrict__ code,
                                                 const Run<VZ>& r, int n1, int n2, int n3,
                                                 float (&c)[VZ + 2], float (&s)[4][VZ]) {
  const long s1 = (long)n2 * n3;
  auto at = [&](long v) { return code[v] ? to_f(x[v]) : 0.f; };
  TX xv[VZ];
  uint8_t cv[VZ];
  load_vec<VZ>(x + r.v0, xv);
  load_vec<VZ>(code + r.v0, cv);
#pragma unroll
  for (int t = 0; t < VZ; ++t) c[t + 1] = cv[t] ? to_f(xv[t]) : 0.f;
  c[0] = r.k0 > 0 ? at(r.v0 - 1) : 0.f;
  c[VZ + 1] = r.k0 + VZ < n3 ? at(r.v0 + VZ) : 0.f;
  const long offs[4] = {-s1, s1, -(long)n3, (long)n3};
  const bool ok[4] = {r.i > 0, r.i < n1 - 1, r.j > 0, r.j < n2 - 1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (ok[q]) {
      load_vec<VZ>(x + r.v0 + offs[q], xv);
      load_vec<VZ>(code + r.v0 + offs[q], cv);
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[q][t] = cv[t] ? to_f(xv[t]) : 0.f;
    } else {
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[q][t] = 0.f;
    }
  }
}

// r = b - A (M x); partials [r.r, b.b]
template <int VZ, typename TX>
__global__ __launch_bounds__(kThreads) void loss_resid_flat_kernel(const TX* __restrict__ x,
                                                                   const float* __restrict__ b,
                                                                   const uint8_t* __restrict__ code,
                                                                   float* __restrict__ r, int n1, int n2,
                                                                   int n3, double* partials) {
  double red[2] = {0.0, 0.0};
  const long nch = (long)n1 * n2 * (n3 / VZ);
  for (long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    Run<VZ> ru;
    run_at<VZ>(cid, n1, n2, n3, ru);
    float c[VZ + 2], s[4][VZ];
    load_masked_star<VZ, TX>(x, code, ru, n1, n2, n3, c, s);
    float bv[VZ], rv[VZ];
    uint8_t cd[VZ];
    load_vec<VZ>(b + ru.v0, bv);
    load_vec<VZ>(code + ru.v0, cd);
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      const float ax = cd[t] ? lap_at<VZ, float>(c, s, t, (int)cd[t]) : 0.f;
      rv[t] = bv[t] - ax;
      red[0] += (double)rv[t] * (double)rv[t];
      red[1] += (double)bv[t] * (double)bv[t];
    }
    store_vec<VZ>(r + ru.v0, rv);
  }
  block_partials<2>(red, partials);
}

// g = ((float)coef * A (M r)) * M, stored in the activation type
template <int VZ, typename TG>
__global__ __launch_bounds__(kThreads) void loss_grad_flat_kernel(const float* __restrict__ r,
                                                                  const uint8_t* __restrict__ code,
                                                                  const double* __restrict__ coef,
                                                                  TG* __restrict__ g, int n1, int n2, int n3) {
  const float cf = (float)coef[0];
  Run<VZ> ru;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!run_at<VZ>(cid, n1, n2, n3, ru)) return;
  float c[VZ + 2], s[4][VZ];
  load_masked_star<VZ, float>(r, code, ru, n1, n2, n3, c, s);
  uint8_t cd[VZ];
  load_vec<VZ>(code + ru.v0, cd);
  TG out[VZ];
#pragma unroll
  for (int t = 0; t < VZ; ++t) {
    const float ax = cd[t] ? lap_at<VZ, float>(c, s, t, (int)cd[t]) : 0.f;
    out[t] = (TG)((cf * ax) * (cd[t] ? 1.f : 0.f));
  }
  store_vec<VZ>(g + ru.v0, out);
}

// sum_i a[i] b[i] in FP64 (drop-in dot for torch-free callers).
__global__ __launch_bounds__(kThreads) void dot_kernel(const float* __restrict__ x,
                                                       const float* __restrict__ y, long n,
                                                       double* partials) {
  double red[1] = {0.0};
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
    red[0] += (double)x[i] * (double)y[i];
  block_partials<1>(red, partials);
}

// Level combination of the SM models' upward pass, y = c0 x0 + c1 x1, with torch's
// rounding of those expressions: each coefficient is first rounded to the element type,
// each product and the sum are rounded to it (FP32 products are single roundings, BF16
// products of BF16 operands are exact in FP32 and rounded once to BF16).
__device__ __forceinline__ float coef_as(const void* c, int ct) {
  return ct == kBF16 ? (float)*(const bf16*)c : *(const float*)c;
}
template <typename T>
__device__ __forceinline__ float rnd(float v) {
  return (float)(T)v;
}
template <int VZ, typename T>
__global__ __launch_bounds__(kThreads) void comb_kernel(const T* __restrict__ x0, const T* __restrict__ x1,
                                                        const void* c0p, const void* c1p, int ct,
                                                        T* __restrict__ y, long nch) {
#pragma clang fp contract(off)  // torch rounds each product before the sum
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (cid >= nch) return;
  const float c0 = rnd<T>(coef_as(c0p, ct)), c1 = rnd<T>(coef_as(c1p, ct));
  float a[VZ], b[VZ], o[VZ];
  load_vec<VZ>(x0 + cid * VZ, a);
  load_vec<VZ>(x1 + cid * VZ, b);
#pragma unroll
  for (int t = 0; t < VZ; ++t) o[t] = rnd<T>(c0 * a[t]) + rnd<T>(c1 * b[t]);
  store_vec<VZ>(y + cid * VZ, o);
}

// Its backward: g0 = c0 g, g1 = c1 g (rounded like the forward's products; either may be
// null) and the coefficient gradients [g.x0, g.x1] accumulated in FP64.
template <int VZ, typename T>
__global__ __launch_bounds__(kThreads) void comb_grad_kernel(const T* __restrict__ g, const T* __restrict__ x0,
                                                             const T* __restrict__ x1, const void* c0p,
                                                             const void* c1p, int ct, T* __restrict__ g0,
                                                             T* __restrict__ g1, long nch, double* partials) {
  const float c0 = rnd<T>(coef_as(c0p, ct)), c1 = rnd<T>(coef_as(c1p, ct));
  double red[2] = {0.0, 0.0};
  for (long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    float gv[VZ], a[VZ], b[VZ], o[VZ];
    load_vec<VZ>(g + cid * VZ, gv);
    load_vec<VZ>(x0 + cid * VZ, a);
    load_vec<VZ>(x1 + cid * VZ, b);
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      red[0] += (double)gv[t] * (double)a[t];
      red[1] += (double)gv[t] * (double)b[t];
    }
    if (g0) {
#pragma unroll
      for (int t = 0; t < VZ; ++t) o[t] = c0 * gv[t];
      store_vec<VZ>(g0 + cid * VZ, o);
    }
    if (g1) {
#pragma unroll
      for (int t = 0; t < VZ; ++t) o[t] = c1 * gv[t];
      store_vec<VZ>(g1 + cid * VZ, o);
    }
  }
  block_partials<2>(red, partials);
}

// ---------------------------------------------------------------------------
// npsd (cg_tests.py:217-281).  With N the newest kept direction and O the older one,
// the reference projects q = P(r) newest first:
//   q1 = q - (q.AN / alfa_N) N,   Q = q1 - (q1.AO / alfa_O) O.
// Since q1.AO = q.AO - g_N (N.AO), both coefficients follow from the two dots q.AN and
// q.AO of one pass plus the cross term N.AO, which was reduced when N was created.  The
// new direction Q is then formed on the fly (neighbour values recomputed from q, N, O)
// by the kernel that applies A to it.
// ---------------------------------------------------------------------------
// [q.AN, q.AO] (AO may be null).
__global__ __launch_bounds__(kThreads) void npsd_proj_kernel(const float* __restrict__ q,
                                                             const float* __restrict__ aN,
                                                             const float* __restrict__ aO, long n,
                                                             SolveState* st, double* partials) {
  if (st->done) return;
  double red[2] = {0.0, 0.0};
  const long nr = n / kDotRun;
  for (long i = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; i < nr;
       i += (long)gridDim.x * kThreads) {
    float qv[kDotRun], nv[kDotRun], ov[kDotRun];
    load_vec<kDotRun>(q + kDotRun * i, qv);
    load_vec<kDotRun>(aN + kDotRun * i, nv);
    if (aO) load_vec<kDotRun>(aO + kDotRun * i, ov);
#pragma unroll
    for (int t = 0; t < kDotRun; ++t) red[0] += (double)qv[t] * (double)nv[t];
    if (aO) {
#pragma unroll
      for (int t = 0; t < kDotRun; ++t) red[1] += (double)qv[t] * (double)ov[t];
    }
  }
  if (blockIdx.x == 0) {
    for (long e = nr * kDotRun + threadIdx.x; e < n; e += kThreads) {
      red[0] += (double)q[e] * (double)aN[e];
      if (aO) red[1] += (double)q[e] * (double)aO[e];
    }
  }
  block_partials<2>(red, partials);
}

// Q = q - g_N N - g_O O, in FP64 with explicit FMAs (the same rounding wherever a value
// is formed, so the stored Q and the neighbour values used for A Q agree bit for bit).
struct DirIn {
  const float* q;
  const float* pN;  // null when no direction is kept
  const float* pO;  // null when fewer than two are kept
  float gN, gO;
};

// Q = q - gN N - gO O in FP32 with fused multiply-adds (one rounding each): Q is stored in
// FP32, and forming it in FP64 at the five rows of a run's neighbourhood cost more VALU
// work than the memory time of the kernel allows
__device__ __forceinline__ float dir_at(const DirIn& d, long v) {
  float u = d.q[v];
  if (d.pN) u = fmaf(-d.gN, d.pN[v], u);
  if (d.pO) u = fmaf(-d.gO, d.pO[v], u);
  return u;
}

// Q at VZ consecutive voxels; nv receives N there (0 without N)
template <int VZ>
__device__ __forceinline__ void dir_vec(const DirIn& d, long v0, float (&out)[VZ], float (&nv)[VZ]) {
  float qv[VZ], ov[VZ];
  load_vec<VZ>(d.q + v0, qv);
  if (d.pN) load_vec<VZ>(d.pN + v0, nv);
  if (d.pO) load_vec<VZ>(d.pO + v0, ov);
#pragma unroll
  for (int t = 0; t < VZ; ++t) {
    float u = qv[t];
    if (d.pN) u = fmaf(-d.gN, nv[t], u);
    else nv[t] = 0.f;
    if (d.pO) u = fmaf(-d.gO, ov[t], u);
    out[t] = u;
  }
}

// Q on the 7-point neighbourhood of a run (layout of load_star); ncen receives N on the
// run itself.
template <int VZ>
__device__ __forceinline__ void load_dir_star(const DirIn& d, const Run<VZ>& r, int n1, int n2,
                                              int n3, float (&c)[VZ + 2], float (&s)[4][VZ],
                                              float (&ncen)[VZ]) {
  const long s1 = (long)n2 * n3;
  const int lane = threadIdx.x & 63;
  float cv[VZ];
  dir_vec<VZ>(d, r.v0, cv, ncen);
#pragma unroll
  for (int t = 0; t < VZ; ++t) c[t + 1] = cv[t];
  float lft = wave_from_prev(c[VZ]), rgt = wave_from_next(c[1]);  // see load_star
  if (lane == 0 && r.k0 > 0) lft = dir_at(d, r.v0 - 1);
  if (lane == 63 && r.k0 + VZ < n3) rgt = dir_at(d, r.v0 + VZ);
  c[0] = r.k0 > 0 ? lft : 0.f;
  c[VZ + 1] = r.k0 + VZ < n3 ? rgt : 0.f;
  const long offs[4] = {-s1, s1, -(long)n3, (long)n3};
  const bool ok[4] = {r.i > 0, r.i < n1 - 1, r.j > 0, r.j < n2 - 1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (ok[q]) {
      float nside[VZ];
      dir_vec<VZ>(d, r.v0 + offs[q], s[q], nside);
    } else {
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[q][t] = 0.f;
    }
  }
}

__device__ __forceinline__ DirIn dir_in(const float* q, const float* pN, const float* pO,
                                        const SolveState* st) {
  DirIn d;
  d.q = q;
  d.pN = pN;
  d.pO = pO;
  d.gN = pN ? (float)st->proj[0] : 0.f;
  d.gO = pO ? (float)st->proj[1] : 0.f;
  return d;
}

// Q -> a.p_new, A Q -> a.w; reduces [Q.AQ, Q.r, N.AQ] (the cross term Q.AN of the next
// projection, formed as N.AQ: A is symmetric on the unknowns and Q, N, AQ, AN vanish off
// them, so the two are equal, and N.AQ needs no read of AN).  a.z = q, a.p_old = O; N
// comes in as pN (may be null).
template <int VZ>
__global__ __launch_bounds__(kThreads) void npsd_dir_flat_kernel(FlatArgs a, const float* pN) {
  SolveState* st = a.st;
  if (st->done) return;
  const DirIn d = dir_in(a.z, pN, pN ? a.p_old : nullptr, st);
  double red[3] = {0.0, 0.0, 0.0};
  const long nch = (long)a.n1 * a.n2 * (a.n3 / VZ);
  for (long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    Run<VZ> r;
    run_at<VZ>(cid, a.n1, a.n2, a.n3, r);
    float c[VZ + 2], s[4][VZ], nv[VZ];
    load_dir_star<VZ>(d, r, a.n1, a.n2, a.n3, c, s, nv);
    float code[VZ], rv[VZ], wv[VZ], qn[VZ];
    // single-use streams and outputs bypass L2 (the Q rows of the neighbourhood stay)
    load_vec_nt<VZ>(a.code + r.v0, code);
    load_vec_nt<VZ>(a.r + r.v0, rv);
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      const int cd = (int)code[t];
      qn[t] = cd ? c[t + 1] : 0.f;
      wv[t] = cd ? lap_at<VZ, float>(c, s, t, cd) : 0.f;
      red[0] += (double)qn[t] * (double)wv[t];
      red[1] += (double)qn[t] * (double)rv[t];
      if (pN) red[2] += (double)nv[t] * (double)wv[t];
    }
    store_vec_nt<VZ>(a.p_new + r.v0, qn);
    store_vec_nt<VZ>(a.w + r.v0, wv);
  }
  block_partials<3>(red, a.partials);
}

__global__ __launch_bounds__(kThreads) void npsd_dir_csr_kernel(CsrArgs a, const float* pN) {
  SolveState* st = a.st;
  if (st->done) return;
  const DirIn d = dir_in(a.z, pN, pN ? a.p_old : nullptr, st);
  double red[3] = {0.0, 0.0, 0.0};
  for (long row = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; row < a.n;
       row += (long)gridDim.x * kThreads) {
    float w = 0.f;
    for (int e = a.rowptr[row]; e < a.rowptr[row + 1]; ++e) w += a.val[e] * dir_at(d, a.col[e]);
    const float qn = dir_at(d, row);
    a.p_new[row] = qn;
    a.w[row] = w;
    red[0] += (double)qn * (double)w;
    red[1] += (double)qn * (double)a.r[row];
    if (pN) red[2] += (double)pN[row] * (double)w;  // N.AQ = Q.AN, as in the flat kernel
  }
  block_partials<3>(red, a.partials);
}

// ---------------------------------------------------------------------------
// Finalisation of the two-kernel reductions: one block sums the partials in block order
// and updates the solver scalars.
// ---------------------------------------------------------------------------
enum FinOp : int {
  kFinResidInit = 0, kFinResid, kFinPW, kFinRZ, kFinDot, kFinProj, kFinNpsd, kFinRZ2, kFinDot2, kFinDot2F32
};

// aux/aux2: the output of kFinDot; alfa of N and O for kFinProj; alfa of the new
// direction for kFinNpsd.
__global__ __launch_bounds__(kFinThreads) void finalize_kernel(int op, int nr,
                                                               const double* __restrict__ partials,
                                                               unsigned nb, SolveState* st,
                                                               double* aux, const double* aux2) {
  if (op != kFinResidInit && op != kFinDot && op != kFinDot2 && op != kFinDot2F32 && st->done)
    return;  // uniform over the block
  // thread 0 fetches the scalars it will read while the block sums the partials
  double rho = 0.0, rho_old = 0.0, norm_b = 0.0, tol = 0.0, atol = 0.0, cross = 0.0, an = 0.0, ao = 0.0;
  int iter = 0, max_it = 0, method = 0;
  double* hist = nullptr;
  if (threadIdx.x == 0 && st) {
    rho = st->rho;
    rho_old = st->rho_old;
    norm_b = st->norm_b;
    tol = st->tol;
    atol = st->atol;
    cross = st->cross;
    iter = st->iter;
    max_it = st->max_it;
    method = st->method;
    hist = st->history;
    if (op == kFinProj) {
      an = aux[0];
      ao = aux2 ? aux2[0] : 1.0;
    }
  }
  double t[3] = {0.0, 0.0, 0.0};
  sum_partials_rows<3>(partials, nb, nr, t);
  if (threadIdx.x != 0) return;
  switch (op) {
    case kFinResidInit:
    case kFinResid: {
      if (op == kFinResidInit) {
        norm_b = sqrt(t[1]);
        st->norm_b = norm_b;
        iter = 0;
        st->done = (t[0] == 0.0) ? 1 : 0;
      } else {
        iter += 1;
        st->rho_old = rho;
      }
      st->iter = iter;
      st->rr = t[0];
      const double norm = sqrt(t[0]) / norm_b;
      st->norm = norm;
      st->scale = t[0] > 0.0 ? 1.0 / sqrt(t[0]) : 0.0;
      if (op == kFinResid) {
        const double stop = fmax(tol, atol / norm_b);
        const bool cont = (method == kNpsd) ? !(norm < stop) : (norm > stop);
        st->done = (cont && iter < max_it) ? 0 : 1;
      }
      if (hist && iter <= max_it) hist[iter] = norm;
      break;
    }
    case kFinPW:
      st->pw = t[0];
      st->alpha = rho / t[0];
      break;
    case kFinRZ:
    case kFinRZ2: {  // kFinRZ: t1 = z.(r - r_old); kFinRZ2: t1 = z.r_old
      const bool flex = (method == kNpcgFlex);
      st->rho = t[0];
      const double num = flex ? (op == kFinRZ ? t[1] : t[0] - t[1]) : t[0];
      st->beta = (iter == 0) ? 0.0 : num / rho_old;
      break;
    }
    case kFinDot:
      aux[0] = t[0];
      break;
    case kFinDot2:
      aux[0] = t[0];
      aux[1] = t[1];
      break;
    case kFinDot2F32:  // the two sums rounded to FP32 (aux points at two floats)
      reinterpret_cast<float*>(aux)[0] = (float)t[0];
      reinterpret_cast<float*>(aux)[1] = (float)t[1];
      break;
    case kFinProj: {
      const double gN = t[0] / an;
      st->proj[0] = gN;
      st->proj[1] = aux2 ? (t[1] - gN * cross) / ao : 0.0;
      break;
    }
    case kFinNpsd:
      aux[0] = t[0];
      st->alpha = t[1] / t[0];
      st->cross = t[2];
      break;
  }
}

static void finalize(int op, int nr, double* partials, unsigned nb, SolveState* st, double* aux,
                     hipStream_t s, const double* aux2 = nullptr) {
  finalize_kernel<<<1, kFinThreads, 0, s>>>(op, nr, partials, nb, st, aux, aux2);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int flat_vz(int n3) { return (n3 % 8 == 0) ? 8 : (n3 % 4 == 0 ? 4 : (n3 % 2 == 0 ? 2 : 1)); }

static unsigned red_blocks(long items) {
  long nb = (items + kThreads - 1) / kThreads;
  if (nb > kMaxRedBlocks) nb = kMaxRedBlocks;
  return (unsigned)(nb < 1 ? 1 : nb);
}

size_t reduction_workspace_bytes() { return (size_t)kMaxRedBlocks * 4 * sizeof(double) + 256; }
long reduction_capacity() { return (long)kMaxRedBlocks * 4; }

int launch_flags_code(const int8_t* flags, uint8_t* code, int n1, int n2, int n3, hipStream_t st) {
  const long V = (long)n1 * n2 * n3;
  flags_code_kernel<<<(unsigned)((V + kThreads - 1) / kThreads), kThreads, 0, st>>>(flags, code, n1,
                                                                                     n2, n3);
  return check_launch("flags_code");
}

static unsigned flat_blocks(const FlatArgs& a) {
  const int vz = flat_vz(a.n3);
  return red_blocks((long)a.n1 * a.n2 * (a.n3 / vz));
}

#define MLPCG_FLAT_DISPATCH(KERN, ITEMS_VZ, ...)                                                \
  do {                                                                                          \
    const int vz_ = flat_vz(a.n3);                                                              \
    const unsigned nb_ = red_blocks((long)a.n1 * a.n2 * (a.n3 / vz_));                          \
    if (vz_ == 8) KERN<8><<<nb_, kThreads, 0, st>>>(__VA_ARGS__);                               \
    else if (vz_ == 4) KERN<4><<<nb_, kThreads, 0, st>>>(__VA_ARGS__);                          \
    else if (vz_ == 2) KERN<2><<<nb_, kThreads, 0, st>>>(__VA_ARGS__);                          \
    else KERN<1><<<nb_, kThreads, 0, st>>>(__VA_ARGS__);                                        \
  } while (0)

int launch_resid_flat(const FlatArgs& a, int init, hipStream_t st) {
  MLPCG_FLAT_DISPATCH(resid_flat_kernel, 0, a, init);
  finalize(init ? kFinResidInit : kFinResid, 2, a.partials, flat_blocks(a), a.st, nullptr, st);
  return check_launch("resid_flat");
}

int launch_p_spmv_flat(const FlatArgs& a, hipStream_t st) {
  MLPCG_FLAT_DISPATCH(p_spmv_flat_kernel, 0, a);
  finalize(kFinPW, 1, a.partials, flat_blocks(a), a.st, nullptr, st);
  return check_launch("p_spmv_flat");
}

int launch_npsd_dir_flat(const FlatArgs& a, const float* pN, double* alfa_out,
                         hipStream_t st) {
  MLPCG_FLAT_DISPATCH(npsd_dir_flat_kernel, 0, a, pN);
  finalize(kFinNpsd, 3, a.partials, flat_blocks(a), a.st, alfa_out, st);
  return check_launch("npsd_dir_flat");
}

int launch_rz_dot(const float* r, const float* z, const float* r_old, long n, SolveState* s,
                  double* partials, hipStream_t st) {
  const unsigned nb = red_blocks(n / kDotRun + 1);
  rz_dot_kernel<<<nb, kThreads, 0, st>>>(r, z, r_old, n, s, partials);
  finalize(kFinRZ, 2, partials, nb, s, nullptr, st);
  return check_launch("rz_dot");
}

int launch_scatter(const int* idx, const float* src, float* grid, long n, const int* done,
                   hipStream_t st) {
  scatter_kernel<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(idx, src, grid, n, done);
  return check_launch("scatter");
}

int launch_gather(const int* idx, const float* grid, float* dst, long n, const int* done,
                  hipStream_t st) {
  gather_kernel<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(idx, grid, dst, n, done);
  return check_launch("gather");
}

int launch_p_spmv_csr(const CsrArgs& a, hipStream_t st) {
  const unsigned nb = red_blocks(a.n);
  p_spmv_csr_kernel<<<nb, kThreads, 0, st>>>(a);
  finalize(kFinPW, 1, a.partials, nb, a.st, nullptr, st);
  return check_launch("p_spmv_csr");
}

int launch_resid_csr(const CsrArgs& a, int init, hipStream_t st) {
  const unsigned nb = red_blocks(a.n);
  resid_csr_kernel<<<nb, kThreads, 0, st>>>(a, init);
  finalize(init ? kFinResidInit : kFinResid, 2, a.partials, nb, a.st, nullptr, st);
  return check_launch("resid_csr");
}

int launch_npsd_dir_csr(const CsrArgs& a, const float* pN, double* alfa_out,
                        hipStream_t st) {
  const unsigned nb = red_blocks(a.n);
  npsd_dir_csr_kernel<<<nb, kThreads, 0, st>>>(a, pN);
  finalize(kFinNpsd, 3, a.partials, nb, a.st, alfa_out, st);
  return check_launch("npsd_dir_csr");
}

int launch_finalize_rz(double* partials, unsigned nb, SolveState* s, hipStream_t st) {
  finalize(kFinRZ2, 2, partials, nb, s, nullptr, st);
  return check_launch("finalize_rz");
}

int launch_finalize_proj(double* partials, unsigned nb, SolveState* s, const double* alfaN,
                         const double* alfaO, hipStream_t st) {
  finalize(kFinProj, 2, partials, nb, s, const_cast<double*>(alfaN), st, alfaO);
  return check_launch("finalize_proj");
}

int launch_npsd_proj(const float* q, const float* aN, const float* aO, const double* alfaN,
                     const double* alfaO, long n, SolveState* s, double* partials, hipStream_t st) {
  const unsigned nb = red_blocks(n / kDotRun + 1);
  npsd_proj_kernel<<<nb, kThreads, 0, st>>>(q, aN, aO, n, s, partials);
  finalize(kFinProj, 2, partials, nb, s, const_cast<double*>(alfaN), st, alfaO);
  return check_launch("npsd_proj");
}

int launch_spmv_csr(const int* rowptr, const int* col, const float* val, const float* x, float* y,
                    long n, hipStream_t st) {
  spmv_csr_kernel<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(rowptr, col, val, x, y, n);
  return check_launch("spmv_csr");
}

int launch_spmv_csr_f64(const int* rowptr, const int* col, const double* val, const double* x,
                        double* y, long n, hipStream_t st) {
  spmv_csr_f64_kernel<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(rowptr, col, val, x, y, n);
  return check_launch("spmv_csr_f64");
}

int launch_spmv_flat(const uint8_t* code, const float* x, float* y, int n1, int n2, int n3,
                     hipStream_t st) {
  const int vz = flat_vz(n3);
  const long items = (long)n1 * n2 * (n3 / vz);
  const unsigned nb = (unsigned)((items + kThreads - 1) / kThreads);
  if (vz == 8) spmv_flat_kernel<8><<<nb, kThreads, 0, st>>>(code, x, y, n1, n2, n3);
  else if (vz == 4) spmv_flat_kernel<4><<<nb, kThreads, 0, st>>>(code, x, y, n1, n2, n3);
  else if (vz == 2) spmv_flat_kernel<2><<<nb, kThreads, 0, st>>>(code, x, y, n1, n2, n3);
  else spmv_flat_kernel<1><<<nb, kThreads, 0, st>>>(code, x, y, n1, n2, n3);
  return check_launch("spmv_flat");
}

int launch_loss_resid(int tx, const void* x, const float* b, const uint8_t* code, float* r, int n1, int n2,
                      int n3, double* partials, double* out2, hipStream_t st) {
  const int vz = flat_vz(n3);
  const long items = (long)n1 * n2 * (n3 / vz);
  const unsigned nb = red_blocks(items);
#define MLPCG_LR(TX)                                                                                   \
  do {                                                                                                 \
    if (vz == 8) loss_resid_flat_kernel<8, TX><<<nb, kThreads, 0, st>>>((const TX*)x, b, code, r, n1, n2, n3, partials); \
    else if (vz == 4) loss_resid_flat_kernel<4, TX><<<nb, kThreads, 0, st>>>((const TX*)x, b, code, r, n1, n2, n3, partials); \
    else if (vz == 2) loss_resid_flat_kernel<2, TX><<<nb, kThreads, 0, st>>>((const TX*)x, b, code, r, n1, n2, n3, partials); \
    else loss_resid_flat_kernel<1, TX><<<nb, kThreads, 0, st>>>((const TX*)x, b, code, r, n1, n2, n3, partials); \
  } while (0)
  if (tx == kBF16) MLPCG_LR(bf16);
  else if (tx == kF32) MLPCG_LR(float);
  else MLPCG_REQUIRE(false, "loss_resid: x must be FP32 or BF16 (got %d)", tx);
#undef MLPCG_LR
  finalize(kFinDot2, 2, partials, nb, nullptr, out2, st);
  return check_launch("loss_resid");
}

int launch_loss_grad(int tg, const float* r, const uint8_t* code, const double* coef, void* g, int n1, int n2,
                     int n3, hipStream_t st) {
  const int vz = flat_vz(n3);
  const long items = (long)n1 * n2 * (n3 / vz);
  const unsigned nb = (unsigned)((items + kThreads - 1) / kThreads);
#define MLPCG_LG(TG)                                                                                  \
  do {                                                                                                \
    if (vz == 8) loss_grad_flat_kernel<8, TG><<<nb, kThreads, 0, st>>>(r, code, coef, (TG*)g, n1, n2, n3);       \
    else if (vz == 4) loss_grad_flat_kernel<4, TG><<<nb, kThreads, 0, st>>>(r, code, coef, (TG*)g, n1, n2, n3);  \
    else if (vz == 2) loss_grad_flat_kernel<2, TG><<<nb, kThreads, 0, st>>>(r, code, coef, (TG*)g, n1, n2, n3);  \
    else loss_grad_flat_kernel<1, TG><<<nb, kThreads, 0, st>>>(r, code, coef, (TG*)g, n1, n2, n3);               \
  } while (0)
  if (tg == kBF16) MLPCG_LG(bf16);
  else if (tg == kF32) MLPCG_LG(float);
  else MLPCG_REQUIRE(false, "loss_grad: g must be FP32 or BF16 (got %d)", tg);
#undef MLPCG_LG
  return check_launch("loss_grad");
}

// VZ of the combination kernels: 16-byte vectors where n and every pointer allow them
static int comb_vz(int dt, long n, std::initializer_list<const void*> ptrs) {
  int vz = dt == kBF16 ? 8 : 4;
  for (const void* p : ptrs)
    if (p && ((uintptr_t)p % 16) != 0) vz = 1;
  return n % vz == 0 ? vz : 1;
}

int launch_comb(int dt, const void* x0, const void* x1, const void* c0, const void* c1, int ct, void* y, long n,
                hipStream_t st) {
  MLPCG_REQUIRE(ct == kF32 || ct == kBF16, "comb: coefficients must be FP32 or BF16 (got %d)", ct);
  if (n == 0) return 0;
  const int vz = comb_vz(dt, n, {x0, x1, y});
  const long nch = n / vz;
  const dim3 grid((unsigned)((nch + kThreads - 1) / kThreads));
#define MLPCG_CB(T, V) comb_kernel<V, T><<<grid, kThreads, 0, st>>>((const T*)x0, (const T*)x1, c0, c1, ct, (T*)y, nch)
  if (dt == kBF16) {
    if (vz == 8) MLPCG_CB(bf16, 8); else MLPCG_CB(bf16, 1);
  } else if (dt == kF32) {
    if (vz == 4) MLPCG_CB(float, 4); else MLPCG_CB(float, 1);
  } else {
    MLPCG_REQUIRE(false, "comb: elements must be FP32 or BF16 (got %d)", dt);
  }
#undef MLPCG_CB
  return check_launch("comb");
}

int launch_comb_grad(int dt, const void* g, const void* x0, const void* x1, const void* c0, const void* c1, int ct,
                     void* g0, void* g1, long n, double* partials, void* out2, int out_dt, hipStream_t st) {
  MLPCG_REQUIRE(ct == kF32 || ct == kBF16, "comb_grad: coefficients must be FP32 or BF16 (got %d)", ct);
  MLPCG_REQUIRE(out_dt == kF64 || out_dt == kF32, "comb_grad: the sums are FP64 or FP32 (got %d)", out_dt);
  const int vz = comb_vz(dt, n, {g, x0, x1, g0, g1});
  const long nch = n / vz;
  const unsigned nb = red_blocks(nch);
#define MLPCG_CG(T, V)                                                                                    \
  comb_grad_kernel<V, T><<<nb, kThreads, 0, st>>>((const T*)g, (const T*)x0, (const T*)x1, c0, c1, ct, (T*)g0, \
                                                  (T*)g1, nch, partials)
  if (dt == kBF16) {
    if (vz == 8) MLPCG_CG(bf16, 8); else MLPCG_CG(bf16, 1);
  } else if (dt == kF32) {
    if (vz == 4) MLPCG_CG(float, 4); else MLPCG_CG(float, 1);
  } else {
    MLPCG_REQUIRE(false, "comb_grad: elements must be FP32 or BF16 (got %d)", dt);
  }
#undef MLPCG_CG
  finalize(out_dt == kF32 ? kFinDot2F32 : kFinDot2, 2, partials, nb, nullptr, (double*)out2, st);
  return check_launch("comb_grad");
}

int launch_dot(const float* x, const float* y, long n, double* partials,
               double* out, hipStream_t st) {
  const unsigned nb = red_blocks(n);
  dot_kernel<<<nb, kThreads, 0, st>>>(x, y, n, partials);
  finalize(kFinDot, 1, partials, nb, nullptr, out, st);
  return check_launch("dot");
}

}  // namespace mlpcg
