// Spatially varying 3x3x3 stencil (SM block) kernels for gfx950.
//
//   stencil_weights  K_p(v) = b_p + sum_{s,e} W[p,s,e] img[s,v+e]                 (27 x V)
//                    (the per-voxel "N x 27 x (27C+1)" contraction; the reference
//                    recomputes it inside every block kernel, sm_block_3d_kernel.cu:46-69)
//                    transposed form Kt_q(v) = K_{26-q}(v + d_q), so that the adjoint
//                    block (sm_block_trans_3d_kernel.cu:46-73) is the same apply kernel.
//   sm_apply         y[b,v] = sum_p K_p(v) x[b, v+d_p]  with a fused epilogue
//                    (input scale, model mask, y = c0*prev + c1*y, output mask).
//   avgpool2 / upsample2 / upsample2_adjoint   level transfers
//                    (F.avg_pool3d(.,2) and F.interpolate(scale 2, align_corners=False)).
//
// Thread mapping everywhere: one thread owns VZ consecutive voxels along the fastest
// axis; a wavefront therefore streams 64*VZ consecutive voxels of every K tap.
#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

// Row segment x[ii][jj][k0-1 .. k0+VZ] (VZ+2 values), zero outside the grid.
template <int VZ, typename A, typename T>
__device__ __forceinline__ void load_row(const T* __restrict__ base, int ii, int jj, int k0,
                                         int n1, int n2, int n3, A* row) {
  if ((unsigned)ii < (unsigned)n1 && (unsigned)jj < (unsigned)n2) {
    const T* r = base + ((long)ii * n2 + jj) * n3;
    load_vec<VZ>(r + k0, row + 1);
    row[0] = k0 > 0 ? (A)r[k0 - 1] : A(0);
    row[VZ + 1] = (k0 + VZ < n3) ? (A)r[k0 + VZ] : A(0);
  } else {
#pragma unroll
    for (int t = 0; t < VZ + 2; ++t) row[t] = A(0);
  }
}

// Row segment x[ii][jj][k0-2 .. k0+VZ+1] (VZ+4 values), zero outside the grid.
template <int VZ, typename A, typename T>
__device__ __forceinline__ void load_row2(const T* __restrict__ base, int ii, int jj, int k0,
                                          int n1, int n2, int n3, A* row) {
  if ((unsigned)ii < (unsigned)n1 && (unsigned)jj < (unsigned)n2) {
    const T* r = base + ((long)ii * n2 + jj) * n3;
    load_vec<VZ>(r + k0, row + 2);
    row[0] = k0 > 1 ? (A)r[k0 - 2] : A(0);
    row[1] = k0 > 0 ? (A)r[k0 - 1] : A(0);
    row[VZ + 2] = (k0 + VZ < n3) ? (A)r[k0 + VZ] : A(0);
    row[VZ + 3] = (k0 + VZ + 1 < n3) ? (A)r[k0 + VZ + 1] : A(0);
  } else {
#pragma unroll
    for (int t = 0; t < VZ + 4; ++t) row[t] = A(0);
  }
}

struct ChunkPos {
  int i, j, k0;
  long v0;
};

template <int VZ>
__device__ __forceinline__ bool chunk_pos(long cid, int n1, int n2, int n3, ChunkPos& c) {
  const int n3c = n3 / VZ;
  const long nch = (long)n1 * n2 * n3c;
  if (cid >= nch) return false;
  const int kc = (int)(cid % n3c);
  const long t = cid / n3c;
  c.j = (int)(t % n2);
  c.i = (int)(t / n2);
  c.k0 = kc * VZ;
  c.v0 = ((long)c.i * n2 + c.j) * n3 + c.k0;
  return true;
}

// ---------------------------------------------------------------------------
// stencil_weights
// ---------------------------------------------------------------------------
template <typename TI, typename TK, int C, int VZ>
__global__ __launch_bounds__(kThreads) void stencil_weights_kernel(
    const TI* __restrict__ img, const typename AccOf<TK>::type* __restrict__ W,
    const typename AccOf<TK>::type* __restrict__ bias, TK* __restrict__ K, int n1, int n2, int n3) {
  typedef typename AccOf<TK>::type A;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, n1, n2, n3, c)) return;
  const long V = (long)n1 * n2 * n3;
  // the 3x3 neighbourhood rows of every channel stay in registers; the taps are a
  // runtime loop so each tap's 27C weights stream through the scalar cache
  A rows[C][3][3][VZ + 2];
#pragma unroll
  for (int s = 0; s < C; ++s)
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
      for (int a2 = 0; a2 < 3; ++a2)
        load_row<VZ>(img + s * V, c.i + a1 - 1, c.j + a2 - 1, c.k0, n1, n2, n3, rows[s][a1][a2]);
#pragma unroll 1
  for (int p = 0; p < 27; ++p) {
    const A* __restrict__ Wp = W + p * C * 27;
    A acc[VZ];
    const A bp = bias[p];
#pragma unroll
    for (int t = 0; t < VZ; ++t) acc[t] = bp;
#pragma unroll
    for (int s = 0; s < C; ++s)
#pragma unroll
      for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
          for (int a3 = 0; a3 < 3; ++a3) {
            const A w = Wp[s * 27 + a1 * 9 + a2 * 3 + a3];
#pragma unroll
            for (int t = 0; t < VZ; ++t) acc[t] += w * rows[s][a1][a2][t + a3];
          }
    store_vec<VZ>(K + p * V + c.v0, acc);
  }
}

// The same stencils with one thread per (tap, voxel): on small grids the per-voxel kernel
// above is a chain of 27 * 28 dependent FMAs per thread on a few waves; here each thread
// computes one K value (same summation order, so the same result).
template <typename TI, typename TK, int C>
__global__ __launch_bounds__(kThreads) void stencil_weights_small_kernel(
    const TI* __restrict__ img, const typename AccOf<TK>::type* __restrict__ W,
    const typename AccOf<TK>::type* __restrict__ bias, TK* __restrict__ K, int n1, int n2, int n3) {
  typedef typename AccOf<TK>::type A;
  const long V = (long)n1 * n2 * n3;
  const long idx = (long)blockIdx.x * kThreads + threadIdx.x;  // p * V + v
  if (idx >= 27 * V) return;
  const int p = (int)(idx / V);
  const long v = idx - (long)p * V;
  const int k = (int)(v % n3);
  const long t = v / n3;
  const int j = (int)(t % n2), i = (int)(t / n2);
  const A* __restrict__ Wp = W + p * C * 27;
  A acc = bias[p];
#pragma unroll
  for (int s = 0; s < C; ++s)
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
      for (int a2 = 0; a2 < 3; ++a2) {
        const int ii = i + a1 - 1, jj = j + a2 - 1;
        const bool rok = (unsigned)ii < (unsigned)n1 && (unsigned)jj < (unsigned)n2;
        const long rb = (long)s * V + ((long)ii * n2 + jj) * n3;
#pragma unroll
        for (int a3 = 0; a3 < 3; ++a3) {
          const int kk = k + a3 - 1;
          const A xv = (rok && (unsigned)kk < (unsigned)n3) ? (A)img[rb + kk] : A(0);
          acc += Wp[s * 27 + a1 * 9 + a2 * 3 + a3] * xv;
        }
      }
  K[idx] = (TK)acc;
}

// Kt_q(v) = K_p(v - d_p) with p = 26 - q (so d_q = -d_p); zero where v + d_q is outside.
template <typename TI, typename TK, int C, int VZ>
__global__ __launch_bounds__(kThreads) void stencil_weights_t_kernel(
    const TI* __restrict__ img, const typename AccOf<TK>::type* __restrict__ W,
    const typename AccOf<TK>::type* __restrict__ bias, TK* __restrict__ K, int n1, int n2, int n3) {
  typedef typename AccOf<TK>::type A;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, n1, n2, n3, c)) return;
  const long V = (long)n1 * n2 * n3;
  // runtime loop over the 9 (dq1, dq2) columns of output taps; for each, the image
  // rows around (i+dq1, j+dq2) are read with a two-voxel z halo that serves all three
  // dq3 taps of the column
#pragma unroll 1
  for (int g = 0; g < 9; ++g) {
    const int dq1 = g / 3 - 1, dq2 = g % 3 - 1;
    A acc[3][VZ];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const A bp = bias[26 - (g * 3 + u)];
#pragma unroll
      for (int t = 0; t < VZ; ++t) acc[u][t] = bp;
    }
#pragma unroll
    for (int s = 0; s < C; ++s) {
#pragma unroll
      for (int e1 = 0; e1 < 3; ++e1) {
#pragma unroll
        for (int e2 = 0; e2 < 3; ++e2) {
          A row[VZ + 4];
          load_row2<VZ>(img + s * V, c.i + dq1 + e1 - 1, c.j + dq2 + e2 - 1, c.k0, n1, n2, n3, row);
#pragma unroll
          for (int u = 0; u < 3; ++u) {  // dq3 = u - 1, q = 3g + u, p = 26 - q
            const A* __restrict__ Wp = W + ((26 - (g * 3 + u)) * C + s) * 27 + e1 * 9 + e2 * 3;
#pragma unroll
            for (int e3 = 0; e3 < 3; ++e3) {
              const A w = Wp[e3];
#pragma unroll
              for (int t = 0; t < VZ; ++t) acc[u][t] += w * row[t + 1 + u + e3 - 1];
            }
          }
        }
      }
    }
    const bool rowok = (unsigned)(c.i + dq1) < (unsigned)n1 && (unsigned)(c.j + dq2) < (unsigned)n2;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
#pragma unroll
      for (int t = 0; t < VZ; ++t) {
        const bool ok = rowok && (unsigned)(c.k0 + t + u - 1) < (unsigned)n3;
        if (!ok) acc[u][t] = A(0);
      }
      store_vec<VZ>(K + (long)(g * 3 + u) * V + c.v0, acc[u]);
    }
  }
}

// ---------------------------------------------------------------------------
// sm_apply_t: the transpose of a stored-stencil apply,
//   y[b, v] = sum_p K_p(v - d_p) x[b, v - d_p]       (sm_block_trans_3d_kernel.cu:46-73)
// gathered from the forward stencils K (no transposed stencils needed): for the source
// row (i - d1, j - d2) the K rows of its three taps and the x row are read over
// [k0 - 1, k0 + VZ] (zero outside the grid), and tap d3 contributes at offset -d3.
// ---------------------------------------------------------------------------
template <typename TK, typename TX, int VZ, int BB>
__global__ __launch_bounds__(kThreads) void sm_apply_t_kernel(ApplyArgs a) {
  typedef typename AccOf<TK>::type A;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, a.n1, a.n2, a.n3, c)) return;
  const long V = (long)a.n1 * a.n2 * a.n3;
  const int b0 = blockIdx.y * BB;
  const TK* __restrict__ K = (const TK*)a.K;
  const TX* __restrict__ x = (const TX*)a.x;
  A acc[BB][VZ];
#pragma unroll
  for (int bb = 0; bb < BB; ++bb)
#pragma unroll
    for (int t = 0; t < VZ; ++t) acc[bb][t] = A(0);
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1) {
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2) {
      const int ri = c.i + 1 - a1, rj = c.j + 1 - a2;  // source row (i - d1, j - d2)
      A kh[3][VZ + 2];
#pragma unroll
      for (int a3 = 0; a3 < 3; ++a3)
        load_row<VZ>(K + (long)(9 * a1 + 3 * a2 + a3) * V, ri, rj, c.k0, a.n1, a.n2, a.n3, kh[a3]);
#pragma unroll
      for (int bb = 0; bb < BB; ++bb) {
        if (BB > 1 && b0 + bb >= a.B) break;
        A row[VZ + 2];
        load_row<VZ>(x + (long)(b0 + bb) * V, ri, rj, c.k0, a.n1, a.n2, a.n3, row);
        // tap a3 (d3 = a3 - 1) reads source z = k0 + t - d3: halo index t + 2 - a3
#pragma unroll
        for (int t = 0; t < VZ; ++t)
          acc[bb][t] += kh[0][t + 2] * row[t + 2] + kh[1][t + 1] * row[t + 1] + kh[2][t] * row[t];
      }
    }
  }
#pragma unroll
  for (int bb = 0; bb < BB; ++bb) {
    if (BB > 1 && b0 + bb >= a.B) break;
    store_vec<VZ>((TX*)a.y + (long)(b0 + bb) * V + c.v0, acc[bb]);
  }
}

// The same gather (B = 1) without halo loads, where every z-row lies within one wave
// (n3 / VZ runs, a power of two up to 64): per source row the thread forms the products
// S_p(z') = K_p(z') x(z') on its own aligned run only (16-byte loads of K and x); tap
// d3 moves them to z = z' + d3, so the run's edge products of taps d3 = -1 and +1 belong to
// the neighbouring runs.  Those are summed over the nine source rows and handed over by
// two DPP wave shifts at the end.  27 + 9 loads per run instead of 3 (27 + 9).
template <typename TK, int VZ>
__global__ __launch_bounds__(kThreads) void sm_apply_t_edge_kernel(ApplyArgs a) {
  typedef typename AccOf<TK>::type A;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  const int L = a.n3 / VZ;  // runs per z-row
  const long nch = (long)a.n1 * a.n2 * L;
  if (cid >= nch) return;
  const int kc = (int)(cid % L);
  const long row = cid / L;
  const int j = (int)(row % a.n2), i = (int)(row / a.n2);
  const int k0 = kc * VZ;
  const long V = (long)a.n1 * a.n2 * a.n3;
  const TK* __restrict__ K = (const TK*)a.K;
  const TK* __restrict__ x = (const TK*)a.x;
  A acc[VZ];
#pragma unroll
  for (int t = 0; t < VZ; ++t) acc[t] = A(0);
  A to_prev = A(0), to_next = A(0);  // edge products for the runs before / after this one
#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1) {
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2) {
      const int ri = i + 1 - a1, rj = j + 1 - a2;  // source row (i - d1, j - d2)
      if ((unsigned)ri >= (unsigned)a.n1 || (unsigned)rj >= (unsigned)a.n2) continue;
      const long off = ((long)ri * a.n2 + rj) * a.n3 + k0;
      A xv[VZ], k0v[VZ], k1v[VZ], k2v[VZ];
      load_vec<VZ>(x + off, xv);
      const long p0 = (long)(9 * a1 + 3 * a2) * V;
      load_vec_nt<VZ>(K + p0 + off, k0v);
      load_vec_nt<VZ>(K + p0 + V + off, k1v);
      load_vec_nt<VZ>(K + p0 + 2 * V + off, k2v);
      // tap d3 = 0 stays, d3 = -1 moves down one cell, d3 = +1 up one cell
#pragma unroll
      for (int t = 0; t < VZ; ++t) acc[t] += k1v[t] * xv[t];
#pragma unroll
      for (int t = 1; t < VZ; ++t) acc[t - 1] += k0v[t] * xv[t];
#pragma unroll
      for (int t = 0; t + 1 < VZ; ++t) acc[t + 1] += k2v[t] * xv[t];
      to_prev += k0v[0] * xv[0];
      to_next += k2v[VZ - 1] * xv[VZ - 1];
    }
  }
  // the row's runs are consecutive lanes of this wave: lane l + 1 holds the next run
  const A from_next = wave_from_next(to_prev), from_prev = wave_from_prev(to_next);
  if (kc + 1 < L) acc[VZ - 1] += from_next;
  if (kc > 0) acc[0] += from_prev;
  store_vec<VZ>((TK*)a.y + ((long)i * a.n2 + j) * a.n3 + k0, acc);
}

// ---------------------------------------------------------------------------
// sm_apply
// ---------------------------------------------------------------------------
template <typename TK, typename TX, typename TY, typename TP, int VZ, int BB>
__global__ __launch_bounds__(kThreads) void sm_apply_kernel(ApplyArgs a) {
  typedef typename AccOf<TK>::type A;
  if (solve_done(a.done)) return;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, a.n1, a.n2, a.n3, c)) return;
  const long V = (long)a.n1 * a.n2 * a.n3;
  const int b0 = blockIdx.y * BB;
  const TK* __restrict__ K = (const TK*)a.K;
  const TX* __restrict__ x = (const TX*)a.x;
  // epilogue operands are loaded first so their latency overlaps the stencil loads (on
  // the small coarse-level launches the kernel is a chain of memory round trips)
  const A scale = a.in_scale ? (A)(*a.in_scale) : A(1);
  const A c0 = a.prev ? *(const A*)a.c0 : A(0);
  const A c1 = a.prev ? *(const A*)a.c1 : A(0);
  uint8_t keep_a[VZ], keep_b[VZ];
  if (a.mask_a) load_vec<VZ>(a.mask_a + c.v0, keep_a);
  if (a.mask_b) load_vec<VZ>(a.mask_b + c.v0, keep_b);
  TP pv[BB][VZ];
  if (a.prev) {
#pragma unroll
    for (int bb = 0; bb < BB; ++bb)
      if (!(BB > 1 && b0 + bb >= a.B)) load_vec<VZ>((const TP*)a.prev + (long)(b0 + bb) * V + c.v0, pv[bb]);
  }
  A acc[BB][VZ];
#pragma unroll
  for (int bb = 0; bb < BB; ++bb)
#pragma unroll
    for (int t = 0; t < VZ; ++t) acc[bb][t] = A(0);

#pragma unroll
  for (int a1 = 0; a1 < 3; ++a1) {
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2) {
      A kv[3][VZ];
#pragma unroll
      for (int a3 = 0; a3 < 3; ++a3)
        load_vec_nt<VZ>(K + (long)(9 * a1 + 3 * a2 + a3) * V + c.v0, kv[a3]);
#pragma unroll
      for (int bb = 0; bb < BB; ++bb) {
        if (BB > 1 && b0 + bb >= a.B) break;
        A row[VZ + 2];
        load_row<VZ>(x + (long)(b0 + bb) * V, c.i + a1 - 1, c.j + a2 - 1, c.k0, a.n1, a.n2,
                     a.n3, row);
#pragma unroll
        for (int t = 0; t < VZ; ++t)
          acc[bb][t] += kv[0][t] * row[t] + kv[1][t] * row[t + 1] + kv[2][t] * row[t + 2];
      }
    }
  }

  // epilogue
#pragma unroll
  for (int bb = 0; bb < BB; ++bb) {
    if (BB > 1 && b0 + bb >= a.B) break;
    const long off = (long)(b0 + bb) * V + c.v0;
    A out[VZ];
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      A val = acc[bb][t] * scale;
      if (a.mask_a && keep_a[t] == 0) val = A(0);
      out[t] = val;
    }
    if (a.prev) {
#pragma unroll
      for (int t = 0; t < VZ; ++t) out[t] = c0 * (A)pv[bb][t] + c1 * out[t];
    }
    if (a.mask_b) {
#pragma unroll
      for (int t = 0; t < VZ; ++t)
        if (keep_b[t] == 0) out[t] = A(0);
    }
    store_vec<VZ>((TY*)a.y + off, out);
  }
}

// ---------------------------------------------------------------------------
// Level transfers
// ---------------------------------------------------------------------------
// out[b][I][J][K] = mean of the 2x2x2 fine cells; (n1, n2, n3) are the coarse dims.
template <typename TI, typename TO, int VZ>
__global__ __launch_bounds__(kThreads) void avgpool2_kernel(const TI* __restrict__ in,
                                                            TO* __restrict__ out, int B, int n1,
                                                            int n2, int n3, const int* done) {
  typedef typename AccOf<TO>::type A;
  if (solve_done(done)) return;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, n1, n2, n3, c)) return;
  const int b = blockIdx.y;
  const long Vc = (long)n1 * n2 * n3;
  const int f2 = 2 * n2, f3 = 2 * n3;
  const long Vf = Vc * 8;
  A s[VZ];
#pragma unroll
  for (int t = 0; t < VZ; ++t) s[t] = A(0);
#pragma unroll
  for (int a1 = 0; a1 < 2; ++a1) {
#pragma unroll
    for (int a2 = 0; a2 < 2; ++a2) {
      A row[2 * VZ];
      load_vec<2 * VZ>(in + b * Vf + ((long)(2 * c.i + a1) * f2 + 2 * c.j + a2) * f3 + 2 * c.k0, row);
#pragma unroll
      for (int t = 0; t < VZ; ++t) s[t] += row[2 * t] + row[2 * t + 1];
    }
  }
#pragma unroll
  for (int t = 0; t < VZ; ++t) s[t] *= A(0.125);
  store_vec<VZ>(out + b * Vc + c.v0, s);
}

// Linear 1-D source indices for the x2 upsample with align_corners=False:
// src = max((f + 0.5) / 2 - 0.5, 0); i0 = floor(src), i1 = min(i0 + 1, nc - 1).
__device__ __forceinline__ void up_axis(int f, int nc, bool nearest, int& i0, int& i1, float& lam) {
  if (nearest) {
    i0 = i1 = f >> 1;
    lam = 0.f;
    return;
  }
  const int m = f >> 1;
  if (f & 1) {
    i0 = m;
    lam = 0.25f;
  } else if (m == 0) {
    i0 = 0;
    lam = 0.f;
  } else {
    i0 = m - 1;
    lam = 0.75f;
  }
  i1 = min(i0 + 1, nc - 1);
}

// out = scale * up(in)      or, with prev:  out = c0 * prev + c1 * scale * up(in)
// (n1, n2, n3) are the FINE dims.
template <typename TI, typename TO, typename TP, int VZ>
__global__ __launch_bounds__(kThreads) void upsample2_kernel(UpArgs a) {
  typedef typename AccOf<TO>::type A;
  if (solve_done(a.done)) return;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, a.n1, a.n2, a.n3, c)) return;
  const int b = blockIdx.y;
  const int m1 = a.n1 >> 1, m2 = a.n2 >> 1, m3 = a.n3 >> 1;
  const long Vf = (long)a.n1 * a.n2 * a.n3, Vc = Vf / 8;
  const TI* __restrict__ in = (const TI*)a.in + b * Vc;
  int i0, i1, j0, j1;
  float li, lj;
  up_axis(c.i, m1, a.nearest, i0, i1, li);
  up_axis(c.j, m2, a.nearest, j0, j1, lj);
  const TI* r00 = in + ((long)i0 * m2 + j0) * m3;
  const TI* r01 = in + ((long)i0 * m2 + j1) * m3;
  const TI* r10 = in + ((long)i1 * m2 + j0) * m3;
  const TI* r11 = in + ((long)i1 * m2 + j1) * m3;
  const A hi0 = A(1) - (A)li, hi1 = (A)li, hj0 = A(1) - (A)lj, hj1 = (A)lj;
  const long off = b * Vf + c.v0;
  // combine operands first (latency overlaps the coarse loads)
  const A c0 = a.prev ? *(const A*)a.c0 : A(0), c1 = a.prev ? *(const A*)a.c1 : A(0);
  TP pv[VZ];
  if (a.prev) load_vec<VZ>((const TP*)a.prev + off, pv);
  A out[VZ];
  if constexpr (VZ >= 2) {
    // Separable: the four coarse rows are combined per coarse cell first (u = the (i, j)
    // interpolation), then u is interpolated along z.  u covers the coarse z segment
    // [k0/2 - 1, k0/2 + VZ/2], indices clamped to the grid: with the clamp the interior
    // weights reproduce the edge cases exactly (f = 0 -> u[0]; f = 2m3-1 -> u[m3-1]), so
    // the z taps below are compile-time.  The two end cells of u come from the
    // neighbouring lanes (consecutive lanes hold consecutive segments of a row) by DPP
    // shifts; lanes 0 / 63 form them from loads, and at the row ends the clamp makes them
    // the segment's own end values.
    constexpr int NS = VZ / 2 + 2;
    const int kc = c.k0 >> 1;
    const int klo = max(kc - 1, 0), khi = min(kc + VZ / 2, m3 - 1);
    const int lane = threadIdx.x & 63;
    const bool row_start = c.k0 == 0, row_end = c.k0 + VZ >= a.n3;
    A seg[4][VZ / 2];
    const TI* rows[4] = {r00, r01, r10, r11};
#pragma unroll
    for (int q = 0; q < 4; ++q) load_vec<VZ / 2>(rows[q] + kc, seg[q]);
    auto comb = [&](A x00, A x01, A x10, A x11) { return hi0 * (hj0 * x00 + hj1 * x01) + hi1 * (hj0 * x10 + hj1 * x11); };
    A u[NS];
#pragma unroll
    for (int s = 0; s < VZ / 2; ++s) u[s + 1] = comb(seg[0][s], seg[1][s], seg[2][s], seg[3][s]);
    const A lo = wave_from_prev(u[VZ / 2]), hi = wave_from_next(u[1]);
    u[0] = row_start ? u[1]
                     : (lane == 0 ? comb((A)r00[klo], (A)r01[klo], (A)r10[klo], (A)r11[klo]) : lo);
    u[NS - 1] = row_end ? u[VZ / 2]
                        : (lane == 63 ? comb((A)r00[khi], (A)r01[khi], (A)r10[khi], (A)r11[khi]) : hi);
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      A z;
      if (a.nearest) z = u[t / 2 + 1];
      else if ((t & 1) == 0) z = A(0.25) * u[t / 2] + A(0.75) * u[t / 2 + 1];
      else z = A(0.75) * u[t / 2 + 1] + A(0.25) * u[t / 2 + 2];
      out[t] = z * (A)a.scale;
    }
  } else {
#pragma unroll
    for (int t = 0; t < VZ; ++t) {
      int k0, k1;
      float lk;
      up_axis(c.k0 + t, m3, a.nearest, k0, k1, lk);
      const A w0 = A(1) - (A)lk, w1 = (A)lk;
      const A v0 = hj0 * (w0 * (A)r00[k0] + w1 * (A)r00[k1]) + hj1 * (w0 * (A)r01[k0] + w1 * (A)r01[k1]);
      const A v1 = hj0 * (w0 * (A)r10[k0] + w1 * (A)r10[k1]) + hj1 * (w0 * (A)r11[k0] + w1 * (A)r11[k1]);
      out[t] = (hi0 * v0 + hi1 * v1) * (A)a.scale;
    }
  }
  if (a.prev) {
#pragma unroll
    for (int t = 0; t < VZ; ++t) out[t] = c0 * (A)pv[t] + c1 * out[t];
  }
  store_vec<VZ>((TO*)a.out + off, out);
}

// Adjoint of upsample2 (its transpose as a linear map), as a gather over the fine
// cells that read each coarse cell.  (n1, n2, n3) are the COARSE dims.
__device__ __forceinline__ float up_weight(int f, int cidx, int nc, bool nearest) {
  int i0, i1;
  float lam;
  up_axis(f, nc, nearest, i0, i1, lam);
  float w = 0.f;
  if (i0 == cidx) w += 1.f - lam;
  if (i1 == cidx) w += lam;
  return w;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kThreads) void upsample2_adjoint_kernel(const TI* __restrict__ gin,
                                                                     TO* __restrict__ gout, int B,
                                                                     int n1, int n2, int n3,
                                                                     int nearest, double scale) {
  typedef typename AccOf<TO>::type A;
  const long Vc = (long)n1 * n2 * n3;
  const long id = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (id >= Vc) return;
  const int b = blockIdx.y;
  const int k = (int)(id % n3);
  const int j = (int)((id / n3) % n2);
  const int i = (int)(id / ((long)n2 * n3));
  const int f1 = 2 * n1, f2 = 2 * n2, f3 = 2 * n3;
  const TI* __restrict__ g = gin + b * Vc * 8;
  A acc = A(0);
  for (int fi = max(2 * i - 1, 0); fi <= min(2 * i + 2, f1 - 1); ++fi) {
    const A wi = (A)up_weight(fi, i, n1, nearest);
    if (wi == A(0)) continue;
    for (int fj = max(2 * j - 1, 0); fj <= min(2 * j + 2, f2 - 1); ++fj) {
      const A wj = (A)up_weight(fj, j, n2, nearest);
      if (wj == A(0)) continue;
      const TI* row = g + ((long)fi * f2 + fj) * f3;
      A s = A(0);
      for (int fk = max(2 * k - 1, 0); fk <= min(2 * k + 2, f3 - 1); ++fk) {
        const A wk = (A)up_weight(fk, k, n3, nearest);
        s += wk * (A)row[fk];
      }
      acc += wi * wj * s;
    }
  }
  gout[b * Vc + id] = (TO)(acc * (A)scale);
}

// The same gather for runs of VZ coarse cells along z (coarse n3 % VZ == 0): per fine row
// of the 4 x 4 window one vector load of the 2 VZ fine cells under the run plus the two
// cells either side, instead of 4 scalar loads per row and coarse cell; the weights and the
// summation order (fine rows ascending, fine z ascending within a row) are those above.
template <typename TI, typename TO, int VZ>
__global__ __launch_bounds__(kThreads) void upsample2_adjoint_run_kernel(const TI* __restrict__ gin,
                                                                         TO* __restrict__ gout, int B,
                                                                         int n1, int n2, int n3,
                                                                         int nearest, double scale) {
  typedef typename AccOf<TO>::type A;
  ChunkPos c;
  const long cid = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (!chunk_pos<VZ>(cid, n1, n2, n3, c)) return;
  const int b = blockIdx.y;
  const long Vc = (long)n1 * n2 * n3;
  const int f1 = 2 * n1, f2 = 2 * n2, f3 = 2 * n3;
  const TI* __restrict__ g = gin + b * Vc * 8;
  // z weights of the run: coarse cell k0 + t reads fine z 2(k0 + t) - 1 + d, d = 0..3
  A wz[VZ][4];
#pragma unroll
  for (int t = 0; t < VZ; ++t)
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int fk = 2 * (c.k0 + t) - 1 + d;
      wz[t][d] = (fk >= 0 && fk < f3) ? (A)up_weight(fk, c.k0 + t, n3, nearest) : A(0);
    }
  const int z0 = 2 * c.k0;  // first fine cell under the run
  A acc[VZ];
#pragma unroll
  for (int t = 0; t < VZ; ++t) acc[t] = A(0);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int fi = 2 * c.i - 1 + a;
    if (fi < 0 || fi >= f1) continue;
    const A wi = (A)up_weight(fi, c.i, n1, nearest);
    if (wi == A(0)) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int fj = 2 * c.j - 1 + e;
      if (fj < 0 || fj >= f2) continue;
      const A wj = (A)up_weight(fj, c.j, n2, nearest);
      if (wj == A(0)) continue;
      const TI* row = g + ((long)fi * f2 + fj) * f3;
      A f[2 * VZ + 2];  // fine z0 - 1 .. z0 + 2 VZ
      load_vec<2 * VZ>(row + z0, f + 1);
      f[0] = z0 > 0 ? (A)row[z0 - 1] : A(0);
      f[2 * VZ + 1] = z0 + 2 * VZ < f3 ? (A)row[z0 + 2 * VZ] : A(0);
#pragma unroll
      for (int t = 0; t < VZ; ++t) {
        A sum = A(0);
#pragma unroll
        for (int d = 0; d < 4; ++d)
          if (wz[t][d] != A(0)) sum += wz[t][d] * f[2 * t + d];
        acc[t] += wi * wj * sum;
      }
    }
  }
  A out[VZ];
#pragma unroll
  for (int t = 0; t < VZ; ++t) out[t] = acc[t] * (A)scale;
  store_vec<VZ>(gout + b * Vc + c.v0, out);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int pick_vz(int n3, int maxvz) {
  for (int vz = maxvz; vz > 1; vz >>= 1)
    if (n3 % vz == 0) return vz;
  return 1;
}

// Small launches (the coarse levels) are latency-bound chains of memory round trips:
// shorter runs per thread give more threads for the same voxels.  Halve the run length
// while a launch would have fewer than kMinItems threads.
constexpr long kMinItems = 1L << 18;

static int pick_vz_small(int n3, int maxvz, long voxels) {
  int vz = pick_vz(n3, maxvz);
  while (vz > 1 && voxels / vz < kMinItems) vz >>= 1;
  return vz;
}

// Launches over fewer voxels than this skip the solver's done flag: their run time is a
// chain of memory round trips, and the flag load would add one; the work a speculative
// iteration then does on them is a few KB.
constexpr long kDoneCheckMinVoxels = 1L << 20;

static inline const int* done_for(const int* done, long voxels) {
  return voxels >= kDoneCheckMinVoxels ? done : nullptr;
}

// grids up to this size (32^3) compute their stencils with one thread per (tap, voxel);
// at 64^3 its 27 gathers per thread cost more than the per-voxel kernel (54 vs 20 us)
constexpr long kSmallWeightsVoxels = 1L << 15;

static inline dim3 grid1(long items, int y = 1) {
  return dim3((unsigned)((items + kThreads - 1) / kThreads), (unsigned)y, 1);
}

template <typename TI, typename TK>
static void launch_weights_t(const void* img, const void* W, const void* b, void* K, int C, int n1,
                             int n2, int n3, bool trans, hipStream_t st) {
  typedef typename AccOf<TK>::type A;
  const long V = (long)n1 * n2 * n3;
  if (!trans && V <= kSmallWeightsVoxels) {
    const long items = 27 * V;
    if (C == 1)
      stencil_weights_small_kernel<TI, TK, 1><<<grid1(items), kThreads, 0, st>>>(
          (const TI*)img, (const A*)W, (const A*)b, (TK*)K, n1, n2, n3);
    else if (C == 2)
      stencil_weights_small_kernel<TI, TK, 2><<<grid1(items), kThreads, 0, st>>>(
          (const TI*)img, (const A*)W, (const A*)b, (TK*)K, n1, n2, n3);
    else
      stencil_weights_small_kernel<TI, TK, 3><<<grid1(items), kThreads, 0, st>>>(
          (const TI*)img, (const A*)W, (const A*)b, (TK*)K, n1, n2, n3);
    return;
  }
  const int maxvz = (C == 1) ? 4 : 2;
  const int vz = pick_vz(n3, maxvz);
  const long items = (long)n1 * n2 * (n3 / vz);
#define MLPCG_SW(CC, VZZ)                                                                      \
  if (C == CC && vz == VZZ) {                                                                  \
    if (trans)                                                                                 \
      stencil_weights_t_kernel<TI, TK, CC, VZZ><<<grid1(items), kThreads, 0, st>>>(           \
          (const TI*)img, (const A*)W, (const A*)b, (TK*)K, n1, n2, n3);                       \
    else                                                                                       \
      stencil_weights_kernel<TI, TK, CC, VZZ><<<grid1(items), kThreads, 0, st>>>(             \
          (const TI*)img, (const A*)W, (const A*)b, (TK*)K, n1, n2, n3);                       \
    return;                                                                                    \
  }
  MLPCG_SW(1, 4) MLPCG_SW(1, 2) MLPCG_SW(1, 1)
  MLPCG_SW(2, 2) MLPCG_SW(2, 1)
  MLPCG_SW(3, 2) MLPCG_SW(3, 1)
#undef MLPCG_SW
}

int launch_stencil_weights(int dtype, const void* img, const void* W, const void* b, void* K, int C,
                           int n1, int n2, int n3, bool trans, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "stencil_weights: C must be 1, 2 or 3 (got %d)", C);
  switch (dtype) {
    case kF64: launch_weights_t<double, double>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    case kF32: launch_weights_t<float, float>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    case kF16: launch_weights_t<f16, f16>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    case kBF16: launch_weights_t<bf16, bf16>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    default: MLPCG_REQUIRE(false, "stencil_weights: bad dtype %d", dtype);
  }
  return check_launch("stencil_weights");
}

int launch_stencil_weights2(int tin, int tk, const void* img, const void* W, const void* b, void* K,
                            int C, int n1, int n2, int n3, bool trans, hipStream_t st) {
  if (tin == tk) return launch_stencil_weights(tk, img, W, b, K, C, n1, n2, n3, trans, st);
  MLPCG_REQUIRE(C >= 1 && C <= 3, "stencil_weights: C must be 1, 2 or 3 (got %d)", C);
  MLPCG_REQUIRE(tin == kF32, "stencil_weights: image type %d with stencil type %d unsupported", tin, tk);
  switch (tk) {
    case kF16: launch_weights_t<float, f16>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    case kBF16: launch_weights_t<float, bf16>(img, W, b, K, C, n1, n2, n3, trans, st); break;
    default: MLPCG_REQUIRE(false, "stencil_weights: bad stencil type %d", tk);
  }
  return check_launch("stencil_weights");
}

template <typename TK, typename TX, typename TY, typename TP>
static void launch_apply_t(const ApplyArgs& a, int vz, hipStream_t st) {
  const long items = (long)a.n1 * a.n2 * (a.n3 / vz);
  const int bb = a.B == 1 ? 1 : 4;
  const int gy = (a.B + bb - 1) / bb;
#define MLPCG_AP(VZZ, BBB)                                                                 \
  if (vz == VZZ && bb == BBB) {                                                            \
    sm_apply_kernel<TK, TX, TY, TP, VZZ, BBB><<<grid1(items, gy), kThreads, 0, st>>>(a);  \
    return;                                                                                \
  }
  MLPCG_AP(8, 1) MLPCG_AP(4, 1) MLPCG_AP(2, 1) MLPCG_AP(1, 1)
  MLPCG_AP(8, 4) MLPCG_AP(4, 4) MLPCG_AP(2, 4) MLPCG_AP(1, 4)
#undef MLPCG_AP
}

template <typename TK>
static void launch_apply_tr_t(const ApplyArgs& a, int vz, hipStream_t st) {
  const long items = (long)a.n1 * a.n2 * (a.n3 / vz);
  const int bb = a.B == 1 ? 1 : 4;
  const int gy = (a.B + bb - 1) / bb;
#define MLPCG_AT(VZZ, BBB)                                                                 \
  if (vz == VZZ && bb == BBB) {                                                            \
    sm_apply_t_kernel<TK, TK, VZZ, BBB><<<grid1(items, gy), kThreads, 0, st>>>(a);        \
    return;                                                                                \
  }
  MLPCG_AT(8, 1) MLPCG_AT(4, 1) MLPCG_AT(2, 1) MLPCG_AT(1, 1)
  MLPCG_AT(8, 4) MLPCG_AT(4, 4) MLPCG_AT(2, 4) MLPCG_AT(1, 4)
#undef MLPCG_AT
}

int launch_sm_apply_transposed(int tk, const ApplyArgs& args, hipStream_t st) {
  MLPCG_REQUIRE(args.B >= 1 && args.n1 >= 1 && args.n2 >= 1 && args.n3 >= 1, "sm_apply_t: bad sizes");
  // runs of 4 cells for every storage but FP64: with 8 16-bit cells the 27 K rows and their
  // halos hold 256 VGPRs and the BF16 level-0 gather at 256^3 took 685 us against 534 us
  // (halo cells from neighbouring lanes by DPP shifts instead of 2-byte loads: 639 us with
  // 8 cells, 632 us with 4)
  if (args.B == 1) {
    // edge-exchange gather with 16-byte runs where each z-row fits one wave
    const int ve = tk == kF64 ? 2 : (tk == kF32 ? 4 : 8);
    const int L = args.n3 / ve;
    if (args.n3 % ve == 0 && L <= 64 && 64 % L == 0) {
      const dim3 g = grid1((long)args.n1 * args.n2 * L);
      switch (tk) {
        case kF64: sm_apply_t_edge_kernel<double, 2><<<g, kThreads, 0, st>>>(args); break;
        case kF32: sm_apply_t_edge_kernel<float, 4><<<g, kThreads, 0, st>>>(args); break;
        case kF16: sm_apply_t_edge_kernel<f16, 8><<<g, kThreads, 0, st>>>(args); break;
        case kBF16: sm_apply_t_edge_kernel<bf16, 8><<<g, kThreads, 0, st>>>(args); break;
        default: MLPCG_REQUIRE(false, "sm_apply_t: bad dtype %d", tk);
      }
      return check_launch("sm_apply_t");
    }
  }
  const int vz = pick_vz_small(args.n3, tk == kF64 ? 2 : 4, (long)args.n1 * args.n2 * args.n3 * args.B);
  switch (tk) {
    case kF64: launch_apply_tr_t<double>(args, vz, st); break;
    case kF32: launch_apply_tr_t<float>(args, vz, st); break;
    case kF16: launch_apply_tr_t<f16>(args, vz, st); break;
    case kBF16: launch_apply_tr_t<bf16>(args, vz, st); break;
    default: MLPCG_REQUIRE(false, "sm_apply_t: bad dtype %d", tk);
  }
  return check_launch("sm_apply_t");
}

int launch_sm_apply(int tk, int tx, int ty, const ApplyArgs& args, hipStream_t st) {
  MLPCG_REQUIRE(args.B >= 1 && args.n1 >= 1 && args.n2 >= 1 && args.n3 >= 1, "sm_apply: bad sizes");
  ApplyArgs a = args;
  a.done = done_for(a.done, (long)a.n1 * a.n2 * a.n3);
  // 16-byte stencil loads per tap and thread: VZ = 8 (16-bit), 4 (FP32), 2 (FP64); wider
  // runs hoist more loads than the register file holds at two waves per SIMD
  const int vz = pick_vz_small(a.n3, tk == kF64 ? 2 : (tk == kF32 ? 4 : 8), (long)a.n1 * a.n2 * a.n3 * a.B);
  if (tk == kF64) {
    MLPCG_REQUIRE(tx == kF64 && ty == kF64, "sm_apply: f64 stencil needs f64 vectors");
    launch_apply_t<double, double, double, double>(a, vz, st);
  } else if (tk == kF32) {
    MLPCG_REQUIRE(tx == kF32 && ty == kF32, "sm_apply: f32 stencil needs f32 vectors");
    launch_apply_t<float, float, float, float>(a, vz, st);
  } else if (tk == kBF16 || tk == kF16) {
    const bool ok = (tx == tk && ty == tk) || (tx == kF32 && ty == tk) || (tx == tk && ty == kF32);
    MLPCG_REQUIRE(ok, "sm_apply: unsupported vector types (%d,%d) for stencil type %d", tx, ty, tk);
    if (tk == kBF16) {
      if (tx == tk && ty == tk) launch_apply_t<bf16, bf16, bf16, bf16>(a, vz, st);
      else if (tx == kF32) launch_apply_t<bf16, float, bf16, bf16>(a, vz, st);
      else launch_apply_t<bf16, bf16, float, bf16>(a, vz, st);
    } else {
      if (tx == tk && ty == tk) launch_apply_t<f16, f16, f16, f16>(a, vz, st);
      else if (tx == kF32) launch_apply_t<f16, float, f16, f16>(a, vz, st);
      else launch_apply_t<f16, f16, float, f16>(a, vz, st);
    }
  } else {
    MLPCG_REQUIRE(false, "sm_apply: bad dtype %d", tk);
  }
  return check_launch("sm_apply");
}

template <typename T>
static void launch_pool_t(const void* in, void* out, int B, int n1, int n2, int n3, const int* done,
                          hipStream_t st) {
  const int vz = pick_vz_small(n3, 4, (long)n1 * n2 * n3 * B);
  const long items = (long)n1 * n2 * (n3 / vz);
  if (vz == 4)
    avgpool2_kernel<T, T, 4><<<grid1(items, B), kThreads, 0, st>>>((const T*)in, (T*)out, B, n1, n2, n3, done);
  else if (vz == 2)
    avgpool2_kernel<T, T, 2><<<grid1(items, B), kThreads, 0, st>>>((const T*)in, (T*)out, B, n1, n2, n3, done);
  else
    avgpool2_kernel<T, T, 1><<<grid1(items, B), kThreads, 0, st>>>((const T*)in, (T*)out, B, n1, n2, n3, done);
}

int launch_avgpool2(int dtype, const void* in, void* out, int B, int n1, int n2, int n3,
                    const int* done, hipStream_t st) {
  MLPCG_REQUIRE(n1 % 2 == 0 && n2 % 2 == 0 && n3 % 2 == 0,
                "avgpool2: grid (%d,%d,%d) must be even in every dimension", n1, n2, n3);
  const int c1 = n1 / 2, c2 = n2 / 2, c3 = n3 / 2;
  done = done_for(done, (long)n1 * n2 * n3);
  switch (dtype) {
    case kF64: launch_pool_t<double>(in, out, B, c1, c2, c3, done, st); break;
    case kF32: launch_pool_t<float>(in, out, B, c1, c2, c3, done, st); break;
    case kF16: launch_pool_t<f16>(in, out, B, c1, c2, c3, done, st); break;
    case kBF16: launch_pool_t<bf16>(in, out, B, c1, c2, c3, done, st); break;
    default: MLPCG_REQUIRE(false, "avgpool2: bad dtype %d", dtype);
  }
  return check_launch("avgpool2");
}

// 16-bit outputs run 16 cells per thread (16-byte coarse loads, two 16-byte stores)
template <typename TI, typename TO>
static void launch_up_t(const UpArgs& a, hipStream_t st) {
  constexpr bool kNarrow = sizeof(TO) == 2 && sizeof(TI) == 2;
  const int vz = pick_vz_small(a.n3, kNarrow ? 16 : 8, (long)a.n1 * a.n2 * a.n3 * a.B);
  const long items = (long)a.n1 * a.n2 * (a.n3 / vz);
  if constexpr (kNarrow) {
    if (vz == 16) {
      upsample2_kernel<TI, TO, TO, 16><<<grid1(items, a.B), kThreads, 0, st>>>(a);
      return;
    }
  }
  if (vz == 8) upsample2_kernel<TI, TO, TO, 8><<<grid1(items, a.B), kThreads, 0, st>>>(a);
  else if (vz == 4) upsample2_kernel<TI, TO, TO, 4><<<grid1(items, a.B), kThreads, 0, st>>>(a);
  else if (vz == 2) upsample2_kernel<TI, TO, TO, 2><<<grid1(items, a.B), kThreads, 0, st>>>(a);
  else upsample2_kernel<TI, TO, TO, 1><<<grid1(items, a.B), kThreads, 0, st>>>(a);
}

int launch_upsample2(int dtype, const UpArgs& args, hipStream_t st) {
  MLPCG_REQUIRE(args.n1 % 2 == 0 && args.n2 % 2 == 0 && args.n3 % 2 == 0,
                "upsample2: fine grid (%d,%d,%d) must be even", args.n1, args.n2, args.n3);
  UpArgs a = args;
  a.done = done_for(a.done, (long)a.n1 * a.n2 * a.n3);
  switch (dtype) {
    case kF64: launch_up_t<double, double>(a, st); break;
    case kF32: launch_up_t<float, float>(a, st); break;
    case kF16: launch_up_t<f16, f16>(a, st); break;
    case kBF16: launch_up_t<bf16, bf16>(a, st); break;
    default: MLPCG_REQUIRE(false, "upsample2: bad dtype %d", dtype);
  }
  return check_launch("upsample2");
}

template <typename T>
static bool launch_up_adj_runs(const void* gin, void* gout, int B, int n1, int n2, int n3, int nearest,
                               double scale, hipStream_t st) {
  const int vz = pick_vz_small(n3, 4, (long)n1 * n2 * n3 * B);
  if (vz < 2) return false;
  const long items = (long)n1 * n2 * (n3 / vz);
  if (vz == 4)
    upsample2_adjoint_run_kernel<T, T, 4><<<grid1(items, B), kThreads, 0, st>>>((const T*)gin, (T*)gout, B, n1, n2,
                                                                              n3, nearest, scale);
  else
    upsample2_adjoint_run_kernel<T, T, 2><<<grid1(items, B), kThreads, 0, st>>>((const T*)gin, (T*)gout, B, n1, n2,
                                                                              n3, nearest, scale);
  return true;
}

int launch_upsample2_adjoint(int dtype, const void* gin, void* gout, int B, int n1, int n2, int n3,
                             int nearest, double scale, hipStream_t st) {
  const long items = (long)n1 * n2 * n3;
  // runs of coarse cells where the coarse rows allow them (16-bit and FP32 vectors)
  if ((dtype == kBF16 && launch_up_adj_runs<bf16>(gin, gout, B, n1, n2, n3, nearest, scale, st)) ||
      (dtype == kF16 && launch_up_adj_runs<f16>(gin, gout, B, n1, n2, n3, nearest, scale, st)) ||
      (dtype == kF32 && launch_up_adj_runs<float>(gin, gout, B, n1, n2, n3, nearest, scale, st)))
    return check_launch("upsample2_adjoint");
  switch (dtype) {
    case kF64:
      upsample2_adjoint_kernel<double, double><<<grid1(items, B), kThreads, 0, st>>>(
          (const double*)gin, (double*)gout, B, n1, n2, n3, nearest, scale);
      break;
    case kF32:
      upsample2_adjoint_kernel<float, float><<<grid1(items, B), kThreads, 0, st>>>(
          (const float*)gin, (float*)gout, B, n1, n2, n3, nearest, scale);
      break;
    case kF16:
      upsample2_adjoint_kernel<f16, f16><<<grid1(items, B), kThreads, 0, st>>>(
          (const f16*)gin, (f16*)gout, B, n1, n2, n3, nearest, scale);
      break;
    case kBF16:
      upsample2_adjoint_kernel<bf16, bf16><<<grid1(items, B), kThreads, 0, st>>>(
          (const bf16*)gin, (bf16*)gout, B, n1, n2, n3, nearest, scale);
      break;
    default: MLPCG_REQUIRE(false, "upsample2_adjoint: bad dtype %d", dtype);
  }
  return check_launch("upsample2_adjoint");
}

}  // namespace mlpcg
