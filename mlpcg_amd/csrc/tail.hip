// The two coarsest levels of the SmallSMModelDn3D V-cycle in one workgroup.
//
// At 256^3, l=5 the levels 16^3 and 8^3 hold 4 096 and 512 voxels: each of their seven
// kernels (pool, pre, pool, l, upsample, post + combine, upsample) is a chain of memory
// round trips of ~5-10 us.  Here one 1024-thread workgroup runs the seven stages with
// barriers between them and keeps the level vectors in LDS (in the plan's storage type,
// so every intermediate is rounded exactly where the per-kernel sequence stores it):
//
//   a  = pool(x_in)                   x_in: output of pre[n-2], level n-2 (global)
//   xf = pre[n-1](a)                  level n-1 = f
//   b  = pool(xf);  xc = l(b)         level n   = c
//   u  = up(xc)
//   zf = c0 * xf + c1 * post[n-1](u)  (sm_model_3d.py:282-286, swap = False)
//   out = up(zf)                      level n-2 (global), the input of post[n-2]
//
// Same formulas as avgpool2_kernel, upsample2_kernel (trilinear or nearest) and
// sm_apply_kernel (stored stencils K [27][V], zero outside the grid).
#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

namespace {

constexpr int kTailThreads = 1024;

struct Dims {
  int n1, n2, n3;
  __device__ int size() const { return n1 * n2 * n3; }
};

__device__ __forceinline__ void decompose(int v, const Dims& d, int& i, int& j, int& k) {
  k = v % d.n3;
  const int t = v / d.n3;
  j = t % d.n2;
  i = t / d.n2;
}

// 1-D x2 upsample source indices and weight (align_corners=False), as up_axis.
__device__ __forceinline__ void up1(int f, int nc, bool nearest, int& i0, int& i1, float& lam) {
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

// coarse (dims c) -> fine value at (i, j, k) of the x2 grid
template <typename T>
__device__ __forceinline__ float up_at(const T* src, const Dims& c, int i, int j, int k, bool nearest) {
  int i0, i1, j0, j1, k0, k1;
  float li, lj, lk;
  up1(i, c.n1, nearest, i0, i1, li);
  up1(j, c.n2, nearest, j0, j1, lj);
  up1(k, c.n3, nearest, k0, k1, lk);
  auto at = [&](int a, int b, int e) { return (float)src[(a * c.n2 + b) * c.n3 + e]; };
  const float w0 = 1.f - lk, w1 = lk;
  const float v0 = (1.f - lj) * (w0 * at(i0, j0, k0) + w1 * at(i0, j0, k1)) + lj * (w0 * at(i0, j1, k0) + w1 * at(i0, j1, k1));
  const float v1 = (1.f - lj) * (w0 * at(i1, j0, k0) + w1 * at(i1, j0, k1)) + lj * (w0 * at(i1, j1, k0) + w1 * at(i1, j1, k1));
  return (1.f - li) * v0 + li * v1;
}

// mean of the 2x2x2 fine cells of coarse voxel (i, j, k); fine dims f
template <typename T>
__device__ __forceinline__ float pool_at(const T* src, const Dims& f, int i, int j, int k) {
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const T* row = src + ((2 * i + a) * f.n2 + 2 * j + b) * f.n3 + 2 * k;
      s += (float)row[0] + (float)row[1];
    }
  return s * 0.125f;
}

// y(v) = sum_p K[p][v] src(v + d_p), zero outside the grid, for the voxels v and v2 (v2
// skipped when >= V): the 27 (54) stencil loads are issued before any use, and the taps
// read LDS at a clamped index and select zero, so nothing branches between the loads.
template <typename T>
__device__ __forceinline__ void apply_pair(const T* __restrict__ K, const T* src, const Dims& d, int v, int v2,
                                           float& y, float& y2) {
  const int V = d.size();
  const bool has2 = v2 < V;
  const int w2 = has2 ? v2 : v;
  T kv[27], kv2[27];  // kept in the storage type (registers)
#pragma unroll
  for (int p = 0; p < 27; ++p) {
    kv[p] = K[(long)p * V + v];
    kv2[p] = K[(long)p * V + w2];
  }
  int i, j, k, i2, j2, k2;
  decompose(v, d, i, j, k);
  decompose(w2, d, i2, j2, k2);
  float acc = 0.f, acc2 = 0.f;
#pragma unroll
  for (int p = 0; p < 27; ++p) {
    const int a1 = p / 9 - 1, a2 = (p / 3) % 3 - 1, a3 = p % 3 - 1;
    {
      const int ii = i + a1, jj = j + a2, kk = k + a3;
      const bool in = ii >= 0 && ii < d.n1 && jj >= 0 && jj < d.n2 && kk >= 0 && kk < d.n3;
      const float xv = (float)src[in ? (ii * d.n2 + jj) * d.n3 + kk : v];
      acc += (float)kv[p] * (in ? xv : 0.f);
    }
    {
      const int ii = i2 + a1, jj = j2 + a2, kk = k2 + a3;
      const bool in = ii >= 0 && ii < d.n1 && jj >= 0 && jj < d.n2 && kk >= 0 && kk < d.n3;
      const float xv = (float)src[in ? (ii * d.n2 + jj) * d.n3 + kk : w2];
      acc2 += (float)kv2[p] * (in ? xv : 0.f);
    }
  }
  y = acc;
  y2 = acc2;
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void coarse_tail_kernel(TailArgs a) {
  // level vectors in LDS: bufA holds pool(x_in) then u; xf; zf; the two coarse vectors
  __shared__ T bufA[kTailMaxFine], xf[kTailMaxFine], zf[kTailMaxFine], bc[kTailMaxFine / 8],
      xc[kTailMaxFine / 8];
  const Dims f{a.f1, a.f2, a.f3};
  const Dims c{a.f1 / 2, a.f2 / 2, a.f3 / 2};
  const Dims g{a.f1 * 2, a.f2 * 2, a.f3 * 2};
  const int Vf = f.size(), Vc = c.size(), Vg = g.size();
  const T* xin = (const T*)a.x_in;
  const T* Kpre = (const T*)a.K_pre;
  const T* Kpost = (const T*)a.K_post;
  const T* Kl = (const T*)a.K_l;
  const bool nearest = a.nearest != 0;
  int i, j, k;
  for (int v = threadIdx.x; v < Vf; v += kTailThreads) {  // a = pool(x_in)
    decompose(v, f, i, j, k);
    bufA[v] = (T)pool_at(xin, g, i, j, k);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < Vf; v += 2 * kTailThreads) {  // xf = pre(a)
    float y, y2;
    apply_pair(Kpre, bufA, f, v, v + kTailThreads, y, y2);
    xf[v] = (T)y;
    if (v + kTailThreads < Vf) xf[v + kTailThreads] = (T)y2;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < Vc; v += kTailThreads) {  // b = pool(xf)
    decompose(v, c, i, j, k);
    bc[v] = (T)pool_at(xf, f, i, j, k);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < Vc; v += 2 * kTailThreads) {  // xc = l(b)
    float y, y2;
    apply_pair(Kl, bc, c, v, v + kTailThreads, y, y2);
    xc[v] = (T)y;
    if (v + kTailThreads < Vc) xc[v + kTailThreads] = (T)y2;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < Vf; v += kTailThreads) {  // u = up(xc)
    decompose(v, f, i, j, k);
    bufA[v] = (T)up_at(xc, c, i, j, k, nearest);
  }
  __syncthreads();
  const float c0 = *a.c0, c1 = *a.c1;
  for (int v = threadIdx.x; v < Vf; v += 2 * kTailThreads) {  // zf = c0 xf + c1 post(u)
    float y, y2;
    apply_pair(Kpost, bufA, f, v, v + kTailThreads, y, y2);
    zf[v] = (T)(c0 * (float)xf[v] + c1 * y);
    if (v + kTailThreads < Vf) zf[v + kTailThreads] = (T)(c0 * (float)xf[v + kTailThreads] + c1 * y2);
  }
  __syncthreads();
  T* out = (T*)a.out;
  for (int v = threadIdx.x; v < Vg; v += kTailThreads) {  // out = up(zf)
    decompose(v, g, i, j, k);
    out[v] = (T)up_at(zf, f, i, j, k, nearest);
  }
}

}  // namespace

bool tail_eligible(int dtype, int f1, int f2, int f3) {
  return dtype != kF64 && f1 % 2 == 0 && f2 % 2 == 0 && f3 % 2 == 0 &&
         (long)f1 * f2 * f3 <= kTailMaxFine;
}

int launch_coarse_tail(int dtype, const TailArgs& a, hipStream_t st) {
  MLPCG_REQUIRE(tail_eligible(dtype, a.f1, a.f2, a.f3), "coarse_tail: level (%d,%d,%d) / dtype %d not supported",
                a.f1, a.f2, a.f3, dtype);
  switch (dtype) {
    case kF32: coarse_tail_kernel<float><<<1, kTailThreads, 0, st>>>(a); break;
    case kF16: coarse_tail_kernel<f16><<<1, kTailThreads, 0, st>>>(a); break;
    case kBF16: coarse_tail_kernel<bf16><<<1, kTailThreads, 0, st>>>(a); break;
    default: MLPCG_REQUIRE(false, "coarse_tail: bad dtype %d", dtype);
  }
  return check_launch("coarse_tail");
}

}  // namespace mlpcg
