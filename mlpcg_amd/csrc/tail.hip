// The two coarsest levels of the SmallSMModelDn3D V-cycle in one workgroup.
//
// At 256^3, l=5 the levels 16^3 and 8^3 hold 4 096 and 512 voxels: each of their seven
// kernels (pool, pre, pool, l, upsample, post + combine, upsample) is a chain of memory
// round trips of ~5-10 us.  Here one 1024-thread workgroup runs the seven stages with
// barriers between them (helper workgroups only warm its L2, see tail_prefetch) and
// keeps the level vectors in LDS as FP32 values rounded to the plan's storage type, so
// every intermediate is rounded exactly where the per-kernel sequence stores it:
//
//   a  = pool(x_in)                   x_in: output of pre[n-2], level n-2 (global)
//   xf = pre[n-1](a)                  level n-1 = f
//   b  = pool(xf);  xc = l(b)         level n   = c
//   u  = up(xc)
//   zf = c0 * xf + c1 * post[n-1](u)  (sm_model_3d.py:282-286, swap = False)
//   out = up(zf)                      level n-2 (global), the input of post[n-2]
//
// The stencil inputs (a, u, b) live in LDS with a one-cell zero halo, so a tap is one
// LDS read at a fixed offset with no bounds test.  Fine-level stages run on pairs of
// consecutive z cells (n3 is even), the final upsample on quadruples.  The formulas and
// summation orders are those of avgpool2_kernel, upsample2_kernel (trilinear or nearest,
// clamped z window) and sm_apply_kernel (per stencil row k0 x0 + k1 x1 + k2 x2).
#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

namespace {

constexpr int kTailThreads = 1024;

// Quotient by a runtime divisor as a high multiply: with m = ceil(2^32 / d),
// floor(v m / 2^32) = floor(v / d) whenever v d < 2^32 (here every index is below 2^15
// and every extent at most 2^13); d = 1 is kept apart (m would need 33 bits).
struct Divider {
  unsigned d, m;
  __device__ explicit Divider(unsigned dd)
      : d(dd), m(dd == 1 ? 0u : (unsigned)((0x100000000ull + dd - 1) / dd)) {}
  __device__ __forceinline__ unsigned quo(unsigned v) const { return d == 1 ? v : __umulhi(v, m); }
};

// v -> (i, j, k) on a grid (., n2, n3) with the dividers of n3 and n2
__device__ __forceinline__ void split(unsigned v, const Divider& d3, const Divider& d2, int& i, int& j,
                                      int& k) {
  const unsigned r = d3.quo(v);
  k = (int)(v - r * d3.d);
  const unsigned ii = d2.quo(r);
  j = (int)(r - ii * d2.d);
  i = (int)ii;
}

template <typename T>
__device__ __forceinline__ float rnd(float x) {
  return (float)(T)x;
}

// Z consecutive values of type T held as loaded (16-bit types stay packed two to a
// register until use, so a run of 27 taps costs 27 Z / 2 registers), read as float.
template <typename T, int Z>
struct Run {
  typedef T vec __attribute__((ext_vector_type(Z)));
  vec r;
  __device__ __forceinline__ void load(const T* p) { r = *reinterpret_cast<const vec*>(p); }
  __device__ __forceinline__ float at(int t) const { return (float)r[t]; }
};
template <typename T, int Z>
struct Run16 {
  unsigned r[Z / 2];
  __device__ __forceinline__ void load(const T* p) {
    if constexpr (Z == 2) {
      r[0] = *reinterpret_cast<const unsigned*>(p);
    } else {
      typedef unsigned uv __attribute__((ext_vector_type(Z / 2)));
      const uv w = *reinterpret_cast<const uv*>(p);
#pragma unroll
      for (int t = 0; t < Z / 2; ++t) r[t] = w[t];
    }
  }
  __device__ __forceinline__ float at(int t) const {
    const unsigned w = r[t >> 1];
    if constexpr (sizeof(T) == 2 && __is_same(T, bf16)) return __uint_as_float((t & 1) ? (w & 0xffff0000u) : (w << 16));
    else return (float)__builtin_bit_cast(f16, (unsigned short)((t & 1) ? (w >> 16) : (w & 0xffffu)));
  }
};
template <int Z> struct Run<bf16, Z> : Run16<bf16, Z> {};
template <int Z> struct Run<f16, Z> : Run16<f16, Z> {};

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

// NZ consecutive fine outputs k = 2m .. 2m+NZ-1 (m = k/2) along z at fine (i, j) from the
// coarse vector src (dims m1, m2, m3, row-major, plain): the z window
// [max(m-1, 0), m .. m+NZ/2-1, min(m+NZ/2, m3-1)] with the compile-time weights of
// upsample2_kernel, then j, then i.
template <int NZ>
__device__ __forceinline__ void up_run(const float* src, int m1, int m2, int m3, int i, int j, int k,
                                       bool nearest, float* out) {
  int i0, i1, j0, j1;
  float li, lj;
  up1(i, m1, nearest, i0, i1, li);
  up1(j, m2, nearest, j0, j1, lj);
  const float hi0 = 1.f - li, hi1 = li, hj0 = 1.f - lj, hj1 = lj;
  constexpr int NS = NZ / 2 + 2;
  const int m = k >> 1;
  const int rows[4] = {(i0 * m2 + j0) * m3, (i0 * m2 + j1) * m3, (i1 * m2 + j0) * m3, (i1 * m2 + j1) * m3};
  float seg[4][NS];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float* r = src + rows[q];
    seg[q][0] = r[max(m - 1, 0)];
#pragma unroll
    for (int s = 0; s < NZ / 2; ++s) seg[q][s + 1] = r[m + s];
    seg[q][NS - 1] = r[min(m + NZ / 2, m3 - 1)];
  }
#pragma unroll
  for (int t = 0; t < NZ; ++t) {
    float z[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (nearest) z[q] = seg[q][t / 2 + 1];
      else if ((t & 1) == 0) z[q] = 0.25f * seg[q][t / 2] + 0.75f * seg[q][t / 2 + 1];
      else z[q] = 0.75f * seg[q][t / 2 + 1] + 0.25f * seg[q][t / 2 + 2];
    }
    const float v0 = hj0 * z[0] + hj1 * z[1];
    const float v1 = hj0 * z[2] + hj1 * z[3];
    out[t] = hi0 * v0 + hi1 * v1;
  }
}

// Stencil product on the fine level for the Z consecutive z cells (k .. k+Z-1) at (i, j):
// kv[p] holds the Z stencil weights of tap p, h points at the haloed cell (i-1, j-1, k-1)
// (row strides hs2, hs3; even offset, so the Z+2 cells of a row are read as float2).
template <int Z, typename R>
__device__ __forceinline__ void apply_run(const R* kv, const float* h, int hs2, int hs3, float* y) {
#pragma unroll
  for (int t = 0; t < Z; ++t) y[t] = 0.f;
#pragma unroll
  for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) {
      float r[Z + 2];
#pragma unroll
      for (int s = 0; s < Z / 2 + 1; ++s) {
        const float2 w = *reinterpret_cast<const float2*>(h + d1 * hs2 + d2 * hs3 + 2 * s);
        r[2 * s] = w.x;
        r[2 * s + 1] = w.y;
      }
      const int p = 9 * d1 + 3 * d2;
#pragma unroll
      for (int t = 0; t < Z; ++t)
        y[t] += kv[p].at(t) * r[t] + kv[p + 1].at(t) * r[t + 1] + kv[p + 2].at(t) * r[t + 2];
    }
}

// One pass: thread t owns the fine run of Z cells starting at Z t (Vf <= Z * kTailThreads)
// and the coarse voxel t.  The global loads of a stage are issued before the barrier that
// precedes it, as soon as registers are free (x_in and, for 16-bit storage, K_pre at
// entry; K_l and K_post after pre; for FP32, K_pre after the first pool and K_post after
// l), so their latency overlaps the LDS stages instead of adding one memory round trip
// per stage.
// Workgroups beside workgroup 0: those on its XCD (workgroups are dispatched to the 8
// XCDs round-robin, so every 8th) read x_in, K_pre, K_l and K_post once in 16-byte
// pieces so that workgroup 0 finds them in its XCD's L2: one CU alone keeps too few
// misses in flight to stream ~0.5 MB from HBM quickly.  The rest exit at once.
constexpr int kTailHelpers = 15;

__device__ __forceinline__ unsigned touch(const void* base, long bytes, int helper) {
  const uint4* p = (const uint4*)base;
  const long n = bytes / 16;
  unsigned acc = 0;
  for (long c = (long)helper * kTailThreads + threadIdx.x; c < n; c += (long)kTailHelpers * kTailThreads) {
    const uint4 w = p[c];
    acc ^= w.x ^ w.w;
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ void tail_prefetch(const TailArgs& a, int helper) {
  const long vf = (long)a.f1 * a.f2 * a.f3, s = sizeof(T);
  unsigned acc = touch(a.x_in, 8 * vf * s, helper);
  acc ^= touch(a.K_pre, 27 * vf * s, helper);
  acc ^= touch(a.K_l, 27 * (vf / 8) * s, helper);
  acc ^= touch(a.K_post, 27 * vf * s, helper);
  asm volatile("" ::"v"(acc));  // keeps the loads
}

template <typename T, int Z>
__global__ __launch_bounds__(kTailThreads) void coarse_tail_kernel(TailArgs a) {
  if (solve_done(a.done)) return;  // uniform over the grid: a queued iteration after convergence
  if (blockIdx.x != 0) {
    if (blockIdx.x % 8 == 0) tail_prefetch<T>(a, blockIdx.x / 8 - 1);
    return;
  }
  __shared__ __align__(16) float H[kTailMaxHalo];         // haloed fine: a, then u
  __shared__ __align__(16) float X[kTailMaxFine];         // xf, then zf in place
  __shared__ __align__(16) float Hc[kTailMaxHaloCoarse];  // haloed coarse: b
  __shared__ __align__(16) float Xc[kTailMaxFine / 8];    // xc
  constexpr bool kNarrow = sizeof(T) == 2;  // K_pre fits in registers beside x_in and K_l
  const int tid = threadIdx.x;
  const int f1 = a.f1, f2 = a.f2, f3 = a.f3;
  const int c1 = f1 >> 1, c2 = f2 >> 1, c3 = f3 >> 1;
  const int g2 = 2 * f2, g3 = 2 * f3;
  const int Vf = f1 * f2 * f3, Vc = Vf >> 3, Vg = Vf << 3;
  const int hs3 = f3 + 2, hs2 = (f2 + 2) * hs3;  // haloed strides (fine)
  const int cs3 = c3 + 2, cs2 = (c2 + 2) * cs3;  // haloed strides (coarse)
  const Divider df3(f3), df2(f2), dc3(c3), dc2(c2), dg3(g3), dg2(g2);
  const T* __restrict__ xin = (const T*)a.x_in;
  const T* __restrict__ Kpre = (const T*)a.K_pre;
  const T* __restrict__ Kpost = (const T*)a.K_post;
  const T* __restrict__ Kl = (const T*)a.K_l;
  const bool nearest = a.nearest != 0;
  const float c0 = *a.c0, c1v = *a.c1;

  const bool own = tid < Vf / Z, ownc = tid < Vc;
  const int v = own ? Z * tid : 0, vc = ownc ? tid : 0;
  int i, j, k, ic, jc, kc;
  split(v, df3, df2, i, j, k);
  split(vc, dc3, dc2, ic, jc, kc);
  const int hf = i * hs2 + j * hs3 + k;  // haloed index of (i-1, j-1, k-1)

  Run<T, 2 * Z> xr[4];
  Run<T, Z> kp[27], kq[27];
  T kl[27];
  if (own) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      xr[r].load(xin + ((long)(2 * i + (r >> 1)) * g2 + 2 * j + (r & 1)) * g3 + 2 * k);
    if constexpr (kNarrow) {
#pragma unroll
      for (int p = 0; p < 27; ++p) kp[p].load(Kpre + (long)p * Vf + v);
    }
  }
  for (int e = tid; e < (f1 + 2) * hs2; e += kTailThreads) H[e] = 0.f;
  for (int e = tid; e < (c1 + 2) * cs2; e += kTailThreads) Hc[e] = 0.f;
  __syncthreads();
  if (own) {  // a = pool(x_in)
    float s[Z];
#pragma unroll
    for (int t = 0; t < Z; ++t) s[t] = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < Z; ++t) s[t] += xr[r].at(2 * t) + xr[r].at(2 * t + 1);
#pragma unroll
    for (int t = 0; t < Z; ++t) H[hf + hs2 + hs3 + 1 + t] = rnd<T>(s[t] * 0.125f);
    if constexpr (!kNarrow) {
#pragma unroll
      for (int p = 0; p < 27; ++p) kp[p].load(Kpre + (long)p * Vf + v);
    }
  }
  __syncthreads();
  if (own) {  // xf = pre(a)
    float y[Z];
    apply_run<Z>(kp, H + hf, hs2, hs3, y);
#pragma unroll
    for (int t = 0; t < Z; ++t) X[v + t] = rnd<T>(y[t]);
  }
  if (ownc) {  // K_l first: l waits for these loads only (the load counter drains in order)
#pragma unroll
    for (int p = 0; p < 27; ++p) kl[p] = Kl[(long)p * Vc + vc];
  }
  if constexpr (kNarrow) {
    if (own) {
#pragma unroll
      for (int p = 0; p < 27; ++p) kq[p].load(Kpost + (long)p * Vf + v);
    }
  }
  __syncthreads();
  if (ownc) {  // b = pool(xf)
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float2 w = *reinterpret_cast<const float2*>(X + ((2 * ic + (r >> 1)) * f2 + 2 * jc + (r & 1)) * f3 + 2 * kc);
      s += w.x + w.y;
    }
    Hc[(ic + 1) * cs2 + (jc + 1) * cs3 + kc + 1] = rnd<T>(s * 0.125f);
  }
  __syncthreads();
  if (ownc) {  // xc = l(b)
    const float* h = Hc + ic * cs2 + jc * cs3 + kc;
    float acc = 0.f;
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) {
        const float* r = h + d1 * cs2 + d2 * cs3;
        const int p = 9 * d1 + 3 * d2;
        acc += (float)kl[p] * r[0] + (float)kl[p + 1] * r[1] + (float)kl[p + 2] * r[2];
      }
    Xc[vc] = rnd<T>(acc);
  }
  if constexpr (!kNarrow) {
    if (own) {
#pragma unroll
      for (int p = 0; p < 27; ++p) kq[p].load(Kpost + (long)p * Vf + v);
    }
  }
  __syncthreads();
  if (own) {  // u = up(xc)
    float u[Z];
    up_run<Z>(Xc, c1, c2, c3, i, j, k, nearest, u);
#pragma unroll
    for (int t = 0; t < Z; ++t) H[hf + hs2 + hs3 + 1 + t] = rnd<T>(u[t]);
  }
  __syncthreads();
  if (own) {  // zf = c0 xf + c1 post(u), in place
    float y[Z];
    apply_run<Z>(kq, H + hf, hs2, hs3, y);
#pragma unroll
    for (int t = 0; t < Z; ++t) X[v + t] = rnd<T>(c0 * X[v + t] + c1v * y[t]);
  }
  __syncthreads();
  T* __restrict__ out = (T*)a.out;
  for (int q = tid; q < Vg / (2 * Z); q += kTailThreads) {  // out = up(zf)
    const int w = 2 * Z * q;
    int io, jo, ko;
    split(w, dg3, dg2, io, jo, ko);
    float o[2 * Z];
    up_run<2 * Z>(X, f1, f2, f3, io, jo, ko, nearest, o);
    store_vec<2 * Z>(out + w, o);
  }
}

// run length along z: 4 cells for 16-bit storage when n3 allows it, else 2 (an FP32 run
// of 4 needs more registers for its 27 taps than a 1024-thread workgroup has)
int tail_run(int dtype, int f3) { return (dtype != kF32 && f3 % 4 == 0) ? 4 : 2; }

template <typename T>
void launch_tail_t(bool z4, const TailArgs& a, hipStream_t st) {
  const unsigned grid = 8 * kTailHelpers + 1;
  if constexpr (sizeof(T) == 2) {
    if (z4) {
      coarse_tail_kernel<T, 4><<<grid, kTailThreads, 0, st>>>(a);
      return;
    }
  }
  coarse_tail_kernel<T, 2><<<grid, kTailThreads, 0, st>>>(a);
}

}  // namespace

bool tail_eligible(int dtype, int f1, int f2, int f3) {
  if (dtype == kF64 || f1 % 2 || f2 % 2 || f3 % 2 || f1 <= 0 || f2 <= 0 || f3 <= 0) return false;
  const long vf = (long)f1 * f2 * f3;
  const long halo = (long)(f1 + 2) * (f2 + 2) * (f3 + 2);
  const long halo_c = (long)(f1 / 2 + 2) * (f2 / 2 + 2) * (f3 / 2 + 2);
  return vf <= kTailMaxFine && vf <= (long)tail_run(dtype, f3) * kTailThreads && halo <= kTailMaxHalo &&
         halo_c <= kTailMaxHaloCoarse;
}

int launch_coarse_tail(int dtype, const TailArgs& a, hipStream_t st) {
  MLPCG_REQUIRE(tail_eligible(dtype, a.f1, a.f2, a.f3), "coarse_tail: level (%d,%d,%d) / dtype %d not supported",
                a.f1, a.f2, a.f3, dtype);
  const bool z4 = tail_run(dtype, a.f3) == 4;
  switch (dtype) {
    case kF32: launch_tail_t<float>(z4, a, st); break;
    case kF16: launch_tail_t<f16>(z4, a, st); break;
    case kBF16: launch_tail_t<bf16>(z4, a, st); break;
    default: MLPCG_REQUIRE(false, "coarse_tail: bad dtype %d", dtype);
  }
  return check_launch("coarse_tail");
}

}  // namespace mlpcg
