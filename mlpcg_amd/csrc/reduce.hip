// Reductions of the SM preconditioner: the SMLinear mean patch (forward) and the SM
// block weight gradient (backward).  Both are deterministic two-pass reductions
// (block partials in a workspace, then one ordered pass), with no global atomics.
//
//   mean patch  ybar[s,e] = (1/V) sum_v img[s, v+e]          sm_linear_3d_kernel.cu:62-83
//               z = sum_{s,e} w[s,e] ybar[s,e] + b           sm_linear_3d_kernel.cu:185-186
//   weight grad G_p(v) = sum_b a[b,v] c[b,v+d_p]              sm_block_3d_kernel.cu:121-229
//               gW[p,s,e] = sum_v G_p(v) img[s,v+e], gb[p] = sum_v G_p(v)
//               i.e. the [27 x V] . [V x (27C+1)] product of per-voxel outer products:
//               on the matrix cores (FP32 operands for one channel; BF16 operands with
//               exact hi/lo products for BF16 data, any C), else tiled through LDS with a
//               3x4 register tile per thread.
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

// ---------------------------------------------------------------------------
// Mean patch.  A voxel u contributes to ybar[s,e] iff u - e is inside the grid, i.e.
// per axis: e=-1 -> u <= n-2, e=0 -> always, e=+1 -> u >= 1.  Each thread sums a
// z-run of VZ voxels once and spreads it over the 27 offsets with those masks.
// ---------------------------------------------------------------------------
template <typename TI, int C, int VZ>
__global__ __launch_bounds__(kThreads) void mean_patch_partial_kernel(const TI* __restrict__ img,
                                                                      int n1, int n2, int n3,
                                                                      double* __restrict__ partials) {
  typedef typename AccOf<TI>::type A;
  constexpr int NV = 27 * C;
  A acc[NV];
#pragma unroll
  for (int o = 0; o < NV; ++o) acc[o] = A(0);
  const long V = (long)n1 * n2 * n3;
  const int n3c = n3 / VZ;
  const long nch = (long)n1 * n2 * n3c;
  for (long cid = (long)blockIdx.x * kThreads + threadIdx.x; cid < nch;
       cid += (long)gridDim.x * kThreads) {
    const int kc = (int)(cid % n3c);
    const long t = cid / n3c;
    const int j = (int)(t % n2), i = (int)(t / n2);
    const int k0 = kc * VZ;
    const long v0 = ((long)i * n2 + j) * n3 + k0;
    const A m1[3] = {i <= n1 - 2 ? A(1) : A(0), A(1), i >= 1 ? A(1) : A(0)};
    const A m2[3] = {j <= n2 - 2 ? A(1) : A(0), A(1), j >= 1 ? A(1) : A(0)};
#pragma unroll
    for (int s = 0; s < C; ++s) {
      A v[VZ];
      load_vec<VZ>(img + s * V + v0, v);
      A cs = A(0);
#pragma unroll
      for (int u = 0; u < VZ; ++u) cs += v[u];
      const A first = (k0 == 0) ? v[0] : A(0);
      const A last = (k0 + VZ == n3) ? v[VZ - 1] : A(0);
      const A s3[3] = {cs - last, cs, cs - first};
#pragma unroll
      for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
          for (int a3 = 0; a3 < 3; ++a3) acc[s * 27 + a1 * 9 + a2 * 3 + a3] += m1[a1] * m2[a2] * s3[a3];
    }
  }
  __shared__ double sh[NV][4];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 0; o < NV; ++o) {
    const double s = wave_sum((double)acc[o]);
    if ((threadIdx.x & 63) == 0) sh[o][w] = s;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < NV; o += kThreads)
    partials[(size_t)blockIdx.x * NV + o] = sh[o][0] + sh[o][1] + sh[o][2] + sh[o][3];
}

// ybar = (sum of the block partials) / V, then z0 = w0 . ybar + b0 and, if w1 is given,
// z1 = w1 . ybar + b1: c0 and c1 of one level share their mean patch (sm_model_3d.py:236-239
// evaluates both on the same image), so one reduction serves both.  One wave per output
// (lanes strided over the partials, a wave sum), all outputs in flight at once.
template <typename A>
__global__ __launch_bounds__(kThreads) void mean_patch_final_kernel(
    const double* __restrict__ partials, int nblk, int NV, double invV, const A* __restrict__ w0,
    const A* __restrict__ b0, A* __restrict__ z0, const A* __restrict__ w1, const A* __restrict__ b1,
    A* __restrict__ z1, A* __restrict__ ybar) {
  __shared__ A ys[27 * 3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = wv; o < NV; o += kThreads / 64) {
    double s = 0.0;
    for (int blk = lane; blk < nblk; blk += 64) s += partials[(size_t)blk * NV + o];
    s = wave_sum(s);
    const A yb = (A)(s * invV);
    if (lane == 0) {
      ys[o] = yb;
      ybar[o] = yb;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double c0 = 0.0, c1 = 0.0;
    for (int o = 0; o < NV; ++o) {
      c0 += (double)w0[o] * (double)ys[o];
      if (w1) c1 += (double)w1[o] * (double)ys[o];
    }
    z0[0] = (A)(c0 + (double)b0[0]);
    if (w1) z1[0] = (A)(c1 + (double)b1[0]);
  }
}

static int mean_patch_blocks(int n1, int n2, int n3, int vz) {
  const long nch = (long)n1 * n2 * (n3 / vz);
  long nb = (nch + kThreads - 1) / kThreads;
  if (nb > 1024) nb = 1024;
  return (int)(nb < 1 ? 1 : nb);
}

static int mean_patch_vz(int n3) { return (n3 % 4 == 0) ? 4 : (n3 % 2 == 0 ? 2 : 1); }

size_t mean_patch_workspace(int /*dtype*/, int C, int n1, int n2, int n3) {
  return (size_t)mean_patch_blocks(n1, n2, n3, mean_patch_vz(n3)) * 27 * C * sizeof(double);
}

template <typename TI>
static void launch_mp_t(const void* img, double* part, int nblk, int C, int n1, int n2, int n3, int vz,
                        hipStream_t st) {
#define MLPCG_MP(CC, VZZ)                                                                         \
  if (C == CC && vz == VZZ) {                                                                     \
    mean_patch_partial_kernel<TI, CC, VZZ><<<nblk, kThreads, 0, st>>>((const TI*)img, n1, n2, n3, \
                                                                      part);                      \
    return;                                                                                       \
  }
  MLPCG_MP(1, 4) MLPCG_MP(1, 2) MLPCG_MP(1, 1)
  MLPCG_MP(2, 4) MLPCG_MP(2, 2) MLPCG_MP(2, 1)
  MLPCG_MP(3, 4) MLPCG_MP(3, 2) MLPCG_MP(3, 1)
#undef MLPCG_MP
}

int launch_sm_linear_pair(int dtype, const void* img, const void* w0, const void* b0, void* z0,
                          const void* w1, const void* b1, void* z1, void* ybar, void* workspace, int C,
                          int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "sm_linear: C must be 1, 2 or 3 (got %d)", C);
  MLPCG_REQUIRE(!w1 || (b1 && z1), "sm_linear: a second weight needs its bias and output");
  const int vz = mean_patch_vz(n3);
  const int nblk = mean_patch_blocks(n1, n2, n3, vz);
  double* part = (double*)workspace;
  const double invV = 1.0 / ((double)n1 * n2 * n3);
  switch (dtype) {
    case kF64: launch_mp_t<double>(img, part, nblk, C, n1, n2, n3, vz, st); break;
    case kF32: launch_mp_t<float>(img, part, nblk, C, n1, n2, n3, vz, st); break;
    case kF16: launch_mp_t<f16>(img, part, nblk, C, n1, n2, n3, vz, st); break;
    case kBF16: launch_mp_t<bf16>(img, part, nblk, C, n1, n2, n3, vz, st); break;
    default: MLPCG_REQUIRE(false, "sm_linear: bad dtype %d", dtype);
  }
  if (dtype == kF64)
    mean_patch_final_kernel<double><<<1, kThreads, 0, st>>>(
        part, nblk, 27 * C, invV, (const double*)w0, (const double*)b0, (double*)z0, (const double*)w1,
        (const double*)b1, (double*)z1, (double*)ybar);
  else
    mean_patch_final_kernel<float><<<1, kThreads, 0, st>>>(
        part, nblk, 27 * C, invV, (const float*)w0, (const float*)b0, (float*)z0, (const float*)w1,
        (const float*)b1, (float*)z1, (float*)ybar);
  return check_launch("sm_linear");
}

int launch_sm_linear(int dtype, const void* img, const void* w, const void* b, void* z, void* ybar,
                     void* workspace, int C, int n1, int n2, int n3, hipStream_t st) {
  return launch_sm_linear_pair(dtype, img, w, b, z, nullptr, nullptr, nullptr, ybar, workspace, C, n1, n2, n3,
                               st);
}

// ---------------------------------------------------------------------------
// SM linear without allocation.  The block partials of the mean patch go to a file-scope
// device array instead of a caller workspace: at most 1024 blocks x 27 C doubles
// (648 KiB for C = 3) for ANY grid, since the partial kernel caps its block count.
// Each GPU holds its own copy of the array until the process ends; host code never
// passes the array's name to a kernel (in host code it names a host-side stand-in) but
// the device address that hipGetSymbolAddress returned, fetched once per device with
// that device current.  Calls that overlap on one GPU share the array and race; nothing
// in it needs clearing (every launched block writes all of its partials before the
// final pass reads them).
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)1024 * 27 * 3 * sizeof(double);
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices] = {};

size_t sm_linear_scratch_bytes() { return kScratchBytes; }

static int scratch_for_current_device(void** out) {
  int dev = -1;
  MLPCG_HIP(hipGetDevice(&dev));
  MLPCG_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "sm_linear_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MLPCG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MLPCG_REQUIRE(p != nullptr, "sm_linear_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int launch_sm_linear_noalloc(int dtype, const void* img, const void* w, const void* b, void* z, void* ybar,
                             int C, int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "sm_linear_noalloc: C must be 1, 2 or 3 (got %d)", C);
  MLPCG_REQUIRE(n1 >= 1 && n2 >= 1 && n3 >= 1, "sm_linear_noalloc: bad grid (%d,%d,%d)", n1, n2, n3);
  const size_t need = mean_patch_workspace(dtype, C, n1, n2, n3);
  MLPCG_REQUIRE(need <= kScratchBytes,
                "sm_linear_noalloc: the reduction needs %zu bytes of scratch, more than the %zu reserved", need,
                kScratchBytes);
  void* ws = nullptr;
  if (scratch_for_current_device(&ws)) return -1;
  return launch_sm_linear(dtype, img, w, b, z, ybar, ws, C, n1, n2, n3, st);
}

// gw = g * ybar, gb = g  (sm_linear_3d_kernel.cu:229-238)
template <typename A>
__global__ void linear_backward_kernel(const A* __restrict__ g, const A* __restrict__ ybar,
                                       A* __restrict__ gw, A* __restrict__ gb, int NV) {
  const A gv = g[0];
  for (int o = threadIdx.x; o < NV; o += blockDim.x) gw[o] = gv * ybar[o];
  if (threadIdx.x == 0) gb[0] = gv;
}

int launch_linear_backward(int dtype, const void* g, const void* ybar, void* gw, void* gb, int C,
                           hipStream_t st) {
  if (dtype == kF64)
    linear_backward_kernel<double><<<1, 128, 0, st>>>((const double*)g, (const double*)ybar,
                                                      (double*)gw, (double*)gb, 27 * C);
  else
    linear_backward_kernel<float><<<1, 128, 0, st>>>((const float*)g, (const float*)ybar, (float*)gw,
                                                     (float*)gb, 27 * C);
  return check_launch("sm_linear_backward");
}

// ---------------------------------------------------------------------------
// Weight gradient.
// ---------------------------------------------------------------------------
template <int C> struct WGShape {
  static constexpr int NQ = 27 * C + 1;          // patch columns + bias column
  static constexpr int RP = 3, RQ = 4;           // register tile
  static constexpr int NPG = 9;                  // 27 / RP
  static constexpr int NQG = (NQ + RQ - 1) / RQ;
  static constexpr int NQP = NQG * RQ;           // padded patch rows in LDS
  static constexpr int NT = NPG * NQG;           // threads per voxel subset
  static constexpr int NSUB = kThreads / NT;     // voxel subsets
  static constexpr int NO = 27 * NQ;             // outputs
};

template <typename TX, typename TI, int C, int TV>
__global__ __launch_bounds__(kThreads) void weight_grad_partial_kernel(
    const TX* __restrict__ a, const TX* __restrict__ cv, int B, const TI* __restrict__ img, int n1,
    int n2, int n3, typename AccOf<TX>::type* __restrict__ partials) {
  typedef typename AccOf<TX>::type A;
  typedef WGShape<C> S;
  constexpr int LD = TV + 1;
  __shared__ A Gs[27 * LD];
  __shared__ A Ps[S::NQP * LD];
  static_assert(S::NSUB >= 1, "tile too large");
  static_assert(27 * LD >= kThreads * S::RP * S::RQ, "reduction buffer too small");
  const int tid = threadIdx.x;
  const int tile_id = tid % S::NT, sub = tid / S::NT;
  const int pg = tile_id / S::NQG, qg = tile_id % S::NQG;
  const long V = (long)n1 * n2 * n3;
  const long s1 = (long)n2 * n3;
  A acc[S::RP][S::RQ];
#pragma unroll
  for (int r = 0; r < S::RP; ++r)
#pragma unroll
    for (int q = 0; q < S::RQ; ++q) acc[r][q] = A(0);
  for (int r = S::NQ; r < S::NQP; ++r)
    for (int u = tid; u < TV; u += kThreads) Ps[r * LD + u] = A(0);

  const long ntiles = (V + TV - 1) / TV;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    if (tid < TV) {
      const long v = tile * TV + tid;
      A G[27];
      A P[27 * C];
#pragma unroll
      for (int p = 0; p < 27; ++p) G[p] = A(0);
#pragma unroll
      for (int q = 0; q < 27 * C; ++q) P[q] = A(0);
      if (v < V) {
        const int k = (int)(v % n3);
        const int j = (int)((v / n3) % n2);
        const int i = (int)(v / s1);
        bool in[27];
        long off[27];
#pragma unroll
        for (int p = 0; p < 27; ++p) {
          const int d1 = p / 9 - 1, d2 = (p / 3) % 3 - 1, d3 = p % 3 - 1;
          in[p] = (unsigned)(i + d1) < (unsigned)n1 && (unsigned)(j + d2) < (unsigned)n2 &&
                  (unsigned)(k + d3) < (unsigned)n3;
          off[p] = d1 * s1 + (long)d2 * n3 + d3;
        }
        for (int b = 0; b < B; ++b) {
          const A av = (A)a[(long)b * V + v];
          const TX* cb = cv + (long)b * V + v;
#pragma unroll
          for (int p = 0; p < 27; ++p)
            if (in[p]) G[p] += av * (A)cb[off[p]];
        }
#pragma unroll
        for (int s = 0; s < C; ++s)
#pragma unroll
          for (int p = 0; p < 27; ++p)
            if (in[p]) P[s * 27 + p] = (A)img[s * V + v + off[p]];
      }
#pragma unroll
      for (int p = 0; p < 27; ++p) Gs[p * LD + tid] = G[p];
#pragma unroll
      for (int q = 0; q < 27 * C; ++q) Ps[q * LD + tid] = P[q];
      Ps[(27 * C) * LD + tid] = (v < V) ? A(1) : A(0);
    }
    __syncthreads();
    if (sub < S::NSUB) {
      for (int u = sub; u < TV; u += S::NSUB) {
        A g[S::RP], pv[S::RQ];
#pragma unroll
        for (int r = 0; r < S::RP; ++r) g[r] = Gs[(pg * S::RP + r) * LD + u];
#pragma unroll
        for (int q = 0; q < S::RQ; ++q) pv[q] = Ps[(qg * S::RQ + q) * LD + u];
#pragma unroll
        for (int r = 0; r < S::RP; ++r)
#pragma unroll
          for (int q = 0; q < S::RQ; ++q) acc[r][q] += g[r] * pv[q];
      }
    }
    __syncthreads();
  }
  A* red = Gs;
  if (sub < S::NSUB) {
#pragma unroll
    for (int r = 0; r < S::RP; ++r)
#pragma unroll
      for (int q = 0; q < S::RQ; ++q)
        red[((sub * S::NT + tile_id) * S::RP + r) * S::RQ + q] = acc[r][q];
  }
  __syncthreads();
  if (tid < S::NT) {
#pragma unroll
    for (int r = 0; r < S::RP; ++r) {
#pragma unroll
      for (int q = 0; q < S::RQ; ++q) {
        A s = A(0);
        for (int sb = 0; sb < S::NSUB; ++sb) s += red[((sb * S::NT + tile_id) * S::RP + r) * S::RQ + q];
        const int op = pg * S::RP + r, oq = qg * S::RQ + q;
        if (oq < S::NQ) partials[(size_t)blockIdx.x * S::NO + op * S::NQ + oq] = s;
      }
    }
  }
}

// One-channel weight gradient on the matrix cores (FP32 MFMA, so FP32 products and sums
// as in the VALU kernel above): gW[p][q] = sum_v G_p(v) P_q(v) with P_q the 27 image taps
// and the bias column (P_27 = 1), a 32 x 32 (padded) product over the voxels.  A wave
// walks one z-row segment of up to 64 voxels, 4 voxels per v_mfma_f32_16x16x4_f32
// (A[i][k]: lane 16k + i; B[k][j]: lane 16k + j; D[4(l/16)+t][l%16]): lane (k, i) feeds
// G and P for taps i and 16 + i of voxel k.  Its neighbour offsets are constants and the
// validity of its two neighbour rows is fixed per segment, so only the z ends need a
// check.  Waves of a block add their 32 x 32 tiles in LDS in a fixed order; the block's
// 27 x 28 sums go to partials[block] like the VALU kernel's.
typedef float wg_f32x4 __attribute__((ext_vector_type(4)));

template <typename TX, typename TI>
__global__ __launch_bounds__(kThreads) void weight_grad_mfma_kernel(
    const TX* __restrict__ a, const TX* __restrict__ cv, int B, const TI* __restrict__ img, int n1,
    int n2, int n3, float* __restrict__ partials) {
  constexpr int NQ = 28, NO = 27 * NQ;
  __shared__ float red[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int kq = lane >> 4, il = lane & 15;
  const long V = (long)n1 * n2 * n3;
  const long s1 = (long)n2 * n3;
  // voxels per segment: the largest of 64, 32, 16, 8, 4 that divides n3 (n3 % 4 == 0)
  const int seg = n3 % 64 == 0 ? 64 : (n3 % 32 == 0 ? 32 : (n3 % 16 == 0 ? 16 : (n3 % 8 == 0 ? 8 : 4)));
  const int nseg3 = n3 / seg;
  const long nseg = (long)n1 * n2 * nseg3;
  // this lane's two taps (rows of G, columns of P): t = il and 16 + il
  int d1[2], d2[2], d3[2];
  bool gok[2], pimg[2], pone[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = 16 * h + il;
    const int tt = t < 27 ? t : 0;
    d1[h] = tt / 9 - 1;
    d2[h] = (tt / 3) % 3 - 1;
    d3[h] = tt % 3 - 1;
    gok[h] = t < 27;
    pimg[h] = t < 27;
    pone[h] = t == 27;
  }
  wg_f32x4 acc[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 2; ++g) acc[h][g] = wg_f32x4{0.f, 0.f, 0.f, 0.f};
  for (long sg = (long)blockIdx.x * 4 + w; sg < nseg; sg += (long)gridDim.x * 4) {
    const int zs = (int)(sg % nseg3) * seg;
    const long rowid = sg / nseg3;
    const int j = (int)(rowid % n2), i = (int)(rowid / n2);
    const long vrow = rowid * n3;
    bool rok[2];
    long roff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      rok[h] = (unsigned)(i + d1[h]) < (unsigned)n1 && (unsigned)(j + d2[h]) < (unsigned)n2;
      roff[h] = (long)d1[h] * s1 + (long)d2[h] * n3 + d3[h];
    }
#pragma unroll 4
    for (int z4 = 0; z4 < seg; z4 += 4) {
      const int z = zs + z4 + kq;
      const long v = vrow + z;
      float gv[2] = {0.f, 0.f}, pv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const bool ok = rok[h] && (unsigned)(z + d3[h]) < (unsigned)n3;
        const long u = ok ? v + roff[h] : v;
        float sacc = 0.f;
        for (int b = 0; b < B; ++b) sacc += (float)a[(long)b * V + v] * (float)cv[(long)b * V + u];
        gv[h] = (gok[h] && ok) ? sacc : 0.f;
        const float iv = (float)img[u];
        pv[h] = pone[h] ? 1.f : ((pimg[h] && ok) ? iv : 0.f);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 2; ++g) acc[h][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(gv[h], pv[g], acc[h][g], 0, 0, 0);
    }
  }
  // D[16h + 4 kq + t][16g + il] of this wave -> LDS, then a fixed-order sum over the waves
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int t = 0; t < 4; ++t) red[w][(16 * h + 4 * kq + t) * 32 + 16 * g + il] = acc[h][g][t];
  __syncthreads();
  for (int o = threadIdx.x; o < NO; o += kThreads) {
    const int pp = o / NQ, qq = o % NQ;
    const int e = pp * 32 + qq;
    partials[(size_t)blockIdx.x * NO + o] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
  }
}

// B == 1 form of weight_grad_mfma_kernel with the neighbour rows staged in LDS: per
// z-row segment (one per wave; the block's four waves run the same number of rounds, so
// the barriers are uniform) the 9 neighbour rows of x and of the image over
// [z0 - 1, z0 + seg] are loaded with coalesced row reads, zero outside the grid, and the
// matrix-core loop reads G and P from LDS without bounds checks.
template <typename TX, typename TI>
__global__ __launch_bounds__(kThreads) void weight_grad_mfma_lds_kernel(
    const TX* __restrict__ a, const TX* __restrict__ cv, const TI* __restrict__ img, int n1, int n2,
    int n3, float* __restrict__ partials) {
  constexpr int NQ = 28, NO = 27 * NQ, RL = 66;  // staged row length (seg <= 64, + halo)
  __shared__ float xs[4][9 * RL], is[4][9 * RL], ac[4][64];
  __shared__ float red[4][32 * 32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int kq = lane >> 4, il = lane & 15;
  const long s1 = (long)n2 * n3;
  const int seg = n3 % 64 == 0 ? 64 : (n3 % 32 == 0 ? 32 : (n3 % 16 == 0 ? 16 : (n3 % 8 == 0 ? 8 : 4)));
  const int nseg3 = n3 / seg;
  const long nseg = (long)n1 * n2 * nseg3;
  int rr[2], cz[2];  // this lane's taps t = il, 16 + il: staged row and z offset (t / 3, t % 3)
  bool gok[2], pimg[2], pone[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = 16 * h + il;
    const int tt = t < 27 ? t : 0;
    rr[h] = tt / 3;
    cz[h] = tt % 3;
    gok[h] = t < 27;
    pimg[h] = t < 27;
    pone[h] = t == 27;
  }
  wg_f32x4 acc[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 2; ++g) acc[h][g] = wg_f32x4{0.f, 0.f, 0.f, 0.f};
  for (long sg0 = (long)blockIdx.x * 4; sg0 < nseg; sg0 += (long)gridDim.x * 4) {
    const long sg = sg0 + w;
    const bool active = sg < nseg;  // uniform per wave
    if (active) {
      const int zs = (int)(sg % nseg3) * seg;
      const long rowid = sg / nseg3;
      const int j = (int)(rowid % n2), i = (int)(rowid / n2);
      for (int e = lane; e < 9 * (seg + 2); e += 64) {
        const int r = e / (seg + 2), cc = e - r * (seg + 2);
        const int ii = i + r / 3 - 1, jj = j + r % 3 - 1, z = zs - 1 + cc;
        const bool ok = (unsigned)ii < (unsigned)n1 && (unsigned)jj < (unsigned)n2 && (unsigned)z < (unsigned)n3;
        const long u = ok ? (long)ii * s1 + (long)jj * n3 + z : 0;
        xs[w][r * RL + cc] = ok ? (float)cv[u] : 0.f;
        is[w][r * RL + cc] = ok ? (float)img[u] : 0.f;
      }
      for (int e = lane; e < seg; e += 64) ac[w][e] = (float)a[rowid * n3 + zs + e];
    }
    __syncthreads();
    if (active) {
      for (int z4 = 0; z4 < seg; z4 += 4) {
        const int z = z4 + kq;
        const float av = ac[w][z];
        float gv[2], pv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int cidx = rr[h] * RL + z + cz[h];
          gv[h] = gok[h] ? av * xs[w][cidx] : 0.f;
          pv[h] = pone[h] ? 1.f : (pimg[h] ? is[w][cidx] : 0.f);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int g = 0; g < 2; ++g)
            acc[h][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(gv[h], pv[g], acc[h][g], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int t = 0; t < 4; ++t) red[w][(16 * h + 4 * kq + t) * 32 + 16 * g + il] = acc[h][g][t];
  __syncthreads();
  for (int o = threadIdx.x; o < NO; o += kThreads) {
    const int pp = o / NQ, qq = o % NQ;
    const int e = pp * 32 + qq;
    partials[(size_t)blockIdx.x * NO + o] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
  }
}

// BF16 weight gradient on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16), B == 1,
// n3 % 64 == 0, any C <= 3.  G_p(v) = a(v) c(v + d_p) of two BF16 values has at most 16
// significant bits, so it is exact in FP32 and splits exactly into hi + lo BF16 parts
// (hi: the top 8 bits by truncation, lo = G - hi: the remaining <= 8 bits).  With
// D[p][q] = sum_v hi_p(v) P_q(v) + lo_p(v) P_q(v) every product is exact and the sums are
// FP32, as in the FP32-operand kernels above; the matrix cores do 32 voxels per product
// instead of 4.
//
// A wave walks a column of z-row segments (plane i, rows j0 .. j0 + J - 1, 64 z cells,
// two K chunks of 32 voxels) and keeps the three neighbour rows along j of each plane
// i - 1, i, i + 1 in a ring of LDS slots, so each step stages one new row per plane
// (loaded into registers during the previous step's products).  x rows are stored as FP32
// (one copy; a lane's 8 taps along z are read at 4-byte alignment), image rows as three
// BF16 copies shifted by one cell so that every B operand is one aligned 16-byte read.
// Operand layout: A[p][k] lane 16 kb + m holds row m, voxels kv .. kv + 7 of the chunk with
// kv = 16 (kb & 1) + 8 (kb >> 1); B[k][q] lane 16 kb + m holds column m of the same voxels;
// D[4 kb + i][m].  Rows p = 27..31 and columns q > 27C read constant zero rows, column 27C
// (bias) a row of ones.  LDS banks (32 x 4 bytes, 32 lanes per cycle): the two lane groups
// of a half-wave read x words 16 apart, and the x rows of a ring (slot pitch kXP, plane
// pitch kXPP words) start at offsets that keep the 16 taps of a half-wave on distinct
// banks mod 16 for most ring positions (a CPU model of the x-operand reads: 107 LDS cycles
// per step, 64 without conflicts, against 213 with the lane groups' voxels side by side and
// 72-word rows).
namespace wgb {
constexpr int kSeg = 64;                       // z cells per row segment
constexpr int kXS = 72;                        // FP32 zero row length; x row: cell z0 - 1 + e at element 3 + e
constexpr int kXP = 70, kXPP = 212;            // x ring: slot and plane pitch (words)
constexpr int kXW = 2 * kXPP + 2 * kXP + 72;   // x ring words per wave
constexpr int kIS = 72;                        // BF16 image row copy (64 cells used)
constexpr int kXBytes = 4 * kXW * 4;           // x rings of the 4 waves
constexpr int kGBytes = 4 * kSeg * 4;          // a rows (FP32) of the 4 waves
constexpr int kZBytes = kXS * 4;               // FP32 zero row
constexpr int kCBytes = 2 * kIS * 2;           // BF16 zero row, ones row
__host__ __device__ constexpr int img_bytes(int C) { return 4 * C * 27 * kIS * 2; }
__host__ __device__ constexpr int lds_bytes(int C) { return kXBytes + kGBytes + kZBytes + kCBytes + img_bytes(C); }
}  // namespace wgb

typedef __bf16 wg_bf16x8 __attribute__((ext_vector_type(8), aligned(16)));
typedef uint32_t wg_u32x4 __attribute__((ext_vector_type(4)));

// hi / lo BF16 parts of an exact FP32 product, packed in pairs (element 0 in the low half)
__device__ __forceinline__ void wg_split2(float p0, float p1, uint32_t& hi, uint32_t& lo) {
  const uint32_t u0 = __float_as_uint(p0), u1 = __float_as_uint(p1);
  const float l0 = p0 - __uint_as_float(u0 & 0xffff0000u);
  const float l1 = p1 - __uint_as_float(u1 & 0xffff0000u);
  hi = __builtin_amdgcn_perm(u1, u0, 0x07060302u);
  lo = __builtin_amdgcn_perm(__float_as_uint(l1), __float_as_uint(l0), 0x07060302u);
}

template <int C>
__global__ __launch_bounds__(kThreads) void weight_grad_bf16_kernel(
    const bf16* __restrict__ a, const bf16* __restrict__ cv, const bf16* __restrict__ img, int n1,
    int n2, int n3, int J, float* __restrict__ partials) {
  using namespace wgb;
  constexpr int NQ = 27 * C + 1, NT = (NQ + 15) / 16, NO = 27 * NQ;
  constexpr int NTASK = 24 + 24 * C + 8, NR = (NTASK + 63) / 64;  // staging tasks per step
  static_assert(4 * 27 * NQ * 4 <= img_bytes(C), "reduction buffer");
  __shared__ __attribute__((aligned(16))) char lds[lds_bytes(C)];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m = lane & 15, kb = lane >> 4;
  // byte offsets in lds
  const int oX = w * kXW * 4;
  const int kv = ((kb & 1) << 4) | ((kb >> 1) << 3);  // first voxel of the lane's 8 in a chunk
  const int oG = kXBytes + w * kSeg * 4;
  const int oZ = kXBytes + kGBytes;
  const int oC = oZ + kZBytes;
  const int oI = oC + kCBytes + w * C * 27 * kIS * 2;
  for (int e = threadIdx.x; e < kXS; e += kThreads) *(float*)(lds + oZ + 4 * e) = 0.f;
  for (int e = threadIdx.x; e < 2 * kIS; e += kThreads)
    *(uint16_t*)(lds + oC + 2 * e) = e < kIS ? (uint16_t)0 : (uint16_t)0x3f80;  // BF16 0, 1
  __syncthreads();

  // this lane's operand rows: A taps p = 16 h + m, B columns q = 16 gt + m.  Offset at
  // step j = base + slot * stride with slot = (j + 2 + e2) mod 3 the ring slot of row
  // j + e2 - 1 (e2 = d2 + 1); constant rows have stride 0.
  int abase[2], ae2[2], astr[2], bbase[NT], be2[NT], bstr[NT];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = 16 * h + m;
    if (p < 27) {
      abase[h] = oX + ((p / 9) * kXPP + 3 + p % 3 + kv) * 4;
      ae2[h] = (p / 3) % 3;
      astr[h] = kXP * 4;
    } else {
      abase[h] = oZ + kv * 4;
      ae2[h] = 0;
      astr[h] = 0;
    }
  }
#pragma unroll
  for (int gt = 0; gt < NT; ++gt) {
    const int q = 16 * gt + m;
    if (q < 27 * C) {
      const int c = q / 27, s = q % 27;
      bbase[gt] = oI + (((c * 9 + (s / 9) * 3) * 3 + s % 3) * kIS + kv) * 2;
      be2[gt] = (s / 3) % 3;
      bstr[gt] = 3 * kIS * 2;
    } else {
      bbase[gt] = oC + ((q == 27 * C ? kIS : 0) + kv) * 2;
      be2[gt] = 0;
      bstr[gt] = 0;
    }
  }

  const long V = (long)n1 * n2 * n3;
  const int nseg3 = n3 / kSeg, njb = (n2 + J - 1) / J;
  const long ncol = (long)n1 * nseg3 * njb;
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 acc[2][NT];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int gt = 0; gt < NT; ++gt) acc[h][gt] = f4{0.f, 0.f, 0.f, 0.f};

  // staging task e of a step: e < 24: x row (plane i + e / 8 - 1, row jr), chunk e % 8;
  // e < 24 + 24C: image channel, same rows; else the a row (i, ja), chunk e % 8.  A task
  // loads the chunk's 8 cells plus the cell pairs before and after it (zero outside the grid).
  struct Regs {
    uint32_t d[NR][6];
  };
  auto task_src = [&](int e, int i, int jr, int ja, int z0, const bf16*& src, bool& ok, int& t) {
    int ii, jj;
    if (e < 24) {
      src = cv;
      ii = i + e / 8 - 1;
      jj = jr;
    } else if (e < 24 + 24 * C) {
      const int f = e - 24;
      src = img + (long)(f / 24) * V;
      ii = i + (f % 24) / 8 - 1;
      jj = jr;
    } else {
      src = a;
      ii = i;
      jj = ja;
    }
    t = e & 7;
    ok = e < NTASK && (unsigned)ii < (unsigned)n1 && (unsigned)jj < (unsigned)n2;
    src += ((long)(ok ? ii : 0) * n2 + (ok ? jj : 0)) * n3 + z0 + 8 * t;
  };
  auto load_step = [&](Regs& r, int i, int jr, int ja, int z0) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int e = lane + 64 * k;
      const bf16* src;
      bool ok;
      int t;
      task_src(e, i, jr, ja, z0, src, ok, t);
      const uint32_t* s32 = (const uint32_t*)src;
      const wg_u32x4 v = ok ? *(const wg_u32x4*)s32 : wg_u32x4{0u, 0u, 0u, 0u};
      const bool okp = ok && z0 + 8 * t > 0, okn = ok && z0 + 8 * t + 8 < n3;
      r.d[k][0] = okp ? s32[-1] : 0u;
      r.d[k][1] = v[0];
      r.d[k][2] = v[1];
      r.d[k][3] = v[2];
      r.d[k][4] = v[3];
      r.d[k][5] = okn ? s32[4] : 0u;
    }
  };
  // slot of row jr in the ring: (jr + 3) mod 3 (jr >= -1)
  auto write_step = [&](const Regs& r, int jr) {
    const int sl = (jr + 3) % 3;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int e = lane + 64 * k;
      const int t = e & 7;
      const uint32_t* d = r.d[k];
      if (e < 24) {
        // rows start at even words (8-byte stores)
        char* dst = lds + oX + ((e / 8) * kXPP + sl * kXP + 4 + 8 * t) * 4;
        typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int u = 0; u < 4; ++u)
          *(f2*)(dst + 8 * u) = f2{__uint_as_float(d[1 + u] << 16), __uint_as_float(d[1 + u] & 0xffff0000u)};
        if (t == 0) *(float*)(dst - 4) = __uint_as_float(d[0] & 0xffff0000u);
        if (t == 7) *(float*)(dst + 32) = __uint_as_float(d[5] << 16);
      } else if (e < 24 + 24 * C) {
        const int f = e - 24;
        char* dst = lds + oI + ((((f / 24) * 9 + ((f % 24) / 8) * 3 + sl) * 3) * kIS + 8 * t) * 2;
        *(wg_u32x4*)dst = wg_u32x4{__builtin_amdgcn_alignbyte(d[1], d[0], 2), __builtin_amdgcn_alignbyte(d[2], d[1], 2),
                                   __builtin_amdgcn_alignbyte(d[3], d[2], 2), __builtin_amdgcn_alignbyte(d[4], d[3], 2)};
        *(wg_u32x4*)(dst + kIS * 2) = wg_u32x4{d[1], d[2], d[3], d[4]};
        *(wg_u32x4*)(dst + 2 * kIS * 2) =
            wg_u32x4{__builtin_amdgcn_alignbyte(d[2], d[1], 2), __builtin_amdgcn_alignbyte(d[3], d[2], 2),
                     __builtin_amdgcn_alignbyte(d[4], d[3], 2), __builtin_amdgcn_alignbyte(d[5], d[4], 2)};
      } else if (e < NTASK) {
        char* dst = lds + oG + 8 * t * 4;
        *(f4*)dst = f4{__uint_as_float(d[1] << 16), __uint_as_float(d[1] & 0xffff0000u),
                       __uint_as_float(d[2] << 16), __uint_as_float(d[2] & 0xffff0000u)};
        *(f4*)(dst + 16) = f4{__uint_as_float(d[3] << 16), __uint_as_float(d[3] & 0xffff0000u),
                              __uint_as_float(d[4] << 16), __uint_as_float(d[4] & 0xffff0000u)};
      }
    }
  };
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  for (long col = (long)blockIdx.x * 4 + w; col < ncol; col += (long)gridDim.x * 4) {
    // column order: plane fastest, so the waves of a block share neighbour planes
    const int i = (int)(col % n1);
    const long rest = col / n1;
    const int jb = (int)(rest % njb);
    const int z0 = (int)(rest / njb) * kSeg;
    const int j0 = jb * J, j1 = min(j0 + J, n2);
    Regs r;
    // prologue: rows j0 - 1 and j0 (the a-row tasks of these two are not used)
    load_step(r, i, j0 - 1, j0, z0);
    write_step(r, j0 - 1);
    load_step(r, i, j0, j0, z0);
    write_step(r, j0);
    load_step(r, i, j0 + 1, j0, z0);
    for (int j = j0; j < j1; ++j) {
      write_step(r, j + 1);  // row j + 1 and the a row of j
      wave_sync();
      if (j + 1 < j1) load_step(r, i, j + 2, j + 1, z0);
      const int sj = (j + 2) % 3;
      int ao[2], bo[NT];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int s = sj + ae2[h];
        s = s >= 3 ? s - 3 : s;
        ao[h] = abase[h] + s * astr[h];
      }
#pragma unroll
      for (int gt = 0; gt < NT; ++gt) {
        int s = sj + be2[gt];
        s = s >= 3 ? s - 3 : s;
        bo[gt] = bbase[gt] + s * bstr[gt];
      }
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const f4 g0 = *(const f4*)(lds + oG + (32 * cc + kv) * 4);
        const f4 g1 = *(const f4*)(lds + oG + (32 * cc + kv + 4) * 4);
        const float gv[8] = {g0[0], g0[1], g0[2], g0[3], g1[0], g1[1], g1[2], g1[3]};
        wg_bf16x8 ahi[2], alo[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float* xp = (const float*)(lds + ao[h] + 128 * cc);
          uint32_t hi[4], lo[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) wg_split2(gv[2 * k] * xp[2 * k], gv[2 * k + 1] * xp[2 * k + 1], hi[k], lo[k]);
          ahi[h] = __builtin_bit_cast(wg_bf16x8, wg_u32x4{hi[0], hi[1], hi[2], hi[3]});
          alo[h] = __builtin_bit_cast(wg_bf16x8, wg_u32x4{lo[0], lo[1], lo[2], lo[3]});
        }
#pragma unroll
        for (int gt = 0; gt < NT; ++gt) {
          const wg_bf16x8 bv = *(const wg_bf16x8*)__builtin_assume_aligned(lds + bo[gt] + 64 * cc, 16);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            acc[h][gt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[h], bv, acc[h][gt], 0, 0, 0);
            acc[h][gt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[h], bv, acc[h][gt], 0, 0, 0);
          }
        }
      }
      wave_sync();
    }
  }
  // D[16 h + 4 kb + t][16 gt + m] of this wave -> LDS (over the image slots), then a
  // fixed-order sum over the four waves
  __syncthreads();
  float* red = (float*)(lds + kXBytes + kGBytes + kZBytes + kCBytes);
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int gt = 0; gt < NT; ++gt)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int p = 16 * h + 4 * kb + t, q = 16 * gt + m;
        if (p < 27 && q < NQ) red[(w * 27 + p) * NQ + q] = acc[h][gt][t];
      }
  __syncthreads();
  for (int o = threadIdx.x; o < NO; o += kThreads)
    partials[(size_t)blockIdx.x * NO + o] = (red[o] + red[NO + o]) + (red[2 * NO + o] + red[3 * NO + o]);
}

template <typename A>
__global__ __launch_bounds__(kThreads) void weight_grad_final_kernel(const A* __restrict__ partials,
                                                                     int nblk, int C,
                                                                     A* __restrict__ gW,
                                                                     A* __restrict__ gb) {
  // one block per output; the block sums the nblk partials (strided, then block-reduced)
  __shared__ double sh[4];
  const int NQ = 27 * C + 1, NO = 27 * NQ;
  const int o = blockIdx.x;
  if (o >= NO) return;
  double s = 0.0;
  for (int blk = threadIdx.x; blk < nblk; blk += kThreads) s += (double)partials[(size_t)blk * NO + o];
  s = block_sum(s, sh);
  if (threadIdx.x != 0) return;
  const int p = o / NQ, q = o % NQ;
  if (q < 27 * C) gW[p * 27 * C + q] = (A)s;
  else gb[p] = (A)s;
}

static int wg_tv(int dtype, int C) { return (dtype == kF64 && C >= 2) ? 128 : 256; }

static int wg_blocks(int dtype, int C, int n1, int n2, int n3) {
  const long V = (long)n1 * n2 * n3;
  const int tv = wg_tv(dtype, C);
  long nb = (V + tv - 1) / tv;
  if (nb > 1024) nb = 1024;
  return (int)(nb < 1 ? 1 : nb);
}

size_t weight_grad_workspace(int dtype, int C, int /*B*/, int n1, int n2, int n3) {
  const size_t acc = dtype == kF64 ? sizeof(double) : sizeof(float);
  return (size_t)wg_blocks(dtype, C, n1, n2, n3) * 27 * (27 * C + 1) * acc;
}

template <typename TX, typename TI>
static void launch_wg_t(int dtype, const void* a, const void* c, const void* img, void* part, int nblk,
                        int C, int B, int n1, int n2, int n3, hipStream_t st) {
  typedef typename AccOf<TX>::type A;
  const int tv = wg_tv(dtype, C);
#define MLPCG_WG(CC, TVV)                                                                  \
  if (C == CC && tv == TVV) {                                                              \
    weight_grad_partial_kernel<TX, TI, CC, TVV><<<nblk, kThreads, 0, st>>>(                \
        (const TX*)a, (const TX*)c, B, (const TI*)img, n1, n2, n3, (A*)part);              \
    return;                                                                                \
  }
  if constexpr (sizeof(A) == 8) {
    MLPCG_WG(1, 256) MLPCG_WG(2, 128) MLPCG_WG(3, 128)
  } else {
    MLPCG_WG(1, 256) MLPCG_WG(2, 256) MLPCG_WG(3, 256)
  }
#undef MLPCG_WG
}

int launch_weight_grad(int dtype, const void* a, const void* c, const void* img, void* gW, void* gb,
                       void* workspace, int C, int B, int n1, int n2, int n3, hipStream_t st) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "weight_grad: C must be 1, 2 or 3 (got %d)", C);
  const int nblk = wg_blocks(dtype, C, n1, n2, n3);
  if (dtype == kBF16 && B == 1 && n3 % wgb::kSeg == 0) {
    // BF16 matrix-core kernel (same partials layout and block count): columns of J row
    // segments, about one column per wave
    const long segs = (long)n1 * n2 * (n3 / wgb::kSeg);
    const long per_wave = (segs + 4L * nblk - 1) / (4L * nblk);
    const int J = (int)(per_wave < 4 ? 4 : (per_wave > 64 ? 64 : per_wave));
    const bf16 *pa = (const bf16*)a, *pc = (const bf16*)c, *pi = (const bf16*)img;
    if (C == 1) weight_grad_bf16_kernel<1><<<nblk, kThreads, 0, st>>>(pa, pc, pi, n1, n2, n3, J, (float*)workspace);
    else if (C == 2) weight_grad_bf16_kernel<2><<<nblk, kThreads, 0, st>>>(pa, pc, pi, n1, n2, n3, J, (float*)workspace);
    else weight_grad_bf16_kernel<3><<<nblk, kThreads, 0, st>>>(pa, pc, pi, n1, n2, n3, J, (float*)workspace);
    weight_grad_final_kernel<float><<<27 * (27 * C + 1), kThreads, 0, st>>>((const float*)workspace, nblk, C,
                                                                           (float*)gW, (float*)gb);
    return check_launch("weight_grad");
  }
  if (C == 1 && dtype != kF64 && n3 % 4 == 0) {
    // matrix-core kernel (same partials layout and block count)
    switch (dtype) {
      case kF32:
        if (B == 1)
          weight_grad_mfma_lds_kernel<float, float><<<nblk, kThreads, 0, st>>>(
              (const float*)a, (const float*)c, (const float*)img, n1, n2, n3, (float*)workspace);
        else
          weight_grad_mfma_kernel<float, float><<<nblk, kThreads, 0, st>>>(
              (const float*)a, (const float*)c, B, (const float*)img, n1, n2, n3, (float*)workspace);
        break;
      case kF16:
        weight_grad_mfma_kernel<f16, f16><<<nblk, kThreads, 0, st>>>(
            (const f16*)a, (const f16*)c, B, (const f16*)img, n1, n2, n3, (float*)workspace);
        break;
      default:
        if (B == 1)
          weight_grad_mfma_lds_kernel<bf16, bf16><<<nblk, kThreads, 0, st>>>(
              (const bf16*)a, (const bf16*)c, (const bf16*)img, n1, n2, n3, (float*)workspace);
        else
          weight_grad_mfma_kernel<bf16, bf16><<<nblk, kThreads, 0, st>>>(
              (const bf16*)a, (const bf16*)c, B, (const bf16*)img, n1, n2, n3, (float*)workspace);
        break;
    }
    weight_grad_final_kernel<float><<<27 * 28, kThreads, 0, st>>>((const float*)workspace, nblk, C,
                                                                  (float*)gW, (float*)gb);
    return check_launch("weight_grad");
  }
  switch (dtype) {
    case kF64: launch_wg_t<double, double>(dtype, a, c, img, workspace, nblk, C, B, n1, n2, n3, st); break;
    case kF32: launch_wg_t<float, float>(dtype, a, c, img, workspace, nblk, C, B, n1, n2, n3, st); break;
    case kF16: launch_wg_t<f16, f16>(dtype, a, c, img, workspace, nblk, C, B, n1, n2, n3, st); break;
    case kBF16: launch_wg_t<bf16, bf16>(dtype, a, c, img, workspace, nblk, C, B, n1, n2, n3, st); break;
    default: MLPCG_REQUIRE(false, "weight_grad: bad dtype %d", dtype);
  }
  const int NO = 27 * (27 * C + 1);
  const int fb = NO;
  if (dtype == kF64)
    weight_grad_final_kernel<double><<<fb, kThreads, 0, st>>>((const double*)workspace, nblk, C,
                                                              (double*)gW, (double*)gb);
  else
    weight_grad_final_kernel<float><<<fb, kThreads, 0, st>>>((const float*)workspace, nblk, C,
                                                             (float*)gW, (float*)gb);
  return check_launch("weight_grad");
}

}  // namespace mlpcg
