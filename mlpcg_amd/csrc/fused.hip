// Fused SM block apply on the matrix cores (16-bit plans).
//
//   y(v) = sum_p K_p(v) x(v + d_p),   K_p(v) = b_p + sum_{c,s} W[p,c,s] img[c, v + d_s]
//
// (sm_block_3d_kernel.cu:46-70).  Instead of streaming a stored K (54 of the 60 bytes per
// voxel of the BF16 apply), K is recomputed per voxel as a small GEMM on the MFMA units:
// for 16 voxels, D[p][v] = sum_k A[p][k] B[k][v] with A = W (27 taps x 27C inputs, padded
// to 32 x 32*NKC) and B = the image patches of the 16 voxels.  D never leaves registers:
// each lane multiplies its taps by x(v + d_p) and a butterfly over the four lane groups
// finishes the sum.  HBM traffic is x, img, y (+ epilogue operands), about 8 B/voxel.
//
// Operand layout of v_mfma_f32_16x16x32_{bf16,f16} (lane l, element j / i):
//   A[l % 16][8 (l / 16) + j],  B[8 (l / 16) + j][l % 16],  D[4 (l / 16) + i][l % 16].
// Precision: W and img are rounded to the 16-bit type, products accumulate in FP32; for
// BF16 the weights are split into hi + lo parts (two MFMAs per operand pair), which keeps
// K as accurate as rounding a FP32-computed K to BF16.  The bias enters as the FP32
// initial accumulator.
//
// Tile: a block of 4 waves covers 8 planes x 8 rows x 64 z-voxels; x and img of the tile
// plus a one-cell halo are staged in LDS; a wave processes 64 consecutive z-voxels (one
// per lane at the end) of one (plane, row) per step: 4 MFMA groups of 16 voxels.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

namespace {

constexpr int kFz = 64, kFj = 8, kFi = 8;
constexpr int kPz = kFz + 2, kPj = kFj + 2, kPi = kFi + 2;
constexpr int kTile = kPi * kPj * kPz;  // halo-tile cells

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename TW> struct Mfma;
template <> struct Mfma<bf16> {
  typedef bf16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<f16> {
  typedef f16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

__host__ __device__ constexpr int nkc(int C) { return (27 * C + 31) / 32; }

// LDS offset of the neighbour d_q (q = 9(d1+1) + 3(d2+1) + (d3+1)) in the halo tile.
__device__ __forceinline__ int tap_off(int q) {
  return ((q / 9 - 1) * kPj + ((q / 3) % 3 - 1)) * kPz + (q % 3 - 1);
}

}  // namespace

size_t fused_weights_bytes(int C) {
  return (size_t)2 * 2 * nkc(C) * 64 * 8 * 2 + (size_t)2 * 64 * 4 * sizeof(float);
}

// apack[part][h][kc][lane][8] (TW), then binit[h][lane][4] (float).
template <typename TW>
__global__ __launch_bounds__(kThreads) void fused_pack_kernel(const float* __restrict__ W,
                                                              const float* __restrict__ b, int C,
                                                              int split, TW* __restrict__ apack,
                                                              float* __restrict__ binit) {
  const int NK = nkc(C);
  const int total = 2 * 2 * NK * 64 * 8;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int j = e % 8, lane = (e / 8) % 64, kc = (e / 512) % NK, h = (e / (512 * NK)) % 2;
    const int part = e / (1024 * NK);
    const int m = 16 * h + lane % 16, k = 32 * kc + 8 * (lane / 16) + j;
    const float w = (m < 27 && k < 27 * C) ? W[((long)m * C + k / 27) * 27 + k % 27] : 0.f;
    const TW hi = (TW)w;
    apack[e] = part == 0 ? hi : (split ? (TW)(w - (float)hi) : (TW)0.f);
  }
  for (int e = threadIdx.x; e < 2 * 64 * 4; e += blockDim.x) {
    const int i = e % 4, lane = (e / 4) % 64, h = e / 256;
    const int m = 16 * h + 4 * (lane / 16) + i;
    binit[e] = m < 27 ? b[m] : 0.f;
  }
}

template <typename TW, typename TX, typename TY, typename TP, int C, bool SPLIT>
__global__ __launch_bounds__(kThreads) void sm_fused_kernel(FusedArgs a) {
  typedef typename Mfma<TW>::v8 v8;
  constexpr int NK = nkc(C);
  constexpr int NP = SPLIT ? 2 : 1;
  __shared__ float xs[kTile];
  __shared__ TW is[C * kTile];
  if (solve_done(a.done)) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, nn = lane & 15;
  const int tz = a.n3 / kFz, tj = a.n2 / kFj;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int z0 = (int)(t % tz) * kFz;
  const int j0 = (int)((t / tz) % tj) * kFj;
  const int i0 = (int)(t / tz / tj) * kFi;
  const long V = (long)a.n1 * a.n2 * a.n3;

  // per-lane constant operands (global loads issued before the tile staging)
  const TW* apack = (const TW*)a.apack;
  v8 A[NP][2][NK];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int kc = 0; kc < NK; ++kc)
        A[p][h][kc] = *(const v8*)(apack + ((((long)p * 2 + h) * NK + kc) * 64 + lane) * 8);
  const float* binit = a.binit;
  f32x4 B0 = *(const f32x4*)(binit + lane * 4);
  f32x4 B1 = *(const f32x4*)(binit + (64 + lane) * 4);
  const float scale = a.in_scale ? (float)(*a.in_scale) : 1.f;
  const float c0 = a.prev ? *(const float*)a.c0 : 0.f;
  const float c1 = a.prev ? *(const float*)a.c1 : 0.f;

  // stage x (as FP32) and img of the halo tile; zero outside the grid
  const TX* __restrict__ x = (const TX*)a.x;
  const TW* __restrict__ img = (const TW*)a.img;
  for (int r = w; r < kPi * kPj; r += 4) {
    const int pi = r / kPj, pj = r - pi * kPj;
    const int gi = i0 + pi - 1, gj = j0 + pj - 1;
    const bool rok = gi >= 0 && gi < a.n1 && gj >= 0 && gj < a.n2;
    const long rb = ((long)gi * a.n2 + gj) * a.n3;
    for (int e = lane; e < kPz; e += 64) {
      const int gz = z0 - 1 + e;
      const bool ok = rok && gz >= 0 && gz < a.n3;
      xs[r * kPz + e] = ok ? (float)x[rb + gz] : 0.f;
#pragma unroll
      for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + e] = ok ? img[c * V + rb + gz] : (TW)0.f;
    }
  }

  // per-lane LDS offsets: image inputs k of this lane's B rows, taps of its D rows
  int boff[NK][8], xoff[2][4];
#pragma unroll
  for (int kc = 0; kc < NK; ++kc)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 32 * kc + 8 * g + j;
      boff[kc][j] = k < 27 * C ? (k / 27) * kTile + tap_off(k % 27) : 0;  // padding reads any cell
    }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = 16 * h + 4 * g + i;
      xoff[h][i] = m < 27 ? tap_off(m) : 0;  // rows m >= 27 of D are exactly zero
    }
  __syncthreads();

  TY* __restrict__ y = (TY*)a.y;
  const TP* __restrict__ prev = (const TP*)a.prev;
  for (int q = 0; q < kFi * 2; ++q) {
    const int ii = q >> 1, jj = w + 4 * (q & 1);
    const int cb = ((ii + 1) * kPj + (jj + 1)) * kPz + 1;
    const long v = ((long)(i0 + ii) * a.n2 + (j0 + jj)) * a.n3 + z0 + lane;
    // epilogue operands first: their latency overlaps the MFMA work
    const uint8_t ka = a.mask_a ? a.mask_a[v] : (uint8_t)1;
    const uint8_t kb = a.mask_b ? a.mask_b[v] : (uint8_t)1;
    const float pv = prev ? (float)prev[v] : 0.f;
    float ymine = 0.f;
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      const int ci = cb + 16 * tt + nn;
      f32x4 d0 = B0, d1 = B1;
#pragma unroll
      for (int kc = 0; kc < NK; ++kc) {
        v8 bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = is[ci + boff[kc][j]];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          d0 = Mfma<TW>::run(A[p][0][kc], bv, d0);
          d1 = Mfma<TW>::run(A[p][1][kc], bv, d1);
        }
      }
      float yp = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) yp += d0[i] * xs[ci + xoff[0][i]] + d1[i] * xs[ci + xoff[1][i]];
      yp += __shfl_xor(yp, 16, 64);
      yp += __shfl_xor(yp, 32, 64);
      ymine = (g == tt) ? yp : ymine;
    }
    float val = ymine * scale;
    if (ka == 0) val = 0.f;
    if (prev) val = c0 * pv + c1 * val;
    if (kb == 0) val = 0.f;
    y[v] = (TY)val;
  }
}

bool fused_eligible(int C, int B, int n1, int n2, int n3) {
  return (C == 1 || C == 2) && B == 1 && n1 % kFi == 0 && n2 % kFj == 0 && n3 % kFz == 0;
}

int launch_fused_pack(int tw, const void* W, const void* b, int C, void* wbuf, hipStream_t st) {
  MLPCG_REQUIRE(C == 1 || C == 2, "fused_pack: C must be 1 or 2");
  float* binit = (float*)((char*)wbuf + (size_t)2 * 2 * nkc(C) * 64 * 8 * 2);
  if (tw == kBF16)
    fused_pack_kernel<bf16><<<1, kThreads, 0, st>>>((const float*)W, (const float*)b, C, 1, (bf16*)wbuf, binit);
  else if (tw == kF16)
    fused_pack_kernel<f16><<<1, kThreads, 0, st>>>((const float*)W, (const float*)b, C, 0, (f16*)wbuf, binit);
  else
    MLPCG_REQUIRE(false, "fused_pack: 16-bit type expected (got %d)", tw);
  return check_launch("fused_pack");
}

template <typename TW, typename TX, typename TY, int C>
static void launch_fused_t(const FusedArgs& a, hipStream_t st) {
  constexpr bool split = std::is_same<TW, bf16>::value;
  const unsigned nb = (unsigned)((a.n1 / kFi) * (a.n2 / kFj) * (a.n3 / kFz));
  sm_fused_kernel<TW, TX, TY, TW, C, split><<<nb, kThreads, 0, st>>>(a);
}

int launch_sm_fused(int tw, int tx, int ty, int C, const FusedArgs& args, hipStream_t st) {
  MLPCG_REQUIRE(fused_eligible(C, 1, args.n1, args.n2, args.n3),
                "sm_fused: grid (%d,%d,%d) / C=%d not supported", args.n1, args.n2, args.n3, C);
  FusedArgs a = args;
#define MLPCG_FU(TW, TXX, TYY)                                 \
  do {                                                         \
    if (C == 1) launch_fused_t<TW, TXX, TYY, 1>(a, st);        \
    else launch_fused_t<TW, TXX, TYY, 2>(a, st);               \
  } while (0)
  if (tw == kBF16) {
    if (tx == kF32 && ty == kBF16) MLPCG_FU(bf16, float, bf16);
    else if (tx == kBF16 && ty == kF32) MLPCG_FU(bf16, bf16, float);
    else if (tx == kBF16 && ty == kBF16) MLPCG_FU(bf16, bf16, bf16);
    else MLPCG_REQUIRE(false, "sm_fused: vector types (%d,%d) unsupported", tx, ty);
  } else if (tw == kF16) {
    if (tx == kF32 && ty == kF16) MLPCG_FU(f16, float, f16);
    else if (tx == kF16 && ty == kF32) MLPCG_FU(f16, f16, float);
    else if (tx == kF16 && ty == kF16) MLPCG_FU(f16, f16, f16);
    else MLPCG_REQUIRE(false, "sm_fused: vector types (%d,%d) unsupported", tx, ty);
  } else {
    MLPCG_REQUIRE(false, "sm_fused: 16-bit stencil type expected (got %d)", tw);
  }
#undef MLPCG_FU
  return check_launch("sm_fused");
}

}  // namespace mlpcg
