// Shared device/host utilities for the gfx950 kernels of the SM preconditioner and PCG.
//
// Layout conventions used by every kernel:
//   grid (n1, n2, n3), last axis fastest, V = n1*n2*n3, index v = (i*n2 + j)*n3 + k;
//   batched vectors are [B][V]; images are [C][V]; stencil weights K are [27][V]
//   (structure of arrays so that a wavefront streams 64 consecutive voxels of one tap);
//   tap p = 9*(d1+1) + 3*(d2+1) + (d3+1) for the offset d in {-1,0,1}^3;
//   values outside the grid are zero.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mlpcg {

typedef __bf16 bf16;
typedef _Float16 f16;

enum DType : int { kF64 = 0, kF32 = 1, kF16 = 2, kBF16 = 3 };

// Accumulation type: FP64 storage accumulates in FP64, every other storage in FP32.
template <typename T> struct AccOf { typedef float type; };
template <> struct AccOf<double> { typedef double type; };

constexpr int kThreads = 256;  // 4 wavefronts of 64 lanes

// ---------------------------------------------------------------------------
// Vector loads/stores of N consecutive elements, split into <=16-byte pieces so that
// the only alignment requirement is min(16, N*sizeof(T)) bytes.
// ---------------------------------------------------------------------------
template <typename T, int N> struct Piece {
  static constexpr int kBytes = (N * (int)sizeof(T) >= 16) ? 16 : N * (int)sizeof(T);
  static constexpr int kElems = kBytes / (int)sizeof(T);
};

template <int N, typename A, typename T>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, A* out) {
  constexpr int E = Piece<T, N>::kElems;
  if constexpr (E >= 2) {
    typedef T vt __attribute__((ext_vector_type(E)));
#pragma unroll
    for (int c = 0; c < N / E; ++c) {
      vt v = *reinterpret_cast<const vt*>(p + c * E);
#pragma unroll
      for (int t = 0; t < E; ++t) out[c * E + t] = (A)v[t];
    }
  } else {
#pragma unroll
    for (int t = 0; t < N; ++t) out[t] = (A)p[t];
  }
}

template <int N, typename T, typename A>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const A* in) {
  constexpr int E = Piece<T, N>::kElems;
  if constexpr (E >= 2) {
    typedef T vt __attribute__((ext_vector_type(E)));
#pragma unroll
    for (int c = 0; c < N / E; ++c) {
      vt v;
#pragma unroll
      for (int t = 0; t < E; ++t) v[t] = (T)in[c * E + t];
      *reinterpret_cast<vt*>(p + c * E) = v;
    }
  } else {
#pragma unroll
    for (int t = 0; t < N; ++t) p[t] = (T)in[t];
  }
}

// Non-temporal variant for single-use streams (stencil weights): keeps the reused
// vectors in L2 instead of the stream.
template <int N, typename A, typename T>
__device__ __forceinline__ void load_vec_nt(const T* __restrict__ p, A* out) {
  constexpr int E = Piece<T, N>::kElems;
  if constexpr (E >= 2) {
    typedef T vt __attribute__((ext_vector_type(E)));
#pragma unroll
    for (int c = 0; c < N / E; ++c) {
      vt v = __builtin_nontemporal_load(reinterpret_cast<const vt*>(p + c * E));
#pragma unroll
      for (int t = 0; t < E; ++t) out[c * E + t] = (A)v[t];
    }
  } else {
#pragma unroll
    for (int t = 0; t < N; ++t) out[t] = (A)__builtin_nontemporal_load(p + t);
  }
}

// Non-temporal stores for outputs no kernel re-reads before they have left L2 (the PCG
// vectors written once per iteration): they do not displace the reused rows from L2.
template <int N, typename A, typename T>
__device__ __forceinline__ void store_vec_nt(T* __restrict__ p, const A* in) {
  constexpr int E = Piece<T, N>::kElems;
  if constexpr (E >= 2) {
    typedef T vt __attribute__((ext_vector_type(E)));
#pragma unroll
    for (int c = 0; c < N / E; ++c) {
      vt v;
#pragma unroll
      for (int t = 0; t < E; ++t) v[t] = (T)in[c * E + t];
      __builtin_nontemporal_store(v, reinterpret_cast<vt*>(p + c * E));
    }
  } else {
#pragma unroll
    for (int t = 0; t < N; ++t) __builtin_nontemporal_store((T)in[t], p + t);
  }
}

// ---------------------------------------------------------------------------
// XCD-aware block order.  Consecutive workgroup ids are dealt round-robin over the
// 8 XCDs; remapping gives every XCD a contiguous range of tiles so the halo rows that
// neighbouring tiles share stay in one L2.
// ---------------------------------------------------------------------------
constexpr int kNumXcd = 8;

__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblk) {
  const unsigned q = nblk / kNumXcd, r = nblk % kNumXcd;
  const unsigned xcd = bid % kNumXcd, slot = bid / kNumXcd;
  // XCDs [0, r) own q+1 tiles, the rest own q.
  const unsigned base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + slot;
}

// ---------------------------------------------------------------------------
// Wavefront / block reductions (64-lane wavefronts).
// ---------------------------------------------------------------------------
template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a 256-thread block; result valid in every thread.  `sh` holds >= 4 values.
template <typename A>
__device__ __forceinline__ A block_sum(A v, A* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  A r = sh[0] + sh[1] + sh[2] + sh[3];
  return r;
}

// Deterministic two-kernel reduction: each block writes its NR partial sums with plain
// stores (partials: [NR][gridDim.x] doubles); a following one-block kernel adds them in
// block order.  No device-scope fences or same-address atomics: on a multi-XCD part an
// agent-scope release writes back the XCD's L2, and thousands of them per kernel cost
// more than the extra launch.  Must be called by all threads of the block.
template <int NR>
__device__ __forceinline__ void block_partials(const double (&v)[NR], double* __restrict__ partials) {
  __shared__ double sh[NR][4];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double s = wave_sum(v[r]);
    if ((threadIdx.x & 63) == 0) sh[r][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nb = gridDim.x;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      partials[(size_t)r * nb + blockIdx.x] = sh[r][0] + sh[r][1] + sh[r][2] + sh[r][3];
  }
}

// Sum of partials[r * nb + 0 .. nb) by one block of kFinThreads threads, in a fixed
// order (per-thread strided sums with 8 independent accumulators, then a tree over the
// threads); the result is valid in thread 0.
constexpr int kFinThreads = 1024;

__device__ __forceinline__ double sum_partials(const double* __restrict__ p, unsigned nb) {
  __shared__ double sh[kFinThreads / 64];
  // up to 8 independent loads per thread in flight (nb <= 8 * kFinThreads in one round)
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  unsigned b = threadIdx.x;
  for (; b + 7 * kFinThreads < nb; b += 8 * kFinThreads) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += p[b + u * kFinThreads];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (b + u * kFinThreads < nb) a[u] += p[b + u * kFinThreads];
  double s = wave_sum(((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7])));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kFinThreads / 64; ++w) t += sh[w];
  return t;
}

// sum_partials for NR rows (row r at p + r * nb) with all rows' loads issued in one round;
// per row the same accumulators and order as sum_partials, so the same results.  Valid
// in thread 0.
template <int NR>
__device__ __forceinline__ void sum_partials_rows(const double* __restrict__ p, unsigned nb, int nr,
                                                  double (&t)[NR]) {
  __shared__ double sh[NR][kFinThreads / 64];
  double a[NR][8];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int u = 0; u < 8; ++u) a[r][u] = 0.0;
  unsigned b = threadIdx.x;
  for (; b + 7 * kFinThreads < nb; b += 8 * kFinThreads) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < nr)
#pragma unroll
        for (int u = 0; u < 8; ++u) a[r][u] += p[(size_t)r * nb + b + u * kFinThreads];
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if (r < nr)
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (b + u * kFinThreads < nb) a[r][u] += p[(size_t)r * nb + b + u * kFinThreads];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double s = wave_sum(((a[r][0] + a[r][1]) + (a[r][2] + a[r][3])) + ((a[r][4] + a[r][5]) + (a[r][6] + a[r][7])));
    if ((threadIdx.x & 63) == 0) sh[r][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double v = 0.0;
      for (int w = 0; w < kFinThreads / 64; ++w) v += sh[r][w];
      t[r] = v;
    }
  }
}

// Where consecutive lanes of a wave own consecutive runs of a z-row (run index = wave
// base + lane), the z-halo of a run is the edge value of the neighbouring lane's run: a
// DPP wave shift instead of a memory load.  wave_from_prev: lane l receives lane l-1's
// value; wave_from_next: lane l+1's.  Lanes 0 / 63 receive garbage and load instead.
__device__ __forceinline__ int dpp_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float wave_from_prev(float v) { return __builtin_bit_cast(float, dpp_prev(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ float wave_from_next(float v) { return __builtin_bit_cast(float, dpp_next(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ double wave_from_prev(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const long long r = ((long long)(unsigned)dpp_prev((int)(u >> 32)) << 32) | (unsigned)dpp_prev((int)u);
  return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double wave_from_next(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const long long r = ((long long)(unsigned)dpp_next((int)(u >> 32)) << 32) | (unsigned)dpp_next((int)u);
  return __builtin_bit_cast(double, r);
}


// Early exit flag shared by all kernels of a running solve: once the solve has
// converged the queued kernels of a speculative iteration return immediately.
__device__ __forceinline__ bool solve_done(const int* done) {
  return done != nullptr && __builtin_nontemporal_load(done) != 0;
}

}  // namespace mlpcg

// ---------------------------------------------------------------------------
// Host-side error plumbing for the C ABI.
// ---------------------------------------------------------------------------
namespace mlpcg {
void set_error(const char* fmt, ...);
int check_launch(const char* what);
}  // namespace mlpcg

#define MLPCG_REQUIRE(cond, ...)            \
  do {                                      \
    if (!(cond)) {                          \
      ::mlpcg::set_error(__VA_ARGS__);      \
      return -1;                            \
    }                                       \
  } while (0)

#define MLPCG_HIP(call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::mlpcg::set_error("%s failed: %s", #call, hipGetErrorString(e_));         \
      return -1;                                                                 \
    }                                                                            \
  } while (0)
