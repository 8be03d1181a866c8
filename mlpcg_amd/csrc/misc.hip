// Small helper kernels of the native runtime.
#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

template <typename A>
__global__ __launch_bounds__(kThreads) void nonzero_mask_kernel(const A* __restrict__ ch,
                                                                uint8_t* __restrict__ m, long V) {
  const long v = (long)blockIdx.x * kThreads + threadIdx.x;
  if (v < V) m[v] = ch[v] != A(0) ? 1 : 0;
}

template <typename A>
__global__ void square_kernel(const A* __restrict__ in, A* __restrict__ out, int n) {
  const int i = threadIdx.x;
  if (i < n) out[i] = in[i] * in[i];
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kThreads) void cast_kernel(const TI* __restrict__ in, TO* __restrict__ out,
                                                        long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = (TO)(typename AccOf<TO>::type)in[i];
}

// 16-bit to 16-bit, 8 elements (16 bytes) per thread: n % 8 == 0, 16-byte aligned buffers
template <typename TI, typename TO>
__global__ __launch_bounds__(kThreads) void cast8_kernel(const TI* __restrict__ in, TO* __restrict__ out,
                                                         long n8) {
  typedef TI vi __attribute__((ext_vector_type(8)));
  typedef TO vo __attribute__((ext_vector_type(8)));
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n8) return;
  const vi a = ((const vi*)in)[i];
  vo b;
#pragma unroll
  for (int k = 0; k < 8; ++k) b[k] = (TO)(float)a[k];
  ((vo*)out)[i] = b;
}

int launch_nonzero_mask(int acc, const void* ch, uint8_t* m, long V, hipStream_t st) {
  const unsigned nb = (unsigned)((V + kThreads - 1) / kThreads);
  if (acc == kF64) nonzero_mask_kernel<double><<<nb, kThreads, 0, st>>>((const double*)ch, m, V);
  else nonzero_mask_kernel<float><<<nb, kThreads, 0, st>>>((const float*)ch, m, V);
  return check_launch("nonzero_mask");
}

int launch_square(int acc, const void* in, void* out, int n, hipStream_t st) {
  if (acc == kF64) square_kernel<double><<<1, 64, 0, st>>>((const double*)in, (double*)out, n);
  else square_kernel<float><<<1, 64, 0, st>>>((const float*)in, (float*)out, n);
  return check_launch("square");
}

int launch_cast(int tin, int tout, const void* in, void* out, long n, hipStream_t st) {
  if (tin == kBF16 && tout == kF16 && n % 8 == 0 && ((uintptr_t)in & 15u) == 0 && ((uintptr_t)out & 15u) == 0) {
    const long n8 = n / 8;
    cast8_kernel<bf16, f16><<<(unsigned)((n8 + kThreads - 1) / kThreads), kThreads, 0, st>>>((const bf16*)in,
                                                                                          (f16*)out, n8);
    return check_launch("cast");
  }
  const unsigned nb = (unsigned)((n + kThreads - 1) / kThreads);
#define MLPCG_CAST(TI_, TIT, TO_, TOT)                                                         \
  if (tin == TI_ && tout == TO_) {                                                             \
    cast_kernel<TIT, TOT><<<nb, kThreads, 0, st>>>((const TIT*)in, (TOT*)out, n);              \
    return check_launch("cast");                                                               \
  }
  MLPCG_CAST(kF32, float, kBF16, bf16)
  MLPCG_CAST(kF32, float, kF16, f16)
  MLPCG_CAST(kF32, float, kF32, float)
  MLPCG_CAST(kF32, float, kF64, double)
  MLPCG_CAST(kBF16, bf16, kF32, float)
  MLPCG_CAST(kBF16, bf16, kF16, f16)
  MLPCG_CAST(kF16, f16, kF32, float)
  MLPCG_CAST(kF64, double, kF32, float)
  MLPCG_CAST(kF64, double, kF64, double)
#undef MLPCG_CAST
  MLPCG_REQUIRE(false, "cast: unsupported pair %d -> %d", tin, tout);
}

}  // namespace mlpcg
