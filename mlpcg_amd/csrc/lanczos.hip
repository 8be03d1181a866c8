// Vector kernels of the Lanczos process that builds the Ritz-vector training right-hand
// sides (lib/create_dataset.py:28-52, FP64 like the reference): an FP64 dot product into a
// device scalar, the three-term update w -= alpha v_j + beta v_{j-1}, and the
// normalisation v = w / ||w|| that also records beta = ||w||.  The matrix product is the
// CSR FP64 SpMV of pcg.hip; the reorthogonalisation against the basis is a plain GEMV.
// All scalars stay on the device.
#include "common.hpp"
#include "pcg.hpp"

namespace mlpcg {

namespace {

constexpr int kLzRun = 8;           // doubles per thread per vector (two 32-byte loads)
constexpr int kLzMaxBlocks = 8192;  // partial rows of the dot (<= reduction workspace)

__global__ __launch_bounds__(kThreads) void dot_f64_kernel(const double* __restrict__ a,
                                                           const double* __restrict__ b, long n,
                                                           double* __restrict__ partials) {
  double red[1] = {0.0};
  const long nr = n / kLzRun;
  for (long q = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x; q < nr;
       q += (long)gridDim.x * kThreads) {
    double av[kLzRun], bv[kLzRun];
    load_vec<kLzRun>(a + kLzRun * q, av);
    load_vec<kLzRun>(b + kLzRun * q, bv);
#pragma unroll
    for (int t = 0; t < kLzRun; ++t) red[0] += av[t] * bv[t];
  }
  // tail (n % kLzRun elements) in block 0
  if (blockIdx.x == 0)
    for (long i = nr * kLzRun + threadIdx.x; i < n; i += kThreads) red[0] += a[i] * b[i];
  block_partials<1>(red, partials);
}

__global__ __launch_bounds__(kFinThreads) void dot_f64_final_kernel(const double* __restrict__ partials,
                                                                   unsigned nb, double* __restrict__ out) {
  const double s = sum_partials(partials, nb);
  if (threadIdx.x == 0) out[0] = s;
}

// w -= (*alpha) v1 + (*beta) v0   (v0 / beta may be null)
__global__ __launch_bounds__(kThreads) void lanczos_update_kernel(double* __restrict__ w,
                                                                  const double* __restrict__ v1,
                                                                  const double* __restrict__ v0,
                                                                  const double* __restrict__ alpha,
                                                                  const double* __restrict__ beta, long n) {
  const double al = *alpha, be = v0 ? *beta : 0.0;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
    const double r = v0 ? v0[i] : 0.0;
    w[i] = w[i] - al * v1[i] - be * r;
  }
}

// out = w / sqrt(*nrm2); *norm_out = sqrt(*nrm2)
__global__ __launch_bounds__(kThreads) void normalize_f64_kernel(double* __restrict__ out,
                                                                 const double* __restrict__ w,
                                                                 const double* __restrict__ nrm2,
                                                                 double* __restrict__ norm_out, long n) {
  const double nr = sqrt(*nrm2);
  if (norm_out && blockIdx.x == 0 && threadIdx.x == 0) norm_out[0] = nr;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
    out[i] = w[i] / nr;
}

unsigned elementwise_blocks(long n) {
  long nb = (n + kThreads - 1) / kThreads;
  if (nb > 16384) nb = 16384;
  return (unsigned)(nb < 1 ? 1 : nb);
}

}  // namespace

size_t dot_f64_workspace_bytes() { return (size_t)kLzMaxBlocks * sizeof(double); }

int launch_dot_f64(const double* a, const double* b, long n, double* out, double* partials, hipStream_t st) {
  MLPCG_REQUIRE(n >= 0, "dot_f64: negative length");
  long nb = (n / kLzRun + kThreads - 1) / kThreads;
  if (nb > kLzMaxBlocks) nb = kLzMaxBlocks;
  if (nb < 1) nb = 1;
  dot_f64_kernel<<<(unsigned)nb, kThreads, 0, st>>>(a, b, n, partials);
  dot_f64_final_kernel<<<1, kFinThreads, 0, st>>>(partials, (unsigned)nb, out);
  return check_launch("dot_f64");
}

int launch_lanczos_update(double* w, const double* v1, const double* v0, const double* alpha,
                          const double* beta, long n, hipStream_t st) {
  MLPCG_REQUIRE(!v0 || beta, "lanczos_update: v0 needs beta");
  lanczos_update_kernel<<<elementwise_blocks(n), kThreads, 0, st>>>(w, v1, v0, alpha, beta, n);
  return check_launch("lanczos_update");
}

int launch_normalize_f64(double* out, const double* w, const double* nrm2, double* norm_out, long n,
                         hipStream_t st) {
  normalize_f64_kernel<<<elementwise_blocks(n), kThreads, 0, st>>>(out, w, nrm2, norm_out, n);
  return check_launch("normalize_f64");
}

}  // namespace mlpcg
