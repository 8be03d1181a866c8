// Kernel argument blocks and host launchers shared by the C ABI and the native runtime.
#pragma once

#include "common.hpp"

namespace mlpcg {

// y[b] = epilogue( sum_p K[p] * x[b](. + d_p) ).  Epilogue, in order:
//   val *= *in_scale              (if in_scale; FP64 scalar on the device)
//   val = 0 where mask_a == 0     (if mask_a; the model's fluid mask)
//   val = c0 * prev + c1 * val    (if prev; c0/c1 are device scalars of the accumulate type)
//   val = 0 where mask_b == 0     (if mask_b; restriction to the unknowns)
struct ApplyArgs {
  const void* K;
  const void* x;
  void* y;
  const void* prev;
  const void* c0;
  const void* c1;
  const uint8_t* mask_a;
  const uint8_t* mask_b;
  const double* in_scale;
  const int* done;
  int B, n1, n2, n3;
};

// out[b] = scale * up(in[b])  (or c0 * prev + c1 * scale * up(in)); fine dims n1..n3.
struct UpArgs {
  const void* in;
  void* out;
  const void* prev;
  const void* c0;
  const void* c1;
  const int* done;
  double scale;
  int nearest;
  int B, n1, n2, n3;
};

// Fused apply for BF16 plans (fused.hip): K recomputed on the matrix cores from the FP16
// image img [C][V] and the packed weights (launch_fused_pack), applied to x with the same
// epilogue as ApplyArgs.  Grids with n1, n2 % 8 == 0 and n3 % 64 == 0, B = 1, C <= 3
// (fused_eligible); fused_blocks(C, ...) workgroups per launch.
struct FusedArgs {
  const void* img;
  const void* apack;  // packed weights and bias (launch_fused_pack)
  const void* x;
  void* y;
  const void* prev;
  const void* c0;
  const void* c1;
  const uint8_t* mask_a;
  const uint8_t* mask_b;
  const double* in_scale;
  const int* done;
  // optional: per-block partial sums [2][blocks] (doubles) of y.dot0 and y.dot1 over the
  // full grid (dot1 may be null), for a following finalize (solver dot products on z)
  const float* dot0;
  const float* dot1;
  double* partials;
  // optional (BF16 output, no epilogue operands): also write the 2x2x2 means of y, the
  // next level's input, as avgpool2 computes them (same sums in the same order)
  void* pool_out;
  int n1, n2, n3;
};

size_t fused_weights_bytes(int C);
unsigned fused_blocks(int C, int n1, int n2, int n3);
bool fused_eligible(int C, int B, int n1, int n2, int n3);
// W [27][C][27] (FP32); fold (C = 3): pack two channels, W[p][c] + W[p][2] for c = 0, 1
int launch_fused_pack(int tw, const void* W, const void* b, int C, void* wbuf, hipStream_t st, bool fold = false);
int launch_sm_fused(int tw, int tx, int ty, int C, const FusedArgs& a, hipStream_t st);
// The stencils K of the fused apply's first product (a.img, a.apack, a.n1..n3), stored as
// BF16 [27][V] (the stored-stencil layout) instead of applied.
int launch_fused_kwrite(int C, const FusedArgs& a, void* K, hipStream_t st);
// Input gradient of a BF16 block (no transpose) through the nine column sums S of
// P_p(v) = g(v) K_p(v) (g in a.x, BF16; K from the first product): dx(u) = sum_c S_c(u + d_c)
// on the matrix-core grids of the fused apply.  work: fused_sgrad_bytes() bytes.
size_t fused_sgrad_bytes(int n1, int n2, int n3);
int launch_fused_sgrad(int C, const FusedArgs& a, void* work, void* dx, hipStream_t st);

// The two coarsest levels of the SmallSM V-cycle in one workgroup (tail.hip): from the
// output x_in of pre[n-2] (level n-2 grid) to up(c0 x + c1 post[n-1](up(l(pool(x))))) on
// the level n-2 grid.  f1..f3: level n-1 dims (even; at most kTailMaxFine voxels, and the
// level n-1 and n grids with a one-cell halo at most kTailMaxHalo / kTailMaxHaloCoarse).
constexpr int kTailMaxFine = 4096;
constexpr int kTailMaxHalo = 8192;
constexpr int kTailMaxHaloCoarse = 2048;
struct TailArgs {
  const void* x_in;
  const void* K_pre;
  const void* K_post;
  const void* K_l;
  const float* c0;
  const float* c1;
  void* out;
  const int* done;  // the solver's stop flag (every workgroup returns at once when set), or null
  int nearest;
  int f1, f2, f3;
};
bool tail_eligible(int dtype, int f1, int f2, int f3);
int launch_coarse_tail(int dtype, const TailArgs& a, hipStream_t st);

int launch_stencil_weights(int dtype, const void* img, const void* W, const void* b, void* K, int C,
                           int n1, int n2, int n3, bool trans, hipStream_t st);
int launch_sm_apply(int tk, int tx, int ty, const ApplyArgs& a, hipStream_t st);
// y[b, v] = sum_p K_p(v - d_p) x[b, v - d_p] from the forward stencils K (K, x, y of type
// tk; uses a.K, a.x, a.y, a.B and the grid only)
int launch_sm_apply_transposed(int tk, const ApplyArgs& a, hipStream_t st);
int launch_avgpool2(int dtype, const void* in, void* out, int B, int n1, int n2, int n3,
                    const int* done, hipStream_t st);
int launch_upsample2(int dtype, const UpArgs& a, hipStream_t st);
int launch_upsample2_adjoint(int dtype, const void* gin, void* gout, int B, int n1, int n2, int n3,
                             int nearest, double scale, hipStream_t st);

// SM linear: ybar[s,e] = mean_v img[s, v+e]; z = w . ybar + b.  ybar/z are of the
// accumulate type; partials needs mean_patch_partials_size() bytes.
size_t mean_patch_workspace(int dtype, int C, int n1, int n2, int n3);
int launch_sm_linear(int dtype, const void* img, const void* w, const void* b, void* z, void* ybar,
                     void* workspace, int C, int n1, int n2, int n3, hipStream_t st);
// Two SM linears on one image (one mean patch): z0 = w0 . ybar + b0, z1 = w1 . ybar + b1
// (w1 may be null: z1 untouched).
int launch_sm_linear_pair(int dtype, const void* img, const void* w0, const void* b0, void* z0,
                          const void* w1, const void* b1, void* z1, void* ybar, void* workspace, int C,
                          int n1, int n2, int n3, hipStream_t st);
// The same with the partials in the library's own device array (no allocation, no
// workspace argument); bytes of that array per GPU.
int launch_sm_linear_noalloc(int dtype, const void* img, const void* w, const void* b, void* z, void* ybar,
                             int C, int n1, int n2, int n3, hipStream_t st);
size_t sm_linear_scratch_bytes();

// Weight gradient of the SM block: gW[p,s,e] = sum_{b,v} a[b,v] c[b,v+d_p] img[s,v+e],
// gb[p] = sum_{b,v} a[b,v] c[b,v+d_p]; (a, c) = (grad_out, x) for the block and
// (x, grad_out) for the transpose.
size_t weight_grad_workspace(int dtype, int C, int B, int n1, int n2, int n3);
int launch_weight_grad(int dtype, const void* a, const void* c, const void* img, void* gW, void* gb,
                       void* workspace, int C, int B, int n1, int n2, int n3, hipStream_t st);

int launch_linear_backward(int dtype, const void* g, const void* ybar, void* gw, void* gb, int C,
                           hipStream_t st);

// Stencil weights with an image of another type (image in the accumulate type, K in
// the storage type): tin in {F32, F64}.
int launch_stencil_weights2(int tin, int tk, const void* img, const void* W, const void* b, void* K,
                            int C, int n1, int n2, int n3, bool trans, hipStream_t st);

int launch_nonzero_mask(int acc, const void* ch, uint8_t* m, long V, hipStream_t st);
int launch_square(int acc, const void* in, void* out, int n, hipStream_t st);
int launch_cast(int tin, int tout, const void* in, void* out, long n, hipStream_t st);

}  // namespace mlpcg
