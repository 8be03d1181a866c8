// extern "C" entry points of libmlpcg_hip.so (declared in include/mlpcg.h).
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mlpcg.h"
#include "runtime.hpp"

namespace mlpcg {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

static size_t elem_size(int dt) { return dt == kF64 ? 8 : (dt == kF32 ? 4 : 2); }

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// BF16 block forwards on grids of at least this many voxels use the fused matrix-core
// apply (as the BF16 plans do); smaller grids stream stored stencils.
constexpr size_t kFusedForwardMinVoxels = (size_t)1 << 18;

}  // namespace mlpcg

using namespace mlpcg;

struct mlpcg_precond {
  Precond* impl;
};
struct mlpcg_solver {
  Solver* impl;
};

#define STREAM(s) ((hipStream_t)(s))

static int check_block_args(const char* who, int C, int B, int n1, int n2, int n3, int dtype) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "%s: C must be 1, 2 or 3 (got %d)", who, C);
  MLPCG_REQUIRE(B >= 1 && n1 >= 1 && n2 >= 1 && n3 >= 1, "%s: bad sizes B=%d grid=(%d,%d,%d)", who, B,
                n1, n2, n3);
  MLPCG_REQUIRE(dtype >= kF64 && dtype <= kBF16, "%s: bad dtype %d", who, dtype);
  MLPCG_REQUIRE((long)n1 * n2 * n3 * 27 < (1L << 31) * 4, "%s: grid too large", who);
  return 0;
}

extern "C" {

const char* mlpcg_last_error(void) { return g_err; }
int mlpcg_version(void) { return 200; }

// --------------------------------------------------------------------------- SM block
static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Workspace: the stencils K [27][V], the weight-gradient partials, and for BF16 an FP16
// copy of the image plus the packed weights of the matrix-core kernels.
size_t mlpcg_sm_block_workspace_size(int C, int B, int n1, int n2, int n3, int dtype) {
  const size_t V = (size_t)n1 * n2 * n3;
  size_t n = align256(27 * V * elem_size(dtype)) + align256(weight_grad_workspace(dtype, C, B, n1, n2, n3));
  if (dtype == kBF16 && C >= 1 && C <= 3) n += align256((size_t)C * V * 2) + align256(fused_weights_bytes(C));
  return n;
}

static int block_fwd(bool trans, const void* image, const void* x, const void* weight,
                     const void* bias, void* y, void* workspace, int C, int B, int n1, int n2,
                     int n3, int dtype, void* stream, const void* image16 = nullptr,
                     const void* wpack_in = nullptr) {
  const char* who = trans ? "sm_block_trans" : "sm_block";
  if (check_block_args(who, C, B, n1, n2, n3, dtype)) return -1;
  MLPCG_REQUIRE(image && x && weight && bias && y && workspace, "%s: null pointer", who);
  MLPCG_REQUIRE(aligned16(x) && aligned16(y) && aligned16(workspace) && aligned16(image),
                "%s: pointers must be 16-byte aligned", who);
  const size_t V = (size_t)n1 * n2 * n3;
  if (!trans && dtype == kBF16 && V >= kFusedForwardMinVoxels && fused_eligible(C, B, n1, n2, n3)) {
    // BF16 block: K recomputed on the matrix cores (the fused apply of the BF16 plans)
    // from an FP16 copy of the image (exact for BF16 image values in the FP16 normal
    // range) and the FP32 weights packed for it; no stencils are stored.  K, x and the
    // sums round as in the stored-stencil apply (K to BF16, FP32 sums).
    char* ws = (char*)workspace;
    const void* img16 = image16 ? image16 : ws;
    void* wbuf = ws + ((C * V * 2 + 255) & ~(size_t)255);
    const void* wpack = wpack_in ? wpack_in : wbuf;
    if (!image16 && launch_cast(kBF16, kF16, image, ws, (long)(C * V), STREAM(stream))) return -1;
    if (!wpack_in && launch_fused_pack(kBF16, weight, bias, C, wbuf, STREAM(stream))) return -1;
    FusedArgs f;
    memset(&f, 0, sizeof(f));
    f.img = img16;
    f.apack = wpack;
    f.x = x;
    f.y = y;
    f.n1 = n1;
    f.n2 = n2;
    f.n3 = n3;
    return launch_sm_fused(kBF16, kBF16, kBF16, C, f, STREAM(stream));
  }
  if (launch_stencil_weights(dtype, image, weight, bias, workspace, C, n1, n2, n3, trans, STREAM(stream)))
    return -1;
  ApplyArgs a;
  memset(&a, 0, sizeof(a));
  a.K = workspace;
  a.x = x;
  a.y = y;
  a.B = B;
  a.n1 = n1;
  a.n2 = n2;
  a.n3 = n3;
  return launch_sm_apply(dtype, dtype, dtype, a, STREAM(stream));
}

static int block_bwd(bool trans, const void* g, const void* image, const void* x, const void* weight,
                     const void* bias, void* gx, void* gw, void* gb, void* workspace, int C, int B,
                     int n1, int n2, int n3, int dtype, void* stream, const void* image16 = nullptr,
                     const void* wpack_in = nullptr) {
  const char* who = trans ? "sm_block_trans_backward" : "sm_block_backward";
  if (check_block_args(who, C, B, n1, n2, n3, dtype)) return -1;
  MLPCG_REQUIRE(g && image && x && weight && bias && gx && gw && gb && workspace, "%s: null pointer", who);
  MLPCG_REQUIRE(aligned16(g) && aligned16(x) && aligned16(gx) && aligned16(workspace) && aligned16(image),
                "%s: pointers must be 16-byte aligned", who);
  const size_t V = (size_t)n1 * n2 * n3;
  const size_t kbytes = align256(27 * V * elem_size(dtype));
  void* part = (char*)workspace + kbytes;
  // the input gradient is the adjoint apply, both from the forward stencils K (recomputed
  // here, not saved from the forward pass): for the block y = sum_p K_p(v) x(v + d_p) it
  // is the transposed gather sum_p K_p(v - d_p) g(v - d_p); for the transpose block it is
  // the block apply itself.  BF16 grids that the fused apply takes get K from its first
  // product (matrix cores, an FP16 image copy): the block sums g K along z-rows into nine
  // column sums and gathers those; the transpose block stores K and applies it.
  if (dtype == kBF16 && V >= kFusedForwardMinVoxels && fused_eligible(C, B, n1, n2, n3)) {
    char* ext = (char*)part + align256(weight_grad_workspace(dtype, C, B, n1, n2, n3));
    const void* img16 = image16 ? image16 : ext;
    void* wbuf = ext + align256((size_t)C * V * 2);
    const void* wpack = wpack_in ? wpack_in : wbuf;
    if (!image16 && launch_cast(kBF16, kF16, image, ext, (long)(C * V), STREAM(stream))) return -1;
    if (!wpack_in && launch_fused_pack(kBF16, weight, bias, C, wbuf, STREAM(stream))) return -1;
    FusedArgs f;
    memset(&f, 0, sizeof(f));
    f.img = img16;
    f.apack = wpack;
    f.n1 = n1;
    f.n2 = n2;
    f.n3 = n3;
    if (!trans) {
      // the block's input gradient through the nine column sums of g K (in the K area)
      MLPCG_REQUIRE(fused_sgrad_bytes(n1, n2, n3) <= kbytes, "%s: workspace layout", who);
      f.x = g;
      if (launch_fused_sgrad(C, f, workspace, gx, STREAM(stream))) return -1;
      const void* av = g;
      return launch_weight_grad(dtype, av, x, image, gw, gb, part, C, B, n1, n2, n3, STREAM(stream));
    }
    if (launch_fused_kwrite(C, f, workspace, STREAM(stream))) return -1;
  } else if (launch_stencil_weights(dtype, image, weight, bias, workspace, C, n1, n2, n3, false,
                                    STREAM(stream))) {
    return -1;
  }
  ApplyArgs a;
  memset(&a, 0, sizeof(a));
  a.K = workspace;
  a.x = g;
  a.y = gx;
  a.B = B;
  a.n1 = n1;
  a.n2 = n2;
  a.n3 = n3;
  if (trans) {
    if (launch_sm_apply(dtype, dtype, dtype, a, STREAM(stream))) return -1;
  } else {
    if (launch_sm_apply_transposed(dtype, a, STREAM(stream))) return -1;
  }
  // block: G_p(v) = sum_b g[b,v] x[b,v+d_p]; transpose: H_p(u) = sum_b x[b,u] g[b,u+d_p]
  const void* av = trans ? x : g;
  const void* cv = trans ? g : x;
  return launch_weight_grad(dtype, av, cv, image, gw, gb, part, C, B, n1, n2, n3, STREAM(stream));
}

int mlpcg_sm_block_inference(const void* image, const void* x, const void* weight, const void* bias,
                             void* y, void* workspace, int C, int B, int n1, int n2, int n3,
                             int dtype, void* stream) {
  return block_fwd(false, image, x, weight, bias, y, workspace, C, B, n1, n2, n3, dtype, stream);
}
int mlpcg_sm_block_forward(const void* image, const void* x, const void* weight, const void* bias,
                           void* y, void* workspace, int C, int B, int n1, int n2, int n3, int dtype,
                           void* stream) {
  return block_fwd(false, image, x, weight, bias, y, workspace, C, B, n1, n2, n3, dtype, stream);
}
int mlpcg_sm_block_backward(const void* grad_out, const void* image, const void* x,
                            const void* weight, const void* bias, void* grad_x, void* grad_w,
                            void* grad_b, void* workspace, int C, int B, int n1, int n2, int n3,
                            int dtype, void* stream) {
  return block_bwd(false, grad_out, image, x, weight, bias, grad_x, grad_w, grad_b, workspace, C, B, n1,
                   n2, n3, dtype, stream);
}
int mlpcg_sm_block_forward_ex(const void* image, const void* image16, const void* wpack, const void* x,
                              const void* weight, const void* bias, void* y, void* workspace, int C, int B, int n1,
                              int n2, int n3, int dtype, void* stream) {
  MLPCG_REQUIRE((!image16 && !wpack) || dtype == kBF16, "sm_block_forward_ex: image16 / wpack need dtype BF16");
  MLPCG_REQUIRE((!image16 || aligned16(image16)) && (!wpack || aligned16(wpack)),
                "sm_block_forward_ex: image16 and wpack must be 16-byte aligned");
  return block_fwd(false, image, x, weight, bias, y, workspace, C, B, n1, n2, n3, dtype, stream, image16, wpack);
}
int mlpcg_sm_block_backward_ex(const void* grad_out, const void* image, const void* image16, const void* wpack,
                               const void* x, const void* weight, const void* bias, void* grad_x, void* grad_w,
                               void* grad_b, void* workspace, int C, int B, int n1, int n2, int n3, int dtype,
                               void* stream) {
  MLPCG_REQUIRE((!image16 && !wpack) || dtype == kBF16, "sm_block_backward_ex: image16 / wpack need dtype BF16");
  MLPCG_REQUIRE((!image16 || aligned16(image16)) && (!wpack || aligned16(wpack)),
                "sm_block_backward_ex: image16 and wpack must be 16-byte aligned");
  return block_bwd(false, grad_out, image, x, weight, bias, grad_x, grad_w, grad_b, workspace, C, B, n1, n2, n3,
                   dtype, stream, image16, wpack);
}

size_t mlpcg_fused_pack_bytes(int C) { return (C >= 1 && C <= 3) ? fused_weights_bytes(C) : 0; }

int mlpcg_fused_pack(const float* weight, const float* bias, int C, void* wpack, void* stream) {
  MLPCG_REQUIRE(weight && bias && wpack, "fused_pack: null pointer");
  MLPCG_REQUIRE(C >= 1 && C <= 3, "fused_pack: C must be 1, 2 or 3 (got %d)", C);
  MLPCG_REQUIRE(aligned16(wpack), "fused_pack: wpack must be 16-byte aligned");
  return launch_fused_pack(kBF16, weight, bias, C, wpack, STREAM(stream));
}
int mlpcg_sm_block_trans_inference(const void* image, const void* x, const void* weight,
                                   const void* bias, void* y, void* workspace, int C, int B, int n1,
                                   int n2, int n3, int dtype, void* stream) {
  return block_fwd(true, image, x, weight, bias, y, workspace, C, B, n1, n2, n3, dtype, stream);
}
int mlpcg_sm_block_trans_forward(const void* image, const void* x, const void* weight,
                                 const void* bias, void* y, void* workspace, int C, int B, int n1,
                                 int n2, int n3, int dtype, void* stream) {
  return block_fwd(true, image, x, weight, bias, y, workspace, C, B, n1, n2, n3, dtype, stream);
}
int mlpcg_sm_block_trans_backward(const void* grad_out, const void* image, const void* x,
                                  const void* weight, const void* bias, void* grad_x, void* grad_w,
                                  void* grad_b, void* workspace, int C, int B, int n1, int n2,
                                  int n3, int dtype, void* stream) {
  return block_bwd(true, grad_out, image, x, weight, bias, grad_x, grad_w, grad_b, workspace, C, B, n1,
                   n2, n3, dtype, stream);
}

// --------------------------------------------------------------------------- SM linear
size_t mlpcg_sm_linear_workspace_size(int C, int n1, int n2, int n3, int dtype) {
  return mean_patch_workspace(dtype, C, n1, n2, n3);
}
int mlpcg_sm_linear_inference(const void* image, const void* weight, const void* bias, void* z,
                              void* ybar, void* workspace, int C, int n1, int n2, int n3,
                              int dtype, void* stream) {
  if (check_block_args("sm_linear", C, 1, n1, n2, n3, dtype)) return -1;
  MLPCG_REQUIRE(image && weight && bias && z && ybar && workspace, "sm_linear: null pointer");
  MLPCG_REQUIRE(aligned16(image), "sm_linear: image must be 16-byte aligned");
  return launch_sm_linear(dtype, image, weight, bias, z, ybar, workspace, C, n1, n2, n3, STREAM(stream));
}
size_t mlpcg_sm_linear_noalloc_scratch_bytes(void) { return sm_linear_scratch_bytes(); }
int mlpcg_sm_linear_inference_noalloc(const void* image, const void* weight, const void* bias, void* z,
                                      void* ybar, int C, int n1, int n2, int n3, int dtype, void* stream) {
  if (check_block_args("sm_linear_noalloc", C, 1, n1, n2, n3, dtype)) return -1;
  MLPCG_REQUIRE(image && weight && bias && z && ybar, "sm_linear_noalloc: null pointer");
  MLPCG_REQUIRE(aligned16(image), "sm_linear_noalloc: image must be 16-byte aligned");
  return launch_sm_linear_noalloc(dtype, image, weight, bias, z, ybar, C, n1, n2, n3, STREAM(stream));
}
int mlpcg_sm_linear_forward(const void* image, const void* weight, const void* bias, void* z,
                            void* ybar, void* workspace, int C, int n1, int n2, int n3, int dtype,
                            void* stream) {
  return mlpcg_sm_linear_inference(image, weight, bias, z, ybar, workspace, C, n1, n2, n3, dtype, stream);
}
int mlpcg_sm_linear_backward(const void* grad_z, const void* ybar, void* grad_w, void* grad_b, int C,
                             int dtype, void* stream) {
  MLPCG_REQUIRE(grad_z && ybar && grad_w && grad_b, "sm_linear_backward: null pointer");
  MLPCG_REQUIRE(C >= 1 && C <= 3, "sm_linear_backward: C must be 1, 2 or 3");
  return launch_linear_backward(dtype, grad_z, ybar, grad_w, grad_b, C, STREAM(stream));
}

// --------------------------------------------------------------------------- building blocks
int mlpcg_stencil_weights(const void* image, int image_dtype, const void* weight, const void* bias,
                          void* K, int K_dtype, int transposed, int C, int n1, int n2, int n3,
                          void* stream) {
  if (check_block_args("stencil_weights", C, 1, n1, n2, n3, K_dtype)) return -1;
  MLPCG_REQUIRE(image && weight && bias && K, "stencil_weights: null pointer");
  MLPCG_REQUIRE(aligned16(image) && aligned16(K), "stencil_weights: pointers must be 16-byte aligned");
  return launch_stencil_weights2(image_dtype, K_dtype, image, weight, bias, K, C, n1, n2, n3,
                                 transposed != 0, STREAM(stream));
}

int mlpcg_sm_apply(const void* K, int K_dtype, const void* x, int x_dtype, void* y, int y_dtype,
                   const void* prev, const void* c0, const void* c1, const uint8_t* mask_a,
                   const uint8_t* mask_b, const double* in_scale, int B, int n1, int n2, int n3,
                   void* stream) {
  MLPCG_REQUIRE(K && x && y, "sm_apply: null pointer");
  MLPCG_REQUIRE(!prev || (c0 && c1), "sm_apply: prev needs c0 and c1");
  MLPCG_REQUIRE(aligned16(K) && aligned16(x) && aligned16(y) && (!prev || aligned16(prev)),
                "sm_apply: pointers must be 16-byte aligned");
  ApplyArgs a;
  a.K = K;
  a.x = x;
  a.y = y;
  a.prev = prev;
  a.c0 = c0;
  a.c1 = c1;
  a.mask_a = mask_a;
  a.mask_b = mask_b;
  a.in_scale = in_scale;
  a.done = nullptr;
  a.B = B;
  a.n1 = n1;
  a.n2 = n2;
  a.n3 = n3;
  return launch_sm_apply(K_dtype, x_dtype, y_dtype, a, STREAM(stream));
}

int mlpcg_avgpool2(const void* in, void* out, int B, int n1, int n2, int n3, int dtype, void* stream) {
  MLPCG_REQUIRE(in && out, "avgpool2: null pointer");
  MLPCG_REQUIRE(B >= 1, "avgpool2: B must be >= 1");
  return launch_avgpool2(dtype, in, out, B, n1, n2, n3, nullptr, STREAM(stream));
}

int mlpcg_upsample2(const void* in, void* out, int B, int n1, int n2, int n3, int nearest, double scale,
                    const void* prev, const void* c0, const void* c1, int dtype, void* stream) {
  MLPCG_REQUIRE(in && out, "upsample2: null pointer");
  MLPCG_REQUIRE(!prev || (c0 && c1), "upsample2: prev needs c0 and c1");
  UpArgs u;
  u.in = in;
  u.out = out;
  u.prev = prev;
  u.c0 = c0;
  u.c1 = c1;
  u.done = nullptr;
  u.scale = scale;
  u.nearest = nearest;
  u.B = B;
  u.n1 = n1;
  u.n2 = n2;
  u.n3 = n3;
  return launch_upsample2(dtype, u, STREAM(stream));
}

int mlpcg_upsample2_adjoint(const void* grad_fine, void* grad_coarse, int B, int n1, int n2, int n3,
                            int nearest, double scale, int dtype, void* stream) {
  MLPCG_REQUIRE(grad_fine && grad_coarse, "upsample2_adjoint: null pointer");
  return launch_upsample2_adjoint(dtype, grad_fine, grad_coarse, B, n1, n2, n3, nearest, scale,
                                  STREAM(stream));
}

// --------------------------------------------------------------------------- Laplacian
int mlpcg_flags_code(const int8_t* flags, uint8_t* code, int n1, int n2, int n3, void* stream) {
  MLPCG_REQUIRE(flags && code, "flags_code: null pointer");
  return launch_flags_code(flags, code, n1, n2, n3, STREAM(stream));
}

int mlpcg_spmv_flags(const uint8_t* code, const float* x, float* y, int n1, int n2, int n3, void* stream) {
  MLPCG_REQUIRE(code && x && y, "spmv_flags: null pointer");
  MLPCG_REQUIRE(aligned16(code) && aligned16(x) && aligned16(y), "spmv_flags: pointers must be 16-byte aligned");
  return launch_spmv_flat(code, x, y, n1, n2, n3, STREAM(stream));
}

int mlpcg_loss_resid(const void* x, int x_dtype, const float* b, const uint8_t* code, float* r, int n1, int n2,
                     int n3, double* out2, void* workspace, void* stream) {
  MLPCG_REQUIRE(x && b && code && r && out2 && workspace, "loss_resid: null pointer");
  MLPCG_REQUIRE(aligned16(x) && aligned16(b) && aligned16(code) && aligned16(r) && aligned16(workspace),
                "loss_resid: pointers must be 16-byte aligned");
  MLPCG_REQUIRE(n1 >= 1 && n2 >= 1 && n3 >= 1, "loss_resid: bad grid (%d,%d,%d)", n1, n2, n3);
  return launch_loss_resid(x_dtype, x, b, code, r, n1, n2, n3, (double*)workspace, out2, STREAM(stream));
}

int mlpcg_loss_grad(const float* r, const uint8_t* code, const double* coef, void* g, int g_dtype, int n1, int n2,
                    int n3, void* stream) {
  MLPCG_REQUIRE(r && code && coef && g, "loss_grad: null pointer");
  MLPCG_REQUIRE(aligned16(r) && aligned16(code) && aligned16(g), "loss_grad: pointers must be 16-byte aligned");
  MLPCG_REQUIRE(n1 >= 1 && n2 >= 1 && n3 >= 1, "loss_grad: bad grid (%d,%d,%d)", n1, n2, n3);
  return launch_loss_grad(g_dtype, r, code, coef, g, n1, n2, n3, STREAM(stream));
}

int mlpcg_comb(const void* x0, const void* x1, const void* c0, const void* c1, int c_dtype, void* y, long n,
               int dtype, void* stream) {
  MLPCG_REQUIRE(n >= 0, "comb: bad length %ld", n);
  MLPCG_REQUIRE(n == 0 || (x0 && x1 && c0 && c1 && y), "comb: null pointer");
  return launch_comb(dtype, x0, x1, c0, c1, c_dtype, y, n, STREAM(stream));
}

int mlpcg_comb_backward(const void* g, const void* x0, const void* x1, const void* c0, const void* c1, int c_dtype,
                        void* g0, void* g1, long n, int dtype, void* out2, int out_dtype, void* workspace,
                        void* stream) {
  MLPCG_REQUIRE(n >= 0, "comb_backward: bad length %ld", n);
  MLPCG_REQUIRE(c0 && c1 && out2 && workspace && (n == 0 || (g && x0 && x1)), "comb_backward: null pointer");
  MLPCG_REQUIRE(aligned16(workspace), "comb_backward: workspace must be 16-byte aligned");
  return launch_comb_grad(dtype, g, x0, x1, c0, c1, c_dtype, g0, g1, n, (double*)workspace, out2, out_dtype,
                          STREAM(stream));
}

int mlpcg_spmv_csr(const int* rowptr, const int* col, const void* val, const void* x, void* y, long n,
                   int dtype, void* stream) {
  MLPCG_REQUIRE(rowptr && col && val && x && y, "spmv_csr: null pointer");
  if (dtype == kF32)
    return launch_spmv_csr(rowptr, col, (const float*)val, (const float*)x, (float*)y, n, STREAM(stream));
  MLPCG_REQUIRE(dtype == kF64, "spmv_csr: dtype must be F32 or F64");
  return launch_spmv_csr_f64(rowptr, col, (const double*)val, (const double*)x, (double*)y, n,
                             STREAM(stream));
}

size_t mlpcg_dot_workspace_size(void) { return reduction_workspace_bytes(); }

int mlpcg_dot(const float* x, const float* y, long n, double* out, void* workspace, void* stream) {
  MLPCG_REQUIRE(x && y && out && workspace, "dot: null pointer");
  // workspace: the per-block partial sums of the two-kernel reduction
  return launch_dot(x, y, n, (double*)workspace, out, STREAM(stream));
}

// --------------------------------------------------------------------------- Lanczos (FP64)
size_t mlpcg_dot_f64_workspace_size(void) { return dot_f64_workspace_bytes(); }

int mlpcg_dot_f64(const double* x, const double* y, long n, double* out, void* workspace, void* stream) {
  MLPCG_REQUIRE(x && y && out && workspace, "dot_f64: null pointer");
  MLPCG_REQUIRE(aligned16(x) && aligned16(y), "dot_f64: vectors must be 16-byte aligned");
  return launch_dot_f64(x, y, n, out, (double*)workspace, STREAM(stream));
}

int mlpcg_lanczos_update(double* w, const double* v_j, const double* v_jm1, const double* alpha,
                         const double* beta, long n, void* stream) {
  MLPCG_REQUIRE(w && v_j && alpha, "lanczos_update: null pointer");
  MLPCG_REQUIRE(!v_jm1 || beta, "lanczos_update: v_jm1 needs beta");
  MLPCG_REQUIRE(n >= 0, "lanczos_update: negative length");
  return launch_lanczos_update(w, v_j, v_jm1, alpha, beta, n, STREAM(stream));
}

int mlpcg_normalize_f64(double* out, const double* w, const double* nrm2, double* norm_out, long n,
                        void* stream) {
  MLPCG_REQUIRE(out && w && nrm2, "normalize_f64: null pointer");
  MLPCG_REQUIRE(n >= 0, "normalize_f64: negative length");
  return launch_normalize_f64(out, w, nrm2, norm_out, n, STREAM(stream));
}

// --------------------------------------------------------------------------- plan
int mlpcg_precond_create(mlpcg_precond** out, int kind, int n_levels, int C, int n1, int n2, int n3,
                         int dtype, int nearest, int mask, int swap, const void* image,
                         const void* const* weights, const void* const* biases, void* stream) {
  return mlpcg_precond_create_ex(out, kind, n_levels, C, n1, n2, n3, dtype, nearest, mask, swap,
                                 kDefaultFusedMode, kDefaultCoarseTail, image, weights, biases, stream);
}

int mlpcg_precond_create_ex(mlpcg_precond** out, int kind, int n_levels, int C, int n1, int n2, int n3,
                            int dtype, int nearest, int mask, int swap, int fused_mode, int coarse_tail,
                            const void* image, const void* const* weights, const void* const* biases,
                            void* stream) {
  MLPCG_REQUIRE(out && image && weights && biases, "precond_create: null pointer");
  Precond* P = nullptr;
  if (precond_create(&P, kind, n_levels, C, n1, n2, n3, dtype, nearest, mask, swap, fused_mode,
                     coarse_tail != 0, image, weights, biases, STREAM(stream)))
    return -1;
  *out = new mlpcg_precond{P};
  return 0;
}

int mlpcg_precond_fused_blocks(const mlpcg_precond* P) {
  if (!P || !P->impl) return -1;
  int nf = 0;
  for (const Precond::Blk& b : P->impl->blocks) nf += b.fused;
  return nf;
}

int mlpcg_precond_coarse_tail(const mlpcg_precond* P) { return (P && P->impl) ? P->impl->tail : -1; }

int mlpcg_precond_update(mlpcg_precond* P, const void* image, const void* const* weights,
                         const void* const* biases, void* stream) {
  MLPCG_REQUIRE(P && P->impl, "precond_update: null plan");
  return P->impl->update(image, weights, biases, STREAM(stream));
}

int mlpcg_precond_apply(mlpcg_precond* P, const void* r, void* z, const double* in_scale,
                        const uint8_t* out_mask, void* stream) {
  MLPCG_REQUIRE(P && P->impl && r && z, "precond_apply: null pointer");
  return P->impl->apply((const float*)r, (float*)z, in_scale, out_mask, nullptr, STREAM(stream));
}

size_t mlpcg_precond_bytes(const mlpcg_precond* P) { return P && P->impl ? P->impl->bytes : 0; }

void mlpcg_precond_destroy(mlpcg_precond* P) {
  if (!P) return;
  delete P->impl;
  delete P;
}

int mlpcg_precond_apply_block(mlpcg_precond* P, int block, const void* x, int x_dtype, void* y,
                              int y_dtype, const double* in_scale, void* stream) {
  MLPCG_REQUIRE(P && P->impl && x && y, "precond_apply_block: null pointer");
  return P->impl->apply_block(block, x, x_dtype, y, y_dtype, in_scale, STREAM(stream));
}

// --------------------------------------------------------------------------- solver
int mlpcg_solver_create_flags(mlpcg_solver** out, const int8_t* flags, int n1, int n2, int n3,
                              void* stream) {
  MLPCG_REQUIRE(out, "solver_create_flags: null output");
  Solver* S = nullptr;
  if (solver_create_flat(&S, flags, n1, n2, n3, STREAM(stream))) return -1;
  *out = new mlpcg_solver{S};
  return 0;
}

int mlpcg_solver_create_csr(mlpcg_solver** out, const int* rowptr, const int* col, const float* val,
                            long n, const int* fluid_idx, int n1, int n2, int n3, void* stream) {
  MLPCG_REQUIRE(out, "solver_create_csr: null output");
  Solver* S = nullptr;
  if (solver_create_csr(&S, rowptr, col, val, n, fluid_idx, n1, n2, n3, STREAM(stream))) return -1;
  *out = new mlpcg_solver{S};
  return 0;
}

int mlpcg_solver_solve(mlpcg_solver* S, mlpcg_precond* P, const float* b, const float* x0, float* x,
                       int method, int max_it, double tol, double atol, double* history,
                       mlpcg_solve_info* info, void* stream) {
  MLPCG_REQUIRE(S && S->impl, "solver_solve: null solver");
  SolveInfo si;
  memset(&si, 0, sizeof(si));
  const int rc = S->impl->solve(P ? P->impl : nullptr, b, x0, x, method, max_it, tol, atol, history,
                                &si, STREAM(stream));
  if (rc) return rc;
  if (info) {
    info->iters = si.iters;
    info->status = si.status;
    info->rel_res = si.rel_res;
    info->norm_b = si.norm_b;
  }
  return 0;
}

int mlpcg_solver_set_graphs(mlpcg_solver* S, int enable) {
  MLPCG_REQUIRE(S && S->impl, "solver_set_graphs: null solver");
  S->impl->use_graphs = enable ? 1 : 0;
  if (!enable) S->impl->drop_graphs();
  return 0;
}

void mlpcg_solver_destroy(mlpcg_solver* S) {
  if (!S) return;
  delete S->impl;
  delete S;
}

}  // extern "C"
