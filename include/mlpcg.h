/*
 * mlpcg.h -- C ABI of the MI355X (gfx950) neural-preconditioned PCG library
 * (libmlpcg_hip.so).
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream (hipStream_t
 * passed as void*, NULL = default stream), launches asynchronously on that stream,
 * and returns 0 on success or -1 with a message available from mlpcg_last_error()
 * (thread-local).  No entry point synchronises the device unless stated.
 *
 * Conventions
 *   grid (n1, n2, n3) with the last axis fastest; V = n1*n2*n3.
 *   image  : [C][V] contiguous, C in {1, 2, 3}
 *   x, y   : [B][V] contiguous (the reference's (B, 1, n1, n2, n3))
 *   weight : [27][C][3][3][3], bias : [27] for SM blocks; [1][C][3][3][3] and [1]
 *            for SM linear -- in the ACCUMULATE type: double for MLPCG_F64, float
 *            for MLPCG_F32 / MLPCG_F16 / MLPCG_BF16.
 *   Values outside the grid are zero.  Unlike the reference, inputs are NOT
 *   pre-padded: the one-voxel halo is handled inside the kernels.
 *   Device pointers must be 16-byte aligned.
 */
#ifndef MLPCG_H
#define MLPCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLPCG_F64 0
#define MLPCG_F32 1
#define MLPCG_F16 2
#define MLPCG_BF16 3

#define MLPCG_NPCG 0      /* cg_tests.py:26   npcg      */
#define MLPCG_NPCG_FLEX 1 /* cg_tests.py:103  npcg_flex */
#define MLPCG_NPSD 2      /* cg_tests.py:189  npsd      */

#define MLPCG_MODEL_SMALL_SM 0  /* SmallSMModelDn3D    sm_model_3d.py:164 */
#define MLPCG_MODEL_SPD 1       /* SPDSMModelDn3D      sm_model_3d.py:293 */
#define MLPCG_MODEL_SMALL_SPD 2 /* SmallSPDSMModelDn3D sm_model_3d.py:388 */

typedef struct mlpcg_precond mlpcg_precond;
typedef struct mlpcg_solver mlpcg_solver;

typedef struct {
  int iters;       /* preconditioner applications, counted as cg_tests.py returns k */
  int status;      /* 0 converged, 1 max_it reached, 2 non-finite residual */
  double rel_res;  /* ||b - A x|| / ||b|| of the last FP64 residual recompute */
  double norm_b;   /* ||b|| */
} mlpcg_solve_info;

/* ---- library ------------------------------------------------------------------ */
const char* mlpcg_last_error(void);
int mlpcg_version(void);

/* ---- SM block (replaces torch_extension/sm_block_3d.cpp) -------------------------
 * y[b,v] = sum_p K_p(v) x[b, v+d_p],  K_p(v) = bias[p] + sum_{s,e} weight[p,s,e] image[s, v+e].
 * workspace: mlpcg_sm_block_workspace_size() bytes (stencil weights + gradient partials; for
 * BF16 also an FP16 image copy and packed weights).  BF16 blocks on grids of >= 2^18 voxels
 * with n1, n2 % 8 == 0, n3 % 64 == 0 (B = 1) run on the matrix cores: image values must lie
 * in the FP16 normal range there. */
size_t mlpcg_sm_block_workspace_size(int C, int B, int n1, int n2, int n3, int dtype);
/* sm_block_3d.cpp:31  smblock3d.inference(image, x, weights, bias) -> [y] */
int mlpcg_sm_block_inference(const void* image, const void* x, const void* weight, const void* bias,
                             void* y, void* workspace, int C, int B, int n1, int n2, int n3,
                             int dtype, void* stream);
/* sm_block_3d.cpp:44  smblock3d.forward(image, x, weights, bias) -> [y] */
int mlpcg_sm_block_forward(const void* image, const void* x, const void* weight, const void* bias,
                           void* y, void* workspace, int C, int B, int n1, int n2, int n3, int dtype,
                           void* stream);
/* sm_block_3d.cpp:57  smblock3d.backward(grad_output, image, x, weights, bias) -> [gx, gW, gb]
 * grad_x = transpose stencil of grad_output; grad_w/grad_b in the accumulate type. */
int mlpcg_sm_block_backward(const void* grad_out, const void* image, const void* x,
                            const void* weight, const void* bias, void* grad_x, void* grad_w,
                            void* grad_b, void* workspace, int C, int B, int n1, int n2, int n3,
                            int dtype, void* stream);

/* No counterpart in the reference: mlpcg_sm_block_forward / _backward with two inputs of
 * the BF16 matrix-core path supplied by the caller, for training loops that reuse a scene
 * image over many steps and a step's weights over its scenes: image16, the FP16 copy of
 * the image (C x V, equal to the BF16 image converted), and wpack, the weights and bias
 * packed by mlpcg_fused_pack (mlpcg_fused_pack_bytes(C) bytes; they must be the packing
 * of weight and bias).  NULL = made in the workspace per call, as by the plain entry
 * points; results are bit-identical.  Grids the matrix-core path does not take ignore
 * both. */
int mlpcg_sm_block_forward_ex(const void* image, const void* image16, const void* wpack, const void* x,
                              const void* weight, const void* bias, void* y, void* workspace, int C, int B, int n1,
                              int n2, int n3, int dtype, void* stream);
int mlpcg_sm_block_backward_ex(const void* grad_out, const void* image, const void* image16, const void* wpack,
                               const void* x, const void* weight, const void* bias, void* grad_x, void* grad_w,
                               void* grad_b, void* workspace, int C, int B, int n1, int n2, int n3, int dtype,
                               void* stream);
size_t mlpcg_fused_pack_bytes(int C);
int mlpcg_fused_pack(const float* weight, const float* bias, int C, void* wpack, void* stream);

/* ---- SM block transpose (replaces torch_extension/sm_block_trans_3d.cpp) ----------
 * y[b,v] = sum_p K_p(v - d_p) x[b, v - d_p]. */
/* sm_block_trans_3d.cpp:30  smblocktrans3d.inference */
int mlpcg_sm_block_trans_inference(const void* image, const void* x, const void* weight,
                                   const void* bias, void* y, void* workspace, int C, int B, int n1,
                                   int n2, int n3, int dtype, void* stream);
/* sm_block_trans_3d.cpp:43  smblocktrans3d.forward */
int mlpcg_sm_block_trans_forward(const void* image, const void* x, const void* weight,
                                 const void* bias, void* y, void* workspace, int C, int B, int n1,
                                 int n2, int n3, int dtype, void* stream);
/* sm_block_trans_3d.cpp:56  smblocktrans3d.backward -> [gx, gW, gb] */
int mlpcg_sm_block_trans_backward(const void* grad_out, const void* image, const void* x,
                                  const void* weight, const void* bias, void* grad_x, void* grad_w,
                                  void* grad_b, void* workspace, int C, int B, int n1, int n2,
                                  int n3, int dtype, void* stream);

/* ---- SM linear (replaces torch_extension/sm_linear_3d.cpp) ------------------------
 * ybar[s,e] = (1/V) sum_v image[s, v+e];  z = sum w*ybar + b.  z ([1]) and ybar
 * ([C*27]) are in the accumulate type. */
size_t mlpcg_sm_linear_workspace_size(int C, int n1, int n2, int n3, int dtype);
/* sm_linear_3d.cpp:36  smlinear3d.inference(image, weights, bias) -> [z, ybar] */
int mlpcg_sm_linear_inference(const void* image, const void* weight, const void* bias, void* z,
                              void* ybar, void* workspace, int C, int n1, int n2, int n3,
                              int dtype, void* stream);
/* sm_linear_3d.cpp:25  smlinear3d.forward(image, weights, bias) -> [z, ybar] */
int mlpcg_sm_linear_forward(const void* image, const void* weight, const void* bias, void* z,
                            void* ybar, void* workspace, int C, int n1, int n2, int n3, int dtype,
                            void* stream);
/* No counterpart in the reference: mlpcg_sm_linear_inference for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same results as mlpcg_sm_linear_inference (same kernels, same
 * order); the reduction partials live in a device array of the library,
 * mlpcg_sm_linear_noalloc_scratch_bytes() bytes per GPU (648 KiB: at most 1024
 * workgroups x 27 C doubles, which covers every grid), held until the process ends.
 * Calls that overlap on one GPU (concurrent streams or threads) share that array and race:
 * serialise them.  Grids whose reduction would not fit are refused with an error. */
size_t mlpcg_sm_linear_noalloc_scratch_bytes(void);
int mlpcg_sm_linear_inference_noalloc(const void* image, const void* weight, const void* bias, void* z,
                                      void* ybar, int C, int n1, int n2, int n3, int dtype,
                                      void* stream);
/* sm_linear_3d.cpp:47  smlinear3d.backward(grad_output, ybar) -> [gw, gb] */
int mlpcg_sm_linear_backward(const void* grad_z, const void* ybar, void* grad_w, void* grad_b,
                             int C, int dtype, void* stream);

/* ---- building blocks of the streamed-stencil path --------------------------------
 * No reference counterpart as separate calls: the reference recomputes K inside each
 * block kernel (sm_block_3d_kernel.cu:46-69) and calls F.avg_pool3d / F.interpolate
 * (sm_model_3d.py:206-230). */
/* K[27][V] (transposed != 0: Kt_q(v) = K_{26-q}(v + d_q), so the transpose block is
 * mlpcg_sm_apply with Kt).  image_dtype in {K_dtype, MLPCG_F32}. */
int mlpcg_stencil_weights(const void* image, int image_dtype, const void* weight, const void* bias,
                          void* K, int K_dtype, int transposed, int C, int n1, int n2, int n3,
                          void* stream);
/* y = epilogue(sum_p K_p x(.+d_p)): *in_scale multiplies the sum (FP64 device scalar),
 * mask_a zeroes where 0, then y = (*c0) prev + (*c1) y if prev != NULL, then mask_b
 * zeroes where 0.  Type combinations: all K_dtype, or x F32 / y K_dtype, or x K_dtype /
 * y F32 (K_dtype in {F16, BF16}). c0/c1 are accumulate-type device scalars. */
int mlpcg_sm_apply(const void* K, int K_dtype, const void* x, int x_dtype, void* y, int y_dtype,
                   const void* prev, const void* c0, const void* c1, const uint8_t* mask_a,
                   const uint8_t* mask_b, const double* in_scale, int B, int n1, int n2, int n3,
                   void* stream);
/* F.avg_pool3d(x, 2): (n1, n2, n3) are the FINE dims (all even). */
int mlpcg_avgpool2(const void* in, void* out, int B, int n1, int n2, int n3, int dtype,
                   void* stream);
/* out = scale * F.interpolate(in, 2, trilinear|nearest, align_corners=False), or
 * out = (*c0) prev + (*c1) scale * up(in) when prev != NULL; FINE dims. */
int mlpcg_upsample2(const void* in, void* out, int B, int n1, int n2, int n3, int nearest,
                    double scale, const void* prev, const void* c0, const void* c1, int dtype,
                    void* stream);
/* Adjoint (transpose) of mlpcg_upsample2 without prev; COARSE dims. */
int mlpcg_upsample2_adjoint(const void* grad_fine, void* grad_coarse, int B, int n1, int n2,
                            int n3, int nearest, double scale, int dtype, void* stream);

/* ---- fluid Laplacian and dots (replaces torch sparse CSR A @ p, cg_tests.py:53) ---- */
/* code[v] = flags[v] == FLUID ? 1 + #non-solid face neighbours : 0 (flag codes of
 * lib/read_data.py:13-16; outside the grid counts as solid). */
int mlpcg_flags_code(const int8_t* flags, uint8_t* code, int n1, int n2, int n3, void* stream);
/* y = A x over the full grid, matrix-free from code; x must vanish off the unknowns. */
int mlpcg_spmv_flags(const uint8_t* code, const float* x, float* y, int n1, int n2, int n3,
                     void* stream);
/* No counterpart in the reference as a call: its training loss is torch arithmetic around
 * a sparse product (loss_functions.py:18-23 residual_loss, train.py:61-72).  One right-hand
 * side of the relative squared residual on the full grid, M = the cells with code != 0:
 * r = b - M A M x (x of dtype MLPCG_F32 or MLPCG_BF16 need not vanish off M) and
 * out2 = [r.r, b.b] in FP64 (workspace: mlpcg_dot_workspace_size() bytes); and the
 * gradient g = ((float)*coef * A M r) * M stored as g_dtype (MLPCG_F32 or MLPCG_BF16). */
int mlpcg_loss_resid(const void* x, int x_dtype, const float* b, const uint8_t* code, float* r, int n1,
                     int n2, int n3, double* out2, void* workspace, void* stream);
int mlpcg_loss_grad(const float* r, const uint8_t* code, const double* coef, void* g, int g_dtype, int n1,
                    int n2, int n3, void* stream);
/* No counterpart in the reference as a call: the level combination of the SM models'
 * upward pass, c0 * x0 + c1 * x1 with one-element coefficient tensors, is torch arithmetic
 * there (sm_model_3d.py:245,258,282,286).  y = c0 x0 + c1 x1 over n elements of dtype
 * MLPCG_F32 or MLPCG_BF16, with torch's roundings (coefficients, of c_dtype MLPCG_F32 or
 * MLPCG_BF16, rounded to dtype; each product and the sum rounded to dtype); its backward
 * g0 = c0 g, g1 = c1 g (g0 or g1 may be NULL) and out2 = [g.x0, g.x1] accumulated in FP64,
 * stored as out_dtype (MLPCG_F64 or MLPCG_F32; workspace: mlpcg_dot_workspace_size() bytes). */
int mlpcg_comb(const void* x0, const void* x1, const void* c0, const void* c1, int c_dtype, void* y, long n,
               int dtype, void* stream);
int mlpcg_comb_backward(const void* g, const void* x0, const void* x1, const void* c0, const void* c1, int c_dtype,
                        void* g0, void* g1, long n, int dtype, void* out2, int out_dtype, void* workspace,
                        void* stream);
/* y = A x, CSR with int32 indices; dtype MLPCG_F32 or MLPCG_F64 (val, x, y alike). */
int mlpcg_spmv_csr(const int* rowptr, const int* col, const void* val, const void* x, void* y,
                   long n, int dtype, void* stream);
/* *out = sum x[i] y[i] accumulated in FP64 (device scalar). */
size_t mlpcg_dot_workspace_size(void);
int mlpcg_dot(const float* x, const float* y, long n, double* out, void* workspace, void* stream);

/* ---- Lanczos vector kernels, FP64 (lib/create_dataset.py:28-52 _lanczos_algorithm) --
 * The reference runs its Lanczos process in numpy; these are the device steps of the
 * same recurrence (the SpMV is mlpcg_spmv_csr with MLPCG_F64, the reorthogonalisation a
 * GEMV).  Scalars are device FP64 values. */
size_t mlpcg_dot_f64_workspace_size(void);
/* *out = x . y (FP64, deterministic two-kernel reduction; x, y 16-byte aligned) -- the
 * w.dot(V[j]) and norm of create_dataset.py:38,42,48 */
int mlpcg_dot_f64(const double* x, const double* y, long n, double* out, void* workspace, void* stream);
/* w -= (*alpha) v_j + (*beta) v_jm1 (v_jm1, beta may be NULL) -- create_dataset.py:40,49 */
int mlpcg_lanczos_update(double* w, const double* v_j, const double* v_jm1, const double* alpha,
                         const double* beta, long n, void* stream);
/* out = w / sqrt(*nrm2); *norm_out = sqrt(*nrm2) if norm_out -- create_dataset.py:42-46 */
int mlpcg_normalize_f64(double* out, const double* w, const double* nrm2, double* norm_out, long n,
                        void* stream);

/* ---- preconditioner plan (SmallSMModelDn3D.eval_forward with K cached) ------------
 * weights/biases: host arrays of device pointers in state_dict order:
 *   SMALL_SM : pre.0..n-1, post.0..n-1, l, c0.0..n-1, c1.0..n-1      (4n+1)
 *   SPD      : l0.0..n-1,  l1.0..n-1,  l, c0.0..n-1, c1.0..n-1       (4n+1)
 *   SMALL_SPD: l0.0..n-1,  l, c0.0..n-1, c1.0..n-1                   (3n+1)
 * image: [C][V] in the accumulate type.  dtype is the storage type of the stencils
 * and level vectors.  Allocates device memory (released by destroy).
 * apply: z = P(r * (*in_scale)) with r, z F32 (F64 for an F64 plan) on the full grid,
 * zeroed where out_mask == 0 if out_mask != NULL. */
int mlpcg_precond_create(mlpcg_precond** out, int kind, int n_levels, int C, int n1, int n2,
                         int n3, int dtype, int nearest, int mask, int swap, const void* image,
                         const void* const* weights, const void* const* biases, void* stream);
/* Recompute everything that depends on the image or the weights (pyramid, stencils,
 * c0/c1) for a new image of the same grid, reusing the plan's memory; same argument
 * conventions as create. */
int mlpcg_precond_update(mlpcg_precond* P, const void* image, const void* const* weights,
                         const void* const* biases, void* stream);
int mlpcg_precond_apply(mlpcg_precond* P, const void* r, void* z, const double* in_scale,
                        const uint8_t* out_mask, void* stream);
size_t mlpcg_precond_bytes(const mlpcg_precond* P);
void mlpcg_precond_destroy(mlpcg_precond* P);
/* mlpcg_precond_create with the plan's execution policies given explicitly (create uses
 * fused_mode 1 and coarse_tail 1).  No counterpart in the reference, whose SM blocks
 * always evaluate the weighted stencil on the fly:
 *   fused_mode: BF16 blocks on grids with n1, n2 % 8 == 0 and n3 % 64 == 0 (C <= 3, not
 *     transposed) recompute K on the matrix cores inside the apply instead of storing it.
 *     0: never (K stored like every other block); 1: on such grids of at least 2^18
 *     voxels (smaller ones are faster with stored K); 2: on every such grid.  Plus 4
 *     (C = 3 only): the image's third channel is the sum of the first two (the
 *     [air, fluid, non-solid] images), so the fused blocks take two channels with the third
 *     channel's weights added to the others' -- the caller vouches for the relation.
 *   coarse_tail: SmallSM plans (swap off, FP32/FP16/BF16) run their two coarsest levels as
 *     one workgroup kernel with the level vectors in LDS (1) instead of seven small kernels
 *     (0), when level n-1 has even dims and at most 4096 voxels (16-bit storage,
 *     n3 % 4 == 0) or 2048 (otherwise).
 * The policies belong to the plan: plans with different policies coexist. */
int mlpcg_precond_create_ex(mlpcg_precond** out, int kind, int n_levels, int C, int n1, int n2,
                            int n3, int dtype, int nearest, int mask, int swap, int fused_mode,
                            int coarse_tail, const void* image, const void* const* weights,
                            const void* const* biases, void* stream);
/* Number of blocks of the plan that use the fused matrix-core apply; whether the plan
 * runs the coarse-tail kernel (1 / 0).  -1 for a null plan.  No reference counterpart. */
int mlpcg_precond_fused_blocks(const mlpcg_precond* P);
int mlpcg_precond_coarse_tail(const mlpcg_precond* P);
/* No counterpart in the reference (it has no plan object): y = block(x * (*in_scale)) for
 * stencil block `block` of the plan alone (block order of the weights array above; not a
 * transposed one), on the grid of its level, without level combine or masks; x and y of
 * the plan's storage type or its vector type.  For measurement and tests. */
int mlpcg_precond_apply_block(mlpcg_precond* P, int block, const void* x, int x_dtype, void* y,
                              int y_dtype, const double* in_scale, void* stream);

/* ---- PCG solver (cg_tests.py npcg / npcg_flex / npsd, test.py:42-51 predict) ------
 * flags solver: vectors live on the full grid (length V, zero off the fluid cells),
 * A is applied matrix-free.  CSR solver: vectors over the n unknowns, P is applied on
 * the grid through scatter/gather with fluid_idx (borrowed; must outlive the solver).
 * solve: stops when ||r||/||b|| <= max(tol, atol/||b||) (npsd: <) or after max_it
 * preconditioner applications; r is recomputed as b - A x in FP64 every iteration;
 * P == NULL means no preconditioning.  history (host, max_it+1 doubles) receives the
 * relative residuals if not NULL.  Synchronises the stream before returning. */
int mlpcg_solver_create_flags(mlpcg_solver** out, const int8_t* flags, int n1, int n2, int n3,
                              void* stream);
int mlpcg_solver_create_csr(mlpcg_solver** out, const int* rowptr, const int* col,
                            const float* val, long n, const int* fluid_idx, int n1, int n2, int n3,
                            void* stream);
int mlpcg_solver_solve(mlpcg_solver* S, mlpcg_precond* P, const float* b, const float* x0,
                       float* x, int method, int max_it, double tol, double atol,
                       double* history, mlpcg_solve_info* info, void* stream);
/* No reference counterpart: replay each iteration as a captured HIP graph (one per
 * buffer parity, re-captured when P, b or the method change) on the solver's own
 * stream, ordered after `stream`.  Same results as direct launches. */
int mlpcg_solver_set_graphs(mlpcg_solver* S, int enable);
void mlpcg_solver_destroy(mlpcg_solver* S);

#ifdef __cplusplus
}
#endif

#endif /* MLPCG_H */
