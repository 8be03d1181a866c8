// Native runtime objects: the preconditioner plan and the PCG solver.
#pragma once

#include <vector>

#include "kernels.hpp"
#include "pcg.hpp"

namespace mlpcg {

constexpr int kMaxLevels = 10;

enum ModelKind : int { kSmallSM = 0, kSPD = 1, kSmallSPD = 2 };

// Everything that depends only on the image is computed once at creation: the image
// pyramid (accumulate precision), the stencil weights K of every block (storage
// precision, [27][V_level]) and the SMLinear coefficients c0/c1 (device scalars).  For
// 16-bit plans, blocks on grids the fused kernel tiles keep no K: they hold the weights
// packed for the matrix cores and a 16-bit copy of their level's image instead.
// This is the caching that SmallSMModelDn3D.eval_forward does for imgs/c0/c1 through
// its list arguments (sm_model_3d.py:180-195), extended to K.
struct Precond {
  unsigned long gen = 0;            // unique per plan of the process (graph caches key on it)
  int kind = kSmallSM;
  int n = 0, C = 1, dtype = kF32, acc = kF32;
  int nearest = 0, mask = 0, swap = 0;
  int dims[kMaxLevels + 1][3];
  long V[kMaxLevels + 1];
  void* img[kMaxLevels + 1] = {};   // [C][V_l] accumulate type
  std::vector<void*> K;             // stencils in block order (see precond.cpp)
  void* coef = nullptr;             // accumulate type: c0[0..n-1], c1[0..n-1] (+ squares)
  void* xd[kMaxLevels + 1] = {};    // downward level vectors (storage type)
  void* xo[kMaxLevels + 1] = {};    // upward level vectors (storage type)
  void* tmp[kMaxLevels + 1] = {};   // transfer scratch (storage type)
  void* xs[kMaxLevels + 1] = {};    // SPD second-half vectors (storage type)
  uint8_t* maskimg = nullptr;       // level-0 model mask: img[1] != 0
  struct Blk {
    int w, level, trans;            // weight index, level, transposed stencil
    int fused;                      // 1: K recomputed on the matrix cores (no K stored)
  };
  std::vector<Blk> blocks;          // one per entry of K (nullptr for fused blocks)
  std::vector<void*> wpack;         // fused blocks: packed MFMA weights + bias init
  void* imgw[kMaxLevels + 1] = {};  // FP16 image of levels with fused blocks (cf channels)
  int cf = 1;                       // image channels of the fused blocks: C, or 2 when the
                                    // third channel is folded into the first two
  int tail = 0;                     // SmallSM: levels n-1 and n in one workgroup kernel
  int nw = 0;                       // number of block weights before c0/c1
  void* lin_ws = nullptr;           // SMLinear reduction workspace
  void* ybar = nullptr;             // SMLinear mean patch scratch
  size_t bytes = 0;
  std::vector<void*> allocs;

  int update(const void* image, const void* const* weights, const void* const* biases,
             hipStream_t st);
  // z = P(r * in_scale).  With dot0 (and optionally dot1) and partials, a plan whose last
  // block is fused also writes per-block partial sums [2][*dot_blocks] of z.dot0, z.dot1
  // (flat layout); *dot_blocks = 0 means they were not computed.
  int apply(const float* r, float* z, const double* in_scale, const uint8_t* out_mask,
            const int* done, hipStream_t st, const float* dot0 = nullptr,
            const float* dot1 = nullptr, double* partials = nullptr,
            unsigned* dot_blocks = nullptr);
  int apply_block(int k, const void* x, int tx, void* y, int ty, const double* in_scale,
                  hipStream_t st);
  int run_block(int k, int level, const void* x, int tx, void* y, int ty, const void* prev,
                const void* c0, const void* c1, const uint8_t* ma, const uint8_t* mb,
                const double* sc, const int* done, hipStream_t st, const float* dot0 = nullptr,
                const float* dot1 = nullptr, double* partials = nullptr, void* pool_out = nullptr);
  ~Precond();
};

// Per-plan policies.  fused_mode: eligible BF16 blocks recompute their stencils on the
// matrix cores inside the apply -- 0 never, 1 (default) on grids of at least
// kFusedMinVoxels voxels, 2 on every eligible grid; plus kFuseFoldThird: for three-channel
// images whose third channel is the sum of the first two (the [air, fluid, non-solid]
// images), the fused blocks take two channels with the third channel's weights added to
// the others' (the caller vouches for the relation).  coarse_tail: the two coarsest SmallSM
// levels run as one workgroup kernel when they fit (default on).
constexpr int kDefaultFusedMode = 1;
constexpr int kFuseFoldThird = 4;
constexpr int kDefaultCoarseTail = 1;

int precond_create(Precond** out, int kind, int n, int C, int n1, int n2, int n3, int dtype,
                   int nearest, int mask, int swap, int fused_mode, int coarse_tail,
                   const void* image, const void* const* weights, const void* const* biases,
                   hipStream_t st);

struct SolveInfo {
  int iters;
  int status;         // 0 converged, 1 max_it reached, 2 breakdown (non-finite)
  double rel_res;     // ||r|| / ||b|| of the last FP64 residual recompute
  double norm_b;
};

struct Solver {
  int layout = 0;     // 0 flat grid (matrix-free), 1 CSR over the unknowns
  int n1 = 0, n2 = 0, n3 = 0;
  long V = 0, n = 0;
  uint8_t* code = nullptr;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const float* val = nullptr;
  const int* fluid_idx = nullptr;
  float* x[2] = {};
  float* r[2] = {};
  float* p[2] = {};
  float* z = nullptr;
  float* w = nullptr;
  float* q_hist[3] = {};   // npsd: directions, ring of 3 (two kept + the one being formed)
  float* aq_hist[2] = {};  // npsd: A-products of the two kept directions
  float* grid_in = nullptr;
  float* grid_out = nullptr;
  SolveState* st = nullptr;
  SolveState* st_host = nullptr;   // pinned ring of status snapshots
  double* partials = nullptr;
  double* history = nullptr;
  int hist_cap = 0;
  hipStream_t side = nullptr;
  hipEvent_t ev_iter[4] = {}, ev_stat[4] = {};
  std::vector<void*> allocs;
  // HIP-graph replay of the iteration (two graphs: even / odd buffer parity; npcg and
  // npcg_flex only, the npsd direction ring has period 6)
  int use_graphs = 0;
  hipStream_t work = nullptr;
  hipEvent_t ev_in = nullptr;
  hipGraphExec_t graph_exec[2] = {};
  unsigned long graph_P = 0;  // generation id of the plan the graphs were captured with
  const float* graph_b = nullptr;
  int graph_method = -1;

  int solve(Precond* P, const float* b, const float* x0, float* x_out, int method, int max_it,
            double tol, double atol, double* history_out, SolveInfo* info, hipStream_t st);
  void drop_graphs();
  ~Solver();
};

int solver_create_flat(Solver** out, const int8_t* flags, int n1, int n2, int n3, hipStream_t st);
int solver_create_csr(Solver** out, const int* rowptr, const int* col, const float* val, long n,
                      const int* fluid_idx, int n1, int n2, int n3, hipStream_t st);

}  // namespace mlpcg
