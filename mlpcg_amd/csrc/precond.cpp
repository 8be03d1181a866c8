// Preconditioner plan: setup (pyramid, stencils, coefficients) and apply.
//
// Block order in Precond::K and in the weights/biases arrays handed to precond_create:
//   kSmallSM  (SmallSMModelDn3D, sm_model_3d.py:164-287):
//             pre[0..n-1], post[0..n-1], l            then c0[0..n-1], c1[0..n-1]
//   kSPD      (SPDSMModelDn3D, :293-383):
//             l0[0..n-1], l1[0..n-1], l               then c0[0..n-1], c1[0..n-1]
//             (K also holds the transposed stencils l1_t[0..n-1], l0_t[0..n-1], l_t)
//   kSmallSPD (SmallSPDSMModelDn3D, :388-444):
//             l0[0..n-1], l                           then c0[0..n-1], c1[0..n-1]
//             (K also holds l0_t[0..n-1], l_t)
#include <atomic>
#include <cstring>

#include "runtime.hpp"

namespace mlpcg {

static size_t dsize(int dt) {
  switch (dt) {
    case kF64: return 8;
    case kF32: return 4;
    default: return 2;
  }
}

// Fused mode 1 uses the matrix-core apply on grids of at least kFusedMinVoxels voxels:
// below that even 2-plane tiles are too few to fill the GPU; at 64^3 with 2-plane tiles
// the fused apply already beats the stored stencils (a 256^3 solve runs 0.75 ms faster
// with level 2 fused).  The coarse tail is on by default: at 256^3, l=5 in BF16 the tail
// kernel takes 21 us against ~40 us of kernel time (plus launch gaps) for the seven
// per-level kernels it replaces.
constexpr long kFusedMinVoxels = 1L << 18;

// plan generation ids: never reused within a process, unlike plan addresses
static std::atomic<unsigned long> g_next_gen{1};

Precond::~Precond() {
  for (void* p : allocs) (void)hipFree(p);
}

static int palloc(Precond* P, void** out, size_t bytes) {
  if (bytes == 0) bytes = 16;
  MLPCG_HIP(hipMalloc(out, bytes));
  P->allocs.push_back(*out);
  P->bytes += bytes;
  return 0;
}

int precond_create(Precond** out, int kind, int n, int C, int n1, int n2, int n3, int dtype,
                   int nearest, int mask, int swap, int fused_mode, int coarse_tail,
                   const void* image, const void* const* weights, const void* const* biases,
                   hipStream_t st) {
  MLPCG_REQUIRE(out != nullptr, "precond_create: null output handle");
  MLPCG_REQUIRE(kind >= kSmallSM && kind <= kSmallSPD, "precond_create: bad model kind %d", kind);
  MLPCG_REQUIRE(n >= 1 && n < kMaxLevels, "precond_create: levels must be in [1,%d)", kMaxLevels);
  MLPCG_REQUIRE(C >= 1 && C <= 3, "precond_create: C must be 1, 2 or 3");
  MLPCG_REQUIRE(dtype >= kF64 && dtype <= kBF16, "precond_create: bad dtype %d", dtype);
  const int div = 1 << n;
  MLPCG_REQUIRE(n1 % div == 0 && n2 % div == 0 && n3 % div == 0,
                "precond_create: grid (%d,%d,%d) must be divisible by 2^levels = %d", n1, n2, n3, div);
  MLPCG_REQUIRE(!mask || C >= 2, "precond_create: mask needs the fluid image channel (C >= 2)");
  MLPCG_REQUIRE(kind == kSmallSM || (!mask && !swap), "precond_create: mask/swap only for SmallSM");
  const int fold = (fused_mode & kFuseFoldThird) ? 1 : 0;
  fused_mode &= ~kFuseFoldThird;
  MLPCG_REQUIRE(fused_mode >= 0 && fused_mode <= 2, "precond_create: fused mode must be 0, 1 or 2 (+4)");
  MLPCG_REQUIRE(!fold || C == 3, "precond_create: folding the third image channel needs C = 3");
  Precond* P = new Precond();
  P->gen = g_next_gen.fetch_add(1);
  P->kind = kind;
  P->n = n;
  P->C = C;
  P->dtype = dtype;
  P->acc = dtype == kF64 ? kF64 : kF32;
  P->nearest = (kind == kSmallSM) ? nearest : 1;
  P->mask = mask;
  P->swap = swap;
  const size_t es = dsize(dtype), as = dsize(P->acc);
  int rc = 0;
#define TRY(x)          \
  do {                  \
    rc = (x);           \
    if (rc) {           \
      delete P;         \
      return rc;        \
    }                   \
  } while (0)
  for (int l = 0; l <= n; ++l) {
    P->dims[l][0] = n1 >> l;
    P->dims[l][1] = n2 >> l;
    P->dims[l][2] = n3 >> l;
    P->V[l] = (long)P->dims[l][0] * P->dims[l][1] * P->dims[l][2];
    TRY(palloc(P, &P->img[l], (size_t)C * P->V[l] * as));
    TRY(palloc(P, &P->xd[l], (size_t)P->V[l] * es));
    TRY(palloc(P, &P->xo[l], (size_t)P->V[l] * es));
    TRY(palloc(P, &P->tmp[l], (size_t)P->V[l] * es));
    if (kind == kSPD) TRY(palloc(P, &P->xs[l], (size_t)P->V[l] * es));
  }
  if (mask) {
    void* m = nullptr;
    TRY(palloc(P, &m, (size_t)P->V[0]));
    P->maskimg = (uint8_t*)m;
  }
  // stencils: (weight index, level, transposed)
  if (kind == kSmallSM) {
    for (int i = 0; i < n; ++i) P->blocks.push_back({i, i, 0, 0});       // pre
    for (int i = 0; i < n; ++i) P->blocks.push_back({n + i, i, 0, 0});   // post
    P->blocks.push_back({2 * n, n, 0, 0});                               // l
    P->nw = 2 * n + 1;
  } else if (kind == kSPD) {
    for (int i = 0; i < n; ++i) P->blocks.push_back({i, i, 0, 0});       // l0
    for (int i = 0; i < n; ++i) P->blocks.push_back({n + i, i, 0, 0});   // l1
    P->blocks.push_back({2 * n, n, 0, 0});                               // l
    for (int i = 0; i < n; ++i) P->blocks.push_back({n + i, i, 1, 0});   // l1_t
    for (int i = 0; i < n; ++i) P->blocks.push_back({i, i, 1, 0});       // l0_t
    P->blocks.push_back({2 * n, n, 1, 0});                               // l_t
    P->nw = 2 * n + 1;
  } else {
    for (int i = 0; i < n; ++i) P->blocks.push_back({i, i, 0, 0});       // l0
    P->blocks.push_back({n, n, 0, 0});                                   // l
    for (int i = 0; i < n; ++i) P->blocks.push_back({i, i, 1, 0});       // l0_t
    P->blocks.push_back({n, n, 1, 0});                                   // l_t
    P->nw = n + 1;
  }
  const bool half = (dtype == kBF16);
  P->cf = fold ? 2 : C;
  for (Precond::Blk& b : P->blocks) {
    const int* d = P->dims[b.level];
    void* K = nullptr;
    void* wp = nullptr;
    const bool big = P->V[b.level] >= kFusedMinVoxels || fused_mode == 2;
    b.fused = (half && !b.trans && fused_mode && big && fused_eligible(P->cf, 1, d[0], d[1], d[2])) ? 1 : 0;
    if (b.fused) {
      TRY(palloc(P, &wp, fused_weights_bytes(P->cf)));
      if (!P->imgw[b.level]) TRY(palloc(P, &P->imgw[b.level], (size_t)P->cf * P->V[b.level] * es));
    } else {
      TRY(palloc(P, &K, (size_t)27 * P->V[b.level] * es));
    }
    P->K.push_back(K);
    P->wpack.push_back(wp);
  }
  // the two coarsest SmallSM levels in one workgroup kernel when they are small enough
  P->tail = (coarse_tail && kind == kSmallSM && !swap && n >= 2 &&
             tail_eligible(dtype, P->dims[n - 1][0], P->dims[n - 1][1], P->dims[n - 1][2]) &&
             !P->blocks[n - 1].fused && !P->blocks[2 * n - 1].fused && !P->blocks[2 * n].fused)
                ? 1
                : 0;
  TRY(palloc(P, &P->coef, (size_t)4 * n * as));
  TRY(palloc(P, &P->lin_ws, mean_patch_workspace(P->acc, C, n1, n2, n3)));
  TRY(palloc(P, &P->ybar, (size_t)27 * C * as));
  TRY(P->update(image, weights, biases, st));
#undef TRY
  *out = P;
  return 0;
}

// Everything that depends on the image (and the weights): pyramid, stencils, SMLinear
// coefficients.  Reuses the buffers allocated at creation.
int Precond::update(const void* image, const void* const* weights, const void* const* biases,
                    hipStream_t st) {
  MLPCG_REQUIRE(image && weights && biases, "precond_update: null pointer");
  const size_t as = dsize(acc);
  int rc;
#define TRY(x)           \
  do {                   \
    rc = (x);            \
    if (rc) return rc;   \
  } while (0)
  // image pyramid in the accumulate precision
  MLPCG_HIP(hipMemcpyAsync(img[0], image, (size_t)C * V[0] * as, hipMemcpyDeviceToDevice, st));
  for (int l = 1; l <= n; ++l)
    TRY(launch_avgpool2(acc, img[l - 1], img[l], C, dims[l - 1][0], dims[l - 1][1], dims[l - 1][2],
                        nullptr, st));
  if (mask) TRY(launch_nonzero_mask(acc, (const char*)img[0] + (size_t)V[0] * as, maskimg, V[0], st));
  for (int l = 0; l <= n; ++l)
    if (imgw[l]) TRY(launch_cast(acc, kF16, img[l], imgw[l], (long)cf * V[l], st));  // FP16 MFMA input
  for (size_t k = 0; k < blocks.size(); ++k) {
    const Blk& b = blocks[k];
    const int* d = dims[b.level];
    if (b.fused)
      TRY(launch_fused_pack(dtype, weights[b.w], biases[b.w], C, wpack[k], st, cf < C));
    else
      TRY(launch_stencil_weights2(acc, dtype, img[b.level], weights[b.w], biases[b.w], K[k], C, d[0],
                                  d[1], d[2], b.trans != 0, st));
  }
  // SMLinear coefficients c0[i], c1[i] from img[i], both from one mean patch of level i
  // (plus their squares for kSmallSPD)
  for (int i = 0; i < n; ++i) {
    const int w0 = nw + i, w1 = nw + n + i;
    char* z0 = (char*)coef + (size_t)i * as;
    char* z1 = (char*)coef + (size_t)(n + i) * as;
    const int* d = dims[i];
    TRY(launch_sm_linear_pair(acc, img[i], weights[w0], biases[w0], z0, weights[w1], biases[w1], z1, ybar,
                              lin_ws, C, d[0], d[1], d[2], st));
  }
  if (kind == kSmallSPD) TRY(launch_square(acc, coef, (char*)coef + (size_t)2 * n * as, 2 * n, st));
#undef TRY
  return 0;
}

// One stencil block: the fused matrix-core kernel or the stored-stencil apply, with the
// level-combine / mask epilogue of ApplyArgs.
int Precond::run_block(int k, int level, const void* x, int tx, void* y, int ty, const void* prev,
                       const void* c0, const void* c1, const uint8_t* ma, const uint8_t* mb,
                       const double* sc, const int* done, hipStream_t st, const float* dot0,
                       const float* dot1, double* partials, void* pool_out) {
  if (blocks[k].fused) {
    FusedArgs f;
    f.pool_out = pool_out;
    f.dot0 = dot0;
    f.dot1 = dot1;
    f.partials = partials;
    f.img = imgw[level];
    f.apack = wpack[k];
    f.x = x;
    f.y = y;
    f.prev = prev;
    f.c0 = c0;
    f.c1 = c1;
    f.mask_a = ma;
    f.mask_b = mb;
    f.in_scale = sc;
    f.done = done;
    f.n1 = dims[level][0];
    f.n2 = dims[level][1];
    f.n3 = dims[level][2];
    return launch_sm_fused(dtype, tx, ty, cf, f, st);
  }
  ApplyArgs a;
  a.K = K[k];
  a.x = x;
  a.y = y;
  a.prev = prev;
  a.c0 = c0;
  a.c1 = c1;
  a.mask_a = ma;
  a.mask_b = mb;
  a.in_scale = sc;
  a.done = done;
  a.B = 1;
  a.n1 = dims[level][0];
  a.n2 = dims[level][1];
  a.n3 = dims[level][2];
  return launch_sm_apply(dtype, tx, ty, a, st);
}

int Precond::apply_block(int k, const void* x, int tx, void* y, int ty, const double* in_scale,
                         hipStream_t st) {
  MLPCG_REQUIRE(k >= 0 && k < (int)blocks.size(), "precond_apply_block: block %d out of range", k);
  MLPCG_REQUIRE(!blocks[k].trans, "precond_apply_block: block %d is a transposed stencil", k);
  MLPCG_REQUIRE((tx == dtype || tx == (dtype == kF64 ? kF64 : kF32)) &&
                    (ty == dtype || ty == (dtype == kF64 ? kF64 : kF32)),
                "precond_apply_block: vector types (%d,%d) do not fit the plan type %d", tx, ty, dtype);
  return run_block(k, blocks[k].level, x, tx, y, ty, nullptr, nullptr, nullptr, nullptr, nullptr, in_scale,
                   nullptr, st);
}

int Precond::apply(const float* r_in, float* z_out, const double* in_scale, const uint8_t* out_mask,
                   const int* done, hipStream_t st, const float* dot0, const float* dot1,
                   double* partials, unsigned* dot_blocks) {
  if (dot_blocks) *dot_blocks = 0;
  const int vdt = dtype == kF64 ? kF64 : kF32;  // type of the solver-side vectors
  const size_t as = dsize(acc);
  auto c0p = [&](int i, int sq) { return (const void*)((const char*)coef + (size_t)(sq * 2 * n + i) * as); };
  auto c1p = [&](int i, int sq) { return (const void*)((const char*)coef + (size_t)(sq * 2 * n + n + i) * as); };
  auto apply_blk = [&](int k, int level, const void* x, int tx, void* y, int ty, const void* prev,
                       const void* c0, const void* c1, const uint8_t* ma, const uint8_t* mb,
                       const double* sc) {
    return run_block(k, level, x, tx, y, ty, prev, c0, c1, ma, mb, sc, done, st);
  };
  auto pool = [&](const void* in, void* outp, int fine_level) {
    const int* d = dims[fine_level];
    return launch_avgpool2(dtype, in, outp, 1, d[0], d[1], d[2], done, st);
  };
  auto up = [&](const void* in, void* outp, int fine_level, double scale, const void* prev,
                const void* c0, const void* c1) {
    UpArgs u;
    u.in = in;
    u.out = outp;
    u.prev = prev;
    u.c0 = c0;
    u.c1 = c1;
    u.done = done;
    u.scale = scale;
    u.nearest = nearest;
    u.B = 1;
    u.n1 = dims[fine_level][0];
    u.n2 = dims[fine_level][1];
    u.n3 = dims[fine_level][2];
    return launch_upsample2(dtype, u, st);
  };
  int rc;
#define TRY(x)           \
  do {                   \
    rc = (x);            \
    if (rc) return rc;   \
  } while (0)

  // ---- downward half (identical for every kind): x_0 = B0(r), x_i = Bi(pool x_{i-1}) ----
  const int ndown = tail ? n - 1 : n;  // with the tail, levels n-1 and n run in one kernel
  // x_i is pooled into tmp[i+1] for the next block (the tail pools its input itself); a
  // fused block without epilogue operands (not level 0 under the model mask) writes that
  // mean in its epilogue, so the pool kernel is left out
  auto pool_into = [&](int i) -> void* {
    const bool pooled = (i + 1 < ndown) || (!tail && i == n - 1);
    return (pooled && blocks[i].fused && !(i == 0 && mask)) ? tmp[i + 1] : nullptr;
  };
  TRY(run_block(0, 0, r_in, vdt, xd[0], dtype, nullptr, nullptr, nullptr, mask ? maskimg : nullptr, nullptr,
                in_scale, done, st, nullptr, nullptr, nullptr, pool_into(0)));
  for (int i = 1; i < ndown; ++i) {
    if (!pool_into(i - 1)) TRY(pool(xd[i - 1], tmp[i], i - 1));
    TRY(run_block(i, i, tmp[i], dtype, xd[i], dtype, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, done,
                  st, nullptr, nullptr, nullptr, pool_into(i)));
  }
  const int kl = (kind == kSmallSPD) ? n : 2 * n;  // index of l
  if (tail) {
    TailArgs t;
    t.x_in = xd[n - 2];
    t.K_pre = K[n - 1];
    t.K_post = K[2 * n - 1];
    t.K_l = K[2 * n];
    t.c0 = (const float*)c0p(n - 1, 0);
    t.c1 = (const float*)c1p(n - 1, 0);
    t.out = tmp[n - 2];
    t.done = done;
    t.nearest = nearest;
    t.f1 = dims[n - 1][0];
    t.f2 = dims[n - 1][1];
    t.f3 = dims[n - 1][2];
    TRY(launch_coarse_tail(dtype, t, st));
  } else {
    if (!pool_into(n - 1)) TRY(pool(xd[n - 1], tmp[n], n - 1));
    TRY(apply_blk(kl, n, tmp[n], dtype, xd[n], dtype, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
  }

  if (kind == kSmallSM) {
    const void* cur = xd[n];
    for (int i = tail ? n - 1 : n; i >= 1; --i) {
      const int lvl = i - 1;
      const bool last = (lvl == 0);
      void* y = last ? (void*)z_out : xo[lvl];
      const int ty = last ? vdt : dtype;
      const uint8_t* ma = (last && mask) ? maskimg : nullptr;
      const uint8_t* mb = last ? out_mask : nullptr;
      if (!swap) {
        if (!(tail && i == n - 1)) TRY(up(cur, tmp[lvl], lvl, 1.0, nullptr, nullptr, nullptr));
        if (last && dot0 && partials && dot_blocks && blocks[n].fused &&
            2L * fused_blocks(cf, dims[0][0], dims[0][1], dims[0][2]) <= reduction_capacity()) {
          // the solver's dot products on z ride in the epilogue of the last block
          TRY(run_block(n, 0, tmp[0], dtype, y, ty, xd[0], c0p(0, 0), c1p(0, 0), ma, mb, nullptr, done, st,
                        dot0, dot1, partials));
          *dot_blocks = fused_blocks(cf, dims[0][0], dims[0][1], dims[0][2]);
          cur = xo[lvl];
          continue;
        }
        TRY(apply_blk(n + lvl, lvl, tmp[lvl], dtype, y, ty, xd[lvl], c0p(lvl, 0), c1p(lvl, 0), ma, mb, nullptr));
      } else {
        TRY(up(cur, tmp[lvl], lvl, 1.0, xd[lvl], c0p(lvl, 0), c1p(lvl, 0)));
        TRY(apply_blk(n + lvl, lvl, tmp[lvl], dtype, y, ty, nullptr, nullptr, nullptr, ma, mb, nullptr));
      }
      cur = xo[lvl];
    }
    return 0;
  }

  if (kind == kSPD) {
    // first half upward: x_i = l1[i-1](up(x_i)/8); x_{i-1} = c0 x_{i-1} + c1 x_i
    const void* cur = xd[n];
    for (int i = n; i >= 1; --i) {
      const int lvl = i - 1;
      TRY(up(cur, tmp[lvl], lvl, 0.125, nullptr, nullptr, nullptr));
      TRY(apply_blk(n + lvl, lvl, tmp[lvl], dtype, xo[lvl], dtype, xd[lvl], c0p(lvl, 0), c1p(lvl, 0),
                    nullptr, nullptr, nullptr));
      cur = xo[lvl];
    }
    // second half downward: xs_0 = x_0 ; xs_{i+1} = pool(l1_t[i](xs_i)); xs_n = l_t(xs_n)
    const int kt1 = 2 * n + 1, kt0 = 3 * n + 1, ktl = 4 * n + 1;
    const void* xs0 = xo[0];
    for (int i = 0; i < n; ++i) {
      const void* src = (i == 0) ? xs0 : xs[i];
      TRY(apply_blk(kt1 + i, i, src, dtype, tmp[i], dtype, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
      TRY(pool(tmp[i], xs[i + 1], i));
    }
    TRY(apply_blk(ktl, n, xs[n], dtype, xd[n], dtype, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    // second half upward: xs_i = l0_t[i](c0 xs_i + c1 up(xs_{i+1})/8)
    cur = xd[n];
    for (int i = n - 1; i >= 0; --i) {
      const void* prev = (i == 0) ? xs0 : xs[i];
      TRY(up(cur, tmp[i], i, 0.125, prev, c0p(i, 0), c1p(i, 0)));
      const bool last = (i == 0);
      void* y = last ? (void*)z_out : xd[i];
      TRY(apply_blk(kt0 + i, i, tmp[i], dtype, y, last ? vdt : dtype, nullptr, nullptr, nullptr, nullptr,
                    last ? out_mask : nullptr, nullptr));
      cur = xd[i];
    }
    return 0;
  }

  // kSmallSPD: x_n = l_t(x_n); x_{i-1} = l0_t[i-1](c0^2 x_{i-1} + c1^2 up(x_i)/8)
  const int kt0 = n + 1, ktl = 2 * n + 1;
  TRY(apply_blk(ktl, n, xd[n], dtype, xo[n], dtype, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
  const void* cur = xo[n];
  for (int i = n; i >= 1; --i) {
    const int lvl = i - 1;
    TRY(up(cur, tmp[lvl], lvl, 0.125, xd[lvl], c0p(lvl, 1), c1p(lvl, 1)));
    const bool last = (lvl == 0);
    void* y = last ? (void*)z_out : xo[lvl];
    TRY(apply_blk(kt0 + lvl, lvl, tmp[lvl], dtype, y, last ? vdt : dtype, nullptr, nullptr, nullptr, nullptr,
                  last ? out_mask : nullptr, nullptr));
    cur = xo[lvl];
  }
  return 0;
#undef TRY
}

}  // namespace mlpcg
