// Fused SM block apply on the matrix cores (BF16 plans).
//
//   y(v) = sum_p K_p(v) x(v + d_p),   K_p(v) = b_p + sum_{c,s} W[p,c,s] img[c, v + d_s]
//
// (sm_block_3d_kernel.cu:46-70).  Instead of streaming a stored K (54 of the 60 bytes per
// voxel of the streamed BF16 apply), K is recomputed on the MFMA units, 16 voxels at a
// time, and applied where the product leaves it:
//   1. D[p][v] = b_p + sum_k W[p][k] P[k][v]: W is 27 taps x 27C inputs padded to
//      32 x 32 NKC, P the image patches of the 16 voxels (an im2col read from LDS), the
//      bias the accumulator;
//   2. y(v) = sum_p x(v + d_p) D[p][v]: every lane holds 8 taps of one voxel after step 1
//      and dots them with the matching x words (BF16 dot2); a cross-lane reduction over
//      the four lane groups completes the sums.
// HBM traffic is x, img and y (+ epilogue operands): about 8 B/voxel instead of 60.
//
// Operand layout of v_mfma_f32_16x16x32_f16 (lane l, element j / i):
//   A[l % 16][8 (l / 16) + j],  B[8 (l / 16) + j][l % 16],  D[4 (l / 16) + i][l % 16].
// Precision: step 1 takes img and W in FP16 (11-bit significands: image values lie in
// [0, 1] and weights are O(1), so the range is ample) with FP32 accumulation and an FP32
// bias, which makes K more accurate than a FP32-computed K rounded to BF16; step 2 rounds
// x and K to BF16 (the storage type of the plan) and accumulates in FP32.
//
// Tile: a block of 4 waves covers FI planes x 8 rows x 64 z-voxels (FI = 8, or 4 / 2
// where 8-plane tiles give too few blocks); x and img of the tile plus a one-cell halo are
// staged in LDS; a wave processes 64 z-voxels of one (plane, row) per step: 4 groups of
// 16 voxels.  A pre block can also write the 2x2x2 means of its output (the next level's
// input) from its epilogue.  The K writer (fused_kwrite_kernel) runs the first product
// alone, operands swapped, and stores K for the input gradient of a BF16 block.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace mlpcg {

namespace {

constexpr int kFz = 64, kFj = 8;  // tile: FI planes (8, or 4 / 2 on small grids) x 8 rows x 64 z
constexpr int kRowCells = kFz + 2, kPj = kFj + 2;  // staged cells per halo row, rows per plane
// LDS row stride (elements): 40 words, so the plane stride is 400 words.  With 32 banks of
// one word, column (d1, d2) of the halo tile then starts at bank 16 d1 + 8 d2 (mod 32),
// which the column-to-lane-group assignment below turns into conflict-free reads.
constexpr int kPz = 80;
static_assert(kPz >= kRowCells && (kPz / 2) % 32 == 8 && (kPj * kPz / 2) % 32 == 16, "bank layout");
__host__ __device__ constexpr int tile_cells(int FI) { return (FI + 2) * kPj * kPz; }

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef bf16 bf16x2 __attribute__((ext_vector_type(2)));

template <typename TW> struct Mfma;
template <> struct Mfma<f16> {
  typedef f16 v8 __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
__host__ __device__ constexpr int nkc(int C) { return C; }

// K slots.  A 64-voxel chunk of a z-row is processed as four groups of 16 voxels that
// share a z parity: group t holds z = 32 (t / 2) + 2n + (t % 2), n = 0..15.  With one
// parity per group, pairs of neighbouring taps along z sit at 4-byte aligned LDS
// addresses for every lane, so a lane's eight K slots come from four 32-bit reads.
// Slot (g, j) of a 32-slot chunk (lane group g = lane / 16 holds slots 8g..8g+7 of the
// MFMA operands) carries tap fused_tap(par, g, j) (q = 9(d1+1) + 3(d2+1) + (d3+1) =
// 3c + d3 + 1 for the (d1, d2) column c, s = 7 the column split over lane groups 0, 2):
//   even z: j = 0..5: (c1,-1) (c1,0) (c1,+1) (c2,+1) (c2,-1) (c2,0); j = 6, 7: (s,-1) (s,0)
//           for g = 0, (s,+1) (spare) for g = 2;
//   odd z:  j = 0..5: (c1,0) (c1,+1) (c1,-1) (c2,-1) (c2,0) (c2,+1); j = 6, 7: (s,0) (s,+1)
//           for g = 0, (spare) (s,-1) for g = 2;
// lane groups 1 and 3 have spares in j = 6, 7.  Registers hold
// slot pairs (2i, 2i+1): the pairs (c1,-1)(c1,0), (c2,-1)(c2,0) of an even voxel and
// (c1,0)(c1,+1), (c2,0)(c2,+1) of an odd voxel are aligned words, and the remaining pair is
// put together from the halves of two words that the other parity reads as its pairs.
// Spare slots read valid cells and get zero weights.  Columns go to lane groups as
// (c1, c2) = (4, 0), (1, 2), (5, 6), (3, 8): the LDS serves a wave's read in two halves
// of 32 lanes (lane groups 0, 1 and 2, 3), and in the padded tile (column c at bank
// 16 d1 + 8 d2 mod 32) the two groups of a half always read columns 16 banks apart, each
// group 16 consecutive words: every read is free of bank conflicts.  Lane groups 1 and 3
// read their spare s-column words at the addresses of groups 0 and 2 (a broadcast).  One
// chunk per image channel; the same table orders the rows of D (the taps of K), so the x
// patches of the second product use the same reads.
constexpr int kSplitCol = 7;
__host__ __device__ constexpr int fused_col(int i) {
  return i == 0 ? 4 : i == 1 ? 0 : i == 2 ? 1 : i == 3 ? 2 : i == 4 ? 5 : i == 5 ? 6 : i == 6 ? 3 : 8;
}
__host__ __device__ constexpr int fused_tap(int par, int g, int j) {
  const int c1 = fused_col(2 * g), c2 = fused_col(2 * g + 1), s = 3 * kSplitCol;
  if (par == 0) {
    return j == 0 ? 3 * c1 : j == 1 ? 3 * c1 + 1 : j == 2 ? 3 * c1 + 2 : j == 3 ? 3 * c2 + 2 :
           j == 4 ? 3 * c2 : j == 5 ? 3 * c2 + 1 : j == 6 ? (g == 2 ? s + 2 : s) : s + 1;
  }
  return j == 0 ? 3 * c1 + 1 : j == 1 ? 3 * c1 + 2 : j == 2 ? 3 * c1 : j == 3 ? 3 * c2 :
         j == 4 ? 3 * c2 + 1 : j == 5 ? 3 * c2 + 2 : j == 6 ? s + 1 : (g == 2 ? s : s + 2);
}
__host__ __device__ constexpr bool fused_slot_valid(int par, int g, int j) {
  return j < 6 || g == 0 || (g == 2 && j == (par == 0 ? 6 : 7));
}
// D row rho is slot (g, j) of the second product's K: rho = 4g + j (j < 4), 16 + 4g + j - 4.
__host__ __device__ constexpr int row_g(int rho) { return (rho & 15) >> 2; }
__host__ __device__ constexpr int row_j(int rho) { return (rho & 3) + (rho >= 16 ? 4 : 0); }
// LDS element offset of the (d1, d2) column c in the halo tile
__host__ __device__ constexpr int col_off(int c) { return (c / 3 - 1) * kPj * kPz + (c % 3 - 1) * kPz; }

// base[off] through a uniform base pointer and a 32-bit lane byte offset.  The empty asm
// makes the offset opaque inside the loop body, so the compiler addresses with the scalar
// base plus the 32-bit offset instead of a per-lane 64-bit address.
template <typename T>
__device__ __forceinline__ T lane_at(const T* base, unsigned boff) {
  asm volatile("" : "+v"(boff));
  return *(const T*)((const char*)base + boff);
}
template <typename T>
__device__ __forceinline__ void lane_store(T* base, unsigned boff, T v) {
  asm volatile("" : "+v"(boff));
  *(T*)((char*)base + boff) = v;
}

// Buffer resource over `bytes` bytes from p (raw, stride 0): loads at byte offsets outside
// [0, bytes) return 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}
// element at lane byte offset boff (range-checked) plus the uniform byte offset soff
__device__ __forceinline__ bf16 rsrc_load(__amdgpu_buffer_rsrc_t r, unsigned boff, int soff, const float*) {
  return (bf16)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, boff, soff, 0));
}
__device__ __forceinline__ bf16 rsrc_load(__amdgpu_buffer_rsrc_t r, unsigned boff, int soff, const bf16*) {
  return __builtin_bit_cast(bf16, __builtin_amdgcn_raw_buffer_load_b16(r, boff, soff, 0));
}
__device__ __forceinline__ f16 rsrc_load(__amdgpu_buffer_rsrc_t r, unsigned boff, int soff, const f16*) {
  return __builtin_bit_cast(f16, __builtin_amdgcn_raw_buffer_load_b16(r, boff, soff, 0));
}

// a + b after v_permlane32_swap (lanes 32..63 of a traded with lanes 0..31 of b) or
// v_permlane16_swap (odd 16-lane rows of a traded with even rows of b).  The ROCm 7.2
// compiler reads the builtin's second result from the first result's register (a + a)
// unless both results pass through an empty asm, which pins them to their own registers.
__device__ __forceinline__ float swap32_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  unsigned u = r[0], v = r[1];
  asm volatile("" : "+v"(u), "+v"(v));
  return __builtin_bit_cast(float, u) + __builtin_bit_cast(float, v);
}
__device__ __forceinline__ float swap16_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  unsigned u = r[0], v = r[1];
  asm volatile("" : "+v"(u), "+v"(v));
  return __builtin_bit_cast(float, u) + __builtin_bit_cast(float, v);
}
// a.lo b.lo + a.hi b.hi of two BF16 pairs in FP32: the first dot of a chain, in the
// three-operand form with a zero accumulator (for a zero the compiler emits the
// accumulate-in-place form after a register initialisation)
__device__ __forceinline__ float dot2_bf16_first(uint32_t a, uint32_t b) {
  float d;
  asm("v_dot2_f32_bf16 %0, %1, %2, 0" : "=v"(d) : "v"(a), "v"(b));
  return d;
}
// low halves of a and b (a in the low half of the result); high halves likewise
__device__ __forceinline__ uint32_t pack_lo(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x07060302u); }
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

}  // namespace

// Two packings of the weights (FP16), then the bias accumulators (FP32), then the packing
// of the K writer (FP16) and its bias accumulators (FP32).
constexpr size_t fused_apack_bytes(int C) { return (size_t)2 * 2 * nkc(C) * 64 * 8 * sizeof(f16); }
constexpr size_t kBpackBytes = (size_t)2 * 2 * 64 * 4 * sizeof(float);
constexpr size_t fused_kw_offset(int C) { return 2 * fused_apack_bytes(C) + kBpackBytes; }
size_t fused_weights_bytes(int C) { return fused_kw_offset(C) + fused_apack_bytes(C) + (size_t)2 * 64 * sizeof(float); }
constexpr int kBiasGroup = 1;               // lane group whose spare slots carry the bias (slot form)
constexpr uint32_t kOnesF16 = 0x3c003c00u;  // two FP16 ones

// apack[form][par][h][kc][lane][8] (FP16): A[m][k] = W[tap of D row 16h + m][channel kc][tap
// of slot k] for z parity par, zero for spare slots.  Form 0 leaves the bias out: the
// kernel starts the accumulation from bpack[par][h][lane][4] (FP32), the bias of the tap of
// D row 16h + 4 (lane / 16) + i, the element i that lane holds of the result (exact, but
// 16 registers).  Form 1 puts the bias on the spare slots (1, 6), (1, 7) of channel 0,
// whose image values the kernel sets to 1, as an FP16 value plus its FP16 remainder (about
// 22 significant bits), for kernels short of registers.  The K writer's packing
// kw[par][hh][kc][lane][8] is the B operand W^T of the swapped product: element j of lane
// l is W[tap 16 hh + l % 16][channel kc][tap of slot (l / 16, j)], its bias accumulators
// kb[hh][lane] = b[16 hh + lane % 16] (taps 27..31 zero).
__global__ __launch_bounds__(kThreads) void fused_pack_kernel(const float* __restrict__ W,
                                                              const float* __restrict__ b, int C,
                                                              int fold, f16* __restrict__ apack) {
  // weight of tap p, packed channel kc, input tap q (fold: C = 2 packed channels from three)
  auto wt = [&](int p, int kc, int q) {
    return fold ? W[((long)p * 3 + kc) * 27 + q] + W[((long)p * 3 + 2) * 27 + q] : W[((long)p * C + kc) * 27 + q];
  };
  const int NK = nkc(C);
  const int total = 2 * 2 * NK * 64 * 8;
  // one element per thread over the grid: the two packings, then the bias accumulators
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 2 * total) {
    const int j = e % 8, lane = (e / 8) % 64, kc = (e / 512) % NK, h = (e / (512 * NK)) % 2;
    const int par = (e / (1024 * NK)) % 2, form = e / total;
    const int rho = 16 * h + lane % 16, gk = lane / 16;
    const bool rok = fused_slot_valid(par, row_g(rho), row_j(rho));
    const int p = fused_tap(par, row_g(rho), row_j(rho)), q = fused_tap(par, gk, j);
    float v = rok && fused_slot_valid(par, gk, j) ? wt(p, kc, q) : 0.f;
    if (form == 1 && rok && kc == 0 && gk == kBiasGroup && j >= 6) {
      const float hi = (float)(f16)b[p];
      v = j == 6 ? hi : b[p] - hi;
    }
    apack[e] = (f16)v;
  } else if (e < 2 * total + 2 * 2 * 64 * 4) {
    float* bpack = (float*)((char*)apack + 2 * fused_apack_bytes(C));
    const int eb = e - 2 * total;
    const int i = eb % 4, lane = (eb / 4) % 64, h = (eb / 256) % 2, par = eb / 512;
    const int rho = 16 * h + 4 * (lane / 16) + i;
    const bool rok = fused_slot_valid(par, row_g(rho), row_j(rho));
    bpack[eb] = rok ? b[fused_tap(par, row_g(rho), row_j(rho))] : 0.f;
  } else if (e < 2 * total + 2 * 2 * 64 * 4 + total) {
    f16* kw = (f16*)((char*)apack + fused_kw_offset(C));
    const int ek = e - 2 * total - 2 * 2 * 64 * 4;
    const int j = ek % 8, lane = (ek / 8) % 64, kc = (ek / 512) % NK, hh = (ek / (512 * NK)) % 2;
    const int par = ek / (1024 * NK);
    const int p = 16 * hh + lane % 16, gk = lane / 16;
    const bool ok = p < 27 && fused_slot_valid(par, gk, j);
    kw[ek] = (f16)(ok ? wt(p, kc, fused_tap(par, gk, j)) : 0.f);
  } else if (e < 2 * total + 2 * 2 * 64 * 4 + total + 2 * 64) {
    float* kb = (float*)((char*)apack + fused_kw_offset(C) + fused_apack_bytes(C));
    const int eb = e - 2 * total - 2 * 2 * 64 * 4 - total;
    const int p = 16 * (eb / 64) + (eb % 64) % 16;
    kb[eb] = p < 27 ? b[p] : 0.f;
  }
}

// The apply: after the first product, lane (g, n) (g = l / 16, n = l % 16) holds D = K for
// the taps of D rows 4g + i and 16 + 4g + i (i < 4) of voxel n of the group, i.e. of the
// lane's own eight K slots (g, 0..7).  The x values of those slots are the same LDS words
// as the image patch, read from the x tile, so each lane forms its share of y(n) with four
// BF16 dot2 instructions; the four lane groups' shares of the four groups of a 64-voxel
// chunk are then summed by two cross-lane swaps (32-lane, then 16-lane), which leave y of
// voxel n of group t = g in lane (g, n): z0 + 32 (g / 2) + 2n + g % 2.
//
// EPI selects the epilogue at compile time (kEpiRuntime: every operand tested at run time):
// each present operand costs its per-chunk load and its VALU work, so the hot launches of
// a solve (pre blocks: none; post blocks: prev, and on level 0 the output mask and the
// solver's dot products) run without the others' dead loads, selects and FP64 FMAs.
constexpr int kEpiPrev = 1, kEpiMaskA = 2, kEpiMaskB = 4, kEpiDot0 = 8, kEpiDot1 = 16;
constexpr int kEpiRuntime = -1;
// Occupancy target: the 8-plane LDS tiles of one channel allow 5 workgroups (20 waves) per
// CU; one-channel kernels without the FP64 dot epilogue fit that in 96 VGPRs (the unrolled
// loop would otherwise hoist LDS reads into more registers), with it in 128.  Three-channel
// kernels (4-plane tiles, 4 workgroups per CU) fit 128 VGPRs except with two solver dots.
// 2-plane tiles (small levels) and the others keep what they need (a tighter target made
// them spill).
__host__ __device__ constexpr int fused_waves(int C, int FI, int EPI) {
  return EPI < 0 || FI < 4 ? 1
         : C == 1 ? ((EPI & (kEpiDot0 | kEpiDot1)) ? 4 : 5)
         : C == 3 ? ((EPI & kEpiDot1) ? 1 : 4)
                  : 1;
}
template <typename TX, typename TY, typename TP, int C, int FI, int EPI, bool PL = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(fused_waves(C, FI, EPI), 8)))
void sm_fused_kernel(FusedArgs a) {
  constexpr int kTile = tile_cells(FI);
  constexpr int kRows = (FI + 2) * kPj / 4;  // halo-tile rows staged per wave
  typedef Mfma<f16>::v8 h8;
  constexpr int NK = nkc(C);
  __shared__ __attribute__((aligned(16))) bf16 xs[kTile];
  __shared__ __attribute__((aligned(16))) f16 is[C * kTile];
  if (solve_done(a.done)) return;
  // the wave index is uniform: keep it (and everything derived from it) in scalar registers
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, nn = lane & 15;
  const int tz = a.n3 / kFz, tj = a.n2 / kFj;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int z0 = (int)(t % tz) * kFz;
  const int j0 = (int)((t / tz) % tj) * kFj;
  const int i0 = (int)(t / tz / tj) * FI;
  const long V = (long)a.n1 * a.n2 * a.n3;

  // per-lane constant operands (global loads issued before the tile staging)
  // the bias as the accumulator (kernels without the FP64 dot epilogue) or on spare slots
  constexpr bool BA = EPI >= 0 && !(EPI & (kEpiDot0 | kEpiDot1));
  const f16* apack = (const f16*)((const char*)a.apack + (BA ? 0 : fused_apack_bytes(C)));
  h8 A[2][2][NK];  // [z parity][row half][channel]
#pragma unroll
  for (int par = 0; par < 2; ++par)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int kc = 0; kc < NK; ++kc)
        A[par][h][kc] = *(const h8*)(apack + ((((long)par * 2 + h) * NK + kc) * 64 + lane) * 8);
  const float* bpack = (const float*)((const char*)a.apack + 2 * fused_apack_bytes(C));
  f32x4 bias[2][2];  // [z parity][row half]
#pragma unroll
  for (int par = 0; par < 2; ++par)
#pragma unroll
    for (int h = 0; h < 2; ++h)
      bias[par][h] = BA ? *(const f32x4*)(bpack + ((par * 2 + h) * 64 + lane) * 4) : f32x4{0.f, 0.f, 0.f, 0.f};
  const bool bias_lane = !BA && g == kBiasGroup;  // its split-column word of channel 0 carries (1, 1)
  const float scale = a.in_scale ? (float)(*a.in_scale) : 1.f;
  const float c0 = a.prev ? *(const float*)a.c0 : 0.f;
  const float c1 = a.prev ? *(const float*)a.c1 : 0.f;

  // stage x and img of the halo tile (BF16 x, FP16 img); zero outside the grid.  Plane pi
  // of the tile (grid plane gi = i0 + pi - 1) is read through a buffer resource over that
  // grid plane, a row of it through the scalar offset and a cell through the lane offset,
  // which is out of range for cells outside [0, n3): the range check turns every halo cell
  // into a zero, with no clamps or selects.  Planes and rows outside the grid get an empty
  // resource.  Wave w stages rows r = w + 4 k (k < kRows) of the (FI + 2) x 10 rows,
  // element e = lane of each row and e = 64 + lane for lanes 0, 1 (interleaving the waves
  // over the planes fetched 8 % fewer bytes from HBM than giving each wave a block of
  // planes).  The staging code is instantiated once per wave index (a uniform branch), so
  // the row geometry is folded at compile time and a row costs a few scalar instructions.
  const TX* __restrict__ x = (const TX*)a.x;
  const f16* __restrict__ img = (const f16*)a.img;
  static_assert((FI + 2) * kPj % 4 == 0 && kRowCells == 66, "staging assumes 4 waves x kRows rows of 66 cells");
  {
    constexpr unsigned kOut = 0x80000000u;  // beyond any plane (planes are under 2^31 bytes)
    unsigned zx[2], zi[2];                  // lane byte offsets of its cells in a row
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int z = z0 - 1 + (u == 0 ? lane : 64 + (lane & 1));
      const bool zok = z >= 0 && z < a.n3;
      zx[u] = zok ? (unsigned)z * (unsigned)sizeof(TX) : kOut;
      zi[u] = zok ? (unsigned)z * 2u : kOut;
    }
    const long plane = (long)a.n2 * a.n3;
    const int xpb = (int)(plane * (long)sizeof(TX)), ipb = (int)(plane * 2);  // plane bytes
    const int xrb = a.n3 * (int)sizeof(TX), irb = a.n3 * 2;                   // row bytes
    const bool lo_i = i0 > 0, hi_i = i0 + FI < a.n1, lo_j = j0 > 0, hi_j = j0 + kFj < a.n2;
    const long p0 = (long)(i0 - 1) * plane;  // grid plane of tile plane 0 (-1 on the first tile)
    auto stage = [&](auto wc) {
      constexpr int W = decltype(wc)::value;
#pragma unroll
      for (int k0 = 0; k0 < kRows; k0 += 13) {
        constexpr int KB = 13;
        bf16 xv[KB][2];
        f16 iv[KB][C][2];
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int k = k0 + kk;
          if (k < kRows) {
            const int r = W + 4 * k, pi = r / kPj, pj = r % kPj;
            // boundary planes and rows are tested at run time, the others are in the grid
            const bool edge_p = pi == 0 || pi == FI + 1, edge_j = pj == 0 || pj == kPj - 1;
            const bool pok = !edge_p || (pi == 0 ? lo_i : hi_i);
            const bool rok = pok && (!edge_j || (pj == 0 ? lo_j : hi_j));
            const long po = edge_p ? (pok ? (long)pi * plane : plane) : (long)pi * plane;
            const int soff = (j0 + pj - 1);
            const auto xr = row_rsrc(x + (p0 + po), rok ? xpb : 0);
#pragma unroll
            for (int u = 0; u < 2; ++u) xv[kk][u] = rsrc_load(xr, zx[u], soff * xrb, (const TX*)nullptr);
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const auto ir = row_rsrc(img + (c * V + p0 + po), rok ? ipb : 0);
#pragma unroll
              for (int u = 0; u < 2; ++u) iv[kk][c][u] = rsrc_load(ir, zi[u], soff * irb, (const f16*)nullptr);
            }
          }
        }
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int r = W + 4 * (k0 + kk);
          if (k0 + kk < kRows) {
            xs[r * kPz + lane] = xv[kk][0];
#pragma unroll
            for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + lane] = iv[kk][c][0];
          }
        }
        if (lane < 2) {
#pragma unroll
          for (int kk = 0; kk < KB; ++kk) {
            const int r = W + 4 * (k0 + kk);
            if (k0 + kk < kRows) {
              xs[r * kPz + 64 + lane] = xv[kk][1];
#pragma unroll
              for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + 64 + lane] = iv[kk][c][1];
            }
          }
        }
      }
    };
    switch (w) {
      case 0: stage(std::integral_constant<int, 0>()); break;
      case 1: stage(std::integral_constant<int, 1>()); break;
      case 2: stage(std::integral_constant<int, 2>()); break;
      default: stage(std::integral_constant<int, 3>()); break;
    }
  }

  // per-lane LDS word offsets (even element indices relative to the chunk's cell cb + 32h):
  // the words of this lane's voxel 2nn (even group; the odd group reads the same words),
  // at the same offsets in the image and in the x tile
  const int o1 = col_off(fused_col(2 * g)) - 1, o2 = col_off(fused_col(2 * g + 1)) - 1;
  const int o8 = col_off(kSplitCol) - 1;
  // split-column word of the even / odd groups (groups 1, 3 read those of groups 0, 2)
  const int r3e = g >= 2 ? 2 : 0, r3o = g >= 2 ? 0 : 2;
  const int iP = o1 + 2 * nn, iQ = o2 + 2 * nn, i8e = o8 + 2 * nn + r3e, i8o = o8 + 2 * nn + r3o;
  // One byte-address register per word (with the wave's row offset folded in): the chunk
  // loop is unrolled, so each read is its own register plus a compile-time offset (the
  // tile array and the chunk).  With distinct registers no two reads can be merged into a
  // ds_read2, whose 8-bit offsets would need a fresh base (a VALU add) per chunk, and
  // whose register pairs would need moves into the operand tuples.
  const int wrow = w * kPz;
  auto lds_base = [&](int elem) {
    unsigned b = (unsigned)(elem + wrow) * 2u;
    asm volatile("" : "+v"(b));
    return b;
  };
  const unsigned bI[6] = {lds_base(iP), lds_base(iP + 2), lds_base(iQ), lds_base(iQ + 2), lds_base(i8e),
                          lds_base(i8o)};
  // after the cross-lane reduction, lane (g, nn) holds voxel nn of group t = g
  const int myz = 32 * (g >> 1) + 2 * nn + (g & 1);
  __syncthreads();

  // epilogue operands through always-valid pointers (x has at least 2V bytes), so the
  // loads are unconditional and can be issued one chunk ahead.  Row pointers are uniform;
  // the lane's element offset myz is the same for every chunk.
  TY* __restrict__ y = (TY*)a.y;
  // LX: the operand is loaded (a compile-time epilogue has it, or the run-time one may);
  // has_x: it is applied
  constexpr bool RT = EPI == kEpiRuntime;
  constexpr bool LA = RT || (EPI & kEpiMaskA), LB = RT || (EPI & kEpiMaskB), LP = RT || (EPI & kEpiPrev);
  constexpr bool LD0 = RT || (EPI & kEpiDot0), LD1 = RT || (EPI & kEpiDot1);
  const uint8_t* mA = a.mask_a ? a.mask_a : (const uint8_t*)a.x;
  const uint8_t* mB = a.mask_b ? a.mask_b : (const uint8_t*)a.x;
  const TP* pP = a.prev ? (const TP*)a.prev : (const TP*)a.x;
  const bool has_a = RT ? a.mask_a != nullptr : LA, has_b = RT ? a.mask_b != nullptr : LB;
  const bool has_p = RT ? a.prev != nullptr : LP;
  const bool dots = RT ? a.partials != nullptr : LD0;
  const float* d0p = a.dot0 ? a.dot0 : (const float*)a.x;
  const float* d1p = a.dot1 ? a.dot1 : (const float*)a.x;
  // without dot1 every lane reads element 0 (a compile-time epilogue with dot1 has it)
  const bool has_d1 = RT ? a.dot1 != nullptr : LD1;
  const long rstep1 = has_d1 ? 1 : 0;
  const unsigned mob = (unsigned)myz, mop = mob * (unsigned)sizeof(TP), mof = mob * 4u;
  const unsigned moy = mob * (unsigned)sizeof(TY), d1off = has_d1 ? mof : 0u;
  double red[2] = {0.0, 0.0};
  auto row = [&](int q) {
    const int ii = q >> 1, jj = w + 4 * (q & 1);
    return ((long)(i0 + ii) * a.n2 + (j0 + jj)) * a.n3 + z0;
  };
  // the LDS words one half-chunk (h) of chunk cell cb needs: per image channel and for x,
  // p0, p2, q0, q2 and the split-column words of the even and odd groups
  struct Words {
    uint32_t im[NK][6];
    uint32_t xw[6];
  };
  // chunk q's cell cb, less the wave's row offset (a compile-time constant in the unrolled loop)
  auto cb_of = [&](int q) { return (((q >> 1) + 1) * kPj + (4 * (q & 1) + 1)) * kPz + 1; };
  auto load_words = [&](int cbq, int h) {
    Words r;
    const int cbyte = 2 * (cbq + 32 * h);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
      const char* ib = (const char*)(is + kc * kTile) + cbyte;
#pragma unroll
      for (int k = 0; k < 6; ++k) r.im[kc][k] = *(const uint32_t*)(ib + bI[k]);
    }
    const char* xb = (const char*)xs + cbyte;
#pragma unroll
    for (int k = 0; k < 6; ++k) r.xw[k] = *(const uint32_t*)(xb + bI[k]);
    return r;
  };
  Words cur = load_words(cb_of(0), 0);
  long vr = row(0);
  float pp[PL ? FI * 2 : 1];  // PL: the output's z-pair sums of each chunk
  uint8_t ka = LA ? lane_at(mA + vr, mob) : uint8_t(1), kb = LB ? lane_at(mB + vr, mob) : uint8_t(1);
  TP pv = LP ? lane_at(pP + vr, mop) : TP(0.f);
  float da = (LD0 && dots) ? lane_at(d0p + vr, mof) : 0.f;
  float db = (LD1 && dots) ? lane_at(d1p + vr * rstep1, d1off) : 0.f;
#pragma unroll
  for (int q = 0; q < FI * 2; ++q) {
    // next chunk's epilogue operands
    const long vn = row(q + 1 < FI * 2 ? q + 1 : q);
    const uint8_t ka_n = LA ? lane_at(mA + vn, mob) : uint8_t(1);
    const uint8_t kb_n = LB ? lane_at(mB + vn, mob) : uint8_t(1);
    const TP pv_n = LP ? lane_at(pP + vn, mop) : TP(0.f);
    const float da_n = (LD0 && dots) ? lane_at(d0p + vn, mof) : 0.f;
    const float db_n = (LD1 && dots) ? lane_at(d1p + vn * rstep1, d1off) : 0.f;
    float part[4];  // this lane's share of y(voxel nn of group t): its 8 K slots
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // LDS words of the next half-chunk ((q, 1) or (q + 1, 0)) are read before this
      // half's products, so their latency overlaps the matrix-core work
      const int qn = h == 0 ? q : (q + 1 < FI * 2 ? q + 1 : q);
      const Words nxt = load_words(cb_of(qn), h ^ 1);
      // image words shared by the even (t = 2h) and odd (t = 2h + 1) groups of this half
      h8 bv[2][NK];
#pragma unroll
      for (int kc = 0; kc < NK; ++kc) {
        const uint32_t* m = cur.im[kc];
        const uint32_t e8 = (kc == 0 && bias_lane) ? kOnesF16 : m[4];
        const uint32_t d8 = (kc == 0 && bias_lane) ? kOnesF16 : m[5];
        bv[0][kc] = __builtin_bit_cast(h8, u32x4{m[0], pack_lo(m[1], m[3]), m[2], e8});
        bv[1][kc] = __builtin_bit_cast(h8, u32x4{m[1], pack_hi(m[0], m[2]), m[3], d8});
      }
      const uint32_t* xw = cur.xw;
      const uint32_t xs_[2][4] = {{xw[0], pack_lo(xw[1], xw[3]), xw[2], xw[4]},
                                  {xw[1], pack_hi(xw[0], xw[2]), xw[3], xw[5]}};
#pragma unroll
      for (int par = 0; par < 2; ++par) {
        // K of this lane's 8 slots (D rows 4g + i and 16 + 4g + i of voxel nn); the biases
        // enter as the accumulator or through the ones slots
        f32x4 d0 = bias[par][0], d1 = bias[par][1];
#pragma unroll
        for (int kc = 0; kc < NK; ++kc) {
          d0 = Mfma<f16>::run(A[par][0][kc], bv[par][kc], d0);
          d1 = Mfma<f16>::run(A[par][1][kc], bv[par][kc], d1);
        }
        // slot pairs (0,1) (2,3) of the first half of D, (4,5) (6,7) of the second, in BF16
        // against the x words of the same slots
        const float kf[8] = {d0[0], d0[1], d0[2], d0[3], d1[0], d1[1], d1[2], d1[3]};
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const bf16x2 kp = {(bf16)kf[2 * k], (bf16)kf[2 * k + 1]};
          acc = k == 0 ? dot2_bf16_first(xs_[par][k], __builtin_bit_cast(uint32_t, kp))
                       : __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, xs_[par][k]), kp, acc, false);
        }
        part[2 * h + par] = acc;
      }
      cur = nxt;
    }
    // sum the four lane groups' shares: afterwards lane (g, nn) holds y of voxel nn of group
    // t = g.  The 32-lane swap pairs groups g and g + 2 (t = 0, 1 kept in the low half,
    // t = 2, 3 in the high half), the 16-lane swap then pairs groups g and g ^ 1.
    const float ymine = swap16_sum(swap32_sum(part[0], part[2]), swap32_sum(part[1], part[3]));
    float val = ymine * scale;
    if (has_a && ka == 0) val = 0.f;
    if (has_p) val = c0 * (float)pv + c1 * val;
    if (has_b && kb == 0) val = 0.f;
    const TY yv = (TY)val;
    lane_store(y + vr, moy, yv);
    if constexpr (PL) {
      // lanes l and l ^ 16 hold the even and odd voxel of a z pair
      const float yr = (float)yv;
      pp[q] = yr + __shfl_xor(yr, 16);
    }
    if (dots) {
      red[0] += (double)(float)yv * (double)da;
      if (LD1) red[1] += (double)(float)yv * (double)db;
    }
    vr = vn;
    ka = ka_n;
    kb = kb_n;
    pv = pv_n;
    da = da_n;
    db = db_n;
  }
  if (dots) block_partials<2>(red, a.partials);
  if constexpr (PL) {
    // the 2x2x2 means of the stored output (the next level's input), summed in the order of
    // avgpool2_kernel: z pair, then (plane 2I, rows 2J, 2J+1), then plane 2I+1's rows.  Rows
    // 2J and 2J+1 belong to different waves, so the pair sums meet in LDS (the tiles' space).
    __syncthreads();
    float* P = (float*)xs;  // [FI][8 rows][32 coarse z]
    static_assert(FI * 8 * 32 * 4 <= kTile * 2, "pair sums fit the x tile");
    if (!(g & 1)) {
#pragma unroll
      for (int q = 0; q < FI * 2; ++q) P[((q >> 1) * 8 + w + 4 * (q & 1)) * 32 + 16 * (g >> 1) + nn] = pp[q];
    }
    __syncthreads();
    TY* __restrict__ out = (TY*)a.pool_out;
    const int c2 = a.n2 >> 1, c3 = a.n3 >> 1;
    for (int e = threadIdx.x; e < (FI / 2) * 4 * 32; e += kThreads) {
      const int kc = e & 31, J = (e >> 5) & 3, I = e >> 7;
      const float* p0 = P + (2 * I * 8 + 2 * J) * 32 + kc;
      float sum = p0[0];
      sum += p0[32];
      sum += p0[8 * 32];
      sum += p0[8 * 32 + 32];
      out[((long)(i0 / 2 + I) * c2 + (j0 / 2 + J)) * c3 + z0 / 2 + kc] = (TY)(sum * 0.125f);
    }
  }
}

// K writer: the stencils K_p(v) of the fused apply's tile, stored in the [27][V] BF16 layout
// of the stored-stencil kernels (the input gradient of a BF16 block needs them).  The first
// product is taken with its operands swapped, D^T = P^T W^T (same registers: the image
// words as A, the packing kw as B), so lane (g, n) holds tap 16 hh + n of voxels 4 g + i of
// a z-parity group; the even and odd groups together give it the 8 consecutive cells
// 32 h + 8 g .. + 7 of one tap, one 16-byte store.  The image tile is staged as in the apply.
template <int C, int FI>
__global__ __launch_bounds__(kThreads) void fused_kwrite_kernel(FusedArgs a, bf16* __restrict__ K) {
  constexpr int kTile = tile_cells(FI);
  constexpr int kRows = (FI + 2) * kPj / 4;
  typedef Mfma<f16>::v8 h8;
  constexpr int NK = nkc(C);
  __shared__ __attribute__((aligned(16))) f16 is[C * kTile];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, nn = lane & 15;
  const int tz = a.n3 / kFz, tj = a.n2 / kFj;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int z0 = (int)(t % tz) * kFz;
  const int j0 = (int)((t / tz) % tj) * kFj;
  const int i0 = (int)(t / tz / tj) * FI;
  const long V = (long)a.n1 * a.n2 * a.n3;

  const f16* kw = (const f16*)((const char*)a.apack + fused_kw_offset(C));
  h8 Wt[2][2][NK];  // [z parity][tap half hh][channel]
#pragma unroll
  for (int par = 0; par < 2; ++par)
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int kc = 0; kc < NK; ++kc)
        Wt[par][hh][kc] = *(const h8*)(kw + ((((long)par * 2 + hh) * NK + kc) * 64 + lane) * 8);
  const float* kbias = (const float*)((const char*)kw + fused_apack_bytes(C));
  const float b0 = kbias[lane], b1 = kbias[64 + lane];

  // image tile (FP16), zero outside the grid: the staging of the apply without x
  const f16* __restrict__ img = (const f16*)a.img;
  {
    constexpr unsigned kOut = 0x80000000u;
    unsigned zi[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int z = z0 - 1 + (u == 0 ? lane : 64 + (lane & 1));
      zi[u] = (z >= 0 && z < a.n3) ? (unsigned)z * 2u : kOut;
    }
    const long plane = (long)a.n2 * a.n3;
    const int ipb = (int)(plane * 2), irb = a.n3 * 2;
    const bool lo_i = i0 > 0, hi_i = i0 + FI < a.n1, lo_j = j0 > 0, hi_j = j0 + kFj < a.n2;
    const long p0 = (long)(i0 - 1) * plane;
    auto stage = [&](auto wc) {
      constexpr int W = decltype(wc)::value;
#pragma unroll
      for (int k0 = 0; k0 < kRows; k0 += 13) {
        constexpr int KB = 13;
        f16 iv[KB][C][2];
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int k = k0 + kk;
          if (k < kRows) {
            const int r = W + 4 * k, pi = r / kPj, pj = r % kPj;
            const bool edge_p = pi == 0 || pi == FI + 1, edge_j = pj == 0 || pj == kPj - 1;
            const bool pok = !edge_p || (pi == 0 ? lo_i : hi_i);
            const bool rok = pok && (!edge_j || (pj == 0 ? lo_j : hi_j));
            const long po = edge_p ? (pok ? (long)pi * plane : plane) : (long)pi * plane;
            const int soff = (j0 + pj - 1);
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const auto ir = row_rsrc(img + (c * V + p0 + po), rok ? ipb : 0);
#pragma unroll
              for (int u = 0; u < 2; ++u) iv[kk][c][u] = rsrc_load(ir, zi[u], soff * irb, (const f16*)nullptr);
            }
          }
        }
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int r = W + 4 * (k0 + kk);
          if (k0 + kk < kRows) {
#pragma unroll
            for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + lane] = iv[kk][c][0];
          }
        }
        if (lane < 2) {
#pragma unroll
          for (int kk = 0; kk < KB; ++kk) {
            const int r = W + 4 * (k0 + kk);
            if (k0 + kk < kRows) {
#pragma unroll
              for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + 64 + lane] = iv[kk][c][1];
            }
          }
        }
      }
    };
    switch (w) {
      case 0: stage(std::integral_constant<int, 0>()); break;
      case 1: stage(std::integral_constant<int, 1>()); break;
      case 2: stage(std::integral_constant<int, 2>()); break;
      default: stage(std::integral_constant<int, 3>()); break;
    }
  }

  // the image words of the lane's slots, addressed as in the apply
  const int o1 = col_off(fused_col(2 * g)) - 1, o2 = col_off(fused_col(2 * g + 1)) - 1;
  const int o8 = col_off(kSplitCol) - 1;
  const int r3e = g >= 2 ? 2 : 0, r3o = g >= 2 ? 0 : 2;
  const int iP = o1 + 2 * nn, iQ = o2 + 2 * nn, i8e = o8 + 2 * nn + r3e, i8o = o8 + 2 * nn + r3o;
  const int wrow = w * kPz;
  auto lds_base = [&](int elem) {
    unsigned b = (unsigned)(elem + wrow) * 2u;
    asm volatile("" : "+v"(b));
    return b;
  };
  const unsigned bI[6] = {lds_base(iP), lds_base(iP + 2), lds_base(iQ), lds_base(iQ + 2), lds_base(i8e),
                          lds_base(i8o)};
  __syncthreads();

  auto row = [&](int q) {
    const int ii = q >> 1, jj = w + 4 * (q & 1);
    return ((long)(i0 + ii) * a.n2 + (j0 + jj)) * a.n3 + z0;
  };
  struct Words {
    uint32_t im[NK][6];
  };
  auto cb_of = [&](int q) { return (((q >> 1) + 1) * kPj + (4 * (q & 1) + 1)) * kPz + 1; };
  auto load_words = [&](int cbq, int h) {
    Words r;
    const int cbyte = 2 * (cbq + 32 * h);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
      const char* ib = (const char*)(is + kc * kTile) + cbyte;
#pragma unroll
      for (int k = 0; k < 6; ++k) r.im[kc][k] = *(const uint32_t*)(ib + bI[k]);
    }
    return r;
  };
  Words cur = load_words(cb_of(0), 0);
#pragma unroll
  for (int q = 0; q < FI * 2; ++q) {
    const long vr = row(q);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qn = h == 0 ? q : (q + 1 < FI * 2 ? q + 1 : q);
      const Words nxt = load_words(cb_of(qn), h ^ 1);
      h8 bv[2][NK];
#pragma unroll
      for (int kc = 0; kc < NK; ++kc) {
        const uint32_t* m = cur.im[kc];
        bv[0][kc] = __builtin_bit_cast(h8, u32x4{m[0], pack_lo(m[1], m[3]), m[2], m[4]});
        bv[1][kc] = __builtin_bit_cast(h8, u32x4{m[1], pack_hi(m[0], m[2]), m[3], m[5]});
      }
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const float bb = hh == 0 ? b0 : b1;
        f32x4 d[2];
#pragma unroll
        for (int par = 0; par < 2; ++par) {
          d[par] = f32x4{bb, bb, bb, bb};
#pragma unroll
          for (int kc = 0; kc < NK; ++kc) d[par] = Mfma<f16>::run(bv[par][kc], Wt[par][hh][kc], d[par]);
        }
        // cells 32 h + 8 g + 2 i (even group) and + 1 (odd group) of tap 16 hh + nn
        u32x4 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const bf16x2 pr = {(bf16)d[0][i], (bf16)d[1][i]};
          o[i] = __builtin_bit_cast(uint32_t, pr);
        }
        const int p = 16 * hh + nn;
        // plain stores: the halves h = 0, 1 of a tap's 128-byte row segment meet in L2
        if (p < 27) *(u32x4*)(K + p * V + vr + 32 * h + 8 * g) = o;
      }
      cur = nxt;
    }
  }
}

// Input gradient of a BF16 block through column sums (the S path).  With P_p(v) = g(v) K_p(v)
// the input gradient is dx(u) = sum_p P_p(u - d_p).  Grouping the 27 taps by their (d1, d2)
// column c = 3 (d1 + 1) + (d2 + 1) and c' = 8 - c,
//   S_c(v) = P_{3c'+2}(v - e3) + P_{3c'+1}(v) + P_{3c'}(v + e3)     (a sum along the z-row of v)
//   dx(u)  = sum_c S_c(u + (c / 3 - 1, c % 3 - 1, 0)).
// fused_swrite_kernel computes K on the matrix cores as the K writer does, multiplies by g,
// forms the nine row sums in LDS and stores S (BF16, 9 x V) -- 18 bytes per voxel instead of
// the 54 of K; the gather then reads 18 bytes per voxel instead of 56, and needs no z
// neighbours.  The two z-end cells of a tile row miss one term from the neighbouring tile:
// their partial sums and the terms the neighbours miss go to a side array (FP32, 36 values
// per tile row) that the gather combines.  S is rounded once to BF16 (the stored-K path
// rounds K instead); the sums are FP32.
constexpr int kPsPitch = 68;  // floats per tap row of a wave's P buffer (64 cells + pad)

template <int C, int FI>
__global__ __launch_bounds__(kThreads) void fused_swrite_kernel(FusedArgs a, bf16* __restrict__ S,
                                                               float* __restrict__ E) {
  constexpr int kTile = tile_cells(FI);
  constexpr int kRows = (FI + 2) * kPj / 4;
  typedef Mfma<f16>::v8 h8;
  constexpr int NK = nkc(C);
  __shared__ __attribute__((aligned(16))) f16 is[C * kTile];
  __shared__ __attribute__((aligned(16))) float ps[4][27 * kPsPitch];
  __shared__ __attribute__((aligned(16))) bf16 gs[4][FI * 2][kFz];  // g of each wave's rows
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, nn = lane & 15;
  const int tz = a.n3 / kFz, tj = a.n2 / kFj;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int zt = (int)(t % tz);
  const int z0 = zt * kFz;
  const int j0 = (int)((t / tz) % tj) * kFj;
  const int i0 = (int)(t / tz / tj) * FI;
  const long V = (long)a.n1 * a.n2 * a.n3;
  const bf16* __restrict__ gsrc = (const bf16*)a.x;

  const f16* kw = (const f16*)((const char*)a.apack + fused_kw_offset(C));
  h8 Wt[2][2][NK];  // [z parity][tap half hh][channel]
#pragma unroll
  for (int par = 0; par < 2; ++par)
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int kc = 0; kc < NK; ++kc)
        Wt[par][hh][kc] = *(const h8*)(kw + ((((long)par * 2 + hh) * NK + kc) * 64 + lane) * 8);
  const float* kbias = (const float*)((const char*)kw + fused_apack_bytes(C));
  const float b0 = kbias[lane], b1 = kbias[64 + lane];

  auto row = [&](int q) {
    const int ii = q >> 1, jj = w + 4 * (q & 1);
    return ((long)(i0 + ii) * a.n2 + (j0 + jj)) * a.n3 + z0;
  };
  // g of the wave's 2 FI rows, staged with the image so the products never wait for HBM
  {
    bf16 gv[FI * 2];
#pragma unroll
    for (int q = 0; q < FI * 2; ++q) gv[q] = gsrc[row(q) + lane];
#pragma unroll
    for (int q = 0; q < FI * 2; ++q) gs[w][q][lane] = gv[q];
  }
  // image tile (FP16), zero outside the grid: the staging of the K writer
  const f16* __restrict__ img = (const f16*)a.img;
  {
    constexpr unsigned kOut = 0x80000000u;
    unsigned zi[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int z = z0 - 1 + (u == 0 ? lane : 64 + (lane & 1));
      zi[u] = (z >= 0 && z < a.n3) ? (unsigned)z * 2u : kOut;
    }
    const long plane = (long)a.n2 * a.n3;
    const int ipb = (int)(plane * 2), irb = a.n3 * 2;
    const bool lo_i = i0 > 0, hi_i = i0 + FI < a.n1, lo_j = j0 > 0, hi_j = j0 + kFj < a.n2;
    const long p0 = (long)(i0 - 1) * plane;
    auto stage = [&](auto wc) {
      constexpr int W = decltype(wc)::value;
#pragma unroll
      for (int k0 = 0; k0 < kRows; k0 += 13) {
        constexpr int KB = 13;
        f16 iv[KB][C][2];
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int k = k0 + kk;
          if (k < kRows) {
            const int r = W + 4 * k, pi = r / kPj, pj = r % kPj;
            const bool edge_p = pi == 0 || pi == FI + 1, edge_j = pj == 0 || pj == kPj - 1;
            const bool pok = !edge_p || (pi == 0 ? lo_i : hi_i);
            const bool rok = pok && (!edge_j || (pj == 0 ? lo_j : hi_j));
            const long po = edge_p ? (pok ? (long)pi * plane : plane) : (long)pi * plane;
            const int soff = (j0 + pj - 1);
#pragma unroll
            for (int c = 0; c < C; ++c) {
              const auto ir = row_rsrc(img + (c * V + p0 + po), rok ? ipb : 0);
#pragma unroll
              for (int u = 0; u < 2; ++u) iv[kk][c][u] = rsrc_load(ir, zi[u], soff * irb, (const f16*)nullptr);
            }
          }
        }
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
          const int r = W + 4 * (k0 + kk);
          if (k0 + kk < kRows) {
#pragma unroll
            for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + lane] = iv[kk][c][0];
          }
        }
        if (lane < 2) {
#pragma unroll
          for (int kk = 0; kk < KB; ++kk) {
            const int r = W + 4 * (k0 + kk);
            if (k0 + kk < kRows) {
#pragma unroll
              for (int c = 0; c < C; ++c) is[c * kTile + r * kPz + 64 + lane] = iv[kk][c][1];
            }
          }
        }
      }
    };
    switch (w) {
      case 0: stage(std::integral_constant<int, 0>()); break;
      case 1: stage(std::integral_constant<int, 1>()); break;
      case 2: stage(std::integral_constant<int, 2>()); break;
      default: stage(std::integral_constant<int, 3>()); break;
    }
  }

  const int o1 = col_off(fused_col(2 * g)) - 1, o2 = col_off(fused_col(2 * g + 1)) - 1;
  const int o8 = col_off(kSplitCol) - 1;
  const int r3e = g >= 2 ? 2 : 0, r3o = g >= 2 ? 0 : 2;
  const int iP = o1 + 2 * nn, iQ = o2 + 2 * nn, i8e = o8 + 2 * nn + r3e, i8o = o8 + 2 * nn + r3o;
  const int wrow = w * kPz;
  auto lds_base = [&](int elem) {
    unsigned b = (unsigned)(elem + wrow) * 2u;
    asm volatile("" : "+v"(b));
    return b;
  };
  const unsigned bI[6] = {lds_base(iP), lds_base(iP + 2), lds_base(iQ), lds_base(iQ + 2), lds_base(i8e),
                          lds_base(i8o)};
  __syncthreads();

  struct Words {
    uint32_t im[NK][6];
  };
  auto cb_of = [&](int q) { return (((q >> 1) + 1) * kPj + (4 * (q & 1) + 1)) * kPz + 1; };
  auto load_words = [&](int cbq, int h) {
    Words r;
    const int cbyte = 2 * (cbq + 32 * h);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
      const char* ib = (const char*)(is + kc * kTile) + cbyte;
#pragma unroll
      for (int k = 0; k < 6; ++k) r.im[kc][k] = *(const uint32_t*)(ib + bI[k]);
    }
    return r;
  };
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  float* __restrict__ pw = ps[w];
  // the row-sum reads of lane z: cells z - 1, z, z + 1 (clamped into the row; the terms of
  // cells outside the tile row are dropped below)
  const int zm = lane > 0 ? lane - 1 : 0, zp = lane < 63 ? lane + 1 : 63;
  const float keep_m = lane > 0 ? 1.f : 0.f, keep_p = lane < 63 ? 1.f : 0.f;
  Words cur = load_words(cb_of(0), 0);
#pragma unroll
  for (int q = 0; q < FI * 2; ++q) {
    const long vr = row(q);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qn = h == 0 ? q : (q + 1 < FI * 2 ? q + 1 : q);
      const Words nxt = load_words(cb_of(qn), h ^ 1);
      // g of the lane's eight cells 32 h + 8 g .. + 7 (word i: cells 2i, 2i + 1)
      const u32x4 gw = *(const u32x4*)&gs[w][q][32 * h + 8 * g];
      h8 bv[2][NK];
#pragma unroll
      for (int kc = 0; kc < NK; ++kc) {
        const uint32_t* m = cur.im[kc];
        bv[0][kc] = __builtin_bit_cast(h8, u32x4{m[0], pack_lo(m[1], m[3]), m[2], m[4]});
        bv[1][kc] = __builtin_bit_cast(h8, u32x4{m[1], pack_hi(m[0], m[2]), m[3], m[5]});
      }
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const float bb = hh == 0 ? b0 : b1;
        f32x4 d[2];
#pragma unroll
        for (int par = 0; par < 2; ++par) {
          d[par] = f32x4{bb, bb, bb, bb};
#pragma unroll
          for (int kc = 0; kc < NK; ++kc) d[par] = Mfma<f16>::run(bv[par][kc], Wt[par][hh][kc], d[par]);
        }
        // P of tap 16 hh + nn at cells 32 h + 8 g + 2 i + par
        float pv[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          pv[2 * i] = d[0][i] * __uint_as_float(gw[i] << 16);
          pv[2 * i + 1] = d[1][i] * __uint_as_float(gw[i] & 0xffff0000u);
        }
        const int p = 16 * hh + nn;
        if (p < 27) {
          float* dst = pw + p * kPsPitch + 32 * h + 8 * g;
          *(f32x4*)dst = f32x4{pv[0], pv[1], pv[2], pv[3]};
          *(f32x4*)(dst + 4) = f32x4{pv[4], pv[5], pv[6], pv[7]};
        }
      }
      cur = nxt;
    }
    wave_sync();
    // the nine row sums of cell z = lane; the end cells store partial sums and export the
    // terms the neighbouring tiles miss
    // side array of tile row rt: [end (0: cell z0, 1: cell z0 + 63)][c][2]: .x the partial
    // sum, .y the term from the neighbouring tile, written by that tile (zero at the grid's
    // z ends, written here)
    const long rt = vr / kFz;  // tile row index: (grid row) * tz + zt
    // all 27 reads before any use (no branch between them), then the stores
    float sv[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const float* r0 = pw + 3 * (8 - c) * kPsPitch;
      sv[c] = (keep_m * r0[2 * kPsPitch + zm] + r0[kPsPitch + lane]) + keep_p * r0[zp];
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) S[c * V + vr + lane] = (bf16)sv[c];
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const float* r0 = pw + 3 * (8 - c) * kPsPitch;
        E[rt * 36 + 2 * c] = sv[c];  // misses P_{3c'+2} of cell z0 - 1
        if (zt > 0) E[(rt - 1) * 36 + 18 + 2 * c + 1] = r0[0];  // the previous tile's cell z0 - 1 misses this
        else E[rt * 36 + 2 * c + 1] = 0.f;
      }
    } else if (lane == 63) {
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const float* r0 = pw + 3 * (8 - c) * kPsPitch;
        E[rt * 36 + 18 + 2 * c] = sv[c];  // misses P_{3c'} of cell z0 + 64
        if (zt + 1 < tz) E[(rt + 1) * 36 + 2 * c + 1] = r0[2 * kPsPitch + 63];  // the next tile's cell z0 + 64 misses this
        else E[rt * 36 + 18 + 2 * c + 1] = 0.f;
      }
    }
    wave_sync();
  }
}

// dx(u) = sum_c S_c(u + (c / 3 - 1, c % 3 - 1, 0)), eight cells along z per thread; rows
// outside the grid contribute zero.  The two end cells of a tile row take their sums from
// the side array instead of S: the partial sum plus the term the neighbouring tile exported
// (none at the grid's z ends).
__global__ __launch_bounds__(kThreads) void fused_sgather_kernel(const bf16* __restrict__ S,
                                                                 const float* __restrict__ E,
                                                                 bf16* __restrict__ dx, int n1, int n2, int n3) {
  const long V = (long)n1 * n2 * n3;
  const long id = (long)xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (id * 8 >= V) return;
  const long v0 = id * 8;
  const int k = (int)(v0 % n3);
  const long rowi = v0 / n3;
  const int j = (int)(rowi % n2), i = (int)(rowi / n2);
  const int kt = k % kFz, tz = n3 / kFz;
  const bool lo_end = kt == 0, hi_end = kt == kFz - 8;
  // every lane loads its side-array pairs (branch-free; lanes inside a tile row read the
  // pairs of the row's first cell and do not use them)
  float2 ev[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const int ii = min(max(i + c / 3 - 1, 0), n1 - 1), jj = min(max(j + c % 3 - 1, 0), n2 - 1);
    const long rt = ((long)ii * n2 + jj) * tz + k / kFz;
    ev[c] = *(const float2*)(E + rt * 36 + (hi_end ? 18 : 0) + 2 * c);
  }
  float acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = 0.f;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const int ii = i + c / 3 - 1, jj = j + c % 3 - 1;
    if (ii < 0 || ii >= n1 || jj < 0 || jj >= n2) continue;
    float sv[8];
    load_vec<8>(S + c * V + ((long)ii * n2 + jj) * n3 + k, sv);
    const float ends = ev[c].x + ev[c].y;
    sv[0] = lo_end ? ends : sv[0];
    sv[7] = hi_end ? ends : sv[7];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] += sv[t];
  }
  store_vec<8>(dx + v0, acc);
}


// Tile depth along i: the deepest of 8, 4 planes that keeps the tile's LDS (x and C image
// channels) within kMaxFusedLds, so at least four workgroups fit a CU, and that still
// gives kMinFusedBlocks workgroups (about eight per CU) to spread over the chip, else 2 (at
// 128^3: 2048 workgroups of 2-plane tiles, 16.0 us, against 1024 of 4-plane tiles, 17.1 us).
constexpr unsigned kMinFusedBlocks = 2048;
constexpr int kMaxFusedLds = 40 * 1024;

static int fused_fi(int C, int n1, int n2, int n3) {
  const unsigned b8 = (unsigned)((n1 / 8) * (n2 / kFj) * (n3 / kFz));
  const int fmax = (C + 1) * tile_cells(8) * 2 <= kMaxFusedLds ? 8 : 4;
  const int f = b8 >= kMinFusedBlocks ? 8 : (2 * b8 >= kMinFusedBlocks ? 4 : 2);
  return f < fmax ? f : fmax;
}

unsigned fused_blocks(int C, int n1, int n2, int n3) {
  return (unsigned)((n1 / fused_fi(C, n1, n2, n3)) * (n2 / kFj) * (n3 / kFz));
}

bool fused_eligible(int C, int B, int n1, int n2, int n3) {
  // a grid plane of FP32 x is addressed with 32-bit buffer offsets (under 2^31 bytes)
  return C >= 1 && C <= 3 && B == 1 && n1 % 8 == 0 && n2 % kFj == 0 && n3 % kFz == 0 &&
         (long)n2 * n3 * 4 < (1L << 31);
}

int launch_fused_pack(int tw, const void* W, const void* b, int C, void* wbuf, hipStream_t st, bool fold) {
  MLPCG_REQUIRE(C >= 1 && C <= 3, "fused_pack: C must be 1, 2 or 3");
  MLPCG_REQUIRE(!fold || C == 3, "fused_pack: folding needs C = 3");
  const int Cp = fold ? 2 : C;  // packed channels
  MLPCG_REQUIRE(tw == kBF16, "fused_pack: BF16 plans only (got %d)", tw);
  const int total = 2 * 2 * nkc(Cp) * 64 * 8;
  const int n = 2 * total + 2 * 2 * 64 * 4 + total + 2 * 64;
  fused_pack_kernel<<<(n + kThreads - 1) / kThreads, kThreads, 0, st>>>((const float*)W, (const float*)b, Cp,
                                                                       fold ? 1 : 0, (f16*)wbuf);
  return check_launch("fused_pack");
}

template <typename TX, typename TY, int C, int EPI, bool PL = false>
static void launch_fused_t(const FusedArgs& a, hipStream_t st) {
  const unsigned nb = fused_blocks(C, a.n1, a.n2, a.n3);
  const int fi = fused_fi(C, a.n1, a.n2, a.n3);
  if constexpr ((C + 1) * tile_cells(8) * 2 <= kMaxFusedLds) {
    if (fi == 8) return sm_fused_kernel<TX, TY, bf16, C, 8, EPI, PL><<<nb, kThreads, 0, st>>>(a);
  }
  if (fi == 4) sm_fused_kernel<TX, TY, bf16, C, 4, EPI, PL><<<nb, kThreads, 0, st>>>(a);
  else sm_fused_kernel<TX, TY, bf16, C, 2, EPI, PL><<<nb, kThreads, 0, st>>>(a);
}

// The epilogues of a preconditioner application get compile-time kernels: the pre blocks
// and the lone-block apply (none; FP32 -> BF16 on level 0, BF16 -> BF16 below), the
// post blocks below level 0 (prev) and level 0 (prev, the output mask and none, one or
// two solver dot products; BF16 -> FP32).  Anything else (the model's mask, other type
// pairs) runs the run-time epilogue.
template <typename TX, typename TY, int C>
static void launch_fused_epi(const FusedArgs& a, hipStream_t st) {
  const int epi = (a.prev ? kEpiPrev : 0) | (a.mask_a ? kEpiMaskA : 0) | (a.mask_b ? kEpiMaskB : 0) |
                  (a.partials ? kEpiDot0 : 0) | (a.partials && a.dot1 ? kEpiDot1 : 0);
  if constexpr (sizeof(TY) == 2) {
    if (epi == 0) return a.pool_out ? launch_fused_t<TX, TY, C, 0, true>(a, st) : launch_fused_t<TX, TY, C, 0>(a, st);
    if constexpr (sizeof(TX) == 2)
      if (epi == kEpiPrev) return launch_fused_t<TX, TY, C, kEpiPrev>(a, st);
  } else {
    if (epi == 0) return launch_fused_t<TX, TY, C, 0>(a, st);
    if (epi == (kEpiPrev | kEpiMaskB)) return launch_fused_t<TX, TY, C, kEpiPrev | kEpiMaskB>(a, st);
    if (epi == (kEpiPrev | kEpiMaskB | kEpiDot0))
      return launch_fused_t<TX, TY, C, kEpiPrev | kEpiMaskB | kEpiDot0>(a, st);
    if (epi == (kEpiPrev | kEpiMaskB | kEpiDot0 | kEpiDot1))
      return launch_fused_t<TX, TY, C, kEpiPrev | kEpiMaskB | kEpiDot0 | kEpiDot1>(a, st);
  }
  launch_fused_t<TX, TY, C, kEpiRuntime>(a, st);
}

int launch_fused_kwrite(int C, const FusedArgs& a, void* K, hipStream_t st) {
  MLPCG_REQUIRE(fused_eligible(C, 1, a.n1, a.n2, a.n3), "fused_kwrite: grid (%d,%d,%d) / C=%d not supported",
                a.n1, a.n2, a.n3, C);
  const int fi = fused_fi(C, a.n1, a.n2, a.n3);
  const unsigned nb = fused_blocks(C, a.n1, a.n2, a.n3);
#define MLPCG_KW(CC)                                                                      \
  if (C == CC) {                                                                          \
    if (fi == 8) fused_kwrite_kernel<CC, 8><<<nb, kThreads, 0, st>>>(a, (bf16*)K);         \
    else if (fi == 4) fused_kwrite_kernel<CC, 4><<<nb, kThreads, 0, st>>>(a, (bf16*)K);    \
    else fused_kwrite_kernel<CC, 2><<<nb, kThreads, 0, st>>>(a, (bf16*)K);                 \
  }
  MLPCG_KW(1) MLPCG_KW(2) MLPCG_KW(3)
#undef MLPCG_KW
  return check_launch("fused_kwrite");
}

size_t fused_sgrad_bytes(int n1, int n2, int n3) {
  const size_t V = (size_t)n1 * n2 * n3;
  return ((9 * V * 2 + 255) & ~(size_t)255) + (size_t)n1 * n2 * (n3 / kFz) * 36 * sizeof(float);
}

int launch_fused_sgrad(int C, const FusedArgs& a, void* work, void* dx, hipStream_t st) {
  MLPCG_REQUIRE(fused_eligible(C, 1, a.n1, a.n2, a.n3), "fused_sgrad: grid (%d,%d,%d) / C=%d not supported",
                a.n1, a.n2, a.n3, C);
  const long V = (long)a.n1 * a.n2 * a.n3;
  bf16* S = (bf16*)work;
  float* E = (float*)((char*)work + ((9 * (size_t)V * 2 + 255) & ~(size_t)255));
  const int fi = fused_fi(C, a.n1, a.n2, a.n3);
  const unsigned nb = fused_blocks(C, a.n1, a.n2, a.n3);
#define MLPCG_SW(CC)                                                                        \
  if (C == CC) {                                                                            \
    if (fi == 8) fused_swrite_kernel<CC, 8><<<nb, kThreads, 0, st>>>(a, S, E);              \
    else if (fi == 4) fused_swrite_kernel<CC, 4><<<nb, kThreads, 0, st>>>(a, S, E);         \
    else fused_swrite_kernel<CC, 2><<<nb, kThreads, 0, st>>>(a, S, E);                      \
  }
  MLPCG_SW(1) MLPCG_SW(2) MLPCG_SW(3)
#undef MLPCG_SW
  fused_sgather_kernel<<<(unsigned)((V / 8 + kThreads - 1) / kThreads), kThreads, 0, st>>>(S, E, (bf16*)dx, a.n1,
                                                                                          a.n2, a.n3);
  return check_launch("fused_sgrad");
}

int launch_sm_fused(int tw, int tx, int ty, int C, const FusedArgs& args, hipStream_t st) {
  MLPCG_REQUIRE(tw == kBF16, "sm_fused: BF16 plans only (got %d)", tw);
  MLPCG_REQUIRE(fused_eligible(C, 1, args.n1, args.n2, args.n3),
                "sm_fused: grid (%d,%d,%d) / C=%d not supported", args.n1, args.n2, args.n3, C);
  MLPCG_REQUIRE(!args.dot1 || args.dot0, "sm_fused: dot1 needs dot0");
  MLPCG_REQUIRE(!args.partials || args.dot0, "sm_fused: partials need dot0");
  MLPCG_REQUIRE(!args.pool_out || (ty == kBF16 && !args.prev && !args.mask_a && !args.mask_b && !args.partials),
                "sm_fused: the pooled output needs a BF16 output and no epilogue operands");
  const FusedArgs& a = args;
#define MLPCG_FU(TXX, TYY)                                 \
  do {                                                     \
    if (C == 1) launch_fused_epi<TXX, TYY, 1>(a, st);      \
    else if (C == 2) launch_fused_epi<TXX, TYY, 2>(a, st); \
    else launch_fused_epi<TXX, TYY, 3>(a, st);             \
  } while (0)
  if (tx == kF32 && ty == kBF16) MLPCG_FU(float, bf16);
  else if (tx == kBF16 && ty == kF32) MLPCG_FU(bf16, float);
  else if (tx == kBF16 && ty == kBF16) MLPCG_FU(bf16, bf16);
  else MLPCG_REQUIRE(false, "sm_fused: vector types (%d,%d) unsupported", tx, ty);
#undef MLPCG_FU
  return check_launch("sm_fused");
}

}  // namespace mlpcg
