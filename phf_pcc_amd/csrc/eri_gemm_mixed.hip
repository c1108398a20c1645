// Reduced-precision forms of the ERI GEMM (Gamma = A P for all grid points).
//
// Compensated split products on the matrix cores:
//   x = s^{-1} (x0 + x1 + x2),  x_i in FP16 or BF16,  s a power of two per row
// of A and per (grid point, re/im) row of P.  The kept terms of A P form one
// contraction over a sequence of segments, each a (P slice, A slice) pair:
//   F16X2  (3 products)  A [a0 | a0 | a1],           P [p0 | p1 | p0]
//   BF16X3 (6 products)  A [a0 a0 a1 a0 a1 a2],      P [p0 p1 p0 p2 p1 p0]
// Each slice is stored once per operand row; the kernels walk the segment
// sequence.  The MFMA accumulates in FP32; every FOLD k-slabs the FP32 tile is folded
// into an FP64 accumulator, and the FP64 epilogue applies s_P^{-1} s_A^{-1}
// (exact powers of two) and writes Gamma in FP64 for the FP64 epilogue of the
// grid pipeline.  FP32 form: one v_mfma_f32_16x16x4f32 product of the FP32
// roundings.
//
// Layout: P is [Mp][nslice*K], A is [Nr][nslice*K] (slices of one row side by
// side), C is [Mp][ldc]; Mp % BM == 0, Nr % BN == 0, K % 64 == 0.
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "phf_common.h"

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef int i4_t __attribute__((ext_vector_type(4)));

namespace {

// One 16-byte fragment per lane and one MFMA per 64-byte k-step: 16x16x32 for
// the 16-bit types (FP32 accumulation), 16x16x64 for int8 (exact int32); for
// FP32 the fragment is four k per lane and feeds four 16x16x4 MFMAs (lane l
// holds k = 4 (l >> 4) + s of the 16-wide k-step for the s-th MFMA, the same
// k for both operands, so the four products cover the step once).
template <typename T> struct MfmaH;
template <> struct MfmaH<float> {
  typedef f4_t v8;
  typedef f4_t acc_t;
  __device__ static acc_t mma(v8 a, v8 b, acc_t c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], c, 0, 0, 0);
  }
};
template <> struct MfmaH<_Float16> {
  typedef h8_t v8;
  typedef f4_t acc_t;
  __device__ static acc_t mma(v8 a, v8 b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct MfmaH<__bf16> {
  typedef b8_t v8;
  typedef f4_t acc_t;
  __device__ static acc_t mma(v8 a, v8 b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <> struct MfmaH<signed char> {
  typedef i4_t v8;
  typedef i4_t acc_t;
  __device__ static acc_t mma(v8 a, v8 b, acc_t c) { return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0); }
};

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// s_waitcnt with only vmcnt <= v (expcnt / lgkmcnt left at their maxima), gfx9 encoding
#define PHF_WAIT_VMCNT(v) __builtin_amdgcn_s_waitcnt(((v) & 0xF) | ((((v) >> 4) & 3) << 14) | (0x7 << 4) | (0xF << 8))

// 16-bit inputs, 16x16x32 MFMA (FP32 inputs: four 16x16x4 MFMAs per 16-byte
// fragment, one slice, no row scales).  Workgroup tile BM x BN (rows of P x rows of
// A), waves of WM x WN arranged (BM/WM) x (BN/WN).  The operands hold nslice
// slices per row ([x0 | x1 | ...], each K long); the contraction runs over a
// sequence of segments, segment sg pairing P slice (patP >> 3 sg) & 7 with A
// slice (patA >> 3 sg) & 7, so duplicated slices are read twice but stored once.
// Operands stream global -> LDS with global_load_lds_dwordx4 (no staging
// registers) through a ring of STAGES k-slabs of 64 elements; a slab row is
// 128 B = 8 chunks of 16 B, eight rows form a 1 KB block written by one wave
// instruction, and chunk c of row r sits at slot c ^ ((r >> 1) & 7) so the 16
// rows of an MFMA fragment read hit 16 distinct bank groups.  Every FOLD slabs
// the FP32 MFMA tile is added into an FP64 tile and restarted, keeping the FP32
// rounding of the long contraction at the level of a FOLD*64 chunk.  A split
// contraction (slab range [sbeg, send) of the segment sequence) adds its FP64
// partial into C with atomics; two partials onto zero give an order-independent
// result.  FOLD64: fold the FP32 MFMA tile into an FP64 tile (else into an FP32
// master tile: half the registers, used by the large per-wave tiles).
//
// The segments are walked one after another rather than fused per k-range
// (both slices of a range loaded once, three products per fragment pair):
// fusing adds the small cross products into the accumulator that holds the
// main product, and the truncating 16-bit MFMA then drops their low bits at
// the main term's ulp (measured: |dE| 3.0e-7 -> 4.7e-7 at n = 64 for an 8%
// faster contraction); a fresh accumulator per segment keeps them at their
// own scale.
template <int BM, int BN, int WM, int WN, int STAGES_, int MINB_, bool FOLD64_>
struct H16Cfg {
  static constexpr int NW = (BM / WM) * (BN / WN);
  static constexpr int NT = NW * 64;
  static constexpr int ROWS = BM + BN;
  static constexpr int STAGE_BYTES = ROWS * 64 * 2;
  static constexpr int STAGES = STAGES_;
  static constexpr int LDS = STAGES * STAGE_BYTES;
  static constexpr int MINB = MINB_;           // workgroups per CU the registers must allow
  static constexpr bool FOLD64 = FOLD64_;
  static_assert(LDS * MINB <= 160 * 1024, "LDS budget");
};

// The spin-free epilogue of one tile (see H16Item) for the fragments of one
// lane: rows row0 + 16 i + r, A rows ar0 + 16 j, values val(i, j, r, row, ar);
// C[row][colmap[ar]] = s1 v + bias.  Per group of eight elements the bias
// values are loaded first and the stores follow, so the group's loads are in
// flight together (a store may alias a later load as far as the compiler can
// tell, which would serialize them).
template <int FM, int FN, class Val>
__device__ __forceinline__ void epi_tile(double* __restrict__ C, long ldc, const double* __restrict__ bias,
                                         const int* __restrict__ colmap, const Epi& e, long row0, long ar0,
                                         Val val) {
  constexpr int JC = FN >= 2 ? 2 : 1;   // fragments per group
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j0 = 0; j0 < FN; j0 += JC) {
      int c1[JC][4];
      double b1[JC][4];
#pragma unroll
      for (int jj = 0; jj < JC; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long row = row0 + i * 16 + r, ar = ar0 + (j0 + jj) * 16;
          c1[jj][r] = colmap[ar];
          b1[jj][r] = (bias && c1[jj][r] >= 0) ? bias[(row & 1) * ldc + c1[jj][r]] : 0.0;
        }
#pragma unroll
      for (int jj = 0; jj < JC; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long row = row0 + i * 16 + r, ar = ar0 + (j0 + jj) * 16;
          if (c1[jj][r] >= 0) C[row * ldc + c1[jj][r]] = b1[jj][r] + e.s1 * val(i, j0 + jj, r, row, ar);
        }
    }
}

// Operand strides (elements): P rows ldP apart with slices sP apart, A rows
// ldA apart with slices sA apart; the contraction runs over the first K
// elements of each slice.  colmap (optional) sends A row r to column
// colmap[r] of C and of the bias (-1: padding, not written).
template <typename T, typename Cfg, int BM, int BN, int WM, int WN>
__device__ __forceinline__ void h16_tile(const T* __restrict__ P, long ldP, long sP, const T* __restrict__ A,
                                         long ldA, long sA, const double* __restrict__ invP,
                                         const double* __restrict__ invA, double* __restrict__ C,
                                         const double* __restrict__ bias, const int* __restrict__ colmap,
                                         long ldc, int K, unsigned patP, unsigned patA, unsigned patW, int mt,
                                         int nt, int sbeg, int send, bool atomic, int fold, unsigned char* smem,
                                         const Epi& epi = Epi()) {
  constexpr int BK = 128 / (int)sizeof(T);    // elements per slab: 128 bytes per slab row
  constexpr int STAGES = Cfg::STAGES, NW = Cfg::NW, ROWS = Cfg::ROWS, STAGE_BYTES = Cfg::STAGE_BYTES;
  constexpr int FM = WM / 16, FN = WN / 16, WGN = BN / WN;
  constexpr int BLOCKS = ROWS / 8;            // 1 KB blocks per slab
  constexpr int BPW = BLOCKS / NW;            // blocks (load instructions) per wave per slab
  static_assert(BPW * NW == BLOCKS && BPW < 64, "blocks per wave");
  typedef typename MfmaH<T>::v8 v8;
  typedef typename MfmaH<T>::acc_t acc_t;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int nk = send - sbeg;
  const int nsl = K / BK;                     // slabs per segment

  const T* src[BPW];
#pragma unroll
  for (int q = 0; q < BPW; ++q) {
    const int b = wave * BPW + q;
    const int r = 8 * b + (lane >> 3);
    const int c = (lane & 7) ^ ((r >> 1) & 7);
    const T* row = (r < BM) ? P + ((long)mt * BM + r) * ldP : A + ((long)nt * BN + (r - BM)) * ldA;
    src[q] = row + c * (16 / (int)sizeof(T));   // chunk c = 16 bytes
  }
  int iseg = sbeg / nsl, ioff = sbeg % nsl;   // segment / slab of the next slab to issue
  auto issue = [&](int slab) {
    unsigned char* st = smem + (slab % STAGES) * STAGE_BYTES;
    const long oP = (long)((patP >> (3 * iseg)) & 7) * sP + ioff * BK;
    const long oA = (long)((patA >> (3 * iseg)) & 7) * sA + ioff * BK;
#pragma unroll
    for (int q = 0; q < BPW; ++q) {
      const bool isP = (wave * BPW + q) < BM / 8;     // wave-uniform
      __builtin_amdgcn_global_load_lds((glb_void_t*)(src[q] + (isP ? oP : oA)),
                                       (lds_void_t*)(st + (wave * BPW + q) * 1024), 16, 0, 0);
    }
    if (++ioff == nsl) { ioff = 0; ++iseg; }
  };
  auto frag = [&](const unsigned char* st, int r, int c) -> v8 {
    return *reinterpret_cast<const v8*>(st + (r >> 3) * 1024 + ((r & 7) * 8 + (c ^ ((r >> 1) & 7))) * 16);
  };

  typedef typename std::conditional<Cfg::FOLD64, d4_t, f4_t>::type facc_t;
  acc_t acc[FM][FN];
  facc_t facc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j] = acc_t{0, 0, 0, 0};
      facc[i][j] = facc_t{0, 0, 0, 0};
    }
  int cseg = sbeg / nsl, coff = sbeg % nsl;   // segment / slab of the slab being multiplied

#pragma unroll
  for (int sl = 0; sl < STAGES - 1; ++sl)
    if (sl < nk) issue(sl);
  for (int kt = 0; kt < nk; ++kt) {
    // own slab-kt loads have landed; the barrier then publishes every wave's
    // slab kt and retires every wave's reads of the stage about to be refilled
    if (STAGES == 3 && kt + 1 < nk) PHF_WAIT_VMCNT(BPW);
    else PHF_WAIT_VMCNT(0);
    __builtin_amdgcn_s_barrier();
    if (kt + STAGES - 1 < nk) issue(kt + STAGES - 1);
    const unsigned char* st = smem + (kt % STAGES) * STAGE_BYTES;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {          // two 64-byte k-steps per slab
      v8 a[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = frag(st, wm * WM + i * 16 + (lane & 15), ks * 4 + (lane >> 4));
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const v8 b = frag(st, BM + wn * WN + j * 16 + (lane & 15), ks * 4 + (lane >> 4));
#pragma unroll
        for (int i = 0; i < FM; ++i) acc[i][j] = MfmaH<T>::mma(a[i], b, acc[i][j]);
      }
    }
    // fold at every fold-th slab, at the end and at every segment end; the
    // segment's weight 2^-(7 class) scales the exact int8 partials (1 for 16-bit)
    const bool seg_end = (++coff == nsl);
    if ((kt + 1) % fold == 0 || kt + 1 == nk || seg_end) {
      const double w = ldexp(1.0, -7 * (int)((patW >> (3 * cseg)) & 7));
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) facc[i][j][r] += (double)acc[i][j][r] * w;
          acc[i][j] = acc_t{0, 0, 0, 0};
        }
    }
    if (seg_end) {
      coff = 0;
      ++cseg;
    }
  }

  if (epi.on) {   // the spin-free epilogue (grouped launches)
    epi_tile<FM, FN>(C, ldc, bias, colmap, epi, (long)mt * BM + wm * WM + 4 * (lane >> 4),
                     (long)nt * BN + wn * WN + (lane & 15), [&](int i, int j, int r, long row, long ar) {
      double v = (double)facc[i][j][r];
      if (invP) v *= invP[row] * invA[ar];   // row scales of the split forms (none for FP32)
      return v;
    });
    return;
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        const long ar = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        const long col = colmap ? colmap[ar] : ar;
        if (col < 0) continue;
        double v = (double)facc[i][j][r];
        if (invP) v *= invP[row] * invA[ar];   // row scales of the split forms (none for FP32)
        if (bias) v += bias[(row & 1) * ldc + col];
        if (atomic)
          atomicAdd(&C[row * ldc + col], v);
        else
          C[row * ldc + col] = v;
      }
}

// One GEMM: workgroup -> (split part, row tile, column tile), XCD-aware.
template <typename T, typename Cfg, int BM, int BN, int WM, int WN, int SPLITK>
__global__ __launch_bounds__((Cfg::NT), (Cfg::MINB))
void eri_gemm_h16_kernel(const T* __restrict__ P, long ldP, long sP, const T* __restrict__ A, long ldA, long sA,
                         const double* __restrict__ invP, const double* __restrict__ invA, double* __restrict__ C,
                         int K, int nseg, unsigned patP, unsigned patA, unsigned patW, long ldc, int tiles_m,
                         const double* __restrict__ bias, const int* __restrict__ colmap, int fold) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int ks_id = t % SPLITK;
  const int tile = t / SPLITK;
  const int mt = tile % tiles_m, nt = tile / tiles_m;
  const int part = nseg * (K / (128 / (int)sizeof(T))) / SPLITK;
  h16_tile<T, Cfg, BM, BN, WM, WN>(P, ldP, sP, A, ldA, sA, invP, invA, C, ks_id == 0 ? bias : nullptr, colmap, ldc, K,
                                   patP, patA, patW, mt, nt, ks_id * part, (ks_id + 1) * part, SPLITK > 1, fold,
                                   smem);
}

// Exact int8 form (I8X4) on a 256 x 256 workgroup tile: eight waves of 128 x
// 64 (two per SIMD), 128 int32 accumulators each, 1 workgroup per CU.
// Integer accumulation is exact, so no FP64 fold tile is needed, and the
// larger tile halves the operand bytes per product of the 128 x 128 tile (the
// operand stream, not the matrix cores, bounds the smaller tile).  The kernels
// split every tile's slab range at each class boundary (i8_tile_parts); one
// range sums the products of one class exactly in int32 and adds
// S_c 2^(-7c) s_P^-1 s_A^-1 to C in FP64 (the first range of an unsplit tile
// stores, with the bias).  One int32 accumulator cannot carry several classes:
// rescaling between classes rounds below a class unit, and carrying the finer
// class instead overflows.  Every slice lies in [-127, 127], so class 3 (four
// products) sums at most 4 * 127^2 * K: K <= I8_TILE_MAX_K keeps every class
// inside int32.  Operand layout, LDS ring, swizzle and split contractions as
// in h16_tile; the loads address a wave-uniform base plus a 32-bit lane
// offset.  add: C += the result (the same lanes wrote C before); atomic:
// atomic adds onto a zero-filled C.
template <int STAGES>
__device__ __forceinline__ void i8_tile(const signed char* __restrict__ P, long ldP, long sP,
                                        const signed char* __restrict__ A, long ldA, long sA,
                                        const double* __restrict__ invP, const double* __restrict__ invA,
                                        double* __restrict__ C, const double* __restrict__ bias,
                                        const int* __restrict__ colmap, long ldc, int K, unsigned patP,
                                        unsigned patA, unsigned patW, int mt, int nt, int sbeg, int send,
                                        bool atomic, bool add, unsigned char* smem, const Epi& epi = Epi()) {
  constexpr int BM = 256, BN = 256, WM = 128, WN = 64, BK = 128;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int STAGE_BYTES = (BM + BN) * 128;
  constexpr int BPW = 8;                      // 1 KB blocks (8 rows) per wave per slab
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 2, wn = wave & 3;
  const int nk = send - sbeg, nsl = K / BK;

  // wave w loads slab rows 64 w .. 64 w + 63: waves 0-3 the P rows, 4-7 the A rows
  const bool isP = wave < 4;
  const long ld = isP ? ldP : ldA, sl = isP ? sP : sA;
  const signed char* wbase = isP ? P + ((long)mt * BM + (wave & 3) * 64) * ldP
                                 : A + ((long)nt * BN + (wave & 3) * 64) * ldA;
  // LDS block b = 8 wave + q holds rows 8 b .. 8 b + 7; lane -> row rl = lane / 8
  // of the block, whose chunk slot swizzle (r >> 1) & 7 = (4 (q & 1) + rl / 2) & 7
  const int rl = lane >> 3;
  const unsigned off_even = (unsigned)(rl * ld) + 16u * ((lane & 7) ^ ((rl >> 1) & 7));
  const unsigned off_odd = (unsigned)(rl * ld) + 16u * ((lane & 7) ^ ((4 + (rl >> 1)) & 7));

  int iseg = sbeg / nsl, ioff = sbeg % nsl;
  auto issue = [&](int slab) {
    unsigned char* st = smem + (slab % STAGES) * STAGE_BYTES + wave * BPW * 1024;
    const signed char* b = wbase + (long)((isP ? (patP >> (3 * iseg)) : (patA >> (3 * iseg))) & 7) * sl + ioff * BK;
#pragma unroll
    for (int q = 0; q < BPW; ++q)
      __builtin_amdgcn_global_load_lds((glb_void_t*)(b + (long)(8 * q) * ld + ((q & 1) ? off_odd : off_even)),
                                       (lds_void_t*)(st + q * 1024), 16, 0, 0);
    if (++ioff == nsl) { ioff = 0; ++iseg; }
  };
  auto frag = [&](const unsigned char* st, int r, int c) -> i4_t {
    return *reinterpret_cast<const i4_t*>(st + (r >> 3) * 1024 + ((r & 7) * 8 + (c ^ ((r >> 1) & 7))) * 16);
  };

  i4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = i4_t{0, 0, 0, 0};

#pragma unroll
  for (int sl0 = 0; sl0 < STAGES - 1; ++sl0)
    if (sl0 < nk) issue(sl0);
  for (int kt = 0; kt < nk; ++kt) {
    if (STAGES == 3 && kt + 1 < nk) PHF_WAIT_VMCNT(BPW);
    else PHF_WAIT_VMCNT(0);
    __builtin_amdgcn_s_barrier();
    if (kt + STAGES - 1 < nk) issue(kt + STAGES - 1);
    const unsigned char* st = smem + (kt % STAGES) * STAGE_BYTES;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      i4_t a[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = frag(st, wm * WM + i * 16 + (lane & 15), ks * 4 + (lane >> 4));
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const i4_t b = frag(st, BM + wn * WN + j * 16 + (lane & 15), ks * 4 + (lane >> 4));
#pragma unroll
        for (int i = 0; i < FM; ++i) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b, acc[i][j], 0, 0, 0);
      }
    }
  }

  const double w = ldexp(1.0, -7 * (int)((patW >> (3 * (sbeg / nsl))) & 7));
  if (epi.on && !add) {   // the spin-free epilogue (grouped launches), first class range
    epi_tile<FM, FN>(C, ldc, bias, colmap, epi, (long)mt * BM + wm * WM + 4 * (lane >> 4),
                     (long)nt * BN + wn * WN + (lane & 15), [&](int i, int j, int r, long row, long ar) {
      double v = (double)acc[i][j][r] * w * invP[row] * invA[ar];
      return v;
    });
    return;
  }
  const double sc = epi.on ? epi.s1 : 1.0;   // later class ranges of a spin-free item add s1 v
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        const long ar = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        const long col = colmap ? colmap[ar] : ar;
        if (col < 0) continue;
        double v = (double)acc[i][j][r] * w * invP[row] * invA[ar] * sc;
        if (bias && !add) v += bias[(row & 1) * ldc + col];
        if (atomic)
          atomicAdd(&C[row * ldc + col], v);
        else if (add)
          C[row * ldc + col] += v;
        else
          C[row * ldc + col] = v;
      }
}

constexpr int I8_TILE_STAGES = 2;
constexpr int I8_TILE_LDS = I8_TILE_STAGES * 512 * 128;

// Slab range [sbeg, send) of one tile, one single-class range after another.
__device__ __forceinline__ void i8_tile_parts(const signed char* P, long ldP, long sP, const signed char* A, long ldA,
                                              long sA, const double* invP, const double* invA, double* C,
                                              const double* bias, const int* colmap, long ldc, int K,
                                              unsigned patP, unsigned patA, unsigned patW, int mt, int nt,
                                              int sbeg, int send, bool atomic, unsigned char* smem,
                                              const Epi& epi = Epi()) {
  const int nsl = K / 128;
  bool add = false;
  for (int b = sbeg; b < send;) {
    // end of the class of slab b
    int sg = b / nsl;
    const unsigned cls = (patW >> (3 * sg)) & 7;
    while (sg + 1 < 10 && ((patW >> (3 * (sg + 1))) & 7) == cls) ++sg;
    const int e = min((sg + 1) * nsl, send);
    if (add) __syncthreads();   // every wave is done with the LDS ring of the previous range
    i8_tile<I8_TILE_STAGES>(P, ldP, sP, A, ldA, sA, invP, invA, C, bias, colmap, ldc, K, patP, patA, patW, mt, nt,
                            b, e, atomic, add, smem, epi);
    add = true;
    b = e;
  }
}

template <int SPLITK>
__global__ __launch_bounds__(512, 1)
void eri_gemm_i8_kernel(const signed char* __restrict__ P, long ldP, long sP, const signed char* __restrict__ A,
                        long ldA, long sA, const double* __restrict__ invP, const double* __restrict__ invA,
                        double* __restrict__ C, int K, int nseg, unsigned patP, unsigned patA, unsigned patW,
                        long ldc, int tiles_m, const double* __restrict__ bias, const int* __restrict__ colmap) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int ks_id = t % SPLITK;
  const int tile = t / SPLITK;
  const int mt = tile % tiles_m, nt = tile / tiles_m;
  const int part = nseg * (K / 128) / SPLITK;
  i8_tile_parts(P, ldP, sP, A, ldA, sA, invP, invA, C, ks_id == 0 ? bias : nullptr, colmap, ldc, K, patP, patA,
                patW, mt, nt, ks_id * part, (ks_id + 1) * part, SPLITK > 1, smem);
}

template <bool ATOMIC>
__global__ __launch_bounds__(512, 1)
void eri_gemm_i8_grouped_kernel(const H16Item* __restrict__ items, int nitems, unsigned patP, unsigned patA,
                                unsigned patW) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const H16Item it = items[w];
    if (it.K == 0) continue;   // empty slot
    __syncthreads();   // every wave is done with the previous tile's LDS ring
    i8_tile_parts((const signed char*)it.P, it.ldP, it.sP, (const signed char*)it.A, it.ldA, it.sA, it.invP,
                  it.invA, it.C, it.bias, it.colmap, it.ldc, it.K, patP, patA, patW, it.mt, it.nt, it.sbeg, it.send,
                  ATOMIC, smem, item_epi(it));
  }
}

// 16-bit forms on the 256 x 256 tile (eight 128 x 64 waves, two per SIMD,
// one workgroup per CU), without an FP64 fold tile.  Each pair of MFMAs (the
// two 32-element k-steps of a slab) starts from C = 0, so the instruction's
// truncating summation acts on a 64-product partial only; the partial is
// added to an FP32 accumulator with a round-to-nearest VALU add.  The 16-bit
// MFMA otherwise truncates against the running sum (a bias that grows with
// the contraction); here the per-step error is unbiased and relative to the
// partial.  Freeing the 128 registers of the FP64 fold tile is what lets a
// wave hold 128 x 64, which halves the operand bytes per product of the
// 128 x 128 tile.  Segments as in h16_tile (all weights 1); operand layout,
// LDS ring and swizzle as in i8_tile.
template <typename T, int STAGES>
__device__ __forceinline__ void rn16_tile(const T* __restrict__ P, long ldP, long sP, const T* __restrict__ A,
                                          long ldA, long sA, const double* __restrict__ invP,
                                          const double* __restrict__ invA, double* __restrict__ C,
                                          const double* __restrict__ bias, const int* __restrict__ colmap, long ldc,
                                          int K, unsigned patP, unsigned patA, int mt, int nt, int sbeg, int send,
                                          bool atomic, unsigned char* smem, const Epi& epi = Epi()) {
  constexpr int BM = 256, BN = 256, WM = 128, WN = 64, BK = 64;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int STAGE_BYTES = (BM + BN) * 128;
  constexpr int BPW = 8;                      // 1 KB blocks (8 rows) per wave per slab
  typedef typename MfmaH<T>::v8 v8;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 2, wn = wave & 3;
  const int nk = send - sbeg, nsl = K / BK;

  // wave w loads slab rows 64 w .. 64 w + 63: waves 0-3 the P rows, 4-7 the A rows
  const bool isP = wave < 4;
  const long ld = isP ? ldP : ldA, sl = isP ? sP : sA;
  const T* wbase = isP ? P + ((long)mt * BM + (wave & 3) * 64) * ldP : A + ((long)nt * BN + (wave & 3) * 64) * ldA;
  const int rl = lane >> 3;
  const unsigned off_even = (unsigned)(rl * ld) + 8u * ((lane & 7) ^ ((rl >> 1) & 7));
  const unsigned off_odd = (unsigned)(rl * ld) + 8u * ((lane & 7) ^ ((4 + (rl >> 1)) & 7));

  int iseg = sbeg / nsl, ioff = sbeg % nsl;
  auto issue = [&](int slab) {
    unsigned char* st = smem + (slab % STAGES) * STAGE_BYTES + wave * BPW * 1024;
    const T* b = wbase + (long)((isP ? (patP >> (3 * iseg)) : (patA >> (3 * iseg))) & 7) * sl + ioff * BK;
#pragma unroll
    for (int q = 0; q < BPW; ++q)
      __builtin_amdgcn_global_load_lds((glb_void_t*)(b + (long)(8 * q) * ld + ((q & 1) ? off_odd : off_even)),
                                       (lds_void_t*)(st + q * 1024), 16, 0, 0);
    if (++ioff == nsl) { ioff = 0; ++iseg; }
  };
  auto frag = [&](const unsigned char* st, int r, int c) -> v8 {
    return *reinterpret_cast<const v8*>(st + (r >> 3) * 1024 + ((r & 7) * 8 + (c ^ ((r >> 1) & 7))) * 16);
  };

  f4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

#pragma unroll
  for (int sl0 = 0; sl0 < STAGES - 1; ++sl0)
    if (sl0 < nk) issue(sl0);
  for (int kt = 0; kt < nk; ++kt) {
    if (STAGES == 3 && kt + 1 < nk) PHF_WAIT_VMCNT(BPW);
    else PHF_WAIT_VMCNT(0);
    __builtin_amdgcn_s_barrier();
    if (kt + STAGES - 1 < nk) issue(kt + STAGES - 1);
    const unsigned char* st = smem + (kt % STAGES) * STAGE_BYTES;
    // the B fragments of both k-steps stay in registers; A row block by row
    // block, so that only FN partials are live per block and the adds of one
    // block overlap the products of the next
    v8 b0[FN], b1[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      b0[j] = frag(st, BM + wn * WN + j * 16 + (lane & 15), lane >> 4);
      b1[j] = frag(st, BM + wn * WN + j * 16 + (lane & 15), 4 + (lane >> 4));
    }
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const v8 a0 = frag(st, wm * WM + i * 16 + (lane & 15), lane >> 4);
      const v8 a1 = frag(st, wm * WM + i * 16 + (lane & 15), 4 + (lane >> 4));
      f4_t t[FN];
#pragma unroll
      for (int j = 0; j < FN; ++j) t[j] = MfmaH<T>::mma(a0, b0[j], f4_t{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        t[j] = MfmaH<T>::mma(a1, b1[j], t[j]);
        acc[i][j] += t[j];
      }
    }
  }

  if (epi.on) {   // the spin-free epilogue (grouped launches)
    epi_tile<FM, FN>(C, ldc, bias, colmap, epi, (long)mt * BM + wm * WM + 4 * (lane >> 4),
                     (long)nt * BN + wn * WN + (lane & 15), [&](int i, int j, int r, long row, long ar) {
      double v = (double)acc[i][j][r] * invP[row] * invA[ar];
      return v;
    });
    return;
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        const long ar = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        const long col = colmap ? colmap[ar] : ar;
        if (col < 0) continue;
        double v = (double)acc[i][j][r] * invP[row] * invA[ar];
        if (bias) v += bias[(row & 1) * ldc + col];
        if (atomic)
          atomicAdd(&C[row * ldc + col], v);
        else
          C[row * ldc + col] = v;
      }
}

template <typename T, int SPLITK>
__global__ __launch_bounds__(512, 1)
void eri_gemm_rn16_kernel(const T* __restrict__ P, long ldP, long sP, const T* __restrict__ A, long ldA, long sA,
                          const double* __restrict__ invP, const double* __restrict__ invA, double* __restrict__ C,
                          int K, int nseg, unsigned patP, unsigned patA, long ldc, int tiles_m,
                          const double* __restrict__ bias, const int* __restrict__ colmap) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int ks_id = t % SPLITK;
  const int tile = t / SPLITK;
  const int mt = tile % tiles_m, nt = tile / tiles_m;
  const int part = nseg * (K / 64) / SPLITK;
  rn16_tile<T, I8_TILE_STAGES>(P, ldP, sP, A, ldA, sA, invP, invA, C, ks_id == 0 ? bias : nullptr, colmap, ldc, K,
                               patP, patA, mt, nt, ks_id * part, (ks_id + 1) * part, SPLITK > 1, smem);
}

template <typename T, bool ATOMIC>
__global__ __launch_bounds__(512, 1)
void eri_gemm_rn16_grouped_kernel(const H16Item* __restrict__ items, int nitems, unsigned patP, unsigned patA) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const H16Item it = items[w];
    if (it.K == 0) continue;   // empty slot
    __syncthreads();   // every wave is done with the previous tile's LDS ring
    rn16_tile<T, I8_TILE_STAGES>((const T*)it.P, it.ldP, it.sP, (const T*)it.A, it.ldA, it.sA, it.invP, it.invA,
                                 it.C, it.bias, it.colmap, it.ldc, it.K, patP, patA, it.mt, it.nt, it.sbeg, it.send,
                                 ATOMIC, smem, item_epi(it));
  }
}

// Many GEMMs of one precision in one persistent launch: workgroup b takes
// items b, b + gridDim.x, ... of a list the host ordered for load balance and
// XCD locality.
template <typename T, typename Cfg, int BM, int BN, int WM, int WN, bool ATOMIC>
__global__ __launch_bounds__((Cfg::NT), (Cfg::MINB))
void eri_gemm_h16_grouped_kernel(const H16Item* __restrict__ items, int nitems, unsigned patP, unsigned patA,
                                 unsigned patW, int fold) {
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const H16Item it = items[w];
    if (it.K == 0) continue;   // empty slot
    __syncthreads();   // every wave is done with the previous tile's LDS ring
    h16_tile<T, Cfg, BM, BN, WM, WN>((const T*)it.P, it.ldP, it.sP, (const T*)it.A, it.ldA, it.sA, it.invP, it.invA,
                                     it.C, it.bias, it.colmap, it.ldc, it.K, patP, patA, patW, it.mt, it.nt,
                                     it.sbeg, it.send, ATOMIC, fold, smem, item_epi(it));
  }
}

// ---------------------------------------------------------------------------
// operand conversion
// ---------------------------------------------------------------------------

template <typename T> __device__ inline T to16(double x);
// Two hardware conversions (f64 -> f32 -> 16-bit).  Without the empty asm the
// compiler folds them into one f64 -> 16-bit conversion, which has no
// instruction and is expanded into ~40 integer operations.  The double
// rounding can move the first slice by one unit in a tie; the next slice
// takes the remainder either way.
template <> __device__ inline _Float16 to16<_Float16>(double x) {
  float f = (float)x;
  asm volatile("" : "+v"(f));
  return (_Float16)f;
}
template <> __device__ inline __bf16 to16<__bf16>(double x) {
  float f = (float)x;
  asm volatile("" : "+v"(f));
  return (__bf16)f;
}
template <> __device__ inline signed char to16<signed char>(double x) { return (signed char)x; }

// One workgroup per row: s = 2^(top - ceil(log2 max|x|)); nslice slices x_i of
// s*x (x_0 the rounding of s*x, x_1 of the remainder, ...), stored as
// [x_0 | x_1 | ...] along the row, each kout long.  With cmap, output column c
// takes source column cmap[c] (0 for -1); without, kout = kt and c maps to c.
// amax (optional): the rows' largest magnitudes, known already (one pass over
// the row instead of two); rows >= rows_data are zero and are not read.
template <typename T, int NS>
__global__ __launch_bounds__(256) void split_rows_kernel(const double* __restrict__ src, long lds, int kt,
                                                         const int* __restrict__ cmap, int kout,
                                                         T* __restrict__ dst, int top,
                                                         double* __restrict__ inv, const double* __restrict__ amax,
                                                         long rows_data) {
  __shared__ double red[256];
  const long row = blockIdx.x;
  const bool data = row < rows_data;
  const double* x = src + row * lds;
  double m = 0.0;
  if (amax) {
    m = data ? amax[row] : 0.0;
  } else {
    for (int c = threadIdx.x; c < kt; c += 256) m = fmax(m, fabs(x[c]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    m = red[0];
  }
  int e = 0;
  if (m > 0.0) {
    frexp(m, &e);             // m = f * 2^e with f in [0.5, 1)
  }
  const double sc = (m > 0.0) ? ldexp(1.0, top - e) : 1.0;
  if (threadIdx.x == 0) inv[row] = 1.0 / sc;
  // consecutive lanes take consecutive columns: the (gathered) reads and the
  // slice stores stay coalesced
  T* d = dst + row * (long)NS * kout;
  const bool vec4 = (kout & 3) == 0 && (reinterpret_cast<uintptr_t>(dst) % (4 * sizeof(T))) == 0 &&
                    (reinterpret_cast<uintptr_t>(cmap) % 16) == 0;
  if (vec4) {
    // four columns per lane and step: one 16-byte map load, four independent
    // gathers, one 4-element store per slice (kout % 4 == 0 and the aligned
    // bases keep every vector access aligned)
    typedef T T4 __attribute__((ext_vector_type(4)));
    for (int c0 = 4 * threadIdx.x; c0 < kout; c0 += 1024) {
      int cols[4];
      if (cmap) {
        const int4 cm = *reinterpret_cast<const int4*>(cmap + c0);
        cols[0] = cm.x; cols[1] = cm.y; cols[2] = cm.z; cols[3] = cm.w;
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) cols[u] = c0 + u;
      }
      double r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) r[u] = (data && cols[u] >= 0 ? x[cols[u]] : 0.0) * sc;
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        T4 v;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if constexpr (std::is_same<T, signed char>::value) {
            const double t = fmin(127.0, fmax(-127.0, rint(r[u])));
            v[u] = (signed char)t;
            r[u] = (r[u] - t) * 128.0;
          } else {
            const T t = to16<T>(r[u]);
            v[u] = t;
            r[u] -= (double)(float)t;
          }
        }
        *reinterpret_cast<T4*>(d + (long)q * kout + c0) = v;
      }
    }
    return;
  }
  for (int c = threadIdx.x; c < kout; c += 256) {
    const int sc_col = cmap ? cmap[c] : c;
    double r = (data && sc_col >= 0 ? x[sc_col] : 0.0) * sc;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if constexpr (std::is_same<T, signed char>::value) {
        // integer slices: x_q = rint(r) in [-127, 127], carry 128 (r - x_q)
        const double v = fmin(127.0, fmax(-127.0, rint(r)));
        d[(long)q * kout + c] = (signed char)v;
        r = (r - v) * 128.0;
      } else {
        const T v = to16<T>(r);
        d[(long)q * kout + c] = v;
        r -= (double)(float)v;
      }
    }
  }
}

// The point's density tile rho^T = conj(C0) X_g^T (128 x 128, contraction nk;
// C0 and X_g column-major NSO x N, complex interleaved) on the FP64 matrix
// cores with the complex GEMM's exact MFMA sequence per output (three real
// products per complex one, 4-deep steps in k order).  On return p1 holds the
// real and p2 the imaginary parts: fragment (i, j), register r is tile row
// wm 16 FM + 16 i + (lane >> 4) + 4 r, tile column wn 16 FN + 16 j + (lane & 15)
// (waves of (16 FM) x (16 FN), WN = 128 / (16 FN) across).  Without maps tile
// row / column q is MO q (q < nso); with MAPPED, rmap[q] / cmap[q] name the MO
// (-1: an empty row or column, zero).
template <int NT, int FM, int FN, bool MAPPED, bool K2>
__device__ __forceinline__ void dens_tile(int nso, int nk, const double* __restrict__ C0,
                                          const double* __restrict__ Xg, const int* __restrict__ rmap,
                                          const int* __restrict__ cmap, d4_t (&p1)[FM][FN], d4_t (&p2)[FM][FN]) {
  constexpr int TB = 128, KC = 16, LD = KC + 2, WN = TB / (16 * FN), CH = TB * LD, NB = K2 ? 2 : 1;
  // K2 (nk <= 2 KC, i.e. N <= 32): both chunks staged at once in two buffers,
  // one barrier and no second load latency; else one chunk after another
  __shared__ double Ar[NB * CH], Ai[NB * CH], Br[NB * CH], Bi[NB * CH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  d4_t p3[FM][FN];
  // every load of NC 16-deep k-chunks is issued before the first LDS store;
  // the first stage is peeled so that its loads are consumed before the
  // accumulators come alive
  constexpr int EL = TB * KC / NT;
  auto stage = [&](int k0, auto nc_tag) {
    constexpr int NC = decltype(nc_tag)::value;
    double lar[NC * EL], lai[NC * EL], lbr[NC * EL], lbi[NC * EL];
    bool la[NC * EL], lb[NC * EL];
#pragma unroll
    for (int e = 0; e < NC * EL; ++e) {
      const int idx = (e % EL) * NT + tid, i = idx % TB, gk = k0 + (e / EL) * KC + idx / TB;
      int ra = i, rb = i;
      if constexpr (MAPPED) {
        ra = rmap[i];
        rb = cmap[i];
      }
      la[e] = (MAPPED ? ra >= 0 : i < nso) && gk < nk;
      lb[e] = (MAPPED ? rb >= 0 : i < nso) && gk < nk;
      const long offa = la[e] ? 2L * (ra + (long)gk * nso) : 0, offb = lb[e] ? 2L * (rb + (long)gk * nso) : 0;
      lar[e] = C0[offa];
      lai[e] = C0[offa + 1];
      lbr[e] = Xg[offb];
      lbi[e] = Xg[offb + 1];
    }
#pragma unroll
    for (int e = 0; e < NC * EL; ++e) {
      const int idx = (e % EL) * NT + tid, i = idx % TB, kk = idx / TB, o = (e / EL) * CH + i * LD + kk;
      Ar[o] = la[e] ? lar[e] : 0.0;
      Ai[o] = la[e] ? -lai[e] : 0.0;
      Br[o] = lb[e] ? lbr[e] : 0.0;
      Bi[o] = lb[e] ? lbi[e] : 0.0;
    }
  };
  auto mfma_chunk = [&](int buf) {
    const double *ar_ = Ar + buf * CH, *ai_ = Ai + buf * CH, *br_ = Br + buf * CH, *bi_ = Bi + buf * CH;
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      double are[FM], aim[FM], bre[FN], bim[FN];
#pragma unroll
      for (int t = 0; t < FM; ++t) {
        const int row = wm * 16 * FM + t * 16 + (lane & 15);
        are[t] = ar_[row * LD + ks + (lane >> 4)];
        aim[t] = ai_[row * LD + ks + (lane >> 4)];
      }
#pragma unroll
      for (int t = 0; t < FN; ++t) {
        const int col = wn * 16 * FN + t * 16 + (lane & 15);
        bre[t] = br_[col * LD + ks + (lane >> 4)];
        bim[t] = bi_[col * LD + ks + (lane >> 4)];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i], bre[j], p1[i][j], 0, 0, 0);
          p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[i], bim[j], p2[i][j], 0, 0, 0);
          p3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i] + aim[i], bre[j] + bim[j], p3[i][j], 0, 0, 0);
        }
    }
  };
  auto zero = [&]() {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        p1[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
        p2[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
        p3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
      }
  };
  if constexpr (K2) {
    stage(0, std::integral_constant<int, 2>());
    __syncthreads();
    zero();
    mfma_chunk(0);
    if (nk > KC) mfma_chunk(1);
  } else {
    stage(0, std::integral_constant<int, 1>());
    __syncthreads();
    zero();
    mfma_chunk(0);
    for (int k0 = KC; k0 < nk; k0 += KC) {
      __syncthreads();
      stage(k0, std::integral_constant<int, 1>());
      __syncthreads();
      mfma_chunk(0);
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double xr = p1[i][j][r] - p2[i][j][r], xi = p3[i][j][r] - p1[i][j][r] - p2[i][j][r];
        p1[i][j][r] = xr;
        p2[i][j][r] = xi;
      }
}

// NS slices of x (scaled) at element c of the row slices d (slice stride kout).
template <typename T, int NS>
__device__ __forceinline__ void put_slices(T* __restrict__ d, long kout, long c, double x) {
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    if constexpr (std::is_same<T, signed char>::value) {
      const double v = fmin(127.0, fmax(-127.0, rint(x)));
      d[(long)q * kout + c] = (signed char)v;
      x = (x - v) * 128.0;
    } else {
      const T v = to16<T>(x);
      d[(long)q * kout + c] = v;
      x -= (double)(float)v;
    }
  }
}

// Density pack fused with the split of its rows, for NSO <= 128: one
// 512-thread workgroup per grid point forms rho^T = conj(C0) X_g^T (dens_tile),
// takes the largest |Re| and |Im| over the point, and writes the NS slices of
// both planes (rows 2g and 2g + 1) straight into the ERI GEMM's operand,
// element (j, l) at column ocol[j NSO + l] (j NSO + l without a map).  The
// slices are those split_rows_kernel makes from the FP64 rows, bit for bit,
// without writing and reading those rows.  Columns no element maps to are
// zeroed (padcols, or [NSO^2, kout) without a map); workgroups past the last
// point clear the padding rows up to rows_total (zero slices, inverse scale 1).
template <typename T, int NS, int NT, int FM, int FN, bool K2>
__global__ __launch_bounds__(NT) void dens_split_kernel(int nso, int nk, const double* __restrict__ C0,
                                                          const double* __restrict__ X, long sX, int npts,
                                                          const int* __restrict__ ocol,
                                                          const int* __restrict__ padcols, int npad, int kout,
                                                          long rows_total, T* __restrict__ dst, int top,
                                                          double* __restrict__ inv) {
  constexpr int WN = 128 / (16 * FN);
  __shared__ double red[2][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x;
  const T zero = static_cast<T>(0.0f);
  if (g >= npts) {
    for (int pr = 0; pr < 2; ++pr) {
      const long row = 2L * npts + 2L * (g - npts) + pr;
      if (row >= rows_total) break;
      T* d = dst + row * (long)NS * kout;
      for (long c = tid; c < (long)NS * kout; c += NT) d[c] = zero;
      if (tid == 0) inv[row] = 1.0;
    }
    return;
  }
  const int wm = wave / WN, wn = wave % WN;
  d4_t p1[FM][FN], p2[FM][FN];
  dens_tile<NT, FM, FN, false, K2>(nso, nk, C0, X + (long)g * sX, nullptr, nullptr, p1, p2);
  // complex results and the point's largest magnitudes
  double mre = 0.0, mim = 0.0;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = wm * 16 * FM + i * 16 + (lane >> 4) + 4 * r, gj = wn * 16 * FN + j * 16 + (lane & 15);
        if (gi < nso && gj < nso) {
          mre = fmax(mre, fabs(p1[i][j][r]));
          mim = fmax(mim, fabs(p2[i][j][r]));
        }
      }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mre = fmax(mre, __shfl_xor(mre, off));
    mim = fmax(mim, __shfl_xor(mim, off));
  }
  if (lane == 0) {
    red[0][wave] = mre;
    red[1][wave] = mim;
  }
  __syncthreads();
  double mx[2] = {0.0, 0.0};
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    mx[0] = fmax(mx[0], red[0][w]);
    mx[1] = fmax(mx[1], red[1][w]);
  }
  double sc[2];
#pragma unroll
  for (int pl = 0; pl < 2; ++pl) {
    int e = 0;
    if (mx[pl] > 0.0) frexp(mx[pl], &e);
    sc[pl] = (mx[pl] > 0.0) ? ldexp(1.0, top - e) : 1.0;
  }
  const long row0 = 2L * g;
  if (tid == 0) {
    inv[row0] = 1.0 / sc[0];
    inv[row0 + 1] = 1.0 / sc[1];
  }
  T* d0 = dst + row0 * (long)NS * kout;
  T* d1 = d0 + (long)NS * kout;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = wm * 16 * FM + i * 16 + (lane >> 4) + 4 * r, gj = wn * 16 * FN + j * 16 + (lane & 15);
        if (gi >= nso || gj >= nso) continue;
        const int src = gi * nso + gj;
        const int c = ocol ? ocol[src] : src;
        if (c < 0) continue;
        put_slices<T, NS>(d0, kout, c, p1[i][j][r] * sc[0]);
        put_slices<T, NS>(d1, kout, c, p2[i][j][r] * sc[1]);
      }
  // output columns no element maps to
  if (ocol) {
    for (int p = tid; p < npad; p += NT) {
      const int c = padcols[p];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        d0[(long)q * kout + c] = zero;
        d1[(long)q * kout + c] = zero;
      }
    }
  } else {
    for (int c = nso * nso + tid; c < kout; c += NT)
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        d0[(long)q * kout + c] = zero;
        d1[(long)q * kout + c] = zero;
      }
  }
}

// Spin-free operand (labels s < n <= 64; smap from dens_split_rjk_maps): the
// staging maps put label La's alpha and beta MOs on the tile rows of fragments
// 2m and 2m + 1 and label Lc's on the columns of fragments 0 and 1, so one
// lane holds P[(alpha b, alpha d)], P[(beta b, beta d)], P[(beta b, alpha d)] and
// P[(alpha b, beta d)] of a label pair (b, d) and writes, at column b n + d,
// set 0 = P_aa + P_bb, set 1 = P_aa - P_bb, set 2 = P_x1, set 3 = P_x2 at
// rows set Mp + 2g + plane (Mk = 3 Mp), each (set, plane) row with its own
// scale.  The values and slices equal those of the density pack,
// rjk_pack_kernel and split_rows_kernel, bit for bit.  Columns [n^2, kj) are
// zeroed; workgroups past the last point clear each set's padding rows
// [2 npts, Mp).
template <typename T, int NS, int NT, int FM, int FN, bool K2>
__global__ __launch_bounds__(NT) void dens_split_rjk_kernel(int n, int nso, int nk, const double* __restrict__ C0,
                                                              const double* __restrict__ X, long sX, int npts,
                                                              const int* __restrict__ smap, int kj, long Mp,
                                                              long Mk, T* __restrict__ dst, int top,
                                                              double* __restrict__ inv) {
  static_assert(FM == 4 && FN == 2, "label layout assumes 4 x 2 fragments per wave");
  constexpr int WN = 128 / (16 * FN);
  __shared__ double red[8][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = blockIdx.x;
  const T zero = static_cast<T>(0.0f);
  const long rs = (long)NS * kj;   // operand row stride
  if (g >= npts) {
    const long pad0 = Mp - 2L * npts;
    for (int pr = 0; pr < 2; ++pr) {
      const long p = 2L * (g - npts) + pr;
      const long set = p / pad0, row = set * Mp + 2L * npts + p % pad0;   // each set's padding rows
      if (set >= 4) break;
      T* d = dst + row * rs;
      for (long c = tid; c < rs; c += NT) d[c] = zero;
      if (tid == 0) inv[row] = 1.0;
    }
    return;
  }
  const int wm = wave / WN, wn = wave % WN;
  d4_t p1[FM][FN], p2[FM][FN];
  dens_tile<NT, FM, FN, true, K2>(nso, nk, C0, X + (long)g * sX, smap, smap + 128, p1, p2);
  const int lc = wn * 16 + (lane & 15);
  // the four sets of one (m, r): value q = set * 2 + plane
  auto sets = [&](int m, int r, double (&v)[8]) {
    const double ar = p1[2 * m][0][r], ai = p2[2 * m][0][r], br = p1[2 * m + 1][1][r], bi = p2[2 * m + 1][1][r];
    v[0] = ar + br;
    v[1] = ai + bi;
    v[2] = ar - br;
    v[3] = ai - bi;
    v[4] = p1[2 * m + 1][0][r];
    v[5] = p2[2 * m + 1][0][r];
    v[6] = p1[2 * m][1][r];
    v[7] = p2[2 * m][1][r];
  };
  double mx[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) mx[q] = 0.0;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int la = wm * 32 + m * 16 + (lane >> 4) + 4 * r;
      if (la >= n || lc >= n) continue;
      double v[8];
      sets(m, r, v);
#pragma unroll
      for (int q = 0; q < 8; ++q) mx[q] = fmax(mx[q], fabs(v[q]));
    }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx[q] = fmax(mx[q], __shfl_xor(mx[q], off));
    if (lane == 0) red[q][wave] = mx[q];
  }
  __syncthreads();
  double sc[8];
  T* drow[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    double m = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) m = fmax(m, red[q][w]);
    int e = 0;
    if (m > 0.0) frexp(m, &e);
    sc[q] = (m > 0.0) ? ldexp(1.0, top - e) : 1.0;
    const int set = q >> 1, pl = q & 1;
    const long row = (long)set * Mp + 2L * g + pl;
    drow[q] = dst + row * rs;
    if (tid == 0) inv[row] = 1.0 / sc[q];
  }
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int la = wm * 32 + m * 16 + (lane >> 4) + 4 * r;
      if (la >= n || lc >= n) continue;
      double v[8];
      sets(m, r, v);
      const long c = (long)la * n + lc;
#pragma unroll
      for (int q = 0; q < 8; ++q) put_slices<T, NS>(drow[q], kj, c, v[q] * sc[q]);
    }
  for (long c = (long)n * n + tid; c < kj; c += NT)
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int s = 0; s < NS; ++s) drow[q][(long)s * kj + c] = zero;
}

__global__ void to_f32_kernel(const double* __restrict__ src, float* __restrict__ dst, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

}  // namespace

// pattern codes: entries are slice indices, segment sg uses bits [3 sg, 3 sg + 3)
static unsigned pack_pattern(const int* p, int n) {
  unsigned v = 0;
  for (int i = 0; i < n; ++i) v |= (unsigned)p[i] << (3 * i);
  return v;
}

// precision modes (values of PHF_PREC_*): stored slices per operand row,
// product segments of the contraction, bytes per stored element
int mixed_slices(int prec) { return prec == 4 ? 4 : (prec == 3 ? 2 : (prec == 2 ? 3 : 1)); }
int mixed_segments(int prec) { return prec == 4 ? 10 : (prec == 3 ? 3 : (prec == 2 ? 6 : 1)); }
int mixed_elem_bytes(int prec) { return prec == 1 ? 4 : (prec == 4 ? 1 : 2); }

//   F16X2  (3 products)  P slices [p0 p1 p0] against A slices [a0 a0 a1]
//   BF16X3 (6 products)  P slices [p0 p1 p0 p2 p1 p0] against A [a0 a0 a1 a0 a1 a2]
//   I8X4  (10 products)  every (p_i, a_j) with i + j <= 3, ordered by class i + j;
//                        class c carries the weight 2^-7c (patW), applied at the
//                        fold of the exact int32 partial sums
// The int8 tile walks the same ten products by decreasing class (3, 2, 1, 0),
// patW holding each segment's class.
static void mixed_patterns_desc(unsigned* patP, unsigned* patA, unsigned* patW) {
  static const int p_d[10] = {0, 1, 2, 3, 0, 1, 2, 0, 1, 0}, a_d[10] = {3, 2, 1, 0, 2, 1, 0, 1, 0, 0};
  static const int w_d[10] = {3, 3, 3, 3, 2, 2, 2, 1, 1, 0};
  *patP = pack_pattern(p_d, 10);
  *patA = pack_pattern(a_d, 10);
  *patW = pack_pattern(w_d, 10);
}

static void mixed_patterns(int prec, unsigned* patP, unsigned* patA, unsigned* patW) {
  static const int a_f16[3] = {0, 0, 1}, p_f16[3] = {0, 1, 0};
  static const int a_b16[6] = {0, 0, 1, 0, 1, 2}, p_b16[6] = {0, 1, 0, 2, 1, 0};
  static const int p_i8[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3}, a_i8[10] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0};
  static const int w_i8[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
  *patW = 0;
  if (prec == 1) {   // one segment: slice 0 against slice 0
    *patP = 0;
    *patA = 0;
  } else if (prec == 3) {
    *patP = pack_pattern(p_f16, 3);
    *patA = pack_pattern(a_f16, 3);
  } else if (prec == 2) {
    *patP = pack_pattern(p_b16, 6);
    *patA = pack_pattern(a_b16, 6);
  } else {
    *patP = pack_pattern(p_i8, 10);
    *patA = pack_pattern(a_i8, 10);
    *patW = pack_pattern(w_i8, 10);
  }
}

void split_operand_map(hipStream_t s, int prec, const double* src, long lds, int rows, int kt, const int* cmap,
                       int kout, void* dst, double* inv, const double* amax, int rows_data) {
  const long rd = rows_data < 0 ? rows : rows_data;
  if (prec == 3)
    split_rows_kernel<_Float16, 2><<<rows, 256, 0, s>>>(src, lds, kt, cmap, kout, (_Float16*)dst, 14, inv, amax, rd);
  else if (prec == 2)
    split_rows_kernel<__bf16, 3><<<rows, 256, 0, s>>>(src, lds, kt, cmap, kout, (__bf16*)dst, 14, inv, amax, rd);
  else
    split_rows_kernel<signed char, 4><<<rows, 256, 0, s>>>(src, lds, kt, cmap, kout, (signed char*)dst, 7, inv,
                                                           amax, rd);
}

// workgroup shape of the fused density pack + split: 512 threads, waves of
// 64 x 32 over the 128 x 128 tile (1024 threads spill the three accumulator
// sets; four real products per complex one at 1024 threads measured within 10%)
constexpr int DS_NT = 512, DS_FM = 4, DS_FN = 2;
int dens_split_max_nso() { return 128; }

void dens_split(hipStream_t s, int prec, int nso, int nk, const double* C0, const double* X, long sX, int npts,
                const int* ocol, const int* padcols, int npad, int kout, long rows_total, void* dst, double* inv) {
  const unsigned grid = (unsigned)(npts + (rows_total - 2L * npts + 1) / 2);
  auto go = [&](auto k2) {
    constexpr bool K2 = decltype(k2)::value;
    if (prec == 3)
      dens_split_kernel<_Float16, 2, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          nso, nk, C0, X, sX, npts, ocol, padcols, npad, kout, rows_total, (_Float16*)dst, 14, inv);
    else if (prec == 2)
      dens_split_kernel<__bf16, 3, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          nso, nk, C0, X, sX, npts, ocol, padcols, npad, kout, rows_total, (__bf16*)dst, 14, inv);
    else
      dens_split_kernel<signed char, 4, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          nso, nk, C0, X, sX, npts, ocol, padcols, npad, kout, rows_total, (signed char*)dst, 7, inv);
  };
  if (nk <= 32) go(std::true_type());
  else go(std::false_type());
}

void dens_split_rjk_maps(const std::vector<int>& mo, int n, int* smap) {
  // tile row q: wave row wm = q / 64, fragment f = (q % 64) / 16 (pair m = f / 2,
  // spin f % 2), x = q % 16 -> label wm 32 + 16 m + x; tile column q: wave
  // column q / 32, fragment (q % 32) / 16 = spin, x = q % 16 -> label (q / 32) 16 + x
  for (int q = 0; q < 128; ++q) {
    const int f = (q % 64) / 16, lr = (q / 64) * 32 + (f / 2) * 16 + q % 16;
    smap[q] = lr < n ? mo[(f % 2) * n + lr] : -1;
    const int lc = (q / 32) * 16 + q % 16;
    smap[128 + q] = lc < n ? mo[((q % 32) / 16) * n + lc] : -1;
  }
}

void dens_split_rjk(hipStream_t s, int prec, int n, int nk, const double* C0, const double* X, long sX, int npts,
                    const int* smap, int kj, long Mp, long Mk, void* dst, double* inv) {
  const long pad = 4 * (Mp - 2L * npts);   // Mk = 3 Mp: four sets of Mp rows
  const unsigned grid = (unsigned)(npts + (pad + 1) / 2);
  const int nso = 2 * n;
  auto go = [&](auto k2) {
    constexpr bool K2 = decltype(k2)::value;
    if (prec == 3)
      dens_split_rjk_kernel<_Float16, 2, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          n, nso, nk, C0, X, sX, npts, smap, kj, Mp, Mk, (_Float16*)dst, 14, inv);
    else if (prec == 2)
      dens_split_rjk_kernel<__bf16, 3, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          n, nso, nk, C0, X, sX, npts, smap, kj, Mp, Mk, (__bf16*)dst, 14, inv);
    else
      dens_split_rjk_kernel<signed char, 4, DS_NT, DS_FM, DS_FN, K2><<<grid, DS_NT, 0, s>>>(
          n, nso, nk, C0, X, sX, npts, smap, kj, Mp, Mk, (signed char*)dst, 7, inv);
  };
  if (nk <= 32) go(std::true_type());
  else go(std::false_type());
}

void split_operand(hipStream_t s, int prec, const double* src, long lds, int rows, int kt, void* dst, double* inv) {
  split_operand_map(s, prec, src, lds, rows, kt, nullptr, kt, dst, inv, nullptr, -1);
}

void convert_f32(hipStream_t s, const double* src, float* dst, long n) {
  long blocks = std::min<long>((n + 255) / 256, 65536);
  to_f32_kernel<<<blocks, 256, 0, s>>>(src, dst, n);
}

// Slabs (64 contraction elements) per FP32 chunk before the fold into FP64.
// Folding more often does not lower the error of these forms: the 16-bit
// MFMA sums its products with truncation inside the instruction, which no
// fold interval removes (measured at 8, 4, 2 and 1 slabs on n = 128).
// PHF_H16_FOLD=<slabs> overrides it (a value above the slab count folds only
// at segment ends).
static int h16_fold() {
  const char* v = getenv("PHF_H16_FOLD");
  const int f = v ? atoi(v) : 0;
  return f > 0 ? f : 8;
}

// Tile configurations of the reduced-precision kernels; the 256 x 256 int8
// tile serves I8X4 only, the 256 x 256 round-to-nearest tile (rn16) the
// 16-bit forms.  fold64: false for the ones that fold into an FP32 master tile
// by truncating MFMA accumulation (the int8 tile accumulates exactly, the
// rn16 tile by round-to-nearest adds of short partials).
struct H16CfgDesc {
  int bm, bn, wm, wn, stages, minb;
  bool fold64, i8big, rn16;
};
static const H16CfgDesc kH16Cfgs[] = {
    {128, 128, 64, 64, 2, 2, true, false, false},  {256, 128, 64, 64, 3, 1, true, false, false},
    {256, 128, 128, 64, 3, 1, false, false, false}, {128, 256, 64, 128, 3, 1, false, false, false},
    {256, 256, 128, 64, 2, 1, true, true, false},   {256, 256, 128, 64, 2, 1, true, false, true},
};
// Longest contraction of the int8 tile: class 3 sums four products of slices
// in [-127, 127] per element, 4 * 127^2 * K < 2^31.
static const int I8_TILE_MAX_K = 33285;
// Longest contraction of the rn16 tile: its FP32 accumulation error grows with
// the contraction (measured |dE| 3.6e-7 at K = 8192, n = 64; 2.8e-6 at
// K = 32768, n = 128, where the FP64-fold tiles give 3.0e-7 and 1.3e-6).
static const int RN16_TILE_MAX_K = 8192;
// Largest NSO of the rn16 tile: the 16-bit products' truncation bias within
// each 64-product partial grows with NSO^2 (F16X2: |dE| 4.0e-7 at n = 64,
// NSO = 128; 1.34e-6 for an n = 80 instance of the 64-instance batch, whose
// contraction K = 6400 is inside the length limit).
static const int RN16_TILE_MAX_NSO = 128;
bool h16_cfg_rn16(int c) { return kH16Cfgs[c].rn16; }
int h16_rn16_max_nso() { return RN16_TILE_MAX_NSO; }
int h16_num_configs() { return (int)(sizeof kH16Cfgs / sizeof kH16Cfgs[0]); }
int h16_cfg_bm(int c) { return kH16Cfgs[c].bm; }
int h16_cfg_bn(int c) { return kH16Cfgs[c].bn; }
bool h16_cfg_fold64(int c) { return kH16Cfgs[c].fold64; }
// whether config c can run precision prec on contractions of length K
bool h16_cfg_serves(int c, int prec, int K) {
  if (kH16Cfgs[c].i8big) return prec == 4 && K <= I8_TILE_MAX_K;
  if (kH16Cfgs[c].rn16) return (prec == 2 || prec == 3) && K <= RN16_TILE_MAX_K;
  return true;
}
// PHF_H16_CFG=<index>[,<splitk>]: the one configuration (and split-K) the
// tuners may use; -1 / 0 when unset
int h16_forced_config() {
  const char* v = getenv("PHF_H16_CFG");
  if (!v || !*v) return -1;
  const int c = atoi(v);
  return (c >= 0 && c < h16_num_configs()) ? c : -1;
}
int h16_forced_splitk() {
  const char* v = getenv("PHF_H16_CFG");
  const char* comma = v ? strchr(v, ',') : nullptr;
  if (!comma) return 0;
  const int sk = atoi(comma + 1);
  return (sk == 1 || sk == 2) ? sk : 0;
}
// slabs of the whole contraction of length K (what split-K divides)
int h16_slabs(int prec, int c, int K) {
  (void)c;
  return mixed_segments(prec) * (K / (128 / mixed_elem_bytes(prec)));
}

// Cost model used to order the autotuner's candidates (and to choose when it
// is off): MFMA rounds over the CUs for cfg c and split-K sk, counting padded
// rows, normalised to one CU's work per round.
double h16_model_cost(int c, int sk, long Mp0, int Nr0, int ncu) {
  const int bm = kH16Cfgs[c].bm, bn = kH16Cfgs[c].bn, per_cu = kH16Cfgs[c].minb;
  const long mp = (Mp0 + bm - 1) / bm * bm;
  const long np = (Nr0 + bn - 1) / bn * bn;
  const long units = (mp / bm) * (np / bn) * sk;
  const long rounds = (units + (long)per_cu * ncu - 1) / ((long)per_cu * ncu);
  return (double)rounds * per_cu * bm * bn / sk;
}

// Mp multiple of the config's BM, Nr multiple of its BN; K (the contraction
// length) a multiple of one 128-byte slab times splitk (64 16-bit, 32 FP32 or
// 128 int8 elements).  P rows hold their
// slices kp elements apart (kp >= K; the contraction reads the first K of
// each), A rows are [nslice][K].  With splitk = 2 the caller zero-fills C
// first.  colmap: optional A row -> C column map (-1: not written).
void eri_gemm_mixed(hipStream_t s, int prec, const void* P, long kp, const void* A, const double* invP,
                    const double* invA, double* C, int Mp, int Nr, int K, long ldc, int cfg, int splitk,
                    const double* bias, const int* colmap) {
  unsigned patP, patA, patW;
  mixed_patterns(prec, &patP, &patA, &patW);
  const int nslice = mixed_slices(prec), nseg = mixed_segments(prec);
  const long ldP = nslice * kp, ldA = (long)nslice * K;
  const int bm = kH16Cfgs[cfg].bm, bn = kH16Cfgs[cfg].bn;
  const int tm = Mp / bm, tn = Nr / bn;
  const unsigned grid = (unsigned)(tm * tn * splitk);
  if (kH16Cfgs[cfg].rn16) {
#define RN_LAUNCH(TT)                                                                                           \
  do {                                                                                                         \
    auto kern = splitk == 2 ? eri_gemm_rn16_kernel<TT, 2> : eri_gemm_rn16_kernel<TT, 1>;                        \
    static bool attr_set = false;                                                                              \
    if (!attr_set) {                                                                                           \
      (void)hipFuncSetAttribute((const void*)eri_gemm_rn16_kernel<TT, 1>,                                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);                      \
      (void)hipFuncSetAttribute((const void*)eri_gemm_rn16_kernel<TT, 2>,                                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);                      \
      attr_set = true;                                                                                         \
    }                                                                                                          \
    kern<<<grid, 512, I8_TILE_LDS, s>>>((const TT*)P, ldP, kp, (const TT*)A, ldA, (long)K, invP, invA, C, K, nseg, \
                                        patP, patA, ldc, tm, bias, colmap);                                     \
  } while (0)
    if (prec == 3) RN_LAUNCH(_Float16); else RN_LAUNCH(__bf16);
#undef RN_LAUNCH
    return;
  }
  if (kH16Cfgs[cfg].i8big) {
    mixed_patterns_desc(&patP, &patA, &patW);
    auto kern = splitk == 2 ? eri_gemm_i8_kernel<2> : eri_gemm_i8_kernel<1>;
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void*)eri_gemm_i8_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                I8_TILE_LDS);
      (void)hipFuncSetAttribute((const void*)eri_gemm_i8_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                I8_TILE_LDS);
      attr_set = true;
    }
    kern<<<grid, 512, I8_TILE_LDS, s>>>((const signed char*)P, ldP, kp, (const signed char*)A, ldA, (long)K, invP,
                                        invA, C, K, nseg, patP, patA, patW, ldc, tm, bias, colmap);
    return;
  }
#define H_LAUNCH(TT, B0, B1, W0, W1, ST, MB, F64, SK)                                                     \
  do {                                                                                                       \
    typedef H16Cfg<B0, B1, W0, W1, ST, MB, F64> CC;                                                          \
    auto kern = eri_gemm_h16_kernel<TT, CC, B0, B1, W0, W1, SK>;                                             \
    static bool attr_set = false;                                                                            \
    if (!attr_set) {                                                                                         \
      (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, CC::LDS);     \
      attr_set = true;                                                                                       \
    }                                                                                                        \
    kern<<<grid, CC::NT, CC::LDS, s>>>((const TT*)P, ldP, kp, (const TT*)A, ldA, (long)K, invP, invA, C, K,   \
                                       nseg, patP, patA, patW, ldc, tm, bias, colmap, h16_fold());           \
  } while (0)
#define H_CFGS(TT, SK)                                                          \
  do {                                                                          \
    switch (cfg) {                                                              \
      case 0: H_LAUNCH(TT, 128, 128, 64, 64, 2, 2, true, SK); break;     \
      case 1: H_LAUNCH(TT, 256, 128, 64, 64, 3, 1, true, SK); break;     \
      case 2: H_LAUNCH(TT, 256, 128, 128, 64, 3, 1, false, SK); break;   \
      default: H_LAUNCH(TT, 128, 256, 64, 128, 3, 1, false, SK); break;  \
    }                                                                           \
  } while (0)
  if (prec == 1) {
    if (splitk == 2) H_CFGS(float, 2); else H_CFGS(float, 1);
  } else if (prec == 2) {
    if (splitk == 2) H_CFGS(__bf16, 2); else H_CFGS(__bf16, 1);
  } else if (prec == 3) {
    if (splitk == 2) H_CFGS(_Float16, 2); else H_CFGS(_Float16, 1);
  } else {
    if (splitk == 2) H_CFGS(signed char, 2); else H_CFGS(signed char, 1);
  }
#undef H_CFGS
#undef H_LAUNCH
}

// Workgroups of the persistent grouped launch for config cfg: one per CU
// times the workgroups per CU the config's registers and LDS allow.
int h16_grouped_grid(int cfg, int ncu) { return ncu * kH16Cfgs[cfg].minb; }

// Items reference operands in the layout of eri_gemm_mixed (slices of length
// it.K); every item's tile lies inside its operands.  atomic: every item adds
// into a C the caller zero-filled (split contractions); else every item stores.
void eri_gemm_mixed_grouped(hipStream_t s, int prec, const H16Item* items, int nitems, int cfg, int grid,
                            bool atomic) {
  unsigned patP, patA, patW;
  mixed_patterns(prec, &patP, &patA, &patW);
  if (kH16Cfgs[cfg].rn16) {
#define RNG_LAUNCH(TT)                                                                                          \
  do {                                                                                                         \
    static bool attr_set = false;                                                                              \
    if (!attr_set) {                                                                                           \
      (void)hipFuncSetAttribute((const void*)eri_gemm_rn16_grouped_kernel<TT, true>,                           \
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);                      \
      (void)hipFuncSetAttribute((const void*)eri_gemm_rn16_grouped_kernel<TT, false>,                          \
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);                      \
      attr_set = true;                                                                                         \
    }                                                                                                          \
    if (atomic)                                                                                                \
      eri_gemm_rn16_grouped_kernel<TT, true><<<grid, 512, I8_TILE_LDS, s>>>(items, nitems, patP, patA);         \
    else                                                                                                       \
      eri_gemm_rn16_grouped_kernel<TT, false><<<grid, 512, I8_TILE_LDS, s>>>(items, nitems, patP, patA);        \
  } while (0)
    if (prec == 3) RNG_LAUNCH(_Float16); else RNG_LAUNCH(__bf16);
#undef RNG_LAUNCH
    return;
  }
  if (kH16Cfgs[cfg].i8big) {
    mixed_patterns_desc(&patP, &patA, &patW);
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void*)eri_gemm_i8_grouped_kernel<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);
      (void)hipFuncSetAttribute((const void*)eri_gemm_i8_grouped_kernel<false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, I8_TILE_LDS);
      attr_set = true;
    }
    if (atomic)
      eri_gemm_i8_grouped_kernel<true><<<grid, 512, I8_TILE_LDS, s>>>(items, nitems, patP, patA, patW);
    else
      eri_gemm_i8_grouped_kernel<false><<<grid, 512, I8_TILE_LDS, s>>>(items, nitems, patP, patA, patW);
    return;
  }
#define G_LAUNCH(TT, B0, B1, W0, W1, ST, MB, F64)                                                         \
  do {                                                                                                       \
    typedef H16Cfg<B0, B1, W0, W1, ST, MB, F64> CC;                                                          \
    auto k_at = eri_gemm_h16_grouped_kernel<TT, CC, B0, B1, W0, W1, true>;                                   \
    auto k_st = eri_gemm_h16_grouped_kernel<TT, CC, B0, B1, W0, W1, false>;                                  \
    static bool attr_set = false;   /* once, before any capture of a launch */                               \
    if (!attr_set) {                                                                                         \
      (void)hipFuncSetAttribute((const void*)k_at, hipFuncAttributeMaxDynamicSharedMemorySize, CC::LDS);     \
      (void)hipFuncSetAttribute((const void*)k_st, hipFuncAttributeMaxDynamicSharedMemorySize, CC::LDS);     \
      attr_set = true;                                                                                       \
    }                                                                                                        \
    if (atomic)                                                                                              \
      k_at<<<grid, CC::NT, CC::LDS, s>>>(items, nitems, patP, patA, patW, h16_fold());                       \
    else                                                                                                     \
      k_st<<<grid, CC::NT, CC::LDS, s>>>(items, nitems, patP, patA, patW, h16_fold());                       \
  } while (0)
#define G_CFGS(TT)                                                             \
  do {                                                                         \
    switch (cfg) {                                                             \
      case 0: G_LAUNCH(TT, 128, 128, 64, 64, 2, 2, true); break;        \
      case 1: G_LAUNCH(TT, 256, 128, 64, 64, 3, 1, true); break;        \
      case 2: G_LAUNCH(TT, 256, 128, 128, 64, 3, 1, false); break;      \
      default: G_LAUNCH(TT, 128, 256, 64, 128, 3, 1, false); break;     \
    }                                                                          \
  } while (0)
  if (prec == 1)
    G_CFGS(float);
  else if (prec == 2)
    G_CFGS(__bf16);
  else if (prec == 3)
    G_CFGS(_Float16);
  else
    G_CFGS(signed char);
#undef G_CFGS
#undef G_LAUNCH
}
