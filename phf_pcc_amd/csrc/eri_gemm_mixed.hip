// Reduced-precision forms of the ERI GEMM (Gamma = A P for all grid points).
//
// Compensated split products on the matrix cores:
//   x = s^{-1} (x0 + x1 + x2),  x_i in FP16 or BF16,  s a power of two per row
// of A and per (grid point, re/im) row of P.  The kept terms of A P are laid
// out along the contraction index so one GEMM pass computes them all:
//   F16X2  (3 products)  A' = [a0 | a0 | a1],           P' = [p0 | p1 | p0]
//   BF16X3 (6 products)  A' = [a0 a0 a1 a0 a1 a2],       P' = [p0 p1 p0 p2 p1 p0]
// The MFMA accumulates in FP32; every FLUSH k-slabs the FP32 tile is folded
// into an FP64 accumulator, and the FP64 epilogue applies s_P^{-1} s_A^{-1}
// (exact powers of two) and writes Gamma in FP64 for the FP64 epilogue of the
// grid pipeline.  FP32 form: one v_mfma_f32_16x16x4f32 product of the FP32
// roundings.
//
// Layout contract as eri_gemm.hip: P' is [Mp][K'], A' is [Nr][K'], C is
// [Mp][ldc]; Mp % BM == 0, Nr % BN == 0, K' % BK == 0.
#include "phf_common.h"

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));

namespace {

template <typename T> struct MfmaH;
template <> struct MfmaH<_Float16> {
  typedef h8_t v8;
  __device__ static f4_t mma(v8 a, v8 b, f4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct MfmaH<__bf16> {
  typedef b8_t v8;
  __device__ static f4_t mma(v8 a, v8 b, f4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// 16-bit inputs, 16x16x32 MFMA.  Workgroup tile BM x 128 with BM/64 x 2 waves
// of 64 x 64 (two waves per SIMD at BM = 256), k-slab of 128 elements in one
// LDS stage (row stride 136 elements = 272 B) with the next slab prefetched
// into registers while the current one is consumed.  SPLITK = 2 halves the
// contraction per workgroup and adds the two FP64 partials with atomics (two
// commutative additions onto zero: the result does not depend on their order).
template <typename T, int BM, int SPLITK>
__global__ __launch_bounds__(BM * 2) void eri_gemm_h16_kernel(const T* __restrict__ P, const T* __restrict__ A,
                                                             const double* __restrict__ invP,
                                                             const double* __restrict__ invA,
                                                             double* __restrict__ C, int K, long ldc, int tiles_m) {
  constexpr int BN = 128, BK = 128, LDH = BK + 8;
  constexpr int NT = BM * 2;                  // threads
  constexpr int WM = 64, WN = 64, FM = 4, FN = 4;
  constexpr int ROWS = BM + BN;               // rows staged per slab
  constexpr int RPP = NT / 16;                // rows per pass: 16 threads x 16 B per 128-element row
  constexpr int NLD = ROWS / RPP;             // 16-byte loads per thread per slab
  static_assert(NLD <= 16 && NLD * RPP == ROWS, "staging registers");
  typedef typename MfmaH<T>::v8 v8;
  __shared__ __attribute__((aligned(16))) T sm[ROWS * LDH];   // rows [0, BM): P tile, [BM, BM+BN): A tile

  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int ks_id = t % SPLITK;
  const int tile = t / SPLITK;
  const int mt = tile % tiles_m, nt = tile / tiles_m;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int kpart = K / SPLITK;
  const int kbeg = ks_id * kpart;
  const int lr = tid >> 4, lc = (tid & 15) * 8;

  // global row pointer of staged row r: P rows of this tile, then A rows
  auto grow = [&](int r) -> const T* {
    return (r < BM) ? P + ((long)mt * BM + r) * K : A + ((long)nt * BN + (r - BM)) * K;
  };

  f4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  uint4 st0, st1, st2, st3, st4, st5, st6, st7, st8, st9, st10, st11, st12, st13, st14, st15;
#define H_LD(dst, e, k0) \
  if constexpr (NLD > (e)) dst = *reinterpret_cast<const uint4*>(grow(lr + RPP * (e)) + (k0) + lc)
#define H_ST(e, src) \
  if constexpr (NLD > (e)) *reinterpret_cast<uint4*>(&sm[(lr + RPP * (e)) * LDH + lc]) = src
#define H_GLOAD(k0)                                                               \
  do {                                                                            \
    H_LD(st0, 0, k0); H_LD(st1, 1, k0); H_LD(st2, 2, k0); H_LD(st3, 3, k0);       \
    H_LD(st4, 4, k0); H_LD(st5, 5, k0); H_LD(st6, 6, k0); H_LD(st7, 7, k0);       \
    H_LD(st8, 8, k0); H_LD(st9, 9, k0); H_LD(st10, 10, k0); H_LD(st11, 11, k0);   \
    H_LD(st12, 12, k0); H_LD(st13, 13, k0); H_LD(st14, 14, k0); H_LD(st15, 15, k0); \
  } while (0)
#define H_LSTORE()                                                                \
  do {                                                                            \
    H_ST(0, st0); H_ST(1, st1); H_ST(2, st2); H_ST(3, st3);                       \
    H_ST(4, st4); H_ST(5, st5); H_ST(6, st6); H_ST(7, st7);                       \
    H_ST(8, st8); H_ST(9, st9); H_ST(10, st10); H_ST(11, st11);                   \
    H_ST(12, st12); H_ST(13, st13); H_ST(14, st14); H_ST(15, st15);               \
  } while (0)

  const int nk = kpart / BK;
  H_GLOAD(kbeg);
  for (int kt = 0; kt < nk; ++kt) {
    H_LSTORE();
    __syncthreads();
    if (kt + 1 < nk) H_GLOAD(kbeg + (kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      v8 a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[i] = *reinterpret_cast<const v8*>(&sm[(wm * WM + i * 16 + (lane & 15)) * LDH + ks + 8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        b[j] = *reinterpret_cast<const v8*>(&sm[(BM + wn * WN + j * 16 + (lane & 15)) * LDH + ks + 8 * (lane >> 4)]);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MfmaH<T>::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#undef H_LD
#undef H_ST
#undef H_GLOAD
#undef H_LSTORE

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        const long col = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        const double v = (double)acc[i][j][r] * invP[row] * invA[col];
        if constexpr (SPLITK == 1)
          C[row * ldc + col] = v;
        else
          atomicAdd(&C[row * ldc + col], v);
      }
}

// FP32 inputs, v_mfma_f32_16x16x4f32, 128 x 128 tile, k-slab 32, FP64 fold every FLUSH slabs.
template <int FLUSH>
__global__ __launch_bounds__(256, 1) void eri_gemm_f32_kernel(const float* __restrict__ P, const float* __restrict__ A,
                                                              double* __restrict__ C, int K, long ldc, int tiles_m) {
  constexpr int BM = 128, BN = 128, BK = 32, LDF = BK + 4;
  constexpr int WM = 64, WN = 64, FM = 4, FN = 4;
  __shared__ __attribute__((aligned(16))) float sP[2][BM * LDF];
  __shared__ __attribute__((aligned(16))) float sA[2][BN * LDF];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int mt = t % tiles_m, nt = t / tiles_m;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const float* Pg = P + (long)mt * BM * K;
  const float* Ag = A + (long)nt * BN * K;
  const int lr = tid >> 3, lc = (tid & 7) * 4;   // 8 threads x 16 B per 32-element row

  f4_t acc[FM][FN];
  d4_t facc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
      facc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    }
  float4 rp0, rp1, rp2, rp3, ra0, ra1, ra2, ra3;
#define F_LD(dst, base, e, k0) dst = *reinterpret_cast<const float4*>((base) + (long)(lr + 32 * (e)) * K + (k0) + lc)
#define F_ST(arr, buf, e, src) *reinterpret_cast<float4*>(&arr[buf][(lr + 32 * (e)) * LDF + lc]) = src
#define F_GLOAD(k0)                                                                          \
  do {                                                                                       \
    F_LD(rp0, Pg, 0, k0); F_LD(rp1, Pg, 1, k0); F_LD(rp2, Pg, 2, k0); F_LD(rp3, Pg, 3, k0); \
    F_LD(ra0, Ag, 0, k0); F_LD(ra1, Ag, 1, k0); F_LD(ra2, Ag, 2, k0); F_LD(ra3, Ag, 3, k0); \
  } while (0)
#define F_LSTORE(buf)                                                                        \
  do {                                                                                       \
    F_ST(sP, buf, 0, rp0); F_ST(sP, buf, 1, rp1); F_ST(sP, buf, 2, rp2); F_ST(sP, buf, 3, rp3); \
    F_ST(sA, buf, 0, ra0); F_ST(sA, buf, 1, ra1); F_ST(sA, buf, 2, ra2); F_ST(sA, buf, 3, ra3); \
  } while (0)
  F_GLOAD(0);
  F_LSTORE(0);
  __syncthreads();
  const int nk = K / BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) F_GLOAD((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      float a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = sP[cur][(wm * WM + i * 16 + (lane & 15)) * LDF + ks + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = sA[cur][(wn * WN + j * 16 + (lane & 15)) * LDF + ks + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if constexpr (FLUSH > 0) {
      if ((kt + 1) % FLUSH == 0 || kt + 1 == nk) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) facc[i][j][r] += (double)acc[i][j][r];
            acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
          }
      }
    }
    if (kt + 1 < nk) F_LSTORE(cur ^ 1);
    __syncthreads();
  }
#undef F_LD
#undef F_ST
#undef F_GLOAD
#undef F_LSTORE
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + 4 * (lane >> 4) + r;
        const long col = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        C[row * ldc + col] = (FLUSH > 0) ? facc[i][j][r] : (double)acc[i][j][r];
      }
}

// ---------------------------------------------------------------------------
// operand conversion
// ---------------------------------------------------------------------------

template <typename T> __device__ inline T to16(double x);
template <> __device__ inline _Float16 to16<_Float16>(double x) { return (_Float16)(float)x; }
template <> __device__ inline __bf16 to16<__bf16>(double x) { return (__bf16)(float)x; }

// One workgroup per row: s = 2^(top - ceil(log2 max|x|)); slices x_i of s*x;
// dst row = concatenation of the slices named by `pattern` (3 bits per entry).
template <typename T>
__global__ __launch_bounds__(256) void split_rows_kernel(const double* __restrict__ src, long lds, int kt,
                                                         T* __restrict__ dst, long ldd, int nseg,
                                                         unsigned pattern, int top, double* __restrict__ inv) {
  __shared__ double red[256];
  const long row = blockIdx.x;
  const double* x = src + row * lds;
  double m = 0.0;
  for (int c = threadIdx.x; c < kt; c += 256) m = fmax(m, fabs(x[c]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  m = red[0];
  int e = 0;
  if (m > 0.0) {
    frexp(m, &e);             // m = f * 2^e with f in [0.5, 1)
  }
  const double sc = (m > 0.0) ? ldexp(1.0, top - e) : 1.0;
  if (threadIdx.x == 0) inv[row] = 1.0 / sc;
  T* d = dst + row * ldd;
  for (int c = threadIdx.x; c < kt; c += 256) {
    double r = x[c] * sc;
    T sl[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      sl[q] = to16<T>(r);
      r -= (double)(float)sl[q];
    }
    for (int sg = 0; sg < nseg; ++sg) d[(long)sg * kt + c] = sl[(pattern >> (3 * sg)) & 7];
  }
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

// precision modes (values of PHF_PREC_*)
int mixed_segments(int prec) { return prec == 3 ? 3 : (prec == 2 ? 6 : 1); }

void split_operand(hipStream_t s, int prec, bool is_a, const double* src, long lds, int rows, int kt, void* dst,
                   double* inv) {
  static const int a_f16[3] = {0, 0, 1}, p_f16[3] = {0, 1, 0};
  static const int a_b16[6] = {0, 0, 1, 0, 1, 2}, p_b16[6] = {0, 1, 0, 2, 1, 0};
  const int nseg = mixed_segments(prec);
  const long ldd = (long)nseg * kt;
  if (prec == 3)
    split_rows_kernel<_Float16><<<rows, 256, 0, s>>>(src, lds, kt, (_Float16*)dst, ldd, nseg,
                                                     pack_pattern(is_a ? a_f16 : p_f16, 3), 14, inv);
  else
    split_rows_kernel<__bf16><<<rows, 256, 0, s>>>(src, lds, kt, (__bf16*)dst, ldd, nseg,
                                                   pack_pattern(is_a ? a_b16 : p_b16, 6), 14, inv);
}

void convert_f32(hipStream_t s, const double* src, float* dst, long n) {
  long blocks = std::min<long>((n + 255) / 256, 65536);
  to_f32_kernel<<<blocks, 256, 0, s>>>(src, dst, n);
}

// Tile height (grid-point rows) and split-K of the 16-bit kernel for a problem
// of Mp0 rows, Nr columns on ncu compute units: fewest padded MFMA cycles per
// CU, counting whole rounds of workgroups (one 256-row or two 128-row
// workgroups per CU).
void mixed_tile_choice(long Mp0, int Nr, int ncu, int* bm, int* splitk) {
  double best = 1e300;
  for (int b : {128, 256})
    for (int sk : {1, 2}) {
      const long mp = (Mp0 + b - 1) / b * b;
      const long units = (mp / b) * (Nr / 128) * sk;
      const int per_round = (b == 256) ? ncu : 2 * ncu;
      const long rounds = (units + per_round - 1) / per_round;
      const double cost = (double)rounds * b / sk * (b == 256 ? 1.0 : 0.5) + (sk == 2 ? 0.02 * rounds : 0.0);
      if (cost < best - 1e-9) { best = cost; *bm = b; *splitk = sk; }
    }
}

// Mp multiple of bm, Nr multiple of 128; Kseg = nseg * kt multiple of 128 * splitk
// (16-bit) or 32 (FP32).  With splitk = 2 the caller zero-fills C first.
void eri_gemm_mixed(hipStream_t s, int prec, const void* P, const void* A, const double* invP, const double* invA,
                    double* C, int Mp, int Nr, int Kseg, long ldc, int bm, int splitk) {
  if (prec == 1) {
    const int tm = Mp / 128, tn = Nr / 128;
    eri_gemm_f32_kernel<16><<<tm * tn, 256, 0, s>>>((const float*)P, (const float*)A, C, Kseg, ldc, tm);
    return;
  }
  const int tm = Mp / bm, tn = Nr / 128;
  const unsigned grid = (unsigned)(tm * tn * splitk);
#define H_LAUNCH(TT, BMV, SK) \
  eri_gemm_h16_kernel<TT, BMV, SK><<<grid, BMV * 2, 0, s>>>((const TT*)P, (const TT*)A, invP, invA, C, Kseg, ldc, tm)
  if (prec == 2) {
    if (bm == 256) { if (splitk == 2) H_LAUNCH(__bf16, 256, 2); else H_LAUNCH(__bf16, 256, 1); }
    else { if (splitk == 2) H_LAUNCH(__bf16, 128, 2); else H_LAUNCH(__bf16, 128, 1); }
  } else {
    if (bm == 256) { if (splitk == 2) H_LAUNCH(_Float16, 256, 2); else H_LAUNCH(_Float16, 256, 1); }
    else { if (splitk == 2) H_LAUNCH(_Float16, 128, 2); else H_LAUNCH(_Float16, 128, 1); }
  }
#undef H_LAUNCH
}
