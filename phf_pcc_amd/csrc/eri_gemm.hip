// The ERI contraction of all grid points as one matrix-core GEMM.
//
// MakeGamma (PHFTools.f90:310-337) computes, per grid point,
//   Gamma_g(i,k) = sum_{j,l} V(i,j,k,l) rho_g(l,j).
// With A[(i,k),(j,l)] = V(i,j,k,l) (NSO^2 x NSO^2, cached on the device) and
// P[(j,l), g] = rho_g(l,j), all grid points together are Gamma = A P.  The
// kernel computes the transposed product
//   C[row][r] = sum_c Pt[row][c] * A[r][c]
// where row runs over (grid point, re/im) and both operands are stored with
// the contraction index c contiguous ("NT"), so one loader serves both.  For a
// complex A the contraction runs over [A_re | A_im] and the density rows are
// [P_re | -P_im] and [P_im | P_re].
//
// Layout contract: Pt is [Mp][ldp] (the contraction uses columns 0..K-1), A is
// [Nr][K], C is [Mp][ldc]; Mp % BM == 0, Nr % BN == 0, K % 16 == 0 (the host
// pads with zeros).  An optional bias [2][ldc] is added to even (real) and odd
// (imaginary) rows: the one-body h, so that C holds h + Gamma for the Fock
// contraction that follows.  An optional column map sends A row r to column
// colmap[r] of C (and of the bias), -1 for padding rows that are not written:
// a diagonal block of a permuted A writes its part of Gamma in place.
#include "phf_common.h"

namespace {

template <int BM, int BN>
__global__ __launch_bounds__(256, 2) void eri_gemm_f64_kernel(const double* __restrict__ P, long ldp,
                                                               const double* __restrict__ A,
                                                               double* __restrict__ C, int K, long ldc,
                                                               int tiles_m, const double* __restrict__ bias,
                                                               const int* __restrict__ colmap) {
  constexpr int BK = 16, LDSW = BK + 2;
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int PL = BM / 32, AL = BN / 32;
  __shared__ double sP[2][BM * LDSW];
  __shared__ double sA[2][BN * LDSW];

  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int mt = t % tiles_m, nt = t / tiles_m;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double* Pg = P + (long)mt * BM * ldp;
  const double* Ag = A + (long)nt * BN * K;
  const int lr = tid >> 3, lc = (tid & 7) * 2;

  d4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  static_assert(PL <= 4 && AL <= 4, "staging registers cover at most 4 rows of 32");
  // global -> registers (next k-slab) and registers -> LDS through named
  // registers: an array here is left in scratch memory by the compiler
  double2 rp0, rp1, rp2, rp3, ra0, ra1, ra2, ra3;
#define ERI_LD(dst, base, ld, e, k0) \
  dst = *reinterpret_cast<const double2*>((base) + (long)(lr + 32 * (e)) * (ld) + (k0) + lc)
#define ERI_ST(arr, buf, e, src) \
  *reinterpret_cast<double2*>(&arr[buf][(lr + 32 * (e)) * LDSW + lc]) = src
#define ERI_GLOAD(k0)                                   \
  do {                                                  \
    ERI_LD(rp0, Pg, ldp, 0, k0);                        \
    if constexpr (PL > 1) ERI_LD(rp1, Pg, ldp, 1, k0);  \
    if constexpr (PL > 2) ERI_LD(rp2, Pg, ldp, 2, k0);  \
    if constexpr (PL > 3) ERI_LD(rp3, Pg, ldp, 3, k0);  \
    ERI_LD(ra0, Ag, K, 0, k0);                          \
    if constexpr (AL > 1) ERI_LD(ra1, Ag, K, 1, k0);    \
    if constexpr (AL > 2) ERI_LD(ra2, Ag, K, 2, k0);    \
    if constexpr (AL > 3) ERI_LD(ra3, Ag, K, 3, k0);    \
  } while (0)
#define ERI_LSTORE(buf)                                 \
  do {                                                  \
    ERI_ST(sP, buf, 0, rp0);                            \
    if constexpr (PL > 1) ERI_ST(sP, buf, 1, rp1);      \
    if constexpr (PL > 2) ERI_ST(sP, buf, 2, rp2);      \
    if constexpr (PL > 3) ERI_ST(sP, buf, 3, rp3);      \
    ERI_ST(sA, buf, 0, ra0);                            \
    if constexpr (AL > 1) ERI_ST(sA, buf, 1, ra1);      \
    if constexpr (AL > 2) ERI_ST(sA, buf, 2, ra2);      \
    if constexpr (AL > 3) ERI_ST(sA, buf, 3, ra3);      \
  } while (0)

  ERI_GLOAD(0);
  ERI_LSTORE(0);
  __syncthreads();
  const int nk = K / BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) ERI_GLOAD((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = sP[cur][(wm * WM + i * 16 + (lane & 15)) * LDSW + ks + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = sA[cur][(wn * WN + j * 16 + (lane & 15)) * LDSW + ks + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) ERI_LSTORE(cur ^ 1);
    __syncthreads();
  }
#undef ERI_GLOAD
#undef ERI_LSTORE
#undef ERI_LD
#undef ERI_ST

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = (long)mt * BM + wm * WM + i * 16 + (lane >> 4) + 4 * r;   // f64 16x16x4 output row
        long col = (long)nt * BN + wn * WN + j * 16 + (lane & 15);
        if (colmap) {
          col = colmap[col];
          if (col < 0) continue;
        }
        C[row * ldc + col] = acc[i][j][r] + (bias ? bias[(row & 1) * ldc + col] : 0.0);
      }
}

}  // namespace

// Tile height for the grid-point dimension: the host pads Mp to this.
int eri_gemm_f64_bm(int rows) { return rows <= 64 * 12 ? 64 : 128; }

void eri_gemm_f64(hipStream_t s, const double* P, long ldp, const double* A, double* C, int Mp, int Nr, int K,
                  long ldc, int bm, const double* bias, const int* colmap) {
  if (bm == 64) {
    const int tm = Mp / 64, tn = Nr / 128;
    eri_gemm_f64_kernel<64, 128><<<tm * tn, 256, 0, s>>>(P, ldp, A, C, K, ldc, tm, bias, colmap);
  } else {
    const int tm = Mp / 128, tn = Nr / 128;
    eri_gemm_f64_kernel<128, 128><<<tm * tn, 256, 0, s>>>(P, ldp, A, C, K, ldc, tm, bias, colmap);
  }
}
