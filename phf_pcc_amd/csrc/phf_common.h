// Shared definitions for the spin-projection grid kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

#define PHF_WAVE 64

// MFMA operand/result lane layouts on gfx950 (checked on the device):
//   f64 16x16x4  : A[i][k] lane i + 16k;  B[k][j] lane j + 16k;
//                  D[i][j] lane j + 16*(i % 4), register i / 4
//   f32 16x16x4, bf16 16x16x32, i8 16x16x64 : A[i][k] lane i + 16*(k / kpl),
//                  element k % kpl (kpl = 1, 8, 16 values per lane);
//                  D[i][j] lane j + 16*(i / 4), register i % 4

// Strided complex operand: element (i, j) of batch b has its real part at
// p[(b / bdiv)*bs + (b % bdiv)*bs2 + i*rs + j*cs] and its imaginary part im
// doubles further on (im = 1 for interleaved complex128, im = plane size for
// split planes).  bdiv = 1, bs2 = 0 (the defaults) give p[b*bs + ...]; a
// batch over pairs (outer, inner) with inner < bdiv strides each separately.
struct ZMat {
  const double* p;
  long rs, cs, bs, im;
  int conj;
  long bs2 = 0;
  int bdiv = 1;
  __host__ __device__ long boff(long b) const { return (b / bdiv) * bs + (b % bdiv) * bs2; }
};

struct ZOut {
  double* p;
  long rs, cs, bs, im;
  // optional: per batch entry b, amax[2 b] / amax[2 b + 1] receive (atomic max
  // of the bit patterns) the largest |Re| / |Im| written, for the row scales
  // of the reduced-precision split; zero-filled by the caller
  unsigned long long* amax = nullptr;
  long bs2 = 0;
  int bdiv = 1;
  __host__ __device__ long boff(long b) const { return (b / bdiv) * bs + (b % bdiv) * bs2; }
};

// Host-side helpers for interleaved complex128 storage.
inline ZMat zmat_colmajor(const double* p, long ld, long bs, int conj = 0) {
  return ZMat{p, 2, 2 * ld, bs, 1, conj};
}
// op(X) = X^T of a column-major X
inline ZMat zmat_colmajor_t(const double* p, long ld, long bs, int conj = 0) {
  return ZMat{p, 2 * ld, 2, bs, 1, conj};
}
inline ZOut zout_colmajor(double* p, long ld, long bs) { return ZOut{p, 2, 2 * ld, bs, 1}; }
// the transpose of a column-major output (element (i, j) at p[2 (j + ld i)]):
// a product C^T = B^T A^T written through it lands where C would, with the
// stores running along the contiguous index
inline ZOut zout_colmajor_t(double* p, long ld, long bs) { return ZOut{p, 2 * ld, 2, bs, 1}; }

// Launch batched C = alpha * op(A) op(B) + beta * C (complex FP64) on stream s.
// m x k times k x n; `batch` problems (blockIdx.z).
void zgemm_batched(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B,
                   double br, double bi, ZOut C, int batch);
// The same product; for small outputs (m, n <= 32) with long contractions
// (k >= 512) the contraction is split into d slices (d divides k, batch d <=
// 4096, batch d m n <= ws_cplx) whose partials go to ws (complex elements)
// and are summed in a fixed order; A and B must not use two-level batching.
void zgemm_batched_splitk(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br,
                          double bi, ZOut C, int batch, double* ws, long ws_cplx);
// The same product with the contraction split into d slices (d divides k;
// batch d m n <= ws_cplx complex elements of ws), for products too few to
// fill the GPU: d times the workgroups, partials summed in a fixed order.  A
// and B must not use two-level batching.
void zgemm_batched_kslices(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br,
                           double bi, ZOut C, int batch, int d, double* ws, long ws_cplx);

// One output tile (or one part of a split contraction) of a 16-bit ERI GEMM in
// a grouped launch: rows mt*BM.. of P against rows nt*BN.. of A over the slabs
// [sbeg, send) (64 elements each) of the segment sequence.  P rows are ldP
// elements apart with their slices sP apart, A rows ldA apart with slices sA
// apart; K is the contraction length; C rows are ldc apart (also the stride
// of the bias planes); colmap (optional) sends A row r to C column colmap[r].
struct H16Item {
  const void* P;
  const void* A;
  const double* invP;
  const double* invA;
  double* C;
  const double* bias;   // null, or [2][ldc] added to row parity planes (first part only)
  const int* colmap;
  long ldP, sP, ldA, sA;
  long ldc;
  int K;                // 0: an empty slot of the work list (skipped)
  int mt, nt;
  int sbeg, send;
  int atomic;           // 1: add into C (zero-filled by the caller)
  // epilogue of the spin-free form's exchange sets (epi = 1): C[colmap[r]] = s1 v (+ bias)
  int epi;
  double s1;
};

// The spin-free epilogue of a work item (see H16Item).
struct Epi {
  int on = 0;
  double s1 = 1.0;
};

__host__ __device__ inline Epi item_epi(const H16Item& it) {
  Epi e;
  e.on = it.epi;
  e.s1 = it.s1;
  return e;
}

// XCD-aware remap of a linear workgroup id: workgroups are dispatched
// round-robin over the 8 XCDs; this gives each XCD a contiguous range of
// tile ids so that tiles sharing an operand panel share that XCD's L2.
__device__ inline int xcd_remap(int bid, int nwg) {
  const int nx = 8;
  int q = nwg / nx, r = nwg % nx;
  int x = bid % nx, loc = bid / nx;
  return (x < r) ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
}
