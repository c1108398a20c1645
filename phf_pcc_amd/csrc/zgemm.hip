// Batched complex FP64 GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b] for arbitrary element
// strides, so transposes, conjugate transposes, row-major host layouts and
// split re/im planes are all expressed through ZMat/ZOut without copies.
// Used for the O(NSO^2 N) per-grid-point stages: the rotation chain
// (Ref^H R1 R2 Rpg Ref C0), the N x N overlap C0^H T, the solve T M^{-1}, the
// transition density rho^T = C0* X^T, the Fock contraction (h + Gamma) X and
// the final Y C0^H.
//
// Tile: (32 FMW) x (32 FNW) per 256-thread workgroup, 2 x 2 waves of
// (16 FMW) x (16 FNW), k-slab 16; the launcher takes FMW/FNW = 1 for sides of
// at most 32 (the N = nocc sides) so no MFMA work is spent on padding.
// Complex product = three real MFMAs (Gauss): P1 = sum ar br, P2 = sum ai bi,
// P3 = sum (ar + ai)(br + bi), then re = P1 - P2, im = P3 - P1 - P2.  The
// imaginary part's rounding error is bounded by eps (|ar| + |ai|)(|br| + |bi|)
// per term instead of eps (|ar bi| + |ai br|): the same order for FP64.
#include <numeric>

#include "phf_common.h"

namespace {

constexpr int ZBK = 16, ZLD = ZBK + 2;

template <int FMW, int FNW>
__global__ __launch_bounds__(256) void zgemm_kernel(int m, int n, int k, double ar, double ai,
                                                    ZMat A, ZMat B, double br, double bi, ZOut C,
                                                    int a_kfast, int b_kfast) {
  constexpr int ZBM = 32 * FMW, ZBN = 32 * FNW;
  __shared__ double As_re[ZBM * ZLD], As_im[ZBM * ZLD];
  __shared__ double Bs_re[ZBN * ZLD], Bs_im[ZBN * ZLD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long b = blockIdx.z;
  const int m0 = blockIdx.x * ZBM, n0 = blockIdx.y * ZBN;
  const double* Ap = A.p + A.boff(b);
  const double* Bp = B.p + B.boff(b);
  const long cb = C.boff(b);

  d4_t p1[FMW][FNW], p2[FMW][FNW], p3[FMW][FNW];
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < FNW; ++j) {
      p1[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
      p2[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
      p3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    }

  // Global loads go to registers first: every load of a slab is issued
  // before any is used (addresses clamped to element (0,0) and the value
  // zeroed when out of range, so no branch separates them), and slab s+1 is
  // loaded while the MFMAs of slab s run.
  constexpr int EA = (ZBM * ZBK) / 256, EB = (ZBN * ZBK) / 256;
  double lar[EA], lai[EA], lbr[EB], lbi[EB];
  bool va[EA], vb[EB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + 256 * e;
      int i, kk;
      if (a_kfast) { kk = idx % ZBK; i = idx / ZBK; }
      else { i = idx % ZBM; kk = idx / ZBM; }
      const int gi = m0 + i, gk = k0 + kk;
      va[e] = gi < m && gk < k;
      const long off = va[e] ? (long)gi * A.rs + (long)gk * A.cs : 0;
      lar[e] = Ap[off];
      lai[e] = Ap[off + A.im];
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + 256 * e;
      int j, kk;
      if (b_kfast) { kk = idx % ZBK; j = idx / ZBK; }
      else { j = idx % ZBN; kk = idx / ZBN; }
      const int gj = n0 + j, gk = k0 + kk;
      vb[e] = gj < n && gk < k;
      const long off = vb[e] ? (long)gk * B.rs + (long)gj * B.cs : 0;
      lbr[e] = Bp[off];
      lbi[e] = Bp[off + B.im];
    }
  };
  if (k > 0) gload(0);
  for (int k0 = 0; k0 < k; k0 += ZBK) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + 256 * e;
      int i, kk;
      if (a_kfast) { kk = idx % ZBK; i = idx / ZBK; }
      else { i = idx % ZBM; kk = idx / ZBM; }
      As_re[i * ZLD + kk] = va[e] ? lar[e] : 0.0;
      As_im[i * ZLD + kk] = va[e] ? (A.conj ? -lai[e] : lai[e]) : 0.0;
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + 256 * e;
      int j, kk;
      if (b_kfast) { kk = idx % ZBK; j = idx / ZBK; }
      else { j = idx % ZBN; kk = idx / ZBN; }
      Bs_re[j * ZLD + kk] = vb[e] ? lbr[e] : 0.0;
      Bs_im[j * ZLD + kk] = vb[e] ? (B.conj ? -lbi[e] : lbi[e]) : 0.0;
    }
    __syncthreads();
    if (k0 + ZBK < k) gload(k0 + ZBK);
#pragma unroll
    for (int ks = 0; ks < ZBK; ks += 4) {
      double are[FMW], aim[FMW], bre[FNW], bim[FNW];
#pragma unroll
      for (int t = 0; t < FMW; ++t) {
        const int row = wm * 16 * FMW + t * 16 + (lane & 15);
        are[t] = As_re[row * ZLD + ks + (lane >> 4)];
        aim[t] = As_im[row * ZLD + ks + (lane >> 4)];
      }
#pragma unroll
      for (int t = 0; t < FNW; ++t) {
        const int col = wn * 16 * FNW + t * 16 + (lane & 15);
        bre[t] = Bs_re[col * ZLD + ks + (lane >> 4)];
        bim[t] = Bs_im[col * ZLD + ks + (lane >> 4)];
      }
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int j = 0; j < FNW; ++j) {
          p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i], bre[j], p1[i][j], 0, 0, 0);
          p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[i], bim[j], p2[i][j], 0, 0, 0);
          p3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[i] + aim[i], bre[j] + bim[j], p3[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }

  const bool use_beta = (br != 0.0 || bi != 0.0);
  double mre = 0.0, mim = 0.0;
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < FNW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm * 16 * FMW + i * 16 + (lane >> 4) + 4 * r;   // f64 16x16x4: row = lane/16 + 4r
        const int gj = n0 + wn * 16 * FNW + j * 16 + (lane & 15);
        if (gi < m && gj < n) {
          const double xr = p1[i][j][r] - p2[i][j][r], xi = p3[i][j][r] - p1[i][j][r] - p2[i][j][r];
          double vr = ar * xr - ai * xi, vi = ar * xi + ai * xr;
          const long off = cb + (long)gi * C.rs + (long)gj * C.cs;
          if (use_beta) {
            const double orr = C.p[off], oi = C.p[off + C.im];
            vr += br * orr - bi * oi;
            vi += br * oi + bi * orr;
          }
          C.p[off] = vr;
          C.p[off + C.im] = vi;
          mre = fmax(mre, fabs(vr));
          mim = fmax(mim, fabs(vi));
        }
      }
  if (C.amax) {   // non-negative doubles order like their bit patterns
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mre = fmax(mre, __shfl_xor(mre, off));
      mim = fmax(mim, __shfl_xor(mim, off));
    }
    if (lane == 0) {
      atomicMax(&C.amax[2 * b], (unsigned long long)__double_as_longlong(mre));
      atomicMax(&C.amax[2 * b + 1], (unsigned long long)__double_as_longlong(mim));
    }
  }
}

// C[b] = alpha sum_s W[b d + s] + beta C[b] for m x n row-major partials W
// (fixed summation order).  grid (ceil(m n / 256), batch).
__global__ void __launch_bounds__(256) zsplit_reduce_kernel(int m, int n, int d, double ar, double ai, double br,
                                                            double bi, const double* __restrict__ W, ZOut C) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  const long b = blockIdx.y;
  const int i = idx / n, j = idx % n;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < d; ++q) {
    const double* w = W + 2 * ((b * d + q) * (long)m * n + idx);
    sr += w[0];
    si += w[1];
  }
  const long off = C.boff(b) + (long)i * C.rs + (long)j * C.cs;
  double vr = ar * sr - ai * si, vi = ar * si + ai * sr;
  if (br != 0.0 || bi != 0.0) {
    const double orr = C.p[off], oi = C.p[off + C.im];
    vr += br * orr - bi * oi;
    vi += br * oi + bi * orr;
  }
  C.p[off] = vr;
  C.p[off + C.im] = vi;
}

}  // namespace

void zgemm_batched_splitk(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br,
                          double bi, ZOut C, int batch, double* ws, long ws_cplx) {
  int d = 1;
  if (ws && m <= 32 && n <= 32 && k >= 512 && A.bdiv == 1 && B.bdiv == 1)
    for (int q = 32; q > 1; --q)
      if (k % q == 0 && k / q >= 128 && (long)batch * q <= 4096 && (long)batch * q * m * n <= ws_cplx) {
        d = q;
        break;
      }
  zgemm_batched_kslices(s, m, n, k, ar, ai, A, B, br, bi, C, batch, d, ws, ws_cplx);
}

void zgemm_batched_kslices(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br,
                           double bi, ZOut C, int batch, int d, double* ws, long ws_cplx) {
  if (d <= 1 || !ws || k % d != 0 || A.bdiv != 1 || B.bdiv != 1 || (long)batch * d > 65535 ||
      (long)batch * d * m * n > ws_cplx) {
    zgemm_batched(s, m, n, k, ar, ai, A, B, br, bi, C, batch);
    return;
  }
  const int ks = k / d;
  ZMat a = A, bb = B;
  a.bs2 = (long)ks * A.cs;   // entry b d + q: batch b, contraction slice q
  a.bdiv = d;
  bb.bs2 = (long)ks * B.rs;
  bb.bdiv = d;
  zgemm_batched(s, m, n, ks, 1, 0, a, bb, 0, 0, ZOut{ws, 2L * n, 2, 2L * m * n, 1}, batch * d);
  dim3 grid((unsigned)((m * n + 255) / 256), (unsigned)batch);
  zsplit_reduce_kernel<<<grid, 256, 0, s>>>(m, n, d, ar, ai, br, bi, ws, C);
}

void zgemm_batched(hipStream_t s, int m, int n, int k, double ar, double ai, ZMat A, ZMat B,
                   double br, double bi, ZOut C, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (batch > 65535) {   // grid.z limit: slices whose starts are multiples of every operand's bdiv
    const long l1 = std::lcm((long)A.bdiv, (long)B.bdiv), L = std::lcm(l1, (long)C.bdiv);
    const long step = 65535 - 65535 % L;   // callers keep every bdiv (a grid dimension) below 65536
    if (step <= 0) return;
    for (long b0 = 0; b0 < batch; b0 += step) {
      ZMat a = A, bb = B;
      ZOut c = C;
      a.p += A.boff(b0);
      bb.p += B.boff(b0);
      c.p += C.boff(b0);
      zgemm_batched(s, m, n, k, ar, ai, a, bb, br, bi, c, (int)std::min<long>(step, batch - b0));
    }
    return;
  }
  if (k <= 0) {
    // C = beta * C handled by a k=0 product of zeros
    k = 0;
  }
  const int a_kfast = (labs(A.cs) < labs(A.rs)) ? 1 : 0;
  const int b_kfast = (labs(B.rs) < labs(B.cs)) ? 1 : 0;
  const int fmw = (m <= 32) ? 1 : 2, fnw = (n <= 32) ? 1 : 2;
  dim3 grid((m + 32 * fmw - 1) / (32 * fmw), (n + 32 * fnw - 1) / (32 * fnw), batch);
  if (fmw == 2 && fnw == 2)
    zgemm_kernel<2, 2><<<grid, 256, 0, s>>>(m, n, k, ar, ai, A, B, br, bi, C, a_kfast, b_kfast);
  else if (fmw == 2)
    zgemm_kernel<2, 1><<<grid, 256, 0, s>>>(m, n, k, ar, ai, A, B, br, bi, C, a_kfast, b_kfast);
  else if (fnw == 2)
    zgemm_kernel<1, 2><<<grid, 256, 0, s>>>(m, n, k, ar, ai, A, B, br, bi, C, a_kfast, b_kfast);
  else
    zgemm_kernel<1, 1><<<grid, 256, 0, s>>>(m, n, k, ar, ai, A, B, br, bi, C, a_kfast, b_kfast);
}
