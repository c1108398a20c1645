// C ABI and host orchestration of the spin-projection grid pipeline.
//
// Per call of phf_buildhsg (BuildHSG, PHFTools.f90:10-201), with grid points
// g = (il, iy, ipg), g = (il*ny + iy)*npg + ipg (the reference's loop order):
//   K1  rotation chain   T_g = Ref^H R1[il] R2[iy] Rpg[ipg] Ref C0   (zgemm, right to left)
//       overlap          M_g = C0^H T_g
//   K2  inverse + det    M_g^{-1}, d_g = det M_g                     (lu_inverse_batched)
//       solve            X_g = T_g M_g^{-1},  rho_g = X_g C0^H
//       density pack     Pt rows (g,re),(g,im) = rho_g^T            (zgemm, split planes)
//   K3  ERI GEMM         Gamma_g = A rho_g for all g at once         (eri_gemm_*)
//   K4  epilogue         F = h + Gamma_g; FX = F X_g; Q = C0^H FX;
//                        e_g = (tr(C0^H h X) + tr Q)/2 = Makeh;
//                        Y_g = FX - X_g Q + e_g X_g, so that the reference's
//                        Gradient (I-rho)(h+Gamma)rho + e rho = Y_g C0^H
//   K5  grid sum         Wd = w * Wigner * weightpg * d_g;
//                        Hmat = sum Wd e, Smat = sum Wd, Ysum = sum Wd Y,
//                        Xsum = sum Wd X (fixed order, deterministic);
//                        Gradmat = Ysum C0^H, Rdmmat = Xsum C0^H.
// The MPI reduce/bcast of the reference (PHFTools.f90:181-188) has no
// counterpart on one GPU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "phf_internal.h"

static thread_local std::string g_create_err;
static int free_grid(phf_ctx* ctx);
static int drop_hamiltonian(phf_ctx* ctx);
static void drop_blocks(phf_ctx* ctx);

int phf_fail(phf_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

namespace {

typedef std::complex<double> cplx;

int fail(phf_ctx* ctx, int code, const std::string& msg) { return phf_fail(ctx, code, msg); }

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

__global__ void fill_zero_kernel(double* p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}


__global__ void imag_nonzero_kernel(const double* stage, long nelem, int* flag) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nelem; i += (long)gridDim.x * blockDim.x)
    if (stage[2 * i + 1] != 0.0) atomicOr(flag, 1);
}

// A[r][c'] for r < nrpad, c' < kt, from the C-order complex V(i,j,k,l):
// r = i*nso + k, c = j*nso + l, c' = half*kpad + c (half 1 = imaginary plane).
__global__ void ham_permute_kernel(const double* __restrict__ stage, double* __restrict__ A, int nso,
                                   int kpad, int kt, long nrpad) {
  const long n2 = (long)nso * nso;
  const long total = nrpad * kt;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / kt;
    const int cp = (int)(idx % kt);
    const int half = cp / kpad, c = cp % kpad;
    double v = 0.0;
    if (r < n2 && c < n2) {
      const long i = r / nso, k = r % nso, j = c / nso, l = c % nso;
      v = stage[((((i * nso + j) * nso + k) * nso) + l) * 2 + half];
    }
    A[idx] = v;
  }
}

// complex A: rows (g,re) = [P_re | -P_im], (g,im) = [P_im | P_re]
__global__ void pack_cplx_kernel(double* Pt, int G, int n2, int kpad, int kt) {
  const int g = blockIdx.y;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n2; c += gridDim.x * blockDim.x) {
    double* re = Pt + (long)(2 * g) * kt;
    double* im = Pt + (long)(2 * g + 1) * kt;
    const double pr = re[c], pi = im[c];
    re[kpad + c] = -pi;
    im[kpad + c] = pr;
  }
}

// Wd[v G + g][M][N] = W[g][M - d bm_v][N - d bn_v] det[v G + g] (zero outside
// the shifted block): variant v of the V density variants of a grid point
// lands in CI block (bm_v, bn_v) of the nci x nci matrices (V = 1: d = nci,
// block (0, 0)).  W holds the d x d weights at the top left of nci x nci.
__global__ void wd_kernel(int G, int V, int nci, int d, const double* W, const double* det, double* Wd) {
  const int nmn = nci * nci;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)V * G * nmn) return;
  const long gv = i / nmn;
  const int mn = (int)(i % nmn), M = mn / nci, N = mn % nci;
  const int v = (int)(gv / G), g = (int)(gv % G);
  // variants: 0 (L C0, B C0) -> (0,0); 1 (L C0*, B C0) -> (0,1); 2 (C0*, C0*) -> (1,1); 3 (C0, C0*) -> (1,0)
  const int bm = (v == 2 || v == 3) ? 1 : 0, bn = (v == 1 || v == 2) ? 1 : 0;
  const int m = M - d * bm, n = N - d * bn;
  double wr = 0.0, wi = 0.0;
  if (m >= 0 && m < d && n >= 0 && n < d) {
    wr = W[2 * ((long)g * nmn + m * nci + n)];
    wi = W[2 * ((long)g * nmn + m * nci + n) + 1];
  }
  const double dr = det[2 * gv], di = det[2 * gv + 1];
  Wd[2 * i] = wr * dr - wi * di;
  Wd[2 * i + 1] = wr * di + wi * dr;
}

// e_g = (sum_{a,p} Hc(a,p) X_g(p,a) + sum_a Q_g(a,a)) / 2; one workgroup per g.
// Hc(a, p) is stored at a nso + p, where X_g(p, a) is, so both reads are
// contiguous 16-byte loads.
__global__ __launch_bounds__(256) void energy_kernel(int nso, int nocc, const double* __restrict__ Hc,
                                                     const double* __restrict__ X, double* __restrict__ Q,
                                                     double* __restrict__ E) {
  __shared__ double sr[256], si[256];
  const long g = blockIdx.x;
  const double2* Xg = reinterpret_cast<const double2*>(X + g * 2L * nso * nocc);
  const double2* H2 = reinterpret_cast<const double2*>(Hc);
  double* Qg = Q + g * 2L * nocc * nocc;
  double ar = 0.0, ai = 0.0;
  for (int idx = threadIdx.x; idx < nso * nocc; idx += 256) {
    const double2 h = H2[idx], x = Xg[idx];
    ar += h.x * x.x - h.y * x.y;
    ai += h.x * x.y + h.y * x.x;
  }
  for (int a = threadIdx.x; a < nocc; a += 256) {
    ar += Qg[2 * (a + a * nocc)];
    ai += Qg[2 * (a + a * nocc) + 1];
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double er = 0.5 * sr[0], ei = 0.5 * si[0];
  if (threadIdx.x == 0) {
    E[2 * g] = er;
    E[2 * g + 1] = ei;
  }
  // Q_g <- Q_g - e_g I, so that the adjoint is one product: Y = FX - X (Q - e I)
  for (int a = threadIdx.x; a < nocc; a += 256) {
    Qg[2 * (a + a * nocc)] -= er;
    Qg[2 * (a + a * nocc) + 1] -= ei;
  }
}

// The epilogue of one grid point in one 256-thread workgroup (NSO <= 128,
// N <= 32, one density per point), fusing the Fock product, Q, the energy and
// the adjoint of the separate launches:
//   FX = (h + Gamma_g) X_g,  Q = C0^H FX,  e_g = (tr(Hc X_g) + tr Q) / 2,
//   Y_g = FX - X_g (Q - e_g I)  (written to Y, the layout of X), E[g] = e_g.
// X_g is staged once in LDS as [k][a] (with the terms of tr(Hc X) formed on
// the way).  Wave w owns the FX rows p in [32 w, 32 w + 32) (all N columns) on
// the FP64 matrix cores, three real products per complex one, with no barrier
// in the k loop: each lane loads its rows of h + Gamma straight into the MFMA
// A layout, the k index of a 16-deep slab permuted so that step s of lane
// group q = lane / 16 takes k = k0 + 4 q + s (16-byte loads of consecutive k),
// and the next slab is loaded while the current one is multiplied.  The FX
// block stays in the wave's registers in the MFMA's D layout, which is also
// its B layout with the row index as the k index: phase 2 forms this wave's
// share of Q = sum over its rows p of conj(C0(p, a)) FX(p, b), the four
// shares are summed through LDS in a fixed order, and every wave forms e_g
// the same way; phase 3 subtracts X (Q - e I) from its FX block.
// Gam: row 2g the real, row 2g + 1 the imaginary plane of h + Gamma_g, element
// (i, k) at i nso + k.  X_g, Y_g: NSO x N column-major complex at g sX; C0
// likewise; Hc(a, p) at a nso + p.
constexpr int FA_LD = 33;   // row stride (complex elements) of X^T and of the Q shares in LDS
__global__ __launch_bounds__(256, 2) void fock_adjoint_kernel(int nso, int nocc, const double* __restrict__ X,
                                                              long sX, const double* __restrict__ Gam, long ldg,
                                                              const double* __restrict__ C0,
                                                              const double* __restrict__ Hc,
                                                              double* __restrict__ Y, double* __restrict__ E) {
  // X_g as [k][a] (128 rows), later the four Q shares (32 rows each)
  __shared__ double2 sm[128 * FA_LD];
  __shared__ double red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l16 = lane & 15, kq = lane >> 4;
  const int pw = 32 * w;
  const long g = blockIdx.x;
  const double* Xg = X + g * sX;
  const double* Gre = Gam + (2 * g) * ldg;
  const double* Gim = Gre + ldg;
  const double2 z2 = make_double2(0.0, 0.0);
  // stage X^T; this thread's terms of tr(Hc X) (Hc(a, k) and X(k, a) share the flat index)
  double hx_re = 0.0, hx_im = 0.0;
  {
    const double2* H2 = reinterpret_cast<const double2*>(Hc);
    const double2* X2 = reinterpret_cast<const double2*>(Xg);
    const int len = nso * nocc;
#pragma unroll 4
    for (int idx = tid; idx < len; idx += 256) {
      const double2 h = H2[idx], x = X2[idx];
      hx_re += h.x * x.x - h.y * x.y;
      hx_im += h.x * x.y + h.y * x.x;
      const int a = idx / nso, k = idx - a * nso;
      sm[k * FA_LD + a] = x;
    }
  }
  d4_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) p1[i][j] = p2[i][j] = p3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  __syncthreads();
  // ---- phase 1: FX(p, a) = sum_k F(p, k) X(k, a); A[p][k] = F(p, k), B[k][a] = X(k, a).
  // Every load is unconditional (rows clamped to nso - 1, columns k >= nso
  // read at 0): the B operand is zero for k >= nso, and FX rows p >= nso are
  // neither stored nor used (C0 is zero there in phase 2).
  if (pw < nso) {
    const long row0 = (long)min(pw + l16, nso - 1) * nso, row1 = (long)min(pw + 16 + l16, nso - 1) * nso;
    // a slab's A operands: fr/fi[pi][h] = F(p, k0 + 4 kq + 2 h + {0, 1}), p = pw + 16 pi + l16
    auto load = [&](int k0, double2 (&ar)[2][2], double2 (&ai)[2][2]) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = k0 + 4 * kq + 2 * h;   // NSO is even: k < nso implies k + 1 < nso
        const int kc = k < nso ? k : 0;
        ar[0][h] = *reinterpret_cast<const double2*>(Gre + row0 + kc);
        ai[0][h] = *reinterpret_cast<const double2*>(Gim + row0 + kc);
        ar[1][h] = *reinterpret_cast<const double2*>(Gre + row1 + kc);
        ai[1][h] = *reinterpret_cast<const double2*>(Gim + row1 + kc);
      }
    };
    auto slab = [&](int k0, const double2 (&ar)[2][2], const double2 (&ai)[2][2]) {
      double2 xb[2][4];   // [aj][s] = X(k0 + 4 kq + s, 16 aj + l16)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = k0 + 4 * kq + s;
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) xb[aj][s] = sm[(k < nso ? k : 0) * FA_LD + 16 * aj + l16];
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool vk = k0 + 4 * kq + s < nso;
        double are[2], aim[2], bre[2], bim[2];
#pragma unroll
        for (int pi = 0; pi < 2; ++pi) {
          are[pi] = (s & 1) ? ar[pi][s >> 1].y : ar[pi][s >> 1].x;
          aim[pi] = (s & 1) ? ai[pi][s >> 1].y : ai[pi][s >> 1].x;
        }
#pragma unroll
        for (int aj = 0; aj < 2; ++aj) {
          const bool v = vk && 16 * aj + l16 < nocc;
          bre[aj] = v ? xb[aj][s].x : 0.0;
          bim[aj] = v ? xb[aj][s].y : 0.0;
        }
#pragma unroll
        for (int pi = 0; pi < 2; ++pi)
#pragma unroll
          for (int aj = 0; aj < 2; ++aj) {
            p1[pi][aj] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[pi], bre[aj], p1[pi][aj], 0, 0, 0);
            p2[pi][aj] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[pi], bim[aj], p2[pi][aj], 0, 0, 0);
            p3[pi][aj] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[pi] + aim[pi], bre[aj] + bim[aj], p3[pi][aj], 0,
                                                              0, 0);
          }
      }
    };
    // two register buffers; slabs in pairs (a slab past nso contributes zero)
    double2 fr0[2][2], fi0[2][2], fr1[2][2], fi1[2][2];
    load(0, fr0, fi0);
    for (int k0 = 0; k0 < nso; k0 += 32) {
      load(k0 + 16, fr1, fi1);
      slab(k0, fr0, fi0);
      load(k0 + 32, fr0, fi0);
      slab(k0 + 16, fr1, fi1);
    }
  }
  // FX block (pi, aj) in this wave's registers: fxr/fxi[pi][aj][r] = FX(p, a),
  // p = pw + 16 pi + kq + 4 r, a = 16 aj + l16
  double fxr[2][2][4], fxi[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        fxr[i][j][r] = p1[i][j][r] - p2[i][j][r];
        fxi[i][j][r] = p3[i][j][r] - p1[i][j][r] - p2[i][j][r];
      }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    hx_re += __shfl_xor(hx_re, off);
    hx_im += __shfl_xor(hx_im, off);
  }
  if (lane == 0) {
    red[0][w] = hx_re;
    red[1][w] = hx_im;
  }
  // ---- phase 2: this wave's share of Q(a, b) = sum_{p in its rows} conj(C0(p, a)) FX(p, b);
  // k-step (pi, r) takes p = pw + 16 pi + 4 r + kq, whose B operand is fx[pi][bj][r]
  d4_t q1[2][2], q2[2][2], q3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) q1[i][j] = q2[i][j] = q3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  if (pw < nso) {
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
      double2 c0v[2][4];   // [ai][r]
#pragma unroll
      for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int a = 16 * ai + l16, p = pw + 16 * pi + 4 * r + kq;
          const bool v = a < nocc && p < nso;
          const double2 c = *reinterpret_cast<const double2*>(C0 + (v ? 2L * (p + (long)a * nso) : 0));
          c0v[ai][r] = v ? c : z2;
        }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int ai = 0; ai < 2; ++ai) {
          const double cr = c0v[ai][r].x, ci = -c0v[ai][r].y;
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const double br = fxr[pi][bj][r], bi = fxi[pi][bj][r];
            q1[ai][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cr, br, q1[ai][bj], 0, 0, 0);
            q2[ai][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ci, bi, q2[ai][bj], 0, 0, 0);
            q3[ai][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(cr + ci, br + bi, q3[ai][bj], 0, 0, 0);
          }
        }
    }
  }
  __syncthreads();   // every wave is done reading X^T
  // share element (a, b) at w 32 FA_LD + a FA_LD + b: a = 16 ai + kq + 4 r, b = 16 bj + l16
  double2* qs = sm;
#pragma unroll
  for (int ai = 0; ai < 2; ++ai)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = 16 * ai + kq + 4 * r, b = 16 * bj + l16;
        qs[(32 * w + a) * FA_LD + b] =
            make_double2(q1[ai][bj][r] - q2[ai][bj][r], q3[ai][bj][r] - q1[ai][bj][r] - q2[ai][bj][r]);
      }
  __syncthreads();
  // Q = (Q_0 + Q_1) + (Q_2 + Q_3), in place over share 0
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int idx = tid + 256 * m, o = (idx >> 5) * FA_LD + (idx & 31), S = 32 * FA_LD;
    const double2 a0 = qs[o], a1 = qs[S + o], a2 = qs[2 * S + o], a3 = qs[3 * S + o];
    qs[o] = make_double2((a0.x + a1.x) + (a2.x + a3.x), (a0.y + a1.y) + (a2.y + a3.y));
  }
  __syncthreads();
  // e_g, formed by every wave the same way: tr(Hc X) + tr Q
  double er, ei;
  {
    double tr_re = 0.0, tr_im = 0.0;
    if (lane < nocc) {
      const double2 d = qs[lane * (FA_LD + 1)];
      tr_re = d.x;
      tr_im = d.y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      tr_re += __shfl_xor(tr_re, off);
      tr_im += __shfl_xor(tr_im, off);
    }
    er = 0.5 * (((red[0][0] + red[0][1]) + (red[0][2] + red[0][3])) + tr_re);
    ei = 0.5 * (((red[1][0] + red[1][1]) + (red[1][2] + red[1][3])) + tr_im);
  }
  if (tid == 0) {
    E[2 * g] = er;
    E[2 * g + 1] = ei;
  }
  if (pw >= nso) return;
  // ---- phase 3: Y(p, b) = FX(p, b) - sum_c X(p, c) Q'(c, b), Q' = Q - e I;
  // A[p][c] = X(p, c), B[c][b] = Q'(c, b), k-step t takes c = 4 t + kq
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) p1[i][j] = p2[i][j] = p3[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int c = 4 * t + kq;
    double are[2], aim[2], bre[2], bim[2];
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
      const int p = pw + 16 * pi + l16;
      const bool v = p < nso && c < nocc;
      const double2 x = *reinterpret_cast<const double2*>(Xg + (v ? 2L * (p + (long)c * nso) : 0));
      are[pi] = v ? x.x : 0.0;
      aim[pi] = v ? x.y : 0.0;
    }
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int b = 16 * bj + l16;
      const double2 q = qs[c * FA_LD + b];
      bre[bj] = q.x - (c == b ? er : 0.0);
      bim[bj] = q.y - (c == b ? ei : 0.0);
    }
#pragma unroll
    for (int pi = 0; pi < 2; ++pi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        p1[pi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[pi], bre[bj], p1[pi][bj], 0, 0, 0);
        p2[pi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[pi], bim[bj], p2[pi][bj], 0, 0, 0);
        p3[pi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[pi] + aim[pi], bre[bj] + bim[bj], p3[pi][bj], 0, 0, 0);
      }
  }
  double* Yg = Y + g * sX;
#pragma unroll
  for (int pi = 0; pi < 2; ++pi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = pw + 16 * pi + kq + 4 * r, b = 16 * bj + l16;
        if (p >= nso || b >= nocc) continue;
        const double xr = p1[pi][bj][r] - p2[pi][bj][r], xi = p3[pi][bj][r] - p1[pi][bj][r] - p2[pi][bj][r];
        *reinterpret_cast<double2*>(Yg + 2L * (p + (long)b * nso)) =
            make_double2(fxr[pi][bj][r] - xr, fxi[pi][bj][r] - xi);
      }
}

// partial[c][mn][e] = sum_{g in chunk c} Wd[g][mn] * V[g][e]   (complex), for
// two sources at once: blockIdx.z = 0 takes (V0, part0), 1 takes (V1, part1)
__global__ void wreduce_partial_kernel(int G, int gchunk, int len, int nmn, const double* __restrict__ Wd,
                                       const double* __restrict__ V0, const double* __restrict__ V1,
                                       double* __restrict__ part0, double* __restrict__ part1) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  const double* V = blockIdx.z ? V1 : V0;
  double* part = blockIdx.z ? part1 : part0;
  if (e >= len) return;
  const int g0 = c * gchunk, g1 = min(G, g0 + gchunk);
  for (int mn = 0; mn < nmn; ++mn) {
    double ar = 0.0, ai = 0.0;
    for (int g = g0; g < g1; ++g) {
      const double wr = Wd[2 * ((long)g * nmn + mn)], wi = Wd[2 * ((long)g * nmn + mn) + 1];
      const double vr = V[2 * ((long)g * len + e)], vi = V[2 * ((long)g * len + e) + 1];
      ar += wr * vr - wi * vi;
      ai += wr * vi + wi * vr;
    }
    const long o = ((long)c * nmn + mn) * len + e;
    part[2 * o] = ar;
    part[2 * o + 1] = ai;
  }
}

// out_z[mn][e] = sum_c part_z[c][mn][e] in chunk order, for both sources
// (blockIdx.y = z)
__device__ void hs_reduce_block(int G, int nmn, const double* Wd, const double* E, double* HS);

// blockIdx.y = 0, 1: out[e] = sum_c part[c][e] for the two partial blocks;
// blockIdx.y = 2: workgroup 0 forms Hmat/Smat (hs_reduce_block), the rest exit.
__global__ __launch_bounds__(256) void wreduce_final_kernel(int nchunk, int len, int nmn,
                                                            const double* __restrict__ part0,
                                                            const double* __restrict__ part1,
                                                            double* __restrict__ out0, double* __restrict__ out1,
                                                            int Ghs, const double* __restrict__ Wd,
                                                            const double* __restrict__ E, double* __restrict__ HS) {
  if (blockIdx.y == 2) {
    if (blockIdx.x == 0) hs_reduce_block(Ghs, nmn, Wd, E, HS);
    return;
  }
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const double* part = blockIdx.y ? part1 : part0;
  double* out = blockIdx.y ? out1 : out0;
  if (i >= (long)len * nmn) return;
  double ar = 0.0, ai = 0.0;
#pragma unroll 8
  for (int c = 0; c < nchunk; ++c) {
    ar += part[2 * ((long)c * nmn * len + i)];
    ai += part[2 * ((long)c * nmn * len + i) + 1];
  }
  out[2 * i] = ar;
  out[2 * i + 1] = ai;
}

// HS[mn] = sum_g Wd[g][mn] E[g],  HS[nmn + mn] = sum_g Wd[g][mn]; one workgroup.
__device__ void hs_reduce_block(int G, int nmn, const double* Wd, const double* E, double* HS) {
  __shared__ double s[4][256];
  for (int mn = 0; mn < nmn; ++mn) {
    double hr = 0, hi = 0, sr = 0, si = 0;
    for (int g = threadIdx.x; g < G; g += 256) {
      const double wr = Wd[2 * ((long)g * nmn + mn)], wi = Wd[2 * ((long)g * nmn + mn) + 1];
      const double er = E[2 * g], ei = E[2 * g + 1];
      hr += wr * er - wi * ei;
      hi += wr * ei + wi * er;
      sr += wr;
      si += wi;
    }
    s[0][threadIdx.x] = hr;
    s[1][threadIdx.x] = hi;
    s[2][threadIdx.x] = sr;
    s[3][threadIdx.x] = si;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st)
        for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      HS[2 * mn] = s[0][0];
      HS[2 * mn + 1] = s[1][0];
      HS[2 * (nmn + mn)] = s[2][0];
      HS[2 * (nmn + mn) + 1] = s[3][0];
    }
    __syncthreads();
  }
}

// LocalG: g(p,q) = (sum_t c_t op_t(Grad_t) - E0 sum_t c_t op_t(Rdm_t)) / ovlp
__global__ void localg_kernel(long n2, int nterm, const int* __restrict__ tmn, const int* __restrict__ tconj,
                              const double* __restrict__ tcoef, const double* __restrict__ grad,
                              const double* __restrict__ rdm, double e0r, double e0i, double ovr,
                              double ovi, double* __restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n2) return;
  double gr = 0, gi = 0, rr = 0, ri = 0;
  for (int t = 0; t < nterm; ++t) {
    const long off = (long)tmn[t] * n2 + e;
    double ar = grad[2 * off], ai = grad[2 * off + 1];
    double br = rdm[2 * off], bi = rdm[2 * off + 1];
    if (tconj[t]) { ai = -ai; bi = -bi; }
    const double cr = tcoef[2 * t], ci = tcoef[2 * t + 1];
    gr += cr * ar - ci * ai;
    gi += cr * ai + ci * ar;
    rr += cr * br - ci * bi;
    ri += cr * bi + ci * br;
  }
  const double nr = gr - (e0r * rr - e0i * ri), ni = gi - (e0r * ri + e0i * rr);
  const double d = ovr * ovr + ovi * ovi;
  out[2 * e] = (nr * ovr + ni * ovi) / d;
  out[2 * e + 1] = (ni * ovr - nr * ovi) / d;
}

// B_so[L] = B_L (x) I_2 in the [alpha..., beta...] spin-orbital order, complex column-major
__global__ void spin_blocks_kernel(const double* __restrict__ b, int n, int nl, double* __restrict__ bso) {
  const int nso = 2 * n;
  const long total = (long)nl * nso * nso;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int L = (int)(idx / ((long)nso * nso));
    const int rem = (int)(idx % ((long)nso * nso));
    const int col = rem / nso, row = rem % nso;     // column-major
    double v = 0.0;
    if ((row < n) == (col < n)) v = b[((long)L * n + row % n) * n + col % n];
    bso[2 * idx] = v;
    bso[2 * idx + 1] = 0.0;
  }
}

// Bt[r][L] = Re Bt_L(i,k), r = i*nso + k (rows >= nso^2 and columns >= nl are zero);
// flags imaginary parts above tol (a complex Ref is not supported on this path)
__global__ void pack_bt_kernel(const double* __restrict__ bt, int nso, int nl, int lpad, long nrpad,
                               double* __restrict__ out, int* __restrict__ cplx_flag) {
  const long n2 = (long)nso * nso;
  const long total = nrpad * lpad;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long r = idx / lpad;
    const int L = (int)(idx % lpad);
    double v = 0.0;
    if (r < n2 && L < nl) {
      const long i = r / nso, k = r % nso;
      const long src = (long)L * n2 + i + k * nso;   // column-major (i, k)
      v = bt[2 * src];
      if (fabs(bt[2 * src + 1]) > 1e-12 * (fabs(v) + 1.0)) atomicOr(cplx_flag, 1);
    }
    out[idx] = v;
  }
}

// In place, for every (i, j): block M[k][l] = A[(i,k),(j,l)] becomes M - M^T,
// i.e. <ij||kl> = (ik|jl) - (il|jk).  32 x 32 tiles of a block and of its
// transpose are handled by one workgroup.
__global__ __launch_bounds__(256) void antisym_kernel(double* __restrict__ A, int nso, long lda) {
  __shared__ double t1[32][33], t2[32][33];
  const int nkb = (nso + 31) / 32;
  const int i = blockIdx.z, j = blockIdx.y / nkb;
  const int kb = blockIdx.y % nkb, lb = blockIdx.x;
  if (lb < kb) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  auto at = [&](int k, int l) -> double& { return A[((long)i * nso + k) * lda + (long)j * nso + l]; };
  for (int y = ty; y < 32; y += 8) {
    const int k1 = kb * 32 + y, l1 = lb * 32 + tx;
    t1[y][tx] = (k1 < nso && l1 < nso) ? at(k1, l1) : 0.0;
    const int k2 = lb * 32 + y, l2 = kb * 32 + tx;
    t2[y][tx] = (k2 < nso && l2 < nso) ? at(k2, l2) : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int k1 = kb * 32 + y, l1 = lb * 32 + tx;
    if (k1 < nso && l1 < nso) at(k1, l1) = t1[y][tx] - t2[tx][y];
    const int k2 = lb * 32 + y, l2 = kb * 32 + tx;
    if (kb != lb && k2 < nso && l2 < nso) at(k2, l2) = t2[y][tx] - t1[tx][y];
  }
}

// G_vo[p - N][a] = (sum_t c_t Ysum_t(p,a) - E0 sum_t c_t Xsum_t(p,a)) / ovlp, p >= N, a < N:
// the (NSO-N, N) block of LocalG, since (Y C0^H)[p, a] = Y(p, a) for a < N.
__global__ void localg_vo_kernel(int nso, int nocc, int nterm, const int* __restrict__ tmn,
                                 const double* __restrict__ tcoef, const double* __restrict__ ysum,
                                 const double* __restrict__ xsum, double e0r, double e0i, double ovr, double ovi,
                                 double* __restrict__ out) {
  const int nv = nso - nocc;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv * nocc) return;
  const int pv = e / nocc, a = e % nocc;
  const long off = (long)(nocc + pv) + (long)a * nso;
  const long blk = (long)nso * nocc;
  double gr = 0, gi = 0, rr = 0, ri = 0;
  for (int t = 0; t < nterm; ++t) {
    const double cr = tcoef[2 * t], ci = tcoef[2 * t + 1];
    const long o = (long)tmn[t] * blk + off;
    const double yr = ysum[2 * o], yi = ysum[2 * o + 1], xr = xsum[2 * o], xi = xsum[2 * o + 1];
    gr += cr * yr - ci * yi;
    gi += cr * yi + ci * yr;
    rr += cr * xr - ci * xi;
    ri += cr * xi + ci * xr;
  }
  const double nr = gr - (e0r * rr - e0i * ri), ni = gi - (e0r * ri + e0i * rr);
  const double d = ovr * ovr + ovi * ovi;
  out[2 * e] = (nr * ovr + ni * ovi) / d;
  out[2 * e + 1] = (ni * ovr - nr * ovi) / d;
}

// nci = 1 (one CI state): EigenSolver, VecDecomp and LocalG on the device
// (PHFTools.py:8-26, PHFTools.f90:203-247).  The 1 x 1 canonical
// orthogonalisation gives E0 = Re H / |Re S| and fsp = |Re S|^{-1/2}, so the
// gradient block is (Ysum - E0 Xsum)[N:, :] f / (f S) with f = 1 / |Re S|.
// out[0..1] = E0, out[2..] = the (NSO - N) x N block, C order.
// out, hs_out and info_out are host memory (coherent, mapped): the kernel
// also copies Hmat/Smat (nhs doubles) and the inverses' info flags there, so
// the call needs no copies back.
__global__ void objective_kernel(int nso, int nocc, const double* __restrict__ hs, const double* __restrict__ ysum,
                                 const double* __restrict__ xsum, double* __restrict__ out, int nhs,
                                 const int* __restrict__ info, long ninfo, double* __restrict__ hs_out,
                                 int* __restrict__ info_out) {
  const int nv = nso - nocc;
  const double as = fabs(hs[2]);
  const double e0 = hs[0] / as, f = 1.0 / as;
  const double ovr = hs[2] * f, ovi = hs[3] * f;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const long nth = (long)gridDim.x * blockDim.x;
  for (long i = e; i < nhs; i += nth) hs_out[i] = hs[i];
  for (long i = e; i < ninfo; i += nth) info_out[i] = info[i];
  if (e == 0) {
    out[0] = e0;
    out[1] = 0.0;
  }
  if (e < nv * nocc) {
    const int pv = e / nocc, a = e % nocc;
    const long o = (long)(nocc + pv) + (long)a * nso;
    const double gr = f * ysum[2 * o], gi = f * ysum[2 * o + 1], rr = f * xsum[2 * o], ri = f * xsum[2 * o + 1];
    const double nr = gr - e0 * rr, ni = gi - e0 * ri;
    const double d = ovr * ovr + ovi * ovi;
    out[2 + 2 * e] = (nr * ovr + ni * ovi) / d;
    out[2 + 2 * e + 1] = (ni * ovr - nr * ovi) / d;
  }
  __threadfence_system();
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

// Wignerfac (PHFTools.f90:350-376)
cplx wignerfac(int j, int n, int m, double alpha, double beta, double gama) {
  const int smax = std::min(j + m, j - n), smin = std::max(0, m - n);
  double fac = std::tgamma(j + n + 1.0) * std::tgamma(j - n + 1.0) * std::tgamma(j + m + 1.0) *
               std::tgamma(j - m + 1.0);
  fac = std::sqrt(fac);
  double d = 0.0;
  for (int s = smin; s <= smax; ++s) {
    const int p1 = n - m + s, p2 = 2 * j + m - n - 2 * s, p3 = n - m + 2 * s;
    const double numer = ((p1 % 2) ? -1.0 : 1.0) * std::pow(std::cos(beta / 2), p2) * std::pow(std::sin(beta / 2), p3);
    const double denom = std::tgamma(j + m - s + 1.0) * std::tgamma(s + 1.0) * std::tgamma(n - m + s + 1.0) *
                         std::tgamma(j - n - s + 1.0);
    d += numer / denom;
  }
  d *= fac;
  return std::exp(cplx(0.0, -alpha * n)) * d * std::exp(cplx(0.0, -gama * m));
}

// Any nonzero A[r][c] outside the spin blocks (row class != column class of
// the permutation) sets *flag.
__global__ void blk_check_kernel(const double* __restrict__ A, long n2, long lda, const signed char* __restrict__ rcls,
                                 const signed char* __restrict__ ccls, int* __restrict__ flag) {
  const long total = n2 * n2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / n2, c = i % n2;
    if (rcls[r] != ccls[c] && A[r * lda + c] != 0.0) *flag = 1;
  }
}

// dst[r'][c'] = A[rowmap[r']][colmap[c']] (0 where either map is -1): block
// rows x k of the permuted A.
__global__ void blk_gather_kernel(const double* __restrict__ A, long lda, const int* __restrict__ rowmap,
                                  const int* __restrict__ colmap, int rows, int k, double* __restrict__ dst) {
  const long total = (long)rows * k;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i / k), c = (int)(i % k);
    const int nr = rowmap[r], nc = colmap[c];
    dst[i] = (nr >= 0 && nc >= 0) ? A[(long)nr * lda + nc] : 0.0;
  }
}

// dst[row][c'] = src[row][cmap[c']] (0 for cmap = -1): density rows in the
// permuted contraction order.
__global__ void gather_cols_kernel(const double* __restrict__ src, long lds, const int* __restrict__ cmap, int kout,
                                   long rows, double* __restrict__ dst) {
  const long total = rows * kout;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / kout;
    const int c = (int)(i % kout), nc = cmap[c];
    dst[i] = nc >= 0 ? src[r * lds + nc] : 0.0;
  }
}

// Spin-free form, built from the dense A (lda columns, pairs r = i*nso + k):
// dst[0] = 2J - K and dst[1] = K over labels, [kj][kj] each (zero padding),
//   J[(a,c),(b,d)] = A[(alpha a, alpha c), (beta b, beta d)]
//   K[(a,c),(b,d)] = -A[(alpha a, beta c), (beta b, alpha d)]
// mo[s] is label s's alpha MO, mo[n + s] its beta MO.
__global__ void rjk_gather_kernel(const double* __restrict__ A, long lda, const int* __restrict__ mo, int n,
                                  int nso, long kj, double* __restrict__ dst) {
  const long n2 = (long)n * n, total = kj * kj;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / kj, c = i % kj;
    double sj = 0.0, kx = 0.0;
    if (r < n2 && c < n2) {
      const int a = (int)(r / n), cc = (int)(r % n), b = (int)(c / n), d = (int)(c % n);
      const double j = A[((long)mo[a] * nso + mo[cc]) * lda + (long)mo[n + b] * nso + mo[n + d]];
      kx = -A[((long)mo[a] * nso + mo[n + cc]) * lda + (long)mo[n + b] * nso + mo[d]];
      sj = 2.0 * j - kx;
    }
    dst[i] = sj;
    dst[total + i] = kx;
  }
}

// Largest deviation of A from the spin-free structure over the entries the
// spin blocks keep (the others were checked to vanish), and the largest |A|
// there: A[aa,aa] = A[bb,bb] = J - K, A[bb,aa] = J, A[ba,ab] = -K (J, K as
// above).  Bit patterns of non-negative doubles order as the values do.
__global__ void rjk_check_kernel(const double* __restrict__ A, long lda, const int* __restrict__ mo, int n, int nso,
                                 unsigned long long* __restrict__ out) {
  const long n2 = (long)n * n, total = n2 * n2;
  double dev = 0.0, amax = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i / n2 / n), cc = (int)(i / n2 % n), b = (int)(i % n2 / n), d = (int)(i % n);
    const int aa = mo[a], ac = mo[cc], ab = mo[b], ad = mo[d];
    const int ba = mo[n + a], bc = mo[n + cc], bb = mo[n + b], bd = mo[n + d];
    auto at = [&](int p, int q, int r, int s) { return A[((long)p * nso + q) * lda + (long)r * nso + s]; };
    const double j = at(aa, ac, bb, bd), kx = -at(aa, bc, bb, ad);
    const double v[4] = {at(aa, ac, ab, ad), at(ba, bc, bb, bd), at(ba, bc, ab, ad), at(ba, ac, ab, bd)};
    const double e[4] = {j - kx, j - kx, j, -kx};
    for (int q = 0; q < 4; ++q) {
      dev = fmax(dev, fabs(v[q] - e[q]));
      amax = fmax(amax, fabs(v[q]));
    }
    amax = fmax(amax, fmax(fabs(j), fabs(kx)));
  }
  for (int off = 32; off > 0; off >>= 1) {
    dev = fmax(dev, __shfl_xor(dev, off));
    amax = fmax(amax, __shfl_xor(amax, off));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, (unsigned long long)__double_as_longlong(dev));
    atomicMax(out + 1, (unsigned long long)__double_as_longlong(amax));
  }
}

// Spin-free density operand from the natural density rows (src, lds columns,
// rows r < rows of data): set s at rows s Mp: dst[r] = P_aa + P_bb,
// dst[Mp + r] = P_aa - P_bb, dst[2 Mp + r] = P_x1, dst[3 Mp + r] = P_x2, kj
// columns (zero beyond n^2).  Padding rows are cleared by the caller.
__global__ void rjk_pack_kernel(const double* __restrict__ src, long lds, long rows, const int* __restrict__ mo,
                                int n, int nso, long kj, long Mp, double* __restrict__ dst) {
  const long n2 = (long)n * n, total = rows * kj;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / kj, c = i % kj;
    double s = 0.0, df = 0.0, x1 = 0.0, x2 = 0.0;
    if (c < n2) {
      const int b = (int)(c / n), d = (int)(c % n);
      const double* x = src + r * lds;
      const double paa = x[(long)mo[b] * nso + mo[d]], pbb = x[(long)mo[n + b] * nso + mo[n + d]];
      s = paa + pbb;
      df = paa - pbb;
      x1 = x[(long)mo[n + b] * nso + mo[d]];
      x2 = x[(long)mo[b] * nso + mo[n + d]];
    }
    dst[r * kj + c] = s;
    dst[(Mp + r) * kj + c] = df;
    dst[(2 * Mp + r) * kj + c] = x1;
    dst[(3 * Mp + r) * kj + c] = x2;
  }
}

// Gamma + h in the natural column order from the spin-free products (row
// order of rjk_pack_kernel): Gamma_aa = (Cs - Cd)/2, Gamma_bb = (Cs + Cd)/2,
// and with all, Gamma_ab = -C_x1, Gamma_ba = -C_x2 (else the GEMM wrote
// those itself); h from the bias planes (row parity).
__global__ void rjk_unpack_kernel(const double* __restrict__ C, long kj, long rows, long Mp,
                                  const int* __restrict__ mo, int n, int nso, const double* __restrict__ hb,
                                  long nrpad, double* __restrict__ gam, int all) {
  const long n2 = (long)n * n, total = rows * n2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / n2, c = i % n2;
    const int a = (int)(c / n), cc = (int)(c % n);
    const double cs = C[r * kj + c], cd = C[(Mp + r) * kj + c];
    const double* h = hb + (r & 1) * nrpad;
    double* g = gam + r * nrpad;
    const long aa = (long)mo[a] * nso + mo[cc], bb = (long)mo[n + a] * nso + mo[n + cc];
    g[aa] = h[aa] + 0.5 * (cs - cd);
    g[bb] = h[bb] + 0.5 * (cs + cd);
    if (all) {
      const double c1 = C[(2 * Mp + r) * kj + c], c2 = C[(3 * Mp + r) * kj + c];
      const long ab = (long)mo[a] * nso + mo[n + cc], ba = (long)mo[n + a] * nso + mo[cc];
      g[ab] = h[ab] - c1;
      g[ba] = h[ba] - c2;
    }
  }
}

}  // namespace

// Wignerfac for the other host files of the library (the PCC grid weights).
std::complex<double> phf_wignerfac(int j, int n, int m, double alpha, double beta, double gama) {
  return wignerfac(j, n, m, alpha, beta, gama);
}

void fill_zero(hipStream_t s, double* p, long n) {
  if (n <= 0) return;
  long blocks = std::min<long>((n + 255) / 256, 65536);
  fill_zero_kernel<<<blocks, 256, 0, s>>>(p, n);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// Device check and stream set-up of a new context: its own non-blocking
// stream, or `shared_stream` (owned by the caller) when not null.
// Staging of phf_invert_batched_noalloc: a file-scope device array, so that
// entry point allocates nothing.  Each GPU the library runs on holds its own
// copy for the life of the process.  In host code the array's name is only a
// stand-in for the symbol: kernels and copies take the device address that
// hipGetSymbolAddress resolves once per device (kept in g_scratch_ptr).
static constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// With `device` current: resolve its copy of the scratch array (once).
static int resolve_scratch(phf_ctx* ctx, int device) {
  if (device < 0 || device >= kMaxDevices) return fail(ctx, PHF_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return PHF_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(ctx, PHF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return PHF_OK;
}

int phf_ctx_init(phf_ctx* ctx, int device, hipStream_t shared_stream) {
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(ctx, PHF_ERR_HIP, std::string("no HIP device (") + hipGetErrorString(e) + ")");
  if (device < 0 || device >= ndev) return fail(ctx, PHF_ERR_ARG, "device index out of range");
  HIPCK(hipSetDevice(device));
  ctx->device = device;
  if (const char* e = getenv("PHF_SPINFREE")) ctx->use_rjk = atoi(e) != 0;
  {
    const int rc = resolve_scratch(ctx, device);
    if (rc) return rc;
  }
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      ctx->ncu = prop.multiProcessorCount;
  }
  if (shared_stream) {
    ctx->stream = shared_stream;
    ctx->owns_stream = false;
  } else {
    HIPCK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->owns_stream = true;
    for (int k = 0; k < 2; ++k) {
      HIPCK(hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
      HIPCK(hipEventCreateWithFlags(&ctx->join_ev[k], kSyncEventFlags));
    }
    HIPCK(hipEventCreateWithFlags(&ctx->fork_ev, kSyncEventFlags));
    HIPCK(hipEventCreateWithFlags(&ctx->hc_fork, kSyncEventFlags));
    HIPCK(hipEventCreateWithFlags(&ctx->hc_join, kSyncEventFlags));
    HIPCK(hipEventCreateWithFlags(&ctx->tq_fork, kSyncEventFlags));
    HIPCK(hipEventCreateWithFlags(&ctx->tq_join, kSyncEventFlags));
    HIPCK(hipEventCreateWithFlags(&ctx->wd_fork, kSyncEventFlags));
  }
  return PHF_OK;
}

extern "C" int phf_ctx_create(int device, phf_ctx** out) {
  if (!out) {
    g_create_err = "phf_ctx_create: out is NULL";
    return PHF_ERR_ARG;
  }
  *out = nullptr;
  phf_ctx* ctx = new phf_ctx_s();
  const int rc = phf_ctx_init(ctx, device, nullptr);
  if (rc) {
    g_create_err = "phf_ctx_create: " + ctx->err;
    delete ctx;
    return rc;
  }
  *out = ctx;
  return PHF_OK;
}

extern "C" int phf_ctx_destroy(phf_ctx* ctx) {
  if (!ctx) return PHF_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->dA) (void)hipFree(ctx->dA);
  if (ctx->dH1) (void)hipFree(ctx->dH1);
  if (ctx->dHb) (void)hipFree(ctx->dHb);
  for (int q = 0; q < 5; ++q) {
    if (ctx->dAm[q]) (void)hipFree(ctx->dAm[q]);
    if (ctx->dAinv[q]) (void)hipFree(ctx->dAinv[q]);
  }
  drop_blocks(ctx);
  phf_free_sc(ctx);
  if (ctx->og_exec) (void)hipGraphExecDestroy(ctx->og_exec);
  if (ctx->og_graph) (void)hipGraphDestroy(ctx->og_graph);
  if (ctx->d_xitems) (void)hipFree(ctx->d_xitems);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->pcc_ws) (void)hipFree(ctx->pcc_ws);
  if (ctx->small) (void)hipFree(ctx->small);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  for (double* q : {ctx->dGr1p, ctx->dGr2, ctx->dGrpgref, ctx->dGw, ctx->dGu, ctx->dGref, ctx->dGP})
    if (q) (void)hipFree(q);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& pr : ctx->bev)
    for (auto& e : pr)
      if (e) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; ++k) {
    if (ctx->side[k]) {
      (void)hipStreamSynchronize(ctx->side[k]);
      (void)hipStreamDestroy(ctx->side[k]);
    }
    if (ctx->join_ev[k]) (void)hipEventDestroy(ctx->join_ev[k]);
  }
  for (hipEvent_t e : {ctx->fork_ev, ctx->hc_fork, ctx->hc_join, ctx->tq_fork, ctx->tq_join, ctx->wd_fork})
    if (e) (void)hipEventDestroy(e);
  if (ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PHF_OK;
}

extern "C" const char* phf_last_error(const phf_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

// Free the workspace arena (the next user allocates it again at its size).
int phf_release_ws(phf_ctx* ctx) {
  ctx->r_valid = false;
  if (!ctx->ws) return PHF_OK;
  HIPCK(hipStreamSynchronize(ctx->stream));
  HIPCK(hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_size = 0;
  return PHF_OK;
}

static int ensure_ws(phf_ctx* ctx, size_t bytes) {
  ctx->r_valid = false;   // every user of the arena may overwrite the resident results
  if (ctx->ws_size >= bytes) return PHF_OK;
  if (ctx->ws) {
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_size = 0;
  }
  HIPCK(hipMalloc(&ctx->ws, bytes));
  ctx->ws_size = bytes;
  return PHF_OK;
}

extern "C" int phf_set_hamiltonian(phf_ctx* ctx, const double* hone, const double* htwo, int nso, int flags) {
  if (!ctx) return PHF_ERR_ARG;
  ++ctx->gen;
  if (!hone || !htwo || nso <= 0 || nso > 512)
    return fail(ctx, PHF_ERR_ARG, "phf_set_hamiltonian: bad arguments");
  HIPCK(hipSetDevice(ctx->device));
  {
    int rc = drop_hamiltonian(ctx);
    if (rc) return rc;
  }
  const long n2 = (long)nso * nso;
  const long n4 = n2 * n2;
  HIPCK(hipMalloc(&ctx->dH1, sizeof(double) * 2 * n2));
  HIPCK(hipMemcpyAsync(ctx->dH1, hone, sizeof(double) * 2 * n2, hipMemcpyHostToDevice, ctx->stream));

  double* stage = nullptr;
  int* dflag = nullptr;
  HIPCK(hipMalloc(&stage, sizeof(double) * 2 * n4));
  HIPCK(hipMalloc(&dflag, sizeof(int)));
  HIPCK(hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
  HIPCK(hipMemcpyAsync(stage, htwo, sizeof(double) * 2 * n4, hipMemcpyHostToDevice, ctx->stream));
  imag_nonzero_kernel<<<4096, 256, 0, ctx->stream>>>(stage, n4, dflag);
  int hflag = 0;
  HIPCK(hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCK(hipStreamSynchronize(ctx->stream));
  ctx->cplx = hflag != 0 || (flags & 1);
  ctx->ham_flags = flags;
  ctx->nso = nso;
  ctx->kpad = (int)round_up(n2, 256);
  ctx->nrpad = (int)round_up(n2, 256);
  ctx->kt = ctx->cplx ? 2 * ctx->kpad : ctx->kpad;
  const long asz = (long)ctx->nrpad * ctx->kt;
  hipError_t e = hipMalloc(&ctx->dA, sizeof(double) * asz);
  if (e != hipSuccess) {
    (void)hipFree(stage);
    (void)hipFree(dflag);
    ctx->dA = nullptr;
    return fail(ctx, PHF_ERR_NOMEM, std::string("phf_set_hamiltonian: ") + hipGetErrorString(e));
  }
  ham_permute_kernel<<<65536, 256, 0, ctx->stream>>>(stage, ctx->dA, nso, ctx->kpad, ctx->kt, ctx->nrpad);
  {
    std::vector<double> hb(2 * (size_t)ctx->nrpad, 0.0);
    for (long r = 0; r < n2; ++r) {
      hb[r] = hone[2 * r];
      hb[ctx->nrpad + r] = hone[2 * r + 1];
    }
    HIPCK(hipMalloc(&ctx->dHb, sizeof(double) * hb.size()));
    HIPCK(hipMemcpyAsync(ctx->dHb, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
  }
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(ctx->stream));
  HIPCK(hipFree(stage));
  HIPCK(hipFree(dflag));
  ctx->has_ham = true;
  return PHF_OK;
}

// Reduced-precision copy of the cached A for mode prec (1 FP32, 2 BF16X3, 3 F16X2).
static int ensure_mixed_a(phf_ctx* ctx, int prec) {
  if (ctx->dAm[prec]) return PHF_OK;
  if (!ctx->dA) return fail(ctx, PHF_ERR_UNSUPPORTED, "the cached Hamiltonian is stored for another precision only");
  const long elems = (long)ctx->nrpad * ctx->kt * mixed_slices(prec);
  HIPCK(hipMalloc(&ctx->dAm[prec], elems * mixed_elem_bytes(prec)));
  if (prec == 1) {
    convert_f32(ctx->stream, ctx->dA, (float*)ctx->dAm[prec], (long)ctx->nrpad * ctx->kt);
  } else {
    HIPCK(hipMalloc(&ctx->dAinv[prec], sizeof(double) * ctx->nrpad));
    split_operand(ctx->stream, prec, ctx->dA, ctx->kt, ctx->nrpad, ctx->kt, ctx->dAm[prec], ctx->dAinv[prec]);
  }
  HIPCK(hipGetLastError());
  return PHF_OK;
}

// ---------------------------------------------------------------------------
// spin-block structure of A
// ---------------------------------------------------------------------------

std::vector<int> ref_spins(const double* ref, int nso) {
  const int n = nso / 2;
  std::vector<int> sp(nso, -1);
  if (nso % 2) return {};
  for (int p = 0; p < nso; ++p) {
    bool a_zero = true, b_zero = true;
    for (int r = 0; r < nso; ++r) {
      const double re = ref[2 * ((long)r * nso + p)], im = ref[2 * ((long)r * nso + p) + 1];
      if (re != 0.0 || im != 0.0) (r < n ? a_zero : b_zero) = false;
    }
    if (b_zero && !a_zero) sp[p] = 0;
    else if (a_zero && !b_zero) sp[p] = 1;
    else return {};
  }
  return sp;
}

// Labels of a spin-restricted Ref: alpha MO p and beta MO q are partners when
// the alpha rows of column p equal the beta rows of column q.  Returns
// mo[0..n) the alpha MOs in increasing order and mo[n..2n) their partners, or
// nothing when some MO has no partner (spin as from ref_spins).
static std::vector<int> rjk_pairing(const double* ref, int nso, const std::vector<int>& spin) {
  const int n = nso / 2;
  if (!ref || (int)spin.size() != nso) return {};
  std::vector<int> alpha, beta;
  for (int p = 0; p < nso; ++p) (spin[p] == 0 ? alpha : beta).push_back(p);
  if ((int)alpha.size() != n || (int)beta.size() != n) return {};
  std::vector<int> mo(2 * n, -1);
  std::vector<char> used(nso, 0);
  for (int s = 0; s < n; ++s) {
    const int p = alpha[s];
    mo[s] = p;
    for (int q : beta) {
      if (used[q]) continue;
      bool same = true;
      for (int r = 0; r < n && same; ++r)
        same = ref[2 * ((long)r * nso + p)] == ref[2 * ((long)(n + r) * nso + q)] &&
               ref[2 * ((long)r * nso + p) + 1] == ref[2 * ((long)(n + r) * nso + q) + 1];
      if (same) {
        mo[n + s] = q;
        used[q] = 1;
        break;
      }
    }
    if (mo[n + s] < 0) return {};
  }
  return mo;
}

static void drop_rjk(phf_ctx* ctx) {
  for (int q = 0; q < 5; ++q) {
    if (ctx->dAjm[q]) (void)hipFree(ctx->dAjm[q]);
    if (ctx->dAjinv[q]) (void)hipFree(ctx->dAjinv[q]);
    ctx->dAjm[q] = nullptr;
    ctx->dAjinv[q] = nullptr;
  }
  if (ctx->dAj) (void)hipFree(ctx->dAj);
  if (ctx->d_rjmo) (void)hipFree(ctx->d_rjmo);
  if (ctx->d_rjsmap) (void)hipFree(ctx->d_rjsmap);
  if (ctx->d_rjcol) (void)hipFree(ctx->d_rjcol);
  ctx->dAj = nullptr;
  ctx->d_rjmo = nullptr;
  ctx->d_rjsmap = nullptr;
  ctx->d_rjcol = nullptr;
  ctx->rjk = false;
  ctx->rj_n = ctx->rj_kj = 0;
  ctx->rj_mo.clear();
}

// Build the spin-free form for the pairing of ref from the dense A (kept only
// when A has that structure to 1e-12 of its largest entry; the spin blocks
// serve otherwise).  Needs the dense A and the blocks.
static int build_rjk(phf_ctx* ctx, const double* ref) {
  const std::vector<int> mo = rjk_pairing(ref, ctx->nso, ctx->blk_spin);
  if (ctx->rjk ? (!mo.empty() && mo == ctx->rj_mo) : mo.empty()) return PHF_OK;
  ++ctx->gen;
  HIPCK(hipStreamSynchronize(ctx->stream));
  drop_rjk(ctx);
  if (mo.empty() || !ctx->dA || !ctx->blocked || ctx->cplx) return PHF_OK;
  hipStream_t s = ctx->stream;
  const int nso = ctx->nso, n = nso / 2;
  const long kj = round_up((long)n * n, 256);
  HIPCK(hipMalloc(&ctx->d_rjmo, sizeof(int) * 2 * n));
  HIPCK(hipMemcpyAsync(ctx->d_rjmo, mo.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, s));
  unsigned long long* dchk = nullptr;
  HIPCK(hipMalloc(&dchk, 2 * sizeof(unsigned long long)));
  fill_zero(s, reinterpret_cast<double*>(dchk), 2);
  rjk_check_kernel<<<8192, 256, 0, s>>>(ctx->dA, ctx->kt, ctx->d_rjmo, n, nso, dchk);
  unsigned long long hchk[2] = {0, 0};
  HIPCK(hipMemcpyAsync(hchk, dchk, sizeof hchk, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  HIPCK(hipFree(dchk));
  double dev, amax;
  std::memcpy(&dev, &hchk[0], sizeof dev);
  std::memcpy(&amax, &hchk[1], sizeof amax);
  if (!(dev <= 1e-12 * amax)) {   // not spin-free to that accuracy (or NaN): keep the blocks
    drop_rjk(ctx);
    return PHF_OK;
  }
  HIPCK(hipMalloc(&ctx->dAj, sizeof(double) * 2 * kj * kj));
  rjk_gather_kernel<<<8192, 256, 0, s>>>(ctx->dA, ctx->kt, ctx->d_rjmo, n, nso, kj, ctx->dAj);
  HIPCK(hipGetLastError());
  {
    // output columns of the direct epilogue: label pair a n + c -> natural
    // (i, k) = i nso + k of Gamma_aa, Gamma_bb, Gamma_ab, Gamma_ba
    std::vector<int> col(4 * kj, -1);
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < n; ++c) {
        const long r = (long)a * n + c;
        col[r] = mo[a] * nso + mo[c];
        col[kj + r] = mo[n + a] * nso + mo[n + c];
        col[2 * kj + r] = mo[a] * nso + mo[n + c];
        col[3 * kj + r] = mo[n + a] * nso + mo[c];
      }
    HIPCK(hipMalloc(&ctx->d_rjcol, sizeof(int) * col.size()));
    HIPCK(hipMemcpyAsync(ctx->d_rjcol, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice, s));
    HIPCK(hipStreamSynchronize(s));
  }
  if (nso <= dens_split_max_nso()) {
    std::vector<int> smap(256);
    dens_split_rjk_maps(mo, n, smap.data());
    HIPCK(hipMalloc(&ctx->d_rjsmap, sizeof(int) * smap.size()));
    HIPCK(hipMemcpyAsync(ctx->d_rjsmap, smap.data(), sizeof(int) * smap.size(), hipMemcpyHostToDevice, s));
  }
  HIPCK(hipStreamSynchronize(s));
  ctx->rjk = true;
  ctx->rj_n = n;
  ctx->rj_kj = (int)kj;
  ctx->rj_mo = mo;
  ctx->tuned.clear();
  return PHF_OK;
}

// Reduced-precision copy of the spin-free form for mode prec.
static int ensure_mixed_rjk(phf_ctx* ctx, int prec) {
  if (ctx->dAjm[prec]) return PHF_OK;
  if (!ctx->dAj) return fail(ctx, PHF_ERR_UNSUPPORTED, "the spin-free form is stored for another precision only");
  const long kj = ctx->rj_kj, elems = 2 * kj * kj;
  const int ns = mixed_slices(prec), eb = mixed_elem_bytes(prec);
  HIPCK(hipMalloc(&ctx->dAjm[prec], elems * ns * eb));
  if (prec == PHF_PREC_FP32) {
    convert_f32(ctx->stream, ctx->dAj, (float*)ctx->dAjm[prec], elems);
  } else {
    HIPCK(hipMalloc(&ctx->dAjinv[prec], sizeof(double) * 2 * kj));
    split_operand(ctx->stream, prec, ctx->dAj, kj, (int)(2 * kj), (int)kj, ctx->dAjm[prec], ctx->dAjinv[prec]);
  }
  HIPCK(hipGetLastError());
  return PHF_OK;
}

static void drop_blocks(phf_ctx* ctx) {
  drop_rjk(ctx);
  for (int q = 0; q < 5; ++q) {
    if (ctx->dAbm[q]) (void)hipFree(ctx->dAbm[q]);
    if (ctx->dAbinv[q]) (void)hipFree(ctx->dAbinv[q]);
    ctx->dAbm[q] = nullptr;
    ctx->dAbinv[q] = nullptr;
  }
  if (ctx->dAb) (void)hipFree(ctx->dAb);
  if (ctx->d_rowmap) (void)hipFree(ctx->d_rowmap);
  if (ctx->d_colperm) (void)hipFree(ctx->d_colperm);
  if (ctx->d_ocol) (void)hipFree(ctx->d_ocol);
  if (ctx->d_padcols) (void)hipFree(ctx->d_padcols);
  ctx->dAb = nullptr;
  ctx->d_rowmap = nullptr;
  ctx->d_colperm = nullptr;
  ctx->d_ocol = nullptr;
  ctx->d_padcols = nullptr;
  ctx->n_padcols = 0;
  ctx->blocked = false;
  ctx->blk_spin.clear();
}

int phf_build_blocks(phf_ctx* ctx, const std::vector<int>& spin, bool release_dense, const double* ref) {
  ++ctx->gen;
  if (!ctx->dA || ctx->cplx || (int)spin.size() != ctx->nso) return PHF_OK;
  if (ctx->blocked && ctx->blk_spin == spin) return build_rjk(ctx, ref);
  HIPCK(hipStreamSynchronize(ctx->stream));
  drop_blocks(ctx);
  const int nso = ctx->nso;
  const long n2 = (long)nso * nso;
  // pair classes: rows (i,k): 0 if s_i == s_k, 1 for (alpha, beta), 2 for (beta, alpha);
  // columns (j,l): 0 if s_j == s_l, 1 for (beta, alpha), 2 for (alpha, beta)
  std::vector<signed char> rcls(n2), ccls(n2);
  std::vector<int> rows[3], cols[3];
  for (int i = 0; i < nso; ++i)
    for (int k = 0; k < nso; ++k) {
      const int r = i * nso + k;
      const int c = (spin[i] == spin[k]) ? 0 : (spin[i] == 0 ? 1 : 2);
      rcls[r] = (signed char)c;
      rows[c].push_back(r);
      const int cc = (spin[i] == spin[k]) ? 0 : (spin[i] == 1 ? 1 : 2);
      ccls[r] = (signed char)cc;
      cols[cc].push_back(r);
    }
  for (int b = 0; b < 3; ++b)
    if (rows[b].size() != cols[b].size()) return PHF_OK;
  hipStream_t s = ctx->stream;
  signed char *drc = nullptr, *dcc = nullptr;
  int* dflag = nullptr;
  HIPCK(hipMalloc(&drc, n2));
  HIPCK(hipMalloc(&dcc, n2));
  HIPCK(hipMalloc(&dflag, sizeof(int)));
  HIPCK(hipMemcpyAsync(drc, rcls.data(), n2, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dcc, ccls.data(), n2, hipMemcpyHostToDevice, s));
  HIPCK(hipMemsetAsync(dflag, 0, sizeof(int), s));
  blk_check_kernel<<<8192, 256, 0, s>>>(ctx->dA, n2, ctx->kt, drc, dcc, dflag);
  int flag = 0;
  HIPCK(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  HIPCK(hipFree(drc));
  HIPCK(hipFree(dcc));
  HIPCK(hipFree(dflag));
  if (flag) return PHF_OK;   // not spin-block structured: stay dense
  std::vector<int> rowmap, colperm;
  long aoff = 0;
  int kbeg = 0;
  for (int b = 0; b < 3; ++b) {
    const int cnt = (int)rows[b].size();
    const int pr = (int)round_up(std::max(cnt, 1), 256);
    ctx->blk_rows[b] = pr;
    ctx->blk_k[b] = pr;
    ctx->blk_roff[b] = (int)rowmap.size();
    ctx->blk_kbeg[b] = kbeg;
    ctx->blk_aoff[b] = aoff;
    for (int q = 0; q < pr; ++q) {
      rowmap.push_back(q < cnt ? rows[b][q] : -1);
      colperm.push_back(q < cnt ? cols[b][q] : -1);
    }
    aoff += (long)pr * pr;
    kbeg += pr;
  }
  ctx->ktp = kbeg;
  HIPCK(hipMalloc(&ctx->d_rowmap, sizeof(int) * rowmap.size()));
  HIPCK(hipMalloc(&ctx->d_colperm, sizeof(int) * colperm.size()));
  HIPCK(hipMemcpyAsync(ctx->d_rowmap, rowmap.data(), sizeof(int) * rowmap.size(), hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(ctx->d_colperm, colperm.data(), sizeof(int) * colperm.size(), hipMemcpyHostToDevice, s));
  // inverse map (natural element -> permuted column) and the padding columns,
  // for the fused density pack + split
  std::vector<int> ocol(n2, -1), padcols;
  for (int c = 0; c < (int)colperm.size(); ++c) {
    if (colperm[c] >= 0) ocol[colperm[c]] = c;
    else padcols.push_back(c);
  }
  HIPCK(hipMalloc(&ctx->d_ocol, sizeof(int) * n2));
  HIPCK(hipMemcpyAsync(ctx->d_ocol, ocol.data(), sizeof(int) * n2, hipMemcpyHostToDevice, s));
  ctx->n_padcols = (int)padcols.size();
  if (!padcols.empty()) {
    HIPCK(hipMalloc(&ctx->d_padcols, sizeof(int) * padcols.size()));
    HIPCK(hipMemcpyAsync(ctx->d_padcols, padcols.data(), sizeof(int) * padcols.size(), hipMemcpyHostToDevice, s));
  }
  HIPCK(hipMalloc(&ctx->dAb, sizeof(double) * aoff));
  for (int b = 0; b < 3; ++b)
    blk_gather_kernel<<<8192, 256, 0, s>>>(ctx->dA, ctx->kt, ctx->d_rowmap + ctx->blk_roff[b],
                                           ctx->d_colperm + ctx->blk_kbeg[b], ctx->blk_rows[b], ctx->blk_k[b],
                                           ctx->dAb + ctx->blk_aoff[b]);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(s));
  ctx->blocked = true;
  ctx->blk_spin = spin;
  ctx->tuned.clear();
  {
    int rc = build_rjk(ctx, ref);
    if (rc) return rc;
  }
  if (release_dense) {
    HIPCK(hipFree(ctx->dA));
    ctx->dA = nullptr;
    for (int q = 0; q < 5; ++q) {
      if (ctx->dAm[q]) { HIPCK(hipFree(ctx->dAm[q])); ctx->dAm[q] = nullptr; }
      if (ctx->dAinv[q]) { HIPCK(hipFree(ctx->dAinv[q])); ctx->dAinv[q] = nullptr; }
    }
  }
  return PHF_OK;
}

// Reduced-precision copy of the spin blocks for mode prec.
static int ensure_mixed_ab(phf_ctx* ctx, int prec) {
  if (ctx->dAbm[prec]) return PHF_OK;
  if (!ctx->dAb) return fail(ctx, PHF_ERR_UNSUPPORTED, "the spin blocks are stored for another precision only");
  const long elems = ctx->blk_aoff[2] + (long)ctx->blk_rows[2] * ctx->blk_k[2];
  const int nrows = ctx->blk_roff[2] + ctx->blk_rows[2];
  const int ns = mixed_slices(prec);
  const int eb = mixed_elem_bytes(prec);
  HIPCK(hipMalloc(&ctx->dAbm[prec], elems * ns * eb));
  if (prec == 1) {
    convert_f32(ctx->stream, ctx->dAb, (float*)ctx->dAbm[prec], elems);
  } else {
    HIPCK(hipMalloc(&ctx->dAbinv[prec], sizeof(double) * nrows));
    for (int b = 0; b < 3; ++b)
      split_operand(ctx->stream, prec, ctx->dAb + ctx->blk_aoff[b], ctx->blk_k[b], ctx->blk_rows[b], ctx->blk_k[b],
                    (char*)ctx->dAbm[prec] + (long)eb * ns * ctx->blk_aoff[b], ctx->dAbinv[prec] + ctx->blk_roff[b]);
  }
  HIPCK(hipGetLastError());
  return PHF_OK;
}

extern "C" int phf_set_blocked(phf_ctx* ctx, int on) {
  if (!ctx) return PHF_ERR_ARG;
  ++ctx->gen;
  if (!on && !ctx->dA && ctx->blocked)
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_set_blocked: the dense A was released");
  ctx->use_blocks = on != 0;
  ctx->r_valid = false;
  return PHF_OK;
}

extern "C" int phf_blocked(const phf_ctx* ctx) { return (ctx && ctx->blocked && ctx->use_blocks) ? 1 : 0; }

extern "C" int phf_set_spinfree(phf_ctx* ctx, int on) {
  if (!ctx) return PHF_ERR_ARG;
  ++ctx->gen;
  ctx->use_rjk = on != 0;
  ctx->r_valid = false;
  return PHF_OK;
}

extern "C" int phf_spinfree(const phf_ctx* ctx) {
  return (ctx && ctx->rjk && ctx->use_rjk && ctx->blocked && ctx->use_blocks) ? 1 : 0;
}

// Keep only the reduced-precision copy of A for mode prec (1..3) and release
// the FP64 A: the Hamiltonian then serves that mode only.
int phf_keep_only_mixed(phf_ctx* ctx, int prec) {
  if (!ctx->has_ham) return fail(ctx, PHF_ERR_NO_HAMILTONIAN, "no Hamiltonian set");
  if (prec < PHF_PREC_FP32 || prec > PHF_PREC_I8X4) return fail(ctx, PHF_ERR_ARG, "not a reduced-precision mode");
  if (ctx->a_only && ctx->a_only != prec) return fail(ctx, PHF_ERR_ARG, "A already reduced to another mode");
  if (ctx->blocked) {
    // the reduced copy of the structure evaluations will use: the spin-free
    // form when it is built and on (the blocks are then not kept), else the blocks
    const bool sf = ctx->rjk && ctx->use_rjk;
    int rc = sf ? ensure_mixed_rjk(ctx, prec) : ensure_mixed_ab(ctx, prec);
    if (rc) return rc;
    if (!sf) {
      HIPCK(hipStreamSynchronize(ctx->stream));
      drop_rjk(ctx);
    }
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (ctx->dAb) {
      HIPCK(hipFree(ctx->dAb));
      ctx->dAb = nullptr;
    }
    if (ctx->dAj) {
      HIPCK(hipFree(ctx->dAj));
      ctx->dAj = nullptr;
    }
    for (int q = 0; q < 5; ++q)
      if (q != prec) {
        if (ctx->dAjm[q]) { HIPCK(hipFree(ctx->dAjm[q])); ctx->dAjm[q] = nullptr; }
        if (ctx->dAjinv[q]) { HIPCK(hipFree(ctx->dAjinv[q])); ctx->dAjinv[q] = nullptr; }
      }
    ctx->use_blocks = true;
  } else {
    int rc = ensure_mixed_a(ctx, prec);
    if (rc) return rc;
  }
  HIPCK(hipStreamSynchronize(ctx->stream));
  if (ctx->dA) {
    HIPCK(hipFree(ctx->dA));
    ctx->dA = nullptr;
  }
  for (int q = 0; q < 5; ++q)
    if (q != prec) {
      if (ctx->dAm[q]) { HIPCK(hipFree(ctx->dAm[q])); ctx->dAm[q] = nullptr; }
      if (ctx->dAinv[q]) { HIPCK(hipFree(ctx->dAinv[q])); ctx->dAinv[q] = nullptr; }
    }
  ctx->a_only = prec;
  ctx->tuned.clear();
  return PHF_OK;
}

// Keep only the FP64 structure evaluations use: with the spin-free form built
// and on, release the spin blocks' FP64 copy; otherwise release the spin-free
// form.  (Batch instances, whose switches do not change after set-up.)
int phf_keep_used_structure(phf_ctx* ctx) {
  if (!ctx->blocked) return PHF_OK;
  HIPCK(hipStreamSynchronize(ctx->stream));
  if (ctx->rjk && ctx->use_rjk) {
    if (ctx->dAb) {
      HIPCK(hipFree(ctx->dAb));
      ctx->dAb = nullptr;
    }
  } else {
    drop_rjk(ctx);
  }
  ctx->tuned.clear();
  return PHF_OK;
}

// Release the cached Hamiltonian (and everything derived from it).
static int drop_hamiltonian(phf_ctx* ctx) {
  ctx->has_ham = false;
  {
    int rc = free_grid(ctx);
    if (rc) return rc;
  }
  for (int q = 0; q < 5; ++q) {
    if (ctx->dAm[q]) { HIPCK(hipFree(ctx->dAm[q])); ctx->dAm[q] = nullptr; }
    if (ctx->dAinv[q]) { HIPCK(hipFree(ctx->dAinv[q])); ctx->dAinv[q] = nullptr; }
  }
  if (ctx->dA) { HIPCK(hipFree(ctx->dA)); ctx->dA = nullptr; }
  drop_blocks(ctx);
  ctx->use_blocks = true;
  ctx->a_only = 0;
  if (ctx->dH1) { HIPCK(hipFree(ctx->dH1)); ctx->dH1 = nullptr; }
  if (ctx->dHb) { HIPCK(hipFree(ctx->dHb)); ctx->dHb = nullptr; }
  ctx->tuned.clear();
  return PHF_OK;
}

extern "C" int phf_set_hamiltonian_factored(phf_ctx* ctx, const double* h, const double* b, int n, int nl,
                                            const double* ref, int flags) {
  if (!ctx) return PHF_ERR_ARG;
  ++ctx->gen;
  if (!h || !b || !ref || n <= 0 || nl <= 0 || 2 * n > 512)
    return fail(ctx, PHF_ERR_ARG, "phf_set_hamiltonian_factored: bad arguments");
  const int nso = 2 * n;
  const long n2 = (long)nso * nso;
  for (long q = 0; q < n2; ++q)
    if (ref[2 * q + 1] != 0.0)
      return fail(ctx, PHF_ERR_UNSUPPORTED,
                  "phf_set_hamiltonian_factored: complex Ref (use phf_set_hamiltonian with the dense HTwo)");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  {
    int rc = drop_hamiltonian(ctx);
    if (rc) return rc;
  }
  const long nrpad = round_up(n2, 256);
  const int lpad = (int)round_up(nl, 16);
  // one-body: hone = Ref^H (h (x) I2) Ref on the host (NSO^3)
  std::vector<double> hso(2 * n2, 0.0), tmp(2 * n2, 0.0), hone(2 * n2, 0.0);
  for (int p = 0; p < nso; ++p)
    for (int q = 0; q < nso; ++q)
      if ((p < n) == (q < n)) hso[2 * ((long)p * nso + q)] = h[(long)(p % n) * n + q % n];
  for (int p = 0; p < nso; ++p)          // tmp = hso Ref  (real Ref)
    for (int q = 0; q < nso; ++q) {
      double acc = 0.0;
      for (int r = 0; r < nso; ++r) acc += hso[2 * ((long)p * nso + r)] * ref[2 * ((long)r * nso + q)];
      tmp[2 * ((long)p * nso + q)] = acc;
    }
  for (int p = 0; p < nso; ++p)          // hone = Ref^T tmp
    for (int q = 0; q < nso; ++q) {
      double acc = 0.0;
      for (int r = 0; r < nso; ++r) acc += ref[2 * ((long)r * nso + p)] * tmp[2 * ((long)r * nso + q)];
      hone[2 * ((long)p * nso + q)] = acc;
    }
  // device scratch: b, Ref, B_so, T1, Bt (complex), Bt packed real
  const size_t bytes = sizeof(double) * ((size_t)nl * n * n + 2 * n2 + 3 * 2 * (size_t)nl * n2 + nrpad * lpad) + 4096;
  {
    int rc = ensure_ws(ctx, bytes);
    if (rc) return rc;
  }
  Arena ar{ctx->ws};
  double* db = ar.take<double>((size_t)nl * n * n);
  double* dref = ar.take<double>(2 * n2);
  double* dbso = ar.take<double>(2 * (size_t)nl * n2);
  double* dt1 = ar.take<double>(2 * (size_t)nl * n2);
  double* dbt = ar.take<double>(2 * (size_t)nl * n2);
  double* dpk = ar.take<double>((size_t)nrpad * lpad);
  int* dflag = reinterpret_cast<int*>(ar.take<double>(1));
  HIPCK(hipMemcpyAsync(db, b, sizeof(double) * nl * n * n, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dref, ref, sizeof(double) * 2 * n2, hipMemcpyHostToDevice, s));
  fill_zero(s, reinterpret_cast<double*>(dflag), 1);
  spin_blocks_kernel<<<4096, 256, 0, s>>>(db, n, nl, dbso);
  // T1[L] = B_so[L] Ref ; Bt[L] = Ref^H T1[L]   (Ref row-major host layout)
  zgemm_batched(s, nso, nso, nso, 1, 0, zmat_colmajor(dbso, nso, 2 * n2), ZMat{dref, 2L * nso, 2, 0, 1, 0}, 0, 0,
                zout_colmajor(dt1, nso, 2 * n2), nl);
  zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{dref, 2, 2L * nso, 0, 1, 1}, zmat_colmajor(dt1, nso, 2 * n2), 0, 0,
                zout_colmajor(dbt, nso, 2 * n2), nl);
  pack_bt_kernel<<<8192, 256, 0, s>>>(dbt, nso, nl, lpad, nrpad, dpk, dflag);
  ctx->nso = nso;
  ctx->cplx = false;
  ctx->kpad = (int)nrpad;
  ctx->nrpad = (int)nrpad;
  ctx->kt = (int)nrpad;
  HIPCK(hipMalloc(&ctx->dA, sizeof(double) * nrpad * nrpad));
  // A1[(i,k),(j,l)] = sum_L Bt_L(i,k) Bt_L(j,l) on the FP64 matrix cores, then the exchange term
  eri_gemm_f64(s, dpk, lpad, dpk, ctx->dA, (int)nrpad, (int)nrpad, lpad, nrpad, 128, nullptr, nullptr);
  {
    dim3 grid((nso + 31) / 32, (unsigned)(((nso + 31) / 32) * nso), (unsigned)nso);
    antisym_kernel<<<grid, 256, 0, s>>>(ctx->dA, nso, nrpad);
  }
  int hflag = 0;
  HIPCK(hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(s));
  if (hflag) {
    (void)drop_hamiltonian(ctx);
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_set_hamiltonian_factored: transformed factors are complex");
  }
  ctx->ham_flags = flags;
  if (!(flags & PHF_HAM_DENSE)) {
    int rc = phf_build_blocks(ctx, ref_spins(ref, nso), (flags & PHF_HAM_BLOCKS_ONLY) != 0, ref);
    if (rc) return rc;
  }
  // one-body copies
  HIPCK(hipMalloc(&ctx->dH1, sizeof(double) * 2 * n2));
  HIPCK(hipMemcpyAsync(ctx->dH1, hone.data(), sizeof(double) * 2 * n2, hipMemcpyHostToDevice, s));
  {
    std::vector<double> hb(2 * (size_t)nrpad, 0.0);
    for (long r = 0; r < n2; ++r) {
      hb[r] = hone[2 * r];
      hb[nrpad + r] = hone[2 * r + 1];
    }
    HIPCK(hipMalloc(&ctx->dHb, sizeof(double) * hb.size()));
    HIPCK(hipMemcpyAsync(ctx->dHb, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, s));
  }
  HIPCK(hipStreamSynchronize(s));
  ctx->has_ham = true;
  return PHF_OK;
}

// Batched InvertC + determinantC (Wrappers.f90:195-231, 295-335) on the
// kernel the grid stages use for the overlaps M_g: host arrays in numpy order,
// complex[batch][n][n] (element (b,i,j) = Matrix_b(i,j)).
extern "C" int phf_invert_batched(phf_ctx* ctx, int n, int batch, const double* m, double* minv, double* det,
                                  int* info) {
  if (!ctx) return PHF_ERR_ARG;
  if (n <= 0 || batch <= 0 || !m || !minv || !det || !info)
    return fail(ctx, PHF_ERR_ARG, "phf_invert_batched: bad arguments");
  if (n > lu_inverse_max_n())
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_invert_batched: n = " + std::to_string(n) + " exceeds " +
                                              std::to_string(lu_inverse_max_n()));
  HIPCK(hipSetDevice(ctx->device));
  const size_t nn = (size_t)n * n, tot = nn * batch;
  std::vector<double> h(2 * tot);
  for (size_t b = 0; b < (size_t)batch; ++b)   // to column-major, as the kernel reads it
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        h[2 * (b * nn + (size_t)j * n + i)] = m[2 * (b * nn + (size_t)i * n + j)];
        h[2 * (b * nn + (size_t)j * n + i) + 1] = m[2 * (b * nn + (size_t)i * n + j) + 1];
      }
  double *dM = nullptr, *dI = nullptr, *dD = nullptr;
  int* dInfo = nullptr;
  int rc = PHF_OK;
  auto run = [&]() -> int {
    HIPCK(hipMalloc(&dM, sizeof(double) * 2 * tot));
    HIPCK(hipMalloc(&dI, sizeof(double) * 2 * tot));
    HIPCK(hipMalloc(&dD, sizeof(double) * 2 * batch));
    HIPCK(hipMalloc(&dInfo, sizeof(int) * batch));
    HIPCK(hipMemcpyAsync(dM, h.data(), sizeof(double) * 2 * tot, hipMemcpyHostToDevice, ctx->stream));
    lu_inverse_batched(ctx->stream, n, dM, 2L * (long)nn, dI, 2L * (long)nn, dD, dInfo, batch);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(h.data(), dI, sizeof(double) * 2 * tot, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipMemcpyAsync(det, dD, sizeof(double) * 2 * batch, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipMemcpyAsync(info, dInfo, sizeof(int) * batch, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return PHF_OK;
  };
  rc = run();
  hipFree(dM);
  hipFree(dI);
  hipFree(dD);
  hipFree(dInfo);
  if (rc) return rc;
  for (size_t b = 0; b < (size_t)batch; ++b)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        minv[2 * (b * nn + (size_t)i * n + j)] = h[2 * (b * nn + (size_t)j * n + i)];
        minv[2 * (b * nn + (size_t)i * n + j) + 1] = h[2 * (b * nn + (size_t)j * n + i) + 1];
      }
  return PHF_OK;
}

// out[b][j][i] = in[b][i][j] for complex n x n matrices (row-major <-> column-major).
__global__ void transpose_batched_kernel(const double2* __restrict__ in, double2* __restrict__ out, int n,
                                         long total) {
  const long nn = (long)n * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / nn, r = idx - b * nn;
    const int i = (int)(r / n), j = (int)(r - (long)i * n);
    out[b * nn + (long)j * n + i] = in[idx];
  }
}

// phf_invert_batched without allocation: the row-major input, its
// column-major copy, the inverse, det and info are staged in the scratch
// array (largest call: 64 MiB of staging), transposed on the device, and the
// same inverse kernels run.  Results equal phf_invert_batched's bit for bit.
extern "C" int phf_invert_batched_noalloc(phf_ctx* ctx, int n, int batch, const double* m, double* minv,
                                          double* det, int* info) {
  if (!ctx) return PHF_ERR_ARG;
  if (n <= 0 || batch <= 0 || !m || !minv || !det || !info)
    return fail(ctx, PHF_ERR_ARG, "phf_invert_batched_noalloc: bad arguments");
  if (n > lu_inverse_max_n())
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_invert_batched_noalloc: n = " + std::to_string(n) + " exceeds " +
                                              std::to_string(lu_inverse_max_n()));
  const size_t tot = (size_t)n * n * batch;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t mat_bytes = up(16 * tot), det_bytes = up(16 * (size_t)batch), info_bytes = up(4 * (size_t)batch);
  const size_t need = 2 * mat_bytes + det_bytes + info_bytes;
  if (need > kNoAllocScratchBytes)
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_invert_batched_noalloc: " + std::to_string(batch) + " matrices of n = " +
                                              std::to_string(n) + " need " + std::to_string(need) +
                                              " bytes of staging, above the scratch limit of " +
                                              std::to_string(kNoAllocScratchBytes) + " bytes");
  HIPCK(hipSetDevice(ctx->device));
  if (!g_scratch_ptr[ctx->device]) {
    const int rc = resolve_scratch(ctx, ctx->device);
    if (rc) return rc;
  }
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[ctx->device]);
  double2* sA = reinterpret_cast<double2*>(base);                 // row-major input, then the inverse
  double2* sB = reinterpret_cast<double2*>(base + mat_bytes);     // column-major input, then row-major inverse
  double* sDet = reinterpret_cast<double*>(base + 2 * mat_bytes);
  int* sInfo = reinterpret_cast<int*>(base + 2 * mat_bytes + det_bytes);
  hipStream_t s = ctx->stream;
  const int blocks = (int)std::min<long>(((long)tot + 255) / 256, 4096);
  const long nn2 = 2L * n * n;
  HIPCK(hipMemcpyAsync(sA, m, 16 * tot, hipMemcpyHostToDevice, s));
  transpose_batched_kernel<<<blocks, 256, 0, s>>>(sA, sB, n, (long)tot);
  lu_inverse_batched(s, n, reinterpret_cast<const double*>(sB), nn2, reinterpret_cast<double*>(sA), nn2, sDet,
                     sInfo, batch);
  transpose_batched_kernel<<<blocks, 256, 0, s>>>(sA, sB, n, (long)tot);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(minv, sB, 16 * tot, hipMemcpyDeviceToHost, s));
  HIPCK(hipMemcpyAsync(det, sDet, 16 * (size_t)batch, hipMemcpyDeviceToHost, s));
  HIPCK(hipMemcpyAsync(info, sInfo, 4 * (size_t)batch, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  return PHF_OK;
}

extern "C" int phf_get_hone(phf_ctx* ctx, double* hone) {
  if (!ctx || !hone) return PHF_ERR_ARG;
  if (!ctx->has_ham) return fail(ctx, PHF_ERR_NO_HAMILTONIAN, "phf_get_hone: no Hamiltonian set");
  HIPCK(hipMemcpyAsync(hone, ctx->dH1, sizeof(double) * 2 * (size_t)ctx->nso * ctx->nso, hipMemcpyDeviceToHost,
                       ctx->stream));
  HIPCK(hipStreamSynchronize(ctx->stream));
  return PHF_OK;
}

static int free_grid(phf_ctx* ctx) {
  HIPCK(hipStreamSynchronize(ctx->stream));
  for (double** q : {&ctx->dGr1p, &ctx->dGr2, &ctx->dGrpgref, &ctx->dGw, &ctx->dGu, &ctx->dGref, &ctx->dGP})
    if (*q) {
      HIPCK(hipFree(*q));
      *q = nullptr;
    }
  ctx->g_spinrot = false;
  ctx->has_grid = false;
  return PHF_OK;
}

extern "C" int phf_set_grid(phf_ctx* ctx, const double* ref, const int* ngrid, int nci, int ncisp,
                            const double* r1, int nr1, const double* r2, int nr2, const double* rpg, int npg,
                            const double* weightsp, const double* weightpg, int ncipg, const double* roota,
                            int na, const double* rootb, int nb, const double* rooty, int ny_, int qj, int sp) {
  if (ctx) ++ctx->gen;
  if (!ctx) return PHF_ERR_ARG;
  if (!ctx->has_ham) return fail(ctx, PHF_ERR_NO_HAMILTONIAN, "phf_set_grid: no Hamiltonian set");
  const int nso = ctx->nso;
  if (!ref || !ngrid || !r1 || !r2 || !rpg || !weightsp || !weightpg)
    return fail(ctx, PHF_ERR_ARG, "phf_set_grid: NULL pointer argument");
  const int nl = ngrid[0], ny = ngrid[1];
  if (nl <= 0 || ny <= 0 || npg <= 0 || nci <= 0 || ncisp <= 0 || ncipg <= 0 || qj < 0)
    return fail(ctx, PHF_ERR_ARG, "phf_set_grid: non-positive grid or CI dimension");
  if (nl > nr1 || ny > nr2) return fail(ctx, PHF_ERR_ARG, "phf_set_grid: ngrid exceeds R1/R2 extent");
  if (qj + qj + (ncipg - 1) * ncisp >= nci)
    return fail(ctx, PHF_ERR_ARG, "phf_set_grid: nci too small for QJ/ncisp/ncipg");
  if (qj > 0) {
    const bool ok = (sp == 2) ? (nl <= na && ny <= nb && nl <= ny_) : (nl <= na && nl <= nb && ny <= ny_);
    if (!roota || !rootb || !rooty || !ok)
      return fail(ctx, PHF_ERR_ARG, "phf_set_grid: Wigner angle arrays too short");
  }
  const long G = (long)nl * ny * npg;
  if (G > 65535) return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_set_grid: more than 65535 grid points");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  {
    int rc = free_grid(ctx);
    if (rc) return rc;
  }
  const int nmn = nci * nci;
  const long n2 = (long)nso * nso;
  // grid weight table W[g][m][n] = weightsp(il,iy) * Wignerfac * weightpg(ipg,p,q) (PHFTools.f90:120-152)
  std::vector<double> wtab(2 * (size_t)G * nmn, 0.0);
  for (int il = 0; il < nl; ++il)
    for (int iy = 0; iy < ny; ++iy)
      for (int ipg = 0; ipg < npg; ++ipg) {
        const long g = ((long)il * ny + iy) * npg + ipg;
        const double w = weightsp[(long)il * ny + iy];
        for (int ni = -qj; ni <= qj; ++ni)
          for (int nj = -qj; nj <= qj; ++nj) {
            cplx wig(1.0, 0.0);
            if (qj > 0)
              wig = (sp == 2) ? wignerfac(qj, ni, nj, roota[il], rootb[iy], rooty[il])
                              : wignerfac(qj, ni, nj, roota[il], rootb[il], rooty[iy]);
            for (int p = 0; p < ncipg; ++p)
              for (int q = 0; q < ncipg; ++q) {
                const cplx ww = weightpg[((long)ipg * ncipg + p) * ncipg + q] * wig;
                const int m = ni + qj + p * ncisp, n = nj + qj + q * ncisp;
                const size_t o = 2 * ((size_t)g * nmn + (size_t)m * nci + n);
                const cplx v = w * ww;
                wtab[o] += v.real();
                wtab[o + 1] += v.imag();
              }
          }
      }
  HIPCK(hipMalloc(&ctx->dGr1p, sizeof(double) * 2 * nl * n2));
  HIPCK(hipMalloc(&ctx->dGr2, sizeof(double) * 2 * ny * n2));
  HIPCK(hipMalloc(&ctx->dGrpgref, sizeof(double) * 2 * npg * n2));
  HIPCK(hipMalloc(&ctx->dGw, sizeof(double) * wtab.size()));
  // staging for Ref, R1, Rpg in the workspace arena
  {
    int rc = ensure_ws(ctx, sizeof(double) * 2 * n2 * (1 + nl + npg) + 1024);
    if (rc) return rc;
  }
  double* dRef = reinterpret_cast<double*>(ctx->ws);
  double* dR1 = dRef + 2 * n2;
  double* dRpg = dR1 + 2 * nl * n2;
  auto up = [&](double* d, const double* h, size_t nd) {
    return hipMemcpyAsync(d, h, nd * sizeof(double), hipMemcpyHostToDevice, s);
  };
  HIPCK(up(dRef, ref, 2 * n2));
  HIPCK(up(dR1, r1, 2 * (size_t)nl * n2));
  HIPCK(up(ctx->dGr2, r2, 2 * (size_t)ny * n2));
  HIPCK(up(dRpg, rpg, 2 * (size_t)npg * n2));
  HIPCK(up(ctx->dGw, wtab.data(), wtab.size()));
  // R1p[il] = Ref^H R1[il]   (row-major host layouts in, column-major out)
  zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{dRef, 2, 2L * nso, 0, 1, 1}, ZMat{dR1, 2L * nso, 2, 2 * n2, 1, 0}, 0,
                0, zout_colmajor(ctx->dGr1p, nso, 2 * n2), nl);
  // RpgRef[ipg] = Rpg[ipg] Ref
  zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{dRpg, 2L * nso, 2, 2 * n2, 1, 0}, ZMat{dRef, 2L * nso, 2, 0, 1, 0}, 0,
                0, zout_colmajor(ctx->dGrpgref, nso, 2 * n2), npg);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(s));
  // Spin-rotation grid (every R1[il], R2[iy] exactly u (x) I_n): then
  // Ref^H R1 R2 W = sum_{s,t} U_g[s][t] Ref_s^H W_t with U_g = u1[il] u2[iy]
  // and Ref_s the rows of spin s, so the rotation chain needs four NSO x N x n
  // products per evaluation instead of NSO^2 N per point.
  {
    std::vector<cplx> u1(4 * (size_t)nl), u2(4 * (size_t)ny);
    auto su2 = [&](const double* R, cplx* u) {
      if (nso % 2) return false;
      const int n = nso / 2;
      for (int st = 0; st < 4; ++st) {
        const long e = (long)((st >> 1) * n) * nso + (st & 1) * n;
        u[st] = cplx(R[2 * e], R[2 * e + 1]);
      }
      for (int i = 0; i < nso; ++i)
        for (int j = 0; j < nso; ++j) {
          const cplx want = (i % n == j % n) ? u[(i / n) * 2 + j / n] : cplx(0.0, 0.0);
          const long e = (long)i * nso + j;
          if (R[2 * e] != want.real() || R[2 * e + 1] != want.imag()) return false;
        }
      return true;
    };
    bool ok = true;
    for (int il = 0; il < nl && ok; ++il) ok = su2(r1 + 2 * (long)il * n2, &u1[4 * il]);
    for (int iy = 0; iy < ny && ok; ++iy) ok = su2(r2 + 2 * (long)iy * n2, &u2[4 * iy]);
    if (ok) {
      std::vector<double> ug(8 * (size_t)G);
      for (int il = 0; il < nl; ++il)
        for (int iy = 0; iy < ny; ++iy)
          for (int ipg = 0; ipg < npg; ++ipg) {
            const long g = ((long)il * ny + iy) * npg + ipg;
            for (int a = 0; a < 2; ++a)
              for (int c = 0; c < 2; ++c) {
                const cplx v = u1[4 * il + 2 * a] * u2[4 * iy + c] + u1[4 * il + 2 * a + 1] * u2[4 * iy + 2 + c];
                ug[8 * g + 2 * (2 * a + c)] = v.real();
                ug[8 * g + 2 * (2 * a + c) + 1] = v.imag();
              }
          }
      HIPCK(hipMalloc(&ctx->dGu, sizeof(double) * ug.size()));
      HIPCK(hipMalloc(&ctx->dGref, sizeof(double) * 2 * n2));
      HIPCK(hipMemcpy(ctx->dGu, ug.data(), sizeof(double) * ug.size(), hipMemcpyHostToDevice));
      HIPCK(hipMemcpy(ctx->dGref, ref, sizeof(double) * 2 * n2, hipMemcpyHostToDevice));
      // P[s][t][ipg] = Ref_s^H (Rpg[ipg] Ref)_t: each evaluation then needs
      // only Q = P L (one batched product) before the combine
      HIPCK(hipMalloc(&ctx->dGP, sizeof(double) * 2 * n2 * 4 * npg));
      const int n = nso / 2;
      for (int st = 0; st < 4; ++st)
        zgemm_batched(s, nso, nso, n, 1, 0, ZMat{ctx->dGref + 2L * (st >> 1) * n * nso, 2, 2L * nso, 0, 1, 1},
                      ZMat{ctx->dGrpgref + 2L * (st & 1) * n, 2, 2L * nso, 2 * n2, 1, 0}, 0, 0,
                      zout_colmajor(ctx->dGP + 2 * n2 * npg * st, nso, 2 * n2), npg);
      HIPCK(hipGetLastError());
      HIPCK(hipStreamSynchronize(s));
    }
    ctx->g_spinrot = ok;
    ctx->tune_log += std::string("{\"grid\": {\"points\": ") + std::to_string(G) + ", \"rotation\": \"" +
                     (ok ? "su2 x identity" : "general") + "\"}}\n";
  }
  // spin-block structure of a dense Hamiltonian, now that Ref is known
  if (ctx->dA && !ctx->cplx && !(ctx->ham_flags & PHF_HAM_DENSE)) {
    int rc = phf_build_blocks(ctx, ref_spins(ref, nso), false, ref);
    if (rc) return rc;
  }
  ctx->g_nl = nl; ctx->g_ny = ny; ctx->g_npg = npg; ctx->g_nci = nci; ctx->g_ncisp = ncisp;
  ctx->g_ncipg = ncipg; ctx->g_qj = qj; ctx->g_sp = sp; ctx->g_G = G; ctx->g_nso = nso;
  ctx->has_grid = true;
  return PHF_OK;
}

// Choose the 16-bit ERI GEMM tile configuration and split-K for one problem
// shape: with autotuning on, every candidate runs on the live operands and the
// fastest (HIP events, best of two after a warm-up) wins; otherwise the cost
// model's choice.  The configurations that fold into FP32 are only candidates
// when PHF_H16_ALLOW_F32FOLD=1; PHF_H16_CFG=<index>[,<splitk>] restricts the
// candidates to that one configuration (and split-K).
template <class Run>
static int tune_h16(phf_ctx* ctx, const std::tuple<int, long, int, int>& key, long rows, int nrpad, Run&& run) {
  const int forced = h16_forced_config();
  const bool allow_f32fold = getenv("PHF_H16_ALLOW_F32FOLD") && atoi(getenv("PHF_H16_ALLOW_F32FOLD")) == 1;
  std::vector<std::pair<int, int>> cands;
  for (int c = 0; c < h16_num_configs(); ++c) {
    if (forced >= 0 ? c != forced : !h16_cfg_fold64(c) && !allow_f32fold) continue;
    if (!h16_cfg_serves(c, std::get<0>(key), std::get<3>(key) / mixed_segments(std::get<0>(key)))) continue;
    if (h16_cfg_rn16(c) && ctx->nso > h16_rn16_max_nso()) continue;
    if (nrpad % h16_cfg_bn(c) != 0) continue;
    for (int sk : {1, 2})
      if (!h16_forced_splitk() || sk == h16_forced_splitk()) cands.push_back({c, sk});
  }
  if (cands.empty())
    return phf_fail(ctx, PHF_ERR_UNSUPPORTED, "no 16-bit ERI GEMM configuration serves this precision and shape");
  std::sort(cands.begin(), cands.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
    return h16_model_cost(a.first, a.second, rows, nrpad, ctx->ncu) <
           h16_model_cost(b.first, b.second, rows, nrpad, ctx->ncu);
  });
  std::pair<int, int> best = cands.front();
  char line[256];
  snprintf(line, sizeof line, "{\"prec\": %d, \"rows\": %ld, \"cols\": %d, \"k\": %d, \"candidates\": [",
           std::get<0>(key), rows, nrpad, std::get<3>(key));
  std::string log = line;
  if (ctx->autotune) {
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (size_t q = 0; q < cands.size(); ++q) {
      run(cands[q].first, cands[q].second);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        HIPCK(hipEventRecord(e0, ctx->stream));
        run(cands[q].first, cands[q].second);
        HIPCK(hipEventRecord(e1, ctx->stream));
        HIPCK(hipEventSynchronize(e1));
        float ms;
        HIPCK(hipEventElapsedTime(&ms, e0, e1));
        ms_min = std::min(ms_min, ms);
      }
      snprintf(line, sizeof line, "%s{\"cfg\": %d, \"splitk\": %d, \"ms\": %.4f}", q ? ", " : "",
               cands[q].first, cands[q].second, ms_min);
      log += line;
      if (ms_min < best_ms) { best_ms = ms_min; best = cands[q]; }
    }
    HIPCK(hipEventDestroy(e0));
    HIPCK(hipEventDestroy(e1));
  }
  snprintf(line, sizeof line, "], \"chosen\": {\"cfg\": %d, \"splitk\": %d}, \"timed\": %s}\n", best.first,
           best.second, ctx->autotune ? "true" : "false");
  log += line;
  ctx->tune_log += log;
  ctx->tuned[key] = best;
  return PHF_OK;
}

extern "C" int phf_autotune(phf_ctx* ctx, int on) {
  if (!ctx) return PHF_ERR_ARG;
  ++ctx->gen;
  ctx->autotune = on != 0;
  ctx->tuned.clear();
  ctx->xgroup.clear();
  return PHF_OK;
}

extern "C" int phf_tuning_report(const phf_ctx* ctx, char* buf, int len) {
  if (!ctx || !buf || len <= 0) return 0;
  const int n = std::min<int>(len - 1, (int)ctx->tune_log.size());
  std::memcpy(buf, ctx->tune_log.data(), n);
  buf[n] = 0;
  return n;
}

void build_c0(const double* z, int nso, int N, int cmplxconj, double* c0) {
  std::fill(c0, c0 + 2 * (size_t)nso * N, 0.0);
  for (int a = 0; a < N; ++a) c0[2 * ((size_t)a + (size_t)a * nso)] = 1.0;
  for (int i = N; i < nso; ++i)
    for (int a = 0; a < N; ++a) {
      const size_t src = 2 * ((size_t)(i - N) * N + a);
      c0[2 * ((size_t)i + (size_t)a * nso)] = z[src];
      c0[2 * ((size_t)i + (size_t)a * nso) + 1] = (cmplxconj == 0) ? 0.0 : z[src + 1];
    }
}

// PHF_FUSED_SPLIT=0 keeps the density pack and the split as two passes over
// FP64 rows (the tests compare both).
static bool no_fused_split() {
  const char* f = getenv("PHF_FUSED_SPLIT");
  return f && atoi(f) == 0;
}

// PHF_FUSED_EPILOGUE=0 keeps the Fock product, Q, the energy and the adjoint
// as separate launches (the tests compare both).
static bool no_fused_epilogue() {
  const char* f = getenv("PHF_FUSED_EPILOGUE");
  return f && atoi(f) == 0;
}
// PHF_FUSED_SOLVE=0 keeps the overlap inverse and the solve X = T M^{-1} as
// separate launches (the tests compare both).
static bool no_fused_solve() {
  const char* f = getenv("PHF_FUSED_SOLVE");
  return f && atoi(f) == 0;
}
// Side stream k for the small pre-stage and grid-sum work (Hc, the T_g
// combine, the determinant weights) with PHF_SIDE_PRE=1; by default that work
// stays on the main stream: each fork or join of a stream costs the main
// stream more idle time (measured 5-10 us in the replayed launch graph) than
// running these 5-17 us kernels in line.
static hipStream_t pre_side(const phf_ctx* ctx, int k) {
  const char* f = getenv("PHF_SIDE_PRE");
  return (f && atoi(f) == 1) ? ctx->side[k] : nullptr;
}
// PHF_SPLIT_RJK=0 keeps the direct grouped GEMM of the spin-free form as one
// launch followed by the unpack (the tests compare both).
static bool no_split_rjk() {
  const char* f = getenv("PHF_SPLIT_RJK");
  return f && atoi(f) == 0;
}

int eval_plan(phf_ctx* ctx, int nocc, int precision, int cmplxconj, EvalBufs& b) {
  if (!ctx->has_ham) return fail(ctx, PHF_ERR_NO_HAMILTONIAN, "phf_buildhsg: no Hamiltonian set");
  if (!ctx->has_grid || ctx->g_nso != ctx->nso)
    return fail(ctx, PHF_ERR_ARG, "phf_buildhsg_z: no spin grid set for this Hamiltonian (phf_set_grid)");
  const int nso = ctx->nso;
  if (nocc <= 0 || nocc > nso) return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: nocc out of range");
  if (nocc > lu_inverse_max_n())
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_buildhsg: nocc above the LDS inverse limit");
  if (precision < PHF_PREC_FP64 || precision > PHF_PREC_I8X4)
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_buildhsg: unknown precision");
  const bool blk = ctx->blocked && ctx->use_blocks;
  const bool rjk = blk && ctx->rjk && ctx->use_rjk &&
                   (precision == PHF_PREC_FP64 ? ctx->dAj != nullptr
                                               : (ctx->dAj != nullptr || ctx->dAjm[precision] != nullptr));
  if (precision == PHF_PREC_FP64 && !rjk && !(blk ? ctx->dAb != nullptr : ctx->dA != nullptr))
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_buildhsg: the Hamiltonian is stored for a reduced precision only");
  const int npg = ctx->g_npg, ny = ctx->g_ny, nci = ctx->g_nci;
  if (cmplxconj == 2 && (nci % 2 != 0 || 2 * ctx->g_ncisp * ctx->g_ncipg > nci))
    return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: CmplxConj=2 needs nci = 2 ncisp ncipg");
  const int N = nocc;
  const int V = (cmplxconj == 2) ? 4 : 1;   // density variants per grid point
  const long G = ctx->g_G;
  const long Gv = V * G;
  const int nmn = nci * nci;
  const long n2 = (long)nso * nso;
  const int kt = ctx->kt, nrpad = ctx->nrpad;
  b.N = N;
  b.prec = precision;
  b.G = G;
  b.V = V;
  b.blk = blk;
  b.rjk = rjk;   // the spin-free form when built (and stored for this precision)
  b.kP = b.rjk ? ctx->rj_kj : (blk ? ctx->ktp : kt);
  const long kP = b.kP;
  // rows are padded for the tallest tile any candidate uses
  b.bm = (precision == PHF_PREC_FP64) ? eri_gemm_f64_bm((int)(2 * Gv)) : 256;
  b.Mp = round_up(2 * Gv, b.bm);
  b.Mk = b.rjk ? 3 * b.Mp : 0;   // the spin-free form's three K row sets, each padded like the first
  const long rows_op = b.Mp + b.Mk;   // rows of the GEMM's density operand
  const bool fused = precision >= PHF_PREC_BF16X3 && V == 1 && !ctx->cplx && nso <= dens_split_max_nso() &&
                     !no_fused_split();
  b.fused = fused;
  b.gchunk = 32;
  b.nchunk = (int)((Gv + b.gchunk - 1) / b.gchunk);
  const long Mp = b.Mp;
  const long lenX = (long)nso * N;
  auto carve = [&](Arena& a, EvalBufs& o) {
    o.C0 = a.take<double>(2 * lenX);
    o.Wpg = a.take<double>(2 * lenX * npg * ((V == 4) ? 2 : 1));
    const int nl_ = (V == 4) ? 2 : 1;          // distinct left factors C0, C0*
    o.W2 = a.take<double>(2 * lenX * npg * std::max(ny, 4) * nl_);   // also the four spin products Q
    o.T = a.take<double>(2 * lenX * G * nl_);
    o.M = a.take<double>(2 * (size_t)N * N * Gv);
    o.Minv = a.take<double>(2 * (size_t)N * N * Gv);
    o.Det = a.take<double>(2 * (size_t)Gv);
    o.Info = reinterpret_cast<int*>(a.take<double>((size_t)Gv));
    o.X = a.take<double>(2 * lenX * Gv);
    // FP64 density rows (not needed when the pack writes the reduced operand);
    // the spin-free products reuse their memory after the operand is packed
    const size_t pt = fused ? 0 : (size_t)Mp * kt, cj = b.rjk ? (size_t)rows_op * kP : 0;
    o.Pt = a.take<double>(std::max(pt, cj));
    o.Cjk = b.rjk ? o.Pt : nullptr;
    o.Gam = a.take<double>((size_t)Mp * nrpad);
    o.FX = a.take<double>(2 * lenX * Gv);
    o.Q = a.take<double>(2 * (size_t)N * N * Gv);
    o.Hc = a.take<double>(2 * lenX * (V == 4 ? 2 : 1));
    o.E = a.take<double>(2 * (size_t)Gv);
    o.Wd = a.take<double>(2 * (size_t)Gv * nmn);
    o.Part = a.take<double>(4 * (size_t)b.nchunk * nmn * lenX);   // two partial blocks (Ysum, Xsum)
    o.Ysum = a.take<double>(2 * (size_t)nmn * lenX);
    o.Xsum = a.take<double>(2 * (size_t)nmn * lenX);
    o.HS = a.take<double>(4 * (size_t)nmn);
    o.Grad = a.take<double>(2 * (size_t)nmn * n2);
    o.Rdm = a.take<double>(2 * (size_t)nmn * n2);
    // reduced-precision density rows and their inverse scales; FP64 rows in
    // the permuted (spin-block) contraction order
    const size_t pm = (precision == PHF_PREC_FP64) ? 0 : (size_t)rows_op * kP * mixed_slices(precision);
    o.Pm = a.take<double>((pm * mixed_elem_bytes(precision) + 7) / 8);
    o.Pinv = a.take<double>(precision == PHF_PREC_FP64 ? 0 : (size_t)rows_op);
    o.Amax = a.take<double>((precision >= PHF_PREC_BF16X3 && !ctx->cplx) ? (size_t)(2 * Gv) : 0);
    const bool fp64_rows = b.rjk ? !fused : (blk && precision <= PHF_PREC_FP32);
    o.Pperm = a.take<double>(fp64_rows ? (size_t)rows_op * kP : 0);
  };
  Arena a0{nullptr};
  EvalBufs dummy;
  carve(a0, dummy);
  {
    int rc = ensure_ws(ctx, a0.off + 256);
    if (rc) return rc;
  }
  Arena a1{ctx->ws};
  carve(a1, b);
  if (precision != PHF_PREC_FP64) {
    int rc = b.rjk ? ensure_mixed_rjk(ctx, precision)
                   : (blk ? ensure_mixed_ab(ctx, precision) : ensure_mixed_a(ctx, precision));
    if (rc) return rc;
  }
  return PHF_OK;
}

// Density variants (V = 4 for CmplxConj = 2, PHFTools.f90:95-119): variant v
// uses the left factor L_v (T = R L_v) and the right factor B_v (M = B_v^H T,
// rho = T M^{-1} B_v^H):  v = 0: (C0, C0) RDM;  1: (C0*, C0) RDMConj;
// 2: (C0*, C0*) RDM of the conjugate;  3: (C0, C0*) RDMConj of the conjugate.
// Virtual point gv = v G + g; V = 1 is variant 0 alone.
static inline int variant_left(int v) { return (v == 1 || v == 2) ? 1 : 0; }
static inline int variant_right(int v) { return (v == 2 || v == 3) ? 1 : 0; }

// Hc_B = B^H h for each right factor B (C0, and C0* when V = 4), N x NSO,
// element (a, p) stored at a NSO + p (the layout of X_g(p, a))
static void hc_products(phf_ctx* ctx, const EvalBufs& b, hipStream_t s) {
  const int nso = ctx->nso, N = b.N, nB = (b.V == 4) ? 2 : 1;
  const long sT = 2L * nso * N;
  for (int bc = 0; bc < nB; ++bc)
    zgemm_batched(s, N, nso, nso, 1, 0, ZMat{b.C0, 2L * nso, 2, 0, 1, bc ? 0 : 1},
                  ZMat{ctx->dH1, 2L * nso, 2, 0, 1, 0}, 0, 0, zout_colmajor_t(b.Hc + bc * sT, nso, 0), 1);
}

namespace {
// T_g = sum_{s,t} U_g[s][t] Q[s][t][g % npg] (NSO x N blocks, complex)
// grid: x over the lenX elements of a point, y over points (grid-stride), so
// no element index is divided
__global__ void spinrot_kernel(long G, int npg, long lenX, const double* __restrict__ U, const double* __restrict__ Q,
                               double* __restrict__ T) {
  for (long g = blockIdx.y; g < G; g += gridDim.y) {
    const int ipg = (int)(g % npg);
    double u[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) u[q] = U[8 * g + q];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < lenX; e += (long)gridDim.x * blockDim.x) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int st = 0; st < 4; ++st) {   // complex elements as 16-byte loads and stores
        const double ur = u[2 * st], ui = u[2 * st + 1];
        const double2 q = reinterpret_cast<const double2*>(Q)[(st * (long)npg + ipg) * lenX + e];
        re += ur * q.x - ui * q.y;
        im += ur * q.y + ui * q.x;
      }
      reinterpret_cast<double2*>(T)[g * lenX + e] = make_double2(re, im);
    }
  }
}

dim3 spinrot_grid(long G, long lenX) {
  return dim3((unsigned)std::min<long>((lenX + 255) / 256, 1024), (unsigned)std::min<long>(G, 65535));
}
}  // namespace

void eval_pre(phf_ctx* ctx, const EvalBufs& b, hipEvent_t* marks) {
  hipStream_t s = ctx->stream;
  const int nso = ctx->nso, N = b.N, npg = ctx->g_npg, ny = ctx->g_ny, nl = ctx->g_nl, V = b.V;
  const long G = b.G, n2 = (long)nso * nso, lenX = (long)nso * N;
  const int kt = ctx->kt;
  const long sT = 2 * lenX;          // batch stride of NSO x N column-major blocks
  const long sNN = 2L * N * N;
  const int nL = (V == 4) ? 2 : 1;
  // spin-rotation grids: the overlaps come from 4 npg small products and a
  // combine; the products are staged in b.Q (G N^2 per variant, free until
  // the epilogue), which holds them when G >= 4 npg
  const bool m_from_spinrot = ctx->g_spinrot && G >= 4L * npg;
  // one density per point, N <= 32: inverse and solve in one kernel per point
  const bool fused_solve = V == 1 && N <= 32 && !no_fused_solve();
  // Hc = B^H h per right factor (used by the energy in eval_post): on a side
  // stream, overlapping the pre-stages, with pre_side
  if (pre_side(ctx, 0)) {
    (void)hipEventRecord(ctx->hc_fork, s);
    (void)hipStreamWaitEvent(ctx->side[0], ctx->hc_fork, 0);
    hc_products(ctx, b, ctx->side[0]);
    (void)hipEventRecord(ctx->hc_join, ctx->side[0]);   // recorded again after the weights below
  } else {
    // first in line: a 16 us launch during which the host's submission of the
    // short pre-stage launches below gets ahead of the device
    hc_products(ctx, b, s);
  }
  if (marks) (void)hipEventRecord(marks[3], s);   // the rotation stage starts after Hc
  // K1: rotation chain, right to left (Ref^H R1 and Rpg Ref are cached with the
  // grid), once per left factor L = C0 (lc = 0) or C0* (lc = 1)
  bool t_forked = false;
  for (int lc = 0; lc < nL; ++lc) {
    double* Wpg = b.Wpg + lc * sT * npg;
    double* W2 = b.W2 + lc * sT * npg * ny;
    double* T = b.T + lc * sT * G;
    if (ctx->g_spinrot) {
      // Q[s][t][ipg] = Ref_s^H (Rpg[ipg] Ref L)_t = P[s][t][ipg] L in one batched
      // product, then T_g = sum U_g[s][t] Q[s][t][ipg]
      // (these few products split their contraction over more workgroups; the
      // epilogue's FX buffer is free scratch until then)
      double* Q = W2;
      const long ws_cplx = lenX * V * G;
      const int dq = (nso % 4 == 0 && nso >= 64 && 4 * npg <= 16) ? 4 : 1;
      zgemm_batched_kslices(s, nso, N, nso, 1, 0, zmat_colmajor(ctx->dGP, nso, 2 * n2),
                            zmat_colmajor(b.C0, nso, 0, lc), 0, 0, zout_colmajor(Q, nso, sT), 4 * npg, dq, b.FX,
                            ws_cplx);
      // T_g (needed by the solve): with pre_side on a side stream, beside the
      // overlaps below
      if (pre_side(ctx, 1)) {
        (void)hipEventRecord(ctx->tq_fork, s);
        (void)hipStreamWaitEvent(ctx->side[1], ctx->tq_fork, 0);
        spinrot_kernel<<<spinrot_grid(G, lenX), 256, 0, ctx->side[1]>>>(G, npg, lenX, ctx->dGu, Q, T);
        (void)hipEventRecord(ctx->tq_join, ctx->side[1]);
        t_forked = true;
      } else {
        spinrot_kernel<<<spinrot_grid(G, lenX), 256, 0, s>>>(G, npg, lenX, ctx->dGu, Q, T);
      }
      // the overlaps of the variants with this left factor the same way:
      // M_g = B^H T_g = sum U_g[s][t] (B^H Q[s][t][ipg]), four N x N products
      // per point-group element (staged in the Q buffer of the epilogue)
      if (m_from_spinrot)
        for (int v = 0; v < V; ++v) {
          if (variant_left(v) != lc) continue;
          double* MQ = b.Q + v * 4L * npg * sNN;
          zgemm_batched_kslices(s, N, N, nso, 1, 0, ZMat{b.C0, 2L * nso, 2, 0, 1, variant_right(v) ? 0 : 1},
                                zmat_colmajor(Q, nso, sT), 0, 0, zout_colmajor(MQ, N, sNN), 4 * npg, dq, b.FX,
                                ws_cplx);
          if (!fused_solve)   // else the fused inverse forms M_g from MQ as it loads it
            spinrot_kernel<<<spinrot_grid(G, (long)N * N), 256, 0, s>>>(G, npg, (long)N * N, ctx->dGu, MQ,
                                                                        b.M + v * sNN * G);
        }
      continue;
    }
    // Wpg[ipg] = (Rpg[ipg] Ref) L
    zgemm_batched(s, nso, N, nso, 1, 0, zmat_colmajor(ctx->dGrpgref, nso, 2 * n2), zmat_colmajor(b.C0, nso, 0, lc),
                  0, 0, zout_colmajor(Wpg, nso, sT), npg);
    // W2[iy] = R2[iy] [Wpg_0 .. Wpg_{npg-1}]
    zgemm_batched(s, nso, N * npg, nso, 1, 0, ZMat{ctx->dGr2, 2L * nso, 2, 2 * n2, 1, 0}, zmat_colmajor(Wpg, nso, 0), 0,
                  0, zout_colmajor(W2, nso, sT * npg), ny);
    // T[il] = (Ref^H R1[il]) W2   -> T_g = columns g*N .. g*N+N-1
    zgemm_batched(s, nso, N * npg * ny, nso, 1, 0, zmat_colmajor(ctx->dGr1p, nso, 2 * n2), zmat_colmajor(W2, nso, 0),
                  0, 0, zout_colmajor(T, nso, sT * npg * ny), nl);
  }
  if (t_forked) (void)hipStreamWaitEvent(s, ctx->tq_join, 0);
  if (marks) (void)hipEventRecord(marks[0], s);
  // M_v = B_v^H T_{L_v}   (B = C0: C0^H; B = C0*: C0^T); formed above from
  // the spin-rotation products when the grid has that structure
  if (!m_from_spinrot)
    for (int v = 0; v < V; ++v)
      zgemm_batched(s, N, N, nso, 1, 0, ZMat{b.C0, 2L * nso, 2, 0, 1, variant_right(v) ? 0 : 1},
                    zmat_colmajor(b.T + variant_left(v) * sT * G, nso, sT), 0, 0,
                    zout_colmajor(b.M + v * sNN * G, N, sNN), (int)G);
  // K2: inverse and determinant of every variant, and X_v = T_{L_v} M_v^{-1}:
  // one kernel per point for one density per point and N <= 32 (the inverse
  // stays in LDS); else the inverse, then X^T = M^{-T} T^T (stores along the
  // NSO index)
  if (fused_solve) {
    if (m_from_spinrot)   // M_g = sum U_g[s][t] M_st, formed by the kernel from the products in b.Q
      lu_inverse_solve_batched(s, N, nso, b.Q, sNN, b.T, sT, b.X, sT, b.Det, b.Info, (int)G, ctx->dGu, npg);
    else
      lu_inverse_solve_batched(s, N, nso, b.M, sNN, b.T, sT, b.X, sT, b.Det, b.Info, (int)G);
  } else {
    lu_inverse_batched(s, N, b.M, sNN, b.Minv, sNN, b.Det, b.Info, (int)(V * G));
    for (int v = 0; v < V; ++v)
      zgemm_batched(s, N, nso, N, 1, 0, zmat_colmajor_t(b.Minv + v * sNN * G, N, sNN),
                    zmat_colmajor_t(b.T + variant_left(v) * sT * G, nso, sT), 0, 0,
                    zout_colmajor_t(b.X + v * sT * G, nso, sT), (int)G);
  }
  // with pre_side, the determinant weights of the grid sums on the side stream
  // (they need only Det), joined through hc_join before the epilogue; else
  // eval_post launches them
  ctx->wd_forked = false;
  if (pre_side(ctx, 0)) {
    const int nci = ctx->g_nci, nmn = nci * nci, d = (V == 4) ? nci / 2 : nci;
    (void)hipEventRecord(ctx->wd_fork, s);
    (void)hipStreamWaitEvent(ctx->side[0], ctx->wd_fork, 0);
    wd_kernel<<<(unsigned)((V * G * nmn + 255) / 256), 256, 0, ctx->side[0]>>>((int)G, V, nci, d, ctx->dGw, b.Det,
                                                                              b.Wd);
    (void)hipEventRecord(ctx->hc_join, ctx->side[0]);
    ctx->wd_forked = true;
  }
  if (marks) (void)hipEventRecord(marks[1], s);
  // density pack: Pt row 2gv = Re rho_gv^T, row 2gv+1 = Im rho_gv^T (flattened
  // j*nso + l); rho^T = conj(B) X^T, one launch per right factor (variants
  // with B = C0 come first, then B = C0*)
  // the pack writes columns [0, NSO^2) of rows [0, 2 V G): clear the padding
  // rows, or everything when the rows are padded too (or complex-packed)
  // 16-bit and int8 forms with one density per point (V = 1), real A and
  // NSO <= 128: the pack writes the slices directly (no FP64 rows)
  if (b.fused) {
    if (b.rjk)
      dens_split_rjk(s, b.prec, ctx->rj_n, N, b.C0, b.X, sT, (int)G, ctx->d_rjsmap, (int)b.kP, b.Mp, b.Mk, b.Pm,
                     b.Pinv);
    else
      dens_split(s, b.prec, nso, N, b.C0, b.X, sT, (int)G, b.blk ? ctx->d_ocol : nullptr,
                 b.blk ? ctx->d_padcols : nullptr, b.blk ? ctx->n_padcols : 0, (int)b.kP, b.Mp, b.Pm, b.Pinv);
    if (marks) (void)hipEventRecord(marks[2], s);
    return;
  }
  if (ctx->cplx || kt != n2)
    fill_zero(s, b.Pt, b.Mp * kt);
  else
    fill_zero(s, b.Pt + 2 * V * G * (long)kt, (b.Mp - 2 * V * G) * (long)kt);
  const int nB = (V == 4) ? 2 : 1;
  const long per = (V == 4) ? 2 * G : G;
  // the reduced forms take the rows' largest magnitudes from the pack itself
  if (b.Amax) fill_zero(s, b.Amax, 2 * V * G);
  for (int bc = 0; bc < nB; ++bc) {
    ZOut out{b.Pt + bc * per * 2L * kt, nso, 1, 2L * kt, kt};
    if (b.Amax) out.amax = reinterpret_cast<unsigned long long*>(b.Amax + bc * per * 2);
    zgemm_batched(s, nso, nso, N, 1, 0, ZMat{b.C0, 2, 2L * nso, 0, 1, bc ? 0 : 1},
                  ZMat{b.X + bc * per * sT, 2L * nso, 2, sT, 1, 0}, 0, 0, out, (int)per);
  }
  if (ctx->cplx) {
    dim3 grid((unsigned)std::min<long>((n2 + 255) / 256, 1024), (unsigned)(V * G));
    pack_cplx_kernel<<<grid, 256, 0, s>>>(b.Pt, (int)(V * G), (int)n2, ctx->kpad, kt);
  }
  if (marks) (void)hipEventRecord(marks[2], s);
  if (b.rjk) {
    // spin-free operand rows in FP64, then the form's rounding or split
    const long kj = b.kP, rows = 2 * V * G, rows_op = b.Mp + b.Mk;
    for (int set = 0; set < 4; ++set) fill_zero(s, b.Pperm + (set * b.Mp + rows) * kj, (b.Mp - rows) * kj);
    rjk_pack_kernel<<<8192, 256, 0, s>>>(b.Pt, kt, rows, ctx->d_rjmo, ctx->rj_n, nso, kj, b.Mp, b.Pperm);
    if (b.prec == PHF_PREC_FP32)
      convert_f32(s, b.Pperm, (float*)b.Pm, rows_op * kj);
    else if (b.prec != PHF_PREC_FP64)
      split_operand_map(s, b.prec, b.Pperm, kj, (int)rows_op, (int)kj, nullptr, (int)kj, b.Pm, b.Pinv, nullptr,
                        (int)rows_op);
    return;
  }
  // operand of the ERI GEMM: FP64 rows as packed, or in the permuted order of
  // the spin blocks; FP32 roundings; 16-bit slices (gathered when permuted)
  const int* cmap = b.blk ? ctx->d_colperm : nullptr;
  if (b.prec == PHF_PREC_FP64) {
    if (b.blk)
      gather_cols_kernel<<<8192, 256, 0, s>>>(b.Pt, kt, ctx->d_colperm, (int)b.kP, b.Mp, b.Pperm);
  } else if (b.prec == PHF_PREC_FP32) {
    if (b.blk) {
      gather_cols_kernel<<<8192, 256, 0, s>>>(b.Pt, kt, ctx->d_colperm, (int)b.kP, b.Mp, b.Pperm);
      convert_f32(s, b.Pperm, (float*)b.Pm, b.Mp * b.kP);
    } else {
      convert_f32(s, b.Pt, (float*)b.Pm, b.Mp * kt);
    }
  } else {
    split_operand_map(s, b.prec, b.Pt, kt, (int)b.Mp, kt, cmap, (int)b.kP, b.Pm, b.Pinv, b.Amax, (int)(2 * V * G));
  }
}

// Work list of a grouped 16-bit ERI GEMM over (context, evaluation buffers)
// units for tile config cfg and split-K sk.  Units are split into their spin
// blocks (or the dense A) and ordered by decreasing contraction length (largest
// first, for the static round-robin over persistent workgroups); within a
// block, column tiles outermost, row tiles and contraction parts innermost, so
// neighbouring items share A panels.  Full rounds of `grid` items are then
// placed so that XCD x (workgroups b with b % 8 == x) takes a contiguous
// eighth of each round.
//
// With split-K 1 the spin-free form's K items over the P_x1 and P_x2 rows
// write -C + h straight into the (alpha, beta) and (beta, alpha) elements of
// Gamma (their own column maps, scale -1, the bias); the (2J - K) and
// K (P_aa - P_bb) items store their products, and the unpack then forms only
// Gamma_aa and Gamma_bb from those two sets.
static bool rjk_direct(const std::vector<std::pair<const phf_ctx*, const EvalBufs*>>& units, int sk) {
  if (sk != 1) return false;
  for (const auto& u : units)
    if (u.second->rjk) return true;
  return false;
}

bool rjk_direct_items(const std::vector<std::pair<const phf_ctx*, const EvalBufs*>>& units, int sk) {
  return rjk_direct(units, sk);
}

void build_h16_items(const std::vector<std::pair<const phf_ctx*, const EvalBufs*>>& units, int prec, int cfg,
                     int sk, int grid, std::vector<H16Item>& out, unsigned blk_mask) {
  const int bm = h16_cfg_bm(cfg), bn = h16_cfg_bn(cfg), ns = mixed_slices(prec);
  const long eb = mixed_elem_bytes(prec);
  const bool direct = rjk_direct(units, sk);
  struct Unit { int u, blk; long cost; };
  std::vector<Unit> work;
  for (size_t i = 0; i < units.size(); ++i) {
    const phf_ctx* c = units[i].first;
    if (units[i].second->rjk)   // spin-free form: 2J - K over rows [0, Mp), K over [Mp, 4 Mp)
      for (int q = 0; q < 2; ++q) {
        if (blk_mask >> q & 1u) work.push_back({(int)i, 8 + q, (long)c->rj_kj});
      }
    else if (units[i].second->blk)
      for (int q = 0; q < 3; ++q) {
        if (blk_mask >> q & 1u) work.push_back({(int)i, q, (long)c->blk_k[q]});
      }
    else
      work.push_back({(int)i, -1, (long)c->kt});
  }
  std::stable_sort(work.begin(), work.end(), [](const Unit& a, const Unit& b) { return a.cost > b.cost; });
  std::vector<H16Item> logical;
  for (const Unit& u : work) {
    const phf_ctx* c = units[u.u].first;
    const EvalBufs& b = *units[u.u].second;
    if (u.blk >= 8 && direct) {
      if (u.blk == 9) continue;   // both units' items come with unit 0
      const long kj = c->rj_kj, Mp = b.Mp;
      const bool f32 = prec == PHF_PREC_FP32;
      const int tm = (int)(Mp / bm), tn = (int)(kj / bn);
      const int nsl = h16_slabs(prec, cfg, (int)kj);
      // set: 0 = P_aa + P_bb against 2J - K, 1..3 = P_aa - P_bb, P_x1, P_x2 against K
      for (int set = 0; set < 4; ++set)
        for (int nt = 0; nt < tn; ++nt)
          for (int mt = 0; mt < tm; ++mt) {
            H16Item it;
            std::memset(&it, 0, sizeof it);
            const int q = set ? 1 : 0;
            it.P = (const char*)b.Pm + eb * ns * kj * (set * Mp);
            it.A = (const char*)c->dAjm[prec] + eb * ns * kj * kj * q;
            it.invP = f32 ? nullptr : b.Pinv + set * Mp;
            it.invA = f32 ? nullptr : c->dAjinv[prec] + q * kj;
            it.ldP = ns * kj;
            it.sP = kj;
            it.ldA = ns * kj;
            it.sA = kj;
            it.K = (int)kj;
            it.mt = mt;
            it.nt = nt;
            it.sbeg = 0;
            it.send = nsl;
            if (set < 2) {   // products for the unpack of Gamma_aa, Gamma_bb
              it.C = b.Cjk + set * Mp * kj;
              it.ldc = kj;
            } else {         // -C + h into Gamma_ab (set 2) or Gamma_ba (set 3)
              it.C = b.Gam;
              it.ldc = c->nrpad;
              it.bias = c->dHb;
              it.colmap = c->d_rjcol + set * kj;
              it.epi = 1;
              it.s1 = -1.0;
            }
            logical.push_back(it);
          }
      continue;
    }
    if (u.blk >= 8) {
      const int q = u.blk - 8;
      const long kj = c->rj_kj, r0 = q ? b.Mp : 0;
      const bool f32 = prec == PHF_PREC_FP32;
      const int tm = (int)((q ? 3 * b.Mp : round_up(2 * b.V * b.G, bm)) / bm), tn = (int)(kj / bn);
      const int part = h16_slabs(prec, cfg, (int)kj) / sk;
      for (int nt = 0; nt < tn; ++nt)
        for (int mt = 0; mt < tm; ++mt)
          for (int ks = 0; ks < sk; ++ks) {
            H16Item it;
            std::memset(&it, 0, sizeof it);
            it.P = (const char*)b.Pm + eb * ns * kj * r0;
            it.A = (const char*)c->dAjm[prec] + eb * ns * kj * kj * q;
            it.invP = f32 ? nullptr : b.Pinv + r0;
            it.invA = f32 ? nullptr : c->dAjinv[prec] + q * kj;
            it.C = b.Cjk + r0 * kj;
            it.bias = nullptr;
            it.colmap = nullptr;
            it.ldP = ns * kj;
            it.sP = kj;
            it.ldA = ns * kj;
            it.sA = kj;
            it.ldc = kj;
            it.K = (int)kj;
            it.mt = mt;
            it.nt = nt;
            it.sbeg = ks * part;
            it.send = (ks + 1) * part;
            it.atomic = sk > 1;
            logical.push_back(it);
          }
      continue;
    }
    const bool blk = u.blk >= 0;
    const int rows = blk ? c->blk_rows[u.blk] : c->nrpad;
    const int K = blk ? c->blk_k[u.blk] : c->kt;
    const long kbeg = blk ? c->blk_kbeg[u.blk] : 0, aoff = blk ? c->blk_aoff[u.blk] : 0;
    const long roff = blk ? c->blk_roff[u.blk] : 0;
    const char* Abase = (const char*)(blk ? c->dAbm[prec] : c->dAm[prec]);
    const bool f32 = prec == PHF_PREC_FP32;   // no row scales
    const double* invA = f32 ? nullptr : (blk ? c->dAbinv[prec] : c->dAinv[prec]) + roff;
    const int tm = (int)(round_up(2 * b.V * b.G, bm) / bm), tn = rows / bn;
    const int part = h16_slabs(prec, cfg, K) / sk;
    for (int nt = 0; nt < tn; ++nt)
      for (int mt = 0; mt < tm; ++mt)
        for (int ks = 0; ks < sk; ++ks) {
          H16Item it;
          std::memset(&it, 0, sizeof it);   // padding too: work lists are compared bytewise
          it.P = (const char*)b.Pm + eb * kbeg;
          it.A = Abase + eb * ns * aoff;
          it.invP = f32 ? nullptr : b.Pinv;
          it.invA = invA;
          it.C = b.Gam;
          it.bias = ks == 0 ? c->dHb : nullptr;
          it.colmap = blk ? c->d_rowmap + roff : nullptr;
          it.ldP = ns * b.kP;
          it.sP = b.kP;
          it.ldA = (long)ns * K;
          it.sA = K;
          it.ldc = c->nrpad;
          it.K = K;
          it.mt = mt;
          it.nt = nt;
          it.sbeg = ks * part;
          it.send = (ks + 1) * part;
          it.atomic = sk > 1;
          logical.push_back(it);
        }
  }
  h16_xcd_order(logical, grid, out);
}

// Work-list order of a persistent launch of `grid` workgroups: workgroup b
// runs on XCD b % 8 and takes items b, b + grid, ...; each XCD gets a
// contiguous run of the logical items in every round.
void h16_xcd_order(const std::vector<H16Item>& logical, int grid, std::vector<H16Item>& out) {
  const size_t n = logical.size();
  out.resize(n);
  const int per = grid / 8;
  if (per == 0 || grid % 8 != 0) {
    out = logical;
    return;
  }
  const size_t full = n / grid;   // full rounds
  for (size_t l = 0; l < full * (size_t)grid; ++l) {
    const size_t r = l / grid, j = l % grid;
    out[r * grid + (j % per) * 8 + j / per] = logical[l];
  }
  // the last, partial round: workgroup b takes slot b (XCD b % 8); give each XCD
  // a contiguous run of the remaining items as well
  const size_t base = full * (size_t)grid, m = n - base;
  size_t l = base;
  for (size_t x = 0; x < 8 && l < n; ++x)
    for (size_t slot = x; slot < m; slot += 8) out[base + slot] = logical[l++];
}

void rjk_unpack(phf_ctx* ctx, const EvalBufs& b, hipStream_t s, bool all) {
  rjk_unpack_kernel<<<8192, 256, 0, s>>>(b.Cjk, b.kP, 2 * b.V * b.G, b.Mp, ctx->d_rjmo, ctx->rj_n, ctx->nso,
                                         ctx->dHb, ctx->nrpad, b.Gam, all ? 1 : 0);
}

// The ERI GEMM of the spin-free form: unit 0 = (2J - K) against the P_aa + P_bb
// rows, unit 1 = K against the P_aa - P_bb, P_x1, P_x2 rows, into b.Cjk, then
// the unpack into Gamma + h.  FP64: two launches.  Reduced forms: each unit's
// tile is tuned like a block's; the first call of a (precision, rows) times
// the two launches against one grouped launch of both units with unit 1's
// tile and with the 256 x 256 tile, and keeps the fastest; a grouped launch
// with split-K 1 writes Gamma_ab and Gamma_ba + h directly (build_h16_items),
// so its unpack forms Gamma_aa and Gamma_bb only.  Profiling events: bev[0]
// around unit 0 (or the grouped launch), bev[1] unit 1 (0 when grouped),
// bev[2] the unpack.
static int rjk_gemm(phf_ctx* ctx, const EvalBufs& b) {
  hipStream_t s = ctx->stream;
  const int prec = b.prec;
  const long kj = ctx->rj_kj, Gv = b.V * b.G, Mp = b.Mp;
  ctx->nblk_timed = 0;
  ctx->rjk_grouped_last = false;
  auto mark = [&](int q, int k) {
    if (!ctx->profiling || !ctx->bev[q][0]) return;
    (void)hipEventRecord(ctx->bev[q][k], s);
    if (k) ctx->nblk_timed = std::max(ctx->nblk_timed, q + 1);
  };
  auto unpack = [&](bool all) -> int {
    mark(2, 0);
    rjk_unpack(ctx, b, s, all);
    mark(2, 1);
    HIPCK(hipGetLastError());
    return PHF_OK;
  };
  if (prec == PHF_PREC_FP64) {
    mark(0, 0);
    eri_gemm_f64(s, b.Pperm, kj, ctx->dAj, b.Cjk, (int)Mp, (int)kj, (int)kj, kj, b.bm, nullptr, nullptr);
    mark(0, 1);
    mark(1, 0);
    eri_gemm_f64(s, b.Pperm + Mp * kj, kj, ctx->dAj + kj * kj, b.Cjk + Mp * kj, (int)b.Mk, (int)kj, (int)kj, kj,
                 b.bm, nullptr, nullptr);
    mark(1, 1);
    return unpack(true);
  }
  const bool f32 = prec == PHF_PREC_FP32;
  const int ns = mixed_slices(prec);
  const long eb = mixed_elem_bytes(prec);
  auto unit_rows = [&](int q) { return q ? 3 * Mp : 2 * Gv; };   // K: the three padded row sets
  auto run_unit = [&](int q, int cfg, int sk) {
    const long r0 = q ? Mp : 0;
    const int mp = (int)round_up(unit_rows(q), h16_cfg_bm(cfg));
    if (sk > 1) fill_zero(s, b.Cjk + r0 * kj, (long)mp * kj);
    eri_gemm_mixed(s, prec, (const char*)b.Pm + eb * ns * kj * r0, kj,
                   (const char*)ctx->dAjm[prec] + eb * ns * kj * kj * q, f32 ? nullptr : b.Pinv + r0,
                   f32 ? nullptr : ctx->dAjinv[prec] + q * kj, b.Cjk + r0 * kj, mp, (int)kj, (int)kj, kj, cfg, sk,
                   nullptr, nullptr);
  };
  std::pair<int, int> choice[2];
  for (int q = 0; q < 2; ++q) {
    const auto key = std::make_tuple(prec, unit_rows(q), (int)kj, (int)kj * mixed_segments(prec));
    auto it = ctx->tuned.find(key);
    if (it == ctx->tuned.end()) {
      int rc = tune_h16(ctx, key, unit_rows(q), (int)kj, [&](int cfg, int sk) { run_unit(q, cfg, sk); });
      if (rc) return rc;
      it = ctx->tuned.find(key);
    }
    choice[q] = it->second;
  }
  int big = -1;
  if (h16_forced_config() < 0)
    for (int c = 0; c < h16_num_configs(); ++c)
      if (h16_cfg_bm(c) == 256 && h16_cfg_bn(c) == 256 && h16_cfg_serves(c, prec, (int)kj) && kj % 256 == 0 &&
          !(h16_cfg_rn16(c) && ctx->nso > h16_rn16_max_nso()))
        big = c;
  auto sel = [&](int g) { return g == 2 ? std::make_pair(big, 1) : choice[1]; };
  auto upload_items = [&](std::vector<H16Item>& items) -> int {
    if (items.size() != ctx->h_xitems.size() ||
        std::memcmp(items.data(), ctx->h_xitems.data(), items.size() * sizeof(H16Item)) != 0) {
      if (items.size() > ctx->xitems_cap) {
        if (ctx->d_xitems) HIPCK(hipFree(ctx->d_xitems));
        ctx->d_xitems = nullptr;
        HIPCK(hipMalloc(&ctx->d_xitems, items.size() * sizeof(H16Item)));
        ctx->xitems_cap = items.size();
      }
      HIPCK(hipMemcpy(ctx->d_xitems, items.data(), items.size() * sizeof(H16Item), hipMemcpyHostToDevice));
      ctx->h_xitems.swap(items);
    }
    return PHF_OK;
  };
  auto upload = [&](int cfg, int sk) -> int {
    std::vector<H16Item> items;
    build_h16_items({{ctx, &b}}, prec, cfg, sk, h16_grouped_grid(cfg, ctx->ncu), items, 3u);
    return upload_items(items);
  };
  auto run_grouped = [&](int cfg, int sk) {
    if (sk > 1) fill_zero(s, b.Cjk, (Mp + b.Mk) * kj);
    eri_gemm_mixed_grouped(s, prec, ctx->d_xitems, (int)ctx->h_xitems.size(), cfg,
                           h16_grouped_grid(cfg, ctx->ncu), sk > 1);
  };
  // grouping choice, keyed like the exchange blocks' with -1 for the block rows
  const auto gkey = std::make_tuple(prec, 2 * Gv, -1, (int)kj);
  auto git = ctx->xgroup.find(gkey);
  if (git == ctx->xgroup.end()) {
    const int ncand = big >= 0 ? 3 : 2;
    float t[3] = {1e30f, 1e30f, 1e30f};
    if (ctx->autotune) {
      hipEvent_t e0, e1;
      HIPCK(hipEventCreate(&e0));
      HIPCK(hipEventCreate(&e1));
      for (int g = 0; g < ncand; ++g) {
        const auto cs = sel(g);
        if (g) {
          int rc = upload(cs.first, cs.second);
          if (rc) return rc;
        }
        auto run_x = [&]() {   // each candidate with its unpack, if it needs one
          if (g) {
            run_grouped(cs.first, cs.second);
            rjk_unpack(ctx, b, s, cs.second > 1);
          } else {
            for (int q = 0; q < 2; ++q) run_unit(q, choice[q].first, choice[q].second);
            rjk_unpack(ctx, b, s, true);
          }
        };
        run_x();
        for (int rep = 0; rep < 2; ++rep) {
          HIPCK(hipEventRecord(e0, s));
          run_x();
          HIPCK(hipEventRecord(e1, s));
          HIPCK(hipEventSynchronize(e1));
          float ms;
          HIPCK(hipEventElapsedTime(&ms, e0, e1));
          t[g] = std::min(t[g], ms);
        }
      }
      HIPCK(hipEventDestroy(e0));
      HIPCK(hipEventDestroy(e1));
    }
    int pick = ncand - 1;   // without timing: one grouped launch
    if (ctx->autotune) {
      pick = 0;
      for (int g = 1; g < ncand; ++g)
        if (t[g] < t[pick]) pick = g;
    }
    if (const char* f = getenv("PHF_XGROUP")) pick = std::max(0, std::min(atoi(f), ncand - 1));
    static const char* names[3] = {"two_launches", "grouped", "grouped_256x256"};
    char line[256];
    const bool half = pick > 0 && sel(pick).second == 1;   // exchange sets written by the grouped launch
    snprintf(line, sizeof line, "{\"prec\": %d, \"rows\": %ld, \"spin_free_units\": {\"two_launches_ms\": %.4f, "
             "\"grouped_ms\": %.4f, \"grouped_256x256_ms\": %.4f, \"chosen\": \"%s\", \"unpack\": \"%s\"}}\n",
             prec, 2 * Gv, t[0], t[1], ncand > 2 ? t[2] : -1.0f, names[pick], half ? "aa_bb" : "all");
    ctx->tune_log += line;
    git = ctx->xgroup.emplace(gkey, pick).first;
  }
  int grouped = git->second;
  if (grouped == 2 && big < 0) grouped = 1;
  auto cs_split = [&](std::pair<int, int> cs) {
    return cs.second == 1 && ctx->side[1] && rjk_direct_items({{ctx, &b}}, 1) && !no_split_rjk();
  };
  if (grouped && cs_split(sel(grouped))) {
    // Two launches of the direct grouped GEMM: the tiles of the Gamma_aa /
    // Gamma_bb product sets on s, the exchange tiles (which write Gamma_ab and
    // Gamma_ba themselves) on side[1].  The unpack follows the first launch on
    // s and runs on the CUs the exchange tiles leave idle in their last round.
    const int cfg = sel(grouped).first, grid = h16_grouped_grid(cfg, ctx->ncu);
    std::vector<H16Item> logical, part[2], ordered, items;
    build_h16_items({{ctx, &b}}, prec, cfg, 1, 1, logical, 3u);   // one workgroup: logical order
    for (const H16Item& it : logical) part[it.epi ? 1 : 0].push_back(it);
    int gq[2];
    for (int q = 0; q < 2; ++q) {
      gq[q] = (int)std::min<long>(grid, round_up((long)part[q].size(), 8));
      h16_xcd_order(part[q], gq[q], ordered);
      items.insert(items.end(), ordered.begin(), ordered.end());
    }
    int rc = upload_items(items);
    if (rc) return rc;
    const int na = (int)part[0].size(), nx = (int)part[1].size();
    mark(0, 0);
    (void)hipEventRecord(ctx->fork_ev, s);
    (void)hipStreamWaitEvent(ctx->side[1], ctx->fork_ev, 0);
    eri_gemm_mixed_grouped(s, prec, ctx->d_xitems, na, cfg, gq[0], false);
    eri_gemm_mixed_grouped(ctx->side[1], prec, ctx->d_xitems + na, nx, cfg, gq[1], false);
    if (ctx->profiling && ctx->bev[0][0]) {   // the GEMM ends with the exchange launch
      (void)hipEventRecord(ctx->bev[0][1], ctx->side[1]);
      ctx->nblk_timed = std::max(ctx->nblk_timed, 1);
    }
    (void)hipEventRecord(ctx->join_ev[1], ctx->side[1]);
    HIPCK(hipGetLastError());
    mark(1, 0);
    mark(1, 1);
    ctx->rjk_grouped_last = true;
    rc = unpack(false);
    (void)hipStreamWaitEvent(s, ctx->join_ev[1], 0);
    return rc;
  }
  if (grouped) {
    const auto cs = sel(grouped);
    int rc = upload(cs.first, cs.second);
    if (rc) return rc;
    mark(0, 0);
    run_grouped(cs.first, cs.second);
    mark(0, 1);
    mark(1, 0);   // unit 1 ran inside the grouped launch (reported as 0)
    mark(1, 1);
    ctx->rjk_grouped_last = true;
    return unpack(cs.second > 1);   // split-K 1: Gamma_ab and Gamma_ba are written already
  } else {
    for (int q = 0; q < 2; ++q) {
      mark(q, 0);
      run_unit(q, choice[q].first, choice[q].second);
      mark(q, 1);
    }
  }
  return unpack(true);
}

int eval_gemm(phf_ctx* ctx, const EvalBufs& b) {
  if (b.rjk) return rjk_gemm(ctx, b);
  ctx->rjk_grouped_last = false;
  hipStream_t s = ctx->stream;
  const int kt = ctx->kt, nrpad = ctx->nrpad, precision = b.prec;
  const long G = b.V * b.G;                  // rows of the GEMM: every variant of every point
  const int nb = b.blk ? 3 : 1;
  // per block: A rows, contraction length, operand offsets, output column map
  auto rows_of = [&](int q) { return b.blk ? ctx->blk_rows[q] : nrpad; };
  auto k_of = [&](int q) { return b.blk ? ctx->blk_k[q] : kt; };
  auto kbeg_of = [&](int q) { return b.blk ? (long)ctx->blk_kbeg[q] : 0L; };
  auto aoff_of = [&](int q) { return b.blk ? ctx->blk_aoff[q] : 0L; };
  auto map_of = [&](int q) { return b.blk ? ctx->d_rowmap + ctx->blk_roff[q] : (const int*)nullptr; };
  auto roff_of = [&](int q) { return b.blk ? (long)ctx->blk_roff[q] : 0L; };
  // the blocks write disjoint columns of Gamma: for FP64, with side streams,
  // they run concurrently (the exchange blocks fill the CUs the direct
  // block's last round of tiles leaves idle; measured 6.41 -> 6.13 ms at
  // n = 64), joined back before the epilogue.  The reduced forms run them one
  // after another: concurrent blocks there compete for L2 over different
  // operand ranges (measured slower: F16X2 1.34 -> 1.49 ms).
  ctx->nblk_timed = 0;
  auto timed = [&](auto&& run_q) {
    return [&, run_q](int q, hipStream_t sq) {
      if (ctx->profiling && ctx->bev[q][0]) (void)hipEventRecord(ctx->bev[q][0], sq);
      run_q(q, sq);
      if (ctx->profiling && ctx->bev[q][0]) {
        (void)hipEventRecord(ctx->bev[q][1], sq);
        ctx->nblk_timed = std::max(ctx->nblk_timed, q + 1);
      }
    };
  };
  auto all_blocks = [&](auto&& run_q0) -> int {
    auto run_q = timed(run_q0);
    if (nb == 3 && ctx->side[0] && precision == PHF_PREC_FP64) {
      HIPCK(hipEventRecord(ctx->fork_ev, s));
      for (int k = 0; k < 2; ++k) HIPCK(hipStreamWaitEvent(ctx->side[k], ctx->fork_ev, 0));
      run_q(0, s);
      run_q(1, ctx->side[0]);
      run_q(2, ctx->side[1]);
      for (int k = 0; k < 2; ++k) {
        HIPCK(hipEventRecord(ctx->join_ev[k], ctx->side[k]));
        HIPCK(hipStreamWaitEvent(s, ctx->join_ev[k], 0));
      }
    } else {
      for (int q = 0; q < nb; ++q) run_q(q, s);
    }
    return PHF_OK;
  };
  if (precision == PHF_PREC_FP64) {
    const double* P = b.blk ? b.Pperm : b.Pt;
    const double* A = b.blk ? ctx->dAb : ctx->dA;
    return all_blocks([&](int q, hipStream_t sq) {
      eri_gemm_f64(sq, P + kbeg_of(q), b.kP, A + aoff_of(q), b.Gam, (int)b.Mp, rows_of(q), k_of(q), nrpad, b.bm,
                   ctx->dHb, map_of(q));
    });
  } else {
    // FP32 and the split forms share the tile routine; FP32 has one slice and no row scales
    const bool f32 = precision == PHF_PREC_FP32;
    const int ns = mixed_slices(precision);
    const long eb = mixed_elem_bytes(precision);
    const char* P = (const char*)b.Pm;
    const char* A = (const char*)(b.blk ? ctx->dAbm[precision] : ctx->dAm[precision]);
    const double* invA = f32 ? nullptr : (b.blk ? ctx->dAbinv[precision] : ctx->dAinv[precision]);
    const double* invP = f32 ? nullptr : b.Pinv;
    auto run_block = [&](int q, int cfg, int sk, hipStream_t sq) {
      const int mp = (int)round_up(2 * G, h16_cfg_bm(cfg));
      eri_gemm_mixed(sq, precision, P + eb * kbeg_of(q), b.kP, A + eb * ns * aoff_of(q), invP,
                     invA ? invA + roff_of(q) : nullptr, b.Gam, mp, rows_of(q), k_of(q), nrpad, cfg, sk, ctx->dHb,
                     map_of(q));
    };
    std::pair<int, int> choice[3];
    bool any_split = false;
    for (int q = 0; q < nb; ++q) {
      const auto key = std::make_tuple(precision, 2 * G, rows_of(q), k_of(q) * mixed_segments(precision));
      auto it = ctx->tuned.find(key);
      if (it == ctx->tuned.end()) {
        auto run = [&](int cfg, int sk) {
          if (sk > 1) fill_zero(s, b.Gam, (long)round_up(2 * G, h16_cfg_bm(cfg)) * nrpad);
          run_block(q, cfg, sk, s);
        };
        int rc = tune_h16(ctx, key, 2 * G, rows_of(q), run);
        if (rc) return rc;
        it = ctx->tuned.find(key);
      }
      choice[q] = it->second;
      any_split = any_split || choice[q].second > 1;
    }
    // X1 and X2 have one shape, hence one tuned (cfg, split); each alone
    // leaves most CUs idle at n = 64 (112 of 256 tiles), so they can run as
    // one grouped persistent launch.  The first call of a (precision, rows)
    // times that against the two launches and keeps the faster.
    const bool can_group = nb == 3 && choice[1] == choice[2];
    // 0: two launches; 1: one grouped launch with the blocks' tuned tile; 2: one grouped launch with the
    // 256 x 256 tile that serves this precision (the blocks' 2 x 112 tiles at n = 64 then nearly fill the CUs)
    int grouped = 0;
    int big = -1;
    if (can_group && h16_forced_config() < 0)
      for (int c = 0; c < h16_num_configs(); ++c)
        if (h16_cfg_bm(c) == 256 && h16_cfg_bn(c) == 256 && h16_cfg_serves(c, precision, k_of(1)) &&
            rows_of(1) % 256 == 0 && !(h16_cfg_rn16(c) && ctx->nso > h16_rn16_max_nso()))
          big = c;
    auto sel = [&](int g) { return g == 2 ? std::make_pair(big, 1) : choice[1]; };
    // the work list is rebuilt per call (a few hundred items) and uploaded only when it differs
    auto upload = [&](int cfg, int sk) -> int {
      std::vector<H16Item> items;
      build_h16_items({{ctx, &b}}, precision, cfg, sk, h16_grouped_grid(cfg, ctx->ncu), items, 6u);
      if (items.size() != ctx->h_xitems.size() ||
          std::memcmp(items.data(), ctx->h_xitems.data(), items.size() * sizeof(H16Item)) != 0) {
        if (items.size() > ctx->xitems_cap) {
          if (ctx->d_xitems) HIPCK(hipFree(ctx->d_xitems));
          ctx->d_xitems = nullptr;
          HIPCK(hipMalloc(&ctx->d_xitems, items.size() * sizeof(H16Item)));
          ctx->xitems_cap = items.size();
        }
        HIPCK(hipMemcpy(ctx->d_xitems, items.data(), items.size() * sizeof(H16Item), hipMemcpyHostToDevice));
        ctx->h_xitems.swap(items);
      }
      return PHF_OK;
    };
    auto run_grouped = [&](int cfg, int sk, hipStream_t sq) {
      eri_gemm_mixed_grouped(sq, precision, ctx->d_xitems, (int)ctx->h_xitems.size(), cfg,
                             h16_grouped_grid(cfg, ctx->ncu), sk > 1);
    };
    if (can_group) {
      const auto gkey = std::make_tuple(precision, 2 * G, rows_of(1), k_of(1));
      auto it = ctx->xgroup.find(gkey);
      if (it == ctx->xgroup.end()) {
        const int ncand = big >= 0 ? 3 : 2;
        float t[3] = {1e30f, 1e30f, 1e30f};
        if (ctx->autotune) {
          hipEvent_t e0, e1;
          HIPCK(hipEventCreate(&e0));
          HIPCK(hipEventCreate(&e1));
          for (int g = 0; g < ncand; ++g) {
            const auto cs = sel(g);
            if (g) {
              int rc = upload(cs.first, cs.second);
              if (rc) return rc;
            }
            auto run_x = [&]() {
              if (cs.second > 1) fill_zero(s, b.Gam, b.Mp * nrpad);
              if (g)
                run_grouped(cs.first, cs.second, s);
              else
                for (int q = 1; q < 3; ++q) run_block(q, cs.first, cs.second, s);
            };
            run_x();
            for (int rep = 0; rep < 2; ++rep) {
              HIPCK(hipEventRecord(e0, s));
              run_x();
              HIPCK(hipEventRecord(e1, s));
              HIPCK(hipEventSynchronize(e1));
              float ms;
              HIPCK(hipEventElapsedTime(&ms, e0, e1));
              t[g] = std::min(t[g], ms);
            }
          }
          HIPCK(hipEventDestroy(e0));
          HIPCK(hipEventDestroy(e1));
        }
        int pick = 0;
        if (ctx->autotune)
          for (int g = 1; g < ncand; ++g)
            if (t[g] < t[pick]) pick = g;
        if (const char* f = getenv("PHF_XGROUP")) {   // pin (tests, experiments): 0, 1, or 2 where it exists
          pick = std::max(0, std::min(atoi(f), ncand - 1));
        }
        static const char* names[3] = {"two_launches", "grouped", "grouped_256x256"};
        char line[256];
        snprintf(line, sizeof line, "{\"prec\": %d, \"rows\": %ld, \"exchange_blocks\": {\"two_launches_ms\": %.4f, "
                 "\"grouped_ms\": %.4f, \"grouped_256x256_ms\": %.4f, \"chosen\": \"%s\"}}\n", precision, 2 * G,
                 t[0], t[1], ncand > 2 ? t[2] : -1.0f, names[pick]);
        ctx->tune_log += line;
        it = ctx->xgroup.emplace(gkey, pick).first;
      }
      grouped = it->second;
      if (grouped == 2 && big < 0) grouped = 1;   // a pinned configuration since the tuning
    }
    const auto gsel = sel(grouped);
    if (grouped) {
      int rc = upload(gsel.first, gsel.second);
      if (rc) return rc;
    }
    // blocks write disjoint columns of Gamma: one zero fill before all of them
    if (any_split || (grouped && gsel.second > 1)) fill_zero(s, b.Gam, b.Mp * nrpad);
    if (grouped) {
      auto run_q = timed([&](int q, hipStream_t sq) {
        if (q == 0)
          run_block(0, choice[0].first, choice[0].second, sq);
        else
          run_grouped(gsel.first, gsel.second, sq);
      });
      run_q(0, s);
      run_q(1, s);
      return PHF_OK;
    }
    return all_blocks([&](int q, hipStream_t sq) { run_block(q, choice[q].first, choice[q].second, sq); });
  }
}

void eval_post(phf_ctx* ctx, const EvalBufs& b, bool resident) {
  hipStream_t s = ctx->stream;
  const int nso = ctx->nso, N = b.N, nci = ctx->g_nci, nrpad = ctx->nrpad, V = b.V;
  const int nmn = nci * nci;
  const long G = b.G, Gv = V * b.G, n2 = (long)nso * nso, lenX = (long)nso * N;
  const long sT = 2 * lenX, sNN = 2L * N * N;
  const int nB = (V == 4) ? 2 : 1;
  const long per = (V == 4) ? 2 * G : G;
  // one density per point, NSO <= 128, N <= 32: the Fock product, Q, the
  // energy and the adjoint in one kernel per point (Y written over FX)
  const bool fused = V == 1 && nso <= 128 && N <= 32 && !no_fused_epilogue();
  if (fused) {
    if (pre_side(ctx, 0)) (void)hipStreamWaitEvent(s, ctx->hc_join, 0);   // else eval_pre formed Hc on s
    fock_adjoint_kernel<<<(unsigned)G, 256, 0, s>>>(nso, N, b.X, sT, b.Gam, nrpad, b.C0, b.Hc, b.FX, b.E);
  }
  // K4: FX = (h + Gamma) X   (the ERI GEMM epilogue added h), formed as
  // FX^T = X^T (h + Gamma)^T so that the stores run along the NSO index
  if (!fused)
    zgemm_batched(s, N, nso, nso, 1, 0, zmat_colmajor_t(b.X, nso, sT), ZMat{b.Gam, 1, nso, 2L * nrpad, nrpad, 0}, 0,
                  0, zout_colmajor_t(b.FX, nso, sT), (int)Gv);
  // per right factor B: Q_v = B^H FX_v, Hc = B^H h (formed by eval_pre), e_v = Makeh
  if (!fused && pre_side(ctx, 0)) (void)hipStreamWaitEvent(s, ctx->hc_join, 0);   // else eval_pre formed Hc on s
  for (int bc = 0; bc < nB && !fused; ++bc) {
    const ZMat bh{b.C0, 2L * nso, 2, 0, 1, bc ? 0 : 1};
    double* Hc = b.Hc + bc * sT;
    zgemm_batched(s, N, N, nso, 1, 0, bh, zmat_colmajor(b.FX + bc * per * sT, nso, sT), 0, 0,
                  zout_colmajor(b.Q + bc * per * sNN, N, sNN), (int)per);
    energy_kernel<<<(unsigned)per, 256, 0, s>>>(nso, N, Hc, b.X + bc * per * sT, b.Q + bc * per * sNN,
                                                b.E + 2 * bc * per);
  }
  // Y_v = FX_v + e_v X_v - X_v Q_v = FX_v - X_v (Q_v - e_v I), in place over FX
  // (energy_kernel shifted the diagonal of Q); Y^T = FX^T - (Q - e I)^T X^T
  double* Y = b.FX;
  if (!fused)
    zgemm_batched(s, N, nso, N, -1, 0, zmat_colmajor_t(b.Q, N, sNN), zmat_colmajor_t(b.X, nso, sT), 1, 0,
                  zout_colmajor_t(Y, nso, sT), (int)Gv);
  // K5: weighted grid sums; every variant lands in its own CI block, so one
  // reduction over all virtual points assembles the nci x nci matrices
  const int d = (V == 4) ? nci / 2 : nci;
  if (ctx->wd_forked)
    ctx->wd_forked = false;   // done on the side stream, joined above (hc_join)
  else
    wd_kernel<<<(unsigned)((Gv * nmn + 255) / 256), 256, 0, s>>>((int)G, V, nci, d, ctx->dGw, b.Det, b.Wd);
  {
    // Ysum and Xsum in one pair of launches (two partial blocks); the final
    // launch also forms Hmat/Smat from variants 0 and 1 only (blocks (0,0),
    // (0,1); PHFTools.f90:153-179) in its third row of workgroups
    double* part1 = b.Part + 2L * b.nchunk * nmn * lenX;
    dim3 grid((unsigned)((lenX + 255) / 256), (unsigned)b.nchunk, 2);
    dim3 fin((unsigned)((lenX * nmn + 255) / 256), 3);
    wreduce_partial_kernel<<<grid, 256, 0, s>>>((int)Gv, b.gchunk, (int)lenX, nmn, b.Wd, Y, b.X, b.Part, part1);
    wreduce_final_kernel<<<fin, 256, 0, s>>>(b.nchunk, (int)lenX, nmn, b.Part, part1, b.Ysum, b.Xsum,
                                            (int)((V == 4) ? 2 * G : G), b.Wd, b.E, b.HS);
  }
  // Gradmat[mn] = Ysum[mn] B^H, Rdmmat[mn] = Xsum[mn] B^H (C-order [nso][nso]);
  // CI block rows m < d come from variants with B = C0, rows m >= d (V = 4)
  // from variants with B = C0*
  if (!resident) {
    const int split = (V == 4) ? d * nci : nmn;
    for (int part = 0; part < 2; ++part) {
      const int m0 = part ? split : 0, cnt = part ? nmn - split : split;
      if (cnt <= 0) continue;
      const ZMat bh{b.C0, 2L * nso, 2, 0, 1, part ? 0 : 1};
      zgemm_batched(s, nso, nso, N, 1, 0, zmat_colmajor(b.Ysum + m0 * sT, nso, sT), bh, 0, 0,
                    ZOut{b.Grad + m0 * 2 * n2, 2L * nso, 2, 2 * n2, 1}, cnt);
      zgemm_batched(s, nso, nso, N, 1, 0, zmat_colmajor(b.Xsum + m0 * sT, nso, sT), bh, 0, 0,
                    ZOut{b.Rdm + m0 * 2 * n2, 2L * nso, 2, 2 * n2, 1}, cnt);
    }
  }
}

static void drop_objective_graph(phf_ctx* ctx) {
  if (ctx->og_exec) (void)hipGraphExecDestroy(ctx->og_exec);
  if (ctx->og_graph) (void)hipGraphDestroy(ctx->og_graph);
  ctx->og_exec = nullptr;
  ctx->og_graph = nullptr;
  ctx->og_key.clear();
}

static int eval_z(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision, double* hmat,
                  double* smat, double* gradmat, double* rdmmat, bool resident, double* objective = nullptr,
                  bool graph_ok = false) {
  if (!ctx) return PHF_ERR_ARG;
  ctx->r_valid = false;
  if (!hmat || !smat) return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: NULL output pointer");
  if (cmplxconj < 0 || cmplxconj > 2) return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: CmplxConj must be 0, 1 or 2");
  if (ctx->has_ham && nocc > 0 && nocc < ctx->nso && !z) return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: z is NULL");
  HIPCK(hipSetDevice(ctx->device));
  EvalBufs b;
  {
    int rc = eval_plan(ctx, nocc, precision, cmplxconj, b);
    if (rc) return rc;
  }
  hipStream_t s = ctx->stream;
  const int nso = ctx->nso, N = nocc, nci = ctx->g_nci, ny = ctx->g_ny, npg = ctx->g_npg;
  const int nmn = nci * nci;
  const long G = b.G, n2 = (long)nso * nso;
  // pinned staging: [C0 | Hmat, Smat | objective | info]
  const long Gv = b.V * G;
  const size_t nobj = 2 + 2 * (size_t)(nso - N) * N;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t c0_bytes = al(sizeof(double) * 2 * (size_t)nso * N), hs_bytes = al(sizeof(double) * 4 * (size_t)nmn);
  const size_t obj_bytes = objective ? al(sizeof(double) * nobj) : 0, info_bytes = al(sizeof(int) * (size_t)Gv);
  const size_t pin_need = c0_bytes + hs_bytes + obj_bytes + info_bytes;
  if (ctx->pin_busy) {
    HIPCK(hipStreamSynchronize(s));
    ctx->pin_busy = false;
  }
  if (ctx->pin_size < pin_need) {
    if (ctx->pin) HIPCK(hipHostFree(ctx->pin));
    ctx->pin = nullptr;
    ctx->pin_size = 0;
    HIPCK(hipHostMalloc((void**)&ctx->pin, pin_need, hipHostMallocCoherent | hipHostMallocMapped));
    ctx->pin_size = pin_need;
    void* dp = nullptr;
    HIPCK(hipHostGetDevicePointer(&dp, ctx->pin, 0));
    ctx->pin_dev = static_cast<char*>(dp);
  }
  double* c0 = reinterpret_cast<double*>(ctx->pin);
  double* hs = reinterpret_cast<double*>(ctx->pin + c0_bytes);
  double* pobj = reinterpret_cast<double*>(ctx->pin + c0_bytes + hs_bytes);
  int* info = reinterpret_cast<int*>(ctx->pin + c0_bytes + hs_bytes + obj_bytes);
  build_c0(z, nso, N, cmplxconj, c0);
  ctx->pin_busy = true;
  // one evaluation's launch sequence: eager, or captured for replay
  auto enqueue = [&]() -> int {
    HIPCK(hipMemcpyAsync(b.C0, c0, sizeof(double) * 2 * (size_t)nso * N, hipMemcpyHostToDevice, s));
    if (ctx->profiling) HIPCK(hipEventRecord(ctx->ev[0], s));
    eval_pre(ctx, b, ctx->profiling ? ctx->ev + 4 : nullptr);
    if (ctx->profiling) HIPCK(hipEventRecord(ctx->ev[1], s));
    {
      int rc = eval_gemm(ctx, b);
      if (rc) return rc;
    }
    if (ctx->profiling) HIPCK(hipEventRecord(ctx->ev[2], s));
    eval_post(ctx, b, resident);
    HIPCK(hipGetLastError());
    if (ctx->profiling) HIPCK(hipEventRecord(ctx->ev[3], s));

    // the objective kernel goes in before any copy back, and the small results
    // come back through pinned memory, so nothing waits on the host until the
    // one synchronization below
    if (objective) {   // results, Hmat/Smat and the info flags stored into pin by the kernel
      const int nel = (nso - N) * N;
      char* pd = ctx->pin_dev;
      objective_kernel<<<(std::max(nel, 1) + 255) / 256, 256, 0, s>>>(
          nso, N, b.HS, b.Ysum, b.Xsum, reinterpret_cast<double*>(pd + c0_bytes + hs_bytes), 4 * nmn, b.Info, Gv,
          reinterpret_cast<double*>(pd + c0_bytes), reinterpret_cast<int*>(pd + c0_bytes + hs_bytes + obj_bytes));
      HIPCK(hipGetLastError());
    } else {
      HIPCK(hipMemcpyAsync(hs, b.HS, sizeof(double) * 4 * (size_t)nmn, hipMemcpyDeviceToHost, s));
      HIPCK(hipMemcpyAsync(info, b.Info, sizeof(int) * (size_t)Gv, hipMemcpyDeviceToHost, s));
    }
    if (!resident) {
      HIPCK(hipMemcpyAsync(gradmat, b.Grad, 2 * (size_t)nmn * n2 * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCK(hipMemcpyAsync(rdmmat, b.Rdm, 2 * (size_t)nmn * n2 * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    return PHF_OK;
  };
  auto make_key = [&]() {
    char kb[512];
    auto env = [](const char* v) { const char* e = getenv(v); return e ? e : "-"; };
    snprintf(kb, sizeof kb, "%d|%d|%d|%d|%d|%ld|%p|%p|%p|%p|%zu|%s|%s|%s|%s|%s|%s|%s|%s|%s", precision, nocc, cmplxconj,
             (int)resident, (int)(objective != nullptr), ctx->gen, (void*)ctx->ws, (void*)ctx->pin,
             (void*)ctx->small, (void*)ctx->d_xitems, ctx->ws_size, env("PHF_H16_CFG"), env("PHF_XGROUP"),
             env("PHF_FUSED_SPLIT"), env("PHF_H16_FOLD"), env("PHF_H16_ALLOW_F32FOLD"), env("PHF_FUSED_EPILOGUE"),
             env("PHF_FUSED_SOLVE"), env("PHF_SPLIT_RJK"), env("PHF_SIDE_PRE"));
    return std::string(kb);
  };
  std::string key = graph_ok ? make_key() : std::string();
  if (graph_ok && ctx->og_exec && key == ctx->og_key) {
    HIPCK(hipGraphLaunch(ctx->og_exec, s));
  } else {
    int rc = enqueue();
    if (rc) return rc;
    if (graph_ok) key = make_key();   // the eager pass may have allocated (the exchange work list)
    if (graph_ok && key != ctx->og_failed_key) {
      // capture the same sequence (nothing runs during capture); the tuning and
      // uploads it depends on were done by the eager pass above
      drop_objective_graph(ctx);
      HIPCK(hipStreamSynchronize(s));
      hipGraph_t g = nullptr;
      int rc2 = PHF_ERR_HIP;
      hipError_t ec = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (ec == hipSuccess) {
        rc2 = enqueue();
        ec = hipStreamEndCapture(s, &g);
      }
      hipGraphExec_t ex = nullptr;
      if (rc2 == PHF_OK && ec == hipSuccess && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess) {
        ctx->og_graph = g;
        ctx->og_exec = ex;
        ctx->og_key = key;
        ctx->tune_log += "{\"objective_graph\": \"captured\"}\n";
      } else {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        ctx->og_failed_key = key;   // stay eager for this key
        ctx->tune_log += "{\"objective_graph\": \"capture failed, eager\"}\n";
      }
    }
  }
  HIPCK(hipStreamSynchronize(s));
  ctx->pin_busy = false;
  if (objective) std::memcpy(objective, pobj, sizeof(double) * nobj);
  if (ctx->profiling) {
    float t01, t12, t23, t03;
    HIPCK(hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
    HIPCK(hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
    HIPCK(hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
    HIPCK(hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]));
    ctx->stage_ms[0] = t03;   // all device work of the call
    ctx->stage_ms[1] = t01;   // rotation, overlap, inverse, density pack
    ctx->stage_ms[2] = t12;   // ERI GEMM
    ctx->stage_ms[3] = t23;   // epilogue and grid sums
    float t04, t45, t56, t61;
    HIPCK(hipEventElapsedTime(&t04, ctx->ev[7], ctx->ev[4]));
    HIPCK(hipEventElapsedTime(&t45, ctx->ev[4], ctx->ev[5]));
    HIPCK(hipEventElapsedTime(&t56, ctx->ev[5], ctx->ev[6]));
    HIPCK(hipEventElapsedTime(&t61, ctx->ev[6], ctx->ev[1]));
    ctx->stage_ms[4] = t04;   // rotation chain T_g
    ctx->stage_ms[5] = t45;   // overlap, LU inverse + det, X_g
    ctx->stage_ms[6] = t56;   // density pack rho_g^T (FP64 planes, or the fused pack + split)
    ctx->stage_ms[7] = t61;   // reduced-precision split / conversion of the densities
    ctx->nstage = 8;
    for (int q = 0; q < ctx->nblk_timed; ++q) {   // [8..10] each spin block's GEMM launch
      float tq;
      HIPCK(hipEventElapsedTime(&tq, ctx->bev[q][0], ctx->bev[q][1]));
      ctx->stage_ms[8 + q] = (q == 1 && ctx->rjk_grouped_last) ? 0.0 : tq;
      ctx->nstage = 9 + q;
    }
  }
  for (long gv = 0; gv < Gv; ++gv)
    if (info[gv] != 0) {
      const long g = gv % G;
      const long il = g / ((long)ny * npg), iy = (g / npg) % ny, ipg = g % npg;
      char buf[200];
      snprintf(buf, sizeof buf, "Error in ZGetRF: singular overlap at grid point (il=%ld, iy=%ld, ipg=%ld)%s",
               il + 1, iy + 1, ipg + 1, b.V > 1 ? (gv / G == 0 ? " (RDM)" : " (conjugate-pair density)") : "");
      return fail(ctx, PHF_ERR_SINGULAR, buf);
    }
  std::memcpy(hmat, hs, 2 * (size_t)nmn * sizeof(double));
  std::memcpy(smat, hs + 2 * nmn, 2 * (size_t)nmn * sizeof(double));
  if (b.V == 4) {
    // Hmat(d:, :d) = Hmat(:d, d:)^H, Hmat(d:, d:) = Hmat(:d, :d), same for Smat (PHFTools.f90:174-179)
    const int d = nci / 2;
    for (double* mat : {hmat, smat})
      for (int m = 0; m < d; ++m)
        for (int n = 0; n < d; ++n) {
          const size_t lo = 2 * ((size_t)(m + d) * nci + n), up = 2 * ((size_t)n * nci + m + d);
          mat[lo] = mat[up];
          mat[lo + 1] = -mat[up + 1];
          const size_t dd = 2 * ((size_t)(m + d) * nci + n + d), tl = 2 * ((size_t)m * nci + n);
          mat[dd] = mat[tl];
          mat[dd + 1] = mat[tl + 1];
        }
  }
  // Hermiticity test (PHFTools.f90:195-200)
  double w1 = 0, w2 = 0, hmax = 0, smax = 0;
  for (int m = 0; m < nci; ++m)
    for (int n = 0; n < nci; ++n) {
      const size_t a = 2 * ((size_t)m * nci + n), c = 2 * ((size_t)n * nci + m);
      w1 = std::max(w1, std::fabs(-hmat[c + 1] - hmat[a + 1]));
      w2 = std::max(w2, std::fabs(-smat[c + 1] - smat[a + 1]));
      hmax = std::max(hmax, std::hypot(hmat[a], hmat[a + 1]));
      smax = std::max(smax, std::hypot(smat[a], smat[a + 1]));
    }
  const double tol = 1e-5;
  if (w1 > hmax * tol || w2 > smax * tol) return fail(ctx, PHF_ERR_NOT_HERMITIAN, "Not Hermitian");
  ctx->r_ysum = b.Ysum;
  ctx->r_xsum = b.Xsum;
  ctx->r_nocc = N;
  ctx->r_valid = true;
  return PHF_OK;
}

extern "C" int phf_eval_objective(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision,
                                  double* e0, double* grad) {
  if (!ctx) return PHF_ERR_ARG;
  if (!e0 || !grad) return fail(ctx, PHF_ERR_ARG, "phf_eval_objective: NULL output pointer");
  if (!ctx->has_grid || ctx->g_nci != 1 || cmplxconj == 2)
    return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_eval_objective: needs a grid with nci = 1 and CmplxConj 0 or 1");
  double h[2], sm[2];
  std::vector<double> obj(2 + 2 * (size_t)(ctx->nso - nocc) * std::max(nocc, 0));
  const char* gv = getenv("PHF_GRAPH");   // PHF_GRAPH=0: always eager
  const bool graph_ok = !ctx->profiling && !(gv && atoi(gv) == 0);
  int rc = eval_z(ctx, z, cmplxconj, nocc, precision, h, sm, nullptr, nullptr, true, obj.data(), graph_ok);
  if (rc) return rc;
  if (!(std::fabs(sm[0]) > 1e-6))   // EigenSolver keeps no state (PHFTools.py:8-18, Tol 1e-6)
    return fail(ctx, PHF_ERR_SINGULAR, "phf_eval_objective: |Re Smat| <= 1e-6, no projected state");
  e0[0] = obj[0];
  std::memcpy(grad, obj.data() + 2, sizeof(double) * (obj.size() - 2));
  return PHF_OK;
}

extern "C" int phf_buildhsg_z(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision, double* hmat,
                              double* smat, double* gradmat, double* rdmmat) {
  if (!gradmat || !rdmmat) return fail(ctx, PHF_ERR_ARG, "phf_buildhsg: NULL output pointer");
  return eval_z(ctx, z, cmplxconj, nocc, precision, hmat, smat, gradmat, rdmmat, false);
}

extern "C" int phf_eval_resident(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision,
                                 double* hmat, double* smat) {
  return eval_z(ctx, z, cmplxconj, nocc, precision, hmat, smat, nullptr, nullptr, true);
}

extern "C" int phf_buildhsg(phf_ctx* ctx, const double* ref, const double* z, const int* ngrid, int nci,
                            int cmplxconj, int ncisp, int nocc, const double* r1, int nr1, const double* r2,
                            int nr2, const double* rpg, int npg, const double* weightsp,
                            const double* weightpg, int ncipg, const double* roota, int na,
                            const double* rootb, int nb, const double* rooty, int ny_, int qj, int sp,
                            int precision, double* hmat, double* smat, double* gradmat, double* rdmmat) {
  int rc = phf_set_grid(ctx, ref, ngrid, nci, ncisp, r1, nr1, r2, nr2, rpg, npg, weightsp, weightpg, ncipg, roota,
                        na, rootb, nb, rooty, ny_, qj, sp);
  if (rc) return rc;
  return phf_buildhsg_z(ctx, z, cmplxconj, nocc, precision, hmat, smat, gradmat, rdmmat);
}

extern "C" int phf_localg_vo(phf_ctx* ctx, const double* smat, const double* e0, const double* fsp, int ncisp,
                             const double* fpg, int ncipg, const double* fk, int nfk, int cmplxconj, double* gvo) {
  if (!ctx) return PHF_ERR_ARG;
  if (!ctx->r_valid) return fail(ctx, PHF_ERR_ARG, "phf_localg_vo: no resident evaluation (phf_eval_resident)");
  if (!smat || !e0 || !fsp || !fpg || !fk || !gvo || ncisp <= 0 || ncipg <= 0 || nfk < 1)
    return fail(ctx, PHF_ERR_ARG, "phf_localg_vo: bad arguments");
  if (cmplxconj == 2) return fail(ctx, PHF_ERR_UNSUPPORTED, "phf_localg_vo: CmplxConj=2");
  const int nci = ctx->g_nci;
  if (ncisp * ncipg > nci) return fail(ctx, PHF_ERR_ARG, "phf_localg_vo: ncisp*ncipg > nci");
  HIPCK(hipSetDevice(ctx->device));
  auto C = [](const double* p, int i) { return cplx(p[2 * i], p[2 * i + 1]); };
  std::vector<int> tmn;
  std::vector<double> tcoef;
  cplx ovlp(0, 0);
  const cplx fk0 = C(fk, 0);
  for (int ni = 0; ni < ncisp; ++ni)
    for (int nj = 0; nj < ncisp; ++nj)
      for (int p = 0; p < ncipg; ++p)
        for (int q = 0; q < ncipg; ++q) {
          const int m = ni + p * ncisp, n = nj + q * ncisp;
          const cplx f = std::conj(C(fsp, ni)) * C(fsp, nj) * std::conj(C(fpg, p)) * C(fpg, q) * std::conj(fk0) * fk0;
          tmn.push_back(m * nci + n);
          tcoef.push_back(f.real());
          tcoef.push_back(f.imag());
          ovlp += f * C(smat, m * nci + n);
        }
  const int nterm = (int)tmn.size();
  const int nso = ctx->nso, N = ctx->r_nocc;
  const size_t need = 256 + sizeof(int) * nterm + 256 + sizeof(double) * 2 * nterm + 256 +
                      sizeof(double) * 2 * (size_t)(nso - N) * N;
  if (ctx->small_size < need) {
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (ctx->small) HIPCK(hipFree(ctx->small));
    ctx->small = nullptr;
    ctx->small_size = 0;
    HIPCK(hipMalloc(&ctx->small, need));
    ctx->small_size = need;
  }
  Arena sa{ctx->small};
  int* dmn = sa.take<int>(nterm);
  double* dcf = sa.take<double>(2 * nterm);
  double* dout = sa.take<double>(2 * (size_t)(nso - N) * N);
  hipStream_t s = ctx->stream;
  HIPCK(hipMemcpyAsync(dmn, tmn.data(), sizeof(int) * nterm, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dcf, tcoef.data(), sizeof(double) * 2 * nterm, hipMemcpyHostToDevice, s));
  const int nel = (nso - N) * N;
  if (nel > 0)
    localg_vo_kernel<<<(nel + 255) / 256, 256, 0, s>>>(nso, N, nterm, dmn, dcf, ctx->r_ysum, ctx->r_xsum, e0[0],
                                                       e0[1], ovlp.real(), ovlp.imag(), dout);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(gvo, dout, sizeof(double) * 2 * (size_t)nel, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  return PHF_OK;
}

extern "C" int phf_set_profiling(phf_ctx* ctx, int on) {
  if (!ctx) return PHF_ERR_ARG;
  HIPCK(hipSetDevice(ctx->device));
  if (on && !ctx->ev[0]) {
    for (auto& e : ctx->ev) HIPCK(hipEventCreateWithFlags(&e, kTimingEventFlags));
    for (auto& pr : ctx->bev)
      for (auto& e : pr) HIPCK(hipEventCreateWithFlags(&e, kTimingEventFlags));
  }
  ctx->profiling = on != 0;
  ctx->nstage = 0;
  return PHF_OK;
}

extern "C" int phf_stage_times(const phf_ctx* ctx, double* ms, int nmax) {
  if (!ctx || !ms) return 0;
  const int n = std::min(nmax, ctx->nstage);
  for (int i = 0; i < n; ++i) ms[i] = ctx->stage_ms[i];
  return n;
}

extern "C" int phf_localg(phf_ctx* ctx, const double* gradmat, const double* rdmmat, const double* smat,
                          const double* e0, const double* fsp, int ncisp, const double* fpg, int ncipg,
                          const double* fk, int nfk, int cmplxconj, int nso, int nci, double* g) {
  if (!ctx) return PHF_ERR_ARG;
  if (!gradmat || !rdmmat || !smat || !e0 || !fsp || !fpg || !fk || !g || nso <= 0 || nci <= 0 || ncisp <= 0 ||
      ncipg <= 0 || nfk < 1)
    return fail(ctx, PHF_ERR_ARG, "phf_localg: bad arguments");
  const int d = ncisp * ncipg;
  if (cmplxconj == 2 && (nfk < 2 || 2 * d > nci))
    return fail(ctx, PHF_ERR_ARG, "phf_localg: CmplxConj=2 needs fk(2) and nci >= 2*ncisp*ncipg");
  if (d > nci) return fail(ctx, PHF_ERR_ARG, "phf_localg: nci < ncisp*ncipg");
  HIPCK(hipSetDevice(ctx->device));
  auto C = [](const double* p, int i) { return cplx(p[2 * i], p[2 * i + 1]); };
  std::vector<int> tmn, tconj;
  std::vector<double> tcoef;
  cplx ovlp(0, 0);
  auto term = [&](int m, int n, cplx c, int cj) {
    tmn.push_back(m * nci + n);
    tconj.push_back(cj);
    tcoef.push_back(c.real());
    tcoef.push_back(c.imag());
  };
  const cplx fk0 = C(fk, 0), fk1 = nfk > 1 ? C(fk, 1) : cplx(0, 0);
  for (int ni = 0; ni < ncisp; ++ni)
    for (int nj = 0; nj < ncisp; ++nj)
      for (int p = 0; p < ncipg; ++p)
        for (int q = 0; q < ncipg; ++q) {
          const int m = ni + p * ncisp, n = nj + q * ncisp;
          const cplx fac = std::conj(C(fsp, ni)) * C(fsp, nj) * std::conj(C(fpg, p)) * C(fpg, q);
          const cplx f11 = fac * std::conj(fk0) * fk0;
          term(m, n, f11, 0);
          ovlp += f11 * C(smat, m * nci + n);
          if (cmplxconj == 2) {
            term(m + d, n + d, fac * std::conj(fk1) * fk1, 1);
            ovlp += fac * std::conj(fk1) * fk1 * C(smat, (m + d) * nci + n + d);
            term(m, n + d, fac * std::conj(fk0) * fk1, 0);
            ovlp += fac * std::conj(fk0) * fk1 * C(smat, m * nci + n + d);
            term(m + d, n, fac * std::conj(fk0) * fk1, 1);
            ovlp += fac * std::conj(fk1) * fk0 * C(smat, (m + d) * nci + n);
          }
        }
  const long n2 = (long)nso * nso;
  const int nterm = (int)tmn.size();
  const size_t mats = 2 * (size_t)nci * nci * n2;
  Arena a0{nullptr};
  a0.take<double>(mats); a0.take<double>(mats); a0.take<double>(2 * n2);
  a0.take<int>(nterm); a0.take<int>(nterm); a0.take<double>(2 * nterm);
  int rc = ensure_ws(ctx, a0.off + 256);
  if (rc) return rc;
  Arena a1{ctx->ws};
  double* dG = a1.take<double>(mats);
  double* dR = a1.take<double>(mats);
  double* dOut = a1.take<double>(2 * n2);
  int* dmn = a1.take<int>(nterm);
  int* dcj = a1.take<int>(nterm);
  double* dcf = a1.take<double>(2 * nterm);
  hipStream_t s = ctx->stream;
  HIPCK(hipMemcpyAsync(dG, gradmat, mats * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dR, rdmmat, mats * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dmn, tmn.data(), nterm * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dcj, tconj.data(), nterm * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dcf, tcoef.data(), 2 * nterm * sizeof(double), hipMemcpyHostToDevice, s));
  localg_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, s>>>(n2, nterm, dmn, dcj, dcf, dG, dR, e0[0], e0[1],
                                                             ovlp.real(), ovlp.imag(), dOut);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(g, dOut, 2 * n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  return PHF_OK;
}
