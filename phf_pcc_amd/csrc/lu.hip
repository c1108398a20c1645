// Batched complex FP64 inverse + determinant of the N x N overlaps M_g = C0^H T_g.
//
// Restates InvertC / determinantC (Wrappers.f90:195-231, 295-335): partial
// (row) pivoting on the largest |Re|+|Im| in the column (IZAMAX's measure,
// first index on ties), det = prod of pivots with a sign flip for every
// PIV(k) != k, info = first exactly-zero pivot.  One workgroup per grid
// point.  N <= 32 runs one wave per matrix with the rows in registers
// (lu_inverse_wave_kernel), 32 < N <= 64 two waves (lu_inverse_wave2_kernel);
// larger N keeps the matrix in LDS, inverted in
// place by Gauss-Jordan elimination (its pivots are the U_kk of the LU
// factorization), then the row interchanges are undone as column
// interchanges in reverse order.
#include "phf_common.h"

namespace {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 crecip(double2 a) {
  // Smith's algorithm for 1/a
  if (fabs(a.x) >= fabs(a.y)) {
    const double r = a.y / a.x, d = a.x + a.y * r;
    return make_double2(1.0 / d, -r / d);
  } else {
    const double r = a.x / a.y, d = a.y + a.x * r;
    return make_double2(r / d, -1.0 / d);
  }
}

template <int NMAX>
__global__ __launch_bounds__(256) void lu_inverse_kernel(int n, const double* __restrict__ M, long mbs,
                                                         double* __restrict__ Minv, long ibs,
                                                         double* __restrict__ det, int* __restrict__ info) {
  __shared__ double2 a[NMAX * NMAX];   // row-major a[i*n + j]
  __shared__ double2 col[NMAX];
  __shared__ int piv[NMAX];
  __shared__ int s_p, s_info;
  const long g = blockIdx.x;
  const int tid = threadIdx.x;
  const double* Mg = M + g * mbs;
  for (int idx = tid; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;   // global is column-major
    a[i * n + j] = make_double2(Mg[2 * idx], Mg[2 * idx + 1]);
  }
  if (tid == 0) s_info = 0;
  __syncthreads();

  double2 detv = make_double2(1.0, 0.0);
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      double best = -1.0;
      int bi = n;
      for (int i = k + tid; i < n; i += 64) {
        const double2 v = a[i * n + k];
        const double c = fabs(v.x) + fabs(v.y);
        if (c > best) { best = c; bi = i; }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (tid == 0) {
        s_p = bi;
        piv[k] = bi;
        if (!(best > 0.0)) s_info = k + 1;
      }
    }
    __syncthreads();
    if (s_info != 0) break;
    const int p = s_p;
    if (p != k) {
      for (int j = tid; j < n; j += 256) {
        const double2 t = a[k * n + j];
        a[k * n + j] = a[p * n + j];
        a[p * n + j] = t;
      }
    }
    __syncthreads();
    const double2 pv = a[k * n + k];
    if (tid == 0) {
      detv = cmul(detv, pv);
      if (p != k) detv = make_double2(-detv.x, -detv.y);
    }
    const double2 pinv = crecip(pv);
    for (int i = tid; i < n; i += 256) col[i] = a[i * n + k];
    __syncthreads();
    for (int j = tid; j < n; j += 256)
      a[k * n + j] = (j == k) ? pinv : cmul(a[k * n + j], pinv);
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += 256) {
      const int i = idx / n, j = idx % n;
      if (i == k) continue;
      if (j == k) {
        const double2 t = cmul(col[i], pinv);
        a[i * n + k] = make_double2(-t.x, -t.y);
      } else {
        const double2 t = cmul(col[i], a[k * n + j]);
        a[i * n + j] = make_double2(a[i * n + j].x - t.x, a[i * n + j].y - t.y);
      }
    }
    __syncthreads();
  }

  if (s_info == 0) {
    for (int k = n - 1; k >= 0; --k) {
      const int p = piv[k];
      if (p != k) {
        for (int i = tid; i < n; i += 256) {
          const double2 t = a[i * n + k];
          a[i * n + k] = a[i * n + p];
          a[i * n + p] = t;
        }
      }
      __syncthreads();
    }
  }
  double* Ig = Minv + g * ibs;
  for (int idx = tid; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    const double2 v = (s_info == 0) ? a[i * n + j] : make_double2(0.0, 0.0);
    Ig[2 * idx] = v.x;
    Ig[2 * idx + 1] = v.y;
  }
  if (tid == 0) {
    det[2 * g] = (s_info == 0) ? detv.x : 0.0;
    det[2 * g + 1] = (s_info == 0) ? detv.y : 0.0;
    info[g] = s_info;
  }
}

// One wave per matrix for n <= NMAX <= 64: each row of M is held by
// LPR = 64 / NMAX lanes, lane (h, i) = h NMAX + i holding the CPL = NMAX / LPR
// storage columns [h CPL, (h+1) CPL) of row i in registers.  Gauss-Jordan
// with the pivot column always in storage column 0: each step writes every
// row's update shifted one place to the left (the last column of part h comes
// from the first of part h+1, one shuffle) and appends the new inverse column
// in storage column NMAX-1, so every register index is a compile-time
// constant; after n steps the inverse's column m sits in storage column
// NMAX-n+m.  Rows never move: each lane tracks the position its row holds in
// the row-interchanged matrix, so the pivot choice (largest |Re|+|Im| among
// the rows at positions k..n-1, first position on ties), the determinant's
// sign flips and info are those of the algorithm above.  The pivot row reaches
// the other lanes through LDS.  Each step's pivot search (a butterfly over the
// wave) starts as soon as the new column 0 is updated and overlaps the rest of
// the row update: with one wave per matrix and fewer matrices than SIMDs,
// nothing else hides its latency.  With B = P M (P the row interchanges), row
// i holds row pos_i of B^{-1}, and M^{-1} = B^{-1} P, so storage column m of
// row i lands in column sigma(m) of M^{-1}, sigma(m) = the row that ended at
// position m.
//
// The routine runs on one wave (lane = threadIdx.x % 64) with LDS scratch
// s_row[NMAX], s_sig[64]; sync() orders the wave's LDS accesses between the
// steps (a workgroup barrier in a one-wave workgroup; in a larger workgroup
// where one wave runs it, a wavefront-scope fence).  put(r, c, v) receives
// M^{-1}(r, c) for r, c < n (zero when M is singular), once each; the
// determinant and info go to *det_out (complex) and *info_out from lane 0.
// One step of the pivot search's wave reduction: each lane compares its
// candidate (c, key) with the lane CTRL names (a DPP row shift or row
// broadcast; lanes without a source, or in rows outside ROWMASK, compare with
// themselves) and keeps the larger c, ties to the smaller key.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void pivot_step(double& c, int& key) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, c);
  const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xf, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xf, false);
  const int ok = __builtin_amdgcn_update_dpp(key, key, CTRL, ROWMASK, 0xf, false);
  const double oc = __builtin_bit_cast(double, ((unsigned long long)(unsigned)ohi << 32) | (unsigned)olo);
  if (oc > c || (oc == c && ok < key)) {
    c = oc;
    key = ok;
  }
}

// The best (c, key) over the wave in every lane: row shifts by 1, 2, 4, 8
// leave each 16-lane row's best in its lane 15, two row broadcasts carry
// them to lane 63, whose pair every lane then reads.  The choice (largest c,
// ties to the smallest key; keys are distinct) does not depend on the order
// of the comparisons, so it equals the butterfly's.
__device__ __forceinline__ void wave_best_pivot(double& c, int& key) {
  pivot_step<0x111, 0xf>(c, key);   // row_shr:1
  pivot_step<0x112, 0xf>(c, key);   // row_shr:2
  pivot_step<0x114, 0xf>(c, key);   // row_shr:4
  pivot_step<0x118, 0xf>(c, key);   // row_shr:8
  pivot_step<0x142, 0xa>(c, key);   // row_bcast:15 into rows 1 and 3
  pivot_step<0x143, 0xc>(c, key);   // row_bcast:31 into rows 2 and 3
  const unsigned long long u = __builtin_bit_cast(unsigned long long, c);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);
  c = __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
  key = __builtin_amdgcn_readlane(key, 63);
}

// get(i, j) returns M(i, j) (complex) for i, j < n.
template <int NMAX, typename Get, typename Sync, typename Put>
__device__ __forceinline__ void lu_wave_inverse(int n, Get get, double2* s_row, int* s_sig, Sync sync, Put put,
                                                double* det_out, int* info_out) {
  constexpr int LPR = 64 / NMAX, CPL = NMAX / LPR;
  const int lane = threadIdx.x & 63, row = lane % NMAX, part = lane / NMAX;
  const bool real_row = row < n;
  double2 a[CPL];
#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    const int j = part * CPL + t;
    a[t] = make_double2(0.0, 0.0);
    if (real_row && j < n) a[t] = get(row, j);
  }
  int pos = row;
  bool done = !real_row;   // row already used as a pivot, or no row of M
  double2 detv = make_double2(1.0, 0.0);
  int sinfo = 0;
  auto search = [&](double& c, int& key) {
    c = (done || part > 0) ? -1.0 : fabs(a[0].x) + fabs(a[0].y);
    if (c != c) c = -1.0;   // a NaN row is never a pivot; keeps the reduction wave-uniform
    key = (pos << 6) | lane;
    wave_best_pivot(c, key);
  };
  double c;
  int key;
  search(c, key);
  for (int k = 0; k < n; ++k) {
    if (!(c > 0.0)) { sinfo = k + 1; break; }
    const int pl = key & 63, pp = key >> 6;   // pl: a part-0 lane, i.e. the pivot row
    const bool is_p = row == pl;
    if (is_p) {
#pragma unroll
      for (int t = 0; t < CPL; ++t) s_row[part * CPL + t] = a[t];
    }
    if (pos == k) pos = pp;
    if (is_p) { pos = k; done = true; }
    // column 0 of this row (part 0's register 0) and the next part's first column
    double2 f = a[0], nxt = a[0];
    if constexpr (LPR == 2) {
      // f: part 1 takes part 0's column 0 (lane - 32), nxt: part 0 takes part
      // 1's (lane + 32): one swap of the wave's halves per 32-bit word
      const unsigned long long ux = __builtin_bit_cast(unsigned long long, a[0].x);
      const unsigned long long uy = __builtin_bit_cast(unsigned long long, a[0].y);
      unsigned w[4] = {(unsigned)ux, (unsigned)(ux >> 32), (unsigned)uy, (unsigned)(uy >> 32)};
      unsigned fw[4], nw[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const auto r = __builtin_amdgcn_permlane32_swap(w[q], w[q], false, false);
        fw[q] = r[0];   // lanes 32-63: lane - 32's word; lanes 0-31: their own
        nw[q] = r[1];   // lanes 0-31: lane + 32's word; lanes 32-63: their own
      }
      f = make_double2(__builtin_bit_cast(double, ((unsigned long long)fw[1] << 32) | fw[0]),
                       __builtin_bit_cast(double, ((unsigned long long)fw[3] << 32) | fw[2]));
      nxt = make_double2(__builtin_bit_cast(double, ((unsigned long long)nw[1] << 32) | nw[0]),
                         __builtin_bit_cast(double, ((unsigned long long)nw[3] << 32) | nw[2]));
    } else if constexpr (LPR > 2) {
      f = make_double2(__shfl(a[0].x, row), __shfl(a[0].y, row));
      const int src = (part < LPR - 1) ? lane + NMAX : lane;
      nxt = make_double2(__shfl(a[0].x, src), __shfl(a[0].y, src));
    }
    sync();
    const double2 pv = s_row[0];
    detv = cmul(detv, pv);
    if (pp != k) detv = make_double2(-detv.x, -detv.y);
    const double2 pinv = crecip(pv);
    // pivot row: a <- pinv * row; other rows: a <- a - (f pinv) * row
    const double2 gm = is_p ? make_double2(-pinv.x, -pinv.y) : cmul(f, pinv);
    const double s0 = is_p ? 0.0 : 1.0;
    // storage column j = part CPL + t takes column j + 1 (register t + 1, or
    // the next part's register 0); column NMAX - 1 takes the new inverse column
    const bool last = part == LPR - 1;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      if (t == CPL - 1 && (LPR == 1 || last)) {
        a[t] = is_p ? pinv : make_double2(-gm.x, -gm.y);
      } else {
        const double2 old = (t < CPL - 1) ? a[t + 1 < CPL ? t + 1 : t] : nxt;
        const double2 r = s_row[part * CPL + t + 1];
        a[t] = make_double2(fma(-gm.x, r.x, fma(gm.y, r.y, s0 * old.x)),
                            fma(-gm.x, r.y, fma(-gm.y, r.x, s0 * old.y)));
      }
      if (t == 0) search(c, key);
    }
    sync();
  }
  if (part == 0) s_sig[pos] = row;
  sync();
  if (real_row) {
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      const int m = part * CPL + t + n - NMAX;
      if (m >= 0) put(pos, s_sig[m], sinfo ? make_double2(0.0, 0.0) : a[t]);
    }
  }
  if (lane == 0) {
    det_out[0] = sinfo ? 0.0 : detv.x;
    det_out[1] = sinfo ? 0.0 : detv.y;
    *info_out = sinfo;
  }
}

template <int NMAX>
__global__ __launch_bounds__(64) void lu_inverse_wave_kernel(int n, const double* __restrict__ M, long mbs,
                                                             double* __restrict__ Minv, long ibs,
                                                             double* __restrict__ det, int* __restrict__ info) {
  __shared__ double2 s_row[NMAX];
  __shared__ int s_sig[64];
  const long g = blockIdx.x;
  double* Ig = Minv + g * ibs;
  // M^{-1} column-major
  const double* Mg = M + g * mbs;   // column-major
  lu_wave_inverse<NMAX>(n, [&](int i, int j) {
                          const long off = 2 * ((long)j * n + i);
                          return make_double2(Mg[off], Mg[off + 1]);
                        },
                        s_row, s_sig, [] { __syncthreads(); },
                        [&](int r, int c, double2 v) {
                          const long off = 2 * ((long)c * n + r);
                          Ig[off] = v.x;
                          Ig[off + 1] = v.y;
                        },
                        det + 2 * g, info + g);
}

// Inverse and solve for one grid point per 128-thread workgroup (n <= 32):
// wave 0 inverts M_g (lu_wave_inverse, M^{-1} kept in LDS), then both waves
// form X_g = T_g M_g^{-1} (NSO x n, column-major complex) on the FP64 matrix
// cores, three real products per complex one, wave w taking the 16-row
// blocks w, w + 2, ...; k-step t takes c = 4 t + lane / 16.  Two waves, not
// more: with two waves per SIMD a CU holds four workgroups, so up to 1024
// points' workgroups are resident at once (the inverse is latency-bound; a
// second round of workgroups would double it).  This replaces the
// inverse's M^{-1} store and the solve's reads of it.  A singular M_g
// (info > 0) gives X_g = 0, as the zero inverse does.
constexpr int LS_LD = 33;   // row stride (complex) of M^{-1} in LDS
__global__ __launch_bounds__(128, 2) void lu_inverse_solve_kernel(int n, int nso, const double* __restrict__ M, long mbs,
                                                               const double* __restrict__ T, long sT,
                                                               double* __restrict__ X, long sX,
                                                               double* __restrict__ det, int* __restrict__ info,
                                                               const double* __restrict__ U, int npg) {
  __shared__ double2 s_row[32];
  __shared__ int s_sig[64];
  __shared__ double2 s_inv[32 * LS_LD];
  const long g = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l16 = lane & 15, kq = lane >> 4;
  const double* Tg = T + g * sT;
  double* Xg = X + g * sX;
  const int nrb = (nso + 15) >> 4;
  // A operands of one 16-row block, k-step t: T(16 rb + l16, 4 t + kq), loaded
  // unconditionally (clamped address; zeroed where used)
  auto load_t = [&](int rb, double2 (&tv)[8]) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int p = 16 * rb + l16, c = 4 * t + kq;
      const bool v = p < nso && c < n;
      tv[t] = *reinterpret_cast<const double2*>(Tg + (v ? 2L * (p + (long)c * nso) : 0));
    }
  };
  // M_g itself, or on spin-rotation grids (U != nullptr) the combination
  // sum_{s,t} U_g[s][t] M_st[g % npg] of the four N x N products M_st (at
  // (st npg + ipg) mbs), in the order of spinrot_kernel
  auto get = [&](int i, int j) {
    const long off = 2 * ((long)j * n + i);
    if (!U) return make_double2(M[g * mbs + off], M[g * mbs + off + 1]);
    const double* u = U + 8 * g;
    const long ipg = g % npg;
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const double ur = u[2 * st], ui = u[2 * st + 1];
      const double* q = M + (st * npg + ipg) * mbs + off;
      re += ur * q[0] - ui * q[1];
      im += ur * q[1] + ui * q[0];
    }
    return make_double2(re, im);
  };
  if (w == 0)
    lu_wave_inverse<32>(n, get, s_row, s_sig,
                        [] {
                          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                          __builtin_amdgcn_wave_barrier();
                        },
                        [&](int r, int c, double2 v) { s_inv[r * LS_LD + c] = v; }, det + 2 * g, info + g);
  __syncthreads();
  // B operands (the same for every row block): M^{-1}(4 t + kq, 16 j + l16)
  double2 bq[8][2];
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = 4 * t + kq, b = 16 * j + l16;
      const double2 q = s_inv[c * LS_LD + b];
      bq[t][j] = (c < n && b < n) ? q : make_double2(0.0, 0.0);
    }
  // row blocks w, w + 2, ...: a block's T loads are issued together, then its products
#pragma unroll 1
  for (int rb = w; rb < nrb; rb += 2) {
    double2 tv[8];
    load_t(rb, tv);
    d4_t p1[2], p2[2], p3[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) p1[j] = p2[j] = p3[j] = d4_t{0.0, 0.0, 0.0, 0.0};
    const int p = 16 * rb + l16;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bool va = p < nso && 4 * t + kq < n;
      const double are = va ? tv[t].x : 0.0, aim = va ? tv[t].y : 0.0;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const double bre = bq[t][j].x, bim = bq[t][j].y;
        p1[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, p1[j], 0, 0, 0);
        p2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bim, p2[j], 0, 0, 0);
        p3[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(are + aim, bre + bim, p3[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int pr = 16 * rb + kq + 4 * r, b = 16 * j + l16;
        if (pr >= nso || b >= n) continue;
        *reinterpret_cast<double2*>(Xg + 2L * (pr + (long)b * nso)) =
            make_double2(p1[j][r] - p2[j][r], p3[j][r] - p1[j][r] - p2[j][r]);
      }
  }
}

// 32 < n <= 64: the same algorithm with two waves per matrix, row i held by
// lane i of both (wave h the storage columns [32 h, 32 h + 32)).  What the
// one-wave form moves by shuffles crosses between the waves through LDS:
// column 0 of every row (wave 0's register 0), the first column of wave 1
// (the shifted-in column of wave 0), and the pivot search, which wave 0 alone
// runs over all rows.  Two barriers per step, as in the one-wave form.
__global__ __launch_bounds__(128) void lu_inverse_wave2_kernel(int n, const double* __restrict__ M, long mbs,
                                                               double* __restrict__ Minv, long ibs,
                                                               double* __restrict__ det, int* __restrict__ info) {
  constexpr int NMAX = 64, CPL = 32;
  __shared__ double2 s_row[NMAX];
  __shared__ double2 s_c0[NMAX], s_c1[NMAX];
  __shared__ int s_sig[NMAX];
  __shared__ double s_best;
  __shared__ int s_key;
  const long g = blockIdx.x;
  const int tid = threadIdx.x, row = tid & 63, part = tid >> 6;
  const double* Mg = M + g * mbs;
  const bool real_row = row < n;
  double2 a[CPL];
#pragma unroll
  for (int t = 0; t < CPL; ++t) {
    const int j = part * CPL + t;
    a[t] = make_double2(0.0, 0.0);
    if (real_row && j < n) {
      const long off = 2 * ((long)j * n + row);   // global is column-major
      a[t] = make_double2(Mg[off], Mg[off + 1]);
    }
  }
  int pos = row;
  bool done = !real_row;
  double2 detv = make_double2(1.0, 0.0);
  int sinfo = 0;
  // wave 0 searches column 0 over all rows and publishes the result
  auto search = [&]() {
    if (part == 0) {
      double c = done ? -1.0 : fabs(a[0].x) + fabs(a[0].y);
      if (c != c) c = -1.0;   // a NaN row is never a pivot; keeps the reduction wave-uniform
      int key = (pos << 6) | row;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double oc = __shfl_xor(c, off);
        const int ok = __shfl_xor(key, off);
        if (oc > c || (oc == c && ok < key)) { c = oc; key = ok; }
      }
      if (row == 0) {
        s_best = c;
        s_key = key;
      }
    }
  };
  search();
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double c = s_best;
    const int key = s_key;
    if (!(c > 0.0)) { sinfo = k + 1; break; }
    const int pl = key & 63, pp = key >> 6;
    const bool is_p = row == pl;
    if (is_p) {
#pragma unroll
      for (int t = 0; t < CPL; ++t) s_row[part * CPL + t] = a[t];
    }
    if (part == 0) s_c0[row] = a[0];
    else s_c1[row] = a[0];
    if (pos == k) pos = pp;
    if (is_p) { pos = k; done = true; }
    __syncthreads();
    const double2 pv = s_row[0];
    detv = cmul(detv, pv);
    if (pp != k) detv = make_double2(-detv.x, -detv.y);
    const double2 pinv = crecip(pv);
    const double2 f = s_c0[row];
    const double2 nxt = s_c1[row];   // used by wave 0 only
    const double2 gm = is_p ? make_double2(-pinv.x, -pinv.y) : cmul(f, pinv);
    const double s0 = is_p ? 0.0 : 1.0;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      if (t == CPL - 1 && part == 1) {
        a[t] = is_p ? pinv : make_double2(-gm.x, -gm.y);
      } else {
        const double2 old = (t < CPL - 1) ? a[t + 1 < CPL ? t + 1 : t] : nxt;
        const double2 r = s_row[part * CPL + t + 1];
        a[t] = make_double2(fma(-gm.x, r.x, fma(gm.y, r.y, s0 * old.x)),
                            fma(-gm.x, r.y, fma(-gm.y, r.x, s0 * old.y)));
      }
      if (t == 0) search();
    }
    __syncthreads();
  }
  if (part == 0) s_sig[pos] = row;
  __syncthreads();
  if (real_row) {
    double* Ig = Minv + g * ibs;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
      const int m = part * CPL + t + n - NMAX;
      if (m >= 0) {
        const long off = 2 * ((long)s_sig[m] * n + pos);
        Ig[off] = sinfo ? 0.0 : a[t].x;
        Ig[off + 1] = sinfo ? 0.0 : a[t].y;
      }
    }
  }
  if (tid == 0) {
    det[2 * g] = sinfo ? 0.0 : detv.x;
    det[2 * g + 1] = sinfo ? 0.0 : detv.y;
    info[g] = sinfo;
  }
}

}  // namespace

int lu_inverse_max_n() { return 96; }

// X_g = T_g M_g^{-1} with the inverse's det and info, for n <= 32 (T_g, X_g:
// NSO x n column-major complex, batch strides sT, sX doubles).  With U, M
// holds the 4 npg spin-rotation products and M_g is their combination.
void lu_inverse_solve_batched(hipStream_t s, int n, int nso, const double* M, long mbs, const double* T, long sT,
                              double* X, long sX, double* det, int* info, int batch, const double* U, int npg) {
  lu_inverse_solve_kernel<<<batch, 128, 0, s>>>(n, nso, M, mbs, T, sT, X, sX, det, info, U, npg > 0 ? npg : 1);
}

void lu_inverse_batched(hipStream_t s, int n, const double* M, long mbs, double* Minv, long ibs,
                        double* det, int* info, int batch) {
  if (n <= 16)
    lu_inverse_wave_kernel<16><<<batch, 64, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
  else if (n <= 32)
    lu_inverse_wave_kernel<32><<<batch, 64, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
  else if (n <= 64)
    lu_inverse_wave2_kernel<<<batch, 128, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
  else
    lu_inverse_kernel<96><<<batch, 256, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
}
