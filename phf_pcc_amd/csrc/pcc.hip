// Projected-CC grid consumer, zeroth order: Olap0 and EOlap0 of PbarHbarOlap
// (SUCCSD_Boson_cmplx/FPUCC.f90:14-272) with the lambda0 / eps0 part of its
// per-point EvalKernel (FPUCC.f90:529-602, TrunODE = 1) on the GPU.
//
// Per grid point g (rotation R_g = R1[ia] R2[iy] Rpg[ipg], times Rk(2) in the
// second pass of ncik = 2):
//   ProjT1 (FPUCC_Tools.f90:676-695): MO2 = R [I; T1]; L = MO2[oo], Y = MO2[vo];
//     U0 = det L, U1 = Y L^-1 (the Thouless split of R e^{T1} |Phi>)
//   TransT2 (PCCCI.f90:98-245): V = e^{-U1} R, U = V T2 V^-1 as four index
//     transforms (IntTran4Sim, IntTrans.f90:486-572; only the vvoo block of
//     the input is nonzero), then E0, F and the second-order CI coefficients
//     C0, C1, C2 of e^U |Phi>
//   CI2CC + ASymm2 (PCCCI.f90:70-96, CCRes.f90:31-69), DSTT2 (W0 U0, W1 + U1)
//   CCEnergy (CCRes.f90:9-29) with the Fock matrix of (HOne, HTwo)
//   lambda0 = W0 U0, eps0 = (ECorr + ESCF + ENuc) lambda0
// and Olap0 = sum_g w_g lambda0_g, EOlap0 = sum_g w_g eps0_g with the
// weights of PbarHbarOlap (Weightsp, Weightpg, Wigner D, fsp, fpg, fk).
//
// Mapping: the rotation products and the four index transforms are batched
// complex FP64 GEMMs (zgemm_batched, MFMA f64); L^-1 and V^-1 the batched
// pivoted LU of the overlap inverses; the second-order terms, CI2CC and the
// CC energy one 256-thread workgroup per grid point (output-stationary loops
// over the transformed tensor, which stays in L2 at these sizes); the grid
// sum one workgroup in a fixed order.  Grid points are processed in chunks
// sized to a few GB of workspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <string>
#include <vector>

#include "phf_internal.h"

namespace {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ inline double2 ld2(const double* p, long i) { return reinterpret_cast<const double2*>(p)[i]; }

constexpr int kThreads = 256;
constexpr long kKws = 4096L * 32 * 32;   // complex elements of the split-contraction workspace

// Block-wide sum of one complex value per thread (fixed order: deterministic).
__device__ double2 block_sum(double2 v, double2* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = cadd(red[t], red[t + s]);
    __syncthreads();
  }
  const double2 out = red[0];
  __syncthreads();
  return out;
}

// Fock of (HOne, HTwo) restricted to what CCEnergy reads: fov[i][a] =
// HOne(i, N+a) + sum_k HTwo(i,k,N+a,k), and ESCF = sum_i HOne(i,i) +
// sum_ij HTwo(i,j,i,j)/2 + ENuc (MKFock_MO, UHF.f90:461-497), conjugated
// for the second pass when conj is set.  One workgroup.
__global__ void pcc_fock_kernel(int nso, int N, const double* __restrict__ h1, const double* __restrict__ h2,
                                double enuc, int conj, double* __restrict__ fov, double* __restrict__ escf) {
  __shared__ double2 red[kThreads];
  const int V = nso - N;
  const double sg = conj ? -1.0 : 1.0;
  const long n2 = (long)nso * nso, n3 = n2 * nso;
  for (int idx = threadIdx.x; idx < N * V; idx += kThreads) {
    const int i = idx / V, a = idx % V;
    double2 acc = ld2(h1, (long)i * nso + N + a);
    for (int k = 0; k < N; ++k) acc = cadd(acc, ld2(h2, i * n3 + k * n2 + (long)(N + a) * nso + k));
    reinterpret_cast<double2*>(fov)[idx] = make_double2(acc.x, sg * acc.y);
  }
  double2 part = make_double2(0.0, 0.0);
  for (int idx = threadIdx.x; idx < N * N; idx += kThreads) {
    const int i = idx / N, j = idx % N;
    part = cadd(part, cscale(ld2(h2, i * n3 + j * n2 + (long)i * nso + j), 0.5));
    if (i == j) part = cadd(part, ld2(h1, (long)i * nso + i));
  }
  const double2 tot = block_sum(part, red);
  if (threadIdx.x == 0) reinterpret_cast<double2*>(escf)[0] = make_double2(tot.x + enuc, sg * tot.y);
}

// ProjT1 set-up per point: MO2[:, :N] = R [:, :N] + R[:, N:] T1, written as
// L (N x N, column-major, packed for the LU kernel) and Y (V x N, column-major).
// R_g is column-major NSO x NSO; t1 row-major [a][i] (conjugated when conj).
__global__ void pcc_thouless_kernel(int nso, int N, const double* __restrict__ R, const double* __restrict__ t1,
                                    int conj, double* __restrict__ L, double* __restrict__ Y) {
  const int g = blockIdx.x, V = nso - N;
  const double* Rg = R + 2L * nso * nso * g;
  const double sg = conj ? -1.0 : 1.0;
  for (int idx = threadIdx.x; idx < nso * N; idx += blockDim.x) {
    const int p = idx % nso, i = idx / nso;
    double2 acc = ld2(Rg, (long)i * nso + p);
    for (int a = 0; a < V; ++a) {
      const double2 t = ld2(t1, (long)a * N + i);
      acc = cadd(acc, cmul(ld2(Rg, (long)(N + a) * nso + p), make_double2(t.x, sg * t.y)));
    }
    if (p < N)
      reinterpret_cast<double2*>(L)[(long)g * N * N + (long)i * N + p] = acc;
    else
      reinterpret_cast<double2*>(Y)[(long)g * V * N + (long)i * V + (p - N)] = acc;
  }
}

// U1 = Y L^-1 (V x N) and V = e^{-U1} R: rows o of V are those of R, rows v
// are R[v, :] - U1 R[o, :] (all column-major).
__global__ void pcc_vmat_kernel(int nso, int N, const double* __restrict__ R, const double* __restrict__ Y,
                                const double* __restrict__ Linv, double* __restrict__ U1, double* __restrict__ Vm) {
  __shared__ double2 u1s[32 * 32];
  const int g = blockIdx.x, V = nso - N;
  const double* Rg = R + 2L * nso * nso * g;
  const double* Yg = Y + 2L * V * N * g;
  const double* Lg = Linv + 2L * N * N * g;
  for (int idx = threadIdx.x; idx < V * N; idx += blockDim.x) {
    const int a = idx % V, i = idx / V;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) acc = cadd(acc, cmul(ld2(Yg, (long)k * V + a), ld2(Lg, (long)i * N + k)));
    u1s[idx] = acc;
    reinterpret_cast<double2*>(U1)[(long)g * V * N + idx] = acc;
  }
  __syncthreads();
  double2* Vg = reinterpret_cast<double2*>(Vm) + (long)g * nso * nso;
  for (int idx = threadIdx.x; idx < nso * nso; idx += blockDim.x) {
    const int p = idx % nso, q = idx / nso;
    double2 v = ld2(Rg, (long)q * nso + p);
    if (p >= N)
      for (int i = 0; i < N; ++i) v = csub(v, cmul(u1s[(long)i * V + (p - N)], ld2(Rg, (long)q * nso + i)));
    Vg[idx] = v;
  }
}

// TransT2's second order (PCCCI.f90:98-245) runs as batched GEMMs over the
// blocks of the transformed tensor U (below, host side); these kernels do the
// rest.  Layouts per point: U row-major [p][q][r][s]; F = sum_i U(p,i,q,i)
// row-major nso x nso; sc = (E0, C0); C1 [a][i]; C2 [a][b][i][j].

// F, E0 = sum_ij U(i,j,i,j)/2 and C0 = 1 + E0 + E0^2/2 + sum fov fvo / 2 +
// sum_ijab U(i,j,a,b) U(a,b,i,j) / 8.  One workgroup per point.
__global__ void __launch_bounds__(kThreads) pcc_tt2_prep_kernel(int nso, int N, const double* __restrict__ Ubuf,
                                                                long ustride, double* __restrict__ Fbuf,
                                                                double* __restrict__ sc) {
  __shared__ double2 red[kThreads];
  const int g = blockIdx.x, V = nso - N, t = threadIdx.x;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso;
  const double* U = Ubuf + ustride * g;
  auto Ue = [&](int p, int q, int r, int s) { return ld2(U, p * n3 + q * n2 + (long)r * n1 + s); };
  double2* F = reinterpret_cast<double2*>(Fbuf) + n2 * g;
  for (int idx = t; idx < nso * nso; idx += kThreads) {
    const int p = idx / nso, q = idx % nso;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < N; ++i) acc = cadd(acc, Ue(p, i, q, i));
    F[idx] = acc;
  }
  double2 part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * N; idx += kThreads) part = cadd(part, Ue(idx / N, idx % N, idx / N, idx % N));
  const double2 e0 = cscale(block_sum(part, red), 0.5);   // also orders the F writes before their reads
  part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * V; idx += kThreads) {
    const int i = idx / V, a = idx % V;
    part = cadd(part, cscale(cmul(F[i * nso + N + a], F[(N + a) * nso + i]), 0.5));
  }
  for (long idx = t; idx < (long)N * N * V * V; idx += kThreads) {   // (i, j, a, b), b fastest
    const int b = idx % V, a = (idx / V) % V, j = (idx / V / V) % N, i = idx / V / V / N;
    part = cadd(part, cscale(cmul(Ue(i, j, N + a, N + b), Ue(N + a, N + b, i, j)), 0.125));
  }
  const double2 q = block_sum(part, red);
  if (t == 0) {
    double2* o = reinterpret_cast<double2*>(sc) + 2L * g;
    o[0] = e0;
    o[1] = cadd(cadd(make_double2(1.0 + e0.x, e0.y), cscale(cmul(e0, e0), 0.5)), q);
  }
}

// Elementwise parts: C2 = vvoo (1 + E0) / 4 - fvo(a,j) fvo(b,i) / 2 for the
// threads of [0, V^2 N^2); for [0, V N) also C1 = fvo (1 + E0) + (fvv fvo -
// fvo foo) / 2 and the compact copies fovT [b][j] = fov(j,b), fvoC [b][j] =
// fvo(b,j) for the C1 products.  grid (ceil(V^2 N^2 / 256), points).
__global__ void __launch_bounds__(256) pcc_tt2_small_kernel(int nso, int N, const double* __restrict__ Ubuf,
                                                            long ustride, const double* __restrict__ Fbuf,
                                                            const double* __restrict__ sc, double* __restrict__ C1,
                                                            double* __restrict__ C2, long c2stride,
                                                            double* __restrict__ fovT, double* __restrict__ fvoC) {
  const int V = nso - N;
  const long S = (long)V * V * N * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long g = blockIdx.y;
  if (idx >= S) return;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso;
  const double* U = Ubuf + ustride * g;
  auto Ue = [&](int p, int q, int r, int s) { return ld2(U, p * n3 + q * n2 + (long)r * n1 + s); };
  const double2* F = reinterpret_cast<const double2*>(Fbuf) + n2 * g;
  auto Fe = [&](int p, int q) { return F[p * nso + q]; };
  const double2 e0 = ld2(sc, 2 * g), one_e0 = make_double2(1.0 + e0.x, e0.y);
  {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / ((long)N * N * V);
    double2 v = cscale(cmul(Ue(N + a, N + b, i, j), one_e0), 0.25);
    v = csub(v, cscale(cmul(Fe(N + a, j), Fe(N + b, i)), 0.5));
    reinterpret_cast<double2*>(C2)[g * c2stride + idx] = v;
  }
  if (idx < (long)V * N) {
    const int a = idx / N, i = idx % N;
    double2 acc = cmul(Fe(N + a, i), one_e0);
    double2 h = make_double2(0.0, 0.0);
    for (int c = 0; c < V; ++c) h = cadd(h, cmul(Fe(N + a, N + c), Fe(N + c, i)));
    for (int k = 0; k < N; ++k) h = csub(h, cmul(Fe(N + a, k), Fe(k, i)));
    reinterpret_cast<double2*>(C1)[g * V * N + idx] = cadd(acc, cscale(h, 0.5));
    // (a, i) read as (b, j)
    reinterpret_cast<double2*>(fovT)[g * V * N + idx] = Fe(i, N + a);
    reinterpret_cast<double2*>(fvoC)[g * V * N + idx] = Fe(N + a, i);
  }
}

// CI2CC + ASymm2 (PCCCI.f90:70-96, CCRes.f90:31-69), DSTT2 (W0 U0, W1 + U1)
// and CCEnergy (CCRes.f90:9-29) for one point: T = C2/C0 - [a != b, i != j]
// t1(a,i) t1(b,j) / 2 with t1 = C1/C0; W1 = t1 + U1; W2 = ASymm2(T);
// lambda0 = C0 U0, eps0 = (ECorr(W1, W2) + ESCF + ENuc) lambda0.  Optional
// outputs for the singles and doubles: W1, W2 and EOvl0 per point.
__global__ void __launch_bounds__(kThreads) pcc_tt2_finish_kernel(
    int nso, int N, const double* __restrict__ C1all, const double* __restrict__ C2all, long c2stride,
    const double* __restrict__ sc, const double* __restrict__ U1all, const double* __restrict__ U0all,
    const double* __restrict__ fovh, const double* __restrict__ escf, const double* __restrict__ h2, int conj,
    double* __restrict__ Tbuf, double* __restrict__ lam_out, double* __restrict__ eps_out,
    double* __restrict__ w1_out, double* __restrict__ w2_out, double* __restrict__ e0_out) {
  __shared__ double2 red[kThreads];
  __shared__ double2 t1s[32 * 32];
  __shared__ double2 w1s[32 * 32];
  const int g = blockIdx.x, V = nso - N;
  const int t = threadIdx.x;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso;
  const double2 c0 = ld2(sc, 2L * g + 1);
  const double2* C1 = reinterpret_cast<const double2*>(C1all) + (long)g * V * N;
  const double2* C2 = reinterpret_cast<const double2*>(C2all) + c2stride * g;
  double2* T = reinterpret_cast<double2*>(Tbuf) + (long)V * V * N * N * g;
  const double2* U1 = reinterpret_cast<const double2*>(U1all) + (long)g * V * N;
  for (int idx = t; idx < V * N; idx += kThreads) {
    const int a = idx / N, i = idx % N;
    t1s[idx] = cdiv(C1[idx], c0);
    w1s[idx] = cadd(t1s[idx], U1[(long)i * V + a]);   // U1 is column-major V x N
    if (w1_out) reinterpret_cast<double2*>(w1_out)[(long)g * V * N + idx] = w1s[idx];
  }
  __syncthreads();
  for (long idx = t; idx < (long)V * V * N * N; idx += kThreads) {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / N / N / V;
    double2 tv = cdiv(C2[idx], c0);
    if (a != b && i != j) tv = csub(tv, cscale(cmul(t1s[a * N + i], t1s[b * N + j]), 0.5));
    T[idx] = tv;
  }
  __syncthreads();
  // CCEnergy(W1, W2) with W2 = ASymm2(T): fov . W1 + 1/4 sum ERI(i,j,a,b) tau(a,b,i,j)
  const double sg = conj ? -1.0 : 1.0;
  const double2* fov = reinterpret_cast<const double2*>(fovh);
  double2 part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * V; idx += kThreads) part = cadd(part, cmul(fov[idx], w1s[(idx % V) * N + idx / V]));
  for (long idx = t; idx < (long)V * V * N * N; idx += kThreads) {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / N / N / V;
    auto Tt = [&](int aa, int bb, int ii, int jj) { return T[(((long)aa * V + bb) * N + ii) * N + jj]; };
    double2 w2 = csub(csub(Tt(a, b, i, j), Tt(a, b, j, i)), csub(Tt(b, a, i, j), Tt(b, a, j, i)));
    if (w2_out) reinterpret_cast<double2*>(w2_out)[(long)V * V * N * N * g + idx] = w2;   // U2t = ASymm2(T)
    w2 = cadd(w2, csub(cmul(w1s[a * N + i], w1s[b * N + j]), cmul(w1s[b * N + i], w1s[a * N + j])));
    double2 eri = ld2(h2, i * n3 + j * n2 + (long)(N + a) * n1 + N + b);
    eri.y *= sg;
    part = cadd(part, cscale(cmul(eri, w2), 0.25));
  }
  const double2 ecorr = block_sum(part, red);
  if (t == 0) {
    const double2 lam = cmul(c0, reinterpret_cast<const double2*>(U0all)[g]);
    const double2 e = cadd(ecorr, reinterpret_cast<const double2*>(escf)[0]);
    reinterpret_cast<double2*>(lam_out)[g] = lam;
    reinterpret_cast<double2*>(eps_out)[g] = cmul(e, lam);
    if (e0_out) reinterpret_cast<double2*>(e0_out)[g] = e;   // EOvl0
  }
}

// ---------------------------------------------------------------------------
// singles and doubles: the CCSD residual (CCRes12 / CCItms) of every point's
// (U1t, U2t) and the BuildKernSD assembly
// ---------------------------------------------------------------------------

// One term of a strided gather: src element for output index (x0, x1, x2, x3)
// of point g at p + 2 (g gs + sum_k x_k st[k]) (complex elements), times
// (ar + i ai), conjugated first when conj.
struct PTerm {
  const double* p;
  long gs;
  long st[4];
  double ar, ai;
  int conj;
};

// out[g][x0][x1][x2][x3] (row-major, dims d, point stride ogs complex
// elements) = (accumulate ? out : 0) + sum of up to four gathered terms.
// Serves block extraction from <pq||rs>, broadcasts of shared blocks into
// per-point buffers, and the permuted antisymmetrisations (APerm) of the
// residual.  grid (ceil(d0 d1 d2 d3 / 256), points).
__global__ void __launch_bounds__(256) pcc_gather_kernel(int d0, int d1, int d2, int d3, int nterm, PTerm t0,
                                                         PTerm t1, PTerm t2, PTerm t3, int accumulate,
                                                         double* __restrict__ out, long ogs) {
  const long total = (long)d0 * d1 * d2 * d3;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long g = blockIdx.y;
  const int x3 = idx % d3, x2 = (idx / d3) % d2, x1 = (idx / d3 / d2) % d1, x0 = idx / ((long)d3 * d2 * d1);
  double2* o = reinterpret_cast<double2*>(out) + g * ogs + idx;
  double2 acc = accumulate ? *o : make_double2(0.0, 0.0);
  const PTerm* ts[4] = {&t0, &t1, &t2, &t3};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= nterm) break;
    const PTerm& t = *ts[q];
    double2 v = ld2(t.p, g * t.gs + x0 * t.st[0] + x1 * t.st[1] + x2 * t.st[2] + x3 * t.st[3]);
    if (t.conj) v.y = -v.y;
    acc = cadd(acc, cmul(make_double2(t.ar, t.ai), v));
  }
  *o = acc;
}

// Full Fock matrix of (HOne, HTwo): F(p,q) = HOne(p,q) + sum_k HTwo(p,k,q,k)
// (MKFock_MO, UHF.f90:461-497), conjugated when conj; row-major.
__global__ void pcc_fock_full_kernel(int nso, int N, const double* __restrict__ h1, const double* __restrict__ h2,
                                     int conj, double* __restrict__ f) {
  const long n2 = (long)nso * nso, n3 = n2 * nso;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n2) return;
  const int p = idx / nso, q = idx % nso;
  double2 acc = ld2(h1, idx);
  for (int k = 0; k < N; ++k) acc = cadd(acc, ld2(h2, p * n3 + k * n2 + (long)q * nso + k));
  if (conj) acc.y = -acc.y;
  reinterpret_cast<double2*>(f)[idx] = acc;
}

// Per-point amplitude tensors of the residual from t1 = U1t [a][i] and
// t2 = U2t [a][b][i][j] (S = V^2 N^2 complex elements per point each):
//   tau  [a][b][i][j] = t2 + t_ai t_bj - t_bi t_aj      tb [a][b][i][j] = t2 + (..)/2
//   tbP  [n][e][b][j] = tb(e,b,j,n)                       t2P [m][e][b][i] = t2(e,b,i,m)
//   i4P  [n][f][b][j] = t2(f,b,j,n)/2 + t_fj t_bn         Y2 [a][i][m][e] = t2(a,e,i,m)
//   Y1   [a][i][m][e] = t_am t_ei
__global__ void __launch_bounds__(256) pcc_amp_kernel(int V, int N, const double* __restrict__ t1all,
                                                      const double* __restrict__ t2all, double* __restrict__ tau,
                                                      double* __restrict__ tb, double* __restrict__ tbP,
                                                      double* __restrict__ t2P, double* __restrict__ i4P,
                                                      double* __restrict__ Y2, double* __restrict__ Y1) {
  const long S = (long)V * V * N * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S) return;
  const long g = blockIdx.y;
  const double* t1 = t1all + 2 * g * V * N;
  const double* t2 = t2all + 2 * g * S;
  auto T1 = [&](int a, int i) { return ld2(t1, (long)a * N + i); };
  auto T2 = [&](int a, int b, int i, int j) { return ld2(t2, (((long)a * V + b) * N + i) * N + j); };
  auto anti = [&](int a, int b, int i, int j) { return csub(cmul(T1(a, i), T1(b, j)), cmul(T1(b, i), T1(a, j))); };
  const long o = g * S + idx;
  {   // natural (a, b, i, j)
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / ((long)N * N * V);
    const double2 x = T2(a, b, i, j), y = anti(a, b, i, j);
    reinterpret_cast<double2*>(tau)[o] = cadd(x, y);
    reinterpret_cast<double2*>(tb)[o] = cadd(x, cscale(y, 0.5));
  }
  {   // (n, e, b, j), j fastest
    const int j = idx % N, b = (idx / N) % V, e = (idx / N / V) % V, n = idx / ((long)N * V * V);
    reinterpret_cast<double2*>(tbP)[o] = cadd(T2(e, b, j, n), cscale(anti(e, b, j, n), 0.5));
    reinterpret_cast<double2*>(t2P)[o] = T2(e, b, j, n);
    reinterpret_cast<double2*>(i4P)[o] = cadd(cscale(T2(e, b, j, n), 0.5), cmul(T1(e, j), T1(b, n)));
  }
  {   // (a, i, m, e), e fastest
    const int e = idx % V, m = (idx / V) % N, i = (idx / V / N) % N, a = idx / ((long)V * N * N);
    reinterpret_cast<double2*>(Y2)[o] = T2(a, e, i, m);
    reinterpret_cast<double2*>(Y1)[o] = cmul(T1(a, m), T1(e, i));
  }
}

// BuildKernSD's assembly (FPUCC_Tools.f90:322-429) and EvalKernel's scaling
// by lambda0 (FPUCC.f90:589-594) for every point of the chunk, summed with
// the point weights into the grid sums (FPUCC.f90:181-186), in point order:
//   lambda1 = l0 U1,  lambda2 = l0 (U2 + U1^U1),
//   eps1 = l0 (R1 + U1 E0),  eps2 = l0 (R2 + (U2 + U1^U1) E0 + P(ij)P(ab) U1 R1)
// with ProjSz (S_z-conserving elements only) when sp = 2.  Threads [0, S)
// take the doubles (a, b, i, j), [S, S + V N) the singles (a, i).
// out: o1 [V N], e1 [V N], o2 [S], e2 [S] (accumulated).
__global__ void __launch_bounds__(256) pcc_sd_sum_kernel(int V, int N, long cg, int sp, int noa, int nva,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ lam,
                                                         const double* __restrict__ e0,
                                                         const double* __restrict__ U1all,
                                                         const double* __restrict__ U2all,
                                                         const double* __restrict__ R1all,
                                                         const double* __restrict__ R2all, double* __restrict__ out) {
  const long S = (long)V * V * N * N, VN = (long)V * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S + VN) return;
  auto sa = [&](int a) { return a < nva ? 1 : -1; };   // virtual alpha first
  auto si = [&](int i) { return i < noa ? 1 : -1; };   // occupied alpha first
  double2* o1 = reinterpret_cast<double2*>(out);
  double2* e1 = o1 + VN;
  double2* o2 = e1 + VN;
  double2* e2 = o2 + S;
  double2 ao = make_double2(0.0, 0.0), ae = make_double2(0.0, 0.0);
  if (idx < S) {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / ((long)N * N * V);
    if (sp == 2 && sa(a) + sa(b) != si(i) + si(j)) return;
    for (long g = 0; g < cg; ++g) {
      const double* u1 = U1all + 2 * g * VN;
      const double* r1 = R1all + 2 * g * VN;
      auto U = [&](int x, int y) { return ld2(u1, (long)x * N + y); };
      auto R = [&](int x, int y) { return ld2(r1, (long)x * N + y); };
      const double2 l0 = ld2(lam, g), en = ld2(e0, g), wg = ld2(w, g);
      const double2 ovl = cadd(ld2(U2all, g * S + idx), csub(cmul(U(a, i), U(b, j)), cmul(U(b, i), U(a, j))));
      double2 x = cadd(ld2(R2all, g * S + idx), cmul(ovl, en));
      x = cadd(x, csub(cmul(U(a, i), R(b, j)), cmul(U(b, i), R(a, j))));
      x = cadd(x, csub(cmul(U(b, j), R(a, i)), cmul(U(a, j), R(b, i))));
      const double2 wl = cmul(wg, l0);
      ao = cadd(ao, cmul(wl, ovl));
      ae = cadd(ae, cmul(wl, x));
    }
    o2[idx] = cadd(o2[idx], ao);
    e2[idx] = cadd(e2[idx], ae);
  } else {
    const long k = idx - S;
    const int a = k / N, i = k % N;
    if (sp == 2 && sa(a) != si(i)) return;
    for (long g = 0; g < cg; ++g) {
      const double2 l0 = ld2(lam, g), en = ld2(e0, g), wg = ld2(w, g);
      const double2 u = ld2(U1all, g * VN + k);
      const double2 x = cadd(ld2(R1all, g * VN + k), cmul(u, en));
      const double2 wl = cmul(wg, l0);
      ao = cadd(ao, cmul(wl, u));
      ae = cadd(ae, cmul(wl, x));
    }
    o1[k] = cadd(o1[k], ao);
    e1[k] = cadd(e1[k], ae);
  }
}

// Olap0 = sum_g w_g lambda0_g, EOlap0 = sum_g w_g eps0_g in a fixed order (one workgroup).
__global__ void pcc_gridsum_kernel(long G, const double* __restrict__ w, const double* __restrict__ lam,
                                   const double* __restrict__ eps, double* __restrict__ out) {
  __shared__ double2 red[kThreads];
  double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
  for (long g = threadIdx.x; g < G; g += kThreads) {
    a = cadd(a, cmul(ld2(w, g), ld2(lam, g)));
    b = cadd(b, cmul(ld2(w, g), ld2(eps, g)));
  }
  const double2 sa = block_sum(a, red);
  const double2 sb = block_sum(b, red);
  if (threadIdx.x == 0) {
    reinterpret_cast<double2*>(out)[0] = sa;
    reinterpret_cast<double2*>(out)[1] = sb;
  }
}


// ---------------------------------------------------------------------------
// host side of the singles and doubles
// ---------------------------------------------------------------------------

inline ZMat RM(const double* p, long ld, long bs) { return ZMat{p, 2 * ld, 2, 2 * bs, 1, 0}; }
inline ZOut RO(double* p, long ld, long bs) { return ZOut{p, 2 * ld, 2, 2 * bs, 1}; }
inline PTerm pterm(const double* p, long gs, long s0, long s1, long s2, long s3, double ar = 1.0, double ai = 0.0,
                   int conj = 0) {
  return PTerm{p, gs, {s0, s1, s2, s3}, ar, ai, conj};
}

void gather(hipStream_t s, int d0, int d1, int d2, int d3, std::initializer_list<PTerm> terms, bool acc, double* out,
            long ogs, long npts) {
  PTerm t[4] = {};
  int n = 0;
  for (const PTerm& x : terms) t[n++] = x;
  const long total = (long)d0 * d1 * d2 * d3;
  dim3 grid((unsigned)((total + 255) / 256), (unsigned)npts);
  pcc_gather_kernel<<<grid, 256, 0, s>>>(d0, d1, d2, d3, n, t[0], t[1], t[2], t[3], acc ? 1 : 0, out, ogs);
}

// The Fock blocks and the blocks of <pq||rs> (conjugated when conj) the
// residual reads, compact and row-major; P1..P11 are permuted copies laid out
// so that every contraction is a plain row-major GEMM.
struct PccShared {
  double *f, *foo, *fov, *fvo, *fvv;
  double *Hoooo, *Hooov, *Hoovv, *Hvvoo, *Hvovv, *Hvvvv, *Hovoo;
  double* P[12];
};

// Per-point buffers of the singles and doubles (chunk points each).
struct PccSD {
  double *W1, *W2, *E0, *tau, *tb, *tbP, *t2P, *i4P, *Y1, *Y2, *WP, *R2, *X, *Z, *Weff, *Wp;
  double *Fvv, *Foo, *Fov, *itm1, *itm3, *r1, *sum;
};

PccShared carve_shared(Arena& a, int nso, int N) {
  const long V = nso - N, O = N;
  PccShared h{};
  h.f = a.take<double>(2L * nso * nso);
  h.foo = a.take<double>(2 * O * O);
  h.fov = a.take<double>(2 * O * V);
  h.fvo = a.take<double>(2 * V * O);
  h.fvv = a.take<double>(2 * V * V);
  h.Hoooo = a.take<double>(2 * O * O * O * O);
  h.Hooov = a.take<double>(2 * O * O * O * V);
  h.Hoovv = a.take<double>(2 * O * O * V * V);
  h.Hvvoo = a.take<double>(2 * V * V * O * O);
  h.Hvovv = a.take<double>(2 * V * O * V * V);
  h.Hvvvv = a.take<double>(2 * V * V * V * V);
  h.Hovoo = a.take<double>(2 * O * V * O * O);
  const long sz[12] = {0, V * V * V * O, V * O * O * V, O * O * V * O, O * V * V * O, O * V * V * O, O * V * V * V,
                       O * O * V * O, O * V * O * V, V * O * V * O, V * O * O * O, V * V * O * V};
  for (int k = 1; k < 12; ++k) h.P[k] = a.take<double>(2 * sz[k]);
  return h;
}

PccSD carve_sd(Arena& a, long cg, int V, int N) {
  const long S = (long)V * V * N * N, VN = (long)V * N, N4 = (long)N * N * N * N;
  PccSD b{};
  b.W1 = a.take<double>(2 * VN * cg);
  b.W2 = a.take<double>(2 * S * cg);
  b.E0 = a.take<double>(2 * cg);
  b.tau = a.take<double>(2 * S * cg);
  b.tb = a.take<double>(2 * S * cg);
  b.tbP = a.take<double>(2 * S * cg);
  b.t2P = a.take<double>(2 * S * cg);
  b.i4P = a.take<double>(2 * S * cg);
  b.Y1 = a.take<double>(2 * S * cg);
  b.Y2 = a.take<double>(2 * S * cg);
  b.WP = a.take<double>(2 * S * cg);
  b.R2 = a.take<double>(2 * S * cg);
  b.X = a.take<double>(2 * S * cg);
  b.Z = a.take<double>(2 * VN * N * N * cg);
  b.Weff = a.take<double>(2 * N4 * cg);
  b.Wp = a.take<double>(2 * N4 * cg);
  b.Fvv = a.take<double>(2L * V * V * cg);
  b.Foo = a.take<double>(2L * N * N * cg);
  b.Fov = a.take<double>(2 * VN * cg);
  b.itm1 = a.take<double>(2L * V * V * cg);
  b.itm3 = a.take<double>(2L * N * N * cg);
  b.r1 = a.take<double>(2 * VN * cg);
  b.sum = a.take<double>(4 * (VN + S));
  return b;
}

// Fock and <pq||rs> blocks of one pass (conj: the second pass of ncik = 2).
void build_shared(hipStream_t s, int nso, int N, const double* h1, const double* h2, int conj, const PccShared& h) {
  const int V = nso - N, O = N;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso;
  pcc_fock_full_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, s>>>(nso, N, h1, h2, conj, h.f);
  gather(s, O, O, 1, 1, {pterm(h.f, 0, n1, 1, 0, 0)}, false, h.foo, 0, 1);
  gather(s, O, V, 1, 1, {pterm(h.f + 2 * N, 0, n1, 1, 0, 0)}, false, h.fov, 0, 1);
  gather(s, V, O, 1, 1, {pterm(h.f + 2 * N * n1, 0, n1, 1, 0, 0)}, false, h.fvo, 0, 1);
  gather(s, V, V, 1, 1, {pterm(h.f + 2 * (N * n1 + N), 0, n1, 1, 0, 0)}, false, h.fvv, 0, 1);
  // block of h2: out dim k is occupied ('o') or virtual ('v') and reads h2 position pos[k]
  const long hst[4] = {n3, n2, n1, 1};
  auto blk = [&](const char* ty, int p0, int p1, int p2, int p3, double* out) {
    const int pos[4] = {p0, p1, p2, p3};
    int d[4];
    long base = 0, st[4];
    for (int k = 0; k < 4; ++k) {
      d[k] = ty[k] == 'v' ? V : O;
      st[k] = hst[pos[k]];
      if (ty[k] == 'v') base += (long)N * hst[pos[k]];
    }
    gather(s, d[0], d[1], d[2], d[3], {pterm(h2 + 2 * base, 0, st[0], st[1], st[2], st[3], 1.0, 0.0, conj)}, false,
           out, 0, 1);
  };
  blk("oooo", 0, 1, 2, 3, h.Hoooo);
  blk("ooov", 0, 1, 2, 3, h.Hooov);
  blk("oovv", 0, 1, 2, 3, h.Hoovv);
  blk("vvoo", 0, 1, 2, 3, h.Hvvoo);
  blk("vovv", 0, 1, 2, 3, h.Hvovv);
  blk("vvvv", 0, 1, 2, 3, h.Hvvvv);
  blk("ovoo", 0, 1, 2, 3, h.Hovoo);
  blk("vvvo", 0, 2, 3, 1, h.P[1]);    // [a][b][f][m] = H(a,m,b,f)
  blk("voov", 3, 0, 1, 2, h.P[2]);    // [b][m][n][c] = H(m,n,c,b)
  blk("oovo", 0, 2, 3, 1, h.P[3]);    // [i][j][e][m] = H(i,m,j,e)
  blk("ovvo", 0, 2, 3, 1, h.P[4]);    // [i][a][e][m] = H(i,m,a,e)
  blk("ovvo", 0, 2, 1, 3, h.P[5]);    // [m][e][b][j] = H(m,b,e,j)
  blk("ovvv", 0, 2, 1, 3, h.P[6]);    // [m][e][b][f] = H(m,b,e,f)
  blk("oovo", 1, 0, 2, 3, h.P[7]);    // [n][m][e][j] = H(m,n,e,j)
  blk("ovov", 0, 2, 1, 3, h.P[8]);    // [m][e][n][f] = H(m,n,e,f)
  blk("vovo", 0, 2, 3, 1, h.P[9]);    // [a][i][e][m] = H(a,m,i,e)
  blk("vooo", 3, 0, 1, 2, h.P[10]);   // [b][m][n][i] = H(m,n,i,b)
  blk("vvov", 0, 1, 3, 2, h.P[11]);   // [a][b][j][e] = H(a,b,e,j)
}

// CCRes12 (CCRes.f90:164-353, intermediates of CCItms :355-536) of every
// point's (U1t, U2t) = (W1, W2) as batched complex GEMMs: a contraction with
// a shared block of <pq||rs> is one GEMM per point (or, where the per-point
// operand is U1t, one GEMM over all points), the rest are per-point products
// and strided gathers.  The particle-particle ladder is kept as
// 1/2 <ab||ef> tau_efij + 1/8 tau_abmn (<mn||ef> tau_efij)
//   - 1/2 P(ab) t_bm (<am||ef> tau_efij)
// (CCItms' Wvvvv contracted with tau, without forming Wvvvv per point).
void pcc_residual(hipStream_t s, int V, int N, long cg, const PccShared& h, const PccSD& b, hipEvent_t* lad,
                  double* kws, long kws_cplx) {
  auto zgb = [&](hipStream_t st, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br, double bi,
                 ZOut C, int batch) {
    zgemm_batched_splitk(st, m, n, k, ar, ai, A, B, br, bi, C, batch, kws, kws_cplx);
  };
  const long O = N, S = (long)V * V * N * N, VN = (long)V * N, N2 = O * O, V2 = (long)V * V, N3 = N2 * O;
  const long N4 = N2 * N2;
  const ZMat W1all{b.W1, 2, 2 * VN, 0, 1, 0};   // [(e, m)][g]: every point's U1t as one k x cg operand
  pcc_amp_kernel<<<dim3((unsigned)((S + 255) / 256), (unsigned)cg), 256, 0, s>>>(V, N, b.W1, b.W2, b.tau, b.tb, b.tbP,
                                                                                  b.t2P, b.i4P, b.Y2, b.Y1);
  // initial values: broadcasts of shared blocks
  gather(s, V, V, 1, 1, {pterm(h.fvv, 0, V, 1, 0, 0)}, false, b.Fvv, V2, cg);
  gather(s, O, O, 1, 1, {pterm(h.foo, 0, O, 1, 0, 0)}, false, b.Foo, N2, cg);
  gather(s, O, V, 1, 1, {pterm(h.fov, 0, V, 1, 0, 0)}, false, b.Fov, VN, cg);
  gather(s, V, O, 1, 1, {pterm(h.fvo, 0, O, 1, 0, 0)}, false, b.r1, VN, cg);
  gather(s, O, V, V, O, {pterm(h.P[5], 0, V * V * O, V * O, O, 1)}, false, b.WP, S, cg);
  gather(s, V, V, O, O, {pterm(h.Hvvoo, 0, V * N2, N2, O, 1)}, false, b.R2, S, cg);
  gather(s, O, O, O, O, {pterm(h.Hoooo, 0, N3, N2, O, 1)}, false, b.Weff, N4, cg);
  // Fvv (CCRes.f90:405-424)
  zgb(s, V, V, N, -0.5, 0, RM(b.W1, N, VN), RM(h.fov, V, 0), 1, 0, RO(b.Fvv, V, V2), (int)cg);
  zgb(s, (int)V2, (int)cg, (int)VN, 1, 0, RM(h.P[1], VN, 0), W1all, 1, 0, ZOut{b.Fvv, 2, 2 * V2, 0, 1}, 1);
  zgb(s, V, V, (int)(VN * O), -0.5, 0, RM(b.tb, VN * O, S), RM(h.P[2], V, 0), 1, 0, RO(b.Fvv, V, V2),
                (int)cg);
  // Foo (:426-441)
  zgb(s, N, N, V, 0.5, 0, RM(h.fov, V, 0), RM(b.W1, N, VN), 1, 0, RO(b.Foo, N, N2), (int)cg);
  zgb(s, (int)N2, (int)cg, (int)VN, 1, 0, RM(h.P[3], VN, 0), W1all, 1, 0, ZOut{b.Foo, 2, 2 * N2, 0, 1}, 1);
  zgb(s, N, N, (int)(O * V2), 0.5, 0, RM(h.Hoovv, O * V2, 0), RM(b.tbP, N, S), 1, 0, RO(b.Foo, N, N2),
                (int)cg);
  // Fov (:443-452)
  zgb(s, (int)VN, (int)cg, (int)VN, 1, 0, RM(h.P[4], VN, 0), W1all, 1, 0, ZOut{b.Fov, 2, 2 * VN, 0, 1}, 1);
  // Wovvo as WP [m][e][b][j] (:487-519)
  zgb(s, (int)(O * V2), N, V, 1, 0, RM(h.P[6], V, 0), RM(b.W1, N, VN), 1, 0, RO(b.WP, N, S), (int)cg);
  zgb(s, V, (int)(VN * O), N, 1, 0, RM(b.W1, N, VN), RM(h.P[7], VN * O, 0), 0, 0, RO(b.X, VN * O, S),
                (int)cg);   // X [b][m][e][j]
  gather(s, O, V, V, O, {pterm(b.X, S, V * O, O, O * V * O, 1, -1.0)}, true, b.WP, S, cg);
  zgb(s, (int)VN, (int)VN, (int)VN, -1, 0, RM(h.P[8], VN, 0), RM(b.i4P, VN, S), 1, 0, RO(b.WP, VN, S),
                (int)cg);
  // res1 (:215-236)
  zgb(s, V, N, V, 1, 0, RM(b.Fvv, V, V2), RM(b.W1, N, VN), 1, 0, RO(b.r1, N, VN), (int)cg);
  zgb(s, V, N, N, -1, 0, RM(b.W1, N, VN), RM(b.Foo, N, N2), 1, 0, RO(b.r1, N, VN), (int)cg);
  zgb(s, (int)VN, 1, (int)VN, 1, 0, RM(b.Y2, VN, S), ZMat{b.Fov, 2, 0, 2 * VN, 1, 0}, 1, 0,
                ZOut{b.r1, 2, 0, 2 * VN, 1}, (int)cg);
  zgb(s, (int)VN, (int)cg, (int)VN, 1, 0, RM(h.P[9], VN, 0), W1all, 1, 0, ZOut{b.r1, 2, 2 * VN, 0, 1}, 1);
  zgb(s, V, N, (int)(VN * O), -0.5, 0, RM(b.W2, VN * O, S), RM(h.P[10], N, 0), 1, 0, RO(b.r1, N, VN),
                (int)cg);
  zgb(s, V, N, (int)(O * V2), 0.5, 0, RM(h.Hvovv, O * V2, 0), RM(b.t2P, N, S), 1, 0, RO(b.r1, N, VN),
                (int)cg);
  // res2 [a][b][i][j], starting from <ab||ij> (:238-352)
  // (1) P(ab) t2(a,e,i,j) itm1(b,e), itm1 = Fvv - t1 Fov / 2
  gather(s, V, V, 1, 1, {pterm(b.Fvv, V2, V, 1, 0, 0)}, false, b.itm1, V2, cg);
  zgb(s, V, V, N, -0.5, 0, RM(b.W1, N, VN), RM(b.Fov, V, VN), 1, 0, RO(b.itm1, V, V2), (int)cg);
  zgb(s, V, (int)N2, V, 1, 0, ZMat{b.itm1, 2L * V, 2, 2 * V2, 1, 0, 0, V},
                ZMat{b.W2, 2 * N2, 2, 2 * S, 1, 0, 2 * VN * O, V}, 0, 0,
                ZOut{b.X, 2 * N2, 2, 2 * S, 1, nullptr, 2 * VN * O, V}, (int)(cg * V));   // X [a][b][i][j]
  gather(s, V, V, O, O, {pterm(b.X, S, VN * O, N2, O, 1), pterm(b.X, S, N2, VN * O, O, 1, -1.0)}, true, b.R2, S, cg);
  // (2) -P(ij) t2(a,b,i,m) itm3(m,j), itm3 = Foo + Fov t1 / 2
  gather(s, O, O, 1, 1, {pterm(b.Foo, N2, O, 1, 0, 0)}, false, b.itm3, N2, cg);
  zgb(s, N, N, V, 0.5, 0, RM(b.Fov, V, VN), RM(b.W1, N, VN), 1, 0, RO(b.itm3, N, N2), (int)cg);
  zgb(s, (int)(V2 * O), N, N, 1, 0, RM(b.W2, N, S), RM(b.itm3, N, N2), 0, 0, RO(b.X, N, S), (int)cg);
  gather(s, V, V, O, O, {pterm(b.X, S, VN * O, N2, O, 1, -1.0), pterm(b.X, S, VN * O, N2, 1, O)}, true, b.R2, S, cg);
  // (3) P(ij)P(ab) [t2(a,e,i,m) Wovvo(m,b,e,j) - t(a,m) t(e,i) <mb||ej>]: X [a][i][b][j]
  zgb(s, (int)VN, (int)VN, (int)VN, 1, 0, RM(b.Y2, VN, S), RM(b.WP, VN, S), 0, 0, RO(b.X, VN, S), (int)cg);
  zgb(s, (int)VN, (int)VN, (int)VN, -1, 0, RM(b.Y1, VN, S), RM(h.P[5], VN, 0), 1, 0, RO(b.X, VN, S),
                (int)cg);
  gather(s, V, V, O, O,
         {pterm(b.X, S, O * VN, O, VN, 1), pterm(b.X, S, O * VN, O, 1, VN, -1.0), pterm(b.X, S, O, O * VN, VN, 1, -1.0),
          pterm(b.X, S, O, O * VN, 1, VN)},
         true, b.R2, S, cg);
  // (4) ladder and hole-hole terms: 1/2 <ab||ef> tau_efij; Weff = <mn||ij> + 1/2 <mn||ef> tau_efij
  //     + P(ij) <mn||ie> t_ej, then 1/2 tau_abmn Weff_mnij; -1/2 P(ab) t_bm (<am||ef> tau_efij)
  if (lad) (void)hipEventRecord(lad[0], s);
  zgb(s, (int)V2, (int)N2, (int)V2, 0.5, 0, RM(h.Hvvvv, V2, 0), RM(b.tau, N2, S), 1, 0, RO(b.R2, N2, S),
                (int)cg);
  if (lad) (void)hipEventRecord(lad[1], s);
  zgb(s, (int)N2, (int)N2, (int)V2, 0.5, 0, RM(h.Hoovv, V2, 0), RM(b.tau, N2, S), 1, 0,
                RO(b.Weff, N2, N4), (int)cg);
  zgb(s, (int)N3, N, V, 1, 0, RM(h.Hooov, V, 0), RM(b.W1, N, VN), 0, 0, RO(b.Wp, N, N4), (int)cg);
  gather(s, O, O, O, O, {pterm(b.Wp, N4, N3, N2, O, 1), pterm(b.Wp, N4, N3, N2, 1, O, -1.0)}, true, b.Weff, N4, cg);
  zgb(s, (int)V2, (int)N2, (int)N2, 0.5, 0, RM(b.tau, N2, S), RM(b.Weff, N2, N4), 1, 0, RO(b.R2, N2, S),
                (int)cg);
  zgb(s, (int)VN, (int)N2, (int)V2, 1, 0, RM(h.Hvovv, V2, 0), RM(b.tau, N2, S), 0, 0, RO(b.Z, N2, VN * N2),
                (int)cg);   // Z [a][m][i][j]
  zgb(s, V, (int)N2, N, 1, 0, ZMat{b.W1, 2L * N, 2, 2 * VN, 1, 0, 0, V},
                ZMat{b.Z, 2 * N2, 2, 2 * VN * N2, 1, 0, 2 * N3, V}, 0, 0,
                ZOut{b.X, 2 * N2, 2, 2 * S, 1, nullptr, 2 * VN * O, V}, (int)(cg * V));   // Q [a][b][i][j]
  gather(s, V, V, O, O, {pterm(b.X, S, VN * O, N2, O, 1, -0.5), pterm(b.X, S, N2, VN * O, O, 1, 0.5)}, true, b.R2, S,
         cg);
  // (6) P(ij) t(e,i) <ab||ej>: X [a][b][j][i] = sum_e P11[(a,b,j), e] t(e, i)
  zgb(s, (int)(V2 * O), N, V, 1, 0, RM(h.P[11], V, 0), RM(b.W1, N, VN), 0, 0, RO(b.X, N, S), (int)cg);
  gather(s, V, V, O, O, {pterm(b.X, S, VN * O, N2, 1, O), pterm(b.X, S, VN * O, N2, O, 1, -1.0)}, true, b.R2, S, cg);
  // (7) -P(ab) t(a,m) <mb||ij>
  zgb(s, V, (int)(VN * O), N, 1, 0, RM(b.W1, N, VN), RM(h.Hovoo, VN * O, 0), 0, 0, RO(b.X, VN * O, S),
                (int)cg);
  gather(s, V, V, O, O, {pterm(b.X, S, VN * O, N2, O, 1, -1.0), pterm(b.X, S, N2, VN * O, O, 1)}, true, b.R2, S, cg);
}

// TransT2's contractions (PCCCI.f90:98-245) for cg points as batched GEMMs
// over compact blocks of U (per point: row-major nso^4 at U + 2 n4 g); the
// blocks, C2 and a product scratch live in W (n4 complex elements per point,
// offsets below; they fit since (V + N)^4 exceeds them by 2 V N (V^2 + N^2)).
// F and sc come from pcc_tt2_prep_kernel; C1 (V N per point) and C2 start
// from pcc_tt2_small_kernel's elementwise terms.  Returns C2's offset in W.
long pcc_transt2_gemm(hipStream_t s, int nso, int N, long cg, const double* U, double* W, const double* F,
                      const double* sc, double* C1, double* fovT, double* fvoC, double* kws, long kws_cplx) {
  auto zgb = [&](hipStream_t st, int m, int n, int k, double ar, double ai, ZMat A, ZMat B, double br, double bi,
                 ZOut C, int batch) {
    zgemm_batched_splitk(st, m, n, k, ar, ai, A, B, br, bi, C, batch, kws, kws_cplx);
  };
  const long V = nso - N, O = N, S = V * V * O * O, VN = V * O;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso, n4 = n2 * n2;
  const long o_vvoo = 0, o_P1 = S, o_P2 = 2 * S, o_ovP = 3 * S, o_C2 = 4 * S, o_X = 5 * S, o_V4 = 6 * S;
  const long o_O4 = o_V4 + V * V * V * V, o_ooov = o_O4 + O * O * O * O, o_ovvv = o_ooov + V * O * O * O;
  const long o_ovoo = o_ovvv + V * V * V * O, o_vvov = o_ovoo + O * V * O * O;
  auto w = [&](long off) { return W + 2 * off; };
  // block of U: out dim k is occupied ('o') or virtual ('v') and reads U position pos[k]
  const long hst[4] = {n3, n2, n1, 1};
  auto ublk = [&](const char* ty, int p0, int p1, int p2, int p3, long off) {
    const int pos[4] = {p0, p1, p2, p3};
    int d[4];
    long base = 0, st[4];
    for (int k = 0; k < 4; ++k) {
      d[k] = ty[k] == 'v' ? (int)V : (int)O;
      st[k] = hst[pos[k]];
      if (ty[k] == 'v') base += O * hst[pos[k]];
    }
    gather(s, d[0], d[1], d[2], d[3], {pterm(U + 2 * base, n4, st[0], st[1], st[2], st[3])}, false, w(off), n4, cg);
  };
  ublk("vvoo", 0, 1, 2, 3, o_vvoo);   // [a][b][i][j] = U(a,b,i,j)
  ublk("vovo", 0, 2, 1, 3, o_P1);     // [a][j][c][k] = U(a,c,j,k)
  ublk("vvoo", 0, 1, 3, 2, o_P2);     // [c][d][j][i] = U(c,d,i,j)
  ublk("vovo", 3, 0, 1, 2, o_ovP);    // [c][k][b][i] = U(k,b,i,c)
  ublk("vvvv", 0, 1, 2, 3, o_V4);
  ublk("oooo", 0, 1, 2, 3, o_O4);
  ublk("vooo", 3, 0, 1, 2, o_ooov);   // [b][i][j][k] = U(i,j,k,b)
  ublk("vvvo", 1, 2, 3, 0, o_ovvv);   // [a][c][d][j] = U(j,a,c,d)
  ublk("ovoo", 0, 1, 2, 3, o_ovoo);   // [k][b][i][j] = U(k,b,i,j)
  ublk("vvov", 0, 1, 2, 3, o_vvov);   // [a][b][i][c] = U(a,b,i,c)
  pcc_tt2_small_kernel<<<dim3((unsigned)((S + 255) / 256), (unsigned)cg), 256, 0, s>>>(nso, N, U, 2 * n4, F, sc, C1,
                                                                                       w(o_C2), n4, fovT, fvoC);
  const ZMat Foo{F, 2 * n1, 2, 2 * n2, 1, 0};
  const ZMat Fvv{F + 2 * (O * n1 + O), 2 * n1, 2, 2 * n2, 1, 0};
  const ZMat Fvo{F + 2 * O * n1, 2 * n1, 2, 2 * n2, 1, 0};
  // C1: 1/2 sum_bj vvoo(a,b,i,j) fov(j,b) - 1/2 sum_bj U(j,a,i,b) fvo(b,j) as matrix-vector products over
  // (b, j) (P1 read as [a][i][b][j], ovP as [b][j][a][i]), then
  // -1/4 sum_bij vvoo(a,b,i,j) U(i,j,k,b) - 1/4 sum_cdj U(j,a,c,d) vvoo(c,d,i,j)
  zgb(s, (int)VN, 1, (int)VN, 0.5, 0, RM(w(o_P1), VN, n4), ZMat{fovT, 2, 0, 2 * VN, 1, 0}, 1, 0,
                ZOut{C1, 2, 0, 2 * VN, 1}, (int)cg);
  zgb(s, (int)VN, 1, (int)VN, -0.5, 0, ZMat{w(o_ovP), 2, 2 * VN, 2 * n4, 1, 0},
                ZMat{fvoC, 2, 0, 2 * VN, 1, 0}, 1, 0, ZOut{C1, 2, 0, 2 * VN, 1}, (int)cg);
  zgb(s, (int)V, N, (int)(VN * O), -0.25, 0, RM(w(o_vvoo), VN * O, n4), RM(w(o_ooov), O, n4), 1, 0,
                RO(C1, O, VN), (int)cg);
  zgb(s, (int)V, N, (int)(V * VN), -0.25, 0, RM(w(o_ovvv), V * VN, n4), RM(w(o_P2), O, n4), 1, 0,
                RO(C1, O, VN), (int)cg);
  // C2: 1/2 sum_kc vvoo(a,c,j,k) U(k,b,i,c) as X [a][j][b][i]
  zgb(s, (int)VN, (int)VN, (int)VN, 1, 0, RM(w(o_P1), VN, n4), RM(w(o_ovP), VN, n4), 0, 0,
                RO(w(o_X), VN, n4), (int)cg);
  gather(s, (int)V, (int)V, N, N, {pterm(w(o_X), n4, O * VN, O, 1, VN, 0.5)}, true, w(o_C2), n4, cg);
  // 1/16 (U_vvvv vvoo + vvoo U_oooo)
  zgb(s, (int)(V * V), (int)(O * O), (int)(V * V), 1.0 / 16, 0, RM(w(o_V4), V * V, n4),
                RM(w(o_vvoo), O * O, n4), 1, 0, RO(w(o_C2), O * O, n4), (int)cg);
  zgb(s, (int)(V * V), (int)(O * O), (int)(O * O), 1.0 / 16, 0, RM(w(o_vvoo), O * O, n4),
                RM(w(o_O4), O * O, n4), 1, 0, RO(w(o_C2), O * O, n4), (int)cg);
  // -1/4 vvoo(a,b,i,k) foo(k,j) + 1/4 fvv(a,c) vvoo(c,b,i,j) - 1/4 fvo(a,k) U(k,b,i,j) + 1/4 U(a,b,i,c) fvo(c,j)
  zgb(s, (int)(V * VN), N, N, -0.25, 0, RM(w(o_vvoo), O, n4), Foo, 1, 0, RO(w(o_C2), O, n4), (int)cg);
  zgb(s, (int)V, (int)(VN * O), (int)V, 0.25, 0, Fvv, RM(w(o_vvoo), VN * O, n4), 1, 0,
                RO(w(o_C2), VN * O, n4), (int)cg);
  zgb(s, (int)V, (int)(VN * O), N, -0.25, 0, Fvo, RM(w(o_ovoo), VN * O, n4), 1, 0, RO(w(o_C2), VN * O, n4),
                (int)cg);
  zgb(s, (int)(V * VN), N, (int)V, 0.25, 0, RM(w(o_vvov), V, n4), Fvo, 1, 0, RO(w(o_C2), O, n4), (int)cg);
  return o_C2;
}

}  // namespace

// PbarHbarOlap on one GPU; sd = NULL: zeroth order only, else every output
// (sd: OlapEx1, OlapEx2, EOlapEx1, EOlapEx2, each NULL-checked by the caller).
static int pcc_run(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1, const double* t2,
                   int nocc, int nso, double enuc, int qj, int sp, const int* npoints, int npg, int ncisp, int ncipg,
                   int ncik, const double* r1, int nr1, const double* r2, int nr2, const double* rpg,
                   const double* rk, const double* roota, int nra, const double* rootb, int nrb, const double* rooty,
                   int nry, const double* weightsp, const double* weightpg, const double* fsp, const double* fpg,
                   const double* fk, double* olap0, double* eolap0, double* const* sd) {
  if (!ctx) return PHF_ERR_ARG;
  const int N = nocc, V = nso - nocc;
  if (!hone || !htwo || !t1 || !t2 || !npoints || !r1 || !r2 || !rpg || !roota || !rootb || !rooty || !weightsp ||
      !weightpg || !fsp || !fpg || !fk || !olap0 || !eolap0)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: null argument");
  if (nso < 2 || N < 1 || V < 1 || qj < 0 || npg < 1 || ncipg < 1 || ncisp != 2 * qj + 1 || (ncik != 1 && ncik != 2))
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: bad dimensions (ncisp must be 2 qj + 1, ncik 1 or 2)");
  if (ncik == 2 && !rk) return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: ncik = 2 needs rk");
  if (nso > 64 || N > 32 || V > 32)
    return phf_fail(ctx, PHF_ERR_UNSUPPORTED, "phf_pcc_olap: nso <= 64 with at most 32 occupied and 32 virtual");
  const int na = npoints[0], ny = npoints[1];
  if (na < 1 || ny < 1 || na > nr1 || ny > nr2)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: npoints exceed the rotation arrays");
  if ((sp == 2 && (nra < na || nrb < ny || nry < na)) || (sp != 2 && (nra < na || nrb < na || nry < ny)))
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: root arrays shorter than the grid");
  HIPCK(hipSetDevice(ctx->device));
  using cplx = std::complex<double>;
  auto C = [](const double* p, long i) { return cplx(p[2 * i], p[2 * i + 1]); };
  // point weights: the sum over (M, N, P, Q) of WeightTot (FPUCC.f90:163-180), times fk(2) in pass 2
  const long G1 = (long)na * ny * npg, G = G1 * ncik;
  std::vector<double> w(2 * G);
  for (int pass = 0; pass < ncik; ++pass)
    for (int ia = 0; ia < na; ++ia)
      for (int iy = 0; iy < ny; ++iy)
        for (int ipg = 0; ipg < npg; ++ipg) {
          cplx tot(0.0, 0.0);
          for (int m = -qj; m <= qj; ++m)
            for (int n = -qj; n <= qj; ++n)
              for (int p = 0; p < ncipg; ++p)
                for (int q = 0; q < ncipg; ++q) {
                  const cplx wig = (sp == 2) ? phf_wignerfac(qj, m, n, roota[ia], rootb[iy], rooty[ia])
                                             : phf_wignerfac(qj, m, n, roota[ia], rootb[ia], rooty[iy]);
                  cplx wt = weightsp[(long)ia * ny + iy] * weightpg[((long)ipg * ncipg + p) * ncipg + q];
                  wt = wig * wt * std::conj(C(fsp, qj + m)) * C(fsp, qj + n);
                  tot += wt * std::conj(C(fpg, p)) * C(fpg, q);
                }
          if (pass == 1) tot *= C(fk, 1);
          const long g = pass * G1 + ((long)ia * ny + iy) * npg + ipg;
          w[2 * g] = tot.real();
          w[2 * g + 1] = tot.imag();
        }
  // chunk of grid points per launch sequence: two NSO^4 buffers per point
  const long n2 = (long)nso * nso, n4 = n2 * n2;
  const long S = (long)V * V * N * N, N4 = (long)N * N * N * N;
  const long sd_point = sd ? 16 * (12 * S + (long)V * N * N * N + 2 * N4 + 2L * V * V + 2L * N * N + 3L * V * N + 1) +
                                 23 * 256
                           : 0;
  const long per_point =
      16 * (2 * n4 + 3 * n2 + (long)V * V * N * N + 2L * N * N + 5L * V * N + 2) + 8 * 256 + sd_point;
  // (the batched index transforms put chunk * V^2 problems on the grid's z dimension)
  const long chunk = std::max<long>(1, std::min<long>({G1, (4L << 30) / per_point, 65535L / ((long)V * V)}));
  Arena a0{nullptr};
  auto carve = [&](Arena& a) {
    struct B {
      double *h1, *h2, *t1, *t2, *r1, *r2, *rpg, *rk, *w, *lam, *eps, *out, *fov, *escf;
      double *L, *Y, *Linv, *detL, *U1, *Vm, *Vinv, *detV, *bufA, *bufB, *F, *T, *sc, *C1, *fovT, *fvoC;
      double *P12, *Rpgk, *Rall, *kws;
      int *infoL, *infoV;
    } b;
    b.h1 = a.take<double>(2 * n2);
    b.h2 = a.take<double>(2 * n4);
    b.t1 = a.take<double>(2L * V * N);
    b.t2 = a.take<double>(2L * V * V * N * N);
    b.r1 = a.take<double>(2 * n2 * na);
    b.r2 = a.take<double>(2 * n2 * ny);
    b.rpg = a.take<double>(2 * n2 * npg);
    b.rk = a.take<double>(ncik == 2 ? 2 * n2 : 0);
    b.w = a.take<double>(2 * G);
    b.lam = a.take<double>(2 * G);
    b.eps = a.take<double>(2 * G);
    b.out = a.take<double>(4);
    b.fov = a.take<double>(2L * N * V);
    b.escf = a.take<double>(2);
    b.L = a.take<double>(2L * N * N * chunk);
    b.Linv = a.take<double>(2L * N * N * chunk);
    b.detL = a.take<double>(2 * chunk);
    b.Y = a.take<double>(2L * V * N * chunk);
    b.U1 = a.take<double>(2L * V * N * chunk);
    b.Vm = a.take<double>(2 * n2 * chunk);
    b.Vinv = a.take<double>(2 * n2 * chunk);
    b.detV = a.take<double>(2 * chunk);
    b.bufA = a.take<double>(2 * n4 * chunk);
    b.bufB = a.take<double>(2 * n4 * chunk);
    b.F = a.take<double>(2 * n2 * chunk);
    b.sc = a.take<double>(4 * chunk);
    b.C1 = a.take<double>(2L * V * N * chunk);
    b.fovT = a.take<double>(2L * V * N * chunk);
    b.fvoC = a.take<double>(2L * V * N * chunk);
    b.P12 = a.take<double>(2 * n2 * na * ny);
    b.Rpgk = a.take<double>(2 * n2 * npg);
    b.Rall = a.take<double>(2 * n2 * G1);
    b.kws = a.take<double>(2 * kKws);   // split-contraction partials
    b.T = a.take<double>(2L * V * V * N * N * chunk);
    b.infoL = a.take<int>(chunk);
    b.infoV = a.take<int>(chunk);
    return b;
  };
  carve(a0);
  if (sd) {
    carve_shared(a0, nso, N);
    carve_sd(a0, chunk, V, N);
  }
  if (ctx->pcc_ws_size < a0.off + 256) {
    if (ctx->pcc_ws) HIPCK(hipFree(ctx->pcc_ws));
    ctx->pcc_ws = nullptr;
    ctx->pcc_ws_size = 0;
    HIPCK(hipMalloc(&ctx->pcc_ws, a0.off + 256));
    ctx->pcc_ws_size = a0.off + 256;
  }
  Arena a1{ctx->pcc_ws};
  auto b = carve(a1);
  PccShared sh{};
  PccSD sdb{};
  if (sd) {
    sh = carve_shared(a1, nso, N);
    sdb = carve_sd(a1, chunk, V, N);
  }
  hipStream_t s = ctx->stream;
  if (sd) fill_zero(s, sdb.sum, 4 * ((long)V * N + S));
  // profiling: [0] start / end of the device work, [1..2] zeroth-order
  // stages of a chunk, [3..4] residual, [5..6] ladder GEMM, [7..8] grid sums
  hipEvent_t ev[9] = {};
  struct EvGuard {   // destroyed on every return path
    hipEvent_t* e;
    ~EvGuard() {
      for (int k = 0; k < 9; ++k)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } ev_guard{ev};
  double tms[5] = {0, 0, 0, 0, 0};
  long nladder = 0;
  const bool prof = ctx->profiling;
  if (prof)
    for (auto& e : ev) HIPCK(hipEventCreateWithFlags(&e, kTimingEventFlags));
  auto span = [&](int a, int b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
    return (double)ms;
  };
  const int noa = N / 2, nva = nso / 2 - noa;   // the ProjSz spin blocks (SP = 2)
  auto up = [&](double* d, const double* h, long nd) {
    return hipMemcpyAsync(d, h, sizeof(double) * nd, hipMemcpyHostToDevice, s);
  };
  HIPCK(up(b.h1, hone, 2 * n2));
  HIPCK(up(b.h2, htwo, 2 * n4));
  HIPCK(up(b.t1, t1, 2L * V * N));
  HIPCK(up(b.t2, t2, 2L * V * V * N * N));
  HIPCK(up(b.r1, r1, 2 * n2 * na));
  HIPCK(up(b.r2, r2, 2 * n2 * ny));
  HIPCK(up(b.rpg, rpg, 2 * n2 * npg));
  if (ncik == 2) HIPCK(up(b.rk, rk + 2 * n2, 2 * n2));   // Rk(2)
  HIPCK(up(b.w, w.data(), 2 * G));
  std::vector<int> hinfo(2 * chunk);
  const long VN = (long)V * N, V2 = (long)V * V;
  if (prof) HIPCK(hipEventRecord(ev[0], s));
  for (int pass = 0; pass < ncik; ++pass) {
    const int cj = pass == 1;
    pcc_fock_kernel<<<1, kThreads, 0, s>>>(nso, N, b.h1, b.h2, enuc, cj, b.fov, b.escf);
    if (sd) build_shared(s, nso, N, b.h1, b.h2, cj, sh);
    // T2 as a (V^2 N) x N row-major matrix, conjugated in pass 2
    const ZMat T2m{b.t2, 2L * N, 2, 0, 1, cj};
    // R_g = R1[ia] R2[iy] Rpg[ipg] (Rk(2)) for every point g = (ia ny + iy) npg + ipg, column-major, in
    // three batched launches: P12 over the (ia, iy) pairs, Rpg Rk (second pass), then P12 Rpg
    zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{b.r1, 2L * nso, 2, 2 * n2, 1, 0, 0, ny},
                  ZMat{b.r2, 2L * nso, 2, 0, 1, 0, 2 * n2, ny}, 0, 0, zout_colmajor(b.P12, nso, 2 * n2), na * ny);
    ZMat Bpg{b.rpg, 2L * nso, 2, 0, 1, 0, 2 * n2, npg};   // row-major Rpg[ipg]
    if (cj) {
      zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{b.rpg, 2L * nso, 2, 2 * n2, 1, 0}, ZMat{b.rk, 2L * nso, 2, 0, 1, 0},
                    0, 0, zout_colmajor(b.Rpgk, nso, 2 * n2), npg);
      Bpg = ZMat{b.Rpgk, 2, 2L * nso, 0, 1, 0, 2 * n2, npg};   // column-major Rpg[ipg] Rk
    }
    zgemm_batched(s, nso, nso, nso, 1, 0, ZMat{b.P12, 2, 2L * nso, 2 * n2, 1, 0, 0, npg}, Bpg, 0, 0,
                  zout_colmajor(b.Rall, nso, 2 * n2), (int)G1);
    for (long g0 = 0; g0 < G1; g0 += chunk) {
      const long cg = std::min<long>(chunk, G1 - g0);
      if (prof) HIPCK(hipEventRecord(ev[1], s));
      const double* R = b.Rall + 2 * n2 * g0;   // this chunk's rotations (column-major)
      // ProjT1: L, Y; L^-1 and det L = U0; U1 = Y L^-1; V = e^{-U1} R; V^-1
      pcc_thouless_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, R, b.t1, cj, b.L, b.Y);
      lu_inverse_batched(s, N, b.L, 2L * N * N, b.Linv, 2L * N * N, b.detL, b.infoL, (int)cg);
      pcc_vmat_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, R, b.Y, b.Linv, b.U1, b.Vm);
      lu_inverse_batched(s, nso, b.Vm, 2 * n2, b.Vinv, 2 * n2, b.detV, b.infoV, (int)cg);
      // U = V T2 V^-1 (IntTran4Sim), index by index:
      // X1[(a,b,i)][s] = sum_j T2[(a,b,i)][j] Vinv(j,s)
      zgemm_batched(s, (int)(V2 * N), nso, N, 1, 0, T2m, zmat_colmajor(b.Vinv, nso, 2 * n2), 0, 0,
                    ZOut{b.bufA, 2L * nso, 2, 2 * V2 * N * nso, 1}, (int)cg);
      // X[(a,b)][r][s] = sum_i Vinv(i,r) X1[(a,b)][i][s]
      zgemm_batched(s, nso, nso, N, 1, 0, ZMat{b.Vinv, 2L * nso, 2, 2 * n2, 1, 0, 0, (int)V2},
                    ZMat{b.bufA, 2L * nso, 2, 2 * V2 * N * nso, 1, 0, 2L * N * nso, (int)V2}, 0, 0,
                    ZOut{b.bufB, 2L * nso, 2, 2 * V2 * n2, 1, nullptr, 2 * n2, (int)V2}, (int)(cg * V2));
      // Y[a][q][(r,s)] = sum_b V(q, N+b) X[a][b][(r,s)]
      zgemm_batched(s, nso, (int)n2, V, 1, 0, ZMat{b.Vm + 2L * N * nso, 2, 2L * nso, 2 * n2, 1, 0, 0, V},
                    ZMat{b.bufB, 2 * n2, 2, 2 * V2 * n2, 1, 0, 2L * V * n2, V}, 0, 0,
                    ZOut{b.bufA, 2 * n2, 2, 2L * V * n2 * nso, 1, nullptr, 2 * n2 * nso, V}, (int)(cg * V));
      // U[p][(q,r,s)] = sum_a V(p, N+a) Y[a][(q,r,s)]
      zgemm_batched(s, nso, (int)(n2 * nso), V, 1, 0, ZMat{b.Vm + 2L * N * nso, 2, 2L * nso, 2 * n2, 1, 0},
                    ZMat{b.bufA, 2 * n2 * nso, 2, 2L * V * n2 * nso, 1, 0}, 0, 0,
                    ZOut{b.bufB, 2 * n2 * nso, 2, 2 * n4, 1}, (int)cg);
      const long goff = pass * G1 + g0;
      // TransT2's second order (bufB holds U; bufA is free scratch), CI2CC, DSTT2, CCEnergy
      pcc_tt2_prep_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, b.bufB, 2 * n4, b.F, b.sc);
      const long oc2 = pcc_transt2_gemm(s, nso, N, cg, b.bufB, b.bufA, b.F, b.sc, b.C1, b.fovT, b.fvoC, b.kws, kKws);
      pcc_tt2_finish_kernel<<<(unsigned)cg, kThreads, 0, s>>>(
          nso, N, b.C1, b.bufA + 2 * oc2, n4, b.sc, b.U1, b.detL, b.fov, b.escf, b.h2, cj, b.T, b.lam + 2 * goff,
          b.eps + 2 * goff, sd ? sdb.W1 : nullptr, sd ? sdb.W2 : nullptr, sd ? sdb.E0 : nullptr);
      if (prof) HIPCK(hipEventRecord(ev[2], s));
      if (sd) {
        if (prof) HIPCK(hipEventRecord(ev[3], s));
        pcc_residual(s, V, N, cg, sh, sdb, prof ? ev + 5 : nullptr, b.kws, kKws);
        if (prof) HIPCK(hipEventRecord(ev[4], s));
        pcc_sd_sum_kernel<<<(unsigned)((S + (long)V * N + 255) / 256), 256, 0, s>>>(
            V, N, cg, sp, noa, nva, b.w + 2 * goff, b.lam + 2 * goff, sdb.E0, sdb.W1, sdb.W2, sdb.r1, sdb.R2,
            sdb.sum);
        if (prof) HIPCK(hipEventRecord(ev[7], s));
      }
      HIPCK(hipGetLastError());
      HIPCK(hipMemcpyAsync(hinfo.data(), b.infoL, sizeof(int) * cg, hipMemcpyDeviceToHost, s));
      HIPCK(hipMemcpyAsync(hinfo.data() + chunk, b.infoV, sizeof(int) * cg, hipMemcpyDeviceToHost, s));
      HIPCK(hipStreamSynchronize(s));
      if (prof) {
        tms[1] += span(1, 2);
        if (sd) {
          tms[2] += span(3, 4);
          tms[3] += span(5, 6);
          tms[4] += span(4, 7);
          ++nladder;
        }
      }
      for (long q = 0; q < cg; ++q)
        if (hinfo[q] != 0 || hinfo[chunk + q] != 0)
          return phf_fail(ctx, PHF_ERR_SINGULAR,
                          std::string("phf_pcc_olap: singular ") + (hinfo[q] ? "occupied block of R e^T1" : "V") +
                              " at grid point " + std::to_string(g0 + q) + (pass ? " (second pass)" : ""));
    }
  }
  pcc_gridsum_kernel<<<1, kThreads, 0, s>>>(G, b.w, b.lam, b.eps, b.out);
  HIPCK(hipGetLastError());
  if (prof) {
    HIPCK(hipEventRecord(ev[8], s));
    HIPCK(hipEventSynchronize(ev[8]));
    tms[0] = span(0, 8);
    for (int k = 0; k < 5; ++k) ctx->stage_ms[k] = tms[k];
    ctx->stage_ms[5] = (double)nladder;
    ctx->nstage = 6;
  }
  double out[4];
  HIPCK(hipMemcpyAsync(out, b.out, sizeof out, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  olap0[0] = out[0];
  olap0[1] = out[1];
  eolap0[0] = out[2];
  eolap0[1] = out[3];
  if (sd) {   // device order o1, e1, o2, e2; host order OlapEx1, OlapEx2, EOlapEx1, EOlapEx2
    const long VN = (long)V * N;
    HIPCK(hipMemcpy(sd[0], sdb.sum, sizeof(double) * 2 * VN, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(sd[2], sdb.sum + 2 * VN, sizeof(double) * 2 * VN, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(sd[1], sdb.sum + 4 * VN, sizeof(double) * 2 * S, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(sd[3], sdb.sum + 4 * VN + 2 * S, sizeof(double) * 2 * S, hipMemcpyDeviceToHost));
  }
  return PHF_OK;
}

extern "C" int phf_pcc_olap0(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1,
                             const double* t2, int nocc, int nso, double enuc, int qj, int sp, const int* npoints,
                             int npg, int ncisp, int ncipg, int ncik, const double* r1, int nr1, const double* r2,
                             int nr2, const double* rpg, const double* rk, const double* roota, int nra,
                             const double* rootb, int nrb, const double* rooty, int nry, const double* weightsp,
                             const double* weightpg, const double* fsp, const double* fpg, const double* fk,
                             double* olap0, double* eolap0) {
  return pcc_run(ctx, hone, htwo, t1, t2, nocc, nso, enuc, qj, sp, npoints, npg, ncisp, ncipg, ncik, r1, nr1, r2, nr2,
                 rpg, rk, roota, nra, rootb, nrb, rooty, nry, weightsp, weightpg, fsp, fpg, fk, olap0, eolap0,
                 nullptr);
}

extern "C" int phf_pcc_olap(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1,
                            const double* t2, int nocc, int nso, double enuc, int qj, int sp, const int* npoints,
                            int npg, int ncisp, int ncipg, int ncik, const double* r1, int nr1, const double* r2,
                            int nr2, const double* rpg, const double* rk, const double* roota, int nra,
                            const double* rootb, int nrb, const double* rooty, int nry, const double* weightsp,
                            const double* weightpg, const double* fsp, const double* fpg, const double* fk,
                            double* olap0, double* olapex1, double* olapex2, double* eolap0, double* eolapex1,
                            double* eolapex2) {
  if (!ctx) return PHF_ERR_ARG;
  if (!olapex1 || !olapex2 || !eolapex1 || !eolapex2)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap: null argument");
  double* const sd[4] = {olapex1, olapex2, eolapex1, eolapex2};
  return pcc_run(ctx, hone, htwo, t1, t2, nocc, nso, enuc, qj, sp, npoints, npg, ncisp, ncipg, ncik, r1, nr1, r2, nr2,
                 rpg, rk, roota, nra, rootb, nrb, rooty, nry, weightsp, weightpg, fsp, fpg, fk, olap0, eolap0, sd);
}
