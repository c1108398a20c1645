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

// The second-order part of TransT2, CI2CC, DSTT2 and CCEnergy for one point.
// U: the transformed T2, row-major [p][q][r][s]; scratch: F (nso^2) and
// T (V^2 N^2) per point.
__global__ void __launch_bounds__(kThreads) pcc_transt2_kernel(
    int nso, int N, const double* __restrict__ Ubuf, long ustride, const double* __restrict__ U1all,
    const double* __restrict__ U0all, const double* __restrict__ fovh, const double* __restrict__ escf,
    const double* __restrict__ h2, int conj, double* __restrict__ Fbuf, double* __restrict__ Tbuf,
    double* __restrict__ lam_out, double* __restrict__ eps_out) {
  __shared__ double2 red[kThreads];
  __shared__ double2 c1s[32 * 32];
  __shared__ double2 w1s[32 * 32];
  const int g = blockIdx.x, V = nso - N;
  const int t = threadIdx.x;
  const long n1 = nso, n2 = n1 * nso, n3 = n2 * nso;
  const double* U = Ubuf + ustride * g;
  auto Ue = [&](int p, int q, int r, int s) { return ld2(U, p * n3 + q * n2 + (long)r * n1 + s); };
  double2* F = reinterpret_cast<double2*>(Fbuf) + n2 * g;
  double2* T = reinterpret_cast<double2*>(Tbuf) + (long)V * V * N * N * g;
  // F(p,q) = sum_i U(p,i,q,i)
  for (int idx = t; idx < nso * nso; idx += kThreads) {
    const int p = idx / nso, q = idx % nso;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < N; ++i) acc = cadd(acc, Ue(p, i, q, i));
    F[idx] = acc;
  }
  double2 part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * N; idx += kThreads) part = cadd(part, Ue(idx / N, idx % N, idx / N, idx % N));
  const double2 e0 = cscale(block_sum(part, red), 0.5);   // also orders the F writes before their reads
  auto Fe = [&](int p, int q) { return F[p * nso + q]; };
  auto vvoo = [&](int a, int b, int i, int j) { return Ue(N + a, N + b, i, j); };
  // C0 = 1 + E0 + E0^2/2 + sum fov fvo / 2 + sum oovv vvoo / 8
  part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * V; idx += kThreads) {
    const int i = idx / V, a = idx % V;
    part = cadd(part, cscale(cmul(Fe(i, N + a), Fe(N + a, i)), 0.5));
  }
  for (long idx = t; idx < (long)N * N * V * V; idx += kThreads) {
    const int b = idx % V, a = (idx / V) % V, j = (idx / V / V) % N, i = idx / V / V / N;
    part = cadd(part, cscale(cmul(Ue(i, j, N + a, N + b), vvoo(a, b, i, j)), 0.125));
  }
  const double2 c0 = cadd(cadd(make_double2(1.0 + e0.x, e0.y), cscale(cmul(e0, e0), 0.5)), block_sum(part, red));
  // C1(a,i)
  for (int idx = t; idx < V * N; idx += kThreads) {
    const int a = idx / N, i = idx % N;
    const double2 fvo_ai = Fe(N + a, i);
    double2 acc = cadd(fvo_ai, cmul(e0, fvo_ai));
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) s = csub(s, cmul(Fe(N + a, k), Fe(k, i)));
    for (int c = 0; c < V; ++c) s = cadd(s, cmul(Fe(N + a, N + c), Fe(N + c, i)));
    acc = cadd(acc, cscale(s, 0.5));
    double2 q = make_double2(0.0, 0.0);   // -1/4 terms
    for (int j = 0; j < N; ++j)
      for (int k = 0; k < N; ++k)
        for (int b = 0; b < V; ++b) q = cadd(q, cmul(vvoo(a, b, j, k), Ue(j, k, i, N + b)));
    for (int j = 0; j < N; ++j)
      for (int c = 0; c < V; ++c)
        for (int d = 0; d < V; ++d) q = cadd(q, cmul(Ue(j, N + a, N + c, N + d), vvoo(c, d, i, j)));
    acc = csub(acc, cscale(q, 0.25));
    double2 h = make_double2(0.0, 0.0);   // +1/2 and -1/2 terms
    for (int b = 0; b < V; ++b)
      for (int j = 0; j < N; ++j) {
        h = cadd(h, cmul(vvoo(a, b, i, j), Fe(j, N + b)));
        h = csub(h, cmul(Ue(j, N + a, i, N + b), Fe(N + b, j)));
      }
    c1s[idx] = cadd(acc, cscale(h, 0.5));
  }
  __syncthreads();
  // C2(a,b,i,j), then T = C2/C0 - [a != b, i != j] t1(a,i) t1(b,j) / 2 with t1 = C1/C0
  for (long idx = t; idx < (long)V * V * N * N; idx += kThreads) {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / N / N / V;
    const double2 v = vvoo(a, b, i, j);
    double2 acc = cadd(cscale(v, 0.25), cscale(cmul(e0, v), 0.25));
    acc = csub(acc, cscale(cmul(Fe(N + a, j), Fe(N + b, i)), 0.5));
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k)
      for (int c = 0; c < V; ++c) s = cadd(s, cmul(vvoo(a, c, j, k), Ue(k, N + b, i, N + c)));
    acc = cadd(acc, cscale(s, 0.5));
    double2 s16 = make_double2(0.0, 0.0);
    for (int c = 0; c < V; ++c)
      for (int d = 0; d < V; ++d) s16 = cadd(s16, cmul(vvoo(c, d, i, j), Ue(N + a, N + b, N + c, N + d)));
    for (int k = 0; k < N; ++k)
      for (int l = 0; l < N; ++l) s16 = cadd(s16, cmul(Ue(k, l, i, j), vvoo(a, b, k, l)));
    acc = cadd(acc, cscale(s16, 1.0 / 16.0));
    double2 s4 = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      s4 = csub(s4, cmul(vvoo(a, b, i, k), Fe(k, j)));
      s4 = csub(s4, cmul(Ue(k, N + b, i, j), Fe(N + a, k)));
    }
    for (int c = 0; c < V; ++c) {
      s4 = cadd(s4, cmul(vvoo(c, b, i, j), Fe(N + a, N + c)));
      s4 = cadd(s4, cmul(Ue(N + a, N + b, i, N + c), Fe(N + c, j)));
    }
    acc = cadd(acc, cscale(s4, 0.25));
    double2 tv = cdiv(acc, c0);
    if (a != b && i != j) {
      const double2 ta = cdiv(c1s[a * N + i], c0), tb = cdiv(c1s[b * N + j], c0);
      tv = csub(tv, cscale(cmul(ta, tb), 0.5));
    }
    T[idx] = tv;
  }
  // W1 = C1/C0 + U1 (DSTT2)
  const double2* U1 = reinterpret_cast<const double2*>(U1all) + (long)g * V * N;
  for (int idx = t; idx < V * N; idx += kThreads) {
    const int a = idx / N, i = idx % N;
    w1s[idx] = cadd(cdiv(c1s[idx], c0), U1[(long)i * V + a]);   // U1 is column-major V x N
  }
  __syncthreads();
  // CCEnergy(W1, W2) with W2 = ASymm2(T): fov . W1 + 1/4 sum ERI(i,j,a,b) tau(a,b,i,j)
  const double sg = conj ? -1.0 : 1.0;
  const double2* fov = reinterpret_cast<const double2*>(fovh);
  part = make_double2(0.0, 0.0);
  for (int idx = t; idx < N * V; idx += kThreads) part = cadd(part, cmul(fov[idx], w1s[(idx % V) * N + idx / V]));
  for (long idx = t; idx < (long)V * V * N * N; idx += kThreads) {
    const int j = idx % N, i = (idx / N) % N, b = (idx / N / N) % V, a = idx / N / N / V;
    auto Tt = [&](int aa, int bb, int ii, int jj) { return T[(((long)aa * V + bb) * N + ii) * N + jj]; };
    double2 w2 = csub(csub(Tt(a, b, i, j), Tt(a, b, j, i)), csub(Tt(b, a, i, j), Tt(b, a, j, i)));
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

}  // namespace

extern "C" int phf_pcc_olap0(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1,
                             const double* t2, int nocc, int nso, double enuc, int qj, int sp, const int* npoints,
                             int npg, int ncisp, int ncipg, int ncik, const double* r1, int nr1, const double* r2,
                             int nr2, const double* rpg, const double* rk, const double* roota, int nra,
                             const double* rootb, int nrb, const double* rooty, int nry, const double* weightsp,
                             const double* weightpg, const double* fsp, const double* fpg, const double* fk,
                             double* olap0, double* eolap0) {
  if (!ctx) return PHF_ERR_ARG;
  const int N = nocc, V = nso - nocc;
  if (!hone || !htwo || !t1 || !t2 || !npoints || !r1 || !r2 || !rpg || !roota || !rootb || !rooty || !weightsp ||
      !weightpg || !fsp || !fpg || !fk || !olap0 || !eolap0)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap0: null argument");
  if (nso < 2 || N < 1 || V < 1 || qj < 0 || npg < 1 || ncipg < 1 || ncisp != 2 * qj + 1 || (ncik != 1 && ncik != 2))
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap0: bad dimensions (ncisp must be 2 qj + 1, ncik 1 or 2)");
  if (ncik == 2 && !rk) return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap0: ncik = 2 needs rk");
  if (nso > 64 || N > 32 || V > 32)
    return phf_fail(ctx, PHF_ERR_UNSUPPORTED, "phf_pcc_olap0: nso <= 64 with at most 32 occupied and 32 virtual");
  const int na = npoints[0], ny = npoints[1];
  if (na < 1 || ny < 1 || na > nr1 || ny > nr2)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap0: npoints exceed the rotation arrays");
  if ((sp == 2 && (nra < na || nrb < ny || nry < na)) || (sp != 2 && (nra < na || nrb < na || nry < ny)))
    return phf_fail(ctx, PHF_ERR_ARG, "phf_pcc_olap0: root arrays shorter than the grid");
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
  const long per_point = 16 * (2 * n4 + 5 * n2 + (long)V * V * N * N + 2L * N * N + 2L * V * N) + 4 * 256;
  // (the batched index transforms put chunk * V^2 problems on the grid's z dimension)
  const long chunk = std::max<long>(1, std::min<long>({G1, (4L << 30) / per_point, 65535L / ((long)V * V)}));
  Arena a0{nullptr};
  auto carve = [&](Arena& a) {
    struct B {
      double *h1, *h2, *t1, *t2, *r1, *r2, *rpg, *rk, *w, *lam, *eps, *out, *fov, *escf;
      double *R, *L, *Y, *Linv, *detL, *U1, *Vm, *Vinv, *detV, *bufA, *bufB, *F, *T, *tmp;
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
    b.R = a.take<double>(2 * n2 * chunk);
    b.tmp = a.take<double>(2 * n2 * chunk);
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
    b.T = a.take<double>(2L * V * V * N * N * chunk);
    b.infoL = a.take<int>(chunk);
    b.infoV = a.take<int>(chunk);
    return b;
  };
  carve(a0);
  if (ctx->pcc_ws_size < a0.off + 256) {
    if (ctx->pcc_ws) HIPCK(hipFree(ctx->pcc_ws));
    ctx->pcc_ws = nullptr;
    ctx->pcc_ws_size = 0;
    HIPCK(hipMalloc(&ctx->pcc_ws, a0.off + 256));
    ctx->pcc_ws_size = a0.off + 256;
  }
  Arena a1{ctx->pcc_ws};
  auto b = carve(a1);
  hipStream_t s = ctx->stream;
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
  for (int pass = 0; pass < ncik; ++pass) {
    const int cj = pass == 1;
    pcc_fock_kernel<<<1, kThreads, 0, s>>>(nso, N, b.h1, b.h2, enuc, cj, b.fov, b.escf);
    // T2 as a (V^2 N) x N row-major matrix, conjugated in pass 2
    const ZMat T2m{b.t2, 2L * N, 2, 0, 1, cj};
    for (long g0 = 0; g0 < G1; g0 += chunk) {
      const long cg = std::min<long>(chunk, G1 - g0);
      // R_g = R1[ia] R2[iy] Rpg[ipg] (Rk(2)): row-major inputs, column-major products
      // tmp = R1 R2 over (ia, iy) pairs of this chunk's points, then tmp Rpg, then Rk
      for (long q = 0; q < cg; ++q) {
        const long gg = g0 + q;
        const long ia = gg / ((long)ny * npg), iy = (gg / npg) % ny, ipg = gg % npg;
        const ZMat A1{b.r1 + 2 * n2 * ia, 2L * nso, 2, 0, 1, 0};
        const ZMat B2{b.r2 + 2 * n2 * iy, 2L * nso, 2, 0, 1, 0};
        zgemm_batched(s, nso, nso, nso, 1, 0, A1, B2, 0, 0, zout_colmajor(b.tmp + 2 * n2 * q, nso, 0), 1);
        const ZMat Bp{b.rpg + 2 * n2 * ipg, 2L * nso, 2, 0, 1, 0};
        if (cj) {
          zgemm_batched(s, nso, nso, nso, 1, 0, zmat_colmajor(b.tmp + 2 * n2 * q, nso, 0), Bp, 0, 0,
                        zout_colmajor(b.Vm + 2 * n2 * q, nso, 0), 1);
          const ZMat Bk{b.rk, 2L * nso, 2, 0, 1, 0};
          zgemm_batched(s, nso, nso, nso, 1, 0, zmat_colmajor(b.Vm + 2 * n2 * q, nso, 0), Bk, 0, 0,
                        zout_colmajor(b.R + 2 * n2 * q, nso, 0), 1);
        } else {
          zgemm_batched(s, nso, nso, nso, 1, 0, zmat_colmajor(b.tmp + 2 * n2 * q, nso, 0), Bp, 0, 0,
                        zout_colmajor(b.R + 2 * n2 * q, nso, 0), 1);
        }
      }
      // ProjT1: L, Y; L^-1 and det L = U0; U1 = Y L^-1; V = e^{-U1} R; V^-1
      pcc_thouless_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, b.R, b.t1, cj, b.L, b.Y);
      lu_inverse_batched(s, N, b.L, 2L * N * N, b.Linv, 2L * N * N, b.detL, b.infoL, (int)cg);
      pcc_vmat_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, b.R, b.Y, b.Linv, b.U1, b.Vm);
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
      pcc_transt2_kernel<<<(unsigned)cg, kThreads, 0, s>>>(nso, N, b.bufB, 2 * n4, b.U1, b.detL, b.fov, b.escf,
                                                            b.h2, cj, b.F, b.T, b.lam + 2 * goff,
                                                            b.eps + 2 * goff);
      HIPCK(hipGetLastError());
      HIPCK(hipMemcpyAsync(hinfo.data(), b.infoL, sizeof(int) * cg, hipMemcpyDeviceToHost, s));
      HIPCK(hipMemcpyAsync(hinfo.data() + chunk, b.infoV, sizeof(int) * cg, hipMemcpyDeviceToHost, s));
      HIPCK(hipStreamSynchronize(s));
      for (long q = 0; q < cg; ++q)
        if (hinfo[q] != 0 || hinfo[chunk + q] != 0)
          return phf_fail(ctx, PHF_ERR_SINGULAR,
                          std::string("phf_pcc_olap0: singular ") + (hinfo[q] ? "occupied block of R e^T1" : "V") +
                              " at grid point " + std::to_string(g0 + q) + (pass ? " (second pass)" : ""));
    }
  }
  pcc_gridsum_kernel<<<1, kThreads, 0, s>>>(G, b.w, b.lam, b.eps, b.out);
  HIPCK(hipGetLastError());
  double out[4];
  HIPCK(hipMemcpyAsync(out, b.out, sizeof out, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  olap0[0] = out[0];
  olap0[1] = out[1];
  eolap0[0] = out[2];
  eolap0[1] = out[3];
  return PHF_OK;
}
