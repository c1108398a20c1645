// Projected SD-space overlap times vector (BuildSC, PHFTools.f90:464-537, with
// BuildSC_Vec, BuildSC_dr.f90:8-1244) on the FP64 matrix cores.
//
// Per grid point g (R = R1[il] R2[iy] Rpg[ipg], C0 the first NOcc unit vectors):
//   Y = R C0 (C0^H R C0)^{-1} C0^H and s = det(C0^H R C0)      (RDM, :249-263)
//   A = (I - Y) R,   B^T = R^H Y  (N x NOcc),   K = R^H Y R.
// For an input (c0, c1, c2) let c2a be c2 antisymmetrised in both index pairs,
// c2a(p,q,r,s) = c2(p,q,r,s) - c2(q,p,r,s) - c2(p,q,s,r) + c2(q,p,s,r).  Then
// BuildSC_Vec on the rotated c1' = R c1 R^H, c2' = ao2mo4(c2, R^H) equals
//   P(q,r) = sum_{p,s} K(s,p) c2a(p,q,r,s)
//   t0     = c0 + tr(K c1) - tr(K P) / 2
//   t1     = t0 Y + A (c1 - P) B^T
//   t2     = c2a x1 A x2 A x3 B x4 B  +  Asym[ Y(a,c) Q(b,d) ],  Q = t1 - t0 Y / 2
// (x_k M applies M to index k; Asym[f] = f(abcd) - f(bacd) - f(abdc) + f(badc));
// the generated sequence's nine NSO^5 passes per point collapse to one pair
// transform per index pair, and every entry of t2 with c or d >= NOcc is zero.
// The result is sum_g w_g t_g / sum_g w_g, w_g = weightsp(il,iy) s_g W(ipg).
//
// c2a is antisymmetric in (p,q) and in (r,s), so the pair transforms act on
// pair spaces: with p < q, r < s, a < b, i < j,
//   t2(a,b,i,j) = sum KA(ab,pq) KB(ij,rs) c2a(pq,rs) + rank term,
//   KA(ab,pq) = A(a,p) A(b,q) - A(a,q) A(b,p),  KB likewise from B,
// 4 x fewer products per pair than the Kronecker A (x) A, and only i < j < NOcc.
//
// Index blocks: indices 1, 2 of c2a run over e12_in values from off12_in
// (the output rows a, b over e12_out from off12_out), indices 3, 4 over e34_in
// from off34_in, the output c, d over NOcc.  The full tensors use (N, 0, N, 0;
// N, 0); the packed LinDep vectors (LinDep.py:156-187), whose c2 lives on
// virtual x virtual x occupied x occupied, use (V, NOcc, V, NOcc; NOcc, 0).
//
// GEMM form over a chunk of Gc points and nv vectors (complex FP64, zgemm):
//   stage 1, batched over g:  X_g = KA_g C2         (rows a<b, columns (v, r<s))
//   stage 2, contraction over (g, r<s), split into groups of points:
//                             T2 += X KBW^T,  KBW_g = w_g KB_g   (rows (a<b, v), columns i<j)
//   P  = C2P Kvec  (all points at once);  RK += YW Qm  (sum_g w_g Y_g (x) Q_g)
// plus one workgroup per (vector, group of points) for t0, t1 and Q.
#include <algorithm>
#include <complex>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "phf_internal.h"

namespace {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.y + b.x * r;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ inline double2 ld2(const double* p, long i) { return make_double2(p[2 * i], p[2 * i + 1]); }
__device__ inline void st2(double* p, long i, double2 v) {
  p[2 * i] = v.x;
  p[2 * i + 1] = v.y;
}

struct ScBlk {
  // element ranges: indices 1, 2 of c2a (in) and rows a, b of the result (out);
  // indices 3, 4 of c2a (in); the result's c, d run over the first O
  int e12_in, off12_in, e12_out, off12_out, e34_in, off34_in, O;
  // pair counts p < q of those ranges (the antisymmetric pair spaces)
  __host__ __device__ long p12_in() const { return (long)e12_in * (e12_in - 1) / 2; }
  __host__ __device__ long p12_out() const { return (long)e12_out * (e12_out - 1) / 2; }
  __host__ __device__ long p34_in() const { return (long)e34_in * (e34_in - 1) / 2; }
  __host__ __device__ long p34_out() const { return (long)O * (O - 1) / 2; }
};

// Pair index t = q (q - 1) / 2 + p of p < q (q-major, the order of LinDep.py:165-170).
__device__ inline void pair_of(long t, int& p, int& q) {
  int b = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
  while ((long)b * (b - 1) / 2 > t) --b;
  while ((long)(b + 1) * b / 2 <= t) ++b;
  q = b;
  p = (int)(t - (long)b * (b - 1) / 2);
}
__device__ inline long pair_index(int p, int q) { return (long)q * (q - 1) / 2 + p; }   // p < q

// M_g = R_g[:O, :O] (column-major, contiguous) for the LU.
__global__ void sc_overlap_kernel(int n, int o, const double* __restrict__ R, double* __restrict__ M) {
  const long g = blockIdx.x;
  for (int idx = threadIdx.x; idx < o * o; idx += blockDim.x) {
    const int i = idx % o, j = idx / o;
    st2(M, g * o * o + idx, ld2(R, g * (long)n * n + i + (long)n * j));
  }
}

// w_g = host weight x det_g; ovlp = sum_g w_g (one workgroup).
__global__ void sc_weight_kernel(int G, const double* __restrict__ wh, const double* __restrict__ det,
                                 double* __restrict__ ws, double* __restrict__ ovlp) {
  __shared__ double red[2][256];
  double sr = 0.0, si = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) {
    const double2 w = cscale(ld2(det, g), wh[g]);
    st2(ws, g, w);
    sr += w.x;
    si += w.y;
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) st2(ovlp, 0, make_double2(red[0][0], red[1][0]));
}

// Per point of a chunk (blockIdx.y): the pair factors
//   KA_g((a,b),(p,q))   = A(a',p') A(b',q') - A(a',q') A(b',p')        a < b, p < q
//   KBW_g((i,j),(r,s))  = w_g [B(i,r') B(j,s') - B(i,s') B(j,r')]      i < j, r < s
// (B(i, x) = BT(x, i)), Kvec((p,s)) = K(s', p') and YW((a,c)) = w_g Y(a', c).
__global__ void sc_factor_kernel(ScBlk k, int n, int g0, int G, const double* __restrict__ A,
                                 const double* __restrict__ BT, const double* __restrict__ Y,
                                 const double* __restrict__ K, const double* __restrict__ ws,
                                 double* __restrict__ KA, double* __restrict__ KBW, double* __restrict__ Kvec,
                                 double* __restrict__ YW) {
  const int gl = blockIdx.y;
  const long g = g0 + gl;
  const bool ok = g < G;
  const int o = k.O;
  const long n2 = (long)n * n, no = (long)n * o;
  const double2 w = ok ? ld2(ws, g) : make_double2(0.0, 0.0);
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double2 z = make_double2(0.0, 0.0);
  {
    const long po = k.p34_out(), pi = k.p34_in(), nkb = po * pi;
    for (long idx = t0; idx < nkb; idx += stride) {
      double2 v = z;
      if (ok) {
        int i, j, r, s;
        pair_of(idx % po, i, j);
        pair_of(idx / po, r, s);
        auto b = [&](int row, int x) { return ld2(BT, g * no + k.off34_in + x + (long)n * row); };
        v = cmul(w, csub(cmul(b(i, r), b(j, s)), cmul(b(i, s), b(j, r))));
      }
      st2(KBW, gl * nkb + idx, v);
    }
  }
  {
    const long po = k.p12_out(), pi = k.p12_in(), nka = po * pi;
    for (long idx = t0; idx < nka; idx += stride) {
      double2 v = z;
      if (ok) {
        int a, b, p, q;
        pair_of(idx % po, a, b);
        pair_of(idx / po, p, q);
        auto m = [&](int row, int col) {
          return ld2(A, g * n2 + k.off12_out + row + (long)n * (k.off12_in + col));
        };
        v = csub(cmul(m(a, p), m(b, q)), cmul(m(a, q), m(b, p)));
      }
      st2(KA, gl * nka + idx, v);
    }
  }
  {
    const long nk = (long)k.e12_in * k.e34_in;
    for (long idx = t0; idx < nk; idx += stride) {
      const int p = idx % k.e12_in, s = idx / k.e12_in;
      st2(Kvec, gl * nk + idx, ok ? ld2(K, g * n2 + k.off34_in + s + (long)n * (k.off12_in + p)) : z);
    }
  }
  {
    const long ny = (long)k.e12_out * o;
    for (long idx = t0; idx < ny; idx += stride) {
      const int a = idx % k.e12_out, c = idx / k.e12_out;
      st2(YW, gl * ny + idx, ok ? cmul(w, ld2(Y, g * no + k.off12_out + a + (long)n * c)) : z);
    }
  }
}

// Full tensors in (host C order, c2[v][p][q][r][s], c1[v][p][q]) -> c2a in
// pair form C2[(p<q) + P12 (v + nv (r<s))] (stage 1) and in full-index form
// C2P[(q,r) + N^2 v + N^2 nv (p,s)] (the P contraction); c1 -> column-major.
__global__ void sc_prep_full_kernel(int n, int nv, const double* __restrict__ c2, const double* __restrict__ c1,
                                    double* __restrict__ C2, double* __restrict__ C2P, double* __restrict__ C1) {
  const long n2 = (long)n * n, tot = n2 * n2 * nv, np = (long)n * (n - 1) / 2;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += stride) {
    const long pq = idx % n2, rem = idx / n2;
    const long v = rem % nv, rs = rem / nv;
    const int p = pq % n, q = pq / n, r = rs % n, s = rs / n;
    const double* cv = c2 + 2 * v * n2 * n2;
    auto x = [&](int i, int j, int kk, int l) { return ld2(cv, ((long)(i * n + j) * n + kk) * n + l); };
    const double2 val = cadd(csub(x(p, q, r, s), x(q, p, r, s)), csub(x(q, p, s, r), x(p, q, s, r)));
    st2(C2P, (q + (long)n * r) + n2 * v + n2 * nv * (p + (long)n * s), val);
    if (p < q && r < s) st2(C2, pair_index(p, q) + np * (v + nv * pair_index(r, s)), val);
    if (rs == 0) st2(C1, pq + n2 * v, ld2(c1, v * n2 + (long)p * n + q));   // C1[(p,q)] = c1[v][p][q]
  }
}

// Packed vectors (vin[v][dim], LinDep.py:156-171: c1[a,i] after c0, then
// c2[a,b,i,j] for i<j, a<b, (i,j)-major) -> the same layouts over (V, V, O, O).
__global__ void sc_prep_packed_kernel(int o, int nvir, int nv, long dim, const double* __restrict__ vin,
                                      double* __restrict__ C2, double* __restrict__ C2P, double* __restrict__ C1,
                                      double* __restrict__ c0v) {
  const long v2 = (long)nvir * nvir, vo = (long)nvir * o, tot = v2 * o * o * nv;
  const long nab = (long)nvir * (nvir - 1) / 2, nij = (long)o * (o - 1) / 2;
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long idx = t0; idx < tot; idx += stride) {
    const long pq = idx % v2, rem = idx / v2;
    const long v = rem % nv, rs = rem / nv;
    const int p = pq % nvir, q = pq / nvir, r = rs % o, s = rs / o;
    double2 val = make_double2(0.0, 0.0);
    if (p != q && r != s) {
      const long kk = 1 + vo + pair_index(min(r, s), max(r, s)) * nab + pair_index(min(p, q), max(p, q));
      val = ld2(vin, v * dim + kk);
      if ((p < q) != (r < s)) val = make_double2(-val.x, -val.y);
    }
    st2(C2P, (q + (long)nvir * r) + vo * v + vo * nv * (p + (long)nvir * s), val);
  }
  for (long idx = t0; idx < nab * nv * nij; idx += stride) {   // C2[(a<b) + nab (v + nv (i<j))]
    const long pab = idx % nab, rem = idx / nab;
    const long v = rem % nv, pij = rem / nv;
    st2(C2, idx, ld2(vin, v * dim + 1 + vo + pij * nab + pab));
  }
  for (long idx = t0; idx < vo * nv; idx += stride) {
    const long v = idx / vo, e = idx % vo;   // e = a + V i
    const int a = e % nvir, i = e / nvir;
    st2(C1, idx, ld2(vin, v * dim + 1 + (long)i * nvir + a));
    if (e == 0) st2(c0v, v, ld2(vin, v * dim));
  }
}

constexpr int SC_T1_REG = 16;   // t1 elements owned per thread: e12_out * O <= 16 NT

// One workgroup of NT threads per (group of points, vector): t0, t1 and Q of
// each point of the group; partial sums of w_g t0 and w_g t1 into slot
// `slot0 + group`.
template <int NT>
__global__ __launch_bounds__(NT) void sc_point_kernel(ScBlk k, int n, int nv, int g0, int gs, int G, int nslot,
                                                       int slot0, const double* __restrict__ c0v,
                                                       const double* __restrict__ C1, const double* __restrict__ Pm,
                                                       int gc, const double* __restrict__ A,
                                                       const double* __restrict__ BT, const double* __restrict__ Y,
                                                       const double* __restrict__ K, const double* __restrict__ ws,
                                                       double* __restrict__ Qm, double* __restrict__ t0p,
                                                       double* __restrict__ t1p) {
  extern __shared__ double2 sm[];
  const int grp = blockIdx.x, v = blockIdx.y, tid = threadIdx.x;
  const int o = k.O, e1 = k.e12_in, e3 = k.e34_in, eo = k.e12_out;
  const long n2 = (long)n * n, no = (long)n * o;
  const long nd = (long)e1 * e3, nt1 = (long)eo * o;
  double2* D = sm;                 // e1 x e3
  double2* T = sm + nd;            // e1 x o   (D B^T)
  double2* red = T + (long)e1 * o; // 8 (reductions)
  __shared__ double2 s_t0;
  const double2 c0 = ld2(c0v, v);
  const double2 z = make_double2(0.0, 0.0);
  double2 acc[SC_T1_REG];
#pragma unroll
  for (int e = 0; e < SC_T1_REG; ++e) acc[e] = z;
  double2 acc0 = z;
  for (int gi = 0; gi < gs; ++gi) {
    const int gl = grp * gs + gi;
    const long g = g0 + gl;
    const long qbase = nt1 * v + nt1 * nv * gl;
    if (gl >= gc) break;
    if (g >= G) {
      for (long e = tid; e < nt1; e += NT) st2(Qm, qbase + e, z);
      continue;
    }
    // D = c1 - P and the traces tr(K c1), tr(K P)
    double2 s1 = z, s2 = z;
    for (long e = tid; e < nd; e += NT) {
      const int q = e % e1, r = e / e1;
      const double2 c1v = ld2(C1, nd * v + e);
      const double2 pv = ld2(Pm, e + nd * v + nd * nv * gl);
      D[e] = csub(c1v, pv);
      const double2 kv = ld2(K, g * n2 + k.off34_in + r + (long)n * (k.off12_in + q));
      s1 = cadd(s1, cmul(kv, c1v));
      s2 = cadd(s2, cmul(kv, pv));
    }
    double2 s = csub(s1, cscale(s2, 0.5));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      s.x += __shfl_xor(s.x, off);
      s.y += __shfl_xor(s.y, off);
    }
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
      double2 t = c0;
      for (int w = 0; w < NT / 64; ++w) t = cadd(t, red[w]);
      s_t0 = t;
    }
    // T = D B^T:  T(p, i) = sum_s D(p, s) BT(s', i)
    for (long e = tid; e < (long)e1 * o; e += NT) {
      const int p = e % e1, i = e / e1;
      double2 t = z;
      for (int s3 = 0; s3 < e3; ++s3) t = cadd(t, cmul(D[p + (long)e1 * s3], ld2(BT, g * no + k.off34_in + s3 + (long)n * i)));
      T[e] = t;
    }
    __syncthreads();
    const double2 t0 = s_t0, w = ld2(ws, g);
    // t1(a, i) = t0 Y(a', i) + sum_p A(a', p') T(p, i);  Q = t1 - t0 Y / 2
#pragma unroll
    for (int r = 0; r < SC_T1_REG; ++r) {
      const long e = tid + (long)NT * r;
      if (e < nt1) {
        const int a = e % eo, i = e / eo;
        const double2 yv = ld2(Y, g * no + k.off12_out + a + (long)n * i);
        double2 t = cmul(t0, yv);
        for (int p = 0; p < e1; ++p) t = cadd(t, cmul(ld2(A, g * n2 + k.off12_out + a + (long)n * (k.off12_in + p)), T[p + (long)e1 * i]));
        st2(Qm, qbase + e, csub(t, cscale(cmul(t0, yv), 0.5)));
        acc[r] = cadd(acc[r], cmul(w, t));
      }
    }
    if (tid == 0) acc0 = cadd(acc0, cmul(w, t0));
    __syncthreads();
  }
  const long slot = (long)v * nslot + slot0 + grp;
#pragma unroll
  for (int r = 0; r < SC_T1_REG; ++r) {
    const long e = tid + (long)NT * r;
    if (e < nt1) st2(t1p, slot * nt1 + e, acc[r]);
  }
  if (tid == 0) st2(t0p, slot, acc0);
}

// sum of the per-slot partials x[(v nslot + sl) * len + e] over sl, eight
// loads in flight
__device__ inline double2 sum_slots(const double* x, long v, int nslot, long len, long e) {
  double2 acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = make_double2(0.0, 0.0);
  const double* base = x + 2 * (v * nslot * len + e);
  int sl = 0;
  for (; sl + 8 <= nslot; sl += 8)
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = cadd(acc[u], ld2(base, (long)(sl + u) * len));
  for (; sl < nslot; ++sl) acc[0] = cadd(acc[0], ld2(base, (long)sl * len));
#pragma unroll
  for (int u = 1; u < 8; ++u) acc[0] = cadd(acc[0], acc[u]);
  return acc[0];
}

// sum over the split partials of stage 2 at T2p element `e` (stride `ps`)
__device__ inline double2 sum_parts(const double* T2p, int nsplit, long ps, long e) {
  double2 t = make_double2(0.0, 0.0);
  for (int j = 0; j < nsplit; ++j) t = cadd(t, ld2(T2p, j * ps + e));
  return t;
}

// Full outputs (host C order t2[v][a][b][c][d], t1[v][a][b], t0[v]).  T2 holds
// the pair-space transform T2[(a<b) + P12 v + P12 nv (c<d)] over c, d < O.
__global__ void sc_final_full_kernel(int n, int o, int nv, int nsplit, const double* __restrict__ T2p,
                                     const double* __restrict__ RK, const double* __restrict__ t1p,
                                     const double* __restrict__ t0p, int nslot, const double* __restrict__ ovlp,
                                     double* __restrict__ t0, double* __restrict__ t1, double* __restrict__ t2) {
  const long n2 = (long)n * n, no = (long)n * o, np = (long)n * (n - 1) / 2, op = (long)o * (o - 1) / 2;
  const long ps = np * nv * op;   // one split partial
  const double2 ov = ld2(ovlp, 0);
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  auto rk = [&](int a, int c, int b, int d, long v) {   // RK((a,c), (b,d,v))
    return ld2(RK, (a + (long)n * c) + no * ((b + (long)n * d) + no * v));
  };
  for (long idx = t0i; idx < n2 * n2 * nv; idx += stride) {
    const int d = idx % n, c = (idx / n) % n, b = (idx / n2) % n, a = (idx / (n2 * n)) % n;
    const long v = idx / (n2 * n2);
    double2 val = make_double2(0.0, 0.0);
    if (c < o && d < o) {
      if (a != b && c != d) {
        val = sum_parts(T2p, nsplit, ps, pair_index(min(a, b), max(a, b)) + np * (v + nv * pair_index(min(c, d), max(c, d))));
        if ((a < b) != (c < d)) val = make_double2(-val.x, -val.y);
      }
      val = cadd(val, csub(rk(a, c, b, d, v), rk(b, c, a, d, v)));
      val = cadd(val, csub(rk(b, d, a, c, v), rk(a, d, b, c, v)));
      val = cdiv(val, ov);
    }
    st2(t2, idx, val);
  }
  for (long idx = t0i; idx < n2 * nv; idx += stride) {
    const int b = idx % n, a = (idx / n) % n;
    const long v = idx / n2;
    double2 val = make_double2(0.0, 0.0);
    if (b < o) {
      val = cdiv(sum_slots(t1p, v, nslot, no, a + (long)n * b), ov);
    }
    st2(t1, idx, val);
  }
  for (long v = t0i; v < nv; v += stride) {
    st2(t0, v, cdiv(sum_slots(t0p, v, nslot, 1, 0), ov));
  }
}

// Packed outputs vout[v][dim] in the order of LinDep.py:173-187.
__global__ void sc_final_packed_kernel(int o, int nvir, int nv, long dim, int nsplit, const double* __restrict__ T2p,
                                       const double* __restrict__ RK, const double* __restrict__ t1p,
                                       const double* __restrict__ t0p, int nslot, const double* __restrict__ ovlp,
                                       double* __restrict__ vout) {
  const long vo = (long)nvir * o;
  const long nab = (long)nvir * (nvir - 1) / 2, nij = (long)o * (o - 1) / 2;
  const long ps = nab * nv * nij;
  const double2 ov = ld2(ovlp, 0);
  const long stride = (long)gridDim.x * blockDim.x;
  auto rk = [&](int a, int c, int b, int d, long v) {
    return ld2(RK, (a + (long)nvir * c) + vo * ((b + (long)nvir * d) + vo * v));
  };
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < dim * nv; idx += stride) {
    const long v = idx / dim, kk = idx % dim;
    double2 val = make_double2(0.0, 0.0);
    if (kk == 0) {
      val = sum_slots(t0p, v, nslot, 1, 0);
    } else if (kk <= vo) {
      const long e = kk - 1;   // = i V + a
      const int a = e % nvir, i = e / nvir;
      val = sum_slots(t1p, v, nslot, vo, a + (long)nvir * i);
    } else {
      const long e = kk - 1 - vo;
      const long pij = e / nab, pab = e % nab;
      int a, b, i, j;
      pair_of(pij, i, j);
      pair_of(pab, a, b);
      val = sum_parts(T2p, nsplit, ps, pab + nab * (v + nv * pij));
      val = cadd(val, csub(rk(a, i, b, j, v), rk(b, i, a, j, v)));
      val = cadd(val, csub(rk(b, j, a, i, v), rk(a, j, b, i, v)));
    }
    st2(vout, idx, cdiv(val, ov));
  }
}

inline int grid_for(long n, int per = 256, int cap = 16384) {
  return (int)std::max<long>(1, std::min<long>((n + per - 1) / per, cap));
}

void free_sc_grid(phf_ctx* ctx) {
  double** ptrs[] = {&ctx->sc_R, &ctx->sc_Y, &ctx->sc_BT, &ctx->sc_K, &ctx->sc_A, &ctx->sc_w, &ctx->sc_ovlp};
  for (double** p : ptrs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  ctx->sc_G = 0;
  ctx->sc_nso = ctx->sc_nocc = 0;
}

int ensure_sc_ws(phf_ctx* ctx, size_t bytes) {
  if (ctx->sc_ws && ctx->sc_ws_size >= bytes) return PHF_OK;
  HIPCK(hipStreamSynchronize(ctx->stream));
  if (ctx->sc_ws) (void)hipFree(ctx->sc_ws);
  ctx->sc_ws = nullptr;
  ctx->sc_ws_size = 0;
  HIPCK(hipMalloc(&ctx->sc_ws, bytes));
  ctx->sc_ws_size = bytes;
  return PHF_OK;
}

// Sizes (complex elements) of one pass over nv vectors with chunks of
// gc = nsplit * gs points.
struct ScSizes {
  int nv = 0, gs = 1, nsplit = 1, gc = 1, nchunk = 1, nslot = 1;
  int pg = 1, npgrp = 1;   // points per t0/t1 workgroup, its groups per chunk
  long c2 = 0, c2p = 0, c1 = 0, kb = 0, kaw = 0, kvec = 0, yw = 0, pm = 0, qm = 0, out1 = 0, t2p = 0, rk = 0,
       t1p = 0, io = 0;
  size_t bytes() const {
    const long e = c2 + c2p + c1 + nv + kb + kaw + kvec + yw + pm + qm + out1 + t2p + rk + t1p + nslot * (long)nv +
                   io + 8;
    return sizeof(double) * 2 * (size_t)e + 16 * 256;
  }
};

ScSizes sc_sizes(const ScBlk& k, int G, int nv, int gs, int nsplit, long io) {
  ScSizes z;
  z.nv = nv;
  z.gs = gs;
  z.nsplit = nsplit;
  z.gc = gs * nsplit;
  z.nchunk = (G + z.gc - 1) / z.gc;
  // one point per one-wave workgroup when its t1 block fits (e12_out O <= 1024),
  // else enough 256-thread (group, vector) workgroups to hide load latencies
  z.pg = ((long)k.e12_out * k.O <= 16 * 64) ? 1 : (int)std::max<long>(1, std::min<long>(z.gc, ((long)z.gc * nv + 8191) / 8192));
  z.npgrp = (z.gc + z.pg - 1) / z.pg;
  z.nslot = z.nchunk * z.npgrp;
  const long e1 = k.e12_in, e3 = k.e34_in, eo = k.e12_out, o = k.O;
  z.c2 = k.p12_in() * k.p34_in() * nv;
  z.c2p = e1 * e1 * e3 * e3 * nv;
  z.c1 = e1 * e3 * nv;
  z.kb = k.p34_out() * k.p34_in() * z.gc;
  z.kaw = k.p12_out() * k.p12_in() * z.gc;
  z.kvec = e1 * e3 * z.gc;
  z.yw = eo * o * z.gc;
  z.pm = e1 * e3 * nv * (long)z.gc;
  z.qm = eo * o * nv * (long)z.gc;
  z.out1 = k.p12_out() * nv * k.p34_in() * z.gc;
  z.t2p = k.p12_out() * nv * k.p34_out() * nsplit;
  z.rk = eo * o * eo * o * nv;
  z.t1p = eo * o * nv * (long)z.nslot;
  z.io = io;
  return z;
}

// Vectors per pass and the chunking: the largest pass that fits `budget`
// bytes with launch grids inside their limits; stage 2 split into groups of
// points so that (tiles x groups) reaches about a thousand workgroups.
bool sc_plan(const ScBlk& k, int G, int nv_total, long io_per_vec, size_t budget, ScSizes& out) {
  for (int nv = std::min(nv_total, 1024); nv >= 1; nv = (nv + 1) / 2) {
    const long m2 = k.p12_out() * nv, n2 = k.p34_out();
    const long tm = (m2 + 63) / 64, tn = (n2 + 31) / 32;
    const long tiles = std::max<long>(1, tm * tn);
    const int target = (int)std::max<long>(1, std::min<long>(256, (1024 + tiles - 1) / tiles));
    const long eoo = (long)k.e12_out * k.O;
    const bool fits_grid = ((long)nv * k.p34_in() + 31) / 32 <= 65535 && (eoo * nv + 31) / 32 <= 65535;
    int gcap = G;
    while (fits_grid && gcap >= 1) {
      const int nsplit = std::min(target, gcap);
      const int gs = (gcap + nsplit - 1) / nsplit;
      ScSizes z = sc_sizes(k, G, nv, gs, nsplit, io_per_vec * nv);
      if (z.bytes() <= budget && z.nslot * (long)nv <= (1L << 31)) {
        out = z;
        return true;
      }
      gcap = (gcap > 8) ? gcap * 3 / 4 : gcap - 1;
    }
    if (nv == 1) break;
  }
  return false;
}

size_t sc_budget() {
  const char* e = getenv("PHF_SC_WS_MB");
  const long mb = e ? atol(e) : 4096;
  return (size_t)std::max<long>(mb, 16) << 20;
}

// HIP-event timing of the launches of one phf_sc_apply* call (phf_set_profiling):
// categories 0 stage-1 GEMM, 1 stage-2 GEMM, 2 P GEMM, 3 rank GEMM, 4 point kernel.
struct ScProf {
  hipStream_t s;
  std::vector<hipEvent_t> ev[5];
  hipEvent_t all0 = nullptr, all1 = nullptr;
  template <class F>
  void run(int cat, F&& f) {
    hipEvent_t a, b;
    (void)hipEventCreateWithFlags(&a, kTimingEventFlags);
    (void)hipEventCreateWithFlags(&b, kTimingEventFlags);
    (void)hipEventRecord(a, s);
    f();
    (void)hipEventRecord(b, s);
    ev[cat].push_back(a);
    ev[cat].push_back(b);
  }
  // stage_ms: [0] all device work, [1..5] the categories, [6] the rest, [7] GEMM launches
  void finish(phf_ctx* ctx) {
    (void)hipEventSynchronize(all1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, all0, all1);
    double rest = t;
    int ng = 0;
    for (int c = 0; c < 5; ++c) {
      double sum = 0.0;
      for (size_t i = 0; i + 1 < ev[c].size(); i += 2) {
        float x = 0.f;
        (void)hipEventElapsedTime(&x, ev[c][i], ev[c][i + 1]);
        sum += x;
      }
      if (c < 4) ng += (int)ev[c].size() / 2;
      ctx->stage_ms[1 + c] = sum;
      rest -= sum;
      for (hipEvent_t e : ev[c]) (void)hipEventDestroy(e);
      ev[c].clear();
    }
    ctx->stage_ms[0] = t;
    ctx->stage_ms[6] = rest;
    ctx->stage_ms[7] = ng;
    ctx->nstage = 8;
    (void)hipEventDestroy(all0);
    (void)hipEventDestroy(all1);
  }
};

template <class F>
inline void timed(ScProf* p, int cat, F&& f) {
  if (p)
    p->run(cat, f);
  else
    f();
}

// One pass over nv vectors whose c2a / c1 / c0 are already in the workspace.
void sc_pass(phf_ctx* ctx, const ScBlk& k, const ScSizes& z, double* C2, double* C2P, double* C1, double* c0v,
             double* kbw, double* ka, double* kvec, double* yw, double* pm, double* qm, double* X, double* t2p,
             double* rk, double* t1p, double* t0p, ScProf* prof) {
  hipStream_t s = ctx->stream;
  const int n = ctx->sc_nso, G = ctx->sc_G, o = k.O, nv = z.nv, gc = z.gc;
  const long e1 = k.e12_in, e3 = k.e34_in, eo = k.e12_out, eoo = eo * o;
  const long p12i = k.p12_in(), p12o = k.p12_out(), p34i = k.p34_in(), p34o = k.p34_out();
  const size_t smem = sizeof(double2) * ((size_t)e1 * e3 + (size_t)e1 * o + 8);
  const bool wave1 = eoo <= 16 * 64;
  if (smem > 65536)
    (void)hipFuncSetAttribute(wave1 ? (const void*)sc_point_kernel<64> : (const void*)sc_point_kernel<256>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  for (int c = 0; c < z.nchunk; ++c) {
    const int g0 = c * gc;
    const double beta = c ? 1.0 : 0.0;
    sc_factor_kernel<<<dim3(grid_for(std::max({p34o * p34i, p12o * p12i, e1 * e3}), 256, 1024), gc), 256, 0, s>>>(
        k, n, g0, G, ctx->sc_A, ctx->sc_BT, ctx->sc_Y, ctx->sc_K, ctx->sc_w, ka, kbw, kvec, yw);
    // P((q,r), v, g) = sum_{(p,s)} C2P((q,r), v; (p,s)) Kvec((p,s), g)
    timed(prof, 2, [&] {
      zgemm_batched(s, (int)(e1 * e3 * nv), gc, (int)(e1 * e3), 1, 0, ZMat{C2P, 2, 2 * e1 * e3 * nv, 0, 1, 0},
                    ZMat{kvec, 2, 2 * e1 * e3, 0, 1, 0}, 0, 0, zout_colmajor(pm, e1 * e3 * nv, 0), 1);
    });
    timed(prof, 4, [&] {
      if (wave1)
        sc_point_kernel<64><<<dim3(z.npgrp, nv), 64, smem, s>>>(k, n, nv, g0, z.pg, G, z.nslot, c * z.npgrp, c0v, C1,
                                                                pm, gc, ctx->sc_A, ctx->sc_BT, ctx->sc_Y, ctx->sc_K,
                                                                ctx->sc_w, qm, t0p, t1p);
      else
        sc_point_kernel<256><<<dim3(z.npgrp, nv), 256, smem, s>>>(k, n, nv, g0, z.pg, G, z.nslot, c * z.npgrp, c0v,
                                                                  C1, pm, gc, ctx->sc_A, ctx->sc_BT, ctx->sc_Y,
                                                                  ctx->sc_K, ctx->sc_w, qm, t0p, t1p);
    });
    // stage 1, batched over the points: X_g[(a<b); (v, r<s)] = sum_{p<q} KA_g((a,b),(p,q)) C2[(p,q); (v, r<s)],
    // formed as X_g^T = C2^T KA_g^T (same storage, a<b fastest) so that the
    // stores run along a<b; the products and their order are the same
    timed(prof, 0, [&] {
      zgemm_batched(s, (int)(nv * p34i), (int)p12o, (int)p12i, 1, 0, ZMat{C2, 2 * p12i, 2, 0, 1, 0},
                    ZMat{ka, 2 * p12o, 2, 2 * p12o * p12i, 1, 0}, 0, 0,
                    ZOut{X, 2 * p12o, 2, 2 * p12o * nv * p34i, 1}, gc);
    });
    // stage 2, one batch entry per group of gs points, contraction over (point, r<s):
    //   T2p[j][(a<b, v); (i<j)] (+)= sum_{g in group j, r<s} X_g[(a<b, v); (r<s)] KBW_g((i,j),(r,s))
    timed(prof, 1, [&] {
      zgemm_batched(s, (int)(p12o * nv), (int)p34o, (int)(z.gs * p34i), 1, 0,
                    ZMat{X, 2, 2 * p12o * nv, 2 * p12o * nv * p34i * z.gs, 1, 0},
                    ZMat{kbw, 2 * p34o, 2, 2 * p34o * p34i * z.gs, 1, 0}, beta, 0,
                    zout_colmajor(t2p, p12o * nv, 2 * p12o * nv * p34o), z.nsplit);
    });
    // rank term: RK((a,c),(b,d,v)) (+)= sum_g YW((a,c), g) Qm(g, (b,d,v))
    timed(prof, 3, [&] {
      zgemm_batched(s, (int)eoo, (int)(eoo * nv), gc, 1, 0, ZMat{yw, 2, 2 * eoo, 0, 1, 0},
                    ZMat{qm, 2 * eoo * nv, 2, 0, 1, 0}, beta, 0, zout_colmajor(rk, eoo, 0), 1);
    });
  }
}

std::unique_ptr<ScProf> sc_prof_begin(phf_ctx* ctx) {
  if (!ctx->profiling) return nullptr;
  std::unique_ptr<ScProf> p(new ScProf);
  p->s = ctx->stream;
  (void)hipEventCreateWithFlags(&p->all0, kTimingEventFlags);
  (void)hipEventCreateWithFlags(&p->all1, kTimingEventFlags);
  (void)hipEventRecord(p->all0, ctx->stream);
  return p;
}

void sc_prof_end(phf_ctx* ctx, ScProf* p) {
  if (!p) return;
  (void)hipEventRecord(p->all1, ctx->stream);
  p->finish(ctx);
}

struct ScBufs {
  double *C2, *C2P, *C1, *c0v, *kb, *kaw, *kvec, *yw, *pm, *qm, *out1, *t2p, *rk, *t1p, *t0p, *io;
};

ScBufs carve(char* base, const ScSizes& z) {
  Arena ar{base};
  ScBufs b;
  b.C2 = ar.take<double>(2 * z.c2);
  b.C2P = ar.take<double>(2 * z.c2p);
  b.C1 = ar.take<double>(2 * z.c1);
  b.c0v = ar.take<double>(2 * z.nv);
  b.kb = ar.take<double>(2 * z.kb);
  b.kaw = ar.take<double>(2 * z.kaw);
  b.kvec = ar.take<double>(2 * z.kvec);
  b.yw = ar.take<double>(2 * z.yw);
  b.pm = ar.take<double>(2 * z.pm);
  b.qm = ar.take<double>(2 * z.qm);
  b.out1 = ar.take<double>(2 * z.out1);
  b.t2p = ar.take<double>(2 * z.t2p);
  b.rk = ar.take<double>(2 * z.rk);
  b.t1p = ar.take<double>(2 * z.t1p);
  b.t0p = ar.take<double>(2 * (long)z.nslot * z.nv);
  b.io = ar.take<double>(2 * z.io);
  return b;
}

}  // namespace

void phf_free_sc(phf_ctx* ctx) {
  free_sc_grid(ctx);
  if (ctx->sc_ws) (void)hipFree(ctx->sc_ws);
  ctx->sc_ws = nullptr;
  ctx->sc_ws_size = 0;
}

extern "C" int phf_sc_set_grid(phf_ctx* ctx, int nso, int nocc, const int* ngrid, int qj, const double* r1,
                               int nr1, const double* r2, int nr2, const double* rpg, int npg,
                               const double* weightsp, const double* weightpg, int ncipg, const double* fsp,
                               int ncisp, const double* fpg) {
  if (!ctx) return PHF_ERR_ARG;
  if (!ngrid || !r1 || !r2 || !rpg || !weightsp || !weightpg || !fsp || !fpg)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_grid: NULL pointer argument");
  const int nl = ngrid[0], ny = ngrid[1];
  if (nso < 2 || nso > 64 || nocc < 1 || nocc > nso)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_grid: need 1 <= nocc <= nso <= 64");
  if (nl <= 0 || ny <= 0 || npg <= 0 || ncipg <= 0 || qj < 0)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_grid: non-positive grid or CI dimension");
  if (nl > nr1 || ny > nr2) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_grid: ngrid exceeds R1/R2 extent");
  if (ncisp < 2 * qj + 1) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_grid: fsp shorter than 2 QJ + 1");
  const long G = (long)nl * ny * npg;
  if (G > 65535) return phf_fail(ctx, PHF_ERR_UNSUPPORTED, "phf_sc_set_grid: more than 65535 grid points");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  HIPCK(hipStreamSynchronize(s));
  free_sc_grid(ctx);
  const int n = nso, o = nocc;
  const long n2 = (long)n * n;
  // Host weight per point: weightsp(il,iy) x sum_{m,n,p,q} w with w REAL, so each
  // complex product assigned to it keeps its real part (PHFTools.f90:484,515-520).
  std::vector<double> wpg(npg, 0.0);
  for (int ipg = 0; ipg < npg; ++ipg)
    for (int m = -qj; m <= qj; ++m)
      for (int nn = -qj; nn <= qj; ++nn)
        for (int p = 0; p < ncipg; ++p)
          for (int q = 0; q < ncipg; ++q) {
            double w = weightpg[((long)ipg * ncipg + p) * ncipg + q];   // * wig, wig = One (:515)
            const std::complex<double> fm(fsp[2 * (m + qj)], fsp[2 * (m + qj) + 1]);
            const std::complex<double> fn(fsp[2 * (nn + qj)], fsp[2 * (nn + qj) + 1]);
            w = (w * std::conj(fm) * fn).real();
            const std::complex<double> fp(fpg[2 * p], fpg[2 * p + 1]), fq(fpg[2 * q], fpg[2 * q + 1]);
            w = (w * std::conj(fp) * fq).real();
            wpg[ipg] += w;
          }
  std::vector<double> wh(G);
  for (int il = 0; il < nl; ++il)
    for (int iy = 0; iy < ny; ++iy)
      for (int ipg = 0; ipg < npg; ++ipg)
        wh[((long)il * ny + iy) * npg + ipg] = weightsp[(long)il * ny + iy] * wpg[ipg];

  HIPCK(hipMalloc(&ctx->sc_R, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_A, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_K, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_Y, sizeof(double) * 2 * G * n * o));
  HIPCK(hipMalloc(&ctx->sc_BT, sizeof(double) * 2 * G * n * o));
  HIPCK(hipMalloc(&ctx->sc_w, sizeof(double) * 2 * G));
  HIPCK(hipMalloc(&ctx->sc_ovlp, sizeof(double) * 2));
  // staging: R1, R2, Rpg (row-major host layouts), R1 R2, M, M^-1, det, info, host weights
  const long nst = 2 * n2 * (nl + ny + npg + (long)nl * ny) + 4 * G * o * o + 2 * G + G + G;
  {
    const int rc = ensure_sc_ws(ctx, sizeof(double) * nst + 8 * 256);
    if (rc) return rc;
  }
  Arena ar{ctx->sc_ws};
  double* dR1 = ar.take<double>(2 * n2 * nl);
  double* dR2 = ar.take<double>(2 * n2 * ny);
  double* dRpg = ar.take<double>(2 * n2 * npg);
  double* dR12 = ar.take<double>(2 * n2 * nl * ny);
  double* dM = ar.take<double>(2 * G * o * o);
  double* dMi = ar.take<double>(2 * G * o * o);
  double* dDet = ar.take<double>(2 * G);
  int* dInfo = ar.take<int>(G);
  double* dWh = ar.take<double>(G);
  HIPCK(hipMemcpyAsync(dR1, r1, sizeof(double) * 2 * n2 * nl, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dR2, r2, sizeof(double) * 2 * n2 * ny, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dRpg, rpg, sizeof(double) * 2 * n2 * npg, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dWh, wh.data(), sizeof(double) * G, hipMemcpyHostToDevice, s));
  // R12[il, iy] = R1[il] R2[iy];  R[g] = R12[il, iy] Rpg[ipg]   (PHFTools.f90:500-501)
  {
    ZMat a{dR1, 2L * n, 2, 2 * n2, 1, 0};
    a.bdiv = ny;
    ZMat b{dR2, 2L * n, 2, 0, 1, 0};
    b.bs2 = 2 * n2;
    b.bdiv = ny;
    zgemm_batched(s, n, n, n, 1, 0, a, b, 0, 0, zout_colmajor(dR12, n, 2 * n2), nl * ny);
    ZMat a2{dR12, 2, 2L * n, 2 * n2, 1, 0};
    a2.bdiv = npg;
    ZMat b2{dRpg, 2L * n, 2, 0, 1, 0};
    b2.bs2 = 2 * n2;
    b2.bdiv = npg;
    zgemm_batched(s, n, n, n, 1, 0, a2, b2, 0, 0, zout_colmajor(ctx->sc_R, n, 2 * n2), (int)G);
  }
  // RDM with newMO = the first NOcc unit vectors: M = R[:O, :O]
  sc_overlap_kernel<<<(unsigned)G, 256, 0, s>>>(n, o, ctx->sc_R, dM);
  lu_inverse_batched(s, o, dM, 2L * o * o, dMi, 2L * o * o, dDet, dInfo, (int)G);
  std::vector<int> info(G);
  HIPCK(hipMemcpyAsync(info.data(), dInfo, sizeof(int) * G, hipMemcpyDeviceToHost, s));
  HIPCK(hipStreamSynchronize(s));
  for (long g = 0; g < G; ++g)
    if (info[g] != 0) {
      const int il = (int)(g / ((long)ny * npg)), iy = (int)((g / npg) % ny), ipg = (int)(g % npg);
      free_sc_grid(ctx);
      return phf_fail(ctx, PHF_ERR_SINGULAR,
                      "phf_sc_set_grid: singular overlap at grid point (il=" + std::to_string(il + 1) +
                          ", iy=" + std::to_string(iy + 1) + ", ipg=" + std::to_string(ipg + 1) + "), zero pivot " +
                          std::to_string(info[g]));
    }
  // Y = R[:, :O] M^-1 (N x O);  B^T = R^H Y;  K = B^T R[:O, :];  A = R - Y R[:O, :]
  zgemm_batched(s, n, o, o, 1, 0, zmat_colmajor(ctx->sc_R, n, 2 * n2), zmat_colmajor(dMi, o, 2L * o * o), 0, 0,
                zout_colmajor(ctx->sc_Y, n, 2L * n * o), (int)G);
  zgemm_batched(s, n, o, n, 1, 0, zmat_colmajor_t(ctx->sc_R, n, 2 * n2, 1), zmat_colmajor(ctx->sc_Y, n, 2L * n * o),
                0, 0, zout_colmajor(ctx->sc_BT, n, 2L * n * o), (int)G);
  zgemm_batched(s, n, n, o, 1, 0, zmat_colmajor(ctx->sc_BT, n, 2L * n * o), zmat_colmajor(ctx->sc_R, n, 2 * n2), 0,
                0, zout_colmajor(ctx->sc_K, n, 2 * n2), (int)G);
  HIPCK(hipMemcpyAsync(ctx->sc_A, ctx->sc_R, sizeof(double) * 2 * G * n2, hipMemcpyDeviceToDevice, s));
  zgemm_batched(s, n, n, o, -1, 0, zmat_colmajor(ctx->sc_Y, n, 2L * n * o), zmat_colmajor(ctx->sc_R, n, 2 * n2), 1,
                0, zout_colmajor(ctx->sc_A, n, 2 * n2), (int)G);
  sc_weight_kernel<<<1, 256, 0, s>>>((int)G, dWh, dDet, ctx->sc_w, ctx->sc_ovlp);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(s));
  ctx->sc_G = (int)G;
  ctx->sc_nso = n;
  ctx->sc_nocc = o;
  return PHF_OK;
}

namespace {
// row-major host matrices -> column-major (batch of n x n complex)
__global__ void sc_colmajor_kernel(int n, long total, const double* __restrict__ in, double* __restrict__ out) {
  const long n2 = (long)n * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long g = idx / n2, e = idx % n2;
    const int i = e % n, j = e / n;
    st2(out, idx, ld2(in, g * n2 + (long)i * n + j));
  }
}
}  // namespace

extern "C" int phf_sc_set_points(phf_ctx* ctx, int nso, int npts, const double* r, const double* y, const double* w) {
  if (!ctx) return PHF_ERR_ARG;
  if (!y || nso < 2 || nso > 64 || npts < 1 || npts > 65535)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_set_points: need Y, 2 <= nso <= 64 and 1 <= npts <= 65535");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  HIPCK(hipStreamSynchronize(s));
  free_sc_grid(ctx);
  const int n = nso;
  const long n2 = (long)n * n, G = npts;
  HIPCK(hipMalloc(&ctx->sc_R, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_A, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_K, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_Y, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_BT, sizeof(double) * 2 * G * n2));
  HIPCK(hipMalloc(&ctx->sc_w, sizeof(double) * 2 * G));
  HIPCK(hipMalloc(&ctx->sc_ovlp, sizeof(double) * 2));
  {
    const int rc = ensure_sc_ws(ctx, sizeof(double) * (2 * G * n2 + 3 * G) + 4 * 256);
    if (rc) return rc;
  }
  Arena ar{ctx->sc_ws};
  double* stage = ar.take<double>(2 * G * n2);
  double* dW = ar.take<double>(2 * G);
  double* dOne = ar.take<double>(G);
  std::vector<double> ones(G, 1.0), wv(2 * G);
  for (long g = 0; g < G; ++g) {
    wv[2 * g] = w ? w[2 * g] : 1.0;
    wv[2 * g + 1] = w ? w[2 * g + 1] : 0.0;
  }
  const int gr = (int)std::min<long>((G * n2 + 255) / 256, 16384);
  if (r) {
    HIPCK(hipMemcpyAsync(stage, r, sizeof(double) * 2 * G * n2, hipMemcpyHostToDevice, s));
    sc_colmajor_kernel<<<gr, 256, 0, s>>>(n, G * n2, stage, ctx->sc_R);
  } else {
    std::vector<double> eye(2 * G * n2, 0.0);
    for (long g = 0; g < G; ++g)
      for (int i = 0; i < n; ++i) eye[2 * (g * n2 + i + (long)n * i)] = 1.0;
    HIPCK(hipMemcpyAsync(ctx->sc_R, eye.data(), sizeof(double) * eye.size(), hipMemcpyHostToDevice, s));
    HIPCK(hipStreamSynchronize(s));
  }
  HIPCK(hipStreamSynchronize(s));
  HIPCK(hipMemcpyAsync(stage, y, sizeof(double) * 2 * G * n2, hipMemcpyHostToDevice, s));
  sc_colmajor_kernel<<<gr, 256, 0, s>>>(n, G * n2, stage, ctx->sc_Y);
  HIPCK(hipMemcpyAsync(dW, wv.data(), sizeof(double) * 2 * G, hipMemcpyHostToDevice, s));
  HIPCK(hipMemcpyAsync(dOne, ones.data(), sizeof(double) * G, hipMemcpyHostToDevice, s));
  // B^T = R^H Y;  K = B^T R;  A = R - Y R  (Y has all N columns here)
  zgemm_batched(s, n, n, n, 1, 0, zmat_colmajor_t(ctx->sc_R, n, 2 * n2, 1), zmat_colmajor(ctx->sc_Y, n, 2 * n2), 0, 0,
                zout_colmajor(ctx->sc_BT, n, 2 * n2), (int)G);
  zgemm_batched(s, n, n, n, 1, 0, zmat_colmajor(ctx->sc_BT, n, 2 * n2), zmat_colmajor(ctx->sc_R, n, 2 * n2), 0, 0,
                zout_colmajor(ctx->sc_K, n, 2 * n2), (int)G);
  HIPCK(hipMemcpyAsync(ctx->sc_A, ctx->sc_R, sizeof(double) * 2 * G * n2, hipMemcpyDeviceToDevice, s));
  zgemm_batched(s, n, n, n, -1, 0, zmat_colmajor(ctx->sc_Y, n, 2 * n2), zmat_colmajor(ctx->sc_R, n, 2 * n2), 1, 0,
                zout_colmajor(ctx->sc_A, n, 2 * n2), (int)G);
  sc_weight_kernel<<<1, 256, 0, s>>>((int)G, dOne, dW, ctx->sc_w, ctx->sc_ovlp);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(s));
  ctx->sc_G = (int)G;
  ctx->sc_nso = n;
  ctx->sc_nocc = n;
  return PHF_OK;
}

extern "C" int phf_sc_apply(phf_ctx* ctx, int nvec, const double* c0, const double* c1, const double* c2, double* t0,
                            double* t1, double* t2) {
  if (!ctx) return PHF_ERR_ARG;
  if (!ctx->sc_G) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply: no grid (phf_sc_set_grid)");
  if (nvec < 1 || !c0 || !c1 || !c2 || !t0 || !t1 || !t2)
    return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply: bad vector count or NULL pointer");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int n = ctx->sc_nso, o = ctx->sc_nocc, G = ctx->sc_G;
  const long n2 = (long)n * n, n4 = n2 * n2;
  const ScBlk k{n, 0, n, 0, n, 0, o};
  ScSizes z;
  // per vector in / out staging: c2, c1, c0 and t2, t1, t0
  if (!sc_plan(k, G, nvec, 2 * (n4 + n2 + 1), sc_budget(), z))
    return phf_fail(ctx, PHF_ERR_NOMEM, "phf_sc_apply: workspace budget too small (PHF_SC_WS_MB)");
  {
    const int rc = ensure_sc_ws(ctx, z.bytes());
    if (rc) return rc;
  }
  std::unique_ptr<ScProf> prof = sc_prof_begin(ctx);
  for (int v0 = 0; v0 < nvec; v0 += z.nv) {
    const int nv = std::min(z.nv, nvec - v0);
    ScSizes zz = sc_sizes(k, G, nv, z.gs, z.nsplit, 2 * (n4 + n2 + 1) * nv);
    ScBufs b = carve(ctx->sc_ws, zz);
    double* in2 = b.io;
    double* in1 = in2 + 2 * n4 * nv;
    double* in0 = in1 + 2 * n2 * nv;
    double* o2 = in0 + 2 * nv;
    double* o1 = o2 + 2 * n4 * nv;
    double* o0 = o1 + 2 * n2 * nv;
    HIPCK(hipMemcpyAsync(in2, c2 + 2 * n4 * v0, sizeof(double) * 2 * n4 * nv, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(in1, c1 + 2 * n2 * v0, sizeof(double) * 2 * n2 * nv, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(b.c0v, c0 + 2 * (long)v0, sizeof(double) * 2 * nv, hipMemcpyHostToDevice, s));
    sc_prep_full_kernel<<<grid_for(n4 * nv), 256, 0, s>>>(n, nv, in2, in1, b.C2, b.C2P, b.C1);
    sc_pass(ctx, k, zz, b.C2, b.C2P, b.C1, b.c0v, b.kb, b.kaw, b.kvec, b.yw, b.pm, b.qm, b.out1, b.t2p, b.rk, b.t1p,
            b.t0p, prof.get());
    sc_final_full_kernel<<<grid_for(n4 * nv), 256, 0, s>>>(n, o, nv, zz.nsplit, b.t2p, b.rk, b.t1p, b.t0p, zz.nslot,
                                                           ctx->sc_ovlp, o0, o1, o2);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(t2 + 2 * n4 * v0, o2, sizeof(double) * 2 * n4 * nv, hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(t1 + 2 * n2 * v0, o1, sizeof(double) * 2 * n2 * nv, hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(t0 + 2 * (long)v0, o0, sizeof(double) * 2 * nv, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
  }
  sc_prof_end(ctx, prof.get());
  return PHF_OK;
}

extern "C" long phf_sd_dim(int nso, int nocc) {
  if (nso < 1 || nocc < 0 || nocc > nso) return -1;
  const long v = nso - nocc;
  return 1 + (long)nocc * v + (long)nocc * (nocc - 1) / 2 * (v * (v - 1) / 2);
}

extern "C" int phf_sc_apply_packed_dev(phf_ctx* ctx, int nvec, const double* vin, double* vout) {
  if (!ctx) return PHF_ERR_ARG;
  if (!ctx->sc_G) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply_packed: no grid (phf_sc_set_grid)");
  if (nvec < 1 || !vin || !vout) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply_packed: bad vector count or NULL");
  const int n = ctx->sc_nso, o = ctx->sc_nocc, G = ctx->sc_G, nvir = n - o;
  if (nvir < 1) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply_packed: no virtual orbitals");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const long dim = phf_sd_dim(n, o);
  const ScBlk k{nvir, o, nvir, o, o, 0, o};
  ScSizes z;
  if (!sc_plan(k, G, nvec, 0, sc_budget(), z))
    return phf_fail(ctx, PHF_ERR_NOMEM, "phf_sc_apply_packed: workspace budget too small (PHF_SC_WS_MB)");
  {
    const int rc = ensure_sc_ws(ctx, z.bytes());
    if (rc) return rc;
  }
  std::unique_ptr<ScProf> prof = sc_prof_begin(ctx);
  for (int v0 = 0; v0 < nvec; v0 += z.nv) {
    const int nv = std::min(z.nv, nvec - v0);
    ScSizes zz = sc_sizes(k, G, nv, z.gs, z.nsplit, 0);
    ScBufs b = carve(ctx->sc_ws, zz);
    const double* vi = vin + 2 * dim * v0;
    sc_prep_packed_kernel<<<grid_for((long)nvir * nvir * o * o * nv), 256, 0, s>>>(o, nvir, nv, dim, vi, b.C2, b.C2P,
                                                                                   b.C1, b.c0v);
    sc_pass(ctx, k, zz, b.C2, b.C2P, b.C1, b.c0v, b.kb, b.kaw, b.kvec, b.yw, b.pm, b.qm, b.out1, b.t2p, b.rk, b.t1p,
            b.t0p, prof.get());
    sc_final_packed_kernel<<<grid_for(dim * nv), 256, 0, s>>>(o, nvir, nv, dim, zz.nsplit, b.t2p, b.rk, b.t1p, b.t0p,
                                                              zz.nslot, ctx->sc_ovlp, vout + 2 * dim * v0);
    HIPCK(hipGetLastError());
  }
  sc_prof_end(ctx, prof.get());
  return PHF_OK;
}

extern "C" int phf_sc_apply_packed(phf_ctx* ctx, int nvec, const double* vin, double* vout) {
  if (!ctx) return PHF_ERR_ARG;
  if (!ctx->sc_G) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply_packed: no grid (phf_sc_set_grid)");
  if (nvec < 1 || !vin || !vout) return phf_fail(ctx, PHF_ERR_ARG, "phf_sc_apply_packed: bad vector count or NULL");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const long dim = phf_sd_dim(ctx->sc_nso, ctx->sc_nocc);
  const size_t bytes = sizeof(double) * 2 * dim * nvec;
  double *di = nullptr, *dout = nullptr;
  HIPCK(hipMalloc(&di, bytes));
  if (hipMalloc(&dout, bytes) != hipSuccess) {
    (void)hipFree(di);
    return phf_fail(ctx, PHF_ERR_NOMEM, "phf_sc_apply_packed: device allocation failed");
  }
  int rc = hipMemcpyAsync(di, vin, bytes, hipMemcpyHostToDevice, s) == hipSuccess ? PHF_OK : PHF_ERR_HIP;
  if (rc == PHF_OK) rc = phf_sc_apply_packed_dev(ctx, nvec, di, dout);
  if (rc == PHF_OK && hipMemcpyAsync(vout, dout, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) rc = PHF_ERR_HIP;
  if (rc == PHF_OK && hipStreamSynchronize(s) != hipSuccess) rc = PHF_ERR_HIP;
  (void)hipStreamSynchronize(s);
  (void)hipFree(di);
  (void)hipFree(dout);
  if (rc == PHF_ERR_HIP) return phf_fail(ctx, rc, "phf_sc_apply_packed: HIP copy failed");
  return rc;
}

extern "C" int phf_buildsc(phf_ctx* ctx, const double* c0, const double* c1, const double* c2, int nocc, int nso,
                           const int* ngrid, int qj, const double* r1, int nr1, const double* r2, int nr2,
                           const double* rpg, int npg, const double* weightsp, const double* weightpg, int ncipg,
                           const double* fsp, int ncisp, const double* fpg, double* t0, double* t1, double* t2) {
  int rc = phf_sc_set_grid(ctx, nso, nocc, ngrid, qj, r1, nr1, r2, nr2, rpg, npg, weightsp, weightpg, ncipg, fsp,
                           ncisp, fpg);
  if (rc) return rc;
  return phf_sc_apply(ctx, 1, c0, c1, c2, t0, t1, t2);
}
