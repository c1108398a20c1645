/*
 * ORACLE -- test infrastructure only (CPU baseline and cross-check).
 *
 * Plain C + OpenMP restatement of BuildHSG (PHF_fort/PHFTools.f90:10-201) for
 * CmplxConj in {0, 1}: per grid point the same algorithm as the reference --
 * R = Ref^H R1 R2 Rpg Ref (ao2mo2, :88-90, :390-398), RDM with an LU inverse
 * and determinant (:249-263, Wrappers.f90:195-231, 295-335), the dense NSO^4
 * MakeGamma (:310-337), Makeh/Kernels (:281-289, :339-348), Gradient
 * (:291-308) and the Wigner-weighted accumulation (:120-166) -- with OpenMP
 * over the Lebedev index as the reference's !$omp do (:79-84).
 *
 * Arrays are C order (as the HIP library takes them): complex arrays are
 * interleaved double pairs.  Used by bench.py's cpu_baseline leg and by the
 * oracle's self-consistency test; never by the product.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

static void matmul(int n, int m, int k, const cplx* a, int a_h, const cplx* b, int b_h, cplx* c) {
  /* c[n x m] = op(a)[n x k] op(b)[k x m]; a_h/b_h: use conjugate transpose of the stored row-major matrix */
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      cplx acc = 0;
      for (int q = 0; q < k; ++q) {
        const cplx av = a_h ? conj(a[q * n + i]) : a[i * k + q];
        const cplx bv = b_h ? conj(b[j * k + q]) : b[q * m + j];
        acc += av * bv;
      }
      c[i * m + j] = acc;
    }
}

/* ZGETRF-style LU with row pivoting on max |Re|+|Im|; returns info (0 ok) */
static int lu(int n, cplx* a, int* piv) {
  int info = 0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = -1;
    for (int i = k; i < n; ++i) {
      const double v = fabs(creal(a[i * n + k])) + fabs(cimag(a[i * n + k]));
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (a[p * n + k] != 0) {
      if (p != k)
        for (int j = 0; j < n; ++j) { cplx t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
      const cplx inv = 1.0 / a[k * n + k];
      for (int i = k + 1; i < n; ++i) a[i * n + k] *= inv;
    } else if (!info) {
      info = k + 1;
    }
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= a[i * n + k] * a[k * n + j];
  }
  return info;
}

static int inverse_det(int n, const cplx* m, cplx* minv, cplx* det, cplx* work, int* piv) {
  memcpy(work, m, sizeof(cplx) * n * n);
  const int info = lu(n, work, piv);
  if (info) return info;
  cplx d = 1;
  for (int i = 0; i < n; ++i) d = (piv[i] == i) ? d * work[i * n + i] : -d * work[i * n + i];
  *det = d;
  /* columns of the inverse: solve L U x = P e_j */
  for (int j = 0; j < n; ++j) {
    cplx* x = minv + j;  /* strided column j of row-major minv */
    for (int i = 0; i < n; ++i) x[i * n] = (i == j) ? 1 : 0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) { cplx t = x[k * n]; x[k * n] = x[piv[k] * n]; x[piv[k] * n] = t; }
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < i; ++q) x[i * n] -= work[i * n + q] * x[q * n];
    for (int i = n - 1; i >= 0; --i) {
      for (int q = i + 1; q < n; ++q) x[i * n] -= work[i * n + q] * x[q * n];
      x[i * n] /= work[i * n + i];
    }
  }
  return 0;
}

static double wigner_d(int j, int n, int m, double beta) {
  const int smax = (j + m < j - n) ? j + m : j - n, smin = (m - n > 0) ? m - n : 0;
  double fac = sqrt(tgamma(j + n + 1.0) * tgamma(j - n + 1.0) * tgamma(j + m + 1.0) * tgamma(j - m + 1.0));
  double d = 0;
  for (int s = smin; s <= smax; ++s) {
    const int p1 = n - m + s, p2 = 2 * j + m - n - 2 * s, p3 = n - m + 2 * s;
    const double num = ((p1 % 2) ? -1.0 : 1.0) * pow(cos(beta / 2), p2) * pow(sin(beta / 2), p3);
    const double den = tgamma(j + m - s + 1.0) * tgamma(s + 1.0) * tgamma(n - m + s + 1.0) * tgamma(j - n - s + 1.0);
    d += num / den;
  }
  return d * fac;
}

/* returns 0, or 1 + flat index of the first grid point with a singular overlap */
int cref_buildhsg(int nso, const double* hone_, const double* htwo_, const double* ref_, const double* z_,
                  int nl, int ny, int npg, int nci, int cmplxconj, int ncisp, int nocc, const double* r1_,
                  const double* r2_, const double* rpg_, const double* weightsp, const double* weightpg, int ncipg,
                  const double* roota, const double* rootb, const double* rooty, int qj, int sp,
                  const int* points, int npoints, double* hmat_, double* smat_, double* grad_, double* rdm_) {
  const cplx* hone = (const cplx*)hone_;
  const cplx* htwo = (const cplx*)htwo_;
  const cplx* ref = (const cplx*)ref_;
  const cplx* z = (const cplx*)z_;
  const cplx* r1 = (const cplx*)r1_;
  const cplx* r2 = (const cplx*)r2_;
  const cplx* rpg = (const cplx*)rpg_;
  cplx* hmat = (cplx*)hmat_;
  cplx* smat = (cplx*)smat_;
  cplx* grad = (cplx*)grad_;
  cplx* rdm = (cplx*)rdm_;
  const long n2 = (long)nso * nso;
  const int N = nocc, nmn = nci * nci;
  memset(hmat, 0, sizeof(cplx) * nmn);
  memset(smat, 0, sizeof(cplx) * nmn);
  memset(grad, 0, sizeof(cplx) * nmn * n2);
  memset(rdm, 0, sizeof(cplx) * nmn * n2);
  cplx* mo = calloc((size_t)nso * N, sizeof(cplx));
  for (int a = 0; a < N; ++a) mo[a * N + a] = 1;
  for (int i = N; i < nso; ++i)
    for (int a = 0; a < N; ++a)
      mo[i * N + a] = (cmplxconj == 0) ? creal(z[(i - N) * N + a]) : z[(i - N) * N + a];
  int err = 0;
  /* work list: all (il, iy, ipg) or the given sample */
  const int total = npoints > 0 ? npoints : nl * ny * npg;

#pragma omp parallel
  {
    cplx *t1 = malloc(sizeof(cplx) * n2), *t2 = malloc(sizeof(cplx) * n2), *rr = malloc(sizeof(cplx) * n2);
    cplx *tmp = malloc(sizeof(cplx) * nso * N), *tmp2 = malloc(sizeof(cplx) * nso * N);
    cplx *m = malloc(sizeof(cplx) * N * N), *minv = malloc(sizeof(cplx) * N * N), *work = malloc(sizeof(cplx) * N * N);
    int* piv = malloc(sizeof(int) * N);
    cplx *rho = malloc(sizeof(cplx) * n2), *gam = malloc(sizeof(cplx) * n2), *g = malloc(sizeof(cplx) * n2);
    cplx *f = malloc(sizeof(cplx) * n2), *lhs = malloc(sizeof(cplx) * n2);
    cplx *lh = calloc(nmn, sizeof(cplx)), *ls = calloc(nmn, sizeof(cplx));
    cplx *lg = calloc((size_t)nmn * n2, sizeof(cplx)), *lr = calloc((size_t)nmn * n2, sizeof(cplx));
#pragma omp for schedule(static)
    for (int w = 0; w < total; ++w) {
      int il, iy, ipg;
      if (npoints > 0) { il = points[3 * w]; iy = points[3 * w + 1]; ipg = points[3 * w + 2]; }
      else { il = w / (ny * npg); iy = (w / npg) % ny; ipg = w % npg; }
      /* R = Ref^H (R1 R2 Rpg) Ref */
      matmul(nso, nso, nso, r1 + il * n2, 0, r2 + iy * n2, 0, t1);
      matmul(nso, nso, nso, t1, 0, rpg + ipg * n2, 0, t2);
      matmul(nso, nso, nso, t2, 0, ref, 0, t1);
      matmul(nso, nso, nso, ref, 1, t1, 0, rr);
      /* RDM */
      matmul(nso, N, nso, rr, 0, mo, 0, tmp);
      matmul(N, N, nso, mo, 1, tmp, 0, m);
      cplx det;
      if (inverse_det(N, m, minv, &det, work, piv)) {
#pragma omp critical
        if (!err) err = 1 + w;
        continue;
      }
      matmul(nso, N, N, tmp, 0, minv, 0, tmp2);
      for (int p = 0; p < nso; ++p)
        for (int q = 0; q < nso; ++q) {
          cplx acc = 0;
          for (int a = 0; a < N; ++a) acc += tmp2[p * N + a] * conj(mo[q * N + a]);
          rho[p * nso + q] = acc;
        }
      /* MakeGamma: gam(i,k) = sum_{j,l} V(i,j,k,l) rho(l,j) */
      for (int i = 0; i < nso; ++i)
        for (int k = 0; k < nso; ++k) {
          cplx acc = 0;
          for (int j = 0; j < nso; ++j) {
            const cplx* v = htwo + (((long)i * nso + j) * nso + k) * nso;
            for (int l = 0; l < nso; ++l) acc += v[l] * rho[l * nso + j];
          }
          gam[i * nso + k] = acc;
        }
      /* Makeh */
      cplx e = 0;
      for (int i = 0; i < nso; ++i)
        for (int j = 0; j < nso; ++j) e += hone[j * nso + i] * rho[i * nso + j] + 0.5 * gam[i * nso + j] * rho[j * nso + i];
      /* Gradient: G = det [ (I - rho)(h + gam) rho + e rho ] */
      for (long q = 0; q < n2; ++q) f[q] = hone[q] + gam[q];
      matmul(nso, nso, nso, f, 0, rho, 0, lhs);
      for (long q = 0; q < n2; ++q) t1[q] = -rho[q];
      for (int i = 0; i < nso; ++i) t1[i * nso + i] += 1;
      matmul(nso, nso, nso, t1, 0, lhs, 0, g);
      const double wsp = weightsp[il * ny + iy];
      for (long q = 0; q < n2; ++q) g[q] = (g[q] + e * rho[q]) * det * wsp;
      const cplx h0 = wsp * e * det, s0 = wsp * det;
      for (int ni = -qj; ni <= qj; ++ni)
        for (int nj = -qj; nj <= qj; ++nj) {
          double a, b, y;
          if (sp == 2) { a = roota[il]; b = rootb[iy]; y = rooty[il]; }
          else { a = roota[il]; b = rootb[il]; y = rooty[iy]; }
          const cplx wig = (qj == 0) ? 1 : cexp(-I * ni * a) * wigner_d(qj, ni, nj, b) * cexp(-I * nj * y);
          for (int p = 0; p < ncipg; ++p)
            for (int qq = 0; qq < ncipg; ++qq) {
              const cplx ww = weightpg[(ipg * ncipg + p) * ncipg + qq] * wig;
              const int mi = ni + qj + p * ncisp, nj2 = nj + qj + qq * ncisp, mn = mi * nci + nj2;
              lh[mn] += h0 * ww;
              ls[mn] += s0 * ww;
              for (long q = 0; q < n2; ++q) {
                lg[mn * n2 + q] += g[q] * ww;
                lr[mn * n2 + q] += rho[q] * wsp * det * ww;
              }
            }
        }
    }
#pragma omp critical
    {
      for (int q = 0; q < nmn; ++q) { hmat[q] += lh[q]; smat[q] += ls[q]; }
      for (long q = 0; q < (long)nmn * n2; ++q) { grad[q] += lg[q]; rdm[q] += lr[q]; }
    }
    free(t1); free(t2); free(rr); free(tmp); free(tmp2); free(m); free(minv); free(work); free(piv);
    free(rho); free(gam); free(g); free(f); free(lhs); free(lh); free(ls); free(lg); free(lr);
  }
  free(mo);
  return err;
}
