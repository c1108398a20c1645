// Batched complex FP64 inverse + determinant of the N x N overlaps M_g = C0^H T_g.
//
// Restates InvertC / determinantC (Wrappers.f90:195-231, 295-335): partial
// (row) pivoting on the largest |Re|+|Im| in the column (IZAMAX's measure,
// first index on ties), det = prod of pivots with a sign flip for every
// PIV(k) != k, info = first exactly-zero pivot.  One workgroup per grid
// point; the matrix lives in LDS and is inverted in place by Gauss-Jordan
// elimination (its pivots are the U_kk of the LU factorization), then the
// row interchanges are undone as column interchanges in reverse order.
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

}  // namespace

int lu_inverse_max_n() { return 96; }

// M: column-major N x N complex matrices with batch stride mbs (doubles).
void lu_inverse_batched(hipStream_t s, int n, const double* M, long mbs, double* Minv, long ibs,
                        double* det, int* info, int batch) {
  if (n <= 32)
    lu_inverse_kernel<32><<<batch, 256, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
  else if (n <= 64)
    lu_inverse_kernel<64><<<batch, 256, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
  else
    lu_inverse_kernel<96><<<batch, 256, 0, s>>>(n, M, mbs, Minv, ibs, det, info);
}
