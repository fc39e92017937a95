// Elementwise kernels over the square-packed alm layout
//   Re a_lm at alm[sim, l, m] (l >= m), Im a_lm at alm[sim, l-m, lmax+1-m] (m >= 1),
// stored Fortran-order (sim fastest).
//
// Replaces norm_alm1 (src/norm_alm.f90:14-33), filt_alm1 (src/filt_alm.f90:16-36)
// and convert_alm_healpy1 / convert_alm_healpy_inv1 (src/convert_alm.f90:34-107).
// Every (row, column) of the square matrix is exactly one real degree of
// freedom, so each kernel maps one thread to one (sim, row, col) entry and
// recovers (l, m, re/im) from (row, col).
#include "fsht_internal.h"

namespace fsht {

// (row, col) of the square matrix -> (l, m, is_imag).
__device__ inline void decode_rc(int row, int col, int lmax, int* l, int* m, int* im) {
  if (row >= col) {  // Re of (l=row, m=col)
    *l = row;
    *m = col;
    *im = 0;
  } else {  // Im of m = lmax+1-col at l = row + m
    *m = lmax + 1 - col;
    *l = row + *m;
    *im = 1;
  }
}

// out = in * fac(l) (* m0 for m=0) + shift(l)
__global__ void k_norm_alm(const double* __restrict__ in, double* __restrict__ out, const double* __restrict__ fac,
                           const double* __restrict__ shift, double m0, int lmax, int nsim) {
  const long long L = lmax + 1;
  const long long total = (long long)nsim * L * L;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long rc = e / nsim;
    const int row = (int)(rc % L), col = (int)(rc / L);
    int l, m, im;
    decode_rc(row, col, lmax, &l, &m, &im);
    double f = fac[l];
    if (m == 0) f = f * m0;
    out[e] = in[e] * f + shift[l];
  }
}

// alm *= win(l)
__global__ void k_filt_alm(double* __restrict__ alm, const double* __restrict__ win, int lmax, int nsim) {
  const long long L = lmax + 1;
  const long long total = (long long)nsim * L * L;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long rc = e / nsim;
    const int row = (int)(rc % L), col = (int)(rc / L);
    int l, m, im;
    decode_rc(row, col, lmax, &l, &m, &im);
    alm[e] = alm[e] * win[l];
  }
}

// Square packing -> healpy order: out(2, nalm, nsim) Fortran order, entry
// (c, idx, sim) at c + 2*(idx + nalm*sim), idx = m*(2*lmax+1-m)/2 + l.
// inverse = 1 goes the other way.
__global__ void k_convert_healpy(const double* __restrict__ src, double* __restrict__ dst, int lmax, int nsim,
                                 int inverse) {
  const long long L = lmax + 1;
  const long long nalm = L * (L + 1) / 2;
  const long long total = nalm * nsim;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long idx = e % nalm;
    const int sim = (int)(e / nalm);
    // m from idx: idx = m*(2L-1-m)/2 + l - ... find m by search on the row starts
    int lo = 0, hi = lmax;
    while (lo < hi) {  // largest m with start(m) <= idx
      const int mid = (lo + hi + 1) / 2;
      const long long st = (long long)mid * (2 * L - mid + 1) / 2;
      if (st <= idx)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int m = lo;
    const long long st = (long long)m * (2 * L - m + 1) / 2;
    const int l = (int)(idx - st) + m;
    const long long re_at = sim + (long long)nsim * (l + L * m);
    const long long im_at = sim + (long long)nsim * ((l - m) + L * (L - m));
    const long long h = 2 * (idx + nalm * sim);
    if (!inverse) {
      dst[h] = src[re_at];
      dst[h + 1] = (m > 0) ? src[im_at] : 0.0;
    } else {
      dst[re_at] = src[h];
      if (m > 0) dst[im_at] = src[h + 1];
    }
  }
}

static dim3 grid_for(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65535LL * 16) b = 65535LL * 16;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

hipError_t launch_norm_alm(const double* in, double* out, const double* fac, const double* shift, double m0,
                           int lmax, int nsim, hipStream_t s) {
  const long long total = (long long)nsim * (lmax + 1) * (lmax + 1);
  hipLaunchKernelGGL(k_norm_alm, grid_for(total), dim3(256), 0, s, in, out, fac, shift, m0, lmax, nsim);
  return hipGetLastError();
}

hipError_t launch_filt_alm(double* alm, const double* win, int lmax, int nsim, hipStream_t s) {
  const long long total = (long long)nsim * (lmax + 1) * (lmax + 1);
  hipLaunchKernelGGL(k_filt_alm, grid_for(total), dim3(256), 0, s, alm, win, lmax, nsim);
  return hipGetLastError();
}

hipError_t launch_convert_healpy(const double* src, double* dst, int lmax, int nsim, int inverse, hipStream_t s) {
  const long long total = (long long)(lmax + 1) * (lmax + 2) / 2 * nsim;
  hipLaunchKernelGGL(k_convert_healpy, grid_for(total), dim3(256), 0, s, src, dst, lmax, nsim, inverse);
  return hipGetLastError();
}

}  // namespace fsht
