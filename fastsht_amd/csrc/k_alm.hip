// Elementwise kernels over the square-packed alm layout
//   Re a_lm at alm[sim, l, m] (l >= m), Im a_lm at alm[sim, l-m, lmax+1-m] (m >= 1),
// stored Fortran-order (sim fastest).
//
// Replaces norm_alm1 (src/norm_alm.f90:14-33), filt_alm1 (src/filt_alm.f90:16-36)
// and convert_alm_healpy1 / convert_alm_healpy_inv1 (src/convert_alm.f90:34-107).
// Every (row, column) of the square matrix is exactly one real degree of
// freedom, so each kernel maps one thread to one (sim, row, col) entry and
// recovers (l, m, re/im) from (row, col).
#include <cstdint>

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

// ---------------------------------------------------------------------------
// Angular power spectra: cl(k, l, c) += sum_m fac_m * x_lm y_lm / (2l+1) over
// Re and Im parts, fac_0 = 1, fac_m = 2, for l >= lmin; components TT only,
// or TT EE BB TE TB EB.  cl is (nsim, lmax+1, ncl), Fortran order.
// Replaces alm2cl_T1 / alm2cl_TEB1 (src/alm2cl_TEB.f90:22-110), keeping their
// summation order (m ascending, the Re term then the Im term) and their
// accumulation into the caller's array.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_alm2cl(const double* __restrict__ aT, const double* __restrict__ aE,
                                                const double* __restrict__ aB, double* __restrict__ cl, int lmin,
                                                int lmax, int nsim, int ncl) {
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  const long long L = lmax + 1;
  if (t >= (long long)nsim * L) return;
  const int k = (int)(t % nsim);
  const int l = (int)(t / nsim);
  if (l < lmin) return;
  const double d = 2.0 * l + 1.0;
  double s[6];
  for (int c = 0; c < 6; ++c) s[c] = (c < ncl) ? cl[(long long)k + nsim * (l + L * c)] : 0.0;
  for (int m = 0; m <= l; ++m) {
    const double fac = (m == 0) ? 1.0 : 2.0;
    for (int part = 0; part < (m > 0 ? 2 : 1); ++part) {
      const long long idx = part ? (long long)k + nsim * ((l - m) + L * (L - m)) : (long long)k + nsim * (l + L * m);
      const double t0 = aT[idx];
      if (ncl == 1) {
        s[0] = s[0] + t0 * t0 * fac / d;
      } else {
        const double e0 = aE[idx], b0 = aB[idx];
        s[0] = s[0] + t0 * t0 * fac / d;
        s[1] = s[1] + e0 * e0 * fac / d;
        s[2] = s[2] + b0 * b0 * fac / d;
        s[3] = s[3] + t0 * e0 * fac / d;
        s[4] = s[4] + t0 * b0 * fac / d;
        s[5] = s[5] + e0 * b0 * fac / d;
      }
    }
  }
  for (int c = 0; c < ncl; ++c) cl[(long long)k + nsim * (l + L * c)] = s[c];
}

hipError_t launch_alm2cl(const double* aT, const double* aE, const double* aB, double* cl, int lmin, int lmax,
                         int nsim, int ncl, hipStream_t s) {
  const long long n = (long long)nsim * (lmax + 1);
  hipLaunchKernelGGL(k_alm2cl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, aT, aE, aB, cl, lmin, lmax, nsim,
                     ncl);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded Gaussian alm with a given (cross) spectrum.  Replaces cl2alm_T1 /
// cl2alm_TEB1 (src/cl2alm_TEB.f90:25-170), which draw from MKL VSL
// (MCG31, inverse-CDF Gaussians); MKL's stream is not reproduced.  Here the
// standard normal of storage element (sim k, row r, column c) of component
// comp is a function of (seed, k, r, c, comp) alone: Philox4x32-10 (Salmon et
// al., SC 2011) on the counter (k, r, c, comp) with key (seed lo, seed hi),
// then Box-Muller on two 53-bit uniforms.  The spectrum is imposed as in the
// reference: variance 1 at m = 0 and 1/2 per part at m > 0, then per l
//   B = c T + v E + w B,  E = b T + u E,  T = a T
// with a = sqrt(TT), b = TE/a, c = TB/a, u = sqrt(EE - b^2),
// v = (TT EB - TE TB)/(u TT), w = sqrt(BB - TB (EE TB - TE EB)/(TT u^2) - v EB/u),
// and every element with l < lmin is 0.
// ---------------------------------------------------------------------------
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ inline double std_normal(unsigned long long seed, int k, int r, int col, int comp) {
  uint32_t c[4] = {(uint32_t)k, (uint32_t)r, (uint32_t)col, (uint32_t)comp};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double two53 = 9007199254740992.0;
  const double u1 = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6) + 1.0) / two53;  // (0, 1]
  const double u2 = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) / two53;        // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// one thread per (sim, storage element); ncomp = 1 (T) or 3 (T, E, B).
// cl is (lmax+1) for T, or (lmax+1, 6) Fortran order: TT EE BB TE TB EB.
__global__ __launch_bounds__(256) void k_cl2alm(const double* __restrict__ cl, double* __restrict__ aT,
                                                double* __restrict__ aE, double* __restrict__ aB, int lmin, int lmax,
                                                int nsim, int ncomp, unsigned long long seed) {
  const long long L = lmax + 1;
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)nsim * L * L) return;
  const int k = (int)(t % nsim);
  const long long rc = t / nsim;
  const int r = (int)(rc % L), col = (int)(rc / L);
  // storage (r, col) -> (l, m): Re of m = col at l = r >= col, else Im of m = L - col at l = r + m
  const bool re = r >= col;
  const int m = re ? col : (int)(L - col);
  const int l = re ? r : r + m;
  const long long idx = (long long)k + nsim * (r + L * col);
  const double fac = (col == 0) ? 1.0 : 0.7071067811865476;  // sqrt(2)/2
  if (l < lmin) {
    aT[idx] = 0.0;
    if (ncomp == 3) aE[idx] = aB[idx] = 0.0;
    return;
  }
  const double gT = std_normal(seed, k, r, col, 0) * fac;
  if (ncomp == 1) {
    aT[idx] = gT * sqrt(cl[l]);
    return;
  }
  const double gE = std_normal(seed, k, r, col, 1) * fac;
  const double gB = std_normal(seed, k, r, col, 2) * fac;
  const double TT = cl[l], EE = cl[l + L], BB = cl[l + 2 * L], TE = cl[l + 3 * L], TB = cl[l + 4 * L],
               EB = cl[l + 5 * L];
  const double a = sqrt(TT);
  const double b = TE / a;
  const double c = TB / a;
  const double u = sqrt(EE - b * b);
  const double v = (TT * EB - TE * TB) / (u * TT);
  const double w = sqrt(BB - TB / TT / (u * u) * (EE * TB - TE * EB) - v * EB / u);
  aB[idx] = gT * c + gE * v + gB * w;
  aE[idx] = gT * b + gE * u;
  aT[idx] = gT * a;
}

hipError_t launch_cl2alm(const double* cl, double* aT, double* aE, double* aB, int lmin, int lmax, int nsim, int ncomp,
                         unsigned long long seed, hipStream_t s) {
  const long long n = (long long)nsim * (lmax + 1) * (lmax + 1);
  hipLaunchKernelGGL(k_cl2alm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cl, aT, aE, aB, lmin, lmax, nsim,
                     ncomp, seed);
  return hipGetLastError();
}

}  // namespace fsht
