// Map-domain helpers of the EB-leakage cleaning (fsht_fix_eb): masking and
// the per-map linear fit of a B map against its leakage template.
//
// Replaces apply_mask_mini / copy_with_mask_mini / lin_resi_mini
// (src/sht_core.f90:209-258), which loop over maps on the host.
#include <algorithm>

#include "fsht_internal.h"

namespace fsht {

// maps[k*npix + p] *= mask[p] for every map k (maps are (npix, nmaps), map-major)
__global__ __launch_bounds__(256) void k_mask_maps(double* __restrict__ maps, const double* __restrict__ mask,
                                                   long long npix, int nmaps) {
  const long long n = npix * nmaps;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const long long p = i % npix;
    maps[i] *= mask[p];
  }
}

// p[0..n) = 0 with a kernel (a memset node recorded into a captured HIP graph
// is not reliable on this platform; a kernel is)
__global__ __launch_bounds__(256) void k_zero(double* __restrict__ p, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0.0;
}

hipError_t launch_zero(double* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long nb = std::min<long long>((n + 255) / 256, 256LL * 64);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_mask_maps(double* maps, const double* mask, long long npix, int nmaps, hipStream_t s) {
  const long long n = npix * nmaps;
  const long long nb = std::min<long long>((n + 255) / 256, 256LL * 64);
  hipLaunchKernelGGL(k_mask_maps, dim3((unsigned)nb), dim3(256), 0, s, maps, mask, npix, nmaps);
  return hipGetLastError();
}

// Block-wide sum of one value per thread (1024 threads, 16 waves).
__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();  // red may still be read by the previous call
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  return s;
}

// One workgroup per map: over the pixels with mask == 1, fit y = a0 + a1*x by
// least squares (two passes: means, then centred moments, as the reference
// does) and write resi = y - a0 - a1*x there, 0 elsewhere.  coef[2k] = a0,
// coef[2k+1] = a1 of map k.
__global__ __launch_bounds__(1024) void k_linfit_resi(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ mask, double* __restrict__ resi,
                                                      double* __restrict__ coef, long long npix) {
  __shared__ double red[16];
  const long long off = (long long)blockIdx.x * npix;
  const double* xs = x + off;
  const double* ys = y + off;
  double sx = 0.0, sy = 0.0, cnt = 0.0;
  for (long long p = threadIdx.x; p < npix; p += blockDim.x)
    if (mask[p] == 1.0) {
      sx += xs[p];
      sy += ys[p];
      cnt += 1.0;
    }
  sx = block_sum(sx, red);
  sy = block_sum(sy, red);
  cnt = block_sum(cnt, red);
  const double mx = sx / cnt, my = sy / cnt;
  double cxx = 0.0, cxy = 0.0;
  for (long long p = threadIdx.x; p < npix; p += blockDim.x)
    if (mask[p] == 1.0) {
      const double dx = xs[p] - mx;
      cxx += dx * dx;
      cxy += (ys[p] - my) * dx;
    }
  cxx = block_sum(cxx, red);
  cxy = block_sum(cxy, red);
  const double a1 = cxy / cxx;
  const double a0 = my - mx * a1;
  for (long long p = threadIdx.x; p < npix; p += blockDim.x)
    resi[off + p] = (mask[p] == 1.0) ? ys[p] - a0 - a1 * xs[p] : 0.0;
  if (threadIdx.x == 0 && coef) {
    coef[2 * blockIdx.x] = a0;
    coef[2 * blockIdx.x + 1] = a1;
  }
}

hipError_t launch_linfit_resi(const double* x, const double* y, const double* mask, double* resi, double* coef,
                              long long npix, int nmaps, hipStream_t s) {
  hipLaunchKernelGGL(k_linfit_resi, dim3(nmaps), dim3(1024), 0, s, x, y, mask, resi, coef, npix);
  return hipGetLastError();
}

}  // namespace fsht
