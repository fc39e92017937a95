// m-major synthesis output -> ring spectra (PERM packing), with aliasing.
//
// Replaces mat2fft_fast and mat2fft_fast_iter (src/fft_mapping.f90:84-190)
// plus the zeroing of unused PERM slots (src/sht_main.f90:137-146).  The
// reference scatters one m at a time and relies on ascending m with
// set-then-accumulate to fold the orders that alias onto the same ring
// frequency (rings shorter than 2(lmax+1)).  Here every output slot gathers
// its aliases explicitly: slot k of a ring of length n receives the orders
// m = k, n-k, n+k, 2n-k, ... (m <= lmax, ring kept for m), rotated by
// e^{+i m phi0}, conjugated for negative frequencies and doubled on the DC and
// Nyquist slots for m > 0 (src/sht_data_init.f90:542-573).
//
// residual = 0: out = X            (alm2t / eb2qu: spectrum of the synthesis)
// residual = 1: out = Xt - n * X   (Jacobi residual in the ring-Fourier domain,
//                                   src/fft_mapping.f90:177-183)
#include "fsht_internal.h"

namespace fsht {

// Thread blocks of the gather / fold kernels: x = 16 spectrum columns (one
// 128-byte line of a pixel-major row; nsimp is a multiple of 16, so column
// blocks are never partly empty), y = rings (gathers) or slots (fold).
#ifndef FSHT_FOLD_BX
#define FSHT_FOLD_BX 16
#endif
struct FoldBlk {
  int x, y;
};
constexpr FoldBlk kFoldBlk{FSHT_FOLD_BX, 256 / FSHT_FOLD_BX};
#ifndef FSHT_GATHER_RINGS
#define FSHT_GATHER_RINGS 32
#endif
constexpr int kGatherRings = FSHT_GATHER_RINGS;  // rings per gather workgroup


template <int NCOMP>
__global__ __launch_bounds__(256) void k_fold(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                              const double2* __restrict__ tw, const double* __restrict__ G,
                                              const long long* __restrict__ goff, const double* __restrict__ Xt,
                                              double* __restrict__ out, int residual, int lmax, int nhr,
                                              int nnorth, int nsim, int nsimp) {
  if (!cols_active(blockIdx.z * kFoldBlk.x, (blockIdx.z + 1) * kFoldBlk.x, nsim, nsimp)) return;  // no map here
  const int i = blockIdx.x;  // north ring
  const int kk = blockIdx.y * kFoldBlk.y + threadIdx.y;
  const int col = blockIdx.z * kFoldBlk.x + threadIdx.x;  // spectrum column: comp*nsimp + sim
  const int ldT = NCOMP * nsimp;
  if (col >= ldT) return;
  const RingGeo g = geo[i];
  const int n = g.nfft;
  if (kk > n / 2) return;
  const bool pair = g.i1S != g.i1N;
  const int comp = col / nsimp, sim = col - comp * nsimp;
  const int ncols = 2 * NCOMP * nsimp;
  const int cre = (comp * 2) * nsimp + sim, cim = (comp * 2 + 1) * nsimp + sim;
  const bool edge = (kk == 0) || (kk == n / 2);

  double XNr = 0.0, XNi = 0.0, XSr = 0.0, XSi = 0.0;
  // aliases in ascending m: k, n-k, n+k, 2n-k, ...  (edge slots: k, k+n, ...)
  for (int base = 0;; base += n) {
    bool stop = false;
    for (int side = 0; side < 2; ++side) {
      int m;
      if (side == 0) {
        m = base + kk;
      } else {
        if (edge) break;
        m = base + n - kk;
      }
      if (m > lmax) {
        stop = true;
        break;
      }
      const int i0 = i0tab[m];
      if (i < i0) {  // ring not kept for this (and, i0 being monotone, any larger) m
        stop = true;
        break;
      }
      const int nr = nhr + 1 - i0;
      const long long gb = goff[m] / (NCOMP == 2 ? 1 : 2) + (long long)(i - i0) * ncols;
      const long long go = (long long)nr * ncols;  // offset of the odd class
      const double er = G[gb + cre], ei = G[gb + cim];
      const double orr = G[gb + go + cre], oi = G[gb + go + cim];
      const double2 cs = tw[(long long)m * nnorth + i];
      // north = even + odd, south = even - odd; rotate by e^{+i m phi0}
      const double nr_ = er + orr, ni_ = ei + oi;
      const double WNr = nr_ * cs.x - ni_ * cs.y, WNi = nr_ * cs.y + ni_ * cs.x;
      const double sr_ = er - orr, si_ = ei - oi;
      const double WSr = sr_ * cs.x - si_ * cs.y, WSi = sr_ * cs.y + si_ * cs.x;
      if (edge) {
        const double f = (m == 0) ? 1.0 : 2.0;
        XNr += f * WNr;
        XSr += f * WSr;
      } else if (side == 0) {
        XNr += WNr;
        XNi += WNi;
        XSr += WSr;
        XSi += WSi;
      } else {
        XNr += WNr;
        XNi -= WNi;
        XSr += WSr;
        XSi -= WSi;
      }
    }
    if (stop) break;
  }
  // write PERM slots
  int s0, s1 = -1;
  if (kk == 0)
    s0 = 0;
  else if (kk == n / 2)
    s0 = 1;
  else {
    s0 = 2 * kk;
    s1 = 2 * kk + 1;
  }
  const double nf = (double)n;
  auto put = [&](long long pix, double v) {
    const long long a = pix * ldT + col;
    out[a] = residual ? Xt[a] - nf * v : v;
  };
  put(g.i1N + s0, XNr);
  if (s1 >= 0) put(g.i1N + s1, XNi);
  if (pair) {
    put(g.i1S + s0, XSr);
    if (s1 >= 0) put(g.i1S + s1, XSi);
  }
}

// Mixed precision: the per-(m, class, column) exponent of max |F| over the
// rings, for the int8 slicing of the analysis operand (k_mixed.hip).  A
// thread tracks its columns' maxima (Re/Im of both parity classes), the
// block reduces them over its rings and folds them into fexp (index
// (2 m + class) * ncols + col) with atomicMax.
struct ColMax {
  double mx[4] = {0.0, 0.0, 0.0, 0.0};  // class 0 Re, class 0 Im, class 1 Re, class 1 Im
  __device__ void add(double v0, double v1, double v2, double v3) {
    mx[0] = fmax(mx[0], fabs(v0));
    mx[1] = fmax(mx[1], fabs(v1));
    mx[2] = fmax(mx[2], fabs(v2));
    mx[3] = fmax(mx[3], fabs(v3));
  }
  __device__ void publish(int* fexp, int m, int ncols, int cre, int cim, bool colok) {
    __shared__ int red[4][kFoldBlk.y][kFoldBlk.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][threadIdx.y][threadIdx.x] = mix_exp(mx[k]);
    __syncthreads();
    if (threadIdx.y < 4 && colok) {
      const int k = threadIdx.y;
      int e = red[k][0][threadIdx.x];
#pragma unroll
      for (int y = 1; y < kFoldBlk.y; ++y) e = max(e, red[k][y][threadIdx.x]);
      if (e != kMixZeroE) atomicMax(&fexp[(long long)(2 * m + (k >> 1)) * ncols + ((k & 1) ? cim : cre)], e);
    }
  }
};

// ---------------------------------------------------------------------------
// Ring spectra -> m-major analysis operand (src/fft_mapping.f90:19-68,
// fft2mat_fast, for all m at once).  For order m and kept north ring i:
//   X_N = slot(m) of ring i rotated by e^{-i m phi0}, conjugated for negative
//   frequencies (only the real part survives on DC / Nyquist slots),
//   X_S the same on the mirror ring, then the parity folds
//   F_even = X_N + X_S, F_odd = X_N - X_S   (equator: both = X_N).
// Layout (same as the synthesis buffer): [m][fold][ring - i0(m)][comp][re/im][sim].
// ---------------------------------------------------------------------------
template <int NCOMP>
__global__ __launch_bounds__(256) void k_gather(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                                const double2* __restrict__ tw, const int* __restrict__ slot,
                                                const double* __restrict__ specT, double* __restrict__ F,
                                                const long long* __restrict__ goff, int nhr, int nnorth,
                                                int nsim, int nsimp, int* __restrict__ fexp) {
  if (!cols_active(blockIdx.z * kFoldBlk.x, (blockIdx.z + 1) * kFoldBlk.x, nsim, nsimp)) return;  // no map here
  const int m = blockIdx.x;
  const int i0 = i0tab[m];
  const int ldT = NCOMP * nsimp;
  const int colx = blockIdx.z * kFoldBlk.x + threadIdx.x;  // spectrum column comp*nsimp + sim
  const bool colok = colx < ldT;
  const int col = colok ? colx : ldT - 1;
  const int comp = col / nsimp, sim = col - comp * nsimp;
  const int nr = nhr + 1 - i0;
  const int ncols = 2 * NCOMP * nsimp;
  const long long fb = goff[m] / (NCOMP == 2 ? 1 : 2);
  const long long odd = (long long)nr * ncols;
  const int cre = (comp * 2) * nsimp + sim, cim = cre + nsimp;
  const long long mrow = (long long)m * nnorth;
  ColMax cm;
  for (int i = i0 + blockIdx.y * kGatherRings + threadIdx.y;
       colok && i <= nhr && i < i0 + ((int)blockIdx.y + 1) * kGatherRings; i += kFoldBlk.y) {
    const int sinfo = slot[mrow + i];
    const int rel = sinfo >> 2, cls = sinfo & 3;
    const double sg = cls == 1 ? 1.0 : (cls == 2 ? -1.0 : 0.0);
    const double2 cs = tw[mrow + i];
    const RingGeo g = geo[i];
    const double* pN = specT + (g.i1N + rel) * (long long)ldT + col;
    const double* pS = specT + (g.i1S + rel) * (long long)ldT + col;
    const double xr = pN[0], xi = pN[ldT], yr = pS[0], yi = pS[ldT];
    const double Nr = xr * cs.x + sg * xi * cs.y;
    const double Ni = sg * xi * cs.x - xr * cs.y;
    const double Sr = yr * cs.x + sg * yi * cs.y;
    const double Si = sg * yi * cs.x - yr * cs.y;
    const double fs = (g.i1S != g.i1N) ? 1.0 : 0.0;
    const long long row = fb + (long long)(i - i0) * ncols;
    const double v0 = fma(fs, Sr, Nr), v1 = fma(fs, Si, Ni), v2 = fma(-fs, Sr, Nr), v3 = fma(-fs, Si, Ni);
    F[row + cre] = v0;
    F[row + cim] = v1;
    F[row + odd + cre] = v2;
    F[row + odd + cim] = v3;
    cm.add(v0, v1, v2, v3);
  }
  if (fexp) cm.publish(fexp, m, ncols, cre, cim, colok);
}

// ---------------------------------------------------------------------------
// Residual analysis operand of the "immediate update" iteration
// (t2alm_iter, src/sht_main.f90:161-264).  There the reference synthesises,
// scatters (mat2fft_fast_iter) and re-analyses one order at a time, m
// ascending, so the residual that order m reads on ring i has only the
// synthesised aliases m' <= m of its slot subtracted:
//   R(m, i) = Xt(slot) - n * sum_{m' <= m, m' ~ slot} X_{m'}(i)
// (every synthesis uses the alm from before the sweep).  One thread computes
// that prefix for its (m, ring, column) from the synthesis buffer G, then
// applies the gather of k_gather.  G and F are distinct buffers.
//
// full = 1 sums the whole alias chain instead (every m' <= lmax whose ring
// set includes ring i): the Jacobi residual of t2alm_iter_old, i.e. k_fold's
// residual followed by k_gather, without writing the residual spectrum.
// ---------------------------------------------------------------------------
template <int NCOMP>
__global__ __launch_bounds__(256) void k_gather_prefix(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                                       const double2* __restrict__ tw, const int* __restrict__ slot,
                                                       const double* __restrict__ G, const double* __restrict__ specT,
                                                       double* __restrict__ F, const long long* __restrict__ goff,
                                                       int nhr, int nnorth, int nsim, int nsimp, int lmax, int full,
                                                       int* __restrict__ fexp) {
  if (!cols_active(blockIdx.z * kFoldBlk.x, (blockIdx.z + 1) * kFoldBlk.x, nsim, nsimp)) return;  // no map here
  const int m = blockIdx.x;
  const int i0 = i0tab[m];
  const int ldT = NCOMP * nsimp;
  const int colx = blockIdx.z * kFoldBlk.x + threadIdx.x;
  const bool colok = colx < ldT;
  const int col = colok ? colx : ldT - 1;
  const int comp = col / nsimp, sim = col - comp * nsimp;
  const int nr = nhr + 1 - i0;
  const int ncols = 2 * NCOMP * nsimp;
  const long long fb = goff[m] / (NCOMP == 2 ? 1 : 2);
  const long long odd = (long long)nr * ncols;
  const int cre = (comp * 2) * nsimp + sim, cim = cre + nsimp;
  const long long mrow = (long long)m * nnorth;
  ColMax cm;
  for (int i = i0 + blockIdx.y * kGatherRings + threadIdx.y;
       colok && i <= nhr && i < i0 + ((int)blockIdx.y + 1) * kGatherRings; i += kFoldBlk.y) {
    const RingGeo g = geo[i];
    const int n = g.nfft;
    const int pos = m % n;
    const int kk = pos <= n / 2 ? pos : n - pos;
    const bool edge = (kk == 0) || (kk == n / 2);
    // alias chain of slot kk in ascending order (k_fold's sum), up to m
    // (prefix) or through every order whose ring set includes ring i (full)
    const int mlast = full ? lmax : m;
    double XNr = 0.0, XNi = 0.0, XSr = 0.0, XSi = 0.0;
    bool stop = false;
    for (int base = 0; base <= mlast && !stop; base += n) {
      for (int side = 0; side < 2; ++side) {
        int mp;
        if (side == 0) {
          mp = base + kk;
        } else {
          if (edge) break;
          mp = base + n - kk;
        }
        if (mp > mlast) {
          stop = true;
          break;
        }
        const int i0p = i0tab[mp];  // i0 is non-decreasing in m
        if (i < i0p) {              // ring dropped for mp and every larger order
          stop = true;
          break;
        }
        const int nrp = nhr + 1 - i0p;
        const long long gb = goff[mp] / (NCOMP == 2 ? 1 : 2) + (long long)(i - i0p) * ncols;
        const long long go = (long long)nrp * ncols;
        const double er = G[gb + cre], ei = G[gb + cim];
        const double orr = G[gb + go + cre], oi = G[gb + go + cim];
        const double2 cs = tw[(long long)mp * nnorth + i];
        const double nr_ = er + orr, ni_ = ei + oi;
        const double WNr = nr_ * cs.x - ni_ * cs.y, WNi = nr_ * cs.y + ni_ * cs.x;
        const double sr_ = er - orr, si_ = ei - oi;
        const double WSr = sr_ * cs.x - si_ * cs.y, WSi = sr_ * cs.y + si_ * cs.x;
        if (edge) {
          const double f = (mp == 0) ? 1.0 : 2.0;
          XNr += f * WNr;
          XSr += f * WSr;
        } else if (side == 0) {
          XNr += WNr;
          XNi += WNi;
          XSr += WSr;
          XSi += WSi;
        } else {
          XNr += WNr;
          XNi -= WNi;
          XSr += WSr;
          XSi -= WSi;
        }
      }
    }
    const int sinfo = slot[mrow + i];
    const int rel = sinfo >> 2, cls = sinfo & 3;
    const double sg = cls == 1 ? 1.0 : (cls == 2 ? -1.0 : 0.0);
    const double nf = (double)n;
    const bool pair = g.i1S != g.i1N;
    const double* pN = specT + (g.i1N + rel) * (long long)ldT + col;
    const double* pS = specT + (g.i1S + rel) * (long long)ldT + col;
    // residual slot pair (the Im slot does not exist on DC / Nyquist: sg = 0 there)
    const double xr = pN[0] - nf * XNr;
    const double xi = edge ? 0.0 : pN[ldT] - nf * XNi;
    const double yr = pair ? pS[0] - nf * XSr : 0.0;
    const double yi = (pair && !edge) ? pS[ldT] - nf * XSi : 0.0;
    const double2 cs = tw[mrow + i];
    const double Nr = xr * cs.x + sg * xi * cs.y;
    const double Ni = sg * xi * cs.x - xr * cs.y;
    const double Sr = yr * cs.x + sg * yi * cs.y;
    const double Si = sg * yi * cs.x - yr * cs.y;
    const double fs = pair ? 1.0 : 0.0;
    const long long row = fb + (long long)(i - i0) * ncols;
    const double v0 = fma(fs, Sr, Nr), v1 = fma(fs, Si, Ni), v2 = fma(-fs, Sr, Nr), v3 = fma(-fs, Si, Ni);
    F[row + cre] = v0;
    F[row + cim] = v1;
    F[row + odd + cre] = v2;
    F[row + odd + cim] = v3;
    cm.add(v0, v1, v2, v3);
  }
  if (fexp) cm.publish(fexp, m, ncols, cre, cim, colok);
}

hipError_t launch_gather_prefix(Handle* h, const double* G, const double* specT, double* F, int full, int* fexp) {
  const int ncomp = h->spin ? 2 : 1;
  dim3 block(kFoldBlk.x, kFoldBlk.y);
  dim3 grid(h->lmax + 1, (h->nnorth + kGatherRings - 1) / kGatherRings,
            (ncomp * h->nsimp + kFoldBlk.x - 1) / kFoldBlk.x);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_gather_prefix<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, G,
                       specT, F, h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, h->lmax, full, fexp);
  else
    hipLaunchKernelGGL(k_gather_prefix<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, G,
                       specT, F, h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, h->lmax, full, fexp);
  return hipGetLastError();
}

hipError_t launch_gather(Handle* h, const double* specT, double* F, int* fexp) {
  const int ncomp = h->spin ? 2 : 1;
  dim3 block(kFoldBlk.x, kFoldBlk.y);
  dim3 grid(h->lmax + 1, (h->nnorth + kGatherRings - 1) / kGatherRings,
            (ncomp * h->nsimp + kFoldBlk.x - 1) / kFoldBlk.x);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_gather<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, specT, F,
                       h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, fexp);
  else
    hipLaunchKernelGGL(k_gather<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, specT, F,
                       h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, fexp);
  return hipGetLastError();
}

hipError_t launch_fold(Handle* h, const double* G, const double* specT_t, double* specT_out, int residual) {
  const int ncomp = h->spin ? 2 : 1;
  const int maxk = 2 * h->nside;  // n/2 for the longest ring
  dim3 block(kFoldBlk.x, kFoldBlk.y);
  dim3 grid(h->nnorth, (maxk + 1 + kFoldBlk.y - 1) / kFoldBlk.y, (ncomp * h->nsimp + kFoldBlk.x - 1) / kFoldBlk.x);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_fold<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, G, h->d_goff, specT_t,
                       specT_out, residual, h->lmax, h->nhr, h->nnorth, h->nsim, h->nsimp);
  else
    hipLaunchKernelGGL(k_fold<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, G, h->d_goff, specT_t,
                       specT_out, residual, h->lmax, h->nhr, h->nnorth, h->nsim, h->nsimp);
  return hipGetLastError();
}

}  // namespace fsht
