// m-major synthesis output -> ring spectra (PERM packing), with aliasing.
//
// Replaces mat2fft_fast (src/fft_mapping.f90:84-190) plus the zeroing of
// unused PERM slots (src/sht_main.f90:137-146).  The reference scatters one m
// at a time and relies on ascending m with set-then-accumulate to fold the
// orders that alias onto the same ring frequency (rings shorter than
// 2(lmax+1)).  Here every output slot gathers its aliases explicitly: slot k
// of a ring of length n receives the orders m = k, n-k, n+k, 2n-k, ...
// (m <= lmax, ring kept for m), rotated by e^{+i m phi0}, conjugated for
// negative frequencies and doubled on the DC and Nyquist slots for m > 0
// (src/sht_data_init.f90:542-573).  The spectra are written map-major (the
// layout the inverse ring FFT reads), so no transpose sits between them.
//
// The gathers read the forward FFT's map-major spectra directly as well, and
// the Jacobi residual is formed from the first gathered operand F0 =
// gather(spectrum of the input maps): by linearity
//   gather(Xt - n * fold(G)) = F0 - n * gather(fold(G)),
// so the refinement sweeps never touch the spectra again.
#include "fsht_internal.h"

namespace fsht {

// Thread block of the residual gather: x = 16 columns (one 128-byte line of
// an F row; nsimp is a multiple of 16, so column blocks are never partly
// empty), y = rings.  The gather and the fold use 256 threads as 16 x 16
// tiles of their own (below).
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
// Grid order of the gathers and the fold: the column blocks are the fastest
// grid dimension (x), so the workgroups that share the rows of one order or
// ring run together and whole rows are fetched at once (alias fold and
// residual gathers 44-47 -> 42-43 ms, gather 8.5-9.0 -> 7.8-8.1 ms per step
// at nside 1024 / 200 maps); FSHT_FOLD_CT_FAST=0 puts the orders / rings first.
#ifndef FSHT_FOLD_CT_FAST
#define FSHT_FOLD_CT_FAST 1
#endif
__device__ inline int blk_major() { return FSHT_FOLD_CT_FAST ? blockIdx.z : blockIdx.x; }
__device__ inline int blk_cols() { return FSHT_FOLD_CT_FAST ? blockIdx.x : blockIdx.z; }
static dim3 fold_grid(unsigned major, unsigned y, unsigned cols) {
  return FSHT_FOLD_CT_FAST ? dim3(cols, y, major) : dim3(major, y, cols);
}
// The fold writes each map's ring segment 16 slots at a time: with its slot
// blocks as the fastest grid dimension, consecutive workgroups write
// consecutive segments of the same maps (fold 10.3 -> 9.1 ms at nside 1024 /
// 200 maps); FSHT_FOLD_SLOT_FAST=0 uses the gathers' order.
#ifndef FSHT_FOLD_SLOT_FAST
#define FSHT_FOLD_SLOT_FAST 1
#endif


template <int NCOMP>
__global__ __launch_bounds__(256) void k_fold(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                              const double2* __restrict__ tw, const double* __restrict__ G,
                                              const long long* __restrict__ goff, double* __restrict__ spec,
                                              int lmax, int nhr, int nnorth, long long npix, int nsim, int nsimp) {
  const int bcol = FSHT_FOLD_SLOT_FAST ? (int)blockIdx.y : blk_cols();
  if (!cols_active(bcol * 16, (bcol + 1) * 16, nsim, nsimp)) return;  // no map here
  const int i = FSHT_FOLD_SLOT_FAST ? (int)blockIdx.z : blk_major();  // north ring
  const RingGeo g = geo[i];
  const int n = g.nfft;
  const int k0 = (FSHT_FOLD_SLOT_FAST ? (int)blockIdx.x : (int)blockIdx.y) * 16;
  if (k0 > n / 2) return;  // whole workgroup
  __shared__ double tile[4][16][17];  // [value][column][slot]
  const bool pair = g.i1S != g.i1N;
  const int ldT = NCOMP * nsimp;
  const int ncols = 2 * NCOMP * nsimp;
  // sum role: lane = column (G rows are read 16 columns at a time), then slot
  {
    const int kk = k0 + (int)(threadIdx.x >> 4);
    const int col = bcol * 16 + (int)(threadIdx.x & 15);  // spectrum column: comp*nsimp + sim
    const int comp = col / nsimp, sim = col - comp * nsimp;
    const int cre = (comp * 2) * nsimp + sim, cim = (comp * 2 + 1) * nsimp + sim;
    const bool edge = (kk == 0) || (kk == n / 2);
    double XNr = 0.0, XNi = 0.0, XSr = 0.0, XSi = 0.0;
    // aliases in ascending m: k, n-k, n+k, 2n-k, ...  (edge slots: k, k+n, ...)
    for (int base = 0; col < ldT && kk <= n / 2; base += n) {
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
    const int tc = threadIdx.x & 15, tk = threadIdx.x >> 4;
    tile[0][tc][tk] = XNr;
    tile[1][tc][tk] = XNi;
    tile[2][tc][tk] = XSr;
    tile[3][tc][tk] = XSi;
  }
  __syncthreads();
  // write role: lane = slot (a map's PERM slots are adjacent), then column
  const int tk = threadIdx.x & 15, tc = threadIdx.x >> 4;
  const int kk = k0 + tk;
  const int col = bcol * 16 + tc;
  const int comp = col / nsimp, sim = col - comp * nsimp;
  if (col >= ldT || sim >= nsim || kk > n / 2) return;  // padding column (no map) or past the ring's last slot
  int s0, s1 = -1;
  if (kk == 0)
    s0 = 0;
  else if (kk == n / 2)
    s0 = 1;
  else {
    s0 = 2 * kk;
    s1 = 2 * kk + 1;
  }
  double* out = spec + ((long long)comp * nsim + sim) * npix;
  out[g.i1N + s0] = tile[0][tc][tk];
  if (s1 >= 0) out[g.i1N + s1] = tile[1][tc][tk];
  if (pair) {
    out[g.i1S + s0] = tile[2][tc][tk];
    if (s1 >= 0) out[g.i1S + s1] = tile[3][tc][tk];
  }
}

// Mixed precision: the per-(m, class, column) exponent of max |F| over the
// rings, for the int8 slicing of the analysis operand (k_mixed.hip).  A
// thread tracks its column's maxima (Re/Im of both parity classes) over its
// rings and folds them into fexp (index (2 m + class) * ncols + col) with
// atomicMax; publish_rings first reduces over the workgroup's rings
// (threadIdx.y) in LDS.
struct ColMax {
  double mx[4] = {0.0, 0.0, 0.0, 0.0};  // class 0 Re, class 0 Im, class 1 Re, class 1 Im
  __device__ void add(double v0, double v1, double v2, double v3) {
    mx[0] = fmax(mx[0], fabs(v0));
    mx[1] = fmax(mx[1], fabs(v1));
    mx[2] = fmax(mx[2], fabs(v2));
    mx[3] = fmax(mx[3], fabs(v3));
  }
  __device__ void publish(int* fexp, int m, int ncols, int cre, int cim) const {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = mix_exp(mx[k]);
      if (e != kMixZeroE) atomicMax(&fexp[(long long)(2 * m + (k >> 1)) * ncols + ((k & 1) ? cim : cre)], e);
    }
  }
  __device__ void publish_rings(int* fexp, int m, int ncols, int cre, int cim, bool colok) {
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
// The spectra are map-major (as the ring FFT writes them).  Workgroup = 16
// orders x 16 spectrum columns over the rings [blockIdx.y, blockIdx.y + 1) *
// kGatherRings; per ring the threads first read with lanes along the orders
// of one map (adjacent PERM slots: whole cache lines), then, after an LDS
// transpose, write with lanes along the columns of one order (one 128-byte F
// line per 16 columns).
// ---------------------------------------------------------------------------
constexpr int kGatherM = 16;  // orders per gather workgroup

template <int NCOMP>
__global__ __launch_bounds__(256) void k_gather(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                                const double2* __restrict__ tw, const int* __restrict__ slot,
                                                const double* __restrict__ spec, double* __restrict__ F,
                                                const long long* __restrict__ goff, int nhr, int nnorth, int lmax,
                                                long long npix, int nsim, int nsimp, int* __restrict__ fexp) {
  if (!cols_active(blk_cols() * 16, (blk_cols() + 1) * 16, nsim, nsimp)) return;  // no map here
  const int m0 = blk_major() * kGatherM, c0 = blk_cols() * 16;
  // i0 is non-decreasing in m: no order of the block keeps a ring below i0(m0)
  const int ibeg = max((int)blockIdx.y * kGatherRings, i0tab[m0]);
  const int iend = min(nhr + 1, ((int)blockIdx.y + 1) * kGatherRings);
  if (ibeg >= iend) return;  // whole workgroup
  __shared__ double tile[4][16][kGatherM + 1];  // [value][column][order]
  const int ldT = NCOMP * nsimp, ncols = 2 * NCOMP * nsimp;
  // read role: lane = order, then column
  const int mr = m0 + (int)(threadIdx.x & 15), cr = c0 + (int)(threadIdx.x >> 4);
  const bool okr = mr <= lmax && cr < ldT;
  const int compr = cr / nsimp, simr = cr - compr * nsimp;
  const bool has = okr && simr < nsim;  // padding columns gather zeros
  const double* sp = spec + ((long long)compr * nsim + (has ? simr : 0)) * npix;
  const int i0r = i0tab[min(mr, lmax)];
  const long long mrow = (long long)min(mr, lmax) * nnorth;
  // write role: lane = column, then order
  const int mw = m0 + (int)(threadIdx.x >> 4), cw = c0 + (int)(threadIdx.x & 15);
  const bool okw = mw <= lmax && cw < ldT;
  const int mwc = min(mw, lmax);
  const int compw = cw / nsimp, simw = cw - compw * nsimp;
  const int i0w = i0tab[mwc];
  const long long fb = goff[mwc] / (NCOMP == 2 ? 1 : 2);
  const long long odd = (long long)(nhr + 1 - i0w) * ncols;
  const int cre = (compw * 2) * nsimp + simw, cim = cre + nsimp;
  ColMax cm;
  for (int i = ibeg; i < iend; ++i) {
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (has && i >= i0r) {
      const int sinfo = slot[mrow + i];
      const int rel = sinfo >> 2, cls = sinfo & 3;
      const double sg = cls == 1 ? 1.0 : (cls == 2 ? -1.0 : 0.0);
      const double2 cs = tw[mrow + i];
      const RingGeo g = geo[i];
      // slot rel and the next one (rel + 1 < n: every ring has at least 4 slots)
      const double xr = sp[g.i1N + rel], xi = sp[g.i1N + rel + 1];
      const double yr = sp[g.i1S + rel], yi = sp[g.i1S + rel + 1];
      const double Nr = xr * cs.x + sg * xi * cs.y;
      const double Ni = sg * xi * cs.x - xr * cs.y;
      const double Sr = yr * cs.x + sg * yi * cs.y;
      const double Si = sg * yi * cs.x - yr * cs.y;
      const double fs = (g.i1S != g.i1N) ? 1.0 : 0.0;
      v0 = fma(fs, Sr, Nr);
      v1 = fma(fs, Si, Ni);
      v2 = fma(-fs, Sr, Nr);
      v3 = fma(-fs, Si, Ni);
    }
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    tile[0][tc][tr] = v0;
    tile[1][tc][tr] = v1;
    tile[2][tc][tr] = v2;
    tile[3][tc][tr] = v3;
    __syncthreads();
    if (okw && i >= i0w) {
      const int wr = threadIdx.x >> 4, wc = threadIdx.x & 15;
      const double w0 = tile[0][wc][wr], w1 = tile[1][wc][wr], w2 = tile[2][wc][wr], w3 = tile[3][wc][wr];
      const long long row = fb + (long long)(i - i0w) * ncols;
      F[row + cre] = w0;
      F[row + cim] = w1;
      F[row + odd + cre] = w2;
      F[row + odd + cim] = w3;
      cm.add(w0, w1, w2, w3);
    }
    __syncthreads();
  }
  if (fexp && okw) cm.publish(fexp, mwc, ncols, cre, cim);
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
// applies the gather of k_gather, as F0 - n * gather(sum) with F0 =
// gather(Xt) written by k_gather.  G, F0 and F are distinct buffers.
//
// full = 1 sums the whole alias chain instead (every m' <= lmax whose ring
// set includes ring i): the Jacobi residual of t2alm_iter_old
// (src/fft_mapping.f90:177-183 followed by fft2mat_fast), without forming the
// residual spectrum.
// ---------------------------------------------------------------------------
template <int NCOMP>
__global__ __launch_bounds__(256) void k_gather_prefix(const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
                                                       const double2* __restrict__ tw, const int* __restrict__ slot,
                                                       const double* __restrict__ G, const double* __restrict__ F0,
                                                       double* __restrict__ F, const long long* __restrict__ goff,
                                                       int nhr, int nnorth, int nsim, int nsimp, int lmax, int full,
                                                       int* __restrict__ fexp) {
  if (!cols_active(blk_cols() * kFoldBlk.x, (blk_cols() + 1) * kFoldBlk.x, nsim, nsimp)) return;  // no map here
  const int m = blk_major();
  const int i0 = i0tab[m];
  const int ldT = NCOMP * nsimp;
  const int colx = blk_cols() * kFoldBlk.x + threadIdx.x;
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
    // gather of the synthesised alias sum (k_gather's arithmetic; the Im
    // slot does not exist on DC / Nyquist, where sg = 0)
    const int cls = slot[mrow + i] & 3;
    const double sg = cls == 1 ? 1.0 : (cls == 2 ? -1.0 : 0.0);
    const double nf = (double)n;
    const double2 cs = tw[mrow + i];
    const double Nr = XNr * cs.x + sg * XNi * cs.y;
    const double Ni = sg * XNi * cs.x - XNr * cs.y;
    const double Sr = XSr * cs.x + sg * XSi * cs.y;
    const double Si = sg * XSi * cs.x - XSr * cs.y;
    const double fs = (g.i1S != g.i1N) ? 1.0 : 0.0;
    const long long row = fb + (long long)(i - i0) * ncols;
    const double v0 = fma(-nf, fma(fs, Sr, Nr), F0[row + cre]);
    const double v1 = fma(-nf, fma(fs, Si, Ni), F0[row + cim]);
    const double v2 = fma(-nf, fma(-fs, Sr, Nr), F0[row + odd + cre]);
    const double v3 = fma(-nf, fma(-fs, Si, Ni), F0[row + odd + cim]);
    F[row + cre] = v0;
    F[row + cim] = v1;
    F[row + odd + cre] = v2;
    F[row + odd + cim] = v3;
    cm.add(v0, v1, v2, v3);
  }
  if (fexp) cm.publish_rings(fexp, m, ncols, cre, cim, colok);
}

hipError_t launch_gather_prefix(Handle* h, const double* G, const double* F0, double* F, int full, int* fexp) {
  const int ncomp = h->spin ? 2 : 1;
  dim3 block(kFoldBlk.x, kFoldBlk.y);
  dim3 grid = fold_grid(h->lmax + 1, (h->nnorth + kGatherRings - 1) / kGatherRings,
                        (ncomp * h->nsimp + kFoldBlk.x - 1) / kFoldBlk.x);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_gather_prefix<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, G, F0,
                       F, h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, h->lmax, full, fexp);
  else
    hipLaunchKernelGGL(k_gather_prefix<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, G, F0,
                       F, h->d_goff, h->nhr, h->nnorth, h->nsim, h->nsimp, h->lmax, full, fexp);
  return hipGetLastError();
}

hipError_t launch_gather(Handle* h, const double* spec, double* F, int* fexp) {
  const int ncomp = h->spin ? 2 : 1;
  dim3 block(256);
  dim3 grid = fold_grid((h->lmax + kGatherM) / kGatherM, (h->nnorth + kGatherRings - 1) / kGatherRings,
                        (ncomp * h->nsimp + 15) / 16);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_gather<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, spec, F,
                       h->d_goff, h->nhr, h->nnorth, h->lmax, h->npix, h->nsim, h->nsimp, fexp);
  else
    hipLaunchKernelGGL(k_gather<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, h->d_slot, spec, F,
                       h->d_goff, h->nhr, h->nnorth, h->lmax, h->npix, h->nsim, h->nsimp, fexp);
  return hipGetLastError();
}

hipError_t launch_fold(Handle* h, const double* G, double* spec) {
  const int ncomp = h->spin ? 2 : 1;
  const int maxk = 2 * h->nside;  // n/2 for the longest ring
  dim3 block(256);
  dim3 grid = FSHT_FOLD_SLOT_FAST ? dim3((maxk + 1 + 15) / 16, (ncomp * h->nsimp + 15) / 16, h->nnorth)
                                  : fold_grid(h->nnorth, (maxk + 1 + 15) / 16, (ncomp * h->nsimp + 15) / 16);
  if (ncomp == 2)
    hipLaunchKernelGGL(k_fold<2>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, G, h->d_goff, spec,
                       h->lmax, h->nhr, h->nnorth, h->npix, h->nsim, h->nsimp);
  else
    hipLaunchKernelGGL(k_fold<1>, grid, block, 0, h->stream, h->d_geo, h->d_i0, h->d_tw, G, h->d_goff, spec,
                       h->lmax, h->nhr, h->nnorth, h->npix, h->nsim, h->nsimp);
  return hipGetLastError();
}

}  // namespace fsht
