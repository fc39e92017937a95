// Mixed-precision Legendre contractions for the temperature transforms.
//
// The analysis alm(l, col) = pa * sum_ring P_lm(ring) F(ring, col) and the
// synthesis G(ring, col) = sum_l P_lm(ring) alm(l, col) of the FP64 kernels
// (k_legendre.hip; the reference's mat2alm / alm2mat dgemm,
// src/sht_core.f90:30-168) are evaluated as sums of exact int8 products:
// every operand row or column is cut into 7-bit fixed-point digits
// ("slices") relative to one power-of-two scale, the digit products run on
// v_mfma_i32_32x32x32_i8 with exact int32 accumulation (one accumulator per
// digit sum), and the sums are combined in FP64 with the two scales at the
// end.  With K slices a result carries about 7K bits relative to the row and
// column maxima.  The Jacobi refinement of t2alm (src/sht_main.f90:276-367)
// keeps its residual and its alm update in FP64, so each pass only needs the
// precision that the later iterations do not damp away.
//
// The Legendre factors do not depend on the data: they are sliced once per
// handle into two tables (analysis and synthesis operand orders, six slices)
// by the same FP64 recurrence, from the same start states, as the FP64
// kernels.  The data operands are sliced per pass (k_quant_F, k_quant_A).
//
// N/S ring symmetry is folded as in the FP64 path: one contraction per parity
// class over the north rings, with the folded operand N +- S.
#include <algorithm>
#include <cstdlib>
#include <tuple>
#include <vector>

#include "fsht_internal.h"

namespace fsht {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* lds_ptr;

// 32x32x32 int8 MFMA operand layout (verified on gfx950 with a probe):
//   A[i][k] in lane i + 32*(k/16), byte k%16;  B[k][j] in lane j + 32*(k/16), byte k%16;
//   D[i][j] in lane j + 32*((i/4)%2), accumulator register 4*(i/8) + i%4.
__device__ inline v16i mfma_i8(v4i a, v4i b, v16i c) { return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0); }

// tile row held by accumulator register r of lane half h (= lane / 32), and
// the position 16 h + r of tile row i in accumulator order
__device__ inline int drow(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }
__device__ inline int dpos(int i) { return 16 * ((i >> 2) & 1) + 4 * (i >> 3) + (i & 3); }

// keeps the scheduler from moving any instruction across this point
__device__ inline void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

constexpr int kNeverStep = 0x3f3f3f3f;  // "no nonzero step" (fill value of the first-step tables)
constexpr int kNeverL = 0x3fffffff;

// K digits of four values x (|x| <= 127/128), one word per slice holding the
// four digits little-endian.  Digit i is the rounded next 7 bits of the
// remainder, so the first k' <= K digits are the k'-digit rounding.
template <int K>
__device__ inline void digits4(double x0, double x1, double x2, double x3, unsigned (&w)[K]) {
  double r[4] = {x0, x1, x2, x3};
#pragma unroll
  for (int i = 0; i < K; ++i) {
    unsigned word = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double t = r[b] * 128.0;
      const double q = rint(t);
      r[b] = t - q;
      word |= ((unsigned)(int)q & 255u) << (8 * b);
    }
    w[i] = word;
  }
}

// 16 values v[0..15] (already scaled) -> K 16-byte slices at dst + s * 1024
template <int K>
__device__ inline void put16(const double* v, signed char* dst) {
  unsigned w[4][K];
#pragma unroll
  for (int q = 0; q < 4; ++q) digits4<K>(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3], w[q]);
#pragma unroll
  for (int s = 0; s < K; ++s)
    *reinterpret_cast<v4i*>(dst + s * 1024) = v4i{(int)w[0][s], (int)w[1][s], (int)w[2][s], (int)w[3][s]};
}

__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// P tables.  One workgroup per (m, block of 256 ring rows); thread = ring,
// running the FP64 recurrence of the FP64 synthesis (entered at its first
// kept l from the resident start state) in tiles of 64 l steps, i.e. 32
// values of each parity class.
//   MODE 0: exponents.  TS: per ring and class, over all l (thread-local).
//           TA: per l, over all rings (workgroup max, then atomicMax across
//           workgroups).  Also the first nonzero step of every tile.
//   MODE 1: digits.  A ring's 32 values of a tile form its TS lane directly;
//           TA lanes (16 rings of one l) are transposed through LDS.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256, 1) void k_ptab(const int2* __restrict__ items, const RingGeo* __restrict__ geo,
                                                 const int* __restrict__ i0tab, const double2* __restrict__ ab,
                                                 const long long* __restrict__ aboff,
                                                 const long long* __restrict__ synoff,
                                                 const int* __restrict__ syn_st, const double2* __restrict__ syn_v,
                                                 const MixOff* __restrict__ mo, signed char* __restrict__ TA,
                                                 int* __restrict__ TAe, signed char* __restrict__ TS,
                                                 int* __restrict__ TSe, int* __restrict__ firstA,
                                                 int* __restrict__ firstS, int lmax, int nhr) {
  constexpr int K = kMixK;
  __shared__ __align__(16) signed char stage[MODE ? 2 : 1][MODE ? K : 1][32][MODE ? 256 + 16 : 16];
  __shared__ int emax[2][32], rmin[2];
  const int2 it = items[blockIdx.x];
  const int m = it.x, rb = it.y, tid = threadIdx.x;
  const int i0 = i0tab[m];
  const int R = nhr + 1 - i0;
  const int row = rb * 256 + tid;
  const bool valid = row < R;
  const long long yo = synoff[m] + (valid ? row : R - 1);
  const int st = valid ? syn_st[yo] : kNeverL;
  const double2 sv = syn_v[yo];  // (P_ls-2, P_ls-1) at the 32-step chunk start ls <= st
  const int ls = (st > lmax) ? kNeverL : m + ((st - m) & ~31);
  const double x = geo[valid ? i0 + row : nhr].z;
  const double2* abm = ab + aboff[m];
  const MixOff o0 = mo[2 * m], o1 = mo[2 * m + 1];
  const bool inrow = row < o0.nrs * 32;  // row of the padded ring tiles (both classes have nrs tiles)
  const long long rpos = (long long)(row >> 5) * 32 + dpos(row & 31);  // within a class's TS exponents
  int eS0 = kMixZeroE, eS1 = kMixZeroE;
  if (MODE == 1 && inrow) {
    eS0 = TSe[o0.tse + rpos];
    eS1 = TSe[o1.tse + rpos];
  }
  if (tid < 64) emax[tid >> 5][tid & 31] = kMixZeroE;
  if (tid < 2) rmin[tid] = kNeverStep;
  __syncthreads();

  double p1 = 0.0, p2 = 0.0, mx0 = 0.0, mx1 = 0.0;
  int myfirst0 = kNeverStep, myfirst1 = kNeverStep;
  const int ntile = (lmax - m) / 64 + 1;
  for (int lt = 0; lt < ntile; ++lt) {
    double v[64];
    const int lb = m + 64 * lt;
#pragma unroll
    for (int s = 0; s < 64; ++s) {
      const int l = lb + s;
      if ((s & 31) == 0 && l == ls) {
        p1 = sv.y;
        p2 = sv.x;
      }
      const double2 c = (l <= lmax) ? abm[l - m] : make_double2(0.0, 0.0);
      const double p = fma(c.x * x, p1, -(c.y * p2));
      p2 = p1;
      p1 = p;
      v[s] = p;
    }
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
      const MixOff& o = cls ? o1 : o0;
      if (lt >= o.nlt) continue;
      double mx = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) mx = fmax(mx, fabs(v[2 * j + cls]));
      if (MODE == 0) {
        if (cls) mx1 = fmax(mx1, mx); else mx0 = fmax(mx0, mx);
        if (mx > 0.0) {
#pragma unroll
          for (int j = 0; j < 32; ++j) atomicMax(&emax[cls][j], mix_exp(fabs(v[2 * j + cls])));
          atomicMin(&rmin[cls], row >> 5);
          if (inrow) {
            if (cls) myfirst1 = min(myfirst1, lt); else myfirst0 = min(myfirst0, lt);
          }
        }
      } else if (inrow) {
        // TS: this ring's lane (row & 31) + 32 g holds values 16 g .. 16 g + 15
        const int e = cls ? eS1 : eS0;
        double sc[32];
#pragma unroll
        for (int j = 0; j < 32; ++j) sc[j] = ldexp(v[2 * j + cls], -e);
        signed char* dst = TS + o.ts + ((long long)(row >> 5) * o.nlt + lt) * K * 1024 + (row & 31) * 16;
        put16<K>(sc, dst);
        put16<K>(sc + 16, dst + 32 * 16);
      }
    }
    if (MODE == 1) {
      // TA digits of the 32 values of each class with the per-l exponents, into LDS
#pragma unroll
      for (int cls = 0; cls < 2; ++cls) {
        const MixOff& o = cls ? o1 : o0;
        if (lt >= o.nlt) continue;
        const int* ea = TAe + o.tae + (long long)lt * 32;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          double r = ldexp(v[2 * j + cls], -ea[dpos(j)]);
#pragma unroll
          for (int s = 0; s < K; ++s) {
            const double t = r * 128.0;
            const double q = rint(t);
            r = t - q;
            stage[cls][s][j][tid] = (signed char)(int)q;
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
      const MixOff& o = cls ? o1 : o0;
      if (lt >= o.nlt) continue;
      if (MODE == 0) {
        if (tid < 32 && emax[cls][tid] != kMixZeroE) atomicMax(&TAe[o.tae + (long long)lt * 32 + dpos(tid)], emax[cls][tid]);
        if (tid == 0 && rmin[cls] != kNeverStep) atomicMin(&firstA[o.fa + lt], rmin[cls]);
      } else {
        const int nrs_here = min(8, o.nrs - rb * 8);
        for (int ch = tid; ch < nrs_here * K * 64; ch += 256) {
          const int L = ch & 63, s = (ch >> 6) % K, rsl = (ch >> 6) / K;
          const v4i val = *reinterpret_cast<const v4i*>(&stage[cls][s][L & 31][rsl * 32 + 16 * (L >> 5)]);
          *reinterpret_cast<v4i*>(TA + o.ta + (((long long)lt * o.nrs + rb * 8 + rsl) * K + s) * 1024 + L * 16) = val;
        }
      }
    }
    __syncthreads();
    if (MODE == 0) {
      if (tid < 64) emax[tid >> 5][tid & 31] = kMixZeroE;
      if (tid < 2) rmin[tid] = kNeverStep;
      __syncthreads();
    }
  }
  if (MODE == 0 && inrow) {
    TSe[o0.tse + rpos] = valid ? mix_exp(mx0) : kMixZeroE;
    TSe[o1.tse + rpos] = valid ? mix_exp(mx1) : kMixZeroE;
    if (myfirst0 != kNeverStep) atomicMin(&firstS[o0.fs + (row >> 5)], myfirst0);
    if (myfirst1 != kNeverStep) atomicMin(&firstS[o1.fs + (row >> 5)], myfirst1);
  }
}

// ---------------------------------------------------------------------------
// Spin-2 tables: F1 (FN = 0) or F2 (FN = 1) of src/get_plm_recursive.f90:105-133,
// from the FP64 recurrence and start states of the P tables, every value
// formed by spin_f from P_l and P_{l-1} (the FP64 spin-2 kernels' own
// expression).  Workgroup = (m, block of 256 ring rows); thread = ring.
// k_ftab_exp (once per handle): the exponents (TSe per ring row and class
// over every l; TAe per l over the rings) and the first nonzero steps, as
// k_ptab<0>, in tiles of 64 l steps (32 values per class).
// ---------------------------------------------------------------------------
template <int FN>
__global__ __launch_bounds__(256, 1) void k_ftab_exp(const int2* __restrict__ items, const RingGeo* __restrict__ geo,
                                                 const int* __restrict__ i0tab, const double2* __restrict__ ab,
                                                 const double2* __restrict__ spin, const long long* __restrict__ aboff,
                                                 const long long* __restrict__ synoff,
                                                 const int* __restrict__ syn_st, const double2* __restrict__ syn_v,
                                                 const MixOff* __restrict__ mo, int* __restrict__ TAe, int* __restrict__ TSe,
                                                 int* __restrict__ firstA, int* __restrict__ firstS, int lmax,
                                                 int nhr) {
  constexpr int MODE = 0;
  constexpr bool TA_OUT = false, TS_OUT = false;
  constexpr int KS = 1;
  __shared__ __align__(16) signed char stage[1][1][32][16];
  __shared__ int emax[2][32], rmin[2];
  const int2 it = items[blockIdx.x];
  const int m = it.x, rb = it.y, tid = threadIdx.x;
  const int i0 = i0tab[m];
  const int R = nhr + 1 - i0;
  const int row = rb * 256 + tid;
  const bool valid = row < R;
  const long long yo = synoff[m] + (valid ? row : R - 1);
  const int st = valid ? syn_st[yo] : kNeverL;
  const double2 sv = syn_v[yo];  // (P_ls-2, P_ls-1) at the 32-step chunk start ls <= st
  const int ls = (st > lmax) ? kNeverL : m + ((st - m) & ~31);
  const RingGeo g = geo[valid ? i0 + row : nhr];
  const double x = g.z;
  const double s2 = g.s * g.s;
  const double bw = x / s2, t2 = 2.0 / s2, cw = (double)m / s2;
  const double2* abm = ab + aboff[m];
  const double2* spm = spin + aboff[m];
  const MixOff o0 = mo[2 * m], o1 = mo[2 * m + 1];
  const bool inrow = row < o0.nrs * 32;  // row of the padded ring tiles (both classes have nrs tiles)
  const long long rpos = (long long)(row >> 5) * 32 + dpos(row & 31);  // within a class's TS exponents
  if (tid < 64) emax[tid >> 5][tid & 31] = kMixZeroE;
  if (tid < 2) rmin[tid] = kNeverStep;
  __syncthreads();

  double p1 = 0.0, p2 = 0.0, mx0 = 0.0, mx1 = 0.0;
  int myfirst0 = kNeverStep, myfirst1 = kNeverStep;
  const int ntile = (lmax - m) / 64 + 1;
  for (int lt = 0; lt < ntile; ++lt) {
    double v[64];
    const int lb = m + 64 * lt;
#pragma unroll
    for (int s = 0; s < 64; ++s) {
      const int l = lb + s;
      if ((s & 31) == 0 && l == ls) {
        p1 = sv.y;
        p2 = sv.x;
      }
      const bool in = l <= lmax;
      const double2 c = in ? abm[l - m] : make_double2(0.0, 0.0);
      const double2 sp = in ? spm[l - m] : make_double2(0.0, 0.0);
      const double p = fma(c.x * x, p1, -(c.y * p2));
      v[s] = spin_f(FN, p, p1, sp, l, m, x, bw, t2, cw);
      p2 = p1;
      p1 = p;
    }
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
      const MixOff& o = cls ? o1 : o0;
      if (lt >= o.nlt) continue;
      double mx = 0.0;
#pragma unroll
      for (int j = 0; j < 32; ++j) mx = fmax(mx, fabs(v[2 * j + cls]));
      if (MODE == 0) {
        if (cls) mx1 = fmax(mx1, mx); else mx0 = fmax(mx0, mx);
        if (mx > 0.0) {
#pragma unroll
          for (int j = 0; j < 32; ++j) atomicMax(&emax[cls][j], mix_exp(fabs(v[2 * j + cls])));
          atomicMin(&rmin[cls], row >> 5);
          if (inrow) {
            if (cls) myfirst1 = min(myfirst1, lt); else myfirst0 = min(myfirst0, lt);
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
      const MixOff& o = cls ? o1 : o0;
      if (lt >= o.nlt) continue;
      if (MODE == 0) {
        if (tid < 32 && emax[cls][tid] != kMixZeroE) atomicMax(&TAe[o.tae + (long long)lt * 32 + dpos(tid)], emax[cls][tid]);
        if (tid == 0 && rmin[cls] != kNeverStep) atomicMin(&firstA[o.fa + lt], rmin[cls]);
      }
    }
    __syncthreads();
    if (MODE == 0) {
      if (tid < 64) emax[tid >> 5][tid & 31] = kMixZeroE;
      if (tid < 2) rmin[tid] = kNeverStep;
      __syncthreads();
    }
  }
  if (MODE == 0 && inrow) {
    TSe[o0.tse + rpos] = valid ? mix_exp(mx0) : kMixZeroE;
    TSe[o1.tse + rpos] = valid ? mix_exp(mx1) : kMixZeroE;
    if (myfirst0 != kNeverStep) atomicMin(&firstS[o0.fs + (row >> 5)], myfirst0);
    if (myfirst1 != kNeverStep) atomicMin(&firstS[o1.fs + (row >> 5)], myfirst1);
  }
}

// k_ftab_dig (per pass): the KS digit slices of one operand order (LAY 0:
// TA, LAY 1: TS) for the orders of one range, into a buffer holding the
// range (the blocks of (m, class) start at byte (o.ta or o.ts) / kMixK * KS -
// tbase).  Half tiles of 32 l steps (16 values per class), 8 steps at a time
// (scheduling fences keep the coefficient loads of one group of 8 live, not
// of the whole tile): TS digits are packed into the ring's 16-byte lanes in
// registers and stored per half tile; TA digits go through LDS and leave as
// 16-byte lanes of 16 rings each.
template <int FN, int LAY, int KS>
__global__ __launch_bounds__(256, 2) void k_ftab_dig(const int2* __restrict__ items, const RingGeo* __restrict__ geo,
                                                     const int* __restrict__ i0tab, const double2* __restrict__ ab,
                                                     const double2* __restrict__ spin,
                                                     const long long* __restrict__ aboff,
                                                     const long long* __restrict__ synoff,
                                                     const int* __restrict__ syn_st,
                                                     const double2* __restrict__ syn_v, const MixOff* __restrict__ mo,
                                                     signed char* __restrict__ T, long long tbase,
                                                     const int* __restrict__ TAe, const int* __restrict__ TSe,
                                                     int lmax, int nhr) {
  constexpr bool TA = LAY == 0;
  __shared__ __align__(16) signed char stage[TA ? 2 : 1][TA ? KS : 1][16][TA ? 256 + 16 : 16];
  __shared__ int ea_s[2][32];
  const int2 it = items[blockIdx.x];
  const int m = it.x, rb = it.y, tid = threadIdx.x;
  const int i0 = i0tab[m];
  const int R = nhr + 1 - i0;
  const int row = rb * 256 + tid;
  const bool valid = row < R;
  const long long yo = synoff[m] + (valid ? row : R - 1);
  const int st = valid ? syn_st[yo] : kNeverL;
  const double2 sv = syn_v[yo];  // (P_ls-2, P_ls-1) at the 32-step chunk start ls <= st
  const int ls = (st > lmax) ? kNeverL : m + ((st - m) & ~31);
  const RingGeo g = geo[valid ? i0 + row : nhr];
  const double x = g.z;
  const double s2 = g.s * g.s;
  const double bw = x / s2, t2 = 2.0 / s2, cw = (double)m / s2;
  const double2* abm = ab + aboff[m];
  const double2* spm = spin + aboff[m];
  const MixOff o0 = mo[2 * m], o1 = mo[2 * m + 1];
  const bool inrow = row < o0.nrs * 32;  // row of the padded ring tiles (both classes have nrs tiles)
  const long long rpos = (long long)(row >> 5) * 32 + dpos(row & 31);
  int eS0 = kMixZeroE, eS1 = kMixZeroE;
  if (!TA && inrow) {
    eS0 = TSe[o0.tse + rpos];
    eS1 = TSe[o1.tse + rpos];
  }
  double p1 = 0.0, p2 = 0.0;
  const int ntile = (lmax - m) / 64 + 1;
  for (int lt = 0; lt < ntile; ++lt) {
    const int lb = m + 64 * lt;
    if (TA) {
      if (tid < 64) {
        const int cls = tid >> 5, j = tid & 31;
        const MixOff& o = cls ? o1 : o0;
        ea_s[cls][j] = lt < o.nlt ? TAe[o.tae + (long long)lt * 32 + dpos(j)] : 0;
      }
      __syncthreads();
    }
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      unsigned w[2][KS][4];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < KS; ++k)
#pragma unroll
          for (int q = 0; q < 4; ++q) w[c][k][q] = 0u;
#pragma unroll
      for (int c8 = 0; c8 < 4; ++c8) {
#pragma unroll
        for (int s8 = 0; s8 < 8; ++s8) {
          const int s = 32 * hh + 8 * c8 + s8;  // step within the tile
          const int l = lb + s;
          if (s8 == 0 && c8 == 0 && l == ls) {
            p1 = sv.y;
            p2 = sv.x;
          }
          const bool in = l <= lmax;
          const double2 cf = in ? abm[l - m] : make_double2(0.0, 0.0);
          const double2 sp = in ? spm[l - m] : make_double2(0.0, 0.0);
          const double p = fma(cf.x * x, p1, -(cf.y * p2));
          const double f = spin_f(FN, p, p1, sp, l, m, x, bw, t2, cw);
          p2 = p1;
          p1 = p;
          const int cls = s & 1, jh = (s >> 1) & 15;  // class, value within the half tile
          double r = ldexp(f, -(TA ? ea_s[cls][(s >> 1)] : (cls ? eS1 : eS0)));
#pragma unroll
          for (int k = 0; k < KS; ++k) {
            const double t = r * 128.0;
            const double q = rint(t);
            r = t - q;
            const unsigned byte = (unsigned)(int)q & 255u;
            if (TA) stage[cls][k][jh][tid] = (signed char)byte;
            else w[cls][k][jh >> 2] |= byte << (8 * (jh & 3));
          }
        }
        sched_fence();
      }
      if (!TA) {
        // this ring's lane (row & 31) + 32 hh holds the half tile's 16 values
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
          const MixOff& o = cls ? o1 : o0;
          if (lt >= o.nlt || !inrow) continue;
          signed char* dst = T + (o.ts / kMixK * KS - tbase) + ((long long)(row >> 5) * o.nlt + lt) * KS * 1024 +
                             ((row & 31) + 32 * hh) * 16;
#pragma unroll
          for (int k = 0; k < KS; ++k)
            *reinterpret_cast<v4i*>(dst + k * 1024) =
                v4i{(int)w[cls][k][0], (int)w[cls][k][1], (int)w[cls][k][2], (int)w[cls][k][3]};
        }
      } else {
        __syncthreads();
        // lane lrow + 32 g of block (l tile, ring step, slice) holds the 16
        // rings 16 g .. 16 g + 15 of the step at l value lrow; this half tile
        // fills the lanes with lrow in [16 hh, 16 hh + 16)
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
          const MixOff& o = cls ? o1 : o0;
          if (lt >= o.nlt) continue;
          const int nrs_here = min(8, o.nrs - rb * 8);
          signed char* base = T + (o.ta / kMixK * KS - tbase);
          for (int ch = tid; ch < nrs_here * KS * 32; ch += 256) {
            const int Lh = ch & 31, k = (ch >> 5) % KS, rsl = (ch >> 5) / KS;
            const int jr = Lh & 15, gg = Lh >> 4;
            const v4i val = *reinterpret_cast<const v4i*>(&stage[cls][k][jr][rsl * 32 + 16 * gg]);
            *reinterpret_cast<v4i*>(base + (((long long)lt * o.nrs + rb * 8 + rsl) * KS + k) * 1024 +
                                    (16 * hh + jr + 32 * gg) * 16) = val;
          }
        }
        __syncthreads();
      }
    }
  }
}

#ifndef FSHT_QUANT_CT_FAST
#define FSHT_QUANT_CT_FAST 1  // 0: the orders are the fastest grid dimension of the operand slicers
#endif

// ---------------------------------------------------------------------------
// Analysis operand: the m-major folded spectrum F (FP64, layout of k_gather,
// ncomp field components) -> FQ slices with the per-column exponents the
// gather left in FQe (index (2 m + class) * ncols + col).  Workgroup = (m,
// class, 8 ring steps, 32-column tile); thread = (column, ring step): 32 rows
// of one column.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_quant_F(const MixOff* __restrict__ mo, const int* __restrict__ i0tab,
                                                 const long long* __restrict__ goff, const double* __restrict__ F,
                                                 const int* __restrict__ FQe, signed char* __restrict__ FQ, int nhr,
                                                 int ncomp, int nsim, int nsimp) {
  const int ncols = 2 * ncomp * nsimp;
  const int m = FSHT_QUANT_CT_FAST ? blockIdx.z : blockIdx.x, cls = blockIdx.y & 1, kb = blockIdx.y >> 1;
  const int ct = FSHT_QUANT_CT_FAST ? blockIdx.x : blockIdx.z;
  if (!cols_active(ct * 32, ct * 32 + 32, nsim, nsimp)) return;  // no map in the tile
  const MixOff o = mo[2 * m + cls];
  const int cl = threadIdx.x & 31, q = threadIdx.x >> 5;
  const int rs = kb * 8 + q;
  if (rs >= o.nrs) return;
  const int R = nhr + 1 - i0tab[m];
  const int col = ct * 32 + cl;
  const int e = FQe[(long long)(2 * m + cls) * ncols + col];
  const double* src = F + goff[m] * ncomp / 2 + (long long)cls * R * ncols + col;
  double v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int r = rs * 32 + j;
    v[j] = r < R ? ldexp(src[(long long)r * ncols], -e) : 0.0;
  }
  signed char* dst = FQ + (long long)ncomp * o.fq + ((long long)ct * o.nrs + rs) * K * 1024 + cl * 16;
  put16<K>(v, dst);
  put16<K>(v + 16, dst + 32 * 16);
}

// ---------------------------------------------------------------------------
// Synthesis operand: packed alm -> AQ slices with one exponent per column
// (AQe, index (2 m + class) * ncols + col).  Columns as in the m-major
// buffer: per component c (alm0, then alm1 when ncomp = 2) the block
// [Re a_lm of the nsimp maps, Im a_lm of the nsimp maps].  Workgroup = (m,
// class, 32-column tile).  HAVE_EXP = 0: the column maxima first (setting
// AQe), then the digits (second read of the same alm, mostly from cache);
// HAVE_EXP = 1: AQe was set by the analysis that wrote alm.
// ---------------------------------------------------------------------------
template <int K, int HAVE_EXP>
__global__ __launch_bounds__(256) void k_quant_A(const MixOff* __restrict__ mo, const double* __restrict__ alm0,
                                                 const double* __restrict__ alm1, signed char* __restrict__ AQ,
                                                 int* __restrict__ AQe, int ncomp, int nsim, int nsimp, int lmax) {
  __shared__ double red[8][32];
  __shared__ int ex[32];
  const int m = FSHT_QUANT_CT_FAST ? blockIdx.z : blockIdx.x, cls = blockIdx.y;
  const int ct = FSHT_QUANT_CT_FAST ? blockIdx.x : blockIdx.z;
  if (!cols_active(ct * 32, ct * 32 + 32, nsim, nsimp)) return;  // no map in the tile
  const MixOff o = mo[2 * m + cls];
  const int cl = threadIdx.x & 31, q = threadIdx.x >> 5;
  const int col = ct * 32 + cl;
  const int ncols = 2 * ncomp * nsimp;
  const int comp = col / (2 * nsimp);
  const int rem = col - comp * 2 * nsimp;
  const int reim = rem >= nsimp ? 1 : 0;
  const int sim = rem - reim * nsimp;
  const bool ok = sim < nsim && !(m == 0 && reim);
  const long long L1 = lmax + 1;
  const double* src = (comp ? alm1 : alm0) + (ok ? sim : 0) + (long long)nsim * (reim ? (L1 * (L1 - m) - m) : (L1 * m));
  auto val = [&](int lst, int j) -> double {
    const int l = m + cls + 2 * (lst * 32 + j);
    return (ok && l <= lmax) ? src[(long long)nsim * l] : 0.0;
  };
  int* pe = AQe + (long long)(2 * m + cls) * ncols + col;
  if (!HAVE_EXP) {
    double mx = 0.0;
    for (int lst = q; lst < o.nlt; lst += 8)
#pragma unroll 8
      for (int j = 0; j < 32; ++j) mx = fmax(mx, fabs(val(lst, j)));
    red[q][cl] = mx;
    __syncthreads();
    if (q == 0) {
      double M = red[0][cl];
#pragma unroll
      for (int k = 1; k < 8; ++k) M = fmax(M, red[k][cl]);
      const int e = mix_exp(M);
      ex[cl] = e;
      *pe = e;
    }
    __syncthreads();
  }
  const int e = HAVE_EXP ? *pe : ex[cl];
  for (int lst = q; lst < o.nlt; lst += 8) {
    double v[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) v[j] = ldexp(val(lst, j), -e);
    signed char* dst = AQ + (long long)ncomp * o.aq + ((long long)ct * o.nlt + lst) * K * 1024 + cl * 16;
    put16<K>(v, dst);
    put16<K>(v + 16, dst + 32 * 16);
  }
}

// ---------------------------------------------------------------------------
// Slice GEMM building blocks
// ---------------------------------------------------------------------------
#ifndef FSHT_SLICE_MIX
#define FSHT_SLICE_MIX 1  // 0: a pass's 4-tile blocks and its last column tiles in separate launches
#endif
#ifndef FSHT_SLICE_PP
#define FSHT_SLICE_PP 1  // 0: the single-group slice GEMM for every slice count
#endif
#if FSHT_SLICE_MIX && !FSHT_SLICE_PP
#error "the merged work lists are launched by the two-group kernel"
#endif
#ifndef FSHT_SLICE_LDS
#define FSHT_SLICE_LDS 147456
#endif
constexpr int kLdsBudget = FSHT_SLICE_LDS;  // bytes of LDS the slice GEMM may use
#ifndef FSHT_SLICE_NSMAX
#define FSHT_SLICE_NSMAX 4
#endif
constexpr int kNsMax = FSHT_SLICE_NSMAX;  // stages of the operand ring, at most

// s_waitcnt with vmcnt <= N and the other counters unconstrained (gfx9 encoding)
template <int N>
__device__ inline void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}
// workgroup barrier that orders memory accesses in the compiler without
// waiting for outstanding loads (the pipeline's own vmcnt waits do that)
__device__ inline void raw_barrier() { asm volatile("s_barrier" ::: "memory"); }

template <int K>
__device__ inline double combine(const v16i (&D)[K], int r) {
  double v = 0.0;
#pragma unroll
  for (int d = K - 1; d >= 0; --d) v = fma((double)D[d][r], __builtin_ldexp(1.0, -7 * (d + 2)), v);
  return v;
}

// Arguments of the slice GEMM kernels.
struct SliceArgs {
  const LegItem* items;
  const MixOff* mo;
  const signed char* TP;  // left operand: the Legendre table (P, F1 or F2) in TA or TS order
  const int* TPe;         // its row exponents
  long long abase;        // the table's operand blocks sit at (o.ta or o.ts) / kMixK * ast - abase
  int ast;                // slices stored per contraction step of the table
  const signed char* OQ;  // right operand (FQ or AQ) and its column exponents
  const int* OQe;
  double* out0;  // analysis: alm of component 0 (T, or E); synthesis: the m-major buffer G
  double* out1;  // analysis, spin 2: the B alm
  const long long* goff;
  const int* i0tab;
  int nsim, nsimp, ncomp, lmax, nhr, nct;
  // 0: spin 0, or F1 (an output column is its input column); 1: F2 (the
  // input column of component c, part r feeds component 1 - c, part 1 - r)
  int fn;
  double pa, beta;
  int* aqe;  // analysis: column exponents of the result for the next synthesis operand (or NULL)
};

// Column map of the spin-2 contractions (QU2EB / EB2QU, src/sht_main.f90:383-436
// and 1079-1147 with the alpha/swap table of sht_data_init.f90:606-632).
// With columns [Q_re, Q_im, U_re, U_im] (analysis) or [E_re, E_im, B_re,
// B_im] (synthesis):
//   analysis  E_re = -pa F1.Q_re + pa F2.U_im   E_im = -pa F1.Q_im - pa F2.U_re
//             B_re = -pa F1.U_re - pa F2.Q_im   B_im = -pa F1.U_im + pa F2.Q_re
//   synthesis Q_re = -F1.E_re + F2.B_im         Q_im = -F1.E_im - F2.B_re
//             U_re = -F1.B_re - F2.E_im         U_im = -F1.B_im + F2.E_re
// so F1 keeps the column with factor -1 and F2 swaps both the component and
// Re/Im with factor +1 where they agree, -1 where they differ.
__device__ inline void spin2_map(int fn, int comp, int reim, int& comp_o, int& reim_o, double& sgn) {
  if (fn == 0) {
    comp_o = comp;
    reim_o = reim;
    sgn = -1.0;
  } else {
    comp_o = 1 - comp;
    reim_o = 1 - reim;
    sgn = comp == reim ? 1.0 : -1.0;
  }
}

// Epilogue of one wave's output tiles (row tile ta, column tiles tb0 ..
// tb0 + TPW - 1): value = 2^(eA(row) + eB(col)) * sum_d 2^-7(d+2) D_d.
//   SYN = 0: alm = pa * sgn * value + beta * alm (the column exponents of
//            the result are folded into aqe when given)
//   SYN = 1: G = sgn * value (+ G when beta != 0), in the fold of the table's
//            parity (F2: the other fold)
template <int K, int SYN, int TPW>
__device__ inline void slice_epilogue(const v16i (&D0)[K], const v16i (&D1)[K], const SliceArgs& A, const MixOff& o,
                                      int m, int cls, int bcls, int ta, int tb0, int L) {
  const int ncols = 2 * A.ncomp * A.nsimp;
  const int h = L >> 5;
  const int* pe = A.TPe + (SYN ? o.tse : o.tae) + (long long)ta * 32 + 16 * h;  // accumulator order
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tb = tb0 + t;
    if (tb >= A.nct) break;
    const int col = tb * 32 + (L & 31);
    const int eB = A.OQe[(long long)(2 * m + bcls) * ncols + col];
    const int comp = col / (2 * A.nsimp);
    const int rem = col - comp * 2 * A.nsimp;
    const int reim = rem >= A.nsimp ? 1 : 0;
    const int sim = rem - reim * A.nsimp;
    int comp_o = comp, reim_o = reim;
    double sgn = 1.0;
    if (A.ncomp == 2) spin2_map(A.fn, comp, reim, comp_o, reim_o, sgn);
    if (!SYN) {
      if (sim >= A.nsim || (m == 0 && reim_o)) continue;
      const long long L1 = A.lmax + 1;
      // this column's entries: sim + nsim (l + L1 m) for Re, sim + nsim (l - m + L1 (L1 - m)) for Im
      double* oc = (comp_o ? A.out1 : A.out0) + sim + (long long)A.nsim * (reim_o ? L1 * (L1 - m) - m : L1 * m);
      // the previous values (beta != 0) are all loaded before any store, so
      // the read-modify-write costs one memory latency, not sixteen
      double prev[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int l = m + cls + 2 * (ta * 32 + drow(r, h));
        prev[r] = (A.beta != 0.0 && l <= A.lmax) ? oc[(long long)A.nsim * l] : 0.0;
      }
      const double f = A.pa * sgn;
      double mx = 0.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int l = m + cls + 2 * (ta * 32 + drow(r, h));
        if (l > A.lmax) continue;
        const double v = f * __builtin_ldexp(combine<K>(t ? D1 : D0, r), pe[r] + eB) + A.beta * prev[r];
        oc[(long long)A.nsim * l] = v;
        mx = fmax(mx, fabs(v));
      }
      // the synthesis operand's exponent of this output column: max over
      // every workgroup writing it (k_quant_A<K, 1> reads it)
      if (A.aqe) {
        const int col_o = comp_o * 2 * A.nsimp + reim_o * A.nsimp + sim;
        atomicMax(A.aqe + (long long)(2 * m + cls) * ncols + col_o, mix_exp(mx));
      }
    } else {
      const int R = A.nhr + 1 - A.i0tab[m];
      const int fold = A.fn ? 1 - cls : cls;
      const int col_o = comp_o * 2 * A.nsimp + reim_o * A.nsimp + sim;
      double* g = A.out0 + A.goff[m] * A.ncomp / 2 + (long long)fold * R * ncols + col_o;
      double prev[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = ta * 32 + drow(r, h);
        prev[r] = (A.beta != 0.0 && row < R) ? g[(long long)row * ncols] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = ta * 32 + drow(r, h);
        if (row < R) g[(long long)row * ncols] = sgn * __builtin_ldexp(combine<K>(t ? D1 : D0, r), pe[r] + eB) + prev[r];
      }
    }
  }
}

// Source of the 1 KB operand block `b` of a stage at step s0 (A tiles
// first: NA * K blocks, then B tiles: NB * K blocks) and its per-step stride.
template <int K, int SYN, int NA, int NB>
__device__ inline void slice_src(const SliceArgs& A, const MixOff& o, const MixOff& ob, const LegItem& it, int nsteps,
                                 int ntA, int s0, int b, const signed char*& src, int& step) {
  if (b < NA * K) {
    const int tile = min(it.blk * NA + b / K, ntA - 1);
    const signed char* baseA = A.TP + (SYN ? o.ts : o.ta) / kMixK * A.ast - A.abase;
    src = baseA + (((long long)tile * nsteps + s0) * A.ast + b % K) * 1024;
    step = A.ast * 1024;
  } else {
    const int bb = b - NA * K;
    const int tile = min(it.colblk * NB + bb / K, A.nct - 1);
    const signed char* baseB = A.OQ + (long long)A.ncomp * (SYN ? ob.aq : ob.fq);
    src = baseB + (((long long)tile * nsteps + s0) * K + bb % K) * 1024;
    step = K * 1024;
  }
}

// ---------------------------------------------------------------------------
// The slice GEMM.  Workgroup = 8 waves (two per SIMD) on (m, class, NA row
// tiles, NB column tiles): NB = 4 with NA = 4 (wave w owns row tile w/2 and
// column tiles 2(w%2), 2(w%2)+1), or NB = 1 with NA = 8 (wave w owns row tile
// w and the column tile) for the last column tiles of a batch whose tile
// count is not a multiple of 4.  Per 32-wide step of the contraction the
// workgroup needs (NA + NB) K blocks of 1 KB (K slices of its A and B tiles);
// the waves copy them global -> LDS with direct-to-LDS loads
// (global_load_lds_dwordx4: 1 KB per wave instruction, lane order kept) into a
// ring of NS stages, NS - 1 steps ahead of the MFMAs.  Each step every wave
// waits for its own loads of the step (vmcnt), one barrier publishes them,
// the stage that every wave finished with is refilled, and the wave reads its
// operands from LDS and issues K(K+1)/2 digit products per B tile into one
// int32 accumulator per digit sum.
//   SYN = 0 (analysis): A = TA (rows l), B = FQ (columns), out = alm
//   SYN = 1 (synthesis): A = TS (rows ring), B = AQ (columns), out = G
// For F2 (A.fn = 1) the analysis reads the operand of the other fold: F2 of
// parity class c pairs with the N -/+ S fold of class 1 - c.
// ---------------------------------------------------------------------------
template <int K, int SYN, int NB>
__global__ __launch_bounds__(512, 1) void k_slice_gemm(const SliceArgs A) {
  static_assert(NB == 4 || NB == 1, "column tiles per workgroup");
  constexpr int NW = 8;                                 // waves
  constexpr int NA = NB == 4 ? 4 : 8;                   // row tiles
  constexpr int TPW = NB == 4 ? 2 : 1;                  // column tiles per wave
  constexpr int NBLK = (NA + NB) * K;                   // 1 KB blocks per step
  constexpr int PER = (NBLK + NW - 1) / NW;             // loads per wave per step (the last may repeat)
  constexpr int NS = kLdsBudget / (NBLK * 1024) < kNsMax ? kLdsBudget / (NBLK * 1024) : kNsMax;  // stages
  static_assert(NS >= 3, "pipeline needs three stages");
  __shared__ __align__(16) signed char st[NS][NBLK * 1024];
  const LegItem it = A.items[blockIdx.x];
  if (it.m < 0) return;  // whole workgroup
  if (!cols_active(it.colblk * NB * 32, (it.colblk + 1) * NB * 32, A.nsim, A.nsimp)) return;  // no map in the block
  const int m = it.m, cls = it.par;
  const int bcls = (!SYN && A.fn) ? 1 - cls : cls;  // class (fold) of the right operand
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int L = threadIdx.x & 63;
  const MixOff o = A.mo[2 * m + cls];
  const MixOff ob = A.mo[2 * m + bcls];
  const int nsteps = SYN ? o.nlt : o.nrs;
  const int ntA = SYN ? o.nrs : o.nlt;

  const int s0 = it.r0;
  // this wave's PER source streams: wave-uniform pointers at step s0 that
  // advance by one step's stride per issue, and their blocks within a stage
  const signed char* gsrc[PER];
  int gstep[PER], lofs[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    int b = w + NW * k;
    if (b >= NBLK) b -= NW;  // this wave's previous block again (same data, same place)
    lofs[k] = b * 1024;
    slice_src<K, SYN, NA, NB>(A, o, ob, it, nsteps, ntA, s0, b, gsrc[k], gstep[k]);
  }
  int fill = 0;  // stage the next issue fills (steps s0, s0 + 1, ... go to stages 0, 1, ..., NS - 1, 0, ...)
  auto issue = [&]() {
    signed char* stage = st[fill];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      __builtin_amdgcn_global_load_lds((const void*)(gsrc[k] + L * 16), (lds_ptr)(stage + lofs[k]), 16, 0, 0);
      gsrc[k] += gstep[k];
    }
    fill = fill + 1 == NS ? 0 : fill + 1;
  };

  const int wa = NB == 4 ? (w >> 1) : w;         // this wave's row tile within the workgroup
  const int wb = NB == 4 ? (w & 1) * 2 : 0;      // its first column tile within the workgroup
  const int ta = it.blk * NA + wa;
  const int tb0 = it.colblk * NB + wb;
  const bool act = ta < ntA && tb0 < A.nct;
  v16i D0[K], D1[K];
#pragma unroll
  for (int d = 0; d < K; ++d) D0[d] = D1[d] = v16i{};
#pragma unroll
  for (int k = 0; k < NS - 1; ++k)
    if (s0 + k < nsteps) issue();
  int use = 0;  // stage holding step s
  for (int s = s0; s < nsteps; ++s) {
    // this wave's loads of step s are done once at most the later issued
    // steps (up to NS - 2 of them) are outstanding
    const int ahead = min(NS - 2, nsteps - 1 - s);
    if (NS > 3 && ahead >= 2) wait_vm<(NS > 3 ? 2 : 1) * PER>();
    else if (ahead >= 1) wait_vm<PER>();
    else wait_vm<0>();
    raw_barrier();
    const signed char* sb = st[use];
    use = use + 1 == NS ? 0 : use + 1;
    // A slices stay in registers for the step; B slices are read one at a
    // time (slice j of the B tiles meets A slices 0 .. K-1-j)
    const signed char* pb = sb + (NA * K + wb * K) * 1024 + L * 16;
    v4i a[K];
    if (act) {
#pragma unroll
      for (int sl = 0; sl < K; ++sl) a[sl] = *reinterpret_cast<const v4i*>(sb + (wa * K + sl) * 1024 + L * 16);
    }
    if (s + NS - 1 < nsteps) issue();  // into the stage read in step s - 1
    if (act) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const v4i b0 = *reinterpret_cast<const v4i*>(pb + j * 1024);
        v4i b1 = b0;
        if (TPW == 2) b1 = *reinterpret_cast<const v4i*>(pb + (K + j) * 1024);
#pragma unroll
        for (int i = 0; i < K - j; ++i) {
          D0[i + j] = mfma_i8(a[i], b0, D0[i + j]);
          if (TPW == 2) D1[i + j] = mfma_i8(a[i], b1, D1[i + j]);
        }
      }
    }
  }
  if (!act) return;
  slice_epilogue<K, SYN, TPW>(D0, D1, A, o, m, cls, bcls, ta, tb0, L);
}

// ---------------------------------------------------------------------------
// The slice GEMM with two wave groups in alternation ("ping-pong").  Same
// work items, tiles and operand streams as k_slice_gemm, but the 8 waves
// form two groups of 4 (waves 0-3 and 4-7, one wave of each group per SIMD)
// whose phases alternate: while one group issues its step's MFMAs from
// registers, the other reads its next operands from LDS, so the matrix pipe
// of every SIMD is fed by one group or the other and the LDS reads no longer
// sit between a barrier and the first MFMA.  Group 0 issues every
// direct-to-LDS load.
// ---------------------------------------------------------------------------
// s_waitcnt lgkmcnt(0) with the other counters unconstrained (gfx9 encoding)
__device__ inline void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xF | (0x3 << 14) | (0x7 << 4)); }

// Shape of the two-group kernel's pipeline for K slices and NB column tiles.
template <int K, int NB>
struct PpShape {
  static_assert(NB == 4 || NB == 1, "column tiles per workgroup");
  static constexpr int NG = 4;                                 // waves per group
  static constexpr int NA = NB == 4 ? 4 : 8;                   // row tiles
  static constexpr int TPW = NB == 4 ? 2 : 1;                  // column tiles per wave
  static constexpr int NBLK = (NA + NB) * K;                   // 1 KB blocks per step
  static constexpr int PER = (NBLK + NG - 1) / NG;             // loads per group-0 wave per step (the last may repeat)
  static constexpr int NS = kLdsBudget / (NBLK * 1024) < kNsMax ? kLdsBudget / (NBLK * 1024) : kNsMax;  // stages
  static constexpr int BYTES = NS * NBLK * 1024;               // LDS of the stage ring
  static_assert(NS >= 3, "pipeline needs three stages");
  static_assert((NS - 1) * PER < 64, "vmcnt range");
};

template <int K, int SYN, int NB>
__device__ __forceinline__ void slice_gemm_pp_body(const SliceArgs& A, const LegItem it,
                                                   signed char* __restrict__ lds) {
  using S = PpShape<K, NB>;
  constexpr int NG = S::NG, NA = S::NA, TPW = S::TPW, NBLK = S::NBLK, PER = S::PER, NS = S::NS;
  auto st = [&](int stage) { return lds + stage * (NBLK * 1024); };
  if (it.m < 0) return;  // whole workgroup
  if (!cols_active(it.colblk * NB * 32, (it.colblk + 1) * NB * 32, A.nsim, A.nsimp)) return;  // no map in the block
  const int m = it.m, cls = it.par;
  const int bcls = (!SYN && A.fn) ? 1 - cls : cls;  // class (fold) of the right operand
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int grp = w >> 2;
  const int L = threadIdx.x & 63;
  const MixOff o = A.mo[2 * m + cls];
  const MixOff ob = A.mo[2 * m + bcls];
  const int nsteps = SYN ? o.nlt : o.nrs;
  const int ntA = SYN ? o.nrs : o.nlt;
  const int s0 = it.r0;

  // group 0's source streams (wave-uniform pointers at step s0, advanced by
  // one step's stride per issue) and their blocks within a stage
  const signed char* gsrc[PER];
  int gstep[PER], lofs[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    int b = (w & 3) + NG * k;
    if (b >= NBLK) b -= NG;  // this wave's previous block again (same data, same place)
    lofs[k] = b * 1024;
    slice_src<K, SYN, NA, NB>(A, o, ob, it, nsteps, ntA, s0, b, gsrc[k], gstep[k]);
  }
  auto issue = [&](int stage) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      __builtin_amdgcn_global_load_lds((const void*)(gsrc[k] + L * 16), (lds_ptr)(st(stage) + lofs[k]), 16, 0, 0);
      gsrc[k] += gstep[k];
    }
  };
  // group 0: wait until the loads of step `s` are done, given that the
  // loads of the steps up to `last` were issued (at most NS - 2 after s)
  auto wait_step = [&](int s, int last) {
    const int ahead = min(NS - 2, last - s);
    if (NS > 3 && ahead >= 2) wait_vm<(NS > 3 ? 2 : 1) * PER>();
    else if (ahead >= 1) wait_vm<PER>();
    else wait_vm<0>();
  };

  const int wa = NB == 4 ? (w >> 1) : w;         // this wave's row tile within the workgroup
  const int wb = NB == 4 ? (w & 1) * 2 : 0;      // its first column tile within the workgroup
  const int ta = it.blk * NA + wa;
  const int tb0 = it.colblk * NB + wb;
  const bool act = ta < ntA && tb0 < A.nct;
  v16i D0[K], D1[K];
#pragma unroll
  for (int d = 0; d < K; ++d) D0[d] = D1[d] = v16i{};
  v4i ra[K], rb[TPW][K];
  auto load_regs = [&](int stage) {
    const signed char* sb = st(stage) + L * 16;
#pragma unroll
    for (int sl = 0; sl < K; ++sl) ra[sl] = *reinterpret_cast<const v4i*>(sb + (wa * K + sl) * 1024);
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int sl = 0; sl < K; ++sl) rb[t][sl] = *reinterpret_cast<const v4i*>(sb + (NA * K + (wb + t) * K + sl) * 1024);
  };
  auto mma = [&]() {
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < K - j; ++i) {
        D0[i + j] = mfma_i8(ra[i], rb[0][j], D0[i + j]);
        if (TPW == 2) D1[i + j] = mfma_i8(ra[i], rb[TPW - 1][j], D1[i + j]);
      }
  };

  // Both groups run the same loop, MEM(t) then MMA(t) per step, group 1
  // one phase behind group 0 (it starts with an idle phase and skips the
  // barrier after its last MMA phase, so both pass the same barriers).
  //   MEM(t): read stage t into registers; group 0 refills stage t - 1
  //           (read by both groups by now) with step t - 1 + NS.
  //   MMA(t): the step's MFMAs; group 0 then waits for its loads of t + 1.
  // Group 0's MEM(t) runs beside group 1's MMA(t - 1), its MMA(t) beside
  // group 1's MEM(t).  Stage t is published by the barrier that ends group
  // 0's MMA(t - 1) (the prologue barrier for t = s0); group 1 completes its
  // LDS reads (lgkmcnt 0) before each barrier, so no refill overtakes them.
  // An item without steps passes no barrier at all (s0 is uniform).
  if (s0 < nsteps) {
    if (grp == 0) {
#pragma unroll
      for (int k = 0; k < NS; ++k)
        if (s0 + k < nsteps) issue(k);
      wait_step(s0, min(s0 + NS - 1, nsteps - 1));
    }
    sched_fence();
    raw_barrier();
    if (grp == 1) raw_barrier();
    sched_fence();
  }
  int cur = 0, prev = NS - 1;  // stages of steps t and t - 1
  for (int t = s0; t < nsteps; ++t) {
    if (act) load_regs(cur);
    if (grp == 0 && t > s0 && t - 1 + NS < nsteps) issue(prev);
    wait_lgkm0();
    sched_fence();
    raw_barrier();
    sched_fence();
    if (act) mma();
    if (grp == 0 && t + 1 < nsteps) wait_step(t + 1, min(t + NS - 1, nsteps - 1));
    sched_fence();
    if (grp == 0 || t + 1 < nsteps) raw_barrier();
    sched_fence();
    prev = cur;
    cur = cur + 1 == NS ? 0 : cur + 1;
  }
  if (!act) return;
  slice_epilogue<K, SYN, TPW>(D0, D1, A, o, m, cls, bcls, ta, tb0, L);
}

template <int K, int SYN, int NB>
__global__ __launch_bounds__(512, 1) void k_slice_gemm_pp(const SliceArgs A) {
  __shared__ __align__(16) signed char st[PpShape<K, NB>::BYTES];
  slice_gemm_pp_body<K, SYN, NB>(A, A.items[blockIdx.x], st);
}

// One launch for a pass's 4-tile blocks and its single remainder tiles
// (colblk >= kSingleTile: the tile colblk - kSingleTile, 8 row tiles), in one
// list ordered by m: the remainder items of an order run beside its 4-tile
// blocks and find its P rows in cache instead of streaming them again.
constexpr int kSingleTile = 1 << 20;

template <int K, int SYN>
__global__ __launch_bounds__(512, 1) void k_slice_gemm_pp_mix(const SliceArgs A) {
  constexpr int B4 = PpShape<K, 4>::BYTES, B1 = PpShape<K, 1>::BYTES;
  __shared__ __align__(16) signed char st[B4 > B1 ? B4 : B1];
  LegItem it = A.items[blockIdx.x];
  if (it.colblk >= kSingleTile) {
    it.colblk -= kSingleTile;
    slice_gemm_pp_body<K, SYN, 1>(A, it, st);
  } else {
    slice_gemm_pp_body<K, SYN, 4>(A, it, st);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

static void free_lists(std::vector<LegItem*>& v) {
  for (LegItem* p : v)
    if (p) (void)hipFree(p);
  v.clear();
}

void mixed_free(Handle* h) {
  void* ptrs[] = {h->d_mix_off, h->d_TA,     h->d_TS,     h->d_AQ,     h->d_TAe,    h->d_TSe,     h->d_FQe,
                  h->d_AQe,     h->d_firstA, h->d_firstS, h->d_pitems, h->d_T2[0], h->d_T2[1], h->d_TAe2[0],
                  h->d_TAe2[1], h->d_TSe2[0], h->d_TSe2[1], h->d_firstA2[0], h->d_firstA2[1], h->d_firstS2[0],
                  h->d_firstS2[1]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->d_FQ && !h->fq_alias) (void)hipFree(h->d_FQ);
  for (int k = 0; k < 3; ++k) {
    if (h->d_anaq_items[k]) (void)hipFree(h->d_anaq_items[k]);
    if (h->d_synq_items[k]) (void)hipFree(h->d_synq_items[k]);
    h->d_anaq_items[k] = h->d_synq_items[k] = nullptr;
    h->n_anaq_items[k] = h->n_synq_items[k] = 0;
  }
  free_lists(h->d_anaq2);
  free_lists(h->d_synq2);
  h->n_anaq2.clear();
  h->n_synq2.clear();
  h->d_mix_off = nullptr;
  h->d_TA = h->d_TS = h->d_FQ = h->d_AQ = nullptr;
  h->d_TAe = h->d_TSe = h->d_FQe = h->d_AQe = nullptr;
  h->d_firstA = h->d_firstS = nullptr;
  h->d_pitems = nullptr;
  for (int f = 0; f < 2; ++f) {
    h->d_T2[f] = nullptr;
    h->d_TAe2[f] = h->d_TSe2[f] = nullptr;
    h->d_firstA2[f] = h->d_firstS2[f] = nullptr;
  }
  h->fq_alias = 0;
  h->mix_ready = h->mix0_ready = h->mix2_ready = 0;
  h->t2_lay = -1;
  h->mix_bytes = 0;
}

static hipError_t fill_int(Handle* h, int* p, long long n, int v) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks), dim3(256), 0, h->stream, p, n, v);
  return hipGetLastError();
}

// Slice-GEMM work lists of the orders [m0, m1) for a batch of nct 32-column
// tiles: [0] full blocks of 4 column tiles (4 row tiles each), [1] the
// remaining column tiles one by one (8 row tiles each), [2] the remaining
// column tiles as one padded block of 4 (six slices).  Every (m, class, row
// tiles, column tiles) writes its outputs, starting at the first step with a
// nonzero table digit of its row tiles.
#if FSHT_SLICE_MIX
// A list of extra items (`rows` times the row tiles of a 4-tile item, colblk
// tagged with `tag`) merged into the 4-tile list, both in (m, class, row
// block) order: each extra item follows the 4-tile items of its row blocks.
static void merge_after(const std::vector<LegItem>& full, const std::vector<double>& full_cost,
                        std::vector<LegItem>& extra, std::vector<double>& extra_cost, int rows, int tag) {
  std::vector<LegItem> out;
  std::vector<double> cost;
  auto key = [](const LegItem& t, int r) { return std::make_tuple(t.m, t.par, t.blk / r); };
  size_t j = 0;
  auto take = [&]() {
    LegItem t = extra[j];
    t.colblk += tag;
    out.push_back(t);
    cost.push_back(extra_cost[j++]);
  };
  for (size_t i = 0; i < full.size(); ++i) {
    while (j < extra.size() && key(extra[j], 1) < key(full[i], rows)) take();
    out.push_back(full[i]);
    cost.push_back(full_cost[i]);
  }
  while (j < extra.size()) take();
  extra.swap(out);
  extra_cost.swap(cost);
}
#endif

static hipError_t build_q_lists(Handle* h, int m0, int m1, int nct, const std::vector<int>& firstA,
                                const std::vector<int>& firstS, LegItem** ana_out, int* nana, LegItem** syn_out,
                                int* nsyn) {
  const int nfull = nct / 4, nrem = nct % 4;
#if FSHT_SLICE_MIX
  std::vector<LegItem> ana0, syn0;
  std::vector<double> ana0_cost, syn0_cost;
#endif
  for (int k = 0; k < 3; ++k) {
    ana_out[k] = syn_out[k] = nullptr;
    nana[k] = nsyn[k] = 0;
  }
  for (int k = 0; k < 3; ++k) {
    if (k > 0 && nrem == 0) break;
    const int NA = k == 1 ? 8 : 4;
    std::vector<LegItem> ana, syn;
    std::vector<double> ana_cost, syn_cost;
    // column blocks of this list (colblk counts blocks of NB tiles) and their tile counts
    std::vector<int> cbs, cnt;
    if (k == 0)
      for (int cb = 0; cb < nfull; ++cb) cbs.push_back(cb), cnt.push_back(4);
    else if (k == 1)
      for (int r = 0; r < nrem; ++r) cbs.push_back(4 * nfull + r), cnt.push_back(1);
    else
      cbs.push_back(nfull), cnt.push_back(4);
    for (int m = m0; m < m1; ++m)
      for (int c = 0; c < 2; ++c) {
        const MixOff& o = h->mix_off[2 * m + c];
        for (int b = 0; NA * b < o.nlt; ++b) {
          int r0 = o.nrs;
          for (int t = NA * b; t < std::min(o.nlt, NA * b + NA); ++t) r0 = std::min(r0, firstA[o.fa + t]);
          for (size_t q = 0; q < cbs.size(); ++q) {
            ana.push_back(LegItem{m, c, b, cbs[q], r0});
            ana_cost.push_back(1.0 + (double)(o.nrs - r0) * NA * cnt[q]);
          }
        }
        for (int b = 0; NA * b < o.nrs; ++b) {
          int l0 = o.nlt;
          for (int t = NA * b; t < std::min(o.nrs, NA * b + NA); ++t) l0 = std::min(l0, firstS[o.fs + t]);
          for (size_t q = 0; q < cbs.size(); ++q) {
            syn.push_back(LegItem{m, c, b, cbs[q], l0});
            syn_cost.push_back(1.0 + (double)(o.nlt - l0) * NA * cnt[q]);
          }
        }
      }
#if FSHT_SLICE_MIX
    // lists 1 and 2 also carry the 4-tile blocks: a pass is one launch
    if (k == 0) {
      ana0 = ana, syn0 = syn, ana0_cost = ana_cost, syn0_cost = syn_cost;
    } else {
      merge_after(ana0, ana0_cost, ana, ana_cost, k == 1 ? 2 : 1, k == 1 ? kSingleTile : 0);
      merge_after(syn0, syn0_cost, syn, syn_cost, k == 1 ? 2 : 1, k == 1 ? kSingleTile : 0);
    }
#endif
    hipError_t e = upload_items(ana, ana_cost, &ana_out[k], &nana[k]);
    if (e == hipSuccess) e = upload_items(syn, syn_cost, &syn_out[k], &nsyn[k]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Host plan of the mixed-precision state (no device memory): per (m, class)
// the offsets into the tables and the sliced operands (room for the handle's
// field components), the (m, ring block) items of the table kernels and the
// totals.  Shared by spin 0 and spin 2.
struct MixPlan {
  std::vector<MixOff> off;
  std::vector<int> pitem_off;
  std::vector<int2> pitems;
  long long ta = 0, fq = 0, aq = 0, tae = 0, tse = 0;
  int fa = 0, fs = 0, nrb_max = 0;
};

static void plan_offsets(const Handle* h, MixPlan& P) {
  const int lmax = h->lmax, nm = lmax + 1;
  const int nct = h->nsimp / 16;  // 32-column tiles of the 2*nsimp columns of one component
  const int ncols = 2 * h->ncomp_ws * h->nsimp;
  P.off.assign(2 * (size_t)nm, MixOff{});
  P.pitem_off.assign((size_t)nm + 1, 0);
  for (int m = 0; m <= lmax; ++m) {
    const int R = h->nhr + 1 - h->i0[m];
    const int nrs = cdiv(R, 32), nrb = cdiv(nrs, 8);
    P.nrb_max = std::max(P.nrb_max, nrb);
    P.pitem_off[m] = (int)P.pitems.size();
    for (int rb = 0; rb < nrb; ++rb) P.pitems.push_back(make_int2(m, rb));
    for (int c = 0; c < 2; ++c) {
      const int nval = (lmax >= m + c) ? (lmax - m - c) / 2 + 1 : 0;
      const int nlt = cdiv(nval, 32);
      MixOff& o = P.off[2 * m + c];
      o.nrs = nrs;
      o.nlt = nlt;
      o.ta = o.ts = P.ta;
      P.ta += (long long)nlt * nrs * kMixK * 1024;
      o.fq = P.fq;
      P.fq += (long long)nct * nrs * kMixK * 1024;
      o.aq = P.aq;
      P.aq += (long long)nct * nlt * kMixK * 1024;
      o.tae = P.tae;
      P.tae += (long long)nlt * 32;
      o.tse = P.tse;
      P.tse += (long long)nrs * 32;
      o.fqe = o.aqe = (long long)(2 * m + c) * ncols;
      o.fa = P.fa;
      P.fa += nlt;
      o.fs = P.fs;
      P.fs += nrs;
    }
  }
  P.pitem_off[nm] = (int)P.pitems.size();
}

// The analysis operand FQ lives in the map-major spectrum buffer when it
// fits there: that buffer only holds data between a transform's ring FFTs
// and its first gather (analysis) or between the fold and the inverse FFTs
// (synthesis), while FQ only lives between a gather and the analysis.
static bool fq_fits_spec(const Handle* h, long long fq_bytes) {
  return fq_bytes <= (long long)h->ncomp_ws * h->nsim_cap * h->npix * 8;
}

static hipError_t mixed_setup_common(Handle* h) {
  if (h->mix_ready) return hipSuccess;
  const int nm = h->lmax + 1;
  const int ncw = h->ncomp_ws;
  const int ncols = 2 * ncw * h->nsimp;
  MixPlan P;
  plan_offsets(h, P);
  h->mix_off = std::move(P.off);
  h->pitem_off = std::move(P.pitem_off);
  const std::vector<int2>& pitems = P.pitems;
  const long long fq = P.fq, aq = P.aq;
  h->mix_nrb_max = P.nrb_max;
  h->mix_ta_total = P.ta;
  h->mix_tae = P.tae;
  h->mix_tse = P.tse;
  h->mix_fa = P.fa;
  h->mix_fs = P.fs;
  const long long nexp = 2LL * nm * ncols;
  const long long fq_bytes = (long long)ncw * fq;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, long long bytes) {
    if (e == hipSuccess) e = hipMalloc(p, (size_t)std::max(bytes, 256LL));
  };
  if (fq_fits_spec(h, fq_bytes)) {
    h->d_FQ = reinterpret_cast<signed char*>(h->ws_spec);
    h->fq_alias = 1;
  } else {
    al((void**)&h->d_FQ, fq_bytes);
  }
  al((void**)&h->d_AQ, (long long)ncw * aq);
  al((void**)&h->d_FQe, nexp * 4);
  al((void**)&h->d_AQe, nexp * 4);
  al((void**)&h->d_mix_off, (long long)h->mix_off.size() * (long long)sizeof(MixOff));
  al((void**)&h->d_pitems, (long long)pitems.size() * (long long)sizeof(int2));
  if (e == hipSuccess)
    e = hipMemcpy(h->d_mix_off, h->mix_off.data(), h->mix_off.size() * sizeof(MixOff), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_pitems, pitems.data(), pitems.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    mixed_free(h);
    return e;
  }
  h->mix_bytes = (h->fq_alias ? 0 : fq_bytes) + (long long)ncw * aq + 8 * nexp;
  h->mix_ready = 1;
  return hipSuccess;
}

// Spin 0: the P tables in both operand orders (six slices), resident.
static hipError_t mixed_setup_spin0(Handle* h) {
  if (h->mix0_ready) return hipSuccess;
  const int lmax = h->lmax;
  const long long ta = h->mix_ta_total, tae = h->mix_tae, tse = h->mix_tse;
  const int fa = h->mix_fa, fs = h->mix_fs;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, long long bytes) {
    if (e == hipSuccess) e = hipMalloc(p, (size_t)std::max(bytes, 256LL));
  };
  al((void**)&h->d_TA, ta);
  al((void**)&h->d_TS, ta);
  al((void**)&h->d_TAe, tae * 4);
  al((void**)&h->d_TSe, tse * 4);
  al((void**)&h->d_firstA, (long long)fa * 4);
  al((void**)&h->d_firstS, (long long)fs * 4);
  if (e == hipSuccess) e = fill_int(h, h->d_firstA, fa, kNeverStep);
  if (e == hipSuccess) e = fill_int(h, h->d_firstS, fs, kNeverStep);
  if (e == hipSuccess) e = fill_int(h, h->d_TAe, tae, kMixZeroE);
  const dim3 pgrid((unsigned)h->pitem_off[lmax + 1]);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ptab<0>, pgrid, dim3(256), 0, h->stream, h->d_pitems, h->d_geo, h->d_i0, h->d_ab, h->d_aboff,
                       h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_mix_off, h->d_TA, h->d_TAe, h->d_TS, h->d_TSe,
                       h->d_firstA, h->d_firstS, lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ptab<1>, pgrid, dim3(256), 0, h->stream, h->d_pitems, h->d_geo, h->d_i0, h->d_ab, h->d_aboff,
                       h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_mix_off, h->d_TA, h->d_TAe, h->d_TS, h->d_TSe,
                       h->d_firstA, h->d_firstS, lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  std::vector<int> firstA(fa), firstS(fs);
  if (e == hipSuccess) e = hipMemcpy(firstA.data(), h->d_firstA, (size_t)fa * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(firstS.data(), h->d_firstS, (size_t)fs * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = build_q_lists(h, 0, lmax + 1, h->nsimp / 16, firstA, firstS, h->d_anaq_items, h->n_anaq_items,
                      h->d_synq_items, h->n_synq_items);
  if (e != hipSuccess) return e;  // the caller frees the mixed-precision state
  h->mix_bytes += 2 * ta + 4 * (tae + tse + fa + fs);
  h->mix0_ready = 1;
  return hipSuccess;
}

// Spin 2: the F1/F2 exponents and first nonzero steps (resident), the plan of
// m ranges whose tables are generated per pass, the chunk buffers and the
// work lists per (function, range).  A range's tables (one function, six
// slices) take at most min(kT2Cap, max(kT2Min, 1/16 of the handle's FP64
// workspace)) bytes: a handle sized for few maps (a serving lane) keeps small
// buffers; FSHT_T2_CAP_MB lowers the limit further (e.g. to exercise many
// ranges on small grids).
constexpr long long kT2Cap = 6LL << 30;
constexpr long long kT2Min = 256LL << 20;
static long long t2_cap_limit(const Handle* h) {
  const long long ws = 8LL * (2LL * h->nsim_cap * h->npix + 2LL * h->npix * h->nsimp + 2LL * h->G_size);
  long long cap = std::min(kT2Cap, std::max(kT2Min, ws / 16));
  const char* v = std::getenv("FSHT_T2_CAP_MB");
  const long long mb = v ? std::atoll(v) : 0;
  return mb > 0 ? std::min(cap, mb << 20) : cap;
}

// m ranges of the spin-2 tables: whole orders, each range's tables (one
// function, one operand order, six slices) within cap_limit bytes unless a
// single order exceeds it.  bounds receives the first order of every range
// and lmax + 1; returns the bytes of the largest range.
static long long plan_t2_ranges(const std::vector<MixOff>& off, long long ta_total, int lmax, long long cap_limit,
                                std::vector<int>& bounds) {
  bounds.assign(1, 0);
  long long cap = 0;
  for (int m = 0; m <= lmax;) {
    int m1 = m;
    const long long start = off[2 * m].ta;
    while (m1 <= lmax && (m1 == m || (m1 + 1 <= lmax ? off[2 * (m1 + 1)].ta : ta_total) - start <= cap_limit)) ++m1;
    const long long end = m1 <= lmax ? off[2 * m1].ta : ta_total;
    cap = std::max(cap, end - start);
    bounds.push_back(m1);
    m = m1;
  }
  return cap;
}

#define FSHT_FTAB_COMMON \
  h->d_geo, h->d_i0, h->d_ab, h->d_spin, h->d_aboff, h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_mix_off

static hipError_t mixed_setup_spin2(Handle* h) {
  if (h->mix2_ready) return hipSuccess;
  const int lmax = h->lmax;
  const long long tae = h->mix_tae, tse = h->mix_tse;
  const int fa = h->mix_fa, fs = h->mix_fs;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, long long bytes) {
    if (e == hipSuccess) e = hipMalloc(p, (size_t)std::max(bytes, 256LL));
  };
  for (int f = 0; f < 2; ++f) {
    al((void**)&h->d_TAe2[f], tae * 4);
    al((void**)&h->d_TSe2[f], tse * 4);
    al((void**)&h->d_firstA2[f], (long long)fa * 4);
    al((void**)&h->d_firstS2[f], (long long)fs * 4);
    if (e == hipSuccess) e = fill_int(h, h->d_firstA2[f], fa, kNeverStep);
    if (e == hipSuccess) e = fill_int(h, h->d_firstS2[f], fs, kNeverStep);
    if (e == hipSuccess) e = fill_int(h, h->d_TAe2[f], tae, kMixZeroE);
  }
  const dim3 pgrid((unsigned)h->pitem_off[lmax + 1]);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ftab_exp<0>, pgrid, dim3(256), 0, h->stream, h->d_pitems, FSHT_FTAB_COMMON, h->d_TAe2[0],
                       h->d_TSe2[0], h->d_firstA2[0], h->d_firstS2[0], h->lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ftab_exp<1>, pgrid, dim3(256), 0, h->stream, h->d_pitems, FSHT_FTAB_COMMON, h->d_TAe2[1],
                       h->d_TSe2[1], h->d_firstA2[1], h->d_firstS2[1], h->lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  std::vector<int> firstA[2], firstS[2];
  for (int f = 0; f < 2 && e == hipSuccess; ++f) {
    firstA[f].resize(fa);
    firstS[f].resize(fs);
    e = hipMemcpy(firstA[f].data(), h->d_firstA2[f], (size_t)fa * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(firstS[f].data(), h->d_firstS2[f], (size_t)fs * 4, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return e;
  const long long cap = plan_t2_ranges(h->mix_off, h->mix_ta_total, lmax, t2_cap_limit(h), h->mix2_chunk);
  const int nchunk = (int)h->mix2_chunk.size() - 1;
  h->t2_cap = cap;
  al((void**)&h->d_T2[0], cap);
  al((void**)&h->d_T2[1], cap);
  if (e != hipSuccess) return e;
  const int nct2 = h->nsimp / 8;  // 32-column tiles of the 4*nsimp spin-2 columns
  h->d_anaq2.assign((size_t)2 * nchunk * 3, nullptr);
  h->d_synq2.assign((size_t)2 * nchunk * 3, nullptr);
  h->n_anaq2.assign((size_t)2 * nchunk * 3, 0);
  h->n_synq2.assign((size_t)2 * nchunk * 3, 0);
  for (int f = 0; f < 2 && e == hipSuccess; ++f)
    for (int c = 0; c < nchunk && e == hipSuccess; ++c) {
      const size_t q = ((size_t)f * nchunk + c) * 3;
      e = build_q_lists(h, h->mix2_chunk[c], h->mix2_chunk[c + 1], nct2, firstA[f], firstS[f], &h->d_anaq2[q],
                        &h->n_anaq2[q], &h->d_synq2[q], &h->n_synq2[q]);
    }
  if (e != hipSuccess) return e;
  h->mix_bytes += 2 * cap + 8 * (tae + tse + fa + fs);
  h->t2_lay = -1;
  h->mix2_ready = 1;
  return hipSuccess;
}

void mixed_planned_bytes(const Handle* h, long long out[3]) {
  MixPlan P;
  plan_offsets(h, P);
  const int ncw = h->ncomp_ws;
  const long long nexp = 2LL * (h->lmax + 1) * 2 * ncw * h->nsimp;
  const long long fq_bytes = (long long)ncw * P.fq;
  out[0] = (fq_fits_spec(h, fq_bytes) ? 0 : fq_bytes) + (long long)ncw * P.aq + 8 * nexp;
  out[1] = 2 * P.ta + 4 * (P.tae + P.tse + P.fa + P.fs);
  std::vector<int> bounds;
  out[2] = 2 * plan_t2_ranges(P.off, P.ta, h->lmax, t2_cap_limit(h), bounds) + 8 * (P.tae + P.tse + P.fa + P.fs);
}

hipError_t mixed_setup(Handle* h) {
  hipError_t e = mixed_setup_common(h);
  if (e == hipSuccess && !h->pol) e = mixed_setup_spin0(h);
  if (e != hipSuccess) mixed_free(h);
  return e;
}

hipError_t mixed_ensure(Handle* h, int spin) {
  hipError_t e = mixed_setup_common(h);
  if (e == hipSuccess) e = spin ? mixed_setup_spin2(h) : mixed_setup_spin0(h);
  if (e != hipSuccess) mixed_free(h);
  return e;
}

static int cur_ncomp(const Handle* h) { return h->spin ? 2 : 1; }

hipError_t launch_fexp_reset(Handle* h) {
  return fill_int(h, h->d_FQe, 2LL * (h->lmax + 1) * 2 * cur_ncomp(h) * h->nsimp, kMixZeroE);
}

// Grid of the operand slicers: (orders, y, column tiles), or with the column
// tiles first (FSHT_QUANT_CT_FAST) so that the tiles of one row run together.
static dim3 quant_grid(int orders, int y, int tiles) {
  return FSHT_QUANT_CT_FAST ? dim3(tiles, y, orders) : dim3(orders, y, tiles);
}

hipError_t launch_quant_F(Handle* h, const double* F, int K) {
  const int ncomp = cur_ncomp(h);
  dim3 grid = quant_grid(h->lmax + 1, 2 * h->mix_nrb_max, ncomp * h->nsimp / 16);
#define FSHT_QF(KV)                                                                                              \
  hipLaunchKernelGGL(k_quant_F<KV>, grid, dim3(256), 0, h->stream, h->d_mix_off, h->d_i0, h->d_goff, F, h->d_FQe, \
                     h->d_FQ, h->nhr, ncomp, h->nsim, h->nsimp)
  switch (K) {
    case 4: FSHT_QF(4); break;
    case 5: FSHT_QF(5); break;
    case 6: FSHT_QF(6); break;
    default: return hipErrorInvalidValue;
  }
#undef FSHT_QF
  return hipGetLastError();
}

hipError_t launch_quant_A(Handle* h, const double* alm0, const double* alm1, int K, int have_exp) {
  const int ncomp = cur_ncomp(h);
  dim3 grid = quant_grid(h->lmax + 1, 2, ncomp * h->nsimp / 16);
#define FSHT_QA(KV, E)                                                                                               \
  hipLaunchKernelGGL((k_quant_A<KV, E>), grid, dim3(256), 0, h->stream, h->d_mix_off, alm0, alm1, h->d_AQ, h->d_AQe, \
                     ncomp, h->nsim, h->nsimp, h->lmax)
  switch (K * 2 + (have_exp ? 1 : 0)) {
    case 8: FSHT_QA(4, 0); break;
    case 9: FSHT_QA(4, 1); break;
    case 10: FSHT_QA(5, 0); break;
    case 11: FSHT_QA(5, 1); break;
    case 12: FSHT_QA(6, 0); break;
    case 13: FSHT_QA(6, 1); break;
    default: return hipErrorInvalidValue;
  }
#undef FSHT_QA
  return hipGetLastError();
}

// One slice-GEMM pass over a set of work lists (full 4-tile column blocks,
// then the remaining column tiles: one by one, or for six slices one padded
// block).
template <int SYN>
static hipError_t launch_gemm(Handle* h, SliceArgs A, int K, LegItem* const* items, const int* nitems) {
#if FSHT_SLICE_MIX && FSHT_SLICE_PP
  if (K <= 5 && nitems[1] > 0) {  // list 1: the 4-tile blocks and the single tiles
    A.items = items[1];
    if (K == 4)
      hipLaunchKernelGGL((k_slice_gemm_pp_mix<4, SYN>), dim3(nitems[1]), dim3(512), 0, h->stream, A);
    else
      hipLaunchKernelGGL((k_slice_gemm_pp_mix<5, SYN>), dim3(nitems[1]), dim3(512), 0, h->stream, A);
    return hipGetLastError();
  }
  if (K == 6 && nitems[2] > 0) {  // list 2: the 4-tile blocks and the padded last block
    A.items = items[2];
    hipLaunchKernelGGL((k_slice_gemm<6, SYN, 4>), dim3(nitems[2]), dim3(512), 0, h->stream, A);
    return hipGetLastError();
  }
#endif
  const int lists[2] = {0, K <= 5 ? 1 : 2};
  for (int q = 0; q < 2; ++q) {
    const int k = lists[q];
    const int n = nitems[k];
    if (n == 0) continue;
    A.items = items[k];
#define FSHT_GQ_K(KERN, KV, NBV) hipLaunchKernelGGL((KERN<KV, SYN, NBV>), dim3(n), dim3(512), 0, h->stream, A)
#if FSHT_SLICE_PP
// four and five slices: the two-group kernel; six slices (whose operand
// registers would not fit beside the accumulators twice over): one group
#define FSHT_GQ(KV, NBV)                       \
  if constexpr (KV <= 5) {                     \
    FSHT_GQ_K(k_slice_gemm_pp, KV, NBV);       \
  } else {                                     \
    FSHT_GQ_K(k_slice_gemm, KV, NBV);          \
  }
#else
#define FSHT_GQ(KV, NBV) FSHT_GQ_K(k_slice_gemm, KV, NBV)
#endif
    if (k == 1) {
      switch (K) {
        case 4: FSHT_GQ(4, 1); break;
        case 5: FSHT_GQ(5, 1); break;
        default: return hipErrorInvalidValue;
      }
    } else {
      switch (K) {
        case 4: FSHT_GQ(4, 4); break;
        case 5: FSHT_GQ(5, 4); break;
        case 6: FSHT_GQ(6, 4); break;
        default: return hipErrorInvalidValue;
      }
    }
#undef FSHT_GQ
#undef FSHT_GQ_K
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static SliceArgs base_args(Handle* h, int SYN, int ncomp) {
  SliceArgs A{};
  A.mo = h->d_mix_off;
  A.OQ = SYN ? h->d_AQ : h->d_FQ;
  A.OQe = SYN ? h->d_AQe : h->d_FQe;
  A.goff = h->d_goff;
  A.i0tab = h->d_i0;
  A.nsim = h->nsim;
  A.nsimp = h->nsimp;
  A.ncomp = ncomp;
  A.lmax = h->lmax;
  A.nhr = h->nhr;
  A.nct = ncomp * h->nsimp / 16;
  A.pa = h->pa;
  return A;
}

int mixed_nrange(const Handle* h) { return h->spin ? (int)h->mix2_chunk.size() - 1 : 1; }

// Spin 2: F1 and F2 slices of the orders of range c, in one operand order
// (synthesis = 0: TA, 1: TS), into the chunk buffers (kept when one range
// covers every order).
static long long range_base(const Handle* h, int c, int K) { return h->mix_off[2 * h->mix2_chunk[c]].ta / kMixK * K; }

hipError_t launch_ftab(Handle* h, int c, int lay, int K) {
  const int nchunk = (int)h->mix2_chunk.size() - 1;
  // a captured stream always generates (a graph replays what was recorded,
  // whatever the buffers hold then) and leaves the buffers' content unknown
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hipError_t ce = hipStreamIsCapturing(h->stream, &cap);
  if (ce != hipSuccess) return ce;
  const bool capturing = cap != hipStreamCaptureStatusNone;
  if (!capturing && nchunk == 1 && h->t2_lay == lay && h->t2_ks == K) return hipSuccess;
  const int m0 = h->mix2_chunk[c];
  const long long tbase = range_base(h, c, K);
  const int i0 = h->pitem_off[m0], i1 = h->pitem_off[h->mix2_chunk[c + 1]];
  if (i1 <= i0) return hipSuccess;
  const int2* items = h->d_pitems + i0;
  const dim3 grid((unsigned)(i1 - i0));
#define FSHT_FT(FN, LAY, KS) \
  hipLaunchKernelGGL((k_ftab_dig<FN, LAY, KS>), grid, dim3(256), 0, h->stream, items, FSHT_FTAB_COMMON, h->d_T2[FN], \
                     tbase, h->d_TAe2[FN], h->d_TSe2[FN], h->lmax, h->nhr)
#define FSHT_FT2(LAY, KS) \
  FSHT_FT(0, LAY, KS);    \
  FSHT_FT(1, LAY, KS)
  switch (lay * 8 + K) {
    case 4: FSHT_FT2(0, 4); break;
    case 5: FSHT_FT2(0, 5); break;
    case 6: FSHT_FT2(0, 6); break;
    case 12: FSHT_FT2(1, 4); break;
    case 13: FSHT_FT2(1, 5); break;
    case 14: FSHT_FT2(1, 6); break;
    default: return hipErrorInvalidValue;
  }
#undef FSHT_FT2
#undef FSHT_FT
  h->t2_lay = capturing ? -1 : lay;
  h->t2_ks = K;
  return hipGetLastError();
}

hipError_t launch_aqe_reset(Handle* h) {
  return fill_int(h, h->d_AQe, 2LL * (h->lmax + 1) * 2 * cur_ncomp(h) * h->nsimp, kMixZeroE);
}

hipError_t launch_analysis_q(Handle* h, int c, double* alm0, double* alm1, double beta, int K, int exps) {
  if (h->spin == 0) {
    SliceArgs A = base_args(h, 0, 1);
    A.TP = h->d_TA;
    A.TPe = h->d_TAe;
    A.ast = kMixK;
    A.out0 = alm0;
    A.beta = beta;
    A.aqe = exps ? h->d_AQe : nullptr;
    return launch_gemm<0>(h, A, K, h->d_anaq_items, h->n_anaq_items);
  }
  // spin 2: F1, then F2 (which adds to F1's result and records the exponents)
  const int nchunk = (int)h->mix2_chunk.size() - 1;
  hipError_t e = hipSuccess;
  for (int f = 0; f < 2 && e == hipSuccess; ++f) {
    SliceArgs A = base_args(h, 0, 2);
    A.TP = h->d_T2[f];
    A.TPe = h->d_TAe2[f];
    A.ast = K;
    A.abase = range_base(h, c, K);
    A.out0 = alm0;
    A.out1 = alm1;
    A.fn = f;
    A.beta = f ? 1.0 : beta;
    A.aqe = (f && exps) ? h->d_AQe : nullptr;
    const size_t q = ((size_t)f * nchunk + c) * 3;
    e = launch_gemm<0>(h, A, K, &h->d_anaq2[q], &h->n_anaq2[q]);
  }
  return e;
}

hipError_t launch_synthesis_q(Handle* h, int c, double* G, int K) {
  if (h->spin == 0) {
    SliceArgs A = base_args(h, 1, 1);
    A.TP = h->d_TS;
    A.TPe = h->d_TSe;
    A.ast = kMixK;
    A.out0 = G;
    return launch_gemm<1>(h, A, K, h->d_synq_items, h->n_synq_items);
  }
  const int nchunk = (int)h->mix2_chunk.size() - 1;
  hipError_t e = hipSuccess;
  for (int f = 0; f < 2 && e == hipSuccess; ++f) {
    SliceArgs A = base_args(h, 1, 2);
    A.TP = h->d_T2[f];
    A.TPe = h->d_TSe2[f];
    A.ast = K;
    A.abase = range_base(h, c, K);
    A.out0 = G;
    A.fn = f;
    A.beta = f ? 1.0 : 0.0;
    const size_t q = ((size_t)f * nchunk + c) * 3;
    e = launch_gemm<1>(h, A, K, &h->d_synq2[q], &h->n_synq2[q]);
  }
  return e;
}

}  // namespace fsht
