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
// Analysis operand: the m-major folded spectrum F (FP64, layout of k_gather)
// -> FQ slices with the per-column exponents the gather left in FQe
// (index (2 m + class) * ncols + col).  Workgroup = (m, class, 8 ring steps,
// 32-column tile); thread = (column, ring step): 32 rows of one column.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_quant_F(const MixOff* __restrict__ mo, const int* __restrict__ i0tab,
                                                 const long long* __restrict__ goff, const double* __restrict__ F,
                                                 const int* __restrict__ FQe, signed char* __restrict__ FQ, int nhr,
                                                 int ncols, int nsim) {
  const int m = blockIdx.x, cls = blockIdx.y & 1, kb = blockIdx.y >> 1, ct = blockIdx.z;
  if (!cols_active(ct * 32, ct * 32 + 32, nsim, ncols / 2)) return;  // no map in the tile
  const MixOff o = mo[2 * m + cls];
  const int cl = threadIdx.x & 31, q = threadIdx.x >> 5;
  const int rs = kb * 8 + q;
  if (rs >= o.nrs) return;
  const int R = nhr + 1 - i0tab[m];
  const int col = ct * 32 + cl;
  const int e = FQe[(long long)(2 * m + cls) * ncols + col];
  const double* src = F + goff[m] / 2 + (long long)cls * R * ncols + col;
  double v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int r = rs * 32 + j;
    v[j] = r < R ? ldexp(src[(long long)r * ncols], -e) : 0.0;
  }
  signed char* dst = FQ + o.fq + ((long long)ct * o.nrs + rs) * K * 1024 + cl * 16;
  put16<K>(v, dst);
  put16<K>(v + 16, dst + 32 * 16);
}

// ---------------------------------------------------------------------------
// Synthesis operand: packed alm -> AQ slices with one exponent per column
// (AQe, index (2 m + class) * ncols + col).  Columns as in the m-major
// buffer: col < nsimp is Re a_lm of map col, nsimp <= col < 2 nsimp is Im.
// Workgroup = (m, class, 32-column tile).  HAVE_EXP = 0: the column maxima
// first (setting AQe), then the digits (second read of the same alm, mostly
// from cache); HAVE_EXP = 1: AQe was set by the analysis that wrote alm.
// ---------------------------------------------------------------------------
template <int K, int HAVE_EXP>
__global__ __launch_bounds__(256) void k_quant_A(const MixOff* __restrict__ mo, const double* __restrict__ alm,
                                                 signed char* __restrict__ AQ, int* __restrict__ AQe, int nsim,
                                                 int nsimp, int lmax) {
  __shared__ double red[8][32];
  __shared__ int ex[32];
  const int m = blockIdx.x, cls = blockIdx.y, ct = blockIdx.z;
  if (!cols_active(ct * 32, ct * 32 + 32, nsim, nsimp)) return;  // no map in the tile
  const MixOff o = mo[2 * m + cls];
  const int cl = threadIdx.x & 31, q = threadIdx.x >> 5;
  const int col = ct * 32 + cl;
  const int ncols = 2 * nsimp;
  const int reim = col >= nsimp ? 1 : 0;
  const int sim = col - reim * nsimp;
  const bool ok = sim < nsim && !(m == 0 && reim);
  const long long L1 = lmax + 1;
  const double* src = alm + (ok ? sim : 0) + (long long)nsim * (reim ? (L1 * (L1 - m) - m) : (L1 * m));
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
    signed char* dst = AQ + o.aq + ((long long)ct * o.nlt + lst) * K * 1024 + cl * 16;
    put16<K>(v, dst);
    put16<K>(v + 16, dst + 32 * 16);
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
// ---------------------------------------------------------------------------
constexpr int kLdsBudget = 147456;  // bytes of LDS the slice GEMM may use

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

template <int K, int SYN, int NB>
__global__ __launch_bounds__(512, 1) void k_slice_gemm(const LegItem* __restrict__ items,
                                                       const MixOff* __restrict__ mo,
                                                       const signed char* __restrict__ TP,
                                                       const int* __restrict__ TPe,
                                                       const signed char* __restrict__ OQ,
                                                       const int* __restrict__ OQe, double* __restrict__ out,
                                                       const long long* __restrict__ goff,
                                                       const int* __restrict__ i0tab, int nsim, int nsimp, int lmax,
                                                       int nhr, int nct, double pa, double beta,
                                                       int* __restrict__ aqe) {
  static_assert(NB == 4 || NB == 1, "column tiles per workgroup");
  constexpr int NW = 8;                                 // waves
  constexpr int NA = NB == 4 ? 4 : 8;                   // row tiles
  constexpr int TPW = NB == 4 ? 2 : 1;                  // column tiles per wave
  constexpr int NBLK = (NA + NB) * K;                   // 1 KB blocks per step
  constexpr int PER = (NBLK + NW - 1) / NW;             // loads per wave per step (the last may repeat)
  constexpr int NS = kLdsBudget / (NBLK * 1024) < 4 ? kLdsBudget / (NBLK * 1024) : 4;  // stages
  static_assert(NS >= 3, "pipeline needs three stages");
  __shared__ __align__(16) signed char st[NS][NBLK * 1024];
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;  // whole workgroup
  if (!cols_active(it.colblk * NB * 32, (it.colblk + 1) * NB * 32, nsim, nsimp)) return;  // no map in the block
  const int m = it.m, cls = it.par;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int L = threadIdx.x & 63;
  const MixOff o = mo[2 * m + cls];
  const int nsteps = SYN ? o.nlt : o.nrs;
  const int ntA = SYN ? o.nrs : o.nlt;
  const int ncols = 2 * nsimp;
  const signed char* baseA = TP + (SYN ? o.ts : o.ta);
  const signed char* baseB = OQ + (SYN ? o.aq : o.fq);

  const int s0 = it.r0;
  // this wave's PER source streams: wave-uniform pointers at step s0 that
  // advance by one step's stride per issue (A tiles keep kMixK slices per
  // step, B tiles K), and their blocks within a stage
  const signed char* gsrc[PER];
  int gstep[PER], lofs[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    int b = w + NW * k;
    if (b >= NBLK) b -= NW;  // this wave's previous block again (same data, same place)
    lofs[k] = b * 1024;
    if (b < NA * K) {
      const int tile = min(it.blk * NA + b / K, ntA - 1);
      gsrc[k] = baseA + (((long long)tile * nsteps + s0) * kMixK + b % K) * 1024;
      gstep[k] = kMixK * 1024;
    } else {
      const int bb = b - NA * K;
      const int tile = min(it.colblk * NB + bb / K, nct - 1);
      gsrc[k] = baseB + (((long long)tile * nsteps + s0) * K + bb % K) * 1024;
      gstep[k] = K * 1024;
    }
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
  const bool act = ta < ntA && tb0 < nct;
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

  // epilogue: out = 2^(eA(row) + eB(col)) * sum_d 2^-7(d+2) D_d
  const int h = L >> 5;
  const int* pe = TPe + (SYN ? o.tse : o.tae) + (long long)ta * 32 + 16 * h;  // accumulator order
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tb = tb0 + t;
    if (tb >= nct) break;
    const int col = tb * 32 + (L & 31);
    const int eB = OQe[(long long)(2 * m + cls) * ncols + col];
    if (!SYN) {
      const int reim = col >= nsimp ? 1 : 0;
      const int sim = col - reim * nsimp;
      if (sim >= nsim || (m == 0 && reim)) continue;
      const long long L1 = lmax + 1;
      // this column's entries: sim + nsim (l + L1 m) for Re, sim + nsim (l - m + L1 (L1 - m)) for Im
      double* oc = out + sim + (long long)nsim * (reim ? L1 * (L1 - m) - m : L1 * m);
      // the previous values (beta != 0) are all loaded before any store, so
      // the read-modify-write costs one memory latency, not sixteen
      double prev[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int l = m + cls + 2 * (ta * 32 + drow(r, h));
        prev[r] = (beta != 0.0 && l <= lmax) ? oc[(long long)nsim * l] : 0.0;
      }
      double mx = 0.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int l = m + cls + 2 * (ta * 32 + drow(r, h));
        if (l > lmax) continue;
        const double v = pa * __builtin_ldexp(combine<K>(t ? D1 : D0, r), pe[r] + eB) + beta * prev[r];
        oc[(long long)nsim * l] = v;
        mx = fmax(mx, fabs(v));
      }
      // the synthesis operand's exponent of this column: max over every
      // workgroup writing it (k_quant_A<K, 1> reads it)
      if (aqe) atomicMax(aqe + (long long)(2 * m + cls) * ncols + col, mix_exp(mx));
    } else {
      const int R = nhr + 1 - i0tab[m];
      const long long gbase = goff[m] / 2 + (long long)cls * R * ncols;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = ta * 32 + drow(r, h);
        if (row < R) out[gbase + (long long)row * ncols + col] = __builtin_ldexp(combine<K>(t ? D1 : D0, r), pe[r] + eB);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

void mixed_free(Handle* h) {
  void* ptrs[] = {h->d_mix_off, h->d_TA, h->d_TS, h->d_FQ, h->d_AQ, h->d_TAe, h->d_TSe, h->d_FQe, h->d_AQe,
                  h->d_firstA,  h->d_firstS};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 3; ++k) {
    if (h->d_anaq_items[k]) (void)hipFree(h->d_anaq_items[k]);
    if (h->d_synq_items[k]) (void)hipFree(h->d_synq_items[k]);
    h->d_anaq_items[k] = h->d_synq_items[k] = nullptr;
    h->n_anaq_items[k] = h->n_synq_items[k] = 0;
  }
  h->d_mix_off = nullptr;
  h->d_TA = h->d_TS = h->d_FQ = h->d_AQ = nullptr;
  h->d_TAe = h->d_TSe = h->d_FQe = h->d_AQe = nullptr;
  h->d_firstA = h->d_firstS = nullptr;
  h->mix_ready = 0;
  h->mix_bytes = 0;
}

static hipError_t fill_int(Handle* h, int* p, long long n, int v) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks), dim3(256), 0, h->stream, p, n, v);
  return hipGetLastError();
}

hipError_t mixed_setup(Handle* h) {
  if (h->mix_ready) return hipSuccess;
  const int lmax = h->lmax, nm = lmax + 1;
  const int nct = h->nsimp / 16;  // 32-column tiles of the 2*nsimp spin-0 columns
  const int ncols = 2 * h->nsimp;
  h->mix_off.assign(2 * (size_t)nm, MixOff{});
  long long ta = 0, fq = 0, aq = 0, tae = 0, tse = 0;
  int fa = 0, fs = 0, nrb_max = 0;
  std::vector<int2> pitems;
  for (int m = 0; m <= lmax; ++m) {
    const int R = h->nhr + 1 - h->i0[m];
    const int nrs = cdiv(R, 32), nrb = cdiv(nrs, 8);
    nrb_max = std::max(nrb_max, nrb);
    for (int rb = 0; rb < nrb; ++rb) pitems.push_back(make_int2(m, rb));
    for (int c = 0; c < 2; ++c) {
      const int nval = (lmax >= m + c) ? (lmax - m - c) / 2 + 1 : 0;
      const int nlt = cdiv(nval, 32);
      MixOff& o = h->mix_off[2 * m + c];
      o.nrs = nrs;
      o.nlt = nlt;
      o.ta = o.ts = ta;
      ta += (long long)nlt * nrs * kMixK * 1024;
      o.fq = fq;
      fq += (long long)nct * nrs * kMixK * 1024;
      o.aq = aq;
      aq += (long long)nct * nlt * kMixK * 1024;
      o.tae = tae;
      tae += (long long)nlt * 32;
      o.tse = tse;
      tse += (long long)nrs * 32;
      o.fqe = o.aqe = (long long)(2 * m + c) * ncols;
      o.fa = fa;
      fa += nlt;
      o.fs = fs;
      fs += nrs;
    }
  }
  h->mix_nrb_max = nrb_max;
  const long long nexp = 2LL * nm * ncols;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, long long bytes) {
    if (e == hipSuccess) e = hipMalloc(p, (size_t)std::max(bytes, 256LL));
  };
  al((void**)&h->d_TA, ta);
  al((void**)&h->d_TS, ta);
  al((void**)&h->d_FQ, fq);
  al((void**)&h->d_AQ, aq);
  al((void**)&h->d_TAe, tae * 4);
  al((void**)&h->d_TSe, tse * 4);
  al((void**)&h->d_FQe, nexp * 4);
  al((void**)&h->d_AQe, nexp * 4);
  al((void**)&h->d_firstA, (long long)fa * 4);
  al((void**)&h->d_firstS, (long long)fs * 4);
  al((void**)&h->d_mix_off, (long long)h->mix_off.size() * (long long)sizeof(MixOff));
  int2* d_pitems = nullptr;
  al((void**)&d_pitems, (long long)pitems.size() * (long long)sizeof(int2));
  if (e != hipSuccess) {
    (void)hipFree(d_pitems);
    mixed_free(h);
    return e;
  }
  h->mix_bytes = 2 * ta + fq + aq + 4 * (tae + tse + 2 * nexp + fa + fs);
  e = hipMemcpy(h->d_mix_off, h->mix_off.data(), h->mix_off.size() * sizeof(MixOff), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_pitems, pitems.data(), pitems.size() * sizeof(int2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = fill_int(h, h->d_firstA, fa, kNeverStep);
  if (e == hipSuccess) e = fill_int(h, h->d_firstS, fs, kNeverStep);
  if (e == hipSuccess) e = fill_int(h, h->d_TAe, tae, kMixZeroE);
  const dim3 pgrid((unsigned)pitems.size());
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ptab<0>, pgrid, dim3(256), 0, h->stream, d_pitems, h->d_geo, h->d_i0, h->d_ab, h->d_aboff,
                       h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_mix_off, h->d_TA, h->d_TAe, h->d_TS, h->d_TSe,
                       h->d_firstA, h->d_firstS, lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ptab<1>, pgrid, dim3(256), 0, h->stream, d_pitems, h->d_geo, h->d_i0, h->d_ab, h->d_aboff,
                       h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_mix_off, h->d_TA, h->d_TAe, h->d_TS, h->d_TSe,
                       h->d_firstA, h->d_firstS, lmax, h->nhr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d_pitems);
  std::vector<int> firstA(fa), firstS(fs);
  if (e == hipSuccess) e = hipMemcpy(firstA.data(), h->d_firstA, (size_t)fa * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(firstS.data(), h->d_firstS, (size_t)fs * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    mixed_free(h);
    return e;
  }
  // work lists: every (m, class, row tiles, column tiles) writes its
  // outputs, starting at the first step with a nonzero P digit of its row tiles
  const int nfull = nct / 4, nrem = nct % 4;
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
    for (int m = 0; m <= lmax; ++m)
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
    e = upload_items(ana, ana_cost, &h->d_anaq_items[k], &h->n_anaq_items[k]);
    if (e == hipSuccess) e = upload_items(syn, syn_cost, &h->d_synq_items[k], &h->n_synq_items[k]);
    if (e != hipSuccess) {
      mixed_free(h);
      return e;
    }
  }
  h->mix_ready = 1;
  return hipSuccess;
}

hipError_t launch_fexp_reset(Handle* h) { return fill_int(h, h->d_FQe, 2LL * (h->lmax + 1) * 2 * h->nsimp, kMixZeroE); }

hipError_t launch_quant_F(Handle* h, const double* F, int K) {
  const int nct = h->nsimp / 16;
  dim3 grid(h->lmax + 1, 2 * h->mix_nrb_max, nct);
  const int ncols = 2 * h->nsimp;
#define FSHT_QF(KV)                                                                                              \
  hipLaunchKernelGGL(k_quant_F<KV>, grid, dim3(256), 0, h->stream, h->d_mix_off, h->d_i0, h->d_goff, F, h->d_FQe, \
                     h->d_FQ, h->nhr, ncols, h->nsim)
  switch (K) {
    case 4: FSHT_QF(4); break;
    case 5: FSHT_QF(5); break;
    case 6: FSHT_QF(6); break;
    default: return hipErrorInvalidValue;
  }
#undef FSHT_QF
  return hipGetLastError();
}

hipError_t launch_quant_A(Handle* h, const double* alm, int K, int have_exp) {
  const int nct = h->nsimp / 16;
  dim3 grid(h->lmax + 1, 2, nct);
#define FSHT_QA(KV, E)                                                                                              \
  hipLaunchKernelGGL((k_quant_A<KV, E>), grid, dim3(256), 0, h->stream, h->d_mix_off, alm, h->d_AQ, h->d_AQe, h->nsim, \
                     h->nsimp, h->lmax)
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

template <int SYN>
static hipError_t launch_gemm(Handle* h, double* out, double beta, int K, int* aqe) {
  const int nct = h->nsimp / 16;
  const signed char* TP = SYN ? h->d_TS : h->d_TA;
  const int* TPe = SYN ? h->d_TSe : h->d_TAe;
  const signed char* OQ = SYN ? h->d_AQ : h->d_FQ;
  const int* OQe = SYN ? h->d_AQe : h->d_FQe;
  // full column blocks, then the remaining tiles (one by one; six slices: one padded block)
  const int lists[2] = {0, K <= 5 ? 1 : 2};
  for (int q = 0; q < 2; ++q) {
    const int k = lists[q];
    const int n = SYN ? h->n_synq_items[k] : h->n_anaq_items[k];
    if (n == 0) continue;
    const LegItem* items = SYN ? h->d_synq_items[k] : h->d_anaq_items[k];
#define FSHT_GQ(KV, NBV)                                                                                          \
  hipLaunchKernelGGL((k_slice_gemm<KV, SYN, NBV>), dim3(n), dim3(512), 0, h->stream, items, h->d_mix_off, TP, TPe, \
                     OQ, OQe, out, h->d_goff, h->d_i0, h->nsim, h->nsimp, h->lmax, h->nhr, nct, h->pa, beta, aqe)
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
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_analysis_q(Handle* h, double* alm, double beta, int K, int exps) {
  if (!exps) return launch_gemm<0>(h, alm, beta, K, nullptr);
  // alm column exponents for the next synthesis operand
  hipError_t e = fill_int(h, h->d_AQe, 2LL * (h->lmax + 1) * 2 * h->nsimp, kMixZeroE);
  return e != hipSuccess ? e : launch_gemm<0>(h, alm, beta, K, h->d_AQe);
}
hipError_t launch_synthesis_q(Handle* h, double* G, int K) { return launch_gemm<1>(h, G, 0.0, K, nullptr); }

}  // namespace fsht
