// Per-m Legendre contractions on FP64 MFMA with on-the-fly P_lm / F1 / F2.
//
// Replaces the reference's per-m loop of get_plm_recursive + fft2mat_fast +
// mat2alm (cuBLAS/MKL dgemm) for the forward transform, and of
// get_plm_recursive + alm2mat + mat2fft for the backward one
// (src/sht_main.f90:65-79, 117-134, 409-430, 1102-1124;
//  src/sht_core.f90:30-168; src/get_plm_recursive.f90:17-136;
//  src/fft_mapping.f90:19-68).
//
// All orders m run in one launch.  North/south ring symmetry is folded:
// P_lm(-z) = (-1)^(l+m) P_lm(z) (and F2 carries the opposite sign), so the
// contraction over 2*nside rings becomes one over the north half for each
// parity class, halving the arithmetic.
//
// v_mfma_f64_16x16x4_f64 operand layout (verified on gfx950):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
//   D[i][j] in lane j + 16*(i%4), accumulator register i/4.
//
// FP64 MFMA and FP64 VALU share one datapath on gfx950, so the recurrence is
// kept lean: coefficients are staged in LDS, every global gather is
// branch-free (clamped addresses, selected-away values) so loads are issued in
// batches, and a ring enters the recurrence at a chunk boundary from a stored
// state (two selects per chunk, none per step).
#include <type_traits>

#include "fsht_internal.h"

namespace fsht {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Source selection.  Output column types: spin 0 (Re, Im); spin 2
// analysis (E_re, E_im, B_re, B_im), synthesis (Q_re, Q_im, U_re, U_im).
// For each output type and Legendre function f (0: P or F1, 1: F2):
// source component, re/im, fold flip (0: same parity class, 1: opposite) and sign.
//   analysis:  E = -pa sum(Q F1 + i U F2),  B = pa sum(i Q F2 - U F1)
//   synthesis: Q = -E F1 - i B F2,          U = i E F2 - B F1
// (signs from the alpha/swap table, src/sht_data_init.f90:606-632, and the
//  type sequence of src/sht_main.f90:409-430, 1102-1124).
// ---------------------------------------------------------------------------
struct Src {
  int comp, reim, flip;
  double sign;
};
__device__ inline Src ana_src(int spin, int otype, int f) {
  if (spin == 0) return Src{0, otype, 0, 1.0};
  switch (otype * 2 + f) {
    case 0: return Src{0, 0, 0, -1.0};  // E_re <- F1 * -Qr
    case 1: return Src{1, 1, 1, 1.0};   // E_re <- F2 * +Ui
    case 2: return Src{0, 1, 0, -1.0};  // E_im <- F1 * -Qi
    case 3: return Src{1, 0, 1, -1.0};  // E_im <- F2 * -Ur
    case 4: return Src{1, 0, 0, -1.0};  // B_re <- F1 * -Ur
    case 5: return Src{0, 1, 1, -1.0};  // B_re <- F2 * -Qi
    case 6: return Src{1, 1, 0, -1.0};  // B_im <- F1 * -Ui
    default: return Src{0, 0, 1, 1.0};  // B_im <- F2 * +Qr
  }
}
__device__ inline Src syn_src(int spin, int otype, int f) {
  if (spin == 0) return Src{0, otype, 0, 1.0};
  switch (otype * 2 + f) {
    case 0: return Src{0, 0, 0, -1.0};  // Q_re <- F1 * -Er
    case 1: return Src{1, 1, 0, 1.0};   // Q_re <- F2 * +Bi
    case 2: return Src{0, 1, 0, -1.0};  // Q_im <- F1 * -Ei
    case 3: return Src{1, 0, 0, -1.0};  // Q_im <- F2 * -Br
    case 4: return Src{1, 0, 0, -1.0};  // U_re <- F1 * -Br
    case 5: return Src{0, 1, 0, -1.0};  // U_re <- F2 * -Ei
    case 6: return Src{1, 1, 0, -1.0};  // U_im <- F1 * -Bi
    default: return Src{0, 0, 0, 1.0};  // U_im <- F2 * +Er
  }
}

// Per-thread recurrence state for one ring.  The recurrence is 0 until the
// chunk that starts at l = ls, where (p1, p2) are loaded from sv.
struct Rec {
  double p1, p2;  // P_{l-1}, P_{l-2}
  double x, bw, t2, cw;
  int ls;         // window start at which the ring enters (0x3fffffff: never)
  double2 sv;     // (P_{ls-1}, P_{ls-2})
};

__device__ inline void enter_chunk(Rec& r, int lc) {
  if (__any(r.ls == lc)) {
    const bool h = r.ls == lc;
    r.p1 = h ? r.sv.x : r.p1;
    r.p2 = h ? r.sv.y : r.p2;
  }
}

// Run 32 recurrence steps l = lc .. lc+31 with coefficients cf[0..31]
// (zero beyond lmax) and hand every value (P, or F_f for spin 2) to emit(s, v).
template <int SPIN, class Emit>
__device__ inline void gen32(Rec& r, const double2* cf, const double2* scf, int lc, int m, int f, Emit&& emit) {
#pragma unroll
  for (int s = 0; s < 32; ++s) {
    const double2 c = cf[s];
    const double p = fma(c.x * r.x, r.p1, -(c.y * r.p2));
    double v = p;
    if (SPIN) v = spin_f(f, p, r.p1, scf[s], lc + s, m, r.x, r.bw, r.t2, r.cw);
    emit(s, v);
    r.p2 = r.p1;
    r.p1 = p;
  }
}

// ---------------------------------------------------------------------------
// Analysis: alm[l in class par, cols] = pa * sum_{north rings} F(l, ring) * X(ring, cols)
// X is the m-major folded spectrum written by k_gather.  Workgroup = (m,
// parity class, l-block of kAnaLB, 64 columns); 4 waves, each owning 16
// columns and all 8 l-tiles of the block.  Rings are consumed in sub-blocks:
// each wave loads its X columns for the sub-block into registers, then the
// workgroup sweeps the l-block generating P (one ring per thread, plain FP64
// recurrence entered from the resident start states) into LDS.  The 32-step
// windows of class par start at l = L0 + 32t + par, so the class's values
// are the even steps of every window.
// ---------------------------------------------------------------------------

// Workgroups per CU the register allocation must allow (spin 0 / spin 2).
#ifndef FSHT_ANA_MINBLK0
#define FSHT_ANA_MINBLK0 1
#endif
#ifndef FSHT_ANA_MINBLK2
#define FSHT_ANA_MINBLK2 2
#endif
// KS = 1: the 4 waves own 4 column tiles and every ring of a sub-block.
// KS = 2 (column blocks of 2 tiles: the last block of a spin-0 batch whose
// 16-column tile count is not a multiple of 4): waves w and w+2 share column
// tile w & 1, each over one half of the sub-block's K-steps, and the two
// partial sums are added through LDS at the end.
template <int SPIN, int KS>
__global__ __launch_bounds__(256, SPIN ? FSHT_ANA_MINBLK2 : FSHT_ANA_MINBLK0) void k_analysis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const long long* __restrict__ seedoff, const int* __restrict__ seed_st, const double2* __restrict__ seed_v,
    const double* __restrict__ seed_p, const double* __restrict__ F, const long long* __restrict__ goff,
    double* __restrict__ alm0,
    double* __restrict__ alm1, int nsim, int nsimp, int lmax, int nhr, double pa, double beta, int want0,
    int want1) {
  constexpr int RS = SPIN ? kAnaRings2 : kAnaRings0;  // rings per sub-block
  constexpr int NCOMP = SPIN ? 2 : 1;
  __shared__ double Pt[256][17];
  __shared__ double2 cf[kAnaLB + 32];  // l = L0 .. L0+287: windows of class 1 end at L0+256
  __shared__ double2 scf[SPIN ? kAnaLB + 32 : 1];
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;
  if (!cols_active(it.colblk * 64 / KS, (it.colblk + 1) * 64 / KS, nsim, nsimp)) return;  // no map in the block
  const int m = it.m, par = it.par;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: operand bases in scalar registers
  const int li = lane & 15, q = lane >> 4;
  const int ncols = 2 * NCOMP * nsimp;
  const int i0 = i0tab[m];
  const int nr = nhr + 1 - i0;
  const int L0 = m + it.blk * kAnaLB;
  const int L1 = min(L0 + kAnaLB, lmax + 1);
  const int nt = (L1 - L0 - par + 31) / 32;  // l-tiles in this block
  // this wave's output column tile and K half
  constexpr int NKB = 64 / KS;                 // K-steps (of 4 rings) per wave and sub-block
  const int kh = (KS == 1) ? 0 : wave / 2;
  const int c0 = (it.colblk * (4 / KS) + wave % (4 / KS)) * 16;  // colblk counts blocks of 4/KS tiles
  const bool colok = c0 < ncols;
  const int otype = colok ? c0 / nsimp : 0;
  const int sim0 = colok ? c0 - otype * nsimp : 0;
  const int ocomp = otype >> 1, oreim = otype & 1;
  bool active = colok && !(m == 0 && oreim == 1);
  if (SPIN) active = active && (ocomp == 0 ? want0 : want1);
  // operand of each Legendre function (spin 2: F1, F2): element (ring row r,
  // column sim0 + li) of the folded spectrum sits at base[r * ncols + li], with
  // a wave-uniform base (source column, fold class).  Held in scalars: the
  // function index of a K-step depends on the wave when KS = 2, and a runtime
  // index into a local array would live in scratch.
  const long long fb = goff[m] / (SPIN ? 1 : 2);
  const double* base0;
  const double* base1 = nullptr;
  double ssign0, ssign1 = 0.0;
  {
    const Src s = ana_src(SPIN, otype, 0);
    base0 = F + fb + (long long)((par ^ s.flip) * nr) * ncols + (s.comp * 2 + s.reim) * nsimp + sim0;
    ssign0 = active ? s.sign : 0.0;
  }
  if (SPIN) {
    const Src s = ana_src(SPIN, otype, 1);
    base1 = F + fb + (long long)((par ^ s.flip) * nr) * ncols + (s.comp * 2 + s.reim) * nsimp + sim0;
    ssign1 = active ? s.sign : 0.0;
  }

  // recurrence coefficients of the block (zero beyond lmax: P stays 0 there)
  {
    const double2* abm = ab + aboff[m];
    const double2* spm = spc + aboff[m];
    for (int k = tid; k < kAnaLB + 32; k += 256) {
      const int l = L0 + k;
      cf[k] = (l <= lmax) ? abm[l - m] : make_double2(0.0, 0.0);
      if (SPIN) scf[k] = (l <= lmax) ? spm[l - m] : make_double2(0.0, 0.0);
    }
  }
  // every wave's recurrence reads coefficients staged by other waves
  __syncthreads();

  d4 D[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) D[t] = d4{0.0, 0.0, 0.0, 0.0};

  const int myf = SPIN ? (tid / RS) : 0;
  const int myr = SPIN ? (tid % RS) : tid;

  // start data of this thread's ring in a sub-block; the next sub-block's are
  // fetched while the current one runs
  struct RecIn {
    int st;
    double2 q;   // P at ls-2, ls-1
    double q2;   // P at ls
    double z, s;
  };
  auto fetch = [&](int rb, RecIn& in) {
    const int i = i0 + rb + myr;
    const int ic = min(i, nhr);
    const long long so = seedoff[m] + (long long)it.blk * nr + (ic - i0);
    const int st = seed_st[so];
    in.st = (i <= nhr) ? st : 0x3fffffff;
    in.q = seed_v[so];
    in.q2 = seed_p[so];
    in.z = geo[ic].z;
    in.s = geo[ic].s;
  };
  RecIn nxt;
  fetch(it.r0, nxt);

  // rings before it.r0 have no kept P_lm in this l-block
  for (int rb = it.r0; rb < nr; rb += RS) {
    // rings of the sub-block that exist; wave w generates (and MFMA group w
    // consumes) rings 64*(w % (RS/64)) .. +63, skipped past the last ring
    const int kvalid = min(RS, nr - rb);
    // ---- this thread's recurrence state (its ring of the sub-block)
    Rec rec;
    {
      const int st = nxt.st;
      // class 1 enters one step later, at ls + 1, from (P_ls, P_ls-1)
      rec.ls = (st > lmax) ? 0x3fffffff : L0 + ((st - L0) & ~31) + par;
      rec.sv = par ? make_double2(nxt.q2, nxt.q.y) : make_double2(nxt.q.y, nxt.q.x);
      rec.x = nxt.z;
      const double s2 = nxt.s * nxt.s;
      rec.bw = nxt.z / s2;
      rec.t2 = 2.0 / s2;
      rec.cw = (double)m / s2;
      rec.p1 = 0.0;
      rec.p2 = 0.0;
    }
    // ---- B fragments: the wave's 16 columns of the folded, rotated ring
    // spectra for the sub-block's rings (one coalesced row read per ring).
    // Rows past the last ring are clamped and read real values: their P rows
    // are 0 (rings that do not exist never enter the recurrence).
    double bfr[NKB];
#pragma unroll
    for (int kj = 0; kj < NKB; ++kj) {
      const int ks = kh * NKB + kj;
      const int krow = 4 * ks + q;
      const int f = SPIN ? (ks >= 32 ? 1 : 0) : 0;
      const int r = SPIN ? (krow % RS) : krow;
      const unsigned row = (unsigned)min(rb + r, nr - 1);
      // 32-bit unsigned byte offset from a uniform base: scalar-base loads
      const unsigned off = 8u * (row * (unsigned)ncols + (unsigned)li);
      const double v = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(f ? base1 : base0) + off);
      bfr[kj] = SPIN ? (f ? ssign1 : ssign0) * v : v;
    }
    if (rb + RS < nr) fetch(rb + RS, nxt);

    // spin 2 keeps the chunk loop rolled (unrolled, the compiler hoists the
    // per-l spin-2 constants of all 8 chunks out of the ring loop and spills);
    // its chunk t accumulates into D[0], and the 8 tiles rotate one slot per
    // chunk, so they stay in registers (a runtime D[t] would live in scratch)
    // and are back in order after each sub-block
    constexpr int kUnrollT = SPIN ? 1 : 8;
#pragma unroll kUnrollT
    for (int t = 0; t < 8; ++t) {
      if (t < nt) {
        const int lc = L0 + 32 * t + par;
        // generate 32 l-steps; keep the 16 values of the parity class (even steps)
        auto emit = [&](int s, double v) {
          if ((s & 1) == 0) Pt[tid][s >> 1] = v;
        };
        if (64 * (wave % (RS / 64)) < kvalid) {
          enter_chunk(rec, lc);
          gen32<SPIN>(rec, cf + 32 * t + par, scf + (SPIN ? 32 * t + par : 0), lc, m, myf, emit);
        }
        __syncthreads();
        if (active) {
#pragma unroll
          for (int gj = 0; gj < 4 / KS; ++gj) {
            const int g = kh * (4 / KS) + gj;  // group of 16 K-steps = 64 rings
            if (64 * (g % (RS / 64)) < kvalid) {
#pragma unroll
              for (int kj = 16 * gj; kj < 16 * gj + 16; ++kj) {
                const double a = Pt[4 * (16 * g + kj - 16 * gj) + q][li];
                D[SPIN ? 0 : t] = mfma64(a, bfr[kj], D[SPIN ? 0 : t]);
              }
              // software pipeline of the A reads: the LDS pair read for MFMAs
              // k+4, k+5 is issued before MFMAs k, k+1 (two pairs in flight)
              __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // 2 LDS reads
#pragma unroll
              for (int j = 0; j < 6; ++j) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // 2 MFMAs
              }
              __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
            }
          }
        }
        __syncthreads();
      }
      if (SPIN) {  // rotate: the next chunk's accumulator moves to slot 0
        const d4 d0 = D[0];
#pragma unroll
        for (int i = 0; i < 7; ++i) D[i] = D[i + 1];
        D[7] = d0;
      }
    }
  }

  if (KS == 2) {
    // waves 2, 3 hand their partial sums to waves 0, 1 (Pt is free: the last
    // chunk ended with a barrier)
    double* red = &Pt[0][0];
    if (kh == 1) {
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[(((wave & 1) * 8 + t) * 4 + v) * 64 + lane] = D[t][v];
    }
    __syncthreads();
    if (kh == 1) return;
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) D[t][v] += red[(((wave & 1) * 8 + t) * 4 + v) * 64 + lane];
  }
  if (!active) return;
  double* alm = ocomp ? alm1 : alm0;
  const int sim = sim0 + li;
  if (sim >= nsim) return;
  const long long L = lmax + 1;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    if (t < nt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = q + 4 * v;
        const int l = L0 + par + 32 * t + 2 * row;
        if (l < L1) {
          const long long idx = oreim ? (long long)sim + nsim * ((l - m) + L * (L - m))
                                      : (long long)sim + nsim * (l + L * m);
          const double val = pa * D[t][v];
          alm[idx] = (beta != 0.0) ? val + beta * alm[idx] : val;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Synthesis: G[class c][ring][cols] = sum_{l in class} F(l, ring) * alm(l, cols)
//
// Workgroup = (m, class, 256 rings, NT 16-column tiles); its 4 waves are
// independent (64 rings each, one ring per lane) and never synchronise: each
// wave stages its own recurrence coefficients and its own P tile in LDS, so
// the two waves a SIMD holds drift apart and one's recurrence runs while the
// other's MFMAs do.  Per 32-step l chunk a wave
//   1. issues the global loads of the next chunk's alm operand (registers),
//   2. runs the plain FP64 recurrence for its rings (each entered at the chunk
//      holding its first kept l from a resident state), keeping the values of its class
//      (spin 2: F1 of its class and F2 of the other one) in LDS,
//   3. contracts them with 4 x NT x NKS v_mfma_f64_16x16x4 (A = P from LDS,
//      B = alm from registers), each accumulator's K chain back to back.
// ---------------------------------------------------------------------------
#ifndef FSHT_SYN_MINBLK0
#define FSHT_SYN_MINBLK0 2
#endif

// Code-motion barrier for LDS traffic between the lanes of one wave: a
// wave's LDS instructions execute in order, so no wait is needed, but the
// compiler must not move accesses across it.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef FSHT_SYN_SEG2
#define FSHT_SYN_SEG2 64
#endif
// spin 2 (4 column tiles per wave, like spin 0) asks for two workgroups per
// CU: the all-VGPR allocation spills a few registers but beats the
// accumulators-in-AGPRs allocation, which fits only one workgroup per CU
// (51 against 36 TFLOP/s; 2 column tiles per wave: 45)
#ifndef FSHT_SYN_MINBLK2
#define FSHT_SYN_MINBLK2 2
#endif
template <int SPIN>
struct SynShared {
  static constexpr int NV = SPIN ? 32 : 16;  // LDS values per ring per chunk
  // l values of recurrence coefficients staged per wave (spin 2 stages two
  // tables and a twice larger P tile, so it keeps a shorter segment to fit
  // two workgroups per CU)
  static constexpr int SEG = SPIN ? FSHT_SYN_SEG2 : kSegL;
  double Pt[4][64][NV + 1];
  double2 cf[4][SEG];
  double2 sf[4][SPIN ? SEG : 1];
};

template <int SPIN, int NT, int CLS>
__device__ inline void synth_wave(SynShared<SPIN>& sh, const LegItem it, const RingGeo* __restrict__ geo,
                                  const double2* __restrict__ abm, const double2* __restrict__ spm,
                                  const int* __restrict__ syn_st, const double2* __restrict__ syn_v,
                                  long long ys0, const double* __restrict__ alm0, const double* __restrict__ alm1,
                                  double* __restrict__ G, long long gbase, int i0, int nsim, int nsimp, int lmax,
                                  int nhr, int rbase) {
  constexpr int NV = SynShared<SPIN>::NV;
  constexpr int NKS = NV / 4;  // K-steps per chunk
  constexpr int NF = SPIN ? 2 : 1;
  const int m = it.m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, q = lane >> 4;
  const int ncols = 2 * (SPIN ? 2 : 1) * nsimp;
  const long long L = lmax + 1;
  double (*Pt)[NV + 1] = sh.Pt[wave];
  double2* cf = sh.cf[wave];
  double2* sf = sh.sf[wave];

  // ---- recurrence state of this lane's ring
  Rec rec;
  {
    const int myring = rbase + lane;
    const bool ring_ok = myring <= nhr;
    const int ic = ring_ok ? myring : nhr;
    const int st = ring_ok ? syn_st[ys0 + (ic - i0)] : 0x3fffffff;
    const double2 q = syn_v[ys0 + (ic - i0)];  // P at ls-2, ls-1
    rec.ls = (st > lmax) ? 0x3fffffff : m + ((st - m) & ~31);
    rec.sv = make_double2(q.y, q.x);
    const RingGeo g = geo[ic];
    rec.x = g.z;
    const double s2 = g.s * g.s;
    rec.bw = g.z / s2;
    rec.t2 = 2.0 / s2;
    rec.cw = (double)m / s2;
    rec.p1 = 0.0;
    rec.p2 = 0.0;
  }

  // ---- alm operand sources.  Element (l, sim) of tile u, function f sits at
  // byte bb[f][u] + 8*(sim + l*nsim): a uniform base (scalar registers) plus a
  // 32-bit lane offset.  Rows are clamped to l <= lmax (P is 0 beyond lmax)
  // and padding columns to sim < nsim, so every read is in bounds.  Spin 2
  // scales by the sign of the term (0 for padding and for Im at m = 0); spin 0
  // needs no scale: padding columns are never read back and Im at m = 0 is
  // stored as 0.
  const char* bb[NF][NT];
  double bs[NF][NT];
  int simc[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int c0 = (it.colblk + u) * 16;
    const bool cok = c0 < ncols;
    const int otype = cok ? c0 / nsimp : 0;
    const int sim = (cok ? c0 - otype * nsimp : 0) + li;
    simc[u] = sim < nsim ? sim : nsim - 1;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const Src s = syn_src(SPIN, otype, f);
      const bool im = s.reim && m > 0;
      const bool ok = cok && sim < nsim && !(m == 0 && s.reim);
      const double* a = s.comp ? alm1 : alm0;
      bb[f][u] = reinterpret_cast<const char*>(a + (long long)nsim * (im ? (L * (L - m) - m) : (L * m)));
      bs[f][u] = ok ? s.sign : 0.0;
    }
  }
  // K-step ks of the chunk at lc holds l = lc + pc(f) + 2*(4*(ks%4) + q)
  auto load_b = [&](int lc, double (&bf)[NKS][NT]) {
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int f = SPIN ? (ks / 4) : 0;
      const int pc = (f == 0) ? CLS : (CLS ^ 1);
      const int l = lc + pc + 8 * (ks % 4) + 2 * q;
      const unsigned row = (unsigned)nsim * (unsigned)(l <= lmax ? l : lmax);
#pragma unroll
      for (int u = 0; u < NT; ++u) {
        const double v = *reinterpret_cast<const double*>(bb[f][u] + 8u * (row + (unsigned)simc[u]));
        bf[ks][u] = SPIN ? v * bs[f][u] : v;
      }
    }
  };
  constexpr int SEG = SynShared<SPIN>::SEG;
  static_assert(SEG % 32 == 0 && SEG >= 32, "segment must hold whole chunks");
  auto load_seg = [&](int l0) {
#pragma unroll
    for (int k = lane; k < SEG; k += 64) {
      const int l = l0 + k;
      cf[k] = (l <= lmax) ? abm[l - m] : make_double2(0.0, 0.0);
      if (SPIN) sf[k] = (l <= lmax) ? spm[l - m] : make_double2(0.0, 0.0);
    }
  };

  d4 D[4][NT];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int u = 0; u < NT; ++u) D[a][u] = d4{0.0, 0.0, 0.0, 0.0};

  // first chunk: the earliest one at which a ring of the wave enters
  int lsmin = rec.ls;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsmin = min(lsmin, __shfl_xor(lsmin, o));
  lsmin = __builtin_amdgcn_readfirstlane(lsmin);  // wave-uniform: keeps the loop scalar
  const int lc0 = (lsmin > lmax) ? lmax + 1 : lsmin;

  // alm operand of the current chunk; the next chunk's is loaded into the
  // same registers right after this chunk's MFMAs are issued, so the loads
  // overlap their execution and the next recurrence (no second buffer, no copy)
  double bcur[NKS][NT];
  if ((lc0 - m) % SEG != 0 && lc0 <= lmax) {  // start inside a segment
    load_seg(lc0 - (lc0 - m) % SEG);
    wave_lds_sync();
  }
  load_b(lc0, bcur);
  for (int lc = lc0; lc <= lmax; lc += 32) {
    const int jseg = (lc - m) % SEG;
    if (jseg == 0) {
      wave_lds_sync();
      load_seg(lc);
      wave_lds_sync();
    }
    const bool more = lc + 32 <= lmax;

    // ---- 32 recurrence steps; step s has l - m parity s & 1 (lc - m is even)
    const double2* c = cf + jseg;
    const double2* sc = sf + (SPIN ? jseg : 0);
    enter_chunk(rec, lc);
    {
#pragma unroll
      for (int s = 0; s < 32; ++s) {
        const double2 cc = c[s];
        const double p = fma(cc.x * rec.x, rec.p1, -(cc.y * rec.p2));
        if (SPIN) {
          // F1 of this class at steps of its parity, F2 of the other class at the others
          if ((s & 1) == CLS)
            Pt[lane][s >> 1] = spin_f(0, p, rec.p1, sc[s], lc + s, m, rec.x, rec.bw, rec.t2, rec.cw);
          else
            Pt[lane][16 + (s >> 1)] = spin_f(1, p, rec.p1, sc[s], lc + s, m, rec.x, rec.bw, rec.t2, rec.cw);
        } else if ((s & 1) == CLS) {
          Pt[lane][s >> 1] = p;
        }
        rec.p2 = rec.p1;
        rec.p1 = p;
      }
    }
    wave_lds_sync();

    // ---- contraction: ring tile a (16 rings) x column tile u, K = NV values
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double av[NKS];
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        const int f = SPIN ? (ks / 4) : 0;
        av[ks] = Pt[16 * a + li][f * 16 + 4 * (ks % 4) + q];
      }
#pragma unroll
      for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) D[a][u] = mfma64(av[ks], bcur[ks][u], D[a][u]);
    }
    if (more) load_b(lc + 32, bcur);
    wave_lds_sync();
  }

#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ring = rbase + 16 * a + q + 4 * v;
      if (ring > nhr) continue;
#pragma unroll
      for (int u = 0; u < NT; ++u) {
        const int col = (it.colblk + u) * 16 + li;
        // spin 0: the Im columns of m = 0 carry nothing
        const bool zero = !SPIN && m == 0 && col >= nsimp;
        if (col < ncols) G[gbase + (long long)(ring - i0) * ncols + col] = zero ? 0.0 : D[a][u][v];
      }
    }
  }
}

template <int SPIN>
__global__ __launch_bounds__(256, SPIN ? FSHT_SYN_MINBLK2 : FSHT_SYN_MINBLK0) void k_synthesis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const long long* __restrict__ synoff, const int* __restrict__ syn_st, const double2* __restrict__ syn_v,
    const double* __restrict__ alm0, const double* __restrict__ alm1, double* __restrict__ G,
    const long long* __restrict__ goff, int nsim, int nsimp, int lmax, int nhr) {
  __shared__ SynShared<SPIN> sh;
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;
  if (!cols_active(it.colblk * 16, it.colblk * 16 + 64, nsim, nsimp)) return;  // no map in the block
  const int m = it.m;
  const int i0 = i0tab[m];
  const int nr = nhr + 1 - i0;
  const int rbase = i0 + it.blk * kSynRings + (threadIdx.x >> 6) * 64;
  // waves never synchronise with each other, so an idle wave may leave
  if (rbase > nhr) return;
  const int ncols = 2 * (SPIN ? 2 : 1) * nsimp;
  const long long gbase = goff[m] / (SPIN ? 1 : 2) + (long long)it.par * nr * ncols;
  const double2* abm = ab + aboff[m];
  const double2* spm = spc + aboff[m];
  const long long ys0 = synoff[m];
  auto run = [&](auto ntc, auto clsc) {
    synth_wave<SPIN, decltype(ntc)::value, decltype(clsc)::value>(sh, it, geo, abm, spm, syn_st, syn_v, ys0, alm0,
                                                                   alm1, G, gbase, i0, nsim, nsimp, lmax, nhr, rbase);
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  if (SPIN) {
    using N4 = std::integral_constant<int, 4>;
    if (it.par) run(N4{}, I1{}); else run(N4{}, I0{});
  } else {
    using N4 = std::integral_constant<int, 4>;
    using N2 = std::integral_constant<int, 2>;
    const bool full = (it.colblk + 4) * 16 <= ncols;
    if (full) {
      if (it.par) run(N4{}, I1{}); else run(N4{}, I0{});
    } else {
      if (it.par) run(N2{}, I1{}); else run(N2{}, I0{});
    }
  }
}

hipError_t launch_analysis(Handle* h, const double* F, double* alm0, double* alm1, double beta, int want0,
                           int want1) {
  const int sp = h->spin ? 1 : 0;
  if (!sp) { want0 = 1; want1 = 0; }
  for (int ks = 0; ks < 2; ++ks) {  // full column blocks, then 2-tile blocks
    const int n = h->n_ana_items[sp][ks];
    if (n == 0) continue;
    const LegItem* items = h->d_ana_items[sp][ks];
#define FSHT_LAUNCH_ANA(SPV, KSV)                                                                               \
  hipLaunchKernelGGL((k_analysis<SPV, KSV>), dim3(n), dim3(256), 0, h->stream, items, h->d_geo, h->d_i0, h->d_ab, \
                     h->d_spin, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, h->d_seed_p, F, h->d_goff,    \
                     alm0, alm1,                                                                                  \
                     h->nsim, h->nsimp, h->lmax, h->nhr, h->pa, beta, want0, want1)
    if (sp) {
      if (ks) FSHT_LAUNCH_ANA(2, 2); else FSHT_LAUNCH_ANA(2, 1);
    } else {
      if (ks) FSHT_LAUNCH_ANA(0, 2); else FSHT_LAUNCH_ANA(0, 1);
    }
#undef FSHT_LAUNCH_ANA
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_synthesis(Handle* h, const double* alm0, const double* alm1, double* G) {
  const int sp = h->spin ? 1 : 0;
  dim3 grid(h->n_syn_items[sp]);
  if (sp)
    hipLaunchKernelGGL(k_synthesis<2>, grid, dim3(256), 0, h->stream, h->d_syn_items[1], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_synoff, h->d_syn_st, h->d_syn_v, alm0, alm1, G, h->d_goff,
                       h->nsim, h->nsimp, h->lmax, h->nhr);
  else
    hipLaunchKernelGGL(k_synthesis<0>, grid, dim3(256), 0, h->stream, h->d_syn_items[0], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_synoff, h->d_syn_st, h->d_syn_v, alm0, alm0, G, h->d_goff,
                       h->nsim, h->nsimp, h->lmax, h->nhr);
  return hipGetLastError();
}

}  // namespace fsht
