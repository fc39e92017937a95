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
// batches, and the seed injection (2 compares + 2 selects per step) only runs
// in the 32-step chunks where some lane of the wave actually starts.
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

// Spin-2 functions from P_l and P_{l-1} (src/get_plm_recursive.f90:111-124).
__device__ inline double spin_f(int f, double P, double Pm1, double2 sp, int l, int m, double x, double bw,
                                double two_on_s2, double cw) {
  const double lam = sp.x, nl = sp.y;
  const double dl = (double)l;
  if (f == 0) {
    const double lm2 = dl - (double)m * (double)m;
    return (Pm1 * lam * bw - P * (two_on_s2 * lm2 + dl * (dl - 1.0))) * nl;
  }
  return (Pm1 * lam - P * (2.0 * x * (dl - 1.0))) * cw * nl;
}

// Per-thread recurrence state for one ring.
struct Rec {
  double p1, p2;  // P_{l-1}, P_{l-2}
  double x, bw, t2, cw;
  int st;         // first non-negligible l (seeds injected at st, st+1)
  double2 sv;     // P_st, P_st+1
};

// Run 32 recurrence steps l = lc .. lc+31 with coefficients cf[0..31]
// (zero beyond lmax) and hand every value (P, or F_f for spin 2) to emit(s, v).
template <int SPIN, bool INJ, class Emit>
__device__ inline void gen32(Rec& r, const double2* cf, const double2* scf, int lc, int m, int f, Emit&& emit) {
#pragma unroll
  for (int s = 0; s < 32; ++s) {
    const double2 c = cf[s];
    double p = fma(c.x * r.x, r.p1, -(c.y * r.p2));
    if (INJ) {
      const int l = lc + s;
      p = (l == r.st) ? r.sv.x : p;
      p = (l == r.st + 1) ? r.sv.y : p;
    }
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
// recurrence restarted from the resident seeds) into LDS.
// ---------------------------------------------------------------------------

// Workgroups per CU the register allocation must allow (spin 0 / spin 2).
#ifndef FSHT_ANA_MINBLK0
#define FSHT_ANA_MINBLK0 1
#endif
#ifndef FSHT_SYN_MINBLK0
#define FSHT_SYN_MINBLK0 2
#endif

template <int SPIN>
__global__ __launch_bounds__(256, SPIN ? 1 : FSHT_ANA_MINBLK0) void k_analysis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const long long* __restrict__ seedoff, const int* __restrict__ seed_st, const double2* __restrict__ seed_v,
    const double* __restrict__ F, const long long* __restrict__ goff, double* __restrict__ alm0,
    double* __restrict__ alm1, int nsim, int nsimp, int lmax, int nhr, double pa, double beta, int want0,
    int want1) {
  constexpr int RS = SPIN ? kAnaRings2 : kAnaRings0;  // rings per sub-block
  constexpr int NCOMP = SPIN ? 2 : 1;
  __shared__ double Pt[256][17];
  __shared__ double2 cf[kAnaLB];
  __shared__ double2 scf[SPIN ? kAnaLB : 1];
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;
  const int m = it.m, par = it.par;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, q = lane >> 4;
  const int ncols = 2 * NCOMP * nsimp;
  const int i0 = i0tab[m];
  const int nr = nhr + 1 - i0;
  const int L0 = m + it.blk * kAnaLB;
  const int L1 = min(L0 + kAnaLB, lmax + 1);
  const int nt = (L1 - L0 - par + 31) / 32;  // l-tiles in this block
  // this wave's output column tile
  const int c0 = (it.colblk * 4 + wave) * 16;
  const bool colok = c0 < ncols;
  const int otype = colok ? c0 / nsimp : 0;
  const int sim0 = colok ? c0 - otype * nsimp : 0;
  const int ocomp = otype >> 1, oreim = otype & 1;
  bool active = colok && !(m == 0 && oreim == 1);
  if (SPIN) active = active && (ocomp == 0 ? want0 : want1);

  // recurrence coefficients of the block (zero beyond lmax: P stays 0 there)
  {
    const double2* abm = ab + aboff[m];
    const double2* spm = spc + aboff[m];
    for (int k = tid; k < kAnaLB; k += 256) {
      const int l = L0 + k;
      cf[k] = (l <= lmax) ? abm[l - m] : make_double2(0.0, 0.0);
      if (SPIN) scf[k] = (l <= lmax) ? spm[l - m] : make_double2(0.0, 0.0);
    }
  }

  d4 D[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) D[t] = d4{0.0, 0.0, 0.0, 0.0};

  const int myf = SPIN ? (tid / RS) : 0;
  const int myr = SPIN ? (tid % RS) : tid;

  const long long fb = goff[m] / (SPIN ? 1 : 2);
  for (int rb = 0; rb < nr; rb += RS) {
    // ---- this thread's recurrence state (its ring of the sub-block)
    Rec rec;
    {
      const int i = i0 + rb + myr;
      const bool valid = i <= nhr;
      const int ic = valid ? i : nhr;
      const RingGeo g = geo[ic];
      const long long so = seedoff[m] + (long long)it.blk * nr + (ic - i0);
      rec.st = valid ? seed_st[so] : 0x3fffffff;
      rec.sv = seed_v[so];
      rec.x = g.z;
      const double s2 = g.s * g.s;
      rec.bw = g.z / s2;
      rec.t2 = 2.0 / s2;
      rec.cw = (double)m / s2;
      rec.p1 = 0.0;
      rec.p2 = 0.0;
    }
    // ---- B fragments: the wave's 16 columns of the folded, rotated ring
    // spectra for the sub-block's rings (one coalesced row read per ring)
    double bfr[64];
#pragma unroll
    for (int ks = 0; ks < 64; ++ks) {
      const int krow = 4 * ks + q;
      const int f = SPIN ? (ks >= 32 ? 1 : 0) : 0;
      const int r = SPIN ? (krow % RS) : krow;
      const int i = i0 + rb + r;
      const bool valid = i <= nhr;
      const int ic = valid ? i : nhr;
      const Src s = ana_src(SPIN, otype, f);
      const int fold = par ^ s.flip;
      const int col = (s.comp * 2 + s.reim) * nsimp + sim0 + li;
      const double v = F[fb + ((long long)fold * nr + (ic - i0)) * ncols + col];
      bfr[ks] = (active && valid) ? s.sign * v : 0.0;
    }

#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t < nt) {
        const int lc = L0 + 32 * t;
        // generate 32 l-steps; keep the 16 values of the parity class
        double keep = 0.0;
        auto emit = [&](int s, double v) {
          if ((s & 1) == 0) {
            keep = v;
          } else {
            Pt[tid][s >> 1] = par ? v : keep;
          }
        };
        const bool inj = (rec.st >= lc - 1) && (rec.st <= lc + 31);
        if (__any(inj))
          gen32<SPIN, true>(rec, cf + 32 * t, scf + (SPIN ? 32 * t : 0), lc, m, myf, emit);
        else
          gen32<SPIN, false>(rec, cf + 32 * t, scf + (SPIN ? 32 * t : 0), lc, m, myf, emit);
        __syncthreads();
        if (active) {
#pragma unroll
          for (int ks = 0; ks < 64; ++ks) {
            const double a = Pt[4 * ks + q][li];
            D[t] = mfma64(a, bfr[ks], D[t]);
          }
        }
        __syncthreads();
      }
    }
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
// Synthesis: G[class c][ring][cols] = sum_{l} F(l, ring) * alm(l, cols)
// Workgroup = (m, class, 256 rings, 64 columns); 4 waves, each owning 64 rings
// (its own recurrence, one ring per lane) and the 64 columns.  The alm
// operands of chunk c+1 are loaded while chunk c's recurrence runs.
// ---------------------------------------------------------------------------
template <int SPIN>
__global__ __launch_bounds__(256, SPIN ? 1 : FSHT_SYN_MINBLK0) void k_synthesis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const long long* __restrict__ seedoff, const int* __restrict__ seed_st, const double2* __restrict__ seed_v,
    const double* __restrict__ alm0, const double* __restrict__ alm1, double* __restrict__ G,
    const long long* __restrict__ goff, int nsim, int nsimp, int lmax, int nhr) {
  constexpr int NCOMP = SPIN ? 2 : 1;
  constexpr int NV = SPIN ? 32 : 16;  // values per lane per 32-step chunk
  constexpr int NKS = NV / 4;         // K-steps per chunk
  __shared__ double Pt[4][64][NV + 1];
  __shared__ double2 cf[2][32];
  __shared__ double2 scf[2][SPIN ? 32 : 1];
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;
  const int m = it.m, cls = it.par;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, q = lane >> 4;
  const int ncols = 2 * NCOMP * nsimp;
  const int i0 = i0tab[m];
  const int nr = nhr + 1 - i0;
  const int rbase = i0 + it.blk * kSynRings + wave * 64;
  const int myring = rbase + lane;
  const bool ring_ok = myring <= nhr;
  const bool wave_ok = rbase <= nhr;
  const double2* abm = ab + aboff[m];
  const double2* spm = spc + aboff[m];
  const long long L = lmax + 1;

  // column tiles of this workgroup and their alm sources per function
  int otype[4], sim0[4];
  bool cok[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int c0 = (it.colblk * 4 + u) * 16;
    cok[u] = c0 < ncols;
    otype[u] = cok[u] ? c0 / nsimp : 0;
    sim0[u] = cok[u] ? c0 - otype[u] * nsimp : 0;
  }

  Rec rec;
  {
    const int ic = ring_ok ? myring : nhr;
    const long long so = seedoff[m] + (ic - i0);
    rec.st = ring_ok ? seed_st[so] : 0x3fffffff;
    rec.sv = seed_v[so];
    const RingGeo g = geo[ic];
    rec.x = g.z;
    const double s2 = g.s * g.s;
    rec.bw = g.z / s2;
    rec.t2 = 2.0 / s2;
    rec.cw = (double)m / s2;
    rec.p1 = 0.0;
    rec.p2 = 0.0;
  }

  d4 D[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) D[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  // branch-free alm operand load for K-step ks of the chunk starting at lc
  auto load_b = [&](int lc, double (&bf)[NKS][4]) {
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int f = SPIN ? (ks / 4) : 0;
      const int vi = 4 * (ks % 4) + q;
      const int pc = (f == 0) ? cls : (cls ^ 1);
      const int l = lc + pc + 2 * vi;
      const bool lok = l <= lmax;
      const int lcl = lok ? l : lmax;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const Src s = syn_src(SPIN, otype[u], f);
        const int sim = sim0[u] + li;
        // Im at m = 0 has no storage: read the (valid) Re address and drop it
        const bool im = s.reim && m > 0;
        const bool ok = cok[u] && lok && sim < nsim && !(m == 0 && s.reim);
        const int simc = sim < nsim ? sim : nsim - 1;
        const double* alm = s.comp ? alm1 : alm0;
        const long long idx = im ? (long long)simc + nsim * ((lcl - m) + L * (L - m))
                                 : (long long)simc + nsim * (lcl + L * m);
        const double v = alm[idx];
        bf[ks][u] = ok ? s.sign * v : 0.0;
      }
    }
  };
  auto load_cf = [&](int lc, int buf) {
    if (tid < 32) {
      const int l = lc + tid;
      cf[buf][tid] = (l <= lmax) ? abm[l - m] : make_double2(0.0, 0.0);
      if (SPIN) scf[buf][tid] = (l <= lmax) ? spm[l - m] : make_double2(0.0, 0.0);
    }
  };

  double bcur[NKS][4], bnext[NKS][4];
  load_b(m, bcur);
  load_cf(m, 0);
  __syncthreads();
  int buf = 0;
  for (int lc = m; lc <= lmax; lc += 32, buf ^= 1) {
    const bool more = lc + 32 <= lmax;
    if (more) {
      load_b(lc + 32, bnext);
      load_cf(lc + 32, buf ^ 1);
    }
    // ---- recurrence for this lane's ring over 32 l-steps; step s has (l - m)
    // parity s & 1, so each (even, odd) pair of steps yields one value per class
    double keep = 0.0;
    auto emit = [&](int s, double v) {
      if ((s & 1) == 0)
        keep = v;
      else
        Pt[wave][lane][s >> 1] = cls ? v : keep;
    };
    if (SPIN) {
      // F1 uses class cls, F2 the other class: generate both functions
      double k1 = 0.0, k2 = 0.0;
      auto emit2 = [&](int s, double v1, double v2) {
        if ((s & 1) == 0) {
          k1 = v1;
          k2 = v2;
        } else {
          Pt[wave][lane][(s >> 1)] = cls ? v1 : k1;            // F1, class cls
          Pt[wave][lane][16 + (s >> 1)] = cls ? k2 : v2;       // F2, other class
        }
      };
      const bool inj = (rec.st >= lc - 1) && (rec.st <= lc + 31);
      const double2* c = cf[buf];
      const double2* sc = scf[buf];
      auto body = [&](auto injc) {
        constexpr bool INJ = decltype(injc)::value;
#pragma unroll
        for (int s = 0; s < 32; ++s) {
          const double2 cc = c[s];
          double p = fma(cc.x * rec.x, rec.p1, -(cc.y * rec.p2));
          if (INJ) {
            const int l = lc + s;
            p = (l == rec.st) ? rec.sv.x : p;
            p = (l == rec.st + 1) ? rec.sv.y : p;
          }
          const double f1 = spin_f(0, p, rec.p1, sc[s], lc + s, m, rec.x, rec.bw, rec.t2, rec.cw);
          const double f2 = spin_f(1, p, rec.p1, sc[s], lc + s, m, rec.x, rec.bw, rec.t2, rec.cw);
          emit2(s, f1, f2);
          rec.p2 = rec.p1;
          rec.p1 = p;
        }
      };
      if (__any(inj))
        body(std::true_type{});
      else
        body(std::false_type{});
    } else {
      const bool inj = (rec.st >= lc - 1) && (rec.st <= lc + 31);
      if (__any(inj))
        gen32<0, true>(rec, cf[buf], nullptr, lc, m, 0, emit);
      else
        gen32<0, false>(rec, cf[buf], nullptr, lc, m, 0, emit);
    }
    __syncthreads();
    if (wave_ok) {
      // one ring tile at a time: its 4 accumulators rotate through the chunk's
      // K-steps.  FP64 MFMA throughput drops when many distinct accumulators
      // are cycled (register-file reads of srcC); 4 in rotation keep it high.
#pragma unroll
      for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          const int f = SPIN ? (ks / 4) : 0;
          const double av = Pt[wave][16 * a + li][f * 16 + 4 * (ks % 4) + q];
#pragma unroll
          for (int u = 0; u < 4; ++u) D[a][u] = mfma64(av, bcur[ks][u], D[a][u]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    if (more) {
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
        for (int u = 0; u < 4; ++u) bcur[ks][u] = bnext[ks][u];
    }
  }

  if (!wave_ok) return;
  const long long gbase = goff[m] / (SPIN ? 1 : 2) + (long long)cls * nr * ncols;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ring = rbase + 16 * a + q + 4 * v;
      if (ring > nhr) continue;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!cok[u]) continue;
        const int col = (it.colblk * 4 + u) * 16 + li;
        G[gbase + (long long)(ring - i0) * ncols + col] = D[a][u][v];
      }
    }
  }
}

hipError_t launch_analysis(Handle* h, const double* F, double* alm0, double* alm1, double beta, int want0,
                           int want1) {
  const int sp = h->spin ? 1 : 0;
  dim3 grid(h->n_ana_items[sp]);
  if (sp)
    hipLaunchKernelGGL(k_analysis<2>, grid, dim3(256), 0, h->stream, h->d_ana_items[1], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, F, h->d_goff, alm0, alm1,
                       h->nsim, h->nsimp, h->lmax, h->nhr, h->pa, beta, want0, want1);
  else
    hipLaunchKernelGGL(k_analysis<0>, grid, dim3(256), 0, h->stream, h->d_ana_items[0], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, F, h->d_goff, alm0, alm1,
                       h->nsim, h->nsimp, h->lmax, h->nhr, h->pa, beta, 1, 0);
  return hipGetLastError();
}

hipError_t launch_synthesis(Handle* h, const double* alm0, const double* alm1, double* G) {
  const int sp = h->spin ? 1 : 0;
  dim3 grid(h->n_syn_items[sp]);
  if (sp)
    hipLaunchKernelGGL(k_synthesis<2>, grid, dim3(256), 0, h->stream, h->d_syn_items[1], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, alm0, alm1, G, h->d_goff,
                       h->nsim, h->nsimp, h->lmax, h->nhr);
  else
    hipLaunchKernelGGL(k_synthesis<0>, grid, dim3(256), 0, h->stream, h->d_syn_items[0], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, alm0, alm0, G, h->d_goff,
                       h->nsim, h->nsimp, h->lmax, h->nhr);
  return hipGetLastError();
}

}  // namespace fsht
