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

// One recurrence step with start-seed injection.
__device__ inline double rec_step(double2 c, double x, double p1, double p2, int l, int st, double2 v) {
  double p = fma(c.x * x, p1, -(c.y * p2));
  p = (l == st) ? v.x : p;
  p = (l == st + 1) ? v.y : p;
  return p;
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

// ---------------------------------------------------------------------------
// Analysis: alm[l in class par, cols] = pa * sum_{north rings} F(l, ring) * X(ring, cols)
// Workgroup = (m, parity class, l-block of kAnaLB, 64 columns); 4 waves, each
// owning 16 columns and all 8 l-tiles of the block.  Rings are consumed in
// sub-blocks: the workgroup gathers its X columns for the sub-block into
// registers (PERM slot gather + e^{-i m phi0} + north/south fold, i.e. the
// reference's fft2mat), then sweeps the l-block generating P (one ring per
// thread, plain FP64 recurrence restarted from the resident seeds) into LDS.
// ---------------------------------------------------------------------------
template <int SPIN>
__global__ __launch_bounds__(256) void k_analysis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const double2* __restrict__ tw, const int* __restrict__ slot, const long long* __restrict__ seedoff,
    const int* __restrict__ seed_st, const double2* __restrict__ seed_v, const double* __restrict__ specT,
    double* __restrict__ alm0, double* __restrict__ alm1, int nsim, int nsimp, int lmax, int nhr, int nnorth,
    double pa, double beta, int want0, int want1) {
  constexpr int RS = SPIN ? kAnaRings2 : kAnaRings0;  // rings per sub-block
  constexpr int NCOMP = SPIN ? 2 : 1;
  __shared__ double Pt[256][17];
  const LegItem it = items[blockIdx.x];
  if (it.m < 0) return;
  const int m = it.m, par = it.par;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, q = lane >> 4;
  const int ldT = NCOMP * nsimp;
  const int ncols = 2 * NCOMP * nsimp;
  const int i0 = i0tab[m];
  const int nr = nhr + 1 - i0;
  const int L0 = m + it.blk * kAnaLB;
  const int L1 = min(L0 + kAnaLB, lmax + 1);
  const int nt = (L1 - L0 - par + 31) / 32;  // l-tiles in this block
  // this wave's output column tile
  const int ct = it.colblk * 4 + wave;
  const int c0 = ct * 16;
  const int otype = c0 / nsimp;
  const int sim0 = c0 - otype * nsimp;
  const int ocomp = otype >> 1, oreim = otype & 1;
  bool active = (c0 < ncols) && !(m == 0 && oreim == 1);
  if (SPIN) active = active && (ocomp == 0 ? want0 : want1);
  const long long mrow = (long long)m * nnorth;
  const double2* abm = ab + aboff[m];
  const double2* spm = spc + aboff[m];

  d4 D[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) D[t] = d4{0.0, 0.0, 0.0, 0.0};

  for (int rb = 0; rb < nr; rb += RS) {
    // ---- gather B fragments (X columns) for the sub-block into registers
    double bfr[64];
#pragma unroll
    for (int ks = 0; ks < 64; ++ks) {
      const int krow = 4 * ks + q;
      const int f = SPIN ? (krow / RS) : 0;
      const int r = SPIN ? (krow % RS) : krow;
      const int i = i0 + rb + r;
      double val = 0.0;
      if (active && i <= nhr) {
        const Src s = ana_src(SPIN, otype, f);
        const int fold = par ^ s.flip;
        const int sinfo = slot[mrow + i];
        const int rel = sinfo >> 2, cls = sinfo & 3;
        const double sg = cls == 1 ? 1.0 : (cls == 2 ? -1.0 : 0.0);
        const double2 cs = tw[mrow + i];
        const RingGeo g = geo[i];
        const int col = s.comp * nsimp + sim0 + li;
        const double* pN = specT + (g.i1N + rel) * (long long)ldT + col;
        const double xr = pN[0];
        const double xi = cls ? pN[ldT] : 0.0;
        // e^{-i m phi0} rotation with conjugation for negative frequencies
        double Fr = xr * cs.x + sg * xi * cs.y;
        double Fi = sg * xi * cs.x - xr * cs.y;
        if (g.i1S != g.i1N) {
          const double* pS = specT + (g.i1S + rel) * (long long)ldT + col;
          const double yr = pS[0];
          const double yi = cls ? pS[ldT] : 0.0;
          const double Gr = yr * cs.x + sg * yi * cs.y;
          const double Gi = sg * yi * cs.x - yr * cs.y;
          if (fold) {
            Fr -= Gr;
            Fi -= Gi;
          } else {
            Fr += Gr;
            Fi += Gi;
          }
        }
        val = s.sign * (s.reim ? Fi : Fr);
      }
      bfr[ks] = val;
    }

    // ---- per-thread recurrence state for its (ring, function)
    const int myf = SPIN ? (tid / RS) : 0;
    const int myr = SPIN ? (tid % RS) : tid;
    const int mi = i0 + rb + myr;
    const bool ring_ok = mi <= nhr;
    int st = 0x3fffffff;
    double2 sv = make_double2(0.0, 0.0);
    double x = 0.0, bw = 0.0, t2 = 0.0, cw = 0.0;
    if (ring_ok) {
      const long long so = seedoff[m] + (long long)it.blk * nr + (mi - i0);
      st = seed_st[so];
      sv = seed_v[so];
      const RingGeo g = geo[mi];
      x = g.z;
      const double s2 = g.s * g.s;
      bw = x / s2;
      t2 = 2.0 / s2;
      cw = (double)m / s2;
    }
    double p1 = 0.0, p2 = 0.0;

#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t < nt) {
        // generate 32 l-steps, keep the 16 values of the parity class
        const int lc = L0 + 32 * t;
#pragma unroll 4
        for (int s = 0; s < 32; ++s) {
          const int l = lc + s;
          double p = 0.0;
          if (l <= lmax) {
            const double2 c = abm[l - m];
            p = rec_step(c, x, p1, p2, l, st, sv);
            if ((s & 1) == par) {
              double val = p;
              if (SPIN) val = spin_f(myf, p, p1, spm[l - m], l, m, x, bw, t2, cw);
              Pt[tid][s >> 1] = val;
            }
          } else if ((s & 1) == par) {
            Pt[tid][s >> 1] = 0.0;
          }
          p2 = p1;
          p1 = p;
        }
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
// (its own recurrence, one ring per lane) and the 64 columns.
// ---------------------------------------------------------------------------
template <int SPIN>
__global__ __launch_bounds__(256) void k_synthesis(
    const LegItem* __restrict__ items, const RingGeo* __restrict__ geo, const int* __restrict__ i0tab,
    const double2* __restrict__ ab, const double2* __restrict__ spc, const long long* __restrict__ aboff,
    const long long* __restrict__ seedoff, const int* __restrict__ seed_st, const double2* __restrict__ seed_v,
    const double* __restrict__ alm0, const double* __restrict__ alm1, double* __restrict__ G,
    const long long* __restrict__ goff, int nsim, int nsimp, int lmax, int nhr) {
  constexpr int NCOMP = SPIN ? 2 : 1;
  constexpr int NV = SPIN ? 32 : 16;  // values per lane per 32-step chunk
  constexpr int NKS = NV / 4;         // K-steps per chunk
  __shared__ double Pt[4][64][NV + 1];
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

  // column tiles of this workgroup
  int otype[4], sim0[4];
  bool cok[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int c0 = (it.colblk * 4 + u) * 16;
    otype[u] = c0 / nsimp;
    sim0[u] = c0 - otype[u] * nsimp;
    cok[u] = c0 < ncols;
  }

  int st = 0x3fffffff;
  double2 sv = make_double2(0.0, 0.0);
  double x = 0.0, bw = 0.0, t2 = 0.0, cw = 0.0;
  if (ring_ok) {
    const long long so = seedoff[m] + (myring - i0);
    st = seed_st[so];
    sv = seed_v[so];
    const RingGeo g = geo[myring];
    x = g.z;
    const double s2 = g.s * g.s;
    bw = x / s2;
    t2 = 2.0 / s2;
    cw = (double)m / s2;
  }
  const double2* abm = ab + aboff[m];
  const double2* spm = spc + aboff[m];
  const long long L = lmax + 1;

  d4 D[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) D[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  double p1 = 0.0, p2 = 0.0;
  for (int lc = m; lc <= lmax; lc += 32) {
    // ---- generate this lane's ring over 32 l-steps
#pragma unroll 4
    for (int s = 0; s < 32; ++s) {
      const int l = lc + s;
      double p = 0.0;
      const int pclass = s & 1;  // (l - m) parity
      if (l <= lmax) {
        p = rec_step(abm[l - m], x, p1, p2, l, st, sv);
        if (SPIN) {
          // F1 in slots [0,16) for l in class cls, F2 in [16,32) for the other class
          const int f = (pclass == cls) ? 0 : 1;
          Pt[wave][lane][f * 16 + (s >> 1)] = spin_f(f, p, p1, spm[l - m], l, m, x, bw, t2, cw);
        } else if (pclass == cls) {
          Pt[wave][lane][s >> 1] = p;
        }
      } else {
        if (SPIN)
          Pt[wave][lane][((pclass == cls) ? 0 : 16) + (s >> 1)] = 0.0;
        else if (pclass == cls)
          Pt[wave][lane][s >> 1] = 0.0;
      }
      p2 = p1;
      p1 = p;
    }
    __syncthreads();
    if (wave_ok) {
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        const int f = SPIN ? (ks / 4) : 0;
        const int vi = 4 * (ks % 4) + q;  // value index within the function's 16
        // l of this K-row: class cls for f=0 (P or F1), the other class for F2
        const int pc = (f == 0) ? cls : (cls ^ 1);
        const int l = lc + pc + 2 * vi;
        double bf[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          double val = 0.0;
          if (cok[u] && l <= lmax) {
            const Src s = syn_src(SPIN, otype[u], f);
            const int sim = sim0[u] + li;
            if (sim < nsim && !(m == 0 && s.reim)) {
              const double* alm = s.comp ? alm1 : alm0;
              const long long idx = s.reim ? (long long)sim + nsim * ((l - m) + L * (L - m))
                                           : (long long)sim + nsim * (l + L * m);
              val = s.sign * alm[idx];
            }
          }
          bf[u] = val;
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const double av = Pt[wave][16 * a + li][f * 16 + 4 * (ks % 4) + q];
#pragma unroll
          for (int u = 0; u < 4; ++u) D[a][u] = mfma64(av, bf[u], D[a][u]);
        }
      }
    }
    __syncthreads();
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

hipError_t launch_analysis(Handle* h, const double* specT, double* alm0, double* alm1, double beta, int want0,
                           int want1) {
  const int sp = h->spin ? 1 : 0;
  dim3 grid(h->n_ana_items[sp]);
  if (sp)
    hipLaunchKernelGGL(k_analysis<2>, grid, dim3(256), 0, h->stream, h->d_ana_items[1], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_tw, h->d_slot, h->d_seedoff, h->d_seed_st, h->d_seed_v, specT,
                       alm0, alm1, h->nsim, h->nsimp, h->lmax, h->nhr, h->nnorth, h->pa, beta, want0, want1);
  else
    hipLaunchKernelGGL(k_analysis<0>, grid, dim3(256), 0, h->stream, h->d_ana_items[0], h->d_geo, h->d_i0, h->d_ab,
                       h->d_spin, h->d_aboff, h->d_tw, h->d_slot, h->d_seedoff, h->d_seed_st, h->d_seed_v, specT,
                       alm0, alm1, h->nsim, h->nsimp, h->lmax, h->nhr, h->nnorth, h->pa, beta, 1, 0);
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
