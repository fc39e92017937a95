// Ring FFTs over HEALPix RING rings.
//
// Replaces map2fft / fft2map (src/compute_fft.f90:15-280), which run MKL DFTI
// on the host ring by ring: forward real FFT, unnormalised, PERM packing
// (slot 0 = Re X_0, slot 1 = Re X_{n/2}, slots 2k,2k+1 = Re/Im X_k), inverse
// unnormalised.  Here one workgroup transforms one north ring together with
// its south mirror for one map (z = x_N + i x_S, one complex FFT, then the two
// real spectra are separated).  Power-of-two rings use a radix-8 FFT in LDS;
// other polar rings (length 4q with q not a power of two, often with a large
// prime factor) are transformed one real ring per workgroup: the even/odd
// pixel pairs form a complex half-length sequence whose DFT comes from
// Bluestein's chirp-z convolution with a power-of-two FFT (filter length
// next_pow2(n - 1): at most 8192 points, 128 KB of LDS, at nside 2048).
#include <cmath>
#include <thread>

#include "fsht_internal.h"

namespace fsht {

static inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
static inline int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}
__host__ __device__ static inline int ilog2(int n) {
  int k = 0;
  while ((1 << k) < n) ++k;
  return k;
}

// ---------------------------------------------------------------------------
// Host: radix-2 FFT used once per handle to transform the Bluestein filters.
// ---------------------------------------------------------------------------
static void host_fft(std::vector<double2>& a, int sign) {
  const int n = (int)a.size();
  const int lg = ilog2(n);
  for (int i = 0; i < n; ++i) {
    unsigned r = 0;
    for (int b = 0; b < lg; ++b)
      if (i & (1 << b)) r |= 1u << (lg - 1 - b);
    if ((int)r > i) std::swap(a[i], a[r]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len / 2;
    for (int pos = 0; pos < half; ++pos) {
      const long double ang = (long double)sign * 2.0L * 3.141592653589793238462643383279502884L * pos / len;
      const double wr = (double)cosl(ang), wi = (double)sinl(ang);
      for (int i = pos; i < n; i += len) {
        const double2 u = a[i];
        const double2 v0 = a[i + half];
        const double2 v = make_double2(v0.x * wr - v0.y * wi, v0.x * wi + v0.y * wr);
        a[i] = make_double2(u.x + v.x, u.y + v.y);
        a[i + half] = make_double2(u.x - v.x, u.y - v.y);
      }
    }
  }
}

// Job list (grouped by LDS size class), twiddles and Bluestein tables.
// kMaxLds is the longest transform held in LDS (8192 complex doubles, 128 KB).
// A power-of-two ring is transformed with its mirror (one job per pair); a
// polar ring of length n = 4q (q not a power of two) is transformed one real
// ring at a time as a complex length-n/2 Bluestein transform of its even/odd
// pixel pairs (filter length next_pow2(n - 1), half the length-n transform's).
static constexpr int kMaxLds = 8192;

hipError_t build_bluestein(Handle* h) {
  const int nside = h->nside;
  const int nn = h->nnorth;
  struct J {
    int ring, n, M, L, half;  // L: Bluestein length (0 for a direct power-of-two transform)
  };
  std::vector<J> jobs;
  for (int i = 0; i < nn; ++i) {
    const int n = (i < nside) ? 4 * (i + 1) : 4 * nside;
    if (is_pow2(n)) {
      jobs.push_back({i, n, n, 0, 0});
    } else {
      const int L = n / 2;  // polar rings are never the equator, so both halves exist
      jobs.push_back({i, n, next_pow2(2 * L - 1), L, 1});
      jobs.push_back({i, n, next_pow2(2 * L - 1), L, 2});
    }
  }
  // Bluestein tables per distinct transform length
  std::vector<int> blue_of_job(jobs.size(), -1);
  h->blue_n.clear();
  h->blue_M.clear();
  h->blue_host.clear();
  long long coff = 0, foff = 0;
  for (size_t ji = 0; ji < jobs.size(); ++ji) {
    const J& j = jobs[ji];
    if (j.L == 0) continue;
    bool found = false;
    for (size_t b = 0; b < h->blue_n.size() && !found; ++b)
      if (h->blue_n[b] == j.L) {
        blue_of_job[ji] = (int)b;
        found = true;
      }
    if (found) continue;
    blue_of_job[ji] = (int)h->blue_n.size();
    h->blue_n.push_back(j.L);
    h->blue_M.push_back(j.M);
    h->blue_host.push_back(BlueInfo{coff, foff});
    coff += 2 * j.L + 1;  // chirp (L) and split roots (L + 1)
    foff += j.M;
  }
  const int nb = (int)h->blue_n.size();
  std::vector<double2> chirp(std::max(1LL, coff)), filt(std::max(1LL, foff));
  {
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t) {
      th.emplace_back([&, t]() {
        for (int b = (int)t; b < nb; b += (int)nth) {
          const int n = h->blue_n[b], M = h->blue_M[b];
          const long long c0 = h->blue_host[b].chirp_off, f0 = h->blue_host[b].filt_off;
          std::vector<double2> bx(M, make_double2(0.0, 0.0));
          for (int j = 0; j < n; ++j) {
            // w_j = exp(-i pi j^2/n), argument reduced exactly: j^2 mod 2n
            const long long r = ((long long)j * j) % (2LL * n);
            const long double ang = 3.141592653589793238462643383279502884L * (long double)r / (long double)n;
            const double c = (double)cosl(ang), s = (double)sinl(ang);
            chirp[c0 + j] = make_double2(c, -s);
            bx[j] = make_double2(c, s);  // conj(w_j)
            if (j) bx[M - j] = make_double2(c, s);
          }
          // split roots exp(-2 pi i k / (2n)) = exp(-pi i k / n), k = 0..n, after the
          // chirp (n here is the half-length transform's L)
          for (int k = 0; k <= n; ++k) {
            const long double ang = 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
            chirp[c0 + n + k] = make_double2((double)cosl(ang), -(double)sinl(ang));
          }
          host_fft(bx, -1);
          // stored in bit-reversed order: the kernels multiply the bit-reversed
          // output of the forward DIF pass position by position (coalesced reads)
          const int lgM = ilog2(M);
          for (int k = 0; k < M; ++k) {
            unsigned rk = 0;
            for (int b = 0; b < lgM; ++b)
              if (k & (1 << b)) rk |= 1u << (lgM - 1 - b);
            filt[f0 + k] = make_double2(bx[rk].x / M, bx[rk].y / M);
          }
        }
      });
    }
    for (auto& x : th) x.join();
  }
  // group jobs by LDS class
  // one class per transform length: kernels are compiled for their exact M
  const int classes[] = {4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192};
  h->fft_jobs.clear();
  h->fft_class_begin.clear();
  h->fft_class_M.clear();
  int lastM = 0;
  for (int ci = 0; ci < (int)(sizeof(classes) / sizeof(classes[0])); ++ci) {
    const int cM = classes[ci];
    h->fft_class_begin.push_back((int)h->fft_jobs.size());
    h->fft_class_M.push_back(cM);
    for (size_t ji = 0; ji < jobs.size(); ++ji) {
      const J& j = jobs[ji];
      if (j.M <= cM && j.M > lastM) h->fft_jobs.push_back(FftJob{j.ring, j.n, j.M, blue_of_job[ji], j.half});
    }
    lastM = cM;
  }
  h->fft_class_begin.push_back((int)h->fft_jobs.size());
  int Mmax = 2;
  for (auto& j : jobs) Mmax = std::max(Mmax, j.M);
  if (Mmax > kMaxLds) {
    h->err = "ring length needs an LDS transform longer than 8192 points (nside > 2048 not supported)";
    return hipErrorInvalidValue;
  }
  h->Mmax = Mmax;
  std::vector<double2> tw(Mmax);  // W_Mmax^k; a pass of span LEN reads k = p Mmax / LEN, p < LEN / 2^R
  for (int k = 0; k < Mmax; ++k) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * k / Mmax;
    tw[k] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  hipError_t e;
  auto up = [&](auto** d, const auto& v) -> hipError_t {
    using T = typename std::remove_reference<decltype(v[0])>::type;
    hipError_t ee = hipMalloc((void**)d, std::max<size_t>(1, v.size()) * sizeof(T));
    if (ee != hipSuccess) return ee;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  };
  if ((e = up(&h->d_fft_jobs, h->fft_jobs)) != hipSuccess) return e;
  if ((e = up(&h->d_blue, h->blue_host)) != hipSuccess) return e;
  if ((e = up(&h->d_chirp, chirp)) != hipSuccess) return e;
  if ((e = up(&h->d_filt, filt)) != hipSuccess) return e;
  if ((e = up(&h->d_twid, tw)) != hipSuccess) return e;
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Device FFT helpers (one workgroup, data in LDS).
// ---------------------------------------------------------------------------
// LDS position of element i of an M-point transform.  One pad element per 8
// puts the lanes of the span-8 pass (elements 8 apart, 128 B without the pad)
// on distinct banks; one more per M/8 does the same for indices read or
// written in bit-reversed order (8 consecutive indices differ only in their
// top 3 bits, M/8 elements apart).  Consecutive elements stay contiguous
// within each aligned group of 8.
template <int M>
__device__ inline int pz(int i) {
  constexpr int LG = (M >= 2) ? __builtin_ctz(M) : 0;
  if constexpr (LG >= 6)
    return i + (i >> 3) + (i >> (LG - 3));
  else
    return i + (i >> 3);
}
// LDS elements of an M-point transform: pz<M>(M - 1) < kLdsLen<M>
template <int M>
constexpr int kLdsLen = M + M / 8 + 8;
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// The transform length of an LDS class is a compile-time constant (every job
// of the class has M = the class size), so the passes are unrolled.  A pass
// applies R consecutive radix-2 stages (R = kFftR while that many remain,
// then the rest) to groups of 2^R elements held in registers: one LDS read
// and one LDS write per element for R stages.  The stages are those of the
// radix-2 transform, so the forward (DIF) output is in bit-reversed order and
// the inverse (DIT) takes bit-reversed input.  tw[k] = exp(-2 pi i k / Mmax)
// for k < Mmax.

__device__ inline double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }  // a * (-i)
__device__ inline double2 mul_pi(double2 a) { return make_double2(-a.y, a.x); }  // a * (+i)

// Stages per pass: 3 (radix 8), with 512 threads and registers capped for
// four waves per SIMD in the 2048- and 4096-point classes, so two 64 KB
// workgroups share a CU: 4 LDS round trips for 4096 points instead of 6 at
// radix 4 (16.6 / 18.0 against 18.6 / 18.5 ms forward / inverse, 200 maps at
// nside 1024).  Radix 16 takes 211-248 VGPRs, which leaves one 512-thread
// workgroup per CU (or idle threads at 256), and runs at 21-24 ms.
#ifndef FSHT_FFT_RMAX
#define FSHT_FFT_RMAX 3
#endif
#ifndef FSHT_FFT_TPB4096
#define FSHT_FFT_TPB4096 512
#endif
#ifndef FSHT_FFT_WPE
#define FSHT_FFT_WPE 4
#endif
#ifndef FSHT_FFT_TPB8192
#define FSHT_FFT_TPB8192 1024
#endif
constexpr int kFftR = FSHT_FFT_RMAX;  // radix-2 stages per pass (at most)

template <int M>
struct FftShape {
  static constexpr int LG = (M >= 2) ? __builtin_ctz(M) : 0;
  // threads per workgroup (the long transforms' LDS footprint leaves one or
  // two workgroups per CU)
  static constexpr int TPB = (M >= 8192) ? FSHT_FFT_TPB8192 : ((M >= 2048) ? FSHT_FFT_TPB4096 : 256);
  // __launch_bounds__ minimum waves per SIMD the register allocation must allow
  static constexpr int MINB = (M >= 8192) ? 1 : ((M >= 2048) ? FSHT_FFT_WPE : 2);
};

// exp(-2 pi i j / (2h)) for j < h <= 8 (the stage roots inside a pass)
__device__ inline double2 stage_root(int h, int j) {
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;  // cos, sin of pi/8
  constexpr double r2 = 0.70710678118654752440;                                // cos pi/4
  if (j == 0) return make_double2(1.0, 0.0);
  const int e = j * (8 / h);  // exponent in units of pi/8
  switch (e) {
    case 1: return make_double2(c1, -s1);
    case 2: return make_double2(r2, -r2);
    case 3: return make_double2(s1, -c1);
    case 4: return make_double2(0.0, -1.0);
    case 5: return make_double2(-s1, -c1);
    case 6: return make_double2(-r2, -r2);
    default: return make_double2(-c1, -s1);
  }
}

// DIF pass: stages of span LEN, LEN/2, ..., LEN/2^(R-1) on the groups
// {b*LEN + p + k*Q : k < 2^R}, Q = LEN / 2^R.  Stage t pairs k and k + h
// (h = 2^(R-1-t), bit h of k clear) with the twiddle W_S^(p + (k mod h) Q)
// = W_S^p * exp(-2 pi i (k mod h) / (2h)), S = LEN / 2^t, W_S^p = (W_LEN^p)^(2^t).
// dif_regs applies them to one group held in registers (x[k] = element k).
template <int M, int R, int LEN>
__device__ inline void dif_regs(double2 (&x)[1 << R], int p, int Mmax, const double2* __restrict__ tw) {
  constexpr int NR = 1 << R;
  double2 wb = tw[p * (Mmax / LEN)];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int h = NR >> (t + 1);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      if (k & h) continue;
      const int j = k & (h - 1);
      const double2 w = (j == 0) ? wb : cmul(wb, stage_root(h, j));
      const double2 u = x[k], v = x[k + h];
      x[k] = cadd(u, v);
      x[k + h] = cmul(csub(u, v), w);
    }
    if (t + 1 < R) wb = cmul(wb, wb);
  }
}

// DIT pass (inverse twiddles): stages of span 2Q, 4Q, ..., LEN on the same
// groups.  Stage t pairs k and k + h (h = 2^t) after multiplying element
// k + h by conj(W_S^(p + (k mod h) Q)), S = 2hQ, W_S^p = (W_LEN^p)^(2^(R-1-t)).
template <int M, int R, int LEN>
__device__ inline void dit_regs(double2 (&x)[1 << R], int p, int Mmax, const double2* __restrict__ tw) {
  constexpr int NR = 1 << R;
  double2 wp[R];
  wp[R - 1] = cconj(tw[p * (Mmax / LEN)]);
#pragma unroll
  for (int t = R - 2; t >= 0; --t) wp[t] = cmul(wp[t + 1], wp[t + 1]);
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int h = 1 << t;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      if (k & h) continue;
      const int j = k & (h - 1);
      const double2 w = (j == 0) ? wp[t] : cmul(wp[t], cconj(stage_root(h, j)));
      const double2 u = x[k], v = cmul(x[k + h], w);
      x[k] = cadd(u, v);
      x[k + h] = csub(u, v);
    }
  }
}

// One pass over every group of the transform in LDS.
template <int M, int R, int LEN, bool DIT>
__device__ inline void lds_pass(double2* buf, int Mmax, const double2* __restrict__ tw) {
  constexpr int NR = 1 << R, Q = LEN / NR, G = M / NR, TPB = FftShape<M>::TPB, NG = (G + TPB - 1) / TPB;
#pragma unroll
  for (int r = 0; r < NG; ++r) {
    const int g = threadIdx.x + TPB * r;
    if (G % TPB == 0 || g < G) {
      const int p = g & (Q - 1), i0 = (g - p) * NR + p;
      double2 x[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] = buf[pz<M>(i0 + k * Q)];
      if constexpr (DIT)
        dit_regs<M, R, LEN>(x, p, Mmax, tw);
      else
        dif_regs<M, R, LEN>(x, p, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR; ++k) buf[pz<M>(i0 + k * Q)] = x[k];
    }
  }
  __syncthreads();
}

// Radix-2 stages of the pass of span LEN: kFftR while that many remain, then the rest.
template <int LEN>
constexpr int pass_r() {
  return __builtin_ctz(LEN) < kFftR ? __builtin_ctz(LEN) : kFftR;
}
// Span of the last DIF pass (= the first DIT pass): every pass of that
// span has Q = 1, so a group is NR consecutive elements.
template <int M>
constexpr int kLastSpan = 1 << (__builtin_ctz(M) - kFftR * ((__builtin_ctz(M) - 1) / kFftR));

// DIF passes of span LEN, LEN >> R, ... while the span exceeds STOP.
template <int M, int LEN, int STOP>
__device__ inline void dif_range(double2* buf, int Mmax, const double2* __restrict__ tw) {
  if constexpr (LEN >= 2 && LEN > STOP) {
    constexpr int R = pass_r<LEN>();
    lds_pass<M, R, LEN, false>(buf, Mmax, tw);
    dif_range<M, (LEN >> R), STOP>(buf, Mmax, tw);
  }
}
// The DIT passes of the spans above START up to LEN, smallest span first.
template <int M, int LEN, int START>
__device__ inline void dit_range(double2* buf, int Mmax, const double2* __restrict__ tw) {
  if constexpr (LEN >= 2 && LEN > START) {
    constexpr int R = pass_r<LEN>();
    dit_range<M, (LEN >> R), START>(buf, Mmax, tw);
    lds_pass<M, R, LEN, true>(buf, Mmax, tw);
  }
}

// Decimation in frequency, forward (e^{-2 pi i jk/M}): natural in, bit-reversed out.
template <int M>
__device__ __attribute__((noinline)) void lds_dif(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dif_range<M, M, 1>(buf, Mmax, tw);
}

// Decimation in time, inverse (e^{+2 pi i jk/M}, unnormalised): bit-reversed in, natural out.
template <int M>
__device__ __attribute__((noinline)) void lds_dit_inv(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dit_range<M, M, 1>(buf, Mmax, tw);
}

// The long classes (first pass with one group per thread) run their first
// DIF pass on the loaded inputs and their last DIT pass into the outputs in
// registers, and the Bluestein transforms their last DIF pass, the filter
// and the first DIT pass in one LDS round trip: 7 instead of 10 for a
// Bluestein transform, 5 instead of 6 for a power-of-two pair.
// The register passes give each thread exactly one group of the first pass.
template <int M>
constexpr bool kRegEnds = M >= 1024 && (M >> pass_r<M>()) == FftShape<M>::TPB;

// DIF passes after the first.
template <int M>
__device__ __attribute__((noinline)) void lds_dif_tail(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dif_range<M, (M >> pass_r<M>()), 1>(buf, Mmax, tw);
}
// DIF passes between the first and the last.
template <int M>
__device__ __attribute__((noinline)) void lds_dif_mid(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dif_range<M, (M >> pass_r<M>()), kLastSpan<M>>(buf, Mmax, tw);
}
// DIT passes after the first (all the remaining ones).
template <int M>
__device__ __attribute__((noinline)) void lds_dit_tail(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dit_range<M, M, kLastSpan<M>>(buf, Mmax, tw);
}
// DIT passes before the last.
template <int M>
__device__ __attribute__((noinline)) void lds_dit_head(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dit_range<M, (M >> pass_r<M>()), 1>(buf, Mmax, tw);
}
// DIT passes between the first and the last.
template <int M>
__device__ __attribute__((noinline)) void lds_dit_mid(double2* buf, int Mmax, const double2* __restrict__ tw) {
  dit_range<M, (M >> pass_r<M>()), kLastSpan<M>>(buf, Mmax, tw);
}
// Last DIF pass, pointwise Bluestein filter (stored in bit-reversed order)
// and first DIT pass on each group of kLastSpan consecutive elements.
template <int M>
__device__ __attribute__((noinline)) void lds_dif_filter_dit(double2* buf, int Mmax, const double2* __restrict__ tw,
                                   const double2* __restrict__ fl) {
  constexpr int LEN = kLastSpan<M>, R = __builtin_ctz(LEN), NR = LEN, G = M / NR, TPB = FftShape<M>::TPB,
                NG = (G + TPB - 1) / TPB;
#pragma unroll
  for (int r = 0; r < NG; ++r) {
    const int g = threadIdx.x + TPB * r;
    if (G % TPB == 0 || g < G) {
      const int i0 = g * NR;
      double2 x[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] = buf[pz<M>(i0 + k)];
      dif_regs<M, R, LEN>(x, 0, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR; ++k) x[k] = cmul(x[k], fl[i0 + k]);
      dit_regs<M, R, LEN>(x, 0, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR; ++k) buf[pz<M>(i0 + k)] = x[k];
    }
  }
  __syncthreads();
}

__device__ inline int brev(int x, int lg) { return (int)(__brev((unsigned)x) >> (32 - lg)); }

// A table pointer the compiler cannot see through: loads through it stay
// after the LDS passes instead of being hoisted above them and held (and
// spilled) across the pass calls.
__device__ inline const double2* after_passes(const double2* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// The forward loops below step by the compile-time workgroup size and are
// unrolled, with their global reads made unconditional (indices clamped
// inside the ring), so all of a thread's loads are in flight together.  The
// inverse loops keep one element per iteration: unrolled, they ran slower.

// One real ring of length n = 2L (L not a power of two) transformed alone:
// y_j = x_{2j} + i x_{2j+1}, j < L, goes through a length-L Bluestein
// transform (chirp w, filter fl), and the real spectrum is split out of it
// (split roots exp(-2 pi i k / n), k = 0..L, stored after the chirp).
//   forward: A_k = (Y_k + conj Y_{L-k})/2, B_k = (Y_k - conj Y_{L-k})/(2i),
//            X_k = A_k + e^{-2 pi i k/n} B_k for k = 0..L (Y indices mod L);
//   inverse: Y'_k = (X_k + conj X_{L-k}) + i e^{+2 pi i k/n} (X_k - conj X_{L-k})
//            for k < L; y = unnormalised inverse DFT_L(Y'), x_{2j} = Re y_j,
//            x_{2j+1} = Im y_j.
// in / out point at the ring's first pixel (PERM spectrum or pixels); L <= M/2.
template <int M>
__device__ void ring_fft_half(double2* buf, int n, const double2* __restrict__ w, const double2* __restrict__ fl,
                              const double2* __restrict__ tw, int Mmax, const double* in, double* out,
                              int inverse) {
  constexpr int TPB = FftShape<M>::TPB, NS = (M + TPB - 1) / TPB, NH = (M / 2 + TPB) / TPB;
  const int L = n >> 1;
  const double2* sr = w + L;  // split roots exp(-2 pi i k / n), k = 0..L
  constexpr int R0 = pass_r<M>(), NR0 = 1 << R0, Q0 = M >> R0;
  if (!inverse) {
    if constexpr (kRegEnds<M>) {
      // first DIF pass on the loaded inputs: thread g holds elements g + k Q0
      const int g = threadIdx.x;
      double2 x[NR0];
#pragma unroll
      for (int k = 0; k < NR0; ++k) {
        const int j = g + k * Q0, jc = min(j, L - 1);
        const double2 y = cmul(make_double2(in[2 * jc], in[2 * jc + 1]), w[jc]);
        x[k] = (j < L) ? y : make_double2(0.0, 0.0);
      }
      dif_regs<M, R0, M>(x, g, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR0; ++k) buf[pz<M>(g + k * Q0)] = x[k];
      __syncthreads();
      lds_dif_mid<M>(buf, Mmax, tw);
      lds_dif_filter_dit<M>(buf, Mmax, tw, fl);
      lds_dit_tail<M>(buf, Mmax, tw);
    } else {
#pragma unroll
      for (int r = 0; r < NS; ++r) {
        const int j = threadIdx.x + r * TPB, jc = min(j, L - 1);
        const double2 y = cmul(make_double2(in[2 * jc], in[2 * jc + 1]), w[jc]);
        if (M % TPB == 0 || j < M) buf[pz<M>(j)] = (j < L) ? y : make_double2(0.0, 0.0);
      }
      __syncthreads();
      lds_dif<M>(buf, Mmax, tw);
#pragma unroll
      for (int r = 0; r < NS; ++r) {
        const int j = threadIdx.x + r * TPB;
        if (M % TPB == 0 || j < M) buf[pz<M>(j)] = cmul(buf[pz<M>(j)], fl[j]);
      }
      __syncthreads();
      lds_dit_inv<M>(buf, Mmax, tw);
    }
    const double2* wa = after_passes(w);
    // Lanes past k = L take k = L and store the value its own lane stores, so
    // the stores need no branch (a branch would hold each chirp load back
    // until the previous store's block).
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const int k = min((int)threadIdx.x + r * TPB, L);
      const int k1 = (k == L) ? 0 : k, k2 = (k == 0) ? 0 : L - k;
      const double2 Yk = cmul(buf[pz<M>(k1)], wa[k1]);
      const double2 Yc = cconj(cmul(buf[pz<M>(k2)], wa[k2]));
      const double2 A = make_double2(0.5 * (Yk.x + Yc.x), 0.5 * (Yk.y + Yc.y));
      const double2 B = make_double2(0.5 * (Yk.y - Yc.y), -0.5 * (Yk.x - Yc.x));
      const double2 X = cadd(A, cmul(B, wa[L + k]));
      // PERM: slot 0 = Re X_0, slot 1 = Re X_L, slots 2k, 2k+1 = Re, Im X_k
      const bool mid = k > 0 && k < L;
      const int ia = (k == 0) ? 0 : (k == L ? 1 : 2 * k);
      out[ia] = X.x;
      out[mid ? 2 * k + 1 : ia] = mid ? X.y : X.x;
    }
  } else {
    // the Bluestein input Z_k, k < M (zero from L on)
    auto zin = [&](int k) -> double2 {
      if (k >= L) return make_double2(0.0, 0.0);
      const int kc = L - k;  // 1..L
      // slot 0 holds Re X_0, slot 1 Re X_L (both with Im 0)
      const double2 Xk = (k == 0) ? make_double2(in[0], 0.0) : make_double2(in[2 * k], in[2 * k + 1]);
      const double2 Xc = (kc == L) ? make_double2(in[1], 0.0) : make_double2(in[2 * kc], -in[2 * kc + 1]);
      const double2 D = cmul(csub(Xk, Xc), cconj(sr[k]));
      const double2 Yp = cadd(cadd(Xk, Xc), mul_pi(D));
      return cmul(cconj(Yp), w[k]);
    };
    if constexpr (kRegEnds<M>) {
      const int g = threadIdx.x;
      double2 x[NR0];
#pragma unroll
      for (int k = 0; k < NR0; ++k) x[k] = zin(g + k * Q0);
      dif_regs<M, R0, M>(x, g, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR0; ++k) buf[pz<M>(g + k * Q0)] = x[k];
      __syncthreads();
      lds_dif_mid<M>(buf, Mmax, tw);
      lds_dif_filter_dit<M>(buf, Mmax, tw, fl);
      lds_dit_mid<M>(buf, Mmax, tw);
      // last DIT pass into the outputs: thread g holds elements g + k Q0
#pragma unroll
      for (int k = 0; k < NR0; ++k) x[k] = buf[pz<M>(g + k * Q0)];
      dit_regs<M, R0, M>(x, g, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR0; ++k) {
        const int j = g + k * Q0;
        if (j < L) {
          const double2 y = cconj(cmul(x[k], w[j]));
          out[2 * j] = y.x;
          out[2 * j + 1] = y.y;
        }
      }
    } else {
      for (int k = threadIdx.x; k < M; k += TPB) buf[pz<M>(k)] = zin(k);
      __syncthreads();
      lds_dif<M>(buf, Mmax, tw);
#pragma unroll
      for (int r = 0; r < NS; ++r) {
        const int j = threadIdx.x + r * TPB;
        if (M % TPB == 0 || j < M) buf[pz<M>(j)] = cmul(buf[pz<M>(j)], fl[j]);
      }
      __syncthreads();
      lds_dit_inv<M>(buf, Mmax, tw);
      for (int j = threadIdx.x; j < L; j += TPB) {
        const double2 y = cconj(cmul(buf[pz<M>(j)], w[j]));
        out[2 * j] = y.x;
        out[2 * j + 1] = y.y;
      }
    }
  }
}

// Forward: maps (sim-major, npix per map) -> spec (same layout, PERM per ring).
// Inverse: spec -> maps.  One workgroup per (map, job within the class).
// Maps staged from the host are transformed in place (in == out), so the map
// and spectrum pointers are not restrict-qualified.  A job is either a
// power-of-two ring with its mirror (n = M, z = x_N + i x_S; the equator
// ring has no mirror and reads itself as S) or one half-length Bluestein ring.
template <int MAXM>
__global__ __launch_bounds__(FftShape<MAXM>::TPB, FftShape<MAXM>::MINB) void k_ring_fft(const FftJob* __restrict__ jobs, int job0,
                                                  const RingGeo* __restrict__ geo, const BlueInfo* __restrict__ blue,
                                                  const double2* __restrict__ chirp, const double2* __restrict__ filt,
                                                  const double2* __restrict__ tw, int Mmax, long long npix, int nsim,
                                                  const double* in0, const double* in1, double* out0,
                                                  double* out1, int inverse,
                                                  int map0) {
  __shared__ double2 buf[kLdsLen<MAXM>];
  constexpr int M = MAXM;  // every job of a class has M equal to the class size
  constexpr int TPB = FftShape<M>::TPB, NS = (M + TPB - 1) / TPB, NH = (M / 2 + TPB) / TPB;
  constexpr int lg = FftShape<M>::LG;
  const FftJob jb = jobs[job0 + blockIdx.y];
  const int mapi = map0 + blockIdx.x;
  const int comp = mapi / nsim, k = mapi - comp * nsim;
  const double* in = (comp ? in1 : in0) + (long long)k * npix;
  double* out = (comp ? out1 : out0) + (long long)k * npix;
  const RingGeo g = geo[jb.ring];
  const bool pair = g.i1S != g.i1N;
  if constexpr (M >= 8) {
    if (jb.half) {
      const long long i1 = (jb.half == 1) ? g.i1N : g.i1S;
      const BlueInfo bi = blue[jb.blue];
      ring_fft_half<M>(buf, jb.n, chirp + bi.chirp_off, filt + bi.filt_off, tw, Mmax, in + i1, out + i1, inverse);
      return;
    }
  }
  // power-of-two pair: n = M
  const double* inN = in + g.i1N;
  const double* inS = in + g.i1S;
  double* outN = out + g.i1N;
  double* outS = out + g.i1S;
  if (!inverse) {
    // ---- load z = x_N + i x_S
    if constexpr (kRegEnds<M>) {
      // first DIF pass on the loaded inputs: thread g holds elements g + k Q0
      constexpr int R0 = pass_r<M>(), NR0 = 1 << R0, Q0 = M >> R0;
      const int g = threadIdx.x;
      double2 x[NR0];
#pragma unroll
      for (int k = 0; k < NR0; ++k) {
        const double xN = inN[g + k * Q0], xS = inS[g + k * Q0];
        x[k] = make_double2(xN, pair ? xS : 0.0);
      }
      dif_regs<M, R0, M>(x, g, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR0; ++k) buf[pz<M>(g + k * Q0)] = x[k];
      __syncthreads();
      lds_dif_tail<M>(buf, Mmax, tw);
    } else {
#pragma unroll
      for (int r = 0; r < NS; ++r) {
        const int j = threadIdx.x + r * TPB;
        if (M % TPB == 0 || j < M) {
          const double xN = inN[j], xS = inS[j];
          buf[pz<M>(j)] = make_double2(xN, pair ? xS : 0.0);
        }
      }
      __syncthreads();
      lds_dif<M>(buf, Mmax, tw);
    }
    // ---- separate the two real spectra (bit-reversed order in LDS), write PERM
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const int kq = (int)threadIdx.x + r * TPB, kk = min(kq, M / 2);
      const int nk = (M - kk) & (M - 1);
      const double2 Zk = buf[pz<M>(brev(kk, lg))];
      const double2 Zn = buf[pz<M>(brev(nk, lg))];
      const double2 XN = make_double2(0.5 * (Zk.x + Zn.x), 0.5 * (Zk.y - Zn.y));
      const double2 XS = make_double2(0.5 * (Zk.y + Zn.y), -0.5 * (Zk.x - Zn.x));
      if (kq <= M / 2) {
        if (kk == 0) {
          outN[0] = XN.x;
          if (pair) outS[0] = XS.x;
        } else if (kk == M / 2) {
          outN[1] = XN.x;
          if (pair) outS[1] = XS.x;
        } else {
          outN[2 * kk] = XN.x;
          outN[2 * kk + 1] = XN.y;
          if (pair) {
            outS[2 * kk] = XS.x;
            outS[2 * kk + 1] = XS.y;
          }
        }
      }
    }
  } else {
    // ---- build Z_k = X_N[k] + i X_S[k], k < M, into bit-reversed positions
    for (int kk = threadIdx.x; kk < M; kk += TPB) {
      const int kh = kk <= M / 2 ? kk : M - kk;
      double2 XN, XS = make_double2(0.0, 0.0);
      if (kh == 0) {  // slot 0 holds Re X_0, slot 1 Re X_{M/2} (both with Im 0)
        XN = make_double2(inN[0], 0.0);
        if (pair) XS = make_double2(inS[0], 0.0);
      } else if (kh == M / 2) {
        XN = make_double2(inN[1], 0.0);
        if (pair) XS = make_double2(inS[1], 0.0);
      } else {
        XN = make_double2(inN[2 * kh], inN[2 * kh + 1]);
        if (pair) XS = make_double2(inS[2 * kh], inS[2 * kh + 1]);
      }
      if (kk > M / 2) {
        XN.y = -XN.y;
        XS.y = -XS.y;
      }
      buf[pz<M>(brev(kk, lg))] = make_double2(XN.x - XS.y, XN.y + XS.x);
    }
    __syncthreads();
    if constexpr (kRegEnds<M>) {
      // last DIT pass into the outputs: thread g holds elements g + k Q0
      constexpr int R0 = pass_r<M>(), NR0 = 1 << R0, Q0 = M >> R0;
      lds_dit_head<M>(buf, Mmax, tw);
      const int g = threadIdx.x;
      double2 x[NR0];
#pragma unroll
      for (int k = 0; k < NR0; ++k) x[k] = buf[pz<M>(g + k * Q0)];
      dit_regs<M, R0, M>(x, g, Mmax, tw);
#pragma unroll
      for (int k = 0; k < NR0; ++k) {
        outN[g + k * Q0] = x[k].x;
        if (pair) outS[g + k * Q0] = x[k].y;
      }
    } else {
      lds_dit_inv<M>(buf, Mmax, tw);
      for (int j = threadIdx.x; j < M; j += TPB) {
        const double2 z = buf[pz<M>(j)];
        outN[j] = z.x;
        if (pair) outS[j] = z.y;
      }
    }
  }
}

hipError_t launch_ring_fft(Handle* h, const double* const* in_comp, double* spec, int inverse,
                           double* const* out_comp, int map0, int nmap) {
  // forward: in_comp = user maps per component, out = spec (+ comp offset)
  // inverse: in = spec, out_comp = user maps per component
  const int ncomp = h->spin ? 2 : 1;
  const long long cstride = (long long)h->nsim * h->npix;
  const double* in0 = inverse ? spec : in_comp[0];
  const double* in1 = inverse ? spec + cstride : (ncomp > 1 ? in_comp[1] : in_comp[0]);
  double* out0 = inverse ? out_comp[0] : spec;
  double* out1 = inverse ? (ncomp > 1 ? out_comp[1] : out_comp[0]) : spec + cstride;
  const int nmaps = nmap > 0 ? nmap : ncomp * h->nsim;
  if (nmap <= 0) map0 = 0;
  for (size_t c = 0; c + 1 < h->fft_class_begin.size(); ++c) {
    const int b = h->fft_class_begin[c], e = h->fft_class_begin[c + 1];
    if (e <= b) continue;
    // maps fastest: consecutive workgroups transform the same ring, so its
    // Bluestein chirp and filter stay in L2 across the maps
    dim3 grid(nmaps, e - b);
    const int M = h->fft_class_M[c];
#define FSHT_LAUNCH_FFT(MM)                                                                                      \
  hipLaunchKernelGGL(k_ring_fft<MM>, grid, dim3(FftShape<MM>::TPB), 0, h->stream, h->d_fft_jobs, b, h->d_geo,    \
                     h->d_blue,                                                                                   \
                     h->d_chirp, h->d_filt, h->d_twid, h->Mmax, h->npix, h->nsim, in0, in1, out0, out1, inverse,   \
                     map0)
    switch (M) {
      case 4: FSHT_LAUNCH_FFT(4); break;
      case 8: FSHT_LAUNCH_FFT(8); break;
      case 16: FSHT_LAUNCH_FFT(16); break;
      case 32: FSHT_LAUNCH_FFT(32); break;
      case 64: FSHT_LAUNCH_FFT(64); break;
      case 128: FSHT_LAUNCH_FFT(128); break;
      case 256: FSHT_LAUNCH_FFT(256); break;
      case 512: FSHT_LAUNCH_FFT(512); break;
      case 1024: FSHT_LAUNCH_FFT(1024); break;
      case 2048: FSHT_LAUNCH_FFT(2048); break;
      case 4096: FSHT_LAUNCH_FFT(4096); break;
      default: FSHT_LAUNCH_FFT(8192); break;
    }
#undef FSHT_LAUNCH_FFT
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

}  // namespace fsht
