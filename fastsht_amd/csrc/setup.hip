// Handle set-up: ring geometry, ring-skip table, Legendre recurrence constants,
// gather/twiddle tables, Legendre start seeds (device kernel), work lists.
//
// Replaces the reference's sht_data_alloc / sht_set_data chain
// (src/sht_data_init.f90:2-311, 383-686) and its plm seed tables
// (plm_table/*.bin read at SHT.py:34-45): everything here is computed from the
// HEALPix closed form and the normalised-Legendre recurrence.
#include <algorithm>
#include <cmath>
#include <thread>

#include "fsht_internal.h"

namespace fsht {

static constexpr double kPi = 3.141592653589793238462643383279502884;

// ---------------------------------------------------------------------------
// HEALPix RING geometry (Gorski et al. 2005).  Ring r = i+1 counted from the
// north pole.  Pixel offsets/lengths as in src/sht_data_init.f90:655-675.
// ---------------------------------------------------------------------------
void ring_geometry(int nside, std::vector<RingGeo>& geo, std::vector<double>& phi0) {
  const int nring = 4 * nside - 1;
  const int nn = 2 * nside;
  geo.resize(nn);
  phi0.resize(nn);
  auto i1_of = [&](long long i) -> long long {
    const long long ns = nside;
    if (i < ns) return 2 * i * (i + 1);
    if (i < 3 * ns - 1) return 2 * ns + 4 * ns * i - 2 * ns * ns;
    return 12 * ns * ns - 2 * (4 * ns - i - 1) * (4 * ns - i);
  };
  for (int i = 0; i < nn; ++i) {
    const int r = i + 1;
    double z, s;
    if (r < nside) {
      const double t = ((double)r / nside) * ((double)r / nside) / 3.0;
      z = 1.0 - t;
      s = std::sqrt(t * (2.0 - t));
      phi0[i] = kPi / (4.0 * r);
    } else {
      z = (4.0 * nside - 2.0 * r) / (3.0 * nside);
      s = std::sqrt((1.0 - z) * (1.0 + z));
      phi0[i] = ((r - nside) % 2 == 0) ? kPi / (4.0 * nside) : 0.0;
    }
    RingGeo g;
    g.i1N = i1_of(i);
    g.i1S = i1_of(nring - 1 - i);
    g.nfft = (i < nside) ? 4 * (i + 1) : 4 * nside;
    g.pad_ = 0;
    g.z = z;
    g.s = s;
    geo[i] = g;
  }
}

// ---------------------------------------------------------------------------
// Recurrence constants (src/get_plm_recursive.f90:31-37, sht_data_init.f90:401-416).
// ---------------------------------------------------------------------------
static void diag_fac(int mmax, std::vector<double>& out) {
  out.resize(mmax + 1);
  out[0] = 1.0;
  for (int m = 1; m <= mmax; ++m) out[m] = out[m - 1] * (2.0 * m + 1.0) / 2.0 / (double)m;
}

static void ab_coeffs(int m, int lmax, double* a, double* b) {
  const double mm = (double)m * (double)m;
  a[0] = 0.0;
  b[0] = -1.0;
  for (int l = m + 1; l <= lmax; ++l) {
    const double dl = (double)l;
    const double c1 = (dl - 1.0) * (dl - 1.0);
    const double c2 = (2.0 * dl - 3.0) * (2.0 * dl - 1.0);
    const double c3 = (2.0 * dl - 1.0) * (2.0 * dl + 1.0);
    const double c4 = dl * dl;
    const double al = std::sqrt(c3 / (c4 - mm));
    a[l - m] = al;
    b[l - m] = std::sqrt((c1 - mm) / c2) * al;
  }
}

// x^n as mantissa in [0.5,1) and binary exponent, by binary powering.
static void pow_scaled_host(double x, int n, double* mant, long* ex) {
  int ex0;
  double base = std::frexp(x, &ex0);
  long be = ex0;
  double r = 1.0;
  long re = 0;
  while (n > 0) {
    if (n & 1) {
      int e;
      r = std::frexp(r * base, &e);
      re += e + be;
    }
    n >>= 1;
    if (n) {
      int e;
      base = std::frexp(base * base, &e);
      be = 2 * be + e;
    }
  }
  *mant = r;
  *ex = re;
}

// max_{m<=l<=lcap} log2|P_lm| for one ring, scaled recurrence.
static double ring_maxlog2(int m, int lcap, double z, double s, const double* dfac, const double* a,
                           const double* b) {
  double ms;
  long e;
  pow_scaled_host(s, m, &ms, &e);
  double p1 = std::sqrt(dfac[m] / 4.0 / kPi) * ms;
  if (m & 1) p1 = -p1;
  int fe;
  p1 = std::frexp(p1, &fe);
  e += fe;
  double p2 = 0.0;
  double best = -1e300;
  for (int l = m; l <= lcap; ++l) {
    double p;
    if (l == m) {
      p = p1;
    } else {
      p = std::fma(a[l - m] * z, p1, -(b[l - m] * p2));
      p2 = p1;
      p1 = p;
    }
    if (e < 0 && std::fabs(p1) > 0x1p100) {
      long sh = (-e < 100) ? -e : 100;
      p1 = std::ldexp(p1, (int)-sh);
      p2 = std::ldexp(p2, (int)-sh);
      e += sh;
      p = p1;
    }
    if (p != 0.0) {
      double lg = std::log2(std::fabs(p)) + (double)e;
      if (lg > best) best = lg;
    }
  }
  return best;
}

// Ring-skip table: first north ring whose max_{m<=l<=lcap}|P_lm| reaches
// 1e-20, lcap = 4*nside (nside <= 1024) or 6143 (nside 2048); these are the
// rules that reproduce SHT_ring_skip_table/*.bin.  Bisection over rings (the
// maximum grows monotonically toward the equator).
void ring_tab(int nside, int nm, const std::vector<RingGeo>& geo, std::vector<int>& out) {
  const int lcap = (nside <= 1024) ? 4 * nside : 6143;
  const double thr2 = -20.0 / 0.30102999566398119521;
  std::vector<double> dfac;
  diag_fac(std::max(nm, 1), dfac);
  out.assign(nm, 0);
  const int nn = 2 * nside;
  unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) {
    th.emplace_back([&, t]() {
      std::vector<double> a(lcap + 2), b(lcap + 2);
      for (int m = (int)t; m < nm; m += (int)nth) {
        const int lc = std::max(lcap, m);
        ab_coeffs(m, lc, a.data(), b.data());
        int lo = 0, hi = nn;
        while (lo < hi) {
          int mid = (lo + hi) / 2;
          double lg = ring_maxlog2(m, lc, geo[mid].z, geo[mid].s, dfac.data(), a.data(), b.data());
          if (lg >= thr2)
            hi = mid;
          else
            lo = mid + 1;
        }
        out[m] = lo;
      }
    });
  }
  for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Device: Legendre start states.  The Legendre kernels run the plain FP64
// recurrence in chunks of 32 l-steps on the grid l - m = 0 (mod 32).  For
// order m, kept north ring i and analysis l-block b (l in [m + b*LB,
// m + (b+1)*LB)), store st, the first l of the block at which P_lm is kept
// (>= 2^-100), and P at ls-2, ls-1, ls with ls the chunk start at or below
// st.  A recurrence entered at ls (or at ls+1 from P_ls, P_ls-1) reproduces
// the exponent-rescaled recurrence bit for bit: rescaling is by powers of two
// and these values are >= 2^-300 (at most 33 steps of growth below 2^-100).
// At ls = m the state is (P_mm, 0, P_mm): the coefficients at l = m are
// (a, b) = (0, -1), so the recurrence yields P_mm there.  The synthesis state
// is the one of the chunk holding the ring's first kept l over the whole range.
// ---------------------------------------------------------------------------
__device__ static void pow_scaled_dev(double x, int n, double* mant, long long* ex) {
  int ex0;
  double base = frexp(x, &ex0);
  long long be = ex0;
  double r = 1.0;
  long long re = 0;
  while (n > 0) {
    if (n & 1) {
      int e;
      r = frexp(r * base, &e);
      re += e + be;
    }
    n >>= 1;
    if (n) {
      int e;
      base = frexp(base * base, &e);
      be = 2 * be + e;
    }
  }
  *mant = r;
  *ex = re;
}

__global__ void k_seeds(int lmax, int nhr, const RingGeo* __restrict__ geo, const int* __restrict__ i0,
                        const double* __restrict__ dfac, const double2* __restrict__ ab,
                        const long long* __restrict__ aboff, const long long* __restrict__ seedoff,
                        int* __restrict__ seed_st, double2* __restrict__ seed_v, double* __restrict__ seed_p,
                        const long long* __restrict__ synoff, int* __restrict__ syn_st,
                        double2* __restrict__ syn_v) {
  const int m = blockIdx.y;
  const int i = i0[m] + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nhr) return;
  const int nr = nhr + 1 - i0[m];
  const int nb = (lmax + 1 - m + kAnaLB - 1) / kAnaLB;
  const long long base = seedoff[m] + (i - i0[m]);
  for (int b = 0; b < nb; ++b) {
    seed_st[base + (long long)b * nr] = 0x3fffffff;
    seed_v[base + (long long)b * nr] = make_double2(0.0, 0.0);
    seed_p[base + (long long)b * nr] = 0.0;
  }
  const long long ys = synoff[m] + (i - i0[m]);
  syn_st[ys] = 0x3fffffff;
  syn_v[ys] = make_double2(0.0, 0.0);
  const double z = geo[i].z, s = geo[i].s;
  const double2* abm = ab + aboff[m];
  double ms;
  long long e;
  pow_scaled_dev(s, m, &ms, &e);
  double p1 = sqrt(dfac[m] / 4.0 / kPi) * ms;
  if (m & 1) p1 = -p1;
  int fe;
  p1 = frexp(p1, &fe);
  e += fe;
  double p2 = 0.0;
  int first = -1;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;  // P at ls-2, ls-1, ls of the latest chunk start ls
  double n0 = 0.0, n1 = 0.0;            // P at the two steps before the next chunk start
  for (int l = m; l <= lmax; ++l) {
    double p;
    if (l == m) {
      p = p1;
    } else {
      const double2 c = abm[l - m];
      p = fma(c.x * z, p1, -(c.y * p2));
      p2 = p1;
      p1 = p;
    }
    if (e < 0 && fabs(p1) > 0x1p100) {
      long long sh = (-e < 100) ? -e : 100;
      p1 = ldexp(p1, (int)-sh);
      p2 = ldexp(p2, (int)-sh);
      e += sh;
      p = p1;
    }
    const int jc = (l - m) & 31;
    if (l == m) {
      q2 = ldexp(p, (int)max(e, -2000LL));
      q0 = q2;
      q1 = 0.0;
    } else if (jc == 30) {
      n0 = ldexp(p, (int)max(e, -2000LL));
    } else if (jc == 31) {
      n1 = ldexp(p, (int)max(e, -2000LL));
    } else if (jc == 0) {
      q0 = n0;
      q1 = n1;
      q2 = ldexp(p, (int)max(e, -2000LL));
    }
    if (first < 0 && p != 0.0) {
      int fp;
      frexp(p, &fp);
      if ((long long)fp + e >= kKeepExp) first = l;
    }
    if (first < 0) continue;
    if (l == first) {
      syn_st[ys] = first;
      syn_v[ys] = make_double2(q0, q1);
    }
    const int b = (l - m) / kAnaLB;
    const int L0 = m + b * kAnaLB;
    if (l == (first > L0 ? first : L0)) {  // start of this block's kept run
      seed_st[base + (long long)b * nr] = l;
      seed_v[base + (long long)b * nr] = make_double2(q0, q1);
      seed_p[base + (long long)b * nr] = q2;
    }
  }
}

hipError_t launch_seeds(Handle* h) {
  dim3 grid((h->nnorth + 255) / 256, h->lmax + 1);
  // diag_fac table lives right after the seeds in a temporary buffer
  std::vector<double> dfac;
  diag_fac(h->lmax + 1, dfac);
  double* d_dfac = nullptr;
  hipError_t e = hipMalloc(&d_dfac, dfac.size() * sizeof(double));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(d_dfac, dfac.data(), dfac.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_seeds, grid, dim3(256), 0, h->stream, h->lmax, h->nhr, h->d_geo, h->d_i0, d_dfac,
                     h->d_ab, h->d_aboff, h->d_seedoff, h->d_seed_st, h->d_seed_v, h->d_seed_p, h->d_synoff, h->d_syn_st,
                     h->d_syn_v);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(h->stream);
  (void)hipFree(d_dfac);
  return e;
}

// Work bounds from the seeds, for the work lists.  Per (m, analysis l-block):
// the first ring (from i0(m)) with a kept P_lm in the block, nr(m) if none.
// Per (m, group of 64 rings = one synthesis wave): the group's first kept l.
__global__ __launch_bounds__(256) void k_starts(int lmax, int nhr, const int* __restrict__ i0,
                                               const long long* __restrict__ seedoff,
                                               const int* __restrict__ seed_st, const long long* __restrict__ synoff,
                                               const int* __restrict__ syn_st, const long long* __restrict__ blkoff,
                                               int* __restrict__ ana_r0, const long long* __restrict__ grpoff,
                                               int* __restrict__ syn_l0) {
  __shared__ int red[256];
  const int m = blockIdx.x;
  const int nr = nhr + 1 - i0[m];
  const int nb = (lmax + 1 - m + kAnaLB - 1) / kAnaLB;
  for (int b = 0; b < nb; ++b) {
    const int* st = seed_st + seedoff[m] + (long long)b * nr;
    int r0 = nr;
    for (int r = threadIdx.x; r < nr; r += 256)
      if (st[r] != 0x3fffffff) r0 = min(r0, r);
    red[threadIdx.x] = r0;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) red[threadIdx.x] = min(red[threadIdx.x], red[threadIdx.x + k]);
      __syncthreads();
    }
    if (threadIdx.x == 0) ana_r0[blkoff[m] + b] = red[0];
    __syncthreads();
  }
  const int ng = (nr + 63) / 64;
  for (int g = threadIdx.x; g < ng; g += 256) {
    int l0 = 0x3fffffff;
    const int e = min(nr, 64 * g + 64);
    for (int r = 64 * g; r < e; ++r) l0 = min(l0, syn_st[synoff[m] + r]);
    syn_l0[grpoff[m] + g] = l0;
  }
}

static hipError_t compute_starts(Handle* h, std::vector<long long>& blkoff, std::vector<int>& ana_r0,
                                 std::vector<long long>& grpoff, std::vector<int>& syn_l0) {
  const int lmax = h->lmax;
  blkoff.assign(lmax + 2, 0);
  grpoff.assign(lmax + 2, 0);
  for (int m = 0; m <= lmax; ++m) {
    const int nr = h->nhr + 1 - h->i0[m];
    blkoff[m + 1] = blkoff[m] + h->nblk[m];
    grpoff[m + 1] = grpoff[m] + (nr + 63) / 64;
  }
  ana_r0.assign(blkoff[lmax + 1], 0);
  syn_l0.assign(grpoff[lmax + 1], 0);
  long long* d_off = nullptr;
  int* d_out = nullptr;
  const size_t noff = blkoff.size() + grpoff.size(), nout = ana_r0.size() + syn_l0.size();
  hipError_t e = hipMalloc(&d_off, noff * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc(&d_out, std::max<size_t>(1, nout) * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(d_off, blkoff.data(), blkoff.size() * sizeof(long long), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_off + blkoff.size(), grpoff.data(), grpoff.size() * sizeof(long long), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_starts, dim3(lmax + 1), dim3(256), 0, h->stream, lmax, h->nhr, h->d_i0, h->d_seedoff,
                       h->d_seed_st, h->d_synoff, h->d_syn_st, d_off, d_out, d_off + blkoff.size(),
                       d_out + ana_r0.size());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && !ana_r0.empty())
    e = hipMemcpy(ana_r0.data(), d_out, ana_r0.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && !syn_l0.empty())
    e = hipMemcpy(syn_l0.data(), d_out + ana_r0.size(), syn_l0.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (d_off) (void)hipFree(d_off);
  if (d_out) (void)hipFree(d_out);
  return e;
}

// ---------------------------------------------------------------------------
// Host tables that depend on (nside, lmax): (a,b), spin constants, gather
// slots and twiddles, seed offsets, synthesis-buffer offsets, work lists.
// ---------------------------------------------------------------------------
struct HostTables {
  std::vector<RingGeo> geo;
  std::vector<double> phi0;
  std::vector<double2> ab, spin, tw;
  std::vector<int> slot;
};

static void fill_ab_spin(Handle* h, HostTables& t) {
  const int lmax = h->lmax;
  h->aboff.resize(lmax + 2);
  long long off = 0;
  for (int m = 0; m <= lmax; ++m) {
    h->aboff[m] = off;
    off += lmax + 1 - m;
  }
  h->aboff[lmax + 1] = off;
  t.ab.resize(off);
  t.spin.resize(off);
  std::vector<double> a(lmax + 1), b(lmax + 1);
  for (int m = 0; m <= lmax; ++m) {
    ab_coeffs(m, lmax, a.data(), b.data());
    const int m1 = std::max(1, m);
    for (int l = m; l <= lmax; ++l) {
      t.ab[h->aboff[m] + l - m] = make_double2(a[l - m], b[l - m]);
      // spin-2 constants (src/get_plm_recursive.f90:105-107, sht_data_init.f90:398-400)
      double lam = 0.0;
      if (l > m1) lam = 2.0 * std::sqrt((double)(2 * l + 1) / (2 * l - 1) * (double)((long long)l * l - (long long)m * m));
      double nl = 0.0;
      if (l >= 2) nl = 1.0 / std::sqrt((double)(l + 2) * (double)(l + 1) * (double)l * (double)(l - 1));
      t.spin[h->aboff[m] + l - m] = make_double2(lam, nl);
    }
  }
}

// Gather slot class per (m, north ring): relative PERM slot and the sign of
// the imaginary part (src/sht_data_init.f90:448-490 tab/fac).
static void fill_slots(Handle* h, HostTables& t) {
  const int lmax = h->lmax, nn = h->nnorth;
  t.slot.resize((size_t)(lmax + 1) * nn);
  t.tw.resize((size_t)(lmax + 1) * nn);
  for (int m = 0; m <= lmax; ++m) {
    for (int i = 0; i < nn; ++i) {
      const int n = t.geo[i].nfft;
      const int pos = m % n;
      int rel, cls;
      if (pos == 0) {
        rel = 0;
        cls = 0;
      } else if (pos < n / 2) {
        rel = 2 * pos;
        cls = 1;
      } else if (pos == n / 2) {
        rel = 1;
        cls = 0;
      } else {
        rel = 2 * (n - pos);
        cls = 2;
      }
      t.slot[(size_t)m * nn + i] = rel * 4 + cls;
      const double ang = (double)m * t.phi0[i];
      t.tw[(size_t)m * nn + i] = make_double2(std::cos(ang), std::sin(ang));
    }
  }
}

static void fill_work(Handle* h) {
  const int lmax = h->lmax;
  h->nblk.resize(lmax + 1);
  h->seedoff.resize(lmax + 2);
  h->synoff.resize(lmax + 2);
  h->goff.resize(lmax + 2);
  long long so = 0, yo = 0;
  for (int m = 0; m <= lmax; ++m) {
    const int nr = h->nhr + 1 - h->i0[m];
    h->nblk[m] = (lmax + 1 - m + kAnaLB - 1) / kAnaLB;
    h->seedoff[m] = so;
    so += (long long)h->nblk[m] * nr;
    h->synoff[m] = yo;
    yo += nr;
  }
  h->seedoff[lmax + 1] = so;
  h->synoff[lmax + 1] = yo;
}

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

// XCD-aware ordering.  The hardware deals workgroups round-robin to the 8
// XCDs (workgroup k runs on XCD k % 8), each with its own L2.  All items of
// one order m read the same operands and seeds, so they are kept on one XCD;
// the m-groups are spread over the XCDs by estimated cost (longest group to
// the least loaded XCD), each XCD's list runs longest groups first, and the
// eight lists are interleaved for the round-robin dispatch (short lists are
// padded with empty items, m = -1).
static std::vector<LegItem> xcd_order(const std::vector<LegItem>& v, const std::vector<double>& cost) {
  constexpr int kXcd = 8;
  struct Group {
    size_t begin, end;
    double cost;
  };
  std::vector<Group> groups;
  for (size_t i = 0; i < v.size();) {
    size_t j = i;
    double c = 0.0;
    while (j < v.size() && v[j].m == v[i].m) c += cost[j++];
    groups.push_back(Group{i, j, c});
    i = j;
  }
  std::stable_sort(groups.begin(), groups.end(), [](const Group& a, const Group& b) { return a.cost > b.cost; });
  std::vector<std::vector<LegItem>> lists(kXcd);
  std::vector<double> load(kXcd, 0.0);
  for (const Group& g : groups) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += g.cost;
    for (size_t i = g.begin; i < g.end; ++i) lists[x].push_back(v[i]);
  }
  size_t per = 0;
  for (const auto& l : lists) per = std::max(per, l.size());
  std::vector<LegItem> out(per * kXcd, LegItem{-1, 0, 0, 0, 0});
  for (int x = 0; x < kXcd; ++x)
    for (size_t j = 0; j < lists[x].size(); ++j) out[j * kXcd + x] = lists[x][j];
  return out;
}

hipError_t upload_items(const std::vector<LegItem>& v, const std::vector<double>& cost, LegItem** d, int* n) {
  std::vector<LegItem> o = xcd_order(v, cost);
  *n = (int)o.size();
  if (o.empty()) return hipSuccess;
  hipError_t e = hipMalloc(d, o.size() * sizeof(LegItem));
  if (e != hipSuccess) return e;
  return hipMemcpy(*d, o.data(), o.size() * sizeof(LegItem), hipMemcpyHostToDevice);
}

static hipError_t build_items(Handle* h) {
  const int lmax = h->lmax;
  std::vector<long long> blkoff, grpoff;
  std::vector<int> ana_r0, syn_l0;
  hipError_t es = compute_starts(h, blkoff, ana_r0, grpoff, syn_l0);
  if (es != hipSuccess) return es;
  for (int sp = 0; sp < 2; ++sp) {
    const int ncomp = sp ? 2 : 1;
    const int ncols = 2 * ncomp * h->nsimp;
    std::vector<LegItem> ana[2], syn;          // analysis: full column blocks, 2-tile blocks
    std::vector<double> ana_cost[2], syn_cost;  // estimated run time of each item
    for (int m = 0; m <= lmax; ++m) {
      const int nr = h->nhr + 1 - h->i0[m];
      for (int par = 0; par < 2; ++par) {
        // analysis: every ring of m against one l-block
        const int ncb = ceil_div(ncols, kAnaCols);
        for (int b = 0; b < h->nblk[m]; ++b) {
          const int nl = std::min(kAnaLB, lmax + 1 - m - b * kAnaLB);
          int r0 = ana_r0[blkoff[m] + b];
          if (r0 < nr) r0 &= ~3;  // whole MFMA K-steps of 4 rings
          const double rings = 64.0 * ceil_div(nr - r0, 64);
          for (int cb = 0; cb < ncb; ++cb) {
            const int ntl = std::min(4, ncols / 16 - 4 * cb);  // 16-column tiles in the block
            if (ntl > 2) {
              ana[0].push_back(LegItem{m, par, b, cb, r0});
              ana_cost[0].push_back(rings * (nl + 32) + 64.0 * 32);
            } else {  // blocks of 2 tiles: K split over wave pairs (colblk counts 2-tile blocks)
              for (int c2 = 0; 2 * c2 < ntl; ++c2) {
                ana[1].push_back(LegItem{m, par, b, 2 * cb + c2, r0});
                ana_cost[1].push_back(rings * (nl + 32) + 64.0 * 32);
              }
            }
          }
        }
        // synthesis: one ring block against the kept l of m (first 16-column
        // tile t0); a wave runs the 32-l chunks from its rings' first kept l
        const int nrb = ceil_div(nr, kSynRings);
        const int ntile = ncols / 16, tpi = kSynCols / 16;
        for (int rb = 0; rb < nrb; ++rb) {
          double chunks = 0.0;
          for (int w = 0; w < kSynRings / 64; ++w) {
            const int g = rb * (kSynRings / 64) + w;
            if (64 * g >= nr) break;
            const int l0 = syn_l0[grpoff[m] + g];
            chunks += 1.0;
            if (l0 <= lmax) chunks += (double)ceil_div(lmax + 1 - (m + (l0 - m) / 32 * 32), 32);
          }
          for (int t0 = 0; t0 < ntile; t0 += tpi) {
            syn.push_back(LegItem{m, par, rb, t0, 0});
            syn_cost.push_back(chunks * std::min(tpi, ntile - t0));
          }
        }
      }
    }
    hipError_t e = hipSuccess;
    for (int k2 = 0; k2 < 2 && e == hipSuccess; ++k2)
      e = upload_items(ana[k2], ana_cost[k2], &h->d_ana_items[sp][k2], &h->n_ana_items[sp][k2]);
    if (e != hipSuccess) return e;
    e = upload_items(syn, syn_cost, &h->d_syn_items[sp], &h->n_syn_items[sp]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class T>
static hipError_t upload(T** d, const std::vector<T>& v) {
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc(d, v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Size of the m-major synthesis buffer for ncomp components.
void synth_offsets(Handle* h, int ncomp) {
  long long go = 0;
  const int ncols = 2 * ncomp * h->nsimp;
  for (int m = 0; m <= h->lmax; ++m) {
    h->goff[m] = go;
    go += 2LL * (h->nhr + 1 - h->i0[m]) * ncols;
  }
  h->goff[h->lmax + 1] = go;
  h->G_size = go;
}

// Per m: the first ring (absolute index) with a kept P_lm at some l <= lmax,
// from the synthesis starts; the equator ring when none is (never the case:
// P_lm is O(1) there).
__global__ __launch_bounds__(256) void k_first_ring(int nhr, const int* __restrict__ i0,
                                                    const long long* __restrict__ synoff,
                                                    const int* __restrict__ syn_st, int* __restrict__ first) {
  __shared__ int red[256];
  const int m = blockIdx.x;
  const int nr = nhr + 1 - i0[m];
  int r0 = nr - 1;
  for (int r = threadIdx.x; r < nr; r += 256)
    if (syn_st[synoff[m] + r] != 0x3fffffff) r0 = min(r0, r);
  red[threadIdx.x] = r0;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] = min(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) first[m] = i0[m] + red[0];
}

// (Re)build the per-(m, ring) seed tables for the current h->i0.
static hipError_t make_seeds(Handle* h) {
  fill_work(h);
  (void)hipFree(h->d_i0);
  (void)hipFree(h->d_seedoff);
  (void)hipFree(h->d_seed_st);
  (void)hipFree(h->d_seed_v);
  (void)hipFree(h->d_seed_p);
  (void)hipFree(h->d_synoff);
  (void)hipFree(h->d_syn_st);
  (void)hipFree(h->d_syn_v);
  h->d_i0 = nullptr;
  h->d_seedoff = h->d_synoff = nullptr;
  h->d_seed_st = h->d_syn_st = nullptr;
  h->d_seed_v = h->d_syn_v = nullptr;
  h->d_seed_p = nullptr;
  hipError_t e;
  if ((e = upload(&h->d_i0, h->i0)) != hipSuccess) return e;
  if ((e = upload(&h->d_seedoff, h->seedoff)) != hipSuccess) return e;
  const long long nseed = h->seedoff[h->lmax + 1];
  if ((e = hipMalloc(&h->d_seed_st, std::max(1LL, nseed) * sizeof(int))) != hipSuccess) return e;
  if ((e = hipMalloc(&h->d_seed_v, std::max(1LL, nseed) * sizeof(double2))) != hipSuccess) return e;
  if ((e = hipMalloc(&h->d_seed_p, std::max(1LL, nseed) * sizeof(double))) != hipSuccess) return e;
  if ((e = upload(&h->d_synoff, h->synoff)) != hipSuccess) return e;
  const long long nsyn = h->synoff[h->lmax + 1];
  if ((e = hipMalloc(&h->d_syn_st, std::max(1LL, nsyn) * sizeof(int))) != hipSuccess) return e;
  if ((e = hipMalloc(&h->d_syn_v, std::max(1LL, nsyn) * sizeof(double2))) != hipSuccess) return e;
  return launch_seeds(h);
}

hipError_t setup_tables(Handle* h) {
  HostTables t;
  ring_geometry(h->nside, t.geo, t.phi0);
  ring_tab(h->nside, h->lmax + 1, t.geo, h->i0_ref);
  h->i0 = h->i0_ref;
  fill_ab_spin(h, t);
  fill_slots(h, t);
  hipError_t e;
  if ((e = upload(&h->d_geo, t.geo)) != hipSuccess) return e;
  if ((e = upload(&h->d_ab, t.ab)) != hipSuccess) return e;
  if ((e = upload(&h->d_spin, t.spin)) != hipSuccess) return e;
  if ((e = upload(&h->d_aboff, h->aboff)) != hipSuccess) return e;
  if ((e = upload(&h->d_tw, t.tw)) != hipSuccess) return e;
  if ((e = upload(&h->d_slot, t.slot)) != hipSuccess) return e;
  if ((e = make_seeds(h)) != hipSuccess) return e;
  // Rings past ring_tab whose P_lm is dropped at every l <= lmax carry no
  // work for m: raise the first ring to the suffix minimum over m' >= m of
  // the first ring with a kept value (non-decreasing in m, never past a ring
  // with a kept value), and rebuild the seeds for the smaller ring sets.
  {
    const int nm = h->lmax + 1;
    std::vector<int> first(nm);
    int* d_first = nullptr;
    if ((e = hipMalloc(&d_first, nm * sizeof(int))) != hipSuccess) return e;
    hipLaunchKernelGGL(k_first_ring, dim3(nm), dim3(256), 0, h->stream, h->nhr, h->d_i0, h->d_synoff, h->d_syn_st,
                       d_first);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(first.data(), d_first, nm * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d_first);
    if (e != hipSuccess) return e;
    std::vector<int> eff(nm);
    int run = h->nhr;
    for (int m = nm - 1; m >= 0; --m) {
      run = std::min(run, first[m]);
      eff[m] = std::max(run, h->i0_ref[m]);
    }
    if (eff != h->i0) {
      h->i0 = eff;
      if ((e = make_seeds(h)) != hipSuccess) return e;
    }
  }
  synth_offsets(h, 2);  // sized for the largest (spin-2) layout; spin-0 uses a prefix
  if ((e = upload(&h->d_goff, h->goff)) != hipSuccess) return e;
  if ((e = build_items(h)) != hipSuccess) return e;
  return hipSuccess;
}

}  // namespace fsht
