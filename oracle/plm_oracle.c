/*
 * ORACLE (test infrastructure only).  CPU restatement of the reference's
 * normalised associated-Legendre machinery, used by tests/ and by bench.py's
 * cpu_baseline leg as the checker.  Nothing in the product links this file.
 *
 * Follows:
 *   - recurrence constants  a(l), b(l):       src/get_plm_recursive.f90:31-37
 *                                             (c1..c4: src/sht_data_init.f90:401-405)
 *   - diagonal seed P_mm = sqrt(diag_fac(m)/4/pi) * (-sin th)^m
 *                                             src/get_plm_recursive.f90:46,73
 *     with diag_fac(m) = prod_{k<=m} (2k+1)/(2k)
 *                                             src/sht_data_init.f90:411-416
 *   - three-term recurrence
 *       P_l = a(l) cos th P_{l-1} - b(l) P_{l-2} src/get_plm_recursive.f90:50-55
 *
 * Where the reference's FP64 recurrence underflows (P_mm below 1e-308 on polar
 * rings) it either loses the values (nside <= 512, get_plm_recursive.f90:41-55)
 * or restarts from tabulated seeds (nside > 512, :68-88).  This restatement
 * instead carries a binary exponent beside the mantissa, so every returned
 * value is the true P_lm rounded to double; values below 2^-900 are returned as
 * exactly 0.  After the point where the scaled value becomes representable,
 * every step is the same fused multiply-add on the same operands as an unscaled
 * recurrence started from (P_l, P_{l+1}), so a device recurrence seeded from
 * those two values reproduces the oracle bit for bit.
 *
 * Build: oracle/Makefile  ->  oracle/_build/libplm_oracle.so (ctypes)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793238462643383279502884
#define ORC_FLUSH_LOG2 (-900) /* values below 2^-900 are returned as 0 */

/* diag_fac(m) for m = 0..mmax, same product order as the reference. */
void orc_diag_fac(int mmax, double* out) {
  out[0] = 1.0;
  for (int m = 1; m <= mmax; ++m) out[m] = out[m - 1] * (2.0 * m + 1.0) / 2.0 / (double)m;
}

/* Recurrence constants for order m, l = m..lmax:
 *   a(l) = sqrt( (2l-1)(2l+1) / (l^2 - m^2) ),
 *   b(l) = sqrt( ((l-1)^2 - m^2) / ((2l-3)(2l-1)) ) * a(l),  a(m)=0, b(m)=-1.  */
void orc_ab(int m, int lmax, double* a, double* b) {
  const double mm = (double)m * (double)m;
  a[0] = 0.0;
  b[0] = -1.0;
  for (int l = m + 1; l <= lmax; ++l) {
    const double dl = (double)l;
    const double c1 = (dl - 1.0) * (dl - 1.0);
    const double c2 = (2.0 * dl - 3.0) * (2.0 * dl - 1.0);
    const double c3 = (2.0 * dl - 1.0) * (2.0 * dl + 1.0);
    const double c4 = dl * dl;
    const double al = sqrt(c3 / (c4 - mm));
    a[l - m] = al;
    b[l - m] = sqrt((c1 - mm) / c2) * al;
  }
}

/* x^n as (mantissa, exponent) with mantissa in [0.5, 1), by binary powering. */
static void pow_scaled(double x, int n, double* mant, long* ex) {
  int ex0;
  double base = frexp(x, &ex0);
  long be = ex0;
  double r = 1.0;
  long re = 0;
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

/* P_mm as (mantissa in [0.5,1) with sign, binary exponent). */
void orc_pmm_scaled(int m, double sinth, const double* diag_fac, double* mant, long* ex) {
  double ms;
  long es;
  pow_scaled(sinth, m, &ms, &es);
  double v = sqrt(diag_fac[m] / 4.0 / ORC_PI) * ms;
  if (m & 1) v = -v;
  int e;
  v = frexp(v, &e);
  *mant = v;
  *ex = es + e;
}

/* Scaled three-term recurrence for one ring.  Calls out[l-m] = P_lm for
 * l = m..lmax (0 below the flush threshold).  Returns the first l whose value
 * is non-zero (lmax+1 if none).  If maxabs != NULL it receives max_l |P_lm|
 * evaluated in log2 space over l = m..lmax (can be far below the flush). */
int orc_plm_ring(int m, int lmax, double costh, double sinth, const double* diag_fac,
                 const double* a, const double* b, double* out, double* maxlog2) {
  double p1;
  long e;
  orc_pmm_scaled(m, sinth, diag_fac, &p1, &e);
  double p2 = 0.0;
  int first = lmax + 1;
  double best = -1e300;
  for (int l = m; l <= lmax; ++l) {
    double p;
    if (l == m) {
      p = p1;
    } else {
      const double ax = a[l - m] * costh;
      p = fma(ax, p1, -(b[l - m] * p2));
      p2 = p1;
      p1 = p;
    }
    /* keep mantissa bounded while the value is tiny */
    if (e < 0) {
      const double ap = fabs(p1);
      if (ap > 0x1p100) {
        long sh = (-e < 100) ? -e : 100;
        p1 = ldexp(p1, (int)-sh);
        p2 = ldexp(p2, (int)-sh);
        e += sh;
        p = p1;
      }
    }
    double lg = (p != 0.0) ? log2(fabs(p)) + (double)e : -1e300;
    if (lg > best) best = lg;
    /* below the first representable value everything is flushed; from there
     * on the true value is returned unflushed */
    if (first > lmax && p != 0.0) {
      int fp;
      frexp(p, &fp);
      /* log2|p| + e >= -900  <=>  frexp-exponent(p) + e >= -899 */
      if ((long)fp + e >= (long)ORC_FLUSH_LOG2 + 1) first = l;
    }
    double val = (first <= l) ? ldexp(p, (int)e) : 0.0;
    if (out) out[l - m] = val;
  }
  if (maxlog2) *maxlog2 = best;
  return first;
}

/* P_lm for a set of rings: out[r*(lmax+1-m) + (l-m)], rings given by cos/sin.
 * first[r] receives the first non-zero l (lmax+1 if none). */
void orc_plm_rings(int m, int lmax, int nr, const double* costh, const double* sinth,
                   double* out, int* first) {
  double* dfac = (double*)malloc(sizeof(double) * (m + 1));
  double* a = (double*)malloc(sizeof(double) * (lmax + 1 - m));
  double* b = (double*)malloc(sizeof(double) * (lmax + 1 - m));
  orc_diag_fac(m, dfac);
  orc_ab(m, lmax, a, b);
  const int n = lmax + 1 - m;
  for (int r = 0; r < nr; ++r) {
    int f = orc_plm_ring(m, lmax, costh[r], sinth[r], dfac, a, b, out + (size_t)r * n, NULL);
    if (first) first[r] = f;
  }
  free(dfac);
  free(a);
  free(b);
}

/* max_{m<=l<=lmax} log10 |P_lm(theta_r)| for each ring r. */
void orc_maxlog10_rings(int m, int lmax, int nr, const double* costh, const double* sinth,
                        double* out) {
  double* dfac = (double*)malloc(sizeof(double) * (m + 1));
  double* a = (double*)malloc(sizeof(double) * (lmax + 1 - m));
  double* b = (double*)malloc(sizeof(double) * (lmax + 1 - m));
  orc_diag_fac(m, dfac);
  orc_ab(m, lmax, a, b);
  for (int r = 0; r < nr; ++r) {
    double lg;
    orc_plm_ring(m, lmax, costh[r], sinth[r], dfac, a, b, NULL, &lg);
    out[r] = lg * 0.30102999566398119521;
  }
  free(dfac);
  free(a);
  free(b);
}

/* Ring-skip table: for each m in [0, nm) the first north ring r (0-based, over
 * the nr rings given, ordered pole -> equator) whose max_{m<=l<=lcap}|P_lm|
 * reaches 10^thr_log10.  Uses bisection over rings (the maximum grows toward
 * the equator); `exhaustive` scans every ring instead, to check that
 * assumption.  nr if no ring qualifies. */
void orc_ring_tab(int nm, int lcap, int nr, const double* costh, const double* sinth,
                  double thr_log10, int exhaustive, int* out) {
  int mmax = nm - 1;
  double* dfac = (double*)malloc(sizeof(double) * (mmax + 1));
  orc_diag_fac(mmax, dfac);
  double* a = (double*)malloc(sizeof(double) * (lcap + 1));
  double* b = (double*)malloc(sizeof(double) * (lcap + 1));
  const double thr2 = thr_log10 / 0.30102999566398119521;
  for (int m = 0; m < nm; ++m) {
    int lc = lcap < m ? m : lcap;
    orc_ab(m, lc, a, b);
    if (exhaustive) {
      int res = nr;
      for (int r = 0; r < nr; ++r) {
        double lg;
        orc_plm_ring(m, lc, costh[r], sinth[r], dfac, a, b, NULL, &lg);
        if (lg >= thr2) {
          res = r;
          break;
        }
      }
      out[m] = res;
    } else {
      int lo = 0, hi = nr; /* answer in [lo, hi] */
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        double lg;
        orc_plm_ring(m, lc, costh[mid], sinth[mid], dfac, a, b, NULL, &lg);
        if (lg >= thr2)
          hi = mid;
        else
          lo = mid + 1;
      }
      out[m] = lo;
    }
  }
  free(dfac);
  free(a);
  free(b);
}
