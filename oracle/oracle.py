"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the reference's batched HEALPix spherical harmonic
transform (liuhao-cn/fastSHT).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker.
The product (fastsht_amd/) never imports it.

Layouts are the reference's (Fortran order):
  maps  (npix, nsim)                     -- each map contiguous
  FFT buffers (nsim, npix), PERM packing -- sim fastest
  alms  (nsim, lmax+1, lmax+1)           -- Re a_lm at [:, l, m], Im a_lm at
                                            [:, l-m, lmax+1-m] (m >= 1)

Each routine cites the reference file:line it restates.  Where the reference's
FP64 Legendre recurrence underflows (polar rings, large m) this restatement
returns the true values instead (see plm_oracle.c); below 2^-900 they are 0.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    """Load (building if needed) the C half of the oracle."""
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "_build", "libplm_oracle.so")
        src = os.path.join(_HERE, "plm_oracle.c")
        if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(so)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.orc_plm_rings.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, ip]
        lib.orc_maxlog10_rings.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        lib.orc_ring_tab.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp,
                                     ctypes.c_double, ctypes.c_int, ip]
        lib.orc_diag_fac.argtypes = [ctypes.c_int, dp]
        lib.orc_ab.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        _LIB = lib
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# --------------------------------------------------------------------------
# Ring geometry
# --------------------------------------------------------------------------
def ring_info(nside):
    """i1 (first pixel), nfft (ring length) per ring.

    src/sht_data_init.f90:655-675 (get_ring_info)."""
    nring = 4 * nside - 1
    i = np.arange(nring, dtype=np.int64)
    i1 = np.empty(nring, np.int64)
    nfft = np.empty(nring, np.int64)
    n = i < nside
    i1[n] = 2 * i[n] * (i[n] + 1)
    nfft[n] = 4 * (i[n] + 1)
    e = (i >= nside) & (i < 3 * nside - 1)
    i1[e] = 2 * nside + 4 * nside * i[e] - 2 * nside ** 2
    nfft[e] = 4 * nside
    s = i >= 3 * nside - 1
    i1[s] = 12 * nside ** 2 - 2 * (4 * nside - i[s] - 1) * (4 * nside - i[s])
    nfft[s] = 4 * (4 * nside - 1 - i[s])
    return i1, nfft


def ring_z_s(nside):
    """cos(theta), sin(theta) of every ring from the HEALPix closed form.

    The reference takes theta of each ring's first pixel from healpy
    (SHT.py:47-50) and uses cos/sin of it (sht_data_init.f90:392-393).  Polar
    ring r (1-based): z = 1 - r^2/(3 nside^2); equatorial: z = (4 nside - 2r) /
    (3 nside); south rings mirror the north (Gorski et al. 2005)."""
    nring = 4 * nside - 1
    z = np.empty(nring)
    s = np.empty(nring)
    for i in range(2 * nside):
        r = i + 1
        if r < nside:
            t = (r / nside) ** 2 / 3.0
            zz = 1.0 - t
            ss = np.sqrt(t * (2.0 - t))
        else:
            zz = (4.0 * nside - 2.0 * r) / (3.0 * nside)
            ss = np.sqrt((1.0 - zz) * (1.0 + zz))
        z[i], s[i] = zz, ss
        j = nring - 1 - i
        z[j], s[j] = -zz, ss
    return z, s


def ring_phi0(nside):
    """Azimuth of the first pixel of every ring (HEALPix RING scheme):
    polar ring r: pi/(4r); equatorial ring r: pi/(4 nside) if (r - nside) even
    else 0; south rings mirror the north.  Reference: SHT.py:47-50 (healpy)."""
    nring = 4 * nside - 1
    phi0 = np.empty(nring)
    for i in range(2 * nside):
        r = i + 1
        if r < nside:
            p = np.pi / (4.0 * r)
        else:
            p = np.pi / (4.0 * nside) if (r - nside) % 2 == 0 else 0.0
        phi0[i] = p
        phi0[nring - 1 - i] = p
    return phi0


def ring_tab(nside, nm=None, lcap=None, thr_log10=-20.0, exhaustive=False):
    """Ring-skip table: first north ring kept for each m.

    Rule (pinned against SHT_ring_skip_table/*.bin): the first north ring whose
    max_{m<=l<=lcap} |P_lm| reaches 1e-20, lcap = 4 nside by default.
    Used as i0 = ring_tab(m) in src/sht_core.f90:189."""
    if nm is None:
        nm = 4 * nside + 1
    if lcap is None:
        lcap = 4 * nside
    z, s = ring_z_s(nside)
    zn = np.ascontiguousarray(z[: 2 * nside])
    sn = np.ascontiguousarray(s[: 2 * nside])
    out = np.zeros(nm, np.int32)
    _lib().orc_ring_tab(nm, lcap, 2 * nside, _dp(zn), _dp(sn), thr_log10, int(exhaustive), _ip(out))
    return out


def plm_north(nside, lmax, m, rings=None):
    """P_lm(theta_i) for north rings (index array) and l = m..lmax.

    Returns (P[len(rings), lmax+1-m], first_nonzero_l[len(rings)]).
    src/get_plm_recursive.f90:30-90."""
    z, s = ring_z_s(nside)
    if rings is None:
        rings = np.arange(2 * nside)
    zz = np.ascontiguousarray(z[rings])
    ss = np.ascontiguousarray(s[rings])
    nr = len(rings)
    n = lmax + 1 - m
    out = np.zeros((nr, n))
    first = np.zeros(nr, np.int32)
    _lib().orc_plm_rings(m, lmax, nr, _dp(zz), _dp(ss), _dp(out), _ip(first))
    return out, first


def plm_generic(m, lmax, costh, sinth):
    """P_lm at arbitrary (cos, sin) pairs: (len, lmax+1-m)."""
    zz = np.ascontiguousarray(np.asarray(costh, dtype=np.float64))
    ss = np.ascontiguousarray(np.asarray(sinth, dtype=np.float64))
    nr = len(zz)
    out = np.zeros((nr, lmax + 1 - m))
    first = np.zeros(nr, np.int32)
    _lib().orc_plm_rings(m, lmax, nr, _dp(zz), _dp(ss), _dp(out), _ip(first))
    return out


def plm_maxlog10(nside, m, lmax, rings):
    z, s = ring_z_s(nside)
    zz = np.ascontiguousarray(z[rings])
    ss = np.ascontiguousarray(s[rings])
    out = np.zeros(len(rings))
    _lib().orc_maxlog10_rings(m, lmax, len(rings), _dp(zz), _dp(ss), _dp(out))
    return out


# --------------------------------------------------------------------------
# The transform object
# --------------------------------------------------------------------------
class OracleSHT:
    """CPU restatement of the reference SHT object (SHT.py:16-227 over
    src/sht_main.f90).  Inputs/outputs use the reference layouts."""

    def __init__(self, nside, lmax, nsim, ring_tab_values=None):
        self.nside, self.lmax, self.nsim = nside, lmax, nsim
        self.nring = 4 * nside - 1
        self.npix = 12 * nside * nside
        self.nhr = 2 * nside - 1
        self.pa = 4.0 * np.pi / self.npix  # sht_data_init.f90:18
        self.i1, self.nfft = ring_info(nside)
        self.z, self.s = ring_z_s(nside)
        self.phi0 = ring_phi0(nside)
        if ring_tab_values is None:
            ring_tab_values = ring_tab(nside, nm=lmax + 1)
        self.rtab = np.asarray(ring_tab_values[: lmax + 1], dtype=np.int64)
        self._tables()
        self._pcache = {}

    # ---- tables (sht_data_init.f90:438-575) ------------------------------
    def _tables(self):
        nring, lmax, nside = self.nring, self.lmax, self.nside
        m = np.arange(lmax + 1)
        nf = self.nfft[:, None]
        pos = m[None, :] % nf
        i1 = self.i1[:, None]
        tab = np.where(pos == 0, i1,
              np.where(pos < nf // 2, i1 + 2 * pos,
              np.where(pos == nf // 2, i1 + 1, i1 + 2 * (nf - pos))))
        self.tab = tab  # forward and backward tables coincide (:453-462, :526-538)
        # forward factors, north rings + equator (:467-490)
        nn = 2 * nside
        ph = self.phi0[:nn, None]
        posn = pos[:nn]
        nfn = nf[:nn]
        cc = np.cos(-m[None, :] * ph)
        ss = np.sin(-m[None, :] * ph)
        dc = posn == 0
        pf = (posn > 0) & (posn < nfn // 2)
        ny = posn == nfn // 2
        ng = posn > nfn // 2
        z = np.zeros_like(cc)
        self.fac1 = cc.copy()
        self.fac2 = np.where(pf, -ss, np.where(ng, ss, z))
        self.fac3 = ss.copy()
        self.fac4 = np.where(pf, cc, np.where(ng, -cc, z))
        # backward factors (:542-573)
        cI = np.cos(m[None, :] * ph)
        sI = np.sin(m[None, :] * ph)
        fm0 = np.where(m[None, :] == 0, 1.0, 2.0)
        dcny = dc | ny
        self.fac1I = np.where(dcny, cI * fm0, cI)
        self.fac2I = np.where(dcny, -sI * fm0, -sI)
        self.fac3I = np.where(dcny, z, np.where(pf, sI, -sI))
        self.fac4I = np.where(dcny, z, np.where(pf, cI, -cI))

    # ---- Legendre functions (get_plm_recursive.f90:17-136) ---------------
    def plm(self, m, pol=False):
        """plm0 (and plm1, plm2 if pol) for rings i0..nring-1-i0, l=m..lmax.
        Returned arrays are (nring, lmax+1-m); rows outside the kept range 0."""
        key = (m, pol)
        if key in self._pcache:
            return self._pcache[key]
        nside, lmax, nring = self.nside, self.lmax, self.nring
        i0 = int(self.rtab[m])
        n = lmax + 1 - m
        north = np.arange(i0, 2 * nside)
        p0n, _ = plm_north(nside, lmax, m, north)
        l = np.arange(m, lmax + 1)
        sgn = np.where((l + m) % 2 == 0, 1.0, -1.0)
        P0 = np.zeros((nring, n))
        P0[i0:2 * nside] = p0n
        south = np.arange(2 * nside, nring - i0)
        P0[south] = P0[nring - 1 - south] * sgn[None, :]  # :93-98
        if not pol:
            res = (P0,)
        else:
            cc = self.z[i0:2 * nside][:, None]
            ss = self.s[i0:2 * nside][:, None]
            b_w = cc / ss ** 2
            two_on_s2 = 2.0 / ss ** 2
            c_w = m / ss ** 2
            lf = l.astype(np.float64)
            normal_l = np.zeros(n)
            k = l >= 2
            normal_l[k] = 1.0 / np.sqrt((lf[k] + 2) * (lf[k] + 1) * lf[k] * (lf[k] - 1))
            m1 = max(1, m)
            lam = np.zeros(n)
            kk = l > m1
            lam[kk] = 2.0 * np.sqrt((2.0 * l[kk] + 1) / (2.0 * l[kk] - 1) * (l[kk] ** 2 - m ** 2))
            normal_m = 2.0 * m * (1 - m)
            P1n = np.zeros_like(p0n)
            P2n = np.zeros_like(p0n)
            P1n[:, 0] = normal_m * normal_l[0] * p0n[:, 0] * (0.5 - 1.0 / ss[:, 0] ** 2)  # :111
            P2n[:, 0] = normal_m * normal_l[0] * p0n[:, 0] * cc[:, 0] / ss[:, 0] ** 2  # :112
            if n > 1:
                pm1 = p0n[:, :-1]
                pl = p0n[:, 1:]
                lv = lf[1:][None, :]
                P1n[:, 1:] = (pm1 * lam[1:][None, :] * b_w
                              - pl * (two_on_s2 * (l[1:][None, :] - m ** 2) + lv * (lv - 1.0))) * normal_l[1:][None, :]
                P2n[:, 1:] = (pm1 * lam[1:][None, :]
                              - pl * (2.0 * cc * (lv - 1.0))) * c_w * normal_l[1:][None, :]
            P1 = np.zeros((nring, n))
            P2 = np.zeros((nring, n))
            P1[i0:2 * nside] = P1n
            P2[i0:2 * nside] = P2n
            P1[south] = P1[nring - 1 - south] * sgn[None, :]  # :127-133
            P2[south] = -P2[nring - 1 - south] * sgn[None, :]
            res = (P0, P1, P2)
        if len(self._pcache) > 4:
            self._pcache.clear()
        self._pcache[key] = res
        return res

    # ---- ring FFT (compute_fft.f90) ----------------------------------------
    def map2fft(self, maps):
        """Real FFT of every ring into PERM slots, transposed to (nsim, npix).
        compute_fft.f90:71-110 (MKL DFTI forward, PERM format, unnormalised)."""
        maps = np.asarray(maps)
        nsim = maps.shape[1]
        out = np.zeros((nsim, self.npix))
        for i in range(self.nring):
            a, n = int(self.i1[i]), int(self.nfft[i])
            X = np.fft.rfft(maps[a:a + n, :], axis=0)  # (n/2+1, nsim)
            blk = np.empty((n, nsim))
            blk[0] = X[0].real
            blk[1] = X[n // 2].real
            blk[2::2] = X[1:n // 2].real
            blk[3::2] = X[1:n // 2].imag
            out[:, a:a + n] = blk.T
        return out

    def fft2map(self, buf):
        """Inverse of map2fft, unnormalised backward (x_j = sum_k X_k e^{+2 pi i jk/n}).
        compute_fft.f90:226-280."""
        nsim = buf.shape[0]
        maps = np.zeros((self.npix, nsim), order="F")
        for i in range(self.nring):
            a, n = int(self.i1[i]), int(self.nfft[i])
            blk = buf[:, a:a + n].T
            X = np.zeros((n // 2 + 1, nsim), complex)
            X[0] = blk[0]
            X[n // 2] = blk[1]
            X[1:n // 2] = blk[2::2] + 1j * blk[3::2]
            maps[a:a + n, :] = np.fft.irfft(X, n=n, axis=0) * n
        return maps

    # ---- mapping (fft_mapping.f90) --------------------------------------
    def _loc(self, m):
        i0 = int(self.rtab[m])
        rlen = self.nring - 2 * i0
        return i0, i0 + rlen - 1

    def fft2mat(self, buf, m):
        """fft_mapping.f90:19-68: gather PERM slots of frequency m, rotate by
        e^{-i m phi0}; south rings use the north factors."""
        i0, i1 = self._loc(m)
        nsim = buf.shape[0]
        c1 = np.zeros((nsim, self.nring))
        c2 = np.zeros((nsim, self.nring))
        rings = np.arange(i0, i1 + 1)
        jf = np.minimum(rings, self.nring - 1 - rings)
        j = self.tab[rings, m]
        x0 = buf[:, j]
        x1 = buf[:, j + 1]
        f1, f2 = self.fac1[jf, m], self.fac2[jf, m]
        f3, f4 = self.fac3[jf, m], self.fac4[jf, m]
        c1[:, rings] = x0 * f1 + x1 * f2
        c2[:, rings] = x0 * f3 + x1 * f4
        return c1, c2

    def mat2fft(self, buf, m, c1, c2):
        """fft_mapping.f90:84-128: scatter with set (first touch) / accumulate.
        Within one m every ring owns a distinct slot pair, so the ring loop is
        vectorised."""
        i0, i1 = self._loc(m)
        rings = np.arange(i0, i1 + 1)
        j = self.tab[rings, m]
        jf = np.minimum(rings, self.nring - 1 - rings)
        fac = np.where(2 * (m + 1) > self.nfft[rings], 1.0, 0.0)[None, :]
        a, b = c1[:, rings], c2[:, rings]
        v0 = a * self.fac1I[jf, m] + b * self.fac2I[jf, m]
        v1 = a * self.fac3I[jf, m] + b * self.fac4I[jf, m]
        buf[:, j] = buf[:, j] * fac + v0
        buf[:, j + 1] = buf[:, j + 1] * fac + v1

    def mat2fft_iter(self, buf1, buf2, m, c1, c2):
        """fft_mapping.f90:142-190: residual X2 = X1 - nfft*synth (first touch)
        or X2 -= nfft*synth (aliased); vectorised over rings as above."""
        i0, i1 = self._loc(m)
        rings = np.arange(i0, i1 + 1)
        j = self.tab[rings, m]
        jf = np.minimum(rings, self.nring - 1 - rings)
        nf = self.nfft[rings].astype(np.float64)
        alias = (2 * (m + 1) > self.nfft[rings])[None, :]
        a, b = c1[:, rings], c2[:, rings]
        s0 = np.where(alias, buf2[:, j], buf1[:, j])
        s1 = np.where(alias, buf2[:, j + 1], buf1[:, j + 1])
        buf2[:, j] = s0 - a * self.fac1I[jf, m] * nf - b * self.fac2I[jf, m] * nf
        buf2[:, j + 1] = s1 - a * self.fac3I[jf, m] * nf - b * self.fac4I[jf, m] * nf

    def zero_unused(self, buf):
        """sht_main.f90:137-146: zero PERM slots >= 2(lmax+1)."""
        for i in range(self.nring):
            n = int(self.nfft[i])
            if 2 * (self.lmax + 1) < n:
                a = int(self.i1[i])
                buf[:, a + 2 * (self.lmax + 1):a + n] = 0.0

    # ---- GEMM core (sht_core.f90) -----------------------------------------
    # alpha/swap table: sht_data_init.f90:606-632
    def _ab(self, inv, typ):
        if inv == 0:
            pa = self.pa
            return {0: (pa, pa, 0), 1: (-pa, -pa, 0), 2: (pa, -pa, 1), 3: (-pa, pa, 1)}[typ]
        return {0: (1.0, 1.0, 0), 1: (-1.0, -1.0, 0), 2: (1.0, -1.0, 1), 3: (-1.0, 1.0, 1)}[typ]

    def _p(self, m):
        lmax = self.lmax
        n = lmax + 1 - m
        return (lmax + 2) * m, (lmax + 1) * n, n  # sht_core.f90:192-193

    def mat2alm(self, almf, P, m, typ, beta, c1, c2):
        """sht_core.f90:62-89.  almf is the (nsim, (lmax+1)^2) column view."""
        i0, i1 = self._loc(m)
        p1, p2, n = self._p(m)
        a1, a2, sw = self._ab(0, typ)
        Pk = P[i0:i1 + 1]
        A = c1 if sw == 0 else c2
        B = c2 if sw == 0 else c1
        almf[:, p1:p1 + n] = a1 * (A[:, i0:i1 + 1] @ Pk) + beta * almf[:, p1:p1 + n]
        if m > 0:
            almf[:, p2:p2 + n] = a2 * (B[:, i0:i1 + 1] @ Pk) + beta * almf[:, p2:p2 + n]

    def alm2mat(self, almf, P, m, typ, beta, c1, c2):
        """sht_core.f90:141-168."""
        i0, i1 = self._loc(m)
        p1, p2, n = self._p(m)
        a1, a2, sw = self._ab(1, typ)
        Pk = P[i0:i1 + 1]
        if sw == 0:
            c1[:, i0:i1 + 1] = a1 * (almf[:, p1:p1 + n] @ Pk.T) + beta * c1[:, i0:i1 + 1]
            if m > 0:
                c2[:, i0:i1 + 1] = a2 * (almf[:, p2:p2 + n] @ Pk.T) + beta * c2[:, i0:i1 + 1]
        else:
            if m == 0:
                c1[:, i0:i1 + 1] = a1 * (almf[:, p1:p1 + n] @ Pk.T) + beta * c1[:, i0:i1 + 1]
            else:
                c2[:, i0:i1 + 1] = a1 * (almf[:, p1:p1 + n] @ Pk.T) + beta * c2[:, i0:i1 + 1]
                c1[:, i0:i1 + 1] = a2 * (almf[:, p2:p2 + n] @ Pk.T) + beta * c1[:, i0:i1 + 1]

    # ---- helpers ----------------------------------------------------------
    def _new_alm(self, nsim):
        a = np.zeros((nsim, self.lmax + 1, self.lmax + 1), order="F")
        return a, a.reshape(nsim, -1, order="F")

    def _caches(self, nsim):
        return np.zeros((nsim, self.nring)), np.zeros((nsim, self.nring))

    # ---- drivers (sht_main.f90) --------------------------------------------
    def t2alm(self, maps, niter=0):
        """niter=0: sht_main.f90:38-89; niter>0: t2alm_iter_old :276-367."""
        nsim = maps.shape[1]
        b1 = self.map2fft(maps)
        alm, af = self._new_alm(nsim)
        c1, c2 = self._caches(nsim)
        for m in range(self.lmax + 1):
            (P,) = self.plm(m)
            c1, c2 = self.fft2mat(b1, m)
            self.mat2alm(af, P, m, 0, 0.0, c1, c2)
        for _ in range(niter):
            b2 = np.zeros_like(b1)
            c1, c2 = self._caches(nsim)
            for m in range(self.lmax + 1):
                (P,) = self.plm(m)
                self.alm2mat(af, P, m, 0, 0.0, c1, c2)
                self.mat2fft_iter(b1, b2, m, c1, c2)
            for m in range(self.lmax + 1):
                (P,) = self.plm(m)
                c1, c2 = self.fft2mat(b2, m)
                self.mat2alm(af, P, m, 0, 1.0, c1, c2)
        return alm

    def t2alm_iter(self, maps, niter=0):
        """Immediate-update refinement, sht_main.f90:161-264: within each sweep
        every order m is synthesised, scattered into the residual buffer
        (mat2fft_iter: set on first touch, accumulate on aliases) and
        re-analysed at once, m ascending."""
        nsim = maps.shape[1]
        b1 = self.map2fft(maps)
        alm, af = self._new_alm(nsim)
        for m in range(self.lmax + 1):
            (P,) = self.plm(m)
            c1, c2 = self.fft2mat(b1, m)
            self.mat2alm(af, P, m, 0, 0.0, c1, c2)
        for _ in range(niter):
            b2 = np.zeros_like(b1)
            c1, c2 = self._caches(nsim)
            for m in range(self.lmax + 1):
                (P,) = self.plm(m)
                self.alm2mat(af, P, m, 0, 0.0, c1, c2)
                self.mat2fft_iter(b1, b2, m, c1, c2)
                d1, d2 = self.fft2mat(b2, m)
                self.mat2alm(af, P, m, 0, 1.0, d1, d2)
        return alm

    def alm2t(self, alm):
        """sht_main.f90:96-155."""
        alm = np.asarray(alm)
        nsim = alm.shape[0]
        af = np.asfortranarray(alm).reshape(nsim, -1, order="F")
        b1 = np.zeros((nsim, self.npix))
        c1, c2 = self._caches(nsim)
        for m in range(self.lmax + 1):
            (P,) = self.plm(m)
            self.alm2mat(af, P, m, 0, 0.0, c1, c2)
            self.mat2fft(b1, m, c1, c2)
        self.zero_unused(b1)
        return self.fft2map(b1)

    def timed_sample(self, maps, ms):
        """Wall-clock parts of one temperature round trip on a sample of orders
        (CPU-baseline helper): full forward ring FFT, per-m analysis
        (get_plm + fft2mat + mat2alm) over `ms`, per-m synthesis (get_plm +
        alm2mat + mat2fft) over `ms`, full inverse ring FFT.  Returns seconds."""
        import time
        nsim = maps.shape[1]
        t0 = time.perf_counter()
        b1 = self.map2fft(maps)
        t1 = time.perf_counter()
        alm, af = self._new_alm(nsim)
        for m in ms:
            (P,) = self.plm(m)
            c1, c2 = self.fft2mat(b1, m)
            self.mat2alm(af, P, m, 0, 0.0, c1, c2)
        t2 = time.perf_counter()
        b2 = np.zeros_like(b1)
        c1, c2 = self._caches(nsim)
        for m in ms:
            (P,) = self.plm(m)
            self.alm2mat(af, P, m, 0, 0.0, c1, c2)
            self.mat2fft(b2, m, c1, c2)
        t3 = time.perf_counter()
        self.fft2map(b2)
        t4 = time.perf_counter()
        return {"fft_fwd": t1 - t0, "analysis": t2 - t1, "synthesis": t3 - t2, "fft_inv": t4 - t3}

    def _qu_forward(self, bq, bu, afE, afB, beta, want_e=True, want_b=True, ms=None):
        """The per-m body shared by QU2EB (:409-430) and qu2eb_iter_old (:596-617, :645-666);
        `ms` restricts it to some orders."""
        for m in (range(self.lmax + 1) if ms is None else ms):
            _, P1, P2 = self.plm(m, pol=True)
            c1, c2 = self.fft2mat(bq, m)
            if want_e:
                self.mat2alm(afE, P1, m, 1, beta, c1, c2)
            if want_b:
                self.mat2alm(afB, P2, m, 3, beta, c1, c2)
            c1, c2 = self.fft2mat(bu, m)
            if want_e:
                self.mat2alm(afE, P2, m, 2, 1.0, c1, c2)
            if want_b:
                self.mat2alm(afB, P1, m, 1, 1.0, c1, c2)

    def qu2eb(self, Q, U, niter=0):
        """niter=0: QU2EB sht_main.f90:383-436; niter>0: qu2eb_iter_old :563-673."""
        nsim = Q.shape[1]
        bq = self.map2fft(Q)
        bu = self.map2fft(U)
        E, afE = self._new_alm(nsim)
        B, afB = self._new_alm(nsim)
        self._qu_forward(bq, bu, afE, afB, 0.0)
        for _ in range(niter):
            b3 = np.zeros_like(bq)
            b4 = np.zeros_like(bu)
            c1, c2 = self._caches(nsim)
            for m in range(self.lmax + 1):  # :621-643
                _, P1, P2 = self.plm(m, pol=True)
                self.alm2mat(afE, P1, m, 1, 0.0, c1, c2)
                self.alm2mat(afB, P2, m, 3, 1.0, c1, c2)
                self.mat2fft_iter(bq, b3, m, c1, c2)
                self.alm2mat(afE, P2, m, 2, 0.0, c1, c2)
                self.alm2mat(afB, P1, m, 1, 1.0, c1, c2)
                self.mat2fft_iter(bu, b4, m, c1, c2)
            self._qu_forward(b3, b4, afE, afB, 1.0)
        return E, B

    def qu2eb_orders(self, Q, U, ms):
        """QU2EB (sht_main.f90:383-436) for the orders `ms` only (each order of
        an unrefined transform is independent); the other orders stay 0."""
        nsim = Q.shape[1]
        bq = self.map2fft(Q)
        bu = self.map2fft(U)
        E, afE = self._new_alm(nsim)
        B, afB = self._new_alm(nsim)
        self._qu_forward(bq, bu, afE, afB, 0.0, ms=ms)
        return E, B

    def qu2e(self, Q, U):
        """QU2E sht_main.f90:955-1005."""
        nsim = Q.shape[1]
        bq = self.map2fft(Q)
        bu = self.map2fft(U)
        E, afE = self._new_alm(nsim)
        for m in range(self.lmax + 1):
            _, P1, P2 = self.plm(m, pol=True)
            c1, c2 = self.fft2mat(bq, m)
            self.mat2alm(afE, P1, m, 1, 0.0, c1, c2)
            c1, c2 = self.fft2mat(bu, m)
            self.mat2alm(afE, P2, m, 2, 1.0, c1, c2)
        return E

    def qu2b(self, Q, U):
        """QU2B sht_main.f90:1016-1064."""
        nsim = Q.shape[1]
        bq = self.map2fft(Q)
        bu = self.map2fft(U)
        B, afB = self._new_alm(nsim)
        for m in range(self.lmax + 1):
            _, P1, P2 = self.plm(m, pol=True)
            c1, c2 = self.fft2mat(bu, m)
            self.mat2alm(afB, P1, m, 1, 0.0, c1, c2)
            c1, c2 = self.fft2mat(bq, m)
            self.mat2alm(afB, P2, m, 3, 1.0, c1, c2)
        return B

    def eb2qu(self, E, B):
        """EB2QU sht_main.f90:1079-1147."""
        nsim = E.shape[0]
        afE = np.asfortranarray(E).reshape(nsim, -1, order="F")
        afB = np.asfortranarray(B).reshape(nsim, -1, order="F")
        b1 = np.zeros((nsim, self.npix))
        b2 = np.zeros((nsim, self.npix))
        c1, c2 = self._caches(nsim)
        for m in range(self.lmax + 1):
            _, P1, P2 = self.plm(m, pol=True)
            self.alm2mat(afE, P1, m, 1, 0.0, c1, c2)
            self.alm2mat(afB, P2, m, 3, 1.0, c1, c2)
            self.mat2fft(b1, m, c1, c2)
            self.alm2mat(afE, P2, m, 2, 0.0, c1, c2)
            self.alm2mat(afB, P1, m, 1, 1.0, c1, c2)
            self.mat2fft(b2, m, c1, c2)
        self.zero_unused(b1)
        self.zero_unused(b2)
        return self.fft2map(b1), self.fft2map(b2)


# --------------------------------------------------------------------------
# alm utilities
# --------------------------------------------------------------------------
def fix_eb(sht, Q, U, mask, niter=0, remask=True):
    """EB-leakage cleaning on a masked sky, src/sht_main.f90:801-941 (with
    copy_with_mask_mini / apply_mask_mini / lin_resi_mini,
    src/sht_core.f90:209-258), on an OracleSHT `sht`.

    Returns dict(Bmap, Qm, Um, almB, Btpl, coef): the cleaned B map (0 outside
    mask == 1), the masked inputs, the B alm of the masked maps, the template
    map and the per-map fit (a0, a1).  remask=False drops the second masking
    (only for the planted-mistake test)."""
    mask = np.asarray(mask, dtype=np.float64)
    Qm = np.asfortranarray(Q * mask[:, None])                 # :846-849
    Um = np.asfortranarray(U * mask[:, None])
    E, B = sht.qu2eb(Qm, Um, niter=niter)                    # :853-860
    Q2, U2 = sht.eb2qu(E, np.zeros_like(E))                   # :866 (E2QU = the E part of EB2QU)
    if remask:
        Q2 = np.asfortranarray(Q2 * mask[:, None])            # :870-871
        U2 = np.asfortranarray(U2 * mask[:, None])
    _, B2 = sht.qu2eb(Q2, U2, niter=niter)                    # :876-881
    y = sht.alm2t(B)                                          # :884-919 (scalar synthesis of B)
    x = sht.alm2t(B2)
    vid = np.nonzero(mask == 1.0)[0]                          # SHT.py:193
    Bmap = np.zeros_like(y)
    coef = np.zeros((2, y.shape[1]))
    for k in range(y.shape[1]):                               # lin_resi_mini, sht_core.f90:239-258
        xs, ys = x[vid, k], y[vid, k]
        mx, my = xs.sum() / len(vid), ys.sum() / len(vid)
        cxx = ((xs - mx) ** 2).sum()
        cxy = ((ys - my) * (xs - mx)).sum()
        a1 = cxy / cxx
        a0 = my - mx * a1
        Bmap[vid, k] = ys - a0 - a1 * xs
        coef[:, k] = (a0, a1)
    return dict(Bmap=Bmap, Qm=Qm, Um=Um, almB=B, Btpl=x, coef=coef)


def galactic_cut_mask(nside, zcut=0.2, holes=0, seed=0):
    """0/1 mask: 0 for |cos theta| < zcut and, optionally, `holes` random
    discs of 0; 1 elsewhere."""
    th, ph = pixel_angles(nside)
    mask = (np.abs(np.cos(th)) >= zcut).astype(np.float64)
    rng = np.random.default_rng(seed)
    for _ in range(holes):
        t0, p0 = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)
        cosd = np.cos(th) * np.cos(t0) + np.sin(th) * np.sin(t0) * np.cos(ph - p0)
        mask[cosd > np.cos(0.3)] = 0.0
    return mask


# --------------------------------------------------------------------------
# Spectra and seeded alm (src/alm2cl_TEB.f90:22-110, src/cl2alm_TEB.f90:25-170)
# --------------------------------------------------------------------------
def alm2cl(almT, almE=None, almB=None, lmin=0, cl=None):
    """cl(k, l, c) += sum_m fac_m (x_lm y_lm)/(2l+1) over Re and Im parts for
    l >= lmin, in the reference's order (m ascending, Re term then Im term).
    T alone: (nsim, lmax+1); with E and B: (nsim, lmax+1, 6) = TT EE BB TE TB EB."""
    nsim, L, _ = almT.shape
    pol = almE is not None
    if cl is None:
        cl = np.zeros((nsim, L, 6) if pol else (nsim, L), order="F")
    comps = [almT, almE, almB] if pol else [almT]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)] if pol else [(0, 0)]
    for m in range(L):
        fac = 1.0 if m == 0 else 2.0
        l1 = max(m, lmin)
        if l1 > L - 1:
            continue
        ls = np.arange(l1, L)
        d = 2.0 * ls + 1.0
        for part in range(2 if m > 0 else 1):
            rows = ls - m if part else ls
            col = L - m if part else m
            vals = [c[:, rows, col] for c in comps]
            for ci, (a, b) in enumerate(pairs):
                term = vals[a] * vals[b] * fac / d
                if pol:
                    cl[:, l1:, ci] = cl[:, l1:, ci] + term
                else:
                    cl[:, l1:] = cl[:, l1:] + term
    return cl


_PHILOX_M = (0xD2511F53, 0xCD9E8D57)
_PHILOX_W = (0x9E3779B9, 0xBB67AE85)


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011) on uint32 arrays:
    ctr (4, n), key (2,) -> (4, n)."""
    c = [np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF for x in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_PHILOX_M[0]) * c[0]
        p1 = np.uint64(_PHILOX_M[1]) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0)) & mask, p1 & mask,
             ((p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1)) & mask, p0 & mask]
        k0 = (k0 + _PHILOX_W[0]) & 0xFFFFFFFF
        k1 = (k1 + _PHILOX_W[1]) & 0xFFFFFFFF
    return c


def std_normal(seed, k, r, col, comp):
    """Standard normal of storage element (sim k, row r, column col) of
    component comp: Philox4x32-10 on the counter (k, r, col, comp) with key
    (seed lo, seed hi), Box-Muller on two 53-bit uniforms (u1 in (0, 1])."""
    k, r, col = np.broadcast_arrays(np.asarray(k), np.asarray(r), np.asarray(col))
    comp = np.full(k.shape, comp)
    c = philox4x32_10([k.ravel(), r.ravel(), col.ravel(), comp.ravel()],
                      (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    u1 = ((c[0] >> np.uint64(5)).astype(np.float64) * 67108864.0 + (c[1] >> np.uint64(6)).astype(np.float64)
          + 1.0) / 9007199254740992.0
    u2 = ((c[2] >> np.uint64(5)).astype(np.float64) * 67108864.0
          + (c[3] >> np.uint64(6)).astype(np.float64)) / 9007199254740992.0
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)).reshape(k.shape)


def teb_factor(cl_l):
    """Lower-triangular factor rows (a; b, u; c, v, w) of the TEB covariance at
    one l, src/cl2alm_TEB.f90:49-56.  cl_l = (TT, EE, BB, TE, TB, EB)."""
    TT, EE, BB, TE, TB, EB = cl_l
    a = np.sqrt(TT)
    b = TE / a
    c = TB / a
    u = np.sqrt(EE - b ** 2)
    v = (TT * EB - TE * TB) / (u * TT)
    w = np.sqrt(BB - TB / TT / u ** 2 * (EE * TB - TE * EB) - v * EB / u)
    return a, b, c, u, v, w


def cl2alm(cl, lmax, nsim, seed=0, lmin=0):
    """Seeded Gaussian alm: cl (lmax+1,) -> almT; cl (lmax+1, 6) -> (T, E, B).
    Variance 1 at m = 0 and 1/2 per part at m > 0, combined per l by
    teb_factor (src/cl2alm_TEB.f90:58-100); l < lmin zeroed (:103-116)."""
    L = lmax + 1
    cl = np.asarray(cl, dtype=np.float64)
    pol = cl.ndim == 2
    kk, rr, cc = np.meshgrid(np.arange(nsim), np.arange(L), np.arange(L), indexing="ij")
    re = rr >= cc
    m = np.where(re, cc, L - cc)
    ll = np.where(re, rr, rr + m)
    fac = np.where(cc == 0, 1.0, 0.7071067811865476)
    out = []
    g = [std_normal(seed, kk, rr, cc, comp) * fac for comp in range(3 if pol else 1)]
    if not pol:
        aT = g[0] * np.sqrt(cl[ll])
        aT[ll < lmin] = 0.0
        return np.asfortranarray(aT)
    with np.errstate(divide="ignore", invalid="ignore"):
        a, b, c, u, v, w = teb_factor([cl[ll, j] for j in range(6)])
    aB = g[0] * c + g[1] * v + g[2] * w
    aE = g[0] * b + g[1] * u
    aT = g[0] * a
    for x in (aT, aE, aB):
        x[ll < lmin] = 0.0
        out.append(np.asfortranarray(x))
    return tuple(out)


def convert_alm_healpy(alm):
    """src/convert_alm.f90:34-56 (convert_alm_healpy1) + SHT.py:204-210:
    square-packed real alm -> healpy m-major complex, shape (nalm, nsim)."""
    alm = np.asarray(alm)
    nsim, L1, _ = alm.shape
    lmax = L1 - 1
    nalm = (lmax + 1) * (lmax + 2) // 2
    out = np.zeros((nalm, nsim), complex)
    pos = 0
    for m in range(lmax + 1):
        n = lmax + 1 - m
        re = alm[:, m:, m].T
        im = alm[:, 0:n, lmax + 1 - m].T if m > 0 else 0.0
        out[pos:pos + n] = re + 1j * im
        pos += n
    return out


def convert_alm_healpy_inv(alm_hp, lmax):
    """src/convert_alm.f90:87-107."""
    nsim = alm_hp.shape[1]
    alm = np.zeros((nsim, lmax + 1, lmax + 1), order="F")
    pos = 0
    for m in range(lmax + 1):
        n = lmax + 1 - m
        alm[:, m:, m] = alm_hp[pos:pos + n].real.T
        if m > 0:
            alm[:, 0:n, lmax + 1 - m] = alm_hp[pos:pos + n].imag.T
        pos += n
    return alm


def norm_alm(alm_in, fac, shift, m0):
    """src/norm_alm.f90:14-33."""
    alm_in = np.asarray(alm_in)
    lmax = alm_in.shape[1] - 1
    out = np.array(alm_in, order="F", copy=True)
    out[:, :, 0] = alm_in[:, :, 0] * fac[None, :] * m0 + shift[None, :]
    for m in range(1, lmax + 1):
        l = np.arange(m, lmax + 1)
        out[:, l, m] = alm_in[:, l, m] * fac[l][None, :] + shift[l][None, :]
        out[:, l - m, lmax + 1 - m] = alm_in[:, l - m, lmax + 1 - m] * fac[l][None, :] + shift[l][None, :]
    return out


def filt_alm(alm, win):
    """src/filt_alm.f90:16-36 (in place on a copy)."""
    alm = np.array(alm, order="F", copy=True)
    lmax = alm.shape[1] - 1
    alm[:, :, 0] *= win[None, :]
    for m in range(1, lmax + 1):
        l = np.arange(m, lmax + 1)
        alm[:, l, m] *= win[l][None, :]
        alm[:, l - m, lmax + 1 - m] *= win[l][None, :]
    return alm


# --------------------------------------------------------------------------
# Independent brute force (for pinning the oracle at tiny sizes)
# --------------------------------------------------------------------------
def pixel_angles(nside):
    """theta, phi of every RING pixel from the closed form."""
    i1, nfft = ring_info(nside)
    z, s = ring_z_s(nside)
    phi0 = ring_phi0(nside)
    npix = 12 * nside * nside
    th = np.empty(npix)
    ph = np.empty(npix)
    for i in range(4 * nside - 1):
        a, n = int(i1[i]), int(nfft[i])
        th[a:a + n] = np.arctan2(s[i], z[i])
        ph[a:a + n] = phi0[i] + 2.0 * np.pi * np.arange(n) / n
    return th, ph


def brute_t2alm(maps, nside, lmax):
    """a_lm = 4pi/npix sum_p t_p conj(Y_lm(p)) with scipy's Y_lm; returns the
    reference square packing.  Pins t2alm(niter=0) at tiny nside."""
    from scipy.special import sph_harm_y
    th, ph = pixel_angles(nside)
    npix = 12 * nside * nside
    nsim = maps.shape[1]
    alm = np.zeros((nsim, lmax + 1, lmax + 1), order="F")
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            Y = sph_harm_y(l, m, th, ph)
            a = (4 * np.pi / npix) * (np.conj(Y)[None, :] @ maps)[0]
            alm[:, l, m] = a.real
            if m > 0:
                alm[:, l - m, lmax + 1 - m] = a.imag
    return alm


def brute_alm2t(alm, nside):
    """t_p = sum_lm a_lm Y_lm(p) (real field: m>0 terms doubled)."""
    from scipy.special import sph_harm_y
    th, ph = pixel_angles(nside)
    nsim, L1, _ = alm.shape
    lmax = L1 - 1
    maps = np.zeros((12 * nside * nside, nsim))
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            Y = sph_harm_y(l, m, th, ph)
            re = alm[:, l, m]
            if m == 0:
                maps += np.outer(Y.real, re)
            else:
                im = alm[:, l - m, lmax + 1 - m]
                maps += 2.0 * (np.outer(Y.real, re) - np.outer(Y.imag, im))
    return maps


# --------------------------------------------------------------------------
# Synthetic inputs (seeded)
# --------------------------------------------------------------------------
def random_alm(lmax, nsim, seed=0, bb_frac=None):
    """Gaussian alm with C_l = 1/(l(l+1)) (l>=2), 0 below; m=0 real.
    Draw order: m-major, then l, Re then Im, all sims at once (PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    alm = np.zeros((nsim, lmax + 1, lmax + 1), order="F")
    l = np.arange(lmax + 1, dtype=np.float64)
    cl = np.zeros(lmax + 1)
    cl[2:] = 1.0 / (l[2:] * (l[2:] + 1))
    if bb_frac is not None:
        cl = cl * bb_frac
    for m in range(lmax + 1):
        n = lmax + 1 - m
        sd = np.sqrt(cl[m:]) if m == 0 else np.sqrt(cl[m:] / 2.0)
        re = rng.standard_normal((n, nsim)) * sd[:, None]
        alm[:, m:, m] = re.T
        if m > 0:
            im = rng.standard_normal((n, nsim)) * sd[:, None]
            alm[:, 0:n, lmax + 1 - m] = im.T
    return alm
