"""GPU parity: the HIP path (through the C ABI / drop-in SHT object) against
the oracle on the same seeded inputs.  Gate: max relative error (max |diff| /
max |oracle|) <= 1e-12 for every FP64 transform."""
import numpy as np
import pytest

from conftest import maxrel
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12

# (128, 40) and (64, 30): lmax far below 2*nside, so most polar rings carry
# no kept P_lm for the larger orders and drop out of the per-order ring sets.
# (128, 257) and (128, 300): more than one 256-l analysis block per order, so
# the analysis restarts its recurrence from the block seeds (the spin-2
# functions at the restart also need P_{l-1})
CASES = [(1, 2, 3), (2, 5, 2), (4, 8, 3), (4, 16, 5), (8, 20, 17), (8, 31, 4), (16, 47, 33),
         (16, 64, 3), (32, 95, 20), (32, 40, 16), (64, 30, 2), (128, 40, 3), (128, 257, 2),
         (128, 300, 2)]


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _maps(rng, ns, nsim):
    return np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim)))


@pytest.mark.parametrize("ns,lmax,nsim", CASES)
def test_spin0_vs_oracle(torch_cuda, ns, lmax, nsim):
    from fastsht_amd import SHT
    print(f"case spin0 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    rng = np.random.default_rng(ns * 1000 + lmax)
    maps = _maps(rng, ns, nsim)
    R = O.OracleSHT(ns, lmax, nsim)
    S0 = SHT(ns, lmax, nsim, niter=0)
    assert np.array_equal(S0.ring_tab(), R.rtab.astype(np.int32))
    assert maxrel(S0.t2alm(maps), R.t2alm(maps)) < TOL
    alm = O.random_alm(lmax, nsim, seed=3)
    assert maxrel(S0.alm2t(alm), R.alm2t(alm)) < TOL
    S3 = SHT(ns, lmax, nsim, niter=3)
    assert maxrel(S3.t2alm(maps), R.t2alm(maps, niter=3)) < TOL


@pytest.mark.parametrize("ns,lmax,nsim", CASES)
def test_spin2_vs_oracle(torch_cuda, ns, lmax, nsim):
    from fastsht_amd import SHT
    print(f"case spin2 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    rng = np.random.default_rng(ns * 7 + lmax)
    Q, U = _maps(rng, ns, nsim), _maps(rng, ns, nsim)
    R = O.OracleSHT(ns, lmax, nsim)
    S0 = SHT(ns, lmax, nsim, niter=0, pol=True)
    E, B = S0.qu2eb(Q, U)
    Eo, Bo = R.qu2eb(Q, U)
    assert maxrel(E, Eo) < TOL and maxrel(B, Bo) < TOL
    assert maxrel(S0.qu2e(Q, U), Eo) < TOL
    assert maxrel(S0.qu2b(Q, U), Bo) < TOL
    Ea, Ba = O.random_alm(lmax, nsim, seed=5), O.random_alm(lmax, nsim, seed=6)
    Qg, Ug = S0.eb2qu(Ea, Ba)
    Qo, Uo = R.eb2qu(Ea, Ba)
    assert maxrel(Qg, Qo) < TOL and maxrel(Ug, Uo) < TOL
    S2 = SHT(ns, lmax, nsim, niter=2, pol=True)
    E2, B2 = S2.qu2eb(Q, U)
    Eo2, Bo2 = R.qu2eb(Q, U, niter=2)
    assert maxrel(E2, Eo2) < TOL and maxrel(B2, Bo2) < TOL


def test_spin0_lmax0_and_nside256(torch_cuda):
    """lmax = 0 (monopole only) and nside 256 / lmax 511, niter 0 and 3."""
    from fastsht_amd import SHT
    for ns, lmax, nsim in [(4, 0, 2), (256, 511, 2)]:
        print(f"case spin0 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
        rng = np.random.default_rng(ns + lmax)
        maps = _maps(rng, ns, nsim)
        R = O.OracleSHT(ns, lmax, nsim)
        S = SHT(ns, lmax, nsim, niter=3)
        assert np.array_equal(S.ring_tab(), R.rtab.astype(np.int32))
        assert maxrel(S.t2alm(maps), R.t2alm(maps, niter=3)) < TOL
        alm = O.random_alm(lmax, nsim, seed=4)
        assert maxrel(S.alm2t(alm), R.alm2t(alm)) < TOL


def test_c0_config_vs_oracle(torch_cuda):
    """nside=64, lmax=191, nsim=100: t2alm(niter=0), t2alm(niter=3), alm2t
    on seeded band-limited + white-noise maps (the reference's CPU config)."""
    from fastsht_amd import SHT
    ns, lmax, nsim = 64, 191, 100
    print("case C0 nside=64 lmax=191 nsim=100", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    alm = O.random_alm(lmax, nsim, seed=0)
    rng = np.random.Generator(np.random.PCG64(0))
    maps = np.asfortranarray(R.alm2t(alm) + rng.standard_normal((12 * ns * ns, nsim)) * np.sqrt(1e-3))
    S = SHT(ns, lmax, nsim, niter=3)
    assert maxrel(S.t2alm(maps), R.t2alm(maps, niter=3)) < TOL
    S0 = SHT(ns, lmax, nsim, niter=0)
    a0 = S0.t2alm(maps)
    assert maxrel(a0, R.t2alm(maps, niter=0)) < TOL
    assert maxrel(S0.alm2t(a0), R.alm2t(a0)) < TOL


def test_device_tensors_match_host(torch_cuda):
    torch = torch_cuda
    from fastsht_amd import SHT, empty_alms, empty_maps
    ns, lmax, nsim = 16, 40, 7
    rng = np.random.default_rng(11)
    maps = _maps(rng, ns, nsim)
    S = SHT(ns, lmax, nsim, niter=2)
    host = S.t2alm(maps)
    dmaps = empty_maps(S.npix, nsim, device=True)
    dmaps.copy_(torch.from_numpy(maps))
    dev = S.t2alm(dmaps)
    torch.cuda.synchronize()
    assert dev.is_cuda
    assert maxrel(dev.cpu().numpy(), host) < 1e-15
    back_h = S.alm2t(host)
    back_d = S.alm2t(dev)
    torch.cuda.synchronize()
    assert maxrel(back_d.cpu().numpy(), back_h) < 1e-15
    # caller-provided outputs return 0 and are filled in place (SHT.py:104-107)
    out = empty_alms(lmax, nsim)
    assert S.t2alm(maps, out) == 0
    assert maxrel(out, host) == 0.0


def test_get_alm_path(torch_cuda):
    from fastsht_amd import SHT
    ns, lmax, nsim = 8, 20, 3
    maps = _maps(np.random.default_rng(1), ns, nsim)
    S = SHT(ns, lmax, nsim, niter=1)
    ref = S.t2alm(maps)
    assert S.t2alm(maps, get_alm=False) is None
    assert maxrel(S.get_alms(), ref) == 0.0


def test_errors(torch_cuda):
    from fastsht_amd import SHT, FastSHTError
    S = SHT(4, 8, 2)
    with pytest.raises(ValueError):
        S.t2alm()  # no maps (SHT.py:77-78)
    with pytest.raises(ValueError):
        S.t2alm(np.zeros((10, 2)))
    with pytest.raises(ValueError):
        S.qu2eb(np.zeros((192, 2)), np.zeros((192, 2)))  # pol=False handle
    with pytest.raises(FastSHTError):
        S.get_alms()  # nothing kept on the device yet
    with pytest.raises(FastSHTError):
        SHT(3, 4, 1)


def test_alm_utilities_vs_oracle(torch_cuda):
    import ctypes

    torch = torch_cuda
    from fastsht_amd import SHT, _lib
    lmax, nsim = 30, 5
    alm = O.random_alm(lmax, nsim, seed=8)
    S = SHT(8, lmax, nsim)
    hp = S.convert_alm_healpy(alm)
    np.testing.assert_array_equal(hp, O.convert_alm_healpy(alm))
    lib = _lib.load()
    L = lmax + 1
    d = torch.from_numpy(np.ascontiguousarray(alm.transpose(2, 1, 0))).cuda()  # storage = Fortran order
    out = torch.empty_like(d)
    fac = torch.linspace(1, 2, L, dtype=torch.float64, device="cuda")
    shift = torch.linspace(0, 0.1, L, dtype=torch.float64, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(lib.fsht_norm_alm(p(d), p(out), p(fac), p(shift), 0.5, lmax, nsim, None))
    torch.cuda.synchronize()
    ref = O.norm_alm(alm, fac.cpu().numpy(), shift.cpu().numpy(), 0.5)
    # in*fac + shift is one fused multiply-add on the GPU: compare at FP64 rounding level
    np.testing.assert_allclose(out.cpu().numpy().transpose(2, 1, 0), ref, rtol=1e-15, atol=1e-16)
    win = torch.linspace(1, 0, L, dtype=torch.float64, device="cuda")
    _lib.check(lib.fsht_filt_alm(p(d), p(win), lmax, nsim, None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(d.cpu().numpy().transpose(2, 1, 0), O.filt_alm(alm, win.cpu().numpy()), rtol=1e-15)
    hpd = torch.from_numpy(np.stack([hp.real.T, hp.imag.T], axis=-1).copy()).cuda()  # (nsim, nalm, 2)
    back = torch.empty_like(d)
    _lib.check(lib.fsht_convert_alm_healpy_inv(p(hpd), p(back), lmax, nsim, _lib.FSHT_DEVICE, None))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(back.cpu().numpy().transpose(2, 1, 0), alm)
    # host arrays through the C ABI (the reference's convert_alm_healpy1 call)
    out_h = np.empty((2, L * (L + 1) // 2, nsim), order="F")
    _lib.check(lib.fsht_convert_alm_healpy(alm.ctypes.data, out_h.ctypes.data, lmax, nsim, _lib.FSHT_HOST, None))
    np.testing.assert_array_equal(out_h[0] + 1j * out_h[1], O.convert_alm_healpy(alm))
    back_h = np.empty_like(alm)
    _lib.check(lib.fsht_convert_alm_healpy_inv(out_h.ctypes.data, back_h.ctypes.data, lmax, nsim, _lib.FSHT_HOST,
                                               None))
    np.testing.assert_array_equal(back_h, alm)


def _single_order_closed_form(nside, lmax, m, a):
    """Map of the field whose only alm are a_lm (l = m..lmax, complex a) for
    one order m: T(theta_r, phi) = g * Re(sum_l P_lm(theta_r) a_l e^{i m phi}),
    g = 1 at m = 0 (a real) else 2; south rings carry (-1)^(l+m).  Also returns
    the per-ring direct sums S_r = sum_{p in ring r} T_p e^{-i m phi_p} used to
    check the analysis."""
    nr = 4 * nside - 1
    P, _ = O.plm_north(nside, lmax, m)              # (2*nside rings, lmax+1-m)
    sgn = (-1.0) ** np.arange(lmax + 1 - m)         # (-1)^(l+m) relative to l = m
    gN = P @ a
    gS = P @ (sgn * a)
    g_ring = np.concatenate([gN, gS[:2 * nside - 1][::-1]])
    # exact phases: phi_p = pi*(2j + h)/n on a ring of n pixels (h = 1 when
    # the ring starts half a pixel off phi = 0), so m*phi_p = pi*k/n with the
    # integer k = (2mj + mh) mod 2n; float64 m*phi would be off by ~m*1e-16
    i1, nfft = O.ring_info(nside)
    nfft = np.asarray(nfft, dtype=np.int64)
    h = np.rint(O.ring_phi0(nside) * nfft / np.pi).astype(np.int64)
    ring_of_pix = np.repeat(np.arange(nr), nfft)
    j = np.arange(12 * nside * nside, dtype=np.int64) - np.repeat(np.asarray(i1, dtype=np.int64), nfft)
    n_p = nfft[ring_of_pix]
    k = (2 * m * j + m * h[ring_of_pix]) % (2 * n_p)
    eph = np.exp(1j * np.pi * k / n_p)
    fac = 1.0 if m == 0 else 2.0
    T = fac * np.real(g_ring[ring_of_pix] * eph)
    S = np.add.reduceat(T * np.conj(eph), np.asarray(i1, dtype=np.int64))
    # north-folded analysis operands: P^T (S_north +/- S_south)
    return T, S, P, sgn


@pytest.mark.parametrize("m", [0, 393, 1200, 3000])
def test_single_order_lmax_4nside(torch_cuda, m):
    """nside=1024, lmax=4*nside: for these orders some kept polar rings only
    rise above the 2^-900 flush level hundreds of l-steps after m (more than
    one 256-l block), and reach O(1) before lmax.  Synthesis must start each
    ring's recurrence at its first kept l; checked against the closed form,
    and the analysis against the direct per-ring sums."""
    from fastsht_amd import SHT
    nside, lmax, nsim = 1024, 4096, 1
    L = lmax + 1
    print(f"case single order nside={nside} lmax={lmax} m={m}", flush=True)
    rng = np.random.default_rng(100 + m)
    n = L - m
    a = rng.standard_normal(n) + (1j * rng.standard_normal(n) if m > 0 else 0.0)
    alm = np.zeros((nsim, L, L), order="F")
    alm[0, m:, m] = a.real
    if m > 0:
        alm[0, :n, L - m] = a.imag
    T, S, P, sgn = _single_order_closed_form(nside, lmax, m, a)
    sht = SHT(nside, lmax, nsim)
    maps = sht.alm2t(alm)
    assert maxrel(maps[:, 0], T) < TOL
    # analysis of the closed-form map, order m only
    out = sht.t2alm(np.asfortranarray(T[:, None]))
    npix = 12 * nside * nside
    SN = S[:2 * nside]
    SS = np.concatenate([S[2 * nside:][::-1], [0.0]])  # mirror of north ring i (equator has none)
    ad = (4 * np.pi / npix) * (P.T @ SN + sgn * (P.T @ SS))
    got = out[0, m:, m] + (1j * out[0, :n, L - m] if m > 0 else 0.0)
    assert maxrel(got, ad) < TOL


def test_nside2048_orders_closed_form(torch_cuda):
    """nside=2048, lmax=2*nside (the largest resolution the reference tabulates).
    Its polar rings longer than 4096 pixels are transformed one real ring at a
    time through a half-length Bluestein transform.  A field made of five
    orders (incl. m = lmax, aliased on every polar ring) is synthesised and
    compared pixel by pixel with the closed form, which checks every ring's
    inverse FFT; the analysis of the closed-form map is compared, order by
    order, with the direct per-ring sums (forward FFT bins m and aliases)."""
    from fastsht_amd import SHT
    nside, lmax, nsim = 2048, 4096, 1
    L = lmax + 1
    ms = [0, 1, 2047, 3001, 4096]
    print(f"case nside={nside} lmax={lmax} orders={ms}", flush=True)
    rng = np.random.default_rng(2048)
    npix = 12 * nside * nside
    nr = 4 * nside - 1
    i1, nfft = O.ring_info(nside)
    i1 = np.asarray(i1, dtype=np.int64)
    nfft = np.asarray(nfft, dtype=np.int64)
    h = np.rint(O.ring_phi0(nside) * nfft / np.pi).astype(np.int64)
    ring_of_pix = np.repeat(np.arange(nr), nfft)
    j = np.arange(npix, dtype=np.int64) - np.repeat(i1, nfft)
    n_p = nfft[ring_of_pix]
    h_p = h[ring_of_pix]
    del ring_of_pix

    def eph(m):  # exact phases e^{i m phi_p}, as in _single_order_closed_form
        return np.exp(1j * np.pi * ((2 * m * j + m * h_p) % (2 * n_p)) / n_p)

    alm = np.zeros((nsim, L, L), order="F")
    T = np.zeros(npix)
    per_m = {}
    for m in ms:
        n = L - m
        a = rng.standard_normal(n) + (1j * rng.standard_normal(n) if m > 0 else 0.0)
        alm[0, m:, m] = a.real
        if m > 0:
            alm[0, :n, L - m] = a.imag
        P, _ = O.plm_north(nside, lmax, m)
        sgn = (-1.0) ** np.arange(n)
        g_ring = np.concatenate([P @ a, (P @ (sgn * a))[:2 * nside - 1][::-1]])
        T += (1.0 if m == 0 else 2.0) * np.real(np.repeat(g_ring, nfft) * eph(m))
        per_m[m] = (a, P, sgn)
    sht = SHT(nside, lmax, nsim)
    assert np.array_equal(sht.ring_tab(), O.ring_tab(nside, nm=L, lcap=6143))
    maps = sht.alm2t(alm)
    assert maxrel(maps[:, 0], T) < TOL
    out = sht.t2alm(np.asfortranarray(T[:, None]))
    for m in ms:
        a, P, sgn = per_m[m]
        n = L - m
        S = np.add.reduceat(T * np.conj(eph(m)), i1)
        SN = S[:2 * nside]
        SS = np.concatenate([S[2 * nside:][::-1], [0.0]])
        ad = (4 * np.pi / npix) * (P.T @ SN + sgn * (P.T @ SS))
        got = out[0, m:, m] + (1j * out[0, :n, L - m] if m > 0 else 0.0)
        assert maxrel(got, ad) < TOL, m


@pytest.mark.parametrize("ns,lmax", [(1024, 64), (2048, 96)])
def test_high_nside_spin2_vs_oracle(torch_cuda, ns, lmax):
    """Spin 2 at nside 1024 and 2048 with a small lmax: per-order ring sets far
    shorter than ring_tab's, and (2048) the long polar rings' half-length
    transforms on Q/U.

    The spin-2 functions F1/F2 (src/get_plm_recursive.f90:115-124) subtract
    terms of size ~ l/sin^2(theta) * P to leave ~ l^2 * P, so near the poles
    they carry a relative rounding error of ~ 2/(l sin^2 theta) * 1e-16 in any
    FP64 evaluation: ~1e-11 on the first rings at nside 2048.  Two correct
    implementations differ by that much there, so maps are held to 1e-12
    away from the 64 most polar rings of each cap, and to 5e-11 overall."""
    from fastsht_amd import SHT
    nsim = 1
    print(f"case spin2 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    Ea, Ba = O.random_alm(lmax, nsim, seed=5), O.random_alm(lmax, nsim, seed=6)
    S = SHT(ns, lmax, nsim, pol=True)
    Qg, Ug = S.eb2qu(Ea, Ba)
    Qo, Uo = R.eb2qu(Ea, Ba)
    K = 64
    cap = 2 * K * (K + 1)  # pixels of the first K rings
    npix = 12 * ns * ns
    body = slice(cap, npix - cap)
    for g, o in ((Qg, Qo), (Ug, Uo)):
        err = np.abs(g - o)
        scale = np.abs(o).max()
        assert err[body].max() / scale < TOL, err[body].max() / scale
        assert err.max() / scale < 5e-11, err.max() / scale
    E, B = S.qu2eb(Qo, Uo)
    Eo, Bo = R.qu2eb(Qo, Uo)
    assert maxrel(E, Eo) < TOL and maxrel(B, Bo) < TOL


@pytest.mark.parametrize("ns,lmax,nsim,niter", [(8, 16, 3, 0), (16, 32, 2, 2), (32, 80, 3, 1)])
def test_fix_eb_vs_oracle(torch_cuda, ns, lmax, nsim, niter):
    """EB-leakage cleaning through the C ABI (host arrays) against the oracle:
    cleaned B map, template, B alm, and the in-place masking of Q and U."""
    from fastsht_amd import SHT, empty_alms, empty_maps
    print(f"case fix_eb nside={ns} lmax={lmax} nsim={nsim} niter={niter}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    rng = np.random.default_rng(ns + niter)
    E, B = O.random_alm(lmax, nsim, seed=7), 0.05 * O.random_alm(lmax, nsim, seed=8)
    Q, U = R.eb2qu(E, B)
    Q = np.asfortranarray(Q + 1e-3 * rng.standard_normal(Q.shape))
    U = np.asfortranarray(U + 1e-3 * rng.standard_normal(U.shape))
    mask = O.galactic_cut_mask(ns, 0.2, holes=2, seed=ns)
    ref = O.fix_eb(R, Q, U, mask, niter=niter)
    sht = SHT(ns, lmax, nsim, niter=niter, pol=True)
    tpl, ab = empty_maps(sht.npix, nsim), empty_alms(lmax, nsim)
    Bmap = sht.fix_eb(Q, U, mask, Btpl_out=tpl, almB_out=ab)
    assert np.array_equal(Q, ref["Qm"]) and np.array_equal(U, ref["Um"])
    assert maxrel(ab, ref["almB"]) < TOL
    assert maxrel(tpl, ref["Btpl"]) < TOL
    assert maxrel(Bmap, ref["Bmap"]) < TOL
    assert np.all(Bmap[mask != 1.0] == 0.0)


def test_fix_eb_device_tensors(torch_cuda):
    """Same cleaning on GPU tensors: Q and U masked in place on the device."""
    torch = torch_cuda
    from fastsht_amd import SHT, empty_maps
    ns, lmax, nsim = 16, 40, 2
    print(f"case fix_eb device nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    Q, U = R.eb2qu(O.random_alm(lmax, nsim, seed=9), O.random_alm(lmax, nsim, seed=10))
    mask = O.galactic_cut_mask(ns, 0.25, holes=1, seed=2)
    ref = O.fix_eb(R, Q, U, mask, niter=1)
    sht = SHT(ns, lmax, nsim, niter=1, pol=True)
    Qd, Ud = empty_maps(sht.npix, nsim, device=True), empty_maps(sht.npix, nsim, device=True)
    Qd.copy_(torch.from_numpy(np.asarray(Q)))
    Ud.copy_(torch.from_numpy(np.asarray(U)))
    Bd = sht.fix_eb(Qd, Ud, torch.from_numpy(mask).cuda())
    torch.cuda.synchronize()
    assert maxrel(Bd.cpu().numpy(), ref["Bmap"]) < TOL
    assert maxrel(Qd.cpu().numpy(), ref["Qm"]) < 1e-15
    # a host mask (NumPy or CPU tensor) with GPU maps is copied to the device
    for host_mask in (mask, torch.from_numpy(mask)):
        Qd.copy_(torch.from_numpy(np.asarray(Q)))
        Ud.copy_(torch.from_numpy(np.asarray(U)))
        Bh = sht.fix_eb(Qd, Ud, host_mask)
        torch.cuda.synchronize()
        assert torch.equal(Bh, Bd)


def test_alm2cl_vs_oracle(torch_cuda):
    """Spectra on the GPU against the oracle (same summation order)."""
    from fastsht_amd import alm2cl
    lmax, nsim = 40, 5
    print(f"case alm2cl lmax={lmax} nsim={nsim}", flush=True)
    T, E, B = (O.random_alm(lmax, nsim, seed=s) for s in (21, 22, 23))
    for lmin in (0, 3):
        got = alm2cl(T, E, B, lmin=lmin)
        ref = O.alm2cl(T, E, B, lmin=lmin)
        assert maxrel(got, ref) < 1e-14
        assert maxrel(alm2cl(T, lmin=lmin), O.alm2cl(T, lmin=lmin)) < 1e-14


def test_cl2alm_vs_oracle(torch_cuda):
    """Seeded alm on the GPU against the oracle: the Philox streams agree bit
    for bit; log/cos may differ in the last ulp."""
    from fastsht_amd import cl2alm
    lmax, nsim = 30, 4
    print(f"case cl2alm lmax={lmax} nsim={nsim}", flush=True)
    L = lmax + 1
    l = np.arange(L)
    cl = np.zeros((L, 6))
    cl[2:, 0] = 1.0 / (l[2:] * (l[2:] + 1.0))
    cl[2:, 1], cl[2:, 2], cl[2:, 3] = 0.5 * cl[2:, 0], 0.05 * cl[2:, 0], 0.3 * cl[2:, 0]
    cl[2:, 4], cl[2:, 5] = 0.01 * cl[2:, 0], 0.02 * cl[2:, 0]
    got = cl2alm(cl, lmax, nsim, seed=99, lmin=2)
    ref = O.cl2alm(cl, lmax, nsim, seed=99, lmin=2)
    for g, r in zip(got, ref):
        assert maxrel(g, r) < 1e-13
    gT = cl2alm(cl[:, 0], lmax, nsim, seed=(1 << 40) + 7)
    with np.errstate(invalid="ignore"):
        rT = O.cl2alm(cl[:, 0], lmax, nsim, seed=(1 << 40) + 7)
    assert maxrel(gT, rT) < 1e-13
    gd = cl2alm(cl[:, 0], lmax, nsim, seed=5, device=True)
    assert gd.is_cuda and maxrel(gd.cpu().numpy(), O.cl2alm(cl[:, 0], lmax, nsim, seed=5)) < 1e-13


@pytest.mark.parametrize("ns,lmax,nsim,niter", [(4, 10, 3, 1), (8, 20, 2, 3), (16, 64, 5, 2), (32, 95, 4, 2)])
def test_t2alm_iter_vs_oracle(torch_cuda, ns, lmax, nsim, niter):
    """method != 'old' (immediate update, t2alm_iter) through the C ABI."""
    from fastsht_amd import SHT
    print(f"case t2alm_iter nside={ns} lmax={lmax} nsim={nsim} niter={niter}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    maps = _maps(np.random.default_rng(ns * 31 + lmax), ns, nsim)
    sht = SHT(ns, lmax, nsim, niter=niter)
    got = sht.t2alm(maps, method="new")
    assert maxrel(got, R.t2alm_iter(maps, niter=niter)) < TOL
    # method='old' on the same handle still runs the Jacobi refinement
    assert maxrel(sht.t2alm(maps), R.t2alm(maps, niter=niter)) < TOL


def test_sht_objects_run_on_torch_current_stream(torch_cuda):
    """Device calls are enqueued on torch's current stream, the default
    (null) stream included, so one SHT object's output is complete before
    another object's call reads it: a long t2alm on one object, read at once
    by a second object, gives what the second object gives after a full
    synchronisation."""
    torch = torch_cuda
    import contextlib
    import ctypes

    from fastsht_amd import SHT, _lib
    lib = _lib.load()
    ns, lmax, n = 256, 511, 40
    print(f"case stream ordering nside={ns} lmax={lmax} nsim={n}", flush=True)
    rng = np.random.default_rng(29)
    maps = torch.from_numpy(rng.standard_normal((n, 12 * ns * ns))).cuda().t()
    A, B = SHT(ns, lmax, n, niter=3), SHT(ns, lmax, n, niter=0)
    for side in (None, torch.cuda.Stream()):
        with (torch.cuda.stream(side) if side is not None else contextlib.nullcontext()):
            a = A.t2alm(maps)
            m = B.alm2t(a)  # enqueued right behind A's t2alm
            for S in (A, B):
                st = ctypes.c_void_p()
                _lib.check(lib.fsht_get_stream(S._h, ctypes.byref(st)), S._h)
                assert (st.value or 0) == torch.cuda.current_stream().cuda_stream
        torch.cuda.synchronize()
        assert torch.equal(m, B.alm2t(a))


def test_repeated_calls_bit_identical(torch_cuda):
    """Every kernel reads only data written before it in the same workgroup
    or launch: repeated transforms of the same device input agree bit for bit,
    for spin 0 and spin 2, with high orders whose l-blocks hold a few l values
    (lmax > 2*nside) and 2-tile column blocks (nsim = 48)."""
    torch = torch_cuda
    from fastsht_amd import SHT, empty_maps
    ns, lmax, nsim = 256, 600, 48
    print(f"case repeated calls nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    npix = 12 * ns * ns
    rng = np.random.default_rng(23)
    maps = empty_maps(npix, nsim, device=True)
    maps.copy_(torch.from_numpy(np.asfortranarray(rng.standard_normal((npix, nsim)))))
    sht = SHT(ns, lmax, nsim, niter=1)
    ref = sht.t2alm(maps).cpu().numpy()
    mref = sht.alm2t(sht.t2alm(maps)).cpu().numpy()
    for _ in range(3):
        assert np.array_equal(sht.t2alm(maps).cpu().numpy(), ref)
    assert np.array_equal(sht.alm2t(sht.t2alm(maps)).cpu().numpy(), mref)
    sp = SHT(ns, lmax, 16, niter=1, pol=True)
    Q = empty_maps(npix, 16, device=True)
    U = empty_maps(npix, 16, device=True)
    Q.copy_(torch.from_numpy(np.asfortranarray(rng.standard_normal((npix, 16)))))
    U.copy_(torch.from_numpy(np.asfortranarray(rng.standard_normal((npix, 16)))))
    E0, B0 = (x.cpu().numpy() for x in sp.qu2eb(Q, U))
    for _ in range(2):
        E, B = sp.qu2eb(Q, U)
        assert np.array_equal(E.cpu().numpy(), E0) and np.array_equal(B.cpu().numpy(), B0)


def test_host_staging_matches_device(torch_cuda):
    """Host arrays move through pinned staging in chunks of whole maps (here
    3 map chunks and 2 alm chunks), each chunk's ring FFTs starting on its
    arrival; pinned host arrays move by direct DMA.  Results are bit-identical
    to the device-tensor path."""
    torch = torch_cuda
    from fastsht_amd import SHT, empty_maps
    ns, lmax, nsim = 256, 600, 48
    print(f"case host staging nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    npix = 12 * ns * ns
    rng = np.random.default_rng(17)
    maps = np.asfortranarray(rng.standard_normal((npix, nsim)))     # pageable
    sht = SHT(ns, lmax, nsim, niter=1)
    dmaps = empty_maps(npix, nsim, device=True)
    dmaps.copy_(torch.from_numpy(maps))
    a_dev = sht.t2alm(dmaps).cpu().numpy()
    a_host = sht.t2alm(maps)
    assert np.array_equal(a_host, a_dev)
    pmaps = empty_maps(npix, nsim)                                    # pinned
    pmaps[...] = maps
    assert np.array_equal(sht.t2alm(pmaps), a_dev)
    m_host = sht.alm2t(np.asfortranarray(a_host))                     # pageable alm in, pinned maps out
    m_dev = sht.alm2t(torch.from_numpy(a_dev).cuda().permute(2, 1, 0).contiguous().permute(2, 1, 0))
    assert np.array_equal(m_host, m_dev.cpu().numpy())
    # spin 2: two components per direction
    sp = SHT(ns, 200, 8, niter=0, pol=True)
    Q = np.asfortranarray(rng.standard_normal((npix, 8)))
    U = np.asfortranarray(rng.standard_normal((npix, 8)))
    E, B = sp.qu2eb(Q, U)
    dQ, dU = empty_maps(npix, 8, device=True), empty_maps(npix, 8, device=True)
    dQ.copy_(torch.from_numpy(Q))
    dU.copy_(torch.from_numpy(U))
    Ed, Bd = sp.qu2eb(dQ, dU)
    assert np.array_equal(E, Ed.cpu().numpy()) and np.array_equal(B, Bd.cpu().numpy())
    Qh, Uh = sp.eb2qu(np.asfortranarray(E), np.asfortranarray(B))
    Qd, Ud = sp.eb2qu(Ed, Bd)
    assert np.array_equal(Qh, Qd.cpu().numpy()) and np.array_equal(Uh, Ud.cpu().numpy())
