"""Mixed precision (int8-slice MFMA Legendre contractions, FP64 refinement)
against the FP64 oracle and the FP64 GPU path.

Gate: max |mixed - FP64| / max |FP64| <= 1e-10 for the alm of t2alm(niter)
and for the maps of alm2t.  The FP64 references are the oracle (small sizes)
and the parity-pinned FP64 GPU path (larger sizes)."""
import numpy as np
import pytest

from conftest import maxrel
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10

# lmax from 2 nside - 1 to 4 nside (every aliasing regime, hence every slice
# ladder), odd nsim (column padding), several scaling blocks at nside 128
CASES = [(1, 2, 3), (4, 8, 3), (4, 16, 5), (8, 20, 17), (8, 32, 17), (16, 47, 33), (32, 63, 40), (32, 95, 20),
         (64, 127, 5), (64, 191, 16), (128, 255, 2), (128, 383, 3)]


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _c0_maps(R, lmax, nsim, seed):
    """Maps built as the benchmark configurations: alm2t of Gaussian alm with
    C_l = 1/(l(l+1)) plus N(0, 1e-3) white noise."""
    alm = O.random_alm(lmax, nsim, seed=seed)
    rng = np.random.default_rng(seed + 1)
    return np.asfortranarray(R.alm2t(alm) + rng.standard_normal((R.npix, nsim)) * np.sqrt(1e-3))


@pytest.mark.parametrize("ns,lmax,nsim", CASES)
def test_mixed_vs_oracle(torch_cuda, ns, lmax, nsim):
    from fastsht_amd import SHT
    print(f"case mixed nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    maps = _c0_maps(R, lmax, nsim, seed=ns + lmax)
    for niter in (0, 1, 2, 3):
        S = SHT(ns, lmax, nsim, niter=niter, precision="mixed")
        ref = R.t2alm(maps, niter=niter)
        err = maxrel(S.t2alm(maps), ref)
        print(f"  t2alm niter={niter}: {err:.2e}", flush=True)
        assert err < TOL, (niter, err)
    alm = O.random_alm(lmax, nsim, seed=3)
    err = maxrel(S.alm2t(alm), R.alm2t(alm))
    print(f"  alm2t: {err:.2e}", flush=True)
    assert err < TOL


def test_mixed_vs_fp64_gpu_nside256(torch_cuda):
    """nside 256, lmax 511 (= 2 nside - 1), 48 maps: several scaling blocks
    along both contractions and three column blocks."""
    from fastsht_amd import SHT
    torch = torch_cuda
    ns, lmax, nsim = 256, 511, 48
    print(f"case mixed vs fp64 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    F = SHT(ns, lmax, nsim, niter=3)
    M = SHT(ns, lmax, nsim, niter=3, precision="mixed")
    alm = torch.from_numpy(O.random_alm(lmax, nsim, seed=11)).cuda()
    alm = alm.permute(2, 1, 0).contiguous().permute(2, 1, 0)
    maps = F.alm2t(alm)
    maps += 0.0316 * torch.randn(maps.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    a64 = F.t2alm(maps)
    amx = M.t2alm(maps)
    torch.cuda.synchronize()
    err = maxrel(amx.cpu().numpy(), a64.cpu().numpy())
    print(f"  t2alm niter=3: {err:.2e}", flush=True)
    assert err < TOL
    m64 = F.alm2t(a64)
    mmx = M.alm2t(a64)
    err = maxrel(mmx.cpu().numpy(), m64.cpu().numpy())
    print(f"  alm2t: {err:.2e}", flush=True)
    assert err < TOL


# ---------------------------------------------------------------------------
# Spin 2: qu2eb(niter) and eb2qu with int8-slice F1/F2 contractions (tables
# generated per pass, one range of orders at a time) against the oracle's
# QU2EB / qu2eb_iter_old / EB2QU, and against the FP64 GPU path.
# ---------------------------------------------------------------------------
# lmax = 2 nside - 1 and the stronger aliasing of 3 nside - 1 and 4 nside (other ladders)
SPIN2_CASES = [(4, 8, 3), (8, 20, 5), (16, 47, 7), (32, 63, 9), (32, 95, 5), (32, 128, 4), (64, 127, 5)]


def _qu_maps(R, lmax, nsim, seed):
    """Q/U of seeded E/B alm (C_l^BB = 0.05 C_l^EE) plus N(0, 1e-3) noise."""
    E = O.random_alm(lmax, nsim, seed=seed)
    B = O.random_alm(lmax, nsim, seed=seed + 1) * np.sqrt(0.05)
    Q, U = R.eb2qu(E, B)
    rng = np.random.default_rng(seed + 2)
    Q = np.asfortranarray(Q + rng.standard_normal(Q.shape) * np.sqrt(1e-3))
    U = np.asfortranarray(U + rng.standard_normal(U.shape) * np.sqrt(1e-3))
    return Q, U


def _spin2_case(ns, lmax, nsim):
    from fastsht_amd import SHT
    R = O.OracleSHT(ns, lmax, nsim)
    Q, U = _qu_maps(R, lmax, nsim, seed=ns + lmax)
    for niter in (0, 1, 2, 3):
        S = SHT(ns, lmax, nsim, niter=niter, pol=True, precision="mixed")
        E, B = S.qu2eb(Q, U)
        Eo, Bo = R.qu2eb(Q, U, niter=niter)
        err = max(maxrel(E, Eo), maxrel(B, Bo))
        print(f"  qu2eb niter={niter}: {err:.2e}", flush=True)
        assert err < TOL, (niter, err)
    Ea, Ba = O.random_alm(lmax, nsim, seed=5), O.random_alm(lmax, nsim, seed=6)
    Qg, Ug = S.eb2qu(Ea, Ba)
    Qo, Uo = R.eb2qu(Ea, Ba)
    err = max(maxrel(Qg, Qo), maxrel(Ug, Uo))
    print(f"  eb2qu: {err:.2e}", flush=True)
    assert err < TOL
    # the same handle's temperature transforms (spin-0 tables built on first use)
    err = maxrel(S.t2alm(Q), R.t2alm(Q, niter=3))
    print(f"  t2alm niter=3 on the spin-2 handle: {err:.2e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("ns,lmax,nsim", SPIN2_CASES)
def test_mixed_spin2_vs_oracle(torch_cuda, ns, lmax, nsim):
    print(f"case mixed spin2 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    _spin2_case(ns, lmax, nsim)


def test_mixed_spin2_many_ranges(torch_cuda, monkeypatch):
    """Tables generated in ranges of orders of at most 1 MB each (10 ranges
    at nside 64 / lmax 127): the same results as one range."""
    monkeypatch.setenv("FSHT_T2_CAP_MB", "1")
    print("case mixed spin2 ranges nside=64 lmax=127 nsim=5 cap=1MB", flush=True)
    _spin2_case(64, 127, 5)


def test_planned_mixed_bytes_match_allocation(torch_cuda):
    """fsht_info's planned mixed-precision bytes (out[36..39)) equal the bytes
    the library then holds (out[9]) after each part is built: the sliced
    operands at fsht_set_precision, the temperature tables at the first mixed
    t2alm, the spin-2 buffers at the first mixed qu2eb (also with ranges
    capped at 1 MB)."""
    import ctypes

    from fastsht_amd import SHT, _lib
    lib = _lib.load()

    def info(S):
        v = (ctypes.c_longlong * 39)()
        _lib.check(lib.fsht_info(S._h, v, 39), S._h)
        return list(v)

    for ns, lmax, nsim, cap in ((32, 64, 20, None), (64, 127, 5, "1"), (128, 383, 40, None)):
        print(f"case planned bytes nside={ns} lmax={lmax} nsim={nsim} range cap={cap}", flush=True)
        with pytest.MonkeyPatch.context() as mp:
            if cap:
                mp.setenv("FSHT_T2_CAP_MB", cap)
            S = SHT(ns, lmax, nsim, niter=1, pol=True)
            v = info(S)
            planned = v[36:39]
            assert v[9] == 0 and min(planned) > 0
            S.set_precision("mixed")
            assert info(S)[9] == planned[0]
            R = O.OracleSHT(ns, lmax, nsim)
            Q, U = _qu_maps(R, lmax, nsim, seed=3)
            S.t2alm(Q)
            assert info(S)[9] == planned[0] + planned[1]
            S.qu2eb(Q, U)
            v = info(S)
            assert v[9] == sum(planned) and v[36:39] == planned


def test_mixed_spin2_vs_fp64_gpu_nside256(torch_cuda):
    """nside 256, lmax 511, 40 Q/U pairs (six 32-column tiles: one full block
    and two single tiles): qu2eb niter 3 and eb2qu against the FP64 GPU path."""
    from fastsht_amd import SHT
    torch = torch_cuda
    ns, lmax, nsim = 256, 511, 40
    print(f"case mixed spin2 vs fp64 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    F = SHT(ns, lmax, nsim, niter=3, pol=True)
    M = SHT(ns, lmax, nsim, niter=3, pol=True, precision="mixed")

    def dev(a):
        t = torch.from_numpy(a).cuda()
        return t.permute(2, 1, 0).contiguous().permute(2, 1, 0)

    E0 = dev(O.random_alm(lmax, nsim, seed=21))
    B0 = dev(O.random_alm(lmax, nsim, seed=22)) * 0.2236
    Q, U = F.eb2qu(E0, B0)
    g = torch.Generator("cuda").manual_seed(7)
    Q += 0.0316 * torch.randn(Q.shape, dtype=torch.float64, device="cuda", generator=g)
    U += 0.0316 * torch.randn(U.shape, dtype=torch.float64, device="cuda", generator=g)
    E64, B64 = F.qu2eb(Q, U)
    Emx, Bmx = M.qu2eb(Q, U)
    torch.cuda.synchronize()
    err = max(maxrel(Emx.cpu().numpy(), E64.cpu().numpy()), maxrel(Bmx.cpu().numpy(), B64.cpu().numpy()))
    print(f"  qu2eb niter=3: {err:.2e}", flush=True)
    assert err < TOL
    Q64, U64 = F.eb2qu(E64, B64)
    Qmx, Umx = M.eb2qu(E64, B64)
    err = max(maxrel(Qmx.cpu().numpy(), Q64.cpu().numpy()), maxrel(Umx.cpu().numpy(), U64.cpu().numpy()))
    print(f"  eb2qu: {err:.2e}", flush=True)
    assert err < TOL
