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
