"""GPU parity on the benchmark configurations and on the reference's lmax
sweep (scripts/test_comprehensive.py:118-119), against the oracle.

- nside 512 / lmax 1535 (the FP64 GPU configuration): t2alm niter 0 and 3 and
  alm2t against the oracle's outputs written by tests/golden/make_c1.py (the
  oracle takes minutes at this size), sampled orders and pixels.
- nside 1024 / lmax 2047 spin 2 with 200 Q/U pairs (the polarised
  configuration's column blocking): qu2eb for sampled orders and maps against
  the oracle's per-order QU2EB.
- nside 32 (spin 0 and 2) and nside 128 (spin 0) over the eleven lmax values
  nside-1, nside, nside+1, 2 nside-1, 2 nside, 2 nside+1, 3 nside-1, 3 nside,
  3 nside+1, 4 nside-1, 4 nside.
Gate: max |diff| / max |oracle| <= 1e-12."""
import os

import numpy as np
import pytest

from conftest import maxrel
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def sweep(ns):
    return [ns - 1, ns, ns + 1, 2 * ns - 1, 2 * ns, 2 * ns + 1, 3 * ns - 1, 3 * ns, 3 * ns + 1, 4 * ns - 1, 4 * ns]


def test_c1_config_vs_oracle_fixture(torch_cuda):
    """nside 512, lmax 1535, 3 maps built as the benchmark configurations."""
    from fastsht_amd import SHT
    print("case C1 nside=512 lmax=1535 nsim=3", flush=True)
    g = np.load(os.path.join(HERE, "golden", "c1_nside512.npz"))
    ns, lmax, nsim = 512, 1535, 3
    L = lmax + 1
    orders = [int(m) for m in g["orders"]]
    alm_in = O.random_alm(lmax, nsim, seed=0)
    rng = np.random.Generator(np.random.PCG64(1))
    noise = np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim))) * np.sqrt(1e-3)
    S = SHT(ns, lmax, nsim)
    clean = S.alm2t(alm_in)
    assert maxrel(clean[::int(g["pix_stride"])], g["alm2t_pix"]) < TOL
    maps = np.asfortranarray(clean + noise)
    for niter in (0, 3):
        S.niter = niter
        a = S.t2alm(maps)
        re = np.stack([a[:, :, m] for m in orders], axis=-1)
        im = np.stack([a[:, :, L - m] if m > 0 else np.zeros_like(a[:, :, 0]) for m in orders], axis=-1)
        err = max(maxrel(re, g[f"t2alm{niter}_re"]), maxrel(im, g[f"t2alm{niter}_im"]))
        print(f"  t2alm niter={niter}: {err:.2e}", flush=True)
        assert err < TOL


def test_c1_shape_nside128_nsim100(torch_cuda):
    """lmax = 3 nside - 1 as in the nside 512 configuration, with 100 maps
    (seven 16-column tiles, both column-block kinds) against the live oracle."""
    from fastsht_amd import SHT
    ns, lmax, nsim = 128, 383, 100
    print(f"case C1 shape nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    R = O.OracleSHT(ns, lmax, nsim)
    alm = O.random_alm(lmax, nsim, seed=7)
    rng = np.random.Generator(np.random.PCG64(8))
    clean = R.alm2t(alm)
    S = SHT(ns, lmax, nsim, niter=3)
    assert maxrel(S.alm2t(alm), clean) < TOL
    maps = np.asfortranarray(clean + rng.standard_normal((12 * ns * ns, nsim)) * np.sqrt(1e-3))
    assert maxrel(S.t2alm(maps), R.t2alm(maps, niter=3)) < TOL
    S.niter = 0
    assert maxrel(S.t2alm(maps), R.t2alm(maps)) < TOL


def test_c2_spin2_sampled_orders_nsim200(torch_cuda):
    """qu2eb on 200 seeded Q/U pairs at nside 1024 / lmax 2047; maps 0, 101
    and 199 and seven orders against the oracle."""
    torch = torch_cuda
    from fastsht_amd import SHT
    ns, lmax, nsim = 1024, 2047, 200
    L = lmax + 1
    print(f"case C2 spin2 nside={ns} lmax={lmax} nsim={nsim}", flush=True)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    npix = 12 * ns * ns
    Q = torch.randn((nsim, npix), dtype=torch.float64, device="cuda", generator=gen).t()
    U = torch.randn((nsim, npix), dtype=torch.float64, device="cuda", generator=gen).t()
    S = SHT(ns, lmax, nsim, pol=True)
    E, B = S.qu2eb(Q, U)
    torch.cuda.synchronize()
    sims = [0, 101, 199]
    orders = [0, 1, 2, 700, 1500, 2046, 2047]
    Qh = np.asfortranarray(Q[:, sims].cpu().numpy())
    Uh = np.asfortranarray(U[:, sims].cpu().numpy())
    del Q, U
    Eg = E[sims].cpu().numpy()
    Bg = B[sims].cpu().numpy()
    del E, B
    R = O.OracleSHT(ns, lmax, len(sims))
    Eo, Bo = R.qu2eb_orders(Qh, Uh, orders)
    mask = np.zeros((L, L), dtype=bool)  # storage entries of the sampled orders
    for m in orders:
        mask[m:, m] = True                # Re a_lm, l >= m
        if m > 0:
            mask[:L - m, L - m] = True    # Im a_lm at row l - m
    err = max(maxrel(Eg[:, mask], Eo[:, mask]), maxrel(Bg[:, mask], Bo[:, mask]))
    print(f"  qu2eb sampled orders: {err:.2e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("lmax", sweep(32))
def test_lmax_sweep_nside32(torch_cuda, lmax):
    from fastsht_amd import SHT
    ns, nsim = 32, 3
    print(f"case sweep nside={ns} lmax={lmax}", flush=True)
    rng = np.random.default_rng(lmax)
    maps = np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim)))
    Q = np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim)))
    U = np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim)))
    R = O.OracleSHT(ns, lmax, nsim)
    S = SHT(ns, lmax, nsim, niter=3)
    assert np.array_equal(S.ring_tab(), R.rtab.astype(np.int32))
    assert maxrel(S.t2alm(maps), R.t2alm(maps, niter=3)) < TOL
    S.niter = 0
    assert maxrel(S.t2alm(maps), R.t2alm(maps)) < TOL
    alm = O.random_alm(lmax, nsim, seed=lmax)
    assert maxrel(S.alm2t(alm), R.alm2t(alm)) < TOL
    P = SHT(ns, lmax, nsim, niter=2, pol=True)
    E, B = P.qu2eb(Q, U)
    Eo, Bo = R.qu2eb(Q, U, niter=2)
    assert maxrel(E, Eo) < TOL and maxrel(B, Bo) < TOL
    Ea, Ba = O.random_alm(lmax, nsim, seed=1), O.random_alm(lmax, nsim, seed=2)
    Qg, Ug = P.eb2qu(Ea, Ba)
    Qo, Uo = R.eb2qu(Ea, Ba)
    assert maxrel(Qg, Qo) < TOL and maxrel(Ug, Uo) < TOL


@pytest.mark.parametrize("lmax", sweep(128))
def test_lmax_sweep_nside128_spin0(torch_cuda, lmax):
    from fastsht_amd import SHT
    ns, nsim = 128, 2
    print(f"case sweep nside={ns} lmax={lmax}", flush=True)
    rng = np.random.default_rng(1000 + lmax)
    maps = np.asfortranarray(rng.standard_normal((12 * ns * ns, nsim)))
    R = O.OracleSHT(ns, lmax, nsim)
    S = SHT(ns, lmax, nsim, niter=3)
    assert np.array_equal(S.ring_tab(), R.rtab.astype(np.int32))
    assert maxrel(S.t2alm(maps), R.t2alm(maps, niter=3)) < TOL
    alm = O.random_alm(lmax, nsim, seed=lmax)
    assert maxrel(S.alm2t(alm), R.alm2t(alm)) < TOL
