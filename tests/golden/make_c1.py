"""Write tests/golden/c1_nside512.npz: the oracle's outputs on the nside=512,
lmax=1535 configuration (3 maps), sampled, for the GPU parity test that runs
where the oracle at this size would take minutes.

Inputs (rebuilt by the test from the same seeds):
  alm_in = oracle.random_alm(1535, 3, seed=0)
  noise  = PCG64(1) standard normal (npix, 3), Fortran order, times sqrt(1e-3)
  maps   = alm2t(alm_in) + noise
Stored: alm2t(alm_in) at every 111th pixel; t2alm(maps, niter=0) and
t2alm(maps, niter=3) for the orders in ORDERS (Re columns alm[:, l, m] and Im
columns alm[:, l-m, lmax+1-m], every l).

  python tests/golden/make_c1.py      (about ten minutes on 8 cores)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle as O  # noqa: E402

NSIDE, LMAX, NSIM = 512, 1535, 3
ORDERS = [0, 1, 2, 3, 100, 511, 512, 1023, 1300, 1534, 1535]
PIX_STRIDE = 111


def columns(alm):
    L = LMAX + 1
    re = np.stack([alm[:, :, m] for m in ORDERS], axis=-1)
    im = np.stack([alm[:, :, L - m] if m > 0 else np.zeros_like(alm[:, :, 0]) for m in ORDERS], axis=-1)
    return re, im


def main():
    R = O.OracleSHT(NSIDE, LMAX, NSIM)
    alm_in = O.random_alm(LMAX, NSIM, seed=0)
    rng = np.random.Generator(np.random.PCG64(1))
    noise = np.asfortranarray(rng.standard_normal((12 * NSIDE * NSIDE, NSIM))) * np.sqrt(1e-3)
    clean = R.alm2t(alm_in)
    maps = np.asfortranarray(clean + noise)
    out = {"orders": np.array(ORDERS, dtype=np.int32), "pix_stride": np.array(PIX_STRIDE),
           "alm2t_pix": clean[::PIX_STRIDE]}
    for niter in (0, 3):
        re, im = columns(R.t2alm(maps, niter=niter))
        out[f"t2alm{niter}_re"], out[f"t2alm{niter}_im"] = re, im
        print("niter", niter, "done", flush=True)
    np.savez_compressed(os.path.join(HERE, "c1_nside512.npz"), **out)


if __name__ == "__main__":
    main()
