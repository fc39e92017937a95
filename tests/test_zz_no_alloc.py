"""The allocation-free EB-leakage cleaning (fsht_fix_eb_noalloc) against the
allocating entry point (fsht_fix_eb, device arrays), called through the C ABI
on the same inputs: two calls in a row on one device, the second with a
smaller configuration (the shared scratch array keeps the first call's
contents).  Results must be identical."""
import ctypes

import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run(torch, nside, lmax, nsim, niter, seed):
    from fastsht_amd import _lib
    from oracle import oracle as O
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.fsht_create(nside, lmax, nsim, niter, 1, ctypes.byref(h)))
    try:
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.fsht_set_stream(h, ctypes.c_void_p(s)), h)
        npix = 12 * nside * nside
        rng = np.random.default_rng(seed)
        Q0 = rng.standard_normal((nsim, npix))
        U0 = rng.standard_normal((nsim, npix))
        mask = torch.from_numpy(O.galactic_cut_mask(nside, zcut=0.3)).cuda().contiguous()
        out = {}
        for name in ("alloc", "noalloc"):
            Q = torch.from_numpy(Q0).cuda().contiguous()  # (nsim, npix) C order == (npix, nsim) Fortran
            U = torch.from_numpy(U0).cuda().contiguous()
            B = torch.empty_like(Q)
            tpl = torch.empty_like(Q)
            coef = torch.empty(2 * nsim, dtype=torch.float64, device="cuda")
            almB = torch.empty(nsim * (lmax + 1) ** 2, dtype=torch.float64, device="cuda")
            args = [ctypes.c_void_p(t.data_ptr()) for t in (Q, U, mask, B, almB, tpl, coef)]
            if name == "alloc":
                _lib.check(lib.fsht_fix_eb(h, *args, niter, _lib.FSHT_DEVICE), h)
            else:
                _lib.check(lib.fsht_fix_eb_noalloc(h, *args, niter), h)
            torch.cuda.synchronize()
            out[name] = [t.cpu().numpy() for t in (Q, U, B, almB, tpl, coef)]
        return out
    finally:
        lib.fsht_destroy(h)


def test_noalloc_matches_fix_eb_twice(torch_cuda):
    for nside, lmax, nsim, niter, seed in [(32, 64, 6, 1, 1), (16, 30, 3, 0, 2)]:
        print(f"case fix_eb_noalloc nside={nside} lmax={lmax} nsim={nsim} niter={niter}", flush=True)
        out = _run(torch_cuda, nside, lmax, nsim, niter, seed)
        for a, b in zip(out["alloc"], out["noalloc"]):
            assert np.array_equal(a, b) or maxrel(b, a) <= 1e-12


def test_noalloc_rejects_oversized_configuration(torch_cuda):
    """A configuration whose temporaries exceed the 128 MiB scratch is an error, not an allocation."""
    from fastsht_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.fsht_create(256, 512, 16, 0, 1, ctypes.byref(h)))
    try:
        t = torch_cuda.zeros(16 * 12 * 256 * 256, dtype=torch_cuda.float64, device="cuda")
        m = torch_cuda.ones(12 * 256 * 256, dtype=torch_cuda.float64, device="cuda")
        p = ctypes.c_void_p(t.data_ptr())
        rc = lib.fsht_fix_eb_noalloc(h, p, p, ctypes.c_void_p(m.data_ptr()), p, None, None, None, 0)
        assert rc == _lib.FSHT_ERR_ARG
        assert b"scratch" in lib.fsht_last_error(h)
    finally:
        lib.fsht_destroy(h)
