"""Angular power spectra of packed alm and seeded Gaussian alm from spectra,
on the GPU (the reference's alm2cl_T1 / alm2cl_TEB1, src/alm2cl_TEB.f90:22-110,
and cl2alm_T1 / cl2alm_TEB1, src/cl2alm_TEB.f90:25-170).

Arrays use the reference's layouts: alm (nsim, lmax+1, lmax+1) Fortran order
(packed square), cl (nsim, lmax+1) for T or (nsim, lmax+1, 6) for TT EE BB TE
TB EB.  NumPy inputs are copied to the GPU and the result comes back as NumPy;
torch CUDA tensors (Fortran-order memory) stay on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .sht import _fortran_strides, _is_torch, _torch

__all__ = ["alm2cl", "cl2alm"]


def _dev_fortran(x, shape, name):
    """(device tensor with Fortran-order memory, was_host)."""
    torch = _torch()
    if _is_torch(x):
        if not x.is_cuda or x.dtype != torch.float64 or tuple(x.shape) != tuple(shape) \
                or tuple(x.stride()) != _fortran_strides(shape):
            raise ValueError(f"{name}: expected a float64 GPU tensor of Fortran-order shape {shape}")
        return x, False
    a = np.asarray(x, dtype=np.float64)
    if a.shape != tuple(shape):
        raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
    t = torch.from_numpy(np.ascontiguousarray(a.transpose(tuple(range(a.ndim))[::-1]))).cuda()
    return t.permute(tuple(range(a.ndim))[::-1]), True


def _empty_fortran(shape):
    torch = _torch()
    return torch.empty(tuple(shape)[::-1], dtype=torch.float64, device="cuda").permute(
        tuple(range(len(shape)))[::-1])


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def alm2cl(almT, almE=None, almB=None, lmin=0):
    """Spectra of packed alm: (nsim, lmax+1) for T alone, or (nsim, lmax+1, 6)
    = TT EE BB TE TB EB when almE and almB are given.  Entries with l < lmin
    are 0."""
    torch = _torch()
    lib = _lib.load()
    nsim, L, _ = almT.shape
    shape = (nsim, L, L)
    aT, host = _dev_fortran(almT, shape, "almT")
    pol = almE is not None
    if pol != (almB is not None):
        raise ValueError("almE and almB go together")
    aE = _dev_fortran(almE, shape, "almE")[0] if pol else None
    aB = _dev_fortran(almB, shape, "almB")[0] if pol else None
    cshape = (nsim, L, 6) if pol else (nsim, L)
    cl = _empty_fortran(cshape)
    cl.zero_()
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.fsht_alm2cl(_ptr(aT), _ptr(aE), _ptr(aB), _ptr(cl), int(lmin), L - 1, nsim, ctypes.c_void_p(s)))
    return cl.cpu().numpy() if host else cl


def cl2alm(cl, lmax, nsim, seed=0, lmin=0, device=False):
    """Seeded Gaussian alm with spectrum cl: cl of shape (lmax+1,) gives almT;
    cl of shape (lmax+1, 6) (TT EE BB TE TB EB) gives (almT, almE, almB).
    The stream is Philox4x32-10 + Box-Muller keyed by `seed` (the reference's
    MKL VSL stream is not reproduced)."""
    torch = _torch()
    lib = _lib.load()
    L = int(lmax) + 1
    c = cl if _is_torch(cl) else torch.as_tensor(np.asarray(cl, dtype=np.float64))
    pol = c.ndim == 2
    if tuple(c.shape) != ((L, 6) if pol else (L,)):
        raise ValueError(f"cl must have shape ({L},) or ({L}, 6)")
    cdev = c.to(device="cuda", dtype=torch.float64).t().contiguous().t() if pol else \
        c.to(device="cuda", dtype=torch.float64).contiguous()
    shape = (int(nsim), L, L)
    outs = [_empty_fortran(shape) for _ in range(3 if pol else 1)]
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.fsht_cl2alm(_ptr(cdev), _ptr(outs[0]), _ptr(outs[1]) if pol else None,
                               _ptr(outs[2]) if pol else None, int(lmin), L - 1, int(nsim),
                               ctypes.c_ulonglong(int(seed) & 0xFFFFFFFFFFFFFFFF), ctypes.c_void_p(s)))
    if not device:
        outs = [np.asfortranarray(o.cpu().numpy()) for o in outs]
    return tuple(outs) if pol else outs[0]
