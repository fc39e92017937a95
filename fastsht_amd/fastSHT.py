"""Stand-in for the reference's compiled `fastSHT` module, backed by the C ABI.

The reference's SHT.py drives a Fortran module (built with f90wrap) through
module-level calls on one process-global configuration (SHT.py:61-62, 90-102,
118, 137-139, 155, 171, 188, 201, 206, 223).  This module offers the same
functions with the same argument order on host (NumPy, Fortran-ordered
float64) arrays, so that SHT.py can run unchanged with

    import sys, fastsht_amd.fastSHT
    sys.modules["fastSHT"] = fastsht_amd.fastSHT

The global configuration becomes one library handle.  Tables that the
reference reads from files and passes to sht_set_data are computed by the
library at sht_data_alloc, so sht_set_data ignores its arguments.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_h = ctypes.c_void_p()
_cfg = {}


def _ok(rc):
    _lib.check(rc, _h if _h.value else None)


def _p(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
        raise ValueError(f"{name}: expected a Fortran-contiguous float64 array")
    return ctypes.c_void_p(a.ctypes.data)


def _need():
    if not _h.value:
        raise RuntimeError("sht_data_alloc has not been called")
    return _lib.load()


def sht_data_alloc(params):
    """sht_data_init.f90 sht_data_alloc: params = (nside, lmax, nring, nsim,
    nbuff, pol).  Replaces any earlier configuration."""
    lib = _lib.load()
    nside, lmax, _nring, nsim, _nbuff, pol = (int(x) for x in np.asarray(params).reshape(-1)[:6])
    if _h.value:
        lib.fsht_destroy(_h)
        _h.value = None
    h = ctypes.c_void_p()
    _lib.check(lib.fsht_create(nside, lmax, nsim, 0, pol, ctypes.byref(h)))
    _h.value = h.value
    _cfg.update(nside=nside, lmax=lmax, nsim=nsim, pol=pol)


def sht_set_data(*tables):
    """sht_data_init.f90 sht_set_data: the ring-skip table and Legendre seeds
    are computed by the library, so the tables passed here are not used."""
    _need()


def _t2alm(maps, niter, alms, fn):
    lib = _need()
    out = None if alms is None else _p(alms, "alms")
    _ok(getattr(lib, fn)(_h, _p(maps, "maps"), out, int(niter), _lib.FSHT_HOST))


def t2alm(maps, alms=None):
    """sht_main.f90 t2alm; without alms the result stays on the GPU (get_alm)."""
    _t2alm(maps, 0, alms, "fsht_t2alm")


def t2alm_iter_old(maps, niter, alms=None):
    """sht_main.f90 t2alm_iter_old (Jacobi refinement)."""
    _t2alm(maps, niter, alms, "fsht_t2alm")


def t2alm_iter(maps, niter, alms=None):
    """sht_main.f90 t2alm_iter (immediate-update refinement)."""
    _t2alm(maps, niter, alms, "fsht_t2alm_iter")


def get_alm(alms):
    """sht_main.f90 get_alm: the alms kept on the GPU by t2alm(maps)."""
    _ok(_need().fsht_get_alm(_h, _p(alms, "alms"), _lib.FSHT_HOST))


def alm2t(alms, maps):
    _ok(_need().fsht_alm2t(_h, _p(alms, "alms"), _p(maps, "maps"), _lib.FSHT_HOST))


def qu2eb(Q, U, almEs, almBs):
    _ok(_need().fsht_qu2eb(_h, _p(Q, "Q"), _p(U, "U"), _p(almEs, "almEs"), _p(almBs, "almBs"), 0, _lib.FSHT_HOST))


def qu2eb_iter_old(Q, U, almEs, almBs, niter, skip_fft=0):
    _ok(_need().fsht_qu2eb(_h, _p(Q, "Q"), _p(U, "U"), _p(almEs, "almEs"), _p(almBs, "almBs"), int(niter),
                           _lib.FSHT_HOST))


def qu2e(Q, U, almEs):
    _ok(_need().fsht_qu2e(_h, _p(Q, "Q"), _p(U, "U"), _p(almEs, "almEs"), _lib.FSHT_HOST))


def qu2b(Q, U, almBs):
    _ok(_need().fsht_qu2b(_h, _p(Q, "Q"), _p(U, "U"), _p(almBs, "almBs"), _lib.FSHT_HOST))


def eb2qu(almEs, almBs, Q, U):
    _ok(_need().fsht_eb2qu(_h, _p(almEs, "almEs"), _p(almBs, "almBs"), _p(Q, "Q"), _p(U, "U"), _lib.FSHT_HOST))


def fix_eb(Q, U, mask, Bmap, alms, B_template, vid=None, niter=0, nv=None, flag=0):
    """sht_main.f90 fix_eb: the fit uses the pixels with mask == 1 (the
    reference passes their indices as vid/nv; the library derives them)."""
    lib = _need()
    mk = np.ascontiguousarray(np.asarray(mask, dtype=np.float64).reshape(-1))
    _ok(lib.fsht_fix_eb(_h, _p(Q, "Q"), _p(U, "U"), ctypes.c_void_p(mk.ctypes.data), _p(Bmap, "Bmap"),
                        _p(alms, "alms"), _p(B_template, "B_template"), None, int(niter), _lib.FSHT_HOST))


def convert_alm_healpy1(alms, alms_hp, nsim, lmax):
    """convert_alm.f90 convert_alm_healpy1: alms_hp (2, (lmax+1)(lmax+2)/2, nsim)."""
    _ok(_lib.load().fsht_convert_alm_healpy(_p(alms, "alms"), _p(alms_hp, "alms_hp"), int(lmax), int(nsim),
                                            _lib.FSHT_HOST, None))


def convert_alm_healpy_inv1(alms_hp, alms, nsim, lmax):
    """convert_alm.f90 convert_alm_healpy_inv1."""
    _ok(_lib.load().fsht_convert_alm_healpy_inv(_p(alms_hp, "alms_hp"), _p(alms, "alms"), int(lmax), int(nsim),
                                                _lib.FSHT_HOST, None))
