"""Drop-in `SHT` object (same constructor, methods, defaults and return
conventions as the reference's SHT.py:16-227) over libfastsht.

Arrays are float64 in the reference's Fortran-order layouts:
  maps (npix, nsim), alms (nsim, lmax+1, lmax+1).
They may be NumPy arrays (host memory: copied to the GPU and back) or torch
CUDA tensors whose memory has that Fortran-order layout (kept on the device;
the work is enqueued on torch's current stream).  `empty_maps` / `empty_alms`
allocate either kind.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib

__all__ = ["SHT", "empty_maps", "empty_alms"]


def _torch():
    import torch  # deferred: only device arrays need it

    return torch


def _is_torch(x):
    return type(x).__module__.startswith("torch") and hasattr(x, "data_ptr")


def _pinned_empty(shape):
    """Fortran-order float64 host array in page-locked memory (freed when the
    array is collected); host-pointer transforms move it by direct DMA."""
    lib = _lib.load()
    n = int(np.prod(shape))
    p = ctypes.c_void_p()
    _lib.check(lib.fsht_host_alloc(n * 8, ctypes.byref(p)))
    buf = (ctypes.c_double * max(n, 1)).from_address(p.value)
    weakref.finalize(buf, lib.fsht_host_free, ctypes.c_void_p(p.value))
    return np.frombuffer(buf, dtype=np.float64, count=n).reshape(shape, order="F")


def empty_maps(npix, nsim, device=False, pinned=True):
    """(npix, nsim) Fortran-order float64 maps on the GPU, or on the host
    (page-locked unless pinned=False)."""
    if device:
        torch = _torch()
        return torch.empty((nsim, npix), dtype=torch.float64, device="cuda").t()
    if pinned:
        return _pinned_empty((npix, nsim))
    return np.empty((npix, nsim), dtype=np.float64, order="F")


def empty_alms(lmax, nsim, device=False, pinned=True):
    """(nsim, lmax+1, lmax+1) Fortran-order float64 alms on the GPU, or on the
    host (page-locked unless pinned=False)."""
    L = lmax + 1
    if device:
        torch = _torch()
        return torch.empty((L, L, nsim), dtype=torch.float64, device="cuda").permute(2, 1, 0)
    if pinned:
        return _pinned_empty((nsim, L, L))
    return np.empty((nsim, L, L), dtype=np.float64, order="F")


def _fortran_strides(shape):
    st, acc = [], 1
    for n in shape:
        st.append(acc)
        acc *= n
    return tuple(st)


class SHT:
    """Batched HEALPix-RING spherical harmonic transforms on one MI355X GPU.

    Reference: SHT.py:16-62 (constructor).  `all_buff` is accepted for
    signature compatibility; workspace is sized by the library itself.
    """

    def __init__(self, nside, lmax, nsim, niter=0, pol=False, all_buff=False, *, precision="fp64"):
        """precision="mixed" runs the Legendre contractions of t2alm / alm2t
        and of qu2eb / eb2qu on int8-slice MFMA, refined in FP64 (within
        1e-10 of the FP64 result); qu2e and qu2b stay FP64."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.nside = int(nside)
        self.lmax = int(lmax)
        self.nsim = int(nsim)
        self.niter = int(niter)
        self.pol = 1 if pol else 0
        self.nring = 4 * self.nside - 1
        self.npix = 12 * self.nside ** 2
        self.maps = None
        self.alms = None
        self._h = None
        self._stream = None
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.fsht_create(self.nside, self.lmax, self.nsim, self.niter, self.pol, ctypes.byref(h)))
        self._h = h
        if precision == "mixed":
            self.set_precision("mixed")

    def set_precision(self, precision):
        """Switch the transforms between "fp64" and "mixed" (the first switch
        to "mixed" builds the int8 Legendre tables of a temperature-only
        handle; a spin-2 capable handle builds its temperature tables and its
        spin-2 state at its first mixed transform of each spin)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        code = _lib.FSHT_PREC_MIXED if precision == "mixed" else _lib.FSHT_PREC_FP64
        self._call(_lib.load().fsht_set_precision, code)
        self.precision = precision

    def set_batch(self, nsim):
        """Run the following calls on nsim maps (1 <= nsim <= the nsim the
        object was created with); no counterpart in the reference."""
        self._call(_lib.load().fsht_set_batch, int(nsim))
        self.nsim = int(nsim)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().fsht_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- argument handling ------------------------------------------------
    def _call(self, fn, *args):
        _lib.check(fn(self._h, *args), self._h)

    def _arg(self, x, shape, name, out=False):
        """(pointer, loc) for an input/output array of the given logical shape."""
        if x is None:
            raise ValueError(f"{name} is None")
        if _is_torch(x):
            torch = _torch()
            if not x.is_cuda or x.dtype != torch.float64:
                raise TypeError(f"{name}: torch tensors must be float64 on the GPU")
            if tuple(x.shape) != tuple(shape) or tuple(x.stride()) != _fortran_strides(shape):
                raise ValueError(f"{name}: expected Fortran-order shape {shape}, got shape "
                                 f"{tuple(x.shape)} strides {tuple(x.stride())}")
            self._use_torch_stream()
            return ctypes.c_void_p(x.data_ptr()), _lib.FSHT_DEVICE, x
        a = x
        if out:
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous
                    and a.shape == tuple(shape)):
                raise ValueError(f"{name}: output must be a Fortran-contiguous float64 array of shape {shape}")
        else:
            a = np.asfortranarray(a, dtype=np.float64)
            if a.shape != tuple(shape):
                raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        return ctypes.c_void_p(a.ctypes.data), _lib.FSHT_HOST, a

    def _use_torch_stream(self):
        s = _torch().cuda.current_stream().cuda_stream
        if s != self._stream:
            _lib.check(_lib.load().fsht_set_stream(self._h, ctypes.c_void_p(s)), self._h)
            self._stream = s

    def _like(self, ref, kind):
        dev = _is_torch(ref)
        if kind == "alm":
            return empty_alms(self.lmax, self.nsim, device=dev)
        return empty_maps(self.npix, self.nsim, device=dev)

    @property
    def _mshape(self):
        return (self.npix, self.nsim)

    @property
    def _ashape(self):
        return (self.nsim, self.lmax + 1, self.lmax + 1)

    # ---- reference API ----------------------------------------------------
    def set_maps(self, maps_in):
        """SHT.py:65-66."""
        self.maps = maps_in

    def set_alms(self, alms_in):
        """SHT.py:69-70 (the reference's version lacks `self`)."""
        self.alms = alms_in

    def t2alm(self, maps_in=None, alms_in=None, method="old", get_alm=True):
        """Temperature analysis (SHT.py:72-107): niter = 0 -> t2alm; niter > 0
        -> t2alm_iter_old (Jacobi) for method='old', else t2alm_iter
        (immediate update).  get_alm=False keeps the result on the GPU for
        get_alms()."""
        maps = self.maps if maps_in is None else maps_in
        if maps is None:
            raise ValueError("No input maps AND no internal maps! Check the parameters or call set_maps(maps_in)!")
        lib = _lib.load()
        fn = lib.fsht_t2alm if (self.niter == 0 or method == "old") else lib.fsht_t2alm_iter
        mp, loc, mkeep = self._arg(maps, self._mshape, "maps")
        if not get_alm:
            self._call(fn, mp, None, self.niter, loc)
            return None
        alms = self._like(maps, "alm") if alms_in is None else alms_in
        ap, aloc, akeep = self._arg(alms, self._ashape, "alms", out=True)
        if aloc != loc:
            raise ValueError("maps and alms must both be host arrays or both be GPU tensors")
        self._call(fn, mp, ap, self.niter, loc)
        return alms if alms_in is None else 0

    def alm2t(self, alms, maps_in=None):
        """Temperature synthesis (SHT.py:109-122)."""
        lib = _lib.load()
        ap, loc, akeep = self._arg(alms, self._ashape, "alms")
        maps = self._like(alms, "map") if maps_in is None else maps_in
        mp, mloc, mkeep = self._arg(maps, self._mshape, "maps", out=True)
        if mloc != loc:
            raise ValueError("alms and maps must both be host arrays or both be GPU tensors")
        self._call(lib.fsht_alm2t, ap, mp, loc)
        return maps if maps_in is None else 0

    def _need_pol(self):
        if not self.pol:
            raise ValueError("spin-2 transforms need SHT(..., pol=True)")

    def qu2eb(self, Q, U, almEs_in=None, almBs_in=None, method="old"):
        """Spin-2 analysis (SHT.py:124-144): niter = 0 -> QU2EB, niter > 0 ->
        qu2eb_iter_old."""
        self._need_pol()
        if self.niter > 0 and method != "old":
            raise NotImplementedError("qu2eb(method != 'old') with niter > 0 has no implementation "
                                      "in the reference either (it returns uninitialised arrays); use method='old'")
        lib = _lib.load()
        qp, loc, qk = self._arg(Q, self._mshape, "Q")
        up, uloc, uk = self._arg(U, self._mshape, "U")
        fresh = almEs_in is None or almBs_in is None
        E = self._like(Q, "alm") if fresh else almEs_in
        B = self._like(Q, "alm") if fresh else almBs_in
        ep, eloc, ek = self._arg(E, self._ashape, "almE", out=True)
        bp, bloc, bk = self._arg(B, self._ashape, "almB", out=True)
        if len({loc, uloc, eloc, bloc}) != 1:
            raise ValueError("all arrays must be host arrays or all GPU tensors")
        self._call(lib.fsht_qu2eb, qp, up, ep, bp, self.niter, loc)
        return (E, B) if fresh else 0

    def qu2e(self, Q, U, almEs_in=None):
        """E only (SHT.py:146-160 -> QU2E)."""
        self._need_pol()
        lib = _lib.load()
        qp, loc, qk = self._arg(Q, self._mshape, "Q")
        up, uloc, uk = self._arg(U, self._mshape, "U")
        E = self._like(Q, "alm") if almEs_in is None else almEs_in
        ep, eloc, ek = self._arg(E, self._ashape, "almE", out=True)
        if len({loc, uloc, eloc}) != 1:
            raise ValueError("all arrays must be host arrays or all GPU tensors")
        self._call(lib.fsht_qu2e, qp, up, ep, loc)
        return E if almEs_in is None else 0

    def qu2b(self, Q, U, almBs_in=None):
        """B only (SHT.py:162-176 -> QU2B)."""
        self._need_pol()
        lib = _lib.load()
        qp, loc, qk = self._arg(Q, self._mshape, "Q")
        up, uloc, uk = self._arg(U, self._mshape, "U")
        B = self._like(Q, "alm") if almBs_in is None else almBs_in
        bp, bloc, bk = self._arg(B, self._ashape, "almB", out=True)
        if len({loc, uloc, bloc}) != 1:
            raise ValueError("all arrays must be host arrays or all GPU tensors")
        self._call(lib.fsht_qu2b, qp, up, bp, loc)
        return B if almBs_in is None else 0

    def eb2qu(self, almEs, almBs):
        """Spin-2 synthesis (SHT.py:179-190 -> EB2QU)."""
        self._need_pol()
        lib = _lib.load()
        ep, loc, ek = self._arg(almEs, self._ashape, "almE")
        bp, bloc, bk = self._arg(almBs, self._ashape, "almB")
        Q = self._like(almEs, "map")
        U = self._like(almEs, "map")
        qp, qloc, qk = self._arg(Q, self._mshape, "Q", out=True)
        up, uloc, uk = self._arg(U, self._mshape, "U", out=True)
        if len({loc, bloc, qloc, uloc}) != 1:
            raise ValueError("all arrays must be host arrays or all GPU tensors")
        self._call(lib.fsht_eb2qu, ep, bp, qp, up, loc)
        return (Q, U)

    def fix_eb(self, Q, U, mask, Btpl_out=None, almB_out=None):
        """EB-leakage cleaning on a masked sky (SHT.py:192-202 ->
        src/sht_main.f90:801-941): returns the cleaned B map (npix, nsim).

        As in the reference, Q and U are masked in place (when they are
        Fortran-ordered float64 NumPy arrays or GPU tensors; other inputs are
        converted and the masked copies are discarded).  The fit uses the
        pixels where mask == 1.  Optional outputs: Btpl_out receives the
        leakage template map, almB_out the B alm of the masked maps."""
        self._need_pol()
        lib = _lib.load()
        dev = _is_torch(Q)
        qp, loc, qk = self._arg(Q, self._mshape, "Q", out=dev or (isinstance(Q, np.ndarray) and Q.flags.f_contiguous
                                                                  and Q.dtype == np.float64))
        up, uloc, uk = self._arg(U, self._mshape, "U", out=dev or (isinstance(U, np.ndarray) and U.flags.f_contiguous
                                                                   and U.dtype == np.float64))
        if len({loc, uloc}) != 1:
            raise ValueError("all arrays must be host arrays or all GPU tensors")
        for name, arr in (("Btpl_out", Btpl_out), ("almB_out", almB_out)):
            if arr is not None and _is_torch(arr) != dev:
                raise ValueError(f"{name} must be {'a GPU tensor' if dev else 'a host array'} like Q and U")
        # the mask goes where Q is: a host mask (NumPy or CPU tensor) is copied
        # to Q's GPU, a GPU mask is brought to the host for host Q
        if dev:
            torch = _torch()
            mk = torch.as_tensor(mask, dtype=torch.float64, device=Q.device).reshape(-1).contiguous()
            if mk.numel() != self.npix:
                raise ValueError(f"mask must have npix={self.npix} entries")
            mp = ctypes.c_void_p(mk.data_ptr())
        else:
            host_mask = mask.detach().cpu().numpy() if _is_torch(mask) else mask
            mk = np.ascontiguousarray(np.asarray(host_mask, dtype=np.float64).reshape(-1))
            if mk.size != self.npix:
                raise ValueError(f"mask must have npix={self.npix} entries")
            mp = ctypes.c_void_p(mk.ctypes.data)
        Bmap = self._like(Q, "map")
        bp, bloc, bk = self._arg(Bmap, self._mshape, "Bmap", out=True)
        tp = ap = None
        for name, arr, shape in (("Btpl_out", Btpl_out, self._mshape), ("almB_out", almB_out, self._ashape)):
            if arr is None:
                continue
            ptr = self._arg(arr, shape, name, out=True)[0]
            if name == "Btpl_out":
                tp = ptr
            else:
                ap = ptr
        self._call(lib.fsht_fix_eb, qp, up, mp, bp, ap, tp, None, self.niter, loc)
        return Bmap

    def convert_alm_healpy(self, alms):
        """Square packing -> healpy complex (nalm, nsim) (SHT.py:204-210,
        convert_alm.f90:34-56), computed on the GPU (host arrays are staged
        through the device by the library)."""
        lib = _lib.load()
        L = self.lmax + 1
        nalm = L * (L + 1) // 2
        if not _is_torch(alms):
            a = np.asfortranarray(alms, dtype=np.float64)
            if a.shape != self._ashape:
                raise ValueError(f"alms: expected shape {self._ashape}, got {a.shape}")
            out = np.empty((2, nalm, self.nsim), order="F")
            _lib.check(lib.fsht_convert_alm_healpy(a.ctypes.data, out.ctypes.data, self.lmax, self.nsim,
                                                   _lib.FSHT_HOST, None))
            return out[0] + 1j * out[1]
        torch = _torch()
        if tuple(alms.stride()) != _fortran_strides(self._ashape):
            raise ValueError("alms must be Fortran-order (nsim, lmax+1, lmax+1)")
        out = torch.empty((self.nsim, nalm, 2), dtype=torch.float64, device=alms.device)
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.fsht_convert_alm_healpy(ctypes.c_void_p(alms.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                               self.lmax, self.nsim, _lib.FSHT_DEVICE, ctypes.c_void_p(s)))
        return torch.complex(out[..., 0], out[..., 1]).t()  # (nalm, nsim)

    def get_alms(self, alms_in=None):
        """Fetch the alms kept on the GPU by t2alm(get_alm=False) (SHT.py:214-227)."""
        lib = _lib.load()
        alms = empty_alms(self.lmax, self.nsim) if alms_in is None else alms_in
        ap, loc, ak = self._arg(alms, self._ashape, "alms", out=True)
        self._call(lib.fsht_get_alm, ap, loc)
        return alms if alms_in is None else 0

    # ---- extras -----------------------------------------------------------
    def ring_tab(self):
        """Ring-skip table (first kept north ring per m) used by this handle."""
        out = np.zeros(self.lmax + 1, dtype=np.int32)
        self._call(_lib.load().fsht_ring_tab, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def synchronize(self):
        self._call(_lib.load().fsht_synchronize)
