"""fastsht_amd: batched HEALPix spherical harmonic transforms on AMD MI355X.

`SHT` is a drop-in for the reference's SHT.py object; the transforms run in
hand-written gfx950 HIP kernels behind the C ABI of include/fastsht.h.
"""
from .sht import SHT, empty_alms, empty_maps  # noqa: F401
from ._lib import FastSHTError  # noqa: F401
from .spectra import alm2cl, cl2alm  # noqa: F401

__version__ = "0.1.0"
