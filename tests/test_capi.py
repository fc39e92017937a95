"""CPU tests of the boundary: the C-ABI library loads, exports every symbol the
header declares, and rejects bad arguments on the host side (no GPU work)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "fastsht.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fsht_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    from fastsht_amd import _lib
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from fastsht_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.fsht_version()


def test_create_rejects_bad_arguments():
    from fastsht_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    # (2048, 0, 40000): one order's folded spectrum would pass 4 GiB (the
    # Legendre kernels address it with 32-bit byte offsets)
    for args in [(3, 8, 1, 0, 0), (4096, 100, 1, 0, 0), (8, 33, 1, 0, 0), (8, 8, 0, 0, 0), (8, 8, 1, -1, 0),
                 (2048, 0, 40000, 0, 0)]:
        rc = lib.fsht_create(*args, ctypes.byref(h))
        assert rc == _lib.FSHT_ERR_ARG, args
        assert h.value is None
        assert lib.fsht_last_error(None)


def test_null_handle_calls_fail_cleanly():
    from fastsht_amd import _lib
    lib = _lib.load()
    assert lib.fsht_t2alm(None, None, None, 0, 0) == _lib.FSHT_ERR_ARG
    assert lib.fsht_synchronize(None) == _lib.FSHT_ERR_ARG
    assert lib.fsht_destroy(None) == _lib.FSHT_OK
    assert lib.fsht_norm_alm(None, None, None, None, 1.0, 4, 1, None) == _lib.FSHT_ERR_ARG


def test_sht_api_surface():
    """The drop-in object keeps SHT.py's method names and defaults."""
    import inspect

    from fastsht_amd import SHT
    sig = inspect.signature(SHT.__init__)
    pos = [n for n, p in sig.parameters.items() if p.kind != inspect.Parameter.KEYWORD_ONLY]
    assert pos[1:] == ["nside", "lmax", "nsim", "niter", "pol", "all_buff"]
    # extras are keyword-only, so every positional call of the reference still binds the same way
    assert [n for n, p in sig.parameters.items() if p.kind == inspect.Parameter.KEYWORD_ONLY] == ["precision"]
    assert sig.parameters["niter"].default == 0 and sig.parameters["pol"].default is False
    for name in ["set_maps", "set_alms", "t2alm", "alm2t", "qu2eb", "qu2e", "qu2b", "eb2qu", "fix_eb",
                 "convert_alm_healpy", "get_alms"]:
        assert callable(getattr(SHT, name)), name
    t2 = inspect.signature(SHT.t2alm)
    assert list(t2.parameters)[1:] == ["maps_in", "alms_in", "method", "get_alm"]
    assert t2.parameters["method"].default == "old" and t2.parameters["get_alm"].default is True


def test_no_cpu_fallback_without_gpu():
    """Without a GPU the product raises; it never computes on the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from fastsht_amd import SHT, FastSHTError
    with pytest.raises(FastSHTError):
        SHT(4, 8, 2)


def test_product_does_not_import_oracle():
    """No product source (every file of the package, the C header and the
    benchmark's timed legs) names the oracle; bench.py imports it only inside
    its CPU-baseline functions."""
    pkg = os.path.join(ROOT, "fastsht_amd")
    files = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    for dirpath, _, names in os.walk(pkg):
        files += [os.path.join(dirpath, f) for f in names if f.endswith((".py", ".hip", ".h", ".cpp"))]
    for path in files:
        assert "oracle" not in open(path).read().lower(), path
    src = open(os.path.join(ROOT, "bench.py")).read()
    imports = [ln for ln in src.splitlines() if "import" in ln and "oracle" in ln]
    assert imports and all(ln.startswith("    ") for ln in imports), imports  # function-local only
    for fn in re.findall(r"^def (\w+)", src, flags=re.M):
        body = src.split(f"def {fn}(", 1)[1].split("\ndef ", 1)[0]
        if "from oracle" in body:
            assert fn in ("cpu_c0", "cpu_sample"), fn


def test_fix_eb_rejects_outputs_on_the_other_side():
    """fix_eb with host Q/U and a torch Btpl_out / almB_out (or the reverse)
    raises before any transfer or kernel; checked without a GPU on an object
    that has no device handle."""
    torch = pytest.importorskip("torch")
    from fastsht_amd import SHT
    S = object.__new__(SHT)
    S.nside, S.lmax, S.nsim, S.niter, S.pol = 2, 4, 1, 0, 1
    S.npix, S.nring = 48, 7
    S._h, S._stream = None, None
    Q = np.zeros((48, 1), order="F")
    U = np.zeros((48, 1), order="F")
    with pytest.raises(ValueError, match="Btpl_out must be a host array"):
        S.fix_eb(Q, U, np.ones(48), Btpl_out=torch.zeros((1, 48), dtype=torch.float64).t())
    with pytest.raises(ValueError, match="almB_out must be a host array"):
        S.fix_eb(Q, U, np.ones(48), almB_out=torch.zeros((5, 5, 1), dtype=torch.float64).permute(2, 1, 0))
