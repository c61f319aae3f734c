"""C-ABI surface checks that need no GPU: the library loads, exports every
symbol include/rgsim.h declares, and the Python binding covers them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rgsim.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for s in ("rg_direct_acc", "rg_bh_create", "rg_bh_compute", "rg_energy", "rg_kick",
              "rg_drift", "rg_direct_vjp", "rg_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from rgsim import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librgsim.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    from rgsim import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("librgsim.so not built")
    L = _lib.load()
    assert L.rg_version() == 1
    # argument validation happens before any device work
    st = L.rg_direct_acc(None, None, None, 10, 1.0, 0.01, 0, None)
    assert st == _lib.RG_ERR_ARG
    assert b"null" in L.rg_last_error()
    assert L.rg_set_param(b"direct_ipt", 3) == _lib.RG_ERR_ARG
    assert L.rg_set_param(b"nope", 1) == _lib.RG_ERR_ARG
    assert L.rg_set_param(b"direct_split", 257) == _lib.RG_ERR_ARG
    assert L.rg_set_param(b"direct_split", -1) == _lib.RG_ERR_ARG
    assert L.rg_set_param(b"direct_split", 0) == _lib.RG_OK
    h = ctypes.c_void_p()
    assert L.rg_bh_create(1, 0.5, 0.1, 1.0, 21, 0, ctypes.byref(h)) == _lib.RG_ERR_CAPACITY
    assert L.rg_bh_create(100, -1.0, 0.1, 1.0, 21, 0, ctypes.byref(h)) == _lib.RG_ERR_ARG


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from rgsim.engine import Engine
    from rgsim.ic import make_plummer
    eng = Engine(softening=0.05)
    with pytest.raises(RuntimeError):
        eng.load_state(make_plummer(300, seed=0))
