"""The C-ABI library loads and exports every function include/hoefem.h declares (CPU only:
no compute call is made without a device)."""

import ctypes
import os
import re

import pytest

from higher_order_embedded_fem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hoefem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hoefem_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhoefem.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == names


def test_host_only_entry_points():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libhoefem.so not built")
    L = _lib.load()
    assert [L.hoefem_nodes_per_element(t) for t in range(6)] == [4, 10, 20, 8, 20, 27]
    assert L.hoefem_nodes_per_element(17) == -1
    mu, lam = ctypes.c_double(), ctypes.c_double()
    L.hoefem_lame_from_young_poisson(1e6, 0.4, ctypes.byref(mu), ctypes.byref(lam))
    from higher_order_embedded_fem_amd.scenes import lame_from_young_poisson

    assert (mu.value, lam.value) == lame_from_young_poisson(1e6, 0.4)
    # invalid descriptors are rejected with a status, not a crash
    h = ctypes.c_void_p()
    assert L.hoefem_model_create(None, ctypes.byref(h)) == 1
    assert b"null" in L.hoefem_last_error()
    assert L.hoefem_model_ndof(None) == 0
    # batches: no sims, a null sim, a null batch
    assert L.hoefem_batch_create(None, 0, ctypes.byref(h)) == 1
    assert L.hoefem_batch_create((ctypes.c_void_p * 1)(None), 1, ctypes.byref(h)) == 1
    assert b"null sim" in L.hoefem_last_error()
    assert L.hoefem_batch_step(None, None, None, None, None, None, None) == 1
    L.hoefem_batch_destroy(None)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.HoefemLibraryMissing):
        saved = _lib._lib
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "nope.so"))
        finally:
            _lib._lib = saved
