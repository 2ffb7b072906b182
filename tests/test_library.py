"""The C-ABI library loads on a machine without a GPU and exports every symbol that
include/diver.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'diver.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(diver_[a-z0-9_]+)\s*\(', src)))


def test_header_and_binding_agree():
    from diver_amd import _lib
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_library_exports_declared_symbols():
    from diver_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libdiver.so not built')
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_version_and_validation_without_gpu():
    from diver_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libdiver.so not built')
    L = _lib.lib()
    assert b'gfx950' in L.diver_version()
    # argument validation happens before any device work
    st = L.diver_masked_intersect(-1, 0, 0.0, 0.0, 1.0, 1.0, 1.0, None, None, None, None, None,
                                  None, None)
    assert st == 1 and b'batch' in L.diver_last_error()
    st = L.diver_aabb_intersect(4, None, None, None, None, None, None, 100, None)
    assert st == 1 and b'multiple of 64' in L.diver_last_error()
    with pytest.raises(_lib.DiverError):
        _lib.call('diver_integrate', None, 5, 6, 3, None, None, None, None)


def test_adam_validation_without_gpu():
    from diver_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libdiver.so not built')
    L = _lib.lib()
    args = (0.9, 0.999, 1e-8, 0.0, 1e-3, 1.0, 0, None)
    assert L.diver_adam_step(None, None, None, None, 0, *args) == 0  # nothing to do
    st = L.diver_adam_step(None, 16, 16, 16, 4, *args)
    assert st == 1 and b'null' in L.diver_last_error()
    st = L.diver_adam_step(16, 18, 16, 16, 4, *args)
    assert st == 1 and b'aligned' in L.diver_last_error()
    st = L.diver_adam_step(16, 16, 16, 16, 4, 1.5, *args[1:])
    assert st == 1 and b'betas' in L.diver_last_error()


def test_decoder_update_range_validation_without_gpu():
    from diver_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libdiver.so not built')
    L = _lib.lib()
    assert L.diver_decoder_update_range(None, None, 0, 0, None) == 1
    assert b'null' in L.diver_last_error()


def test_decoder_validation_without_gpu():
    import ctypes as C
    import numpy as np
    from diver_amd import _lib
    from diver_amd.decoder import CanonicalDecoder
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libdiver.so not built')
    canon = CanonicalDecoder(48, 32, 0, False, 1, 1, np.zeros(10, np.float32))
    d = canon.desc()
    h = C.c_void_p()
    p = canon.params
    st = _lib.lib().diver_decoder_create(C.byref(d), p.ctypes.data_as(C.POINTER(C.c_float)),
                                         p.size, C.byref(h))
    assert st == 4  # DIVER_EUNSUPPORTED
