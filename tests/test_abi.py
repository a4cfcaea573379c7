"""CPU checks of the C ABI: the library loads, exports every function that
include/afm.h declares, rejects bad arguments before touching a GPU, and its
host build of delta equals the oracle's correctly rounded delta."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "afm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(afm_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(afm):
    from afmibins_amd import _lib
    L = _lib.load()
    names = declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name


def test_python_binding_covers_header(afm):
    from afmibins_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared())


def test_bad_geometry_rejected_without_gpu(afm):
    import ctypes as C
    from afmibins_amd import _lib
    L = _lib.load()
    h = C.c_void_p()
    for g in (_lib.Geom(3, 64, 1, 1), _lib.Geom(2, 48, 1, 1), _lib.Geom(2, 64, 2, 1), _lib.Geom(2, 4, 1, 1)):
        assert L.afm_ctx_create(C.byref(g), 0, C.byref(h)) == -1
        assert L.afm_last_error()


def test_host_delta_equals_oracle(afm, oracle):
    rng = np.random.default_rng(0)
    for n in (8, 64, 2048, 4096):
        r = np.concatenate([rng.uniform(-3.2 / n, 3.2 / n, 200000), np.arange(-13, 14) / (4.0 * n)])
        assert np.array_equal(afm.delta_host(r, n), oracle.delta(r, n, 1))


def test_host_delta_argument_matches_x87(afm, oracle):
    # the integer-exact argument reproduces the reference's long-double
    # sequence: compare delta at points where cos is exactly known
    n = 16
    r = np.array([0.0, 1.0, 2.0, -1.0, -2.0]) / n
    assert np.array_equal(afm.delta_host(r, n), oracle.delta(r, n, 1))
    assert afm.delta_host(np.array([0.0]), n)[0] == 0.5 * n


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from afmibins_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.AfmError):
        _lib.load()
