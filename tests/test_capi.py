"""The C ABI library loads on the host and exports every declared symbol.

No compute calls here: only argument validation that returns before any
device work."""

import ctypes

import pytest

from lattice_boltzmann_parallel_solver_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the header
    assert sorted(_lib._SIGNATURES) == names


def test_version():
    assert _lib.load().lbm_version() == 1


def test_argument_validation_without_device():
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.lbm_create(ctypes.byref(h), 0, 5, 0, 0) == _lib.LBM_ERR_ARG
    assert "lattice" in L.lbm_last_error().decode()
    assert L.lbm_create(ctypes.byref(h), 4, 5, 7, 0) == _lib.LBM_ERR_ARG
    assert L.lbm_stream(None, None, 4, 4, 0, None) == _lib.LBM_ERR_ARG
    assert L.lbm_equilibrium(None, None, None, 4, 5, None) == _lib.LBM_ERR_ARG
    assert L.lbm_collide(None, None, None, 2.5, None, 4, 0, None) == _lib.LBM_ERR_ARG
    assert L.lbm_run(None, 1.0, 1) == _lib.LBM_ERR_ARG


def test_check_raises_with_message():
    L = _lib.load()
    h = ctypes.c_void_p()
    with pytest.raises(_lib.LbmError, match="lattice"):
        _lib.check(L.lbm_create(ctypes.byref(h), -1, 5, 0, 0))


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """With the HIP library absent the package raises at first use; there is
    no CPU fallback for any entry point."""
    import lattice_boltzmann_parallel_solver_amd._lib as lib

    monkeypatch.setattr(lib, "_lib", None)
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "liblbm_hip.so"))
    with pytest.raises(RuntimeError, match="missing"):
        lib.load()
    from lattice_boltzmann_parallel_solver_amd import Lattice

    with pytest.raises(RuntimeError, match="missing"):
        Lattice(8, 8)
