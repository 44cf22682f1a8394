"""The C-ABI library loads (no GPU needed) and exports every entry point
that include/sktopt_amd.h declares, with the bindings' signatures."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sktopt_amd.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(skt_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 30
    assert "skt_oc_bisect" in names and "skt_solver_solve" in names


def test_library_exports_every_declared_symbol():
    from sktopt_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.skt_abi_version() == 1


def test_bindings_cover_the_header():
    from sktopt_amd import _lib
    assert sorted(_lib.exported_symbols()) == _declared()


def test_error_mapping():
    from sktopt_amd import _lib
    import pytest
    with pytest.raises(ValueError):
        _lib.check(_lib.SKT_EINVAL)
    with pytest.raises(_lib.ConvergenceError):
        _lib.check(_lib.SKT_ENOCONV)
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.SKT_ENOTIMPL)


def test_argument_validation_without_gpu():
    """Argument checks run before any device work, so they work on CPU."""
    from sktopt_amd import _lib
    lib = _lib.load()
    st = lib.skt_interpolate(7, 10, None, 1.0, 0.0, 3.0, None, 0, None)
    assert st == _lib.SKT_EINVAL and "SIMP or RAMP" in _lib.last_error()
    h = ctypes.c_void_p()
    st = lib.skt_hex8_create(0, 0, 1, 1, 1.0, 1.0, 1.0, 0.3, 2, None, ctypes.byref(h))
    assert st == _lib.SKT_EINVAL
