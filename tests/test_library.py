"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import re
from pathlib import Path

from diffswe2d_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "swe2d_hip.h").read_text()
    return sorted(set(re.findall(r"^\S.*?\b((?:swe|sur)_\w+)\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.exported_symbols())


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    loaded = _lib.load()
    assert loaded.swe_abi_version() == _lib.SWE_ABI_VERSION
    assert loaded.swe_control_bytes() == ctypes.sizeof(_lib.SweControl)


def test_workspace_size_and_argument_checks():
    lib = _lib.load()
    g = _lib.make_grid(2, 64, 96, 1, 1, 0.1, 0.1)
    assert lib.swe_workspace_bytes(g, 4) >= 2 * 3 * 64 * 96 * 4
    bad = _lib.make_grid(1, 1, 96, 1, 1, 0.1, 0.1)
    assert lib.swe_workspace_bytes(bad, 4) == 0
    assert b"ng cannot exceed" in lib.swe_last_error_message()
    assert lib.swe_strerror(-5) == b"The calculated timestep is non-finite."
