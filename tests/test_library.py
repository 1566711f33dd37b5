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


def test_struct_layouts_match_the_header(tmp_path):
    """Every ABI struct the binding mirrors has the header's size and field offsets
    (compiled with the host C compiler; the header needs only the C library)."""
    import shutil
    import subprocess

    import pytest

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = {"swe_grid": _lib.SweGrid, "swe_config": _lib.SweConfigC, "swe_forcing": _lib.SweForcing,
             "swe_control": _lib.SweControl, "swe_dynamic_level": _lib.SweDynamicLevel,
             "swe_scenario": _lib.SweScenario}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "swe2d_hip.h"', "int main(void) {"]
    for c_name, py in pairs.items():
        lines.append(f'  printf("{c_name} size %zu\\n", sizeof({c_name}));')
        for field, _ in py._fields_:
            if field.startswith("pad"):
                continue
            lines.append(f'  printf("{c_name} {field} %zu\\n", offsetof({c_name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        c_name, key, value = line.split()
        got[(c_name, key)] = int(value)
    for c_name, py in pairs.items():
        assert got[(c_name, "size")] == ctypes.sizeof(py), c_name
        for field, _ in py._fields_:
            if not field.startswith("pad"):
                assert got[(c_name, field)] == getattr(py, field).offset, (c_name, field)
