"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU calls)."""
import ctypes
import os
import re

from svg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(svg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"svg_build", "svg_sssd", "svg_msad", "svg_server_start"} <= names
    assert names == set(_lib.EXPORTS), sorted(names ^ set(_lib.EXPORTS))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_errors_without_device_are_status_codes():
    L = _lib.lib()
    # argument errors are reported before any device work
    assert L.svg_build(None, None, None) == -1
    assert b"null" in L.svg_last_error()
