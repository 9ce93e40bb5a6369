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


def test_ctypes_mirrors_match_the_header_layout(tmp_path):
    """Every ctypes structure has the size and field offsets of the C struct
    it mirrors (compiled from include/svg_amd.h with the host C compiler)."""
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no host C compiler")
    pairs = [(_lib.BuildParams, "svg_build_params"), (_lib.BuildStats, "svg_build_stats"),
             (_lib.ServerCfg, "svg_server_cfg"), (_lib.Request, "svg_request"),
             (_lib.ServerStats, "svg_server_stats")]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svg_amd.h"', "int main(void) {"]
    for cls, cname in pairs:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                          text=True).stdout.split("\n") if l)
    for cls, cname in pairs:
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)
