"""The C ABI library loads and exports every entry point include/xpg_integrate.h declares, and the ctypes structs of
xpg_amd/_lib.py have the C layout (checked against gcc on the header).  No GPU needed."""
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
STRUCTS = dict(Detector="xpg_detector", Call="xpg_call", Refl="xpg_refl", Pass1Out="xpg_pass1_out",
               Pass2Out="xpg_pass2_out", ProfileAcc="xpg_profile_acc", Frames="xpg_frames", Widths="xpg_widths")


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(xpg_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_declared_symbols():
    from xpg_amd import _lib
    h = _lib.lib()
    declared = _declared()
    assert {"xpg_pass1", "xpg_pass2", "xpg_extract", "xpg_accumulate_profiles", "xpg_normalise_profiles",
            "xpg_pass1_frames", "xpg_pass2_frames", "xpg_decode_bins"} <= declared
    for name in declared:
        assert hasattr(h, name), name
    assert declared <= set(_lib.EXPORTS)
    assert h.xpg_version().decode().startswith("xpg_amd")
    assert h.xpg_lds_bytes(7, 11, 11) > 7 * 11 * 11 * 5


def test_ctypes_structs_match_c_layout(tmp_path):
    import ctypes
    from xpg_amd import _lib
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "xpg_integrate.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in getattr(_lib, py)._fields_:
            lines.append(f'printf("{py} {f} %zu\\n", offsetof({c}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        py, f, v = line.split()
        st = getattr(_lib, py)
        got = ctypes.sizeof(st) if f == "size" else getattr(st, f).offset
        assert got == int(v), (py, f, got, v)
        seen += 1
    assert seen > 60
