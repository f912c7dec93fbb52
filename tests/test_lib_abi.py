"""The C ABI library loads and exports every entry point include/*.h declares, and the ctypes structs of
xpg_amd/_lib.py have the C layout (checked against gcc on the header).  No GPU needed."""
import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
STRUCTS = dict(Detector="xpg_detector", Call="xpg_call", Refl="xpg_refl", Pass1Out="xpg_pass1_out",
               Pass2Out="xpg_pass2_out", ProfileAcc="xpg_profile_acc", Frames="xpg_frames", Widths="xpg_widths",
               Prepared="xpg_prepared", Bslz4Batch="xpg_bslz4_batch", CbfBatch="xpg_cbf_batch",
               FmSupports="xpg_fm_supports", FmWindow="xpg_fm_window", LpGeom="xpg_lp_geom", FinalIn="xpg_final_in",
               FinalOut="xpg_final_out", FmRows="xpg_fm_rows", FmModel="xpg_fm_model", RfSpots="xpg_rf_spots",
               RfSetup="xpg_rf_setup", RfJob="xpg_rf_job")


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(xpg_\w+)\s*\(", txt, flags=re.M))
    return names


def _prototypes():
    """{name: [parameter kinds]} of every function the headers declare ('ptr', 'f64', 'i32', 'i64', 'u32')."""
    protos = {}
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(xpg_\w+)\s*\(([^)]*)\)\s*;", txt, flags=re.M):
            params = [q.strip() for q in m.group(2).split(",") if q.strip() and q.strip() != "void"]
            kinds = []
            for q in params:
                if "*" in q:
                    kinds.append("ptr")
                else:
                    kinds.append({"double": "f64", "int32_t": "i32", "int64_t": "i64", "uint32_t": "u32",
                                  "int": "i32", "xpg_lz4_compress_fn": "ptr"}[q.split()[0]])
            protos[m.group(1)] = kinds
    return protos


def test_ctypes_signatures_match_prototypes():
    import ctypes as C
    from xpg_amd import _lib

    def kind(t):
        if t in (C.c_void_p, C.c_char_p) or isinstance(t, type(C.POINTER(C.c_int))):
            return "ptr"
        return {C.c_double: "f64", C.c_int32: "i32", C.c_int64: "i64", C.c_uint32: "u32"}[t]
    protos = _prototypes()
    assert set(protos) == set(_lib.EXPORTS)
    for name, (args, _) in _lib._SIGS.items():
        assert [kind(t) for t in args] == protos[name], name


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
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "xpg_integrate.h"', '#include "xpg_decode.h"', '#include "xpg_cbf.h"',
             '#include "xpg_forward.h"', '#include "xpg_finalise.h"', '#include "xpg_refine.h"', "int main(void) {"]
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


def test_static_scratch_decode_refuses_oversized_batches_before_any_launch():
    """xpg_cbf_decode_static checks the batch's workspace against its fixed scratch before it touches a device (so this
    runs without a GPU: the pointers are never dereferenced)."""
    import ctypes as C
    from xpg_amd import _lib
    h = _lib.lib()
    b = _lib.CbfBatch()
    for k in ("src", "stream_off", "stream_len", "n_elem", "chunk0", "group0", "out", "out_off", "status"):
        setattr(b, k, C.c_void_p(0x1000))
    b.n_frames = 1
    limit = 64 << 20
    nc = 1
    while h.xpg_cbf_workspace_bytes(nc * 2, (nc * 2 + 63) // 64) <= limit:
        nc *= 2
    lo, hi = nc, nc * 2                       # the largest chunk count that fits lies in [lo, hi)
    while hi - lo > 1:
        mid = (lo + hi) // 2
        lo, hi = (mid, hi) if h.xpg_cbf_workspace_bytes(mid, (mid + 63) // 64) <= limit else (lo, mid)
    assert 400_000 < lo < 450_000
    b.n_chunks, b.n_groups = hi, (hi + 63) // 64
    assert h.xpg_cbf_decode_static(C.byref(b), None) == -3              # XPG_ERR_TOO_LARGE
    assert b"static scratch" in h.xpg_last_error()
    b.n_chunks = -1
    assert h.xpg_cbf_decode_static(C.byref(b), None) == -1              # XPG_ERR_ARG
