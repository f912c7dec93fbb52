"""Oracle (test infrastructure only): ctypes wrapper of oracle/codec.c, the CPU restatement of the bitshuffle-LZ4
frame decode (xpg/io/bslz4.py:28-108) and of the CBF byte-offset decode (xpg/io/cbf.py:12-44).  Built by __graft_entry__.build() into oracle/libxpg_oracle_codec.so; only the
tests, the smoke check and bench.py's CPU-baseline leg use it, as the checker."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "libxpg_oracle_codec.so"
_h = None


def lib():
    global _h
    if _h is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: __graft_entry__.build() compiles it")
        h = C.CDLL(str(LIB))
        h.lz4_block_decode.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
        h.lz4_block_decode.restype = C.c_int
        h.bitunshuffle.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_void_p]
        h.bitunshuffle.restype = None
        h.bslz4_decode_chunk.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        h.bslz4_decode_chunk.restype = C.c_int64
        h.bslz4_codes.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int64, C.c_void_p,
                                  C.c_void_p]
        h.bslz4_codes.restype = None
        h.cbf_decode.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
        h.cbf_decode.restype = C.c_int
        _h = h
    return _h


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lz4_decode(block: bytes, n: int):
    """(code, bytes): code n on success, -1 malformed, -2 offset-0 match."""
    src = np.frombuffer(bytes(block), np.uint8).copy() if len(block) else np.zeros(1, np.uint8)
    out = np.zeros(max(n, 1), np.uint8)
    r = lib().lz4_block_decode(_p(src), len(block), _p(out), n)
    return r, out[:n].tobytes()


def bitunshuffle(buf: np.ndarray, n: int, esize: int) -> np.ndarray:
    buf = np.ascontiguousarray(buf, np.uint8)
    out = np.zeros(n * esize, np.uint8)
    lib().bitunshuffle(_p(buf), n, esize, _p(out))
    return out


def decode_chunk(chunk: bytes, nel: int, esize: int):
    """(block_elems or negative code, element bytes (nel * esize) as uint8)."""
    src = np.frombuffer(bytes(chunk), np.uint8).copy() if len(chunk) else np.zeros(1, np.uint8)
    out = np.zeros(nel * esize, np.uint8)
    scratch = np.zeros(1 << 17, np.uint8)
    r = lib().bslz4_decode_chunk(_p(src), len(chunk), nel, esize, _p(out), _p(scratch), scratch.size)
    return int(r), out


def codes(raw: np.ndarray, esize: int, overload=None, min_valid=0.0, static_bad=None) -> np.ndarray:
    """float32 frame from element bytes, as unshuffle_frame_torch (xpg/io/bslz4.py:100-108)."""
    raw = np.ascontiguousarray(raw, np.uint8)
    nel = raw.size // esize
    sb = None if static_bad is None else np.ascontiguousarray(np.asarray(static_bad).reshape(-1), np.uint8)
    out = np.zeros(nel, np.float32)
    lib().bslz4_codes(_p(raw), nel, esize, int(min_valid > 0), math.ceil(min_valid) if min_valid > 0 else 0,
                      int(overload is not None), math.ceil(overload) if overload is not None else 0,
                      None if sb is None else _p(sb), _p(out))
    return out


def cbf_decode(buf: bytes, n: int):
    """(status, int32 values): decode_byte_offset (xpg/io/cbf.py:12-44) restated in C; status 1 = stream too short."""
    src = np.frombuffer(bytes(buf), np.uint8).copy() if len(buf) else np.zeros(1, np.uint8)
    out = np.zeros(max(n, 1), np.int32)
    r = lib().cbf_decode(_p(src), len(buf), n, _p(out))
    return int(r), out[:n]
