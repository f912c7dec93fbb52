"""Bitshuffle-LZ4 frames (HDF5 filter 32008, the Eiger / NeXus chunk format) decoded on the device.

The reference splits the decode: host worker processes undo the LZ4 stage (``lz4_stage``, xpg/io/bslz4.py:34-63,
``_worker_bitshuffled``, xpg/io/images.py:466-477) and the device undoes the bit shuffle and applies the value codes
(``unshuffle_frame_torch``, xpg/io/bslz4.py:80-108, called per frame from xpg/spotfind.py:243-248).  Here the
compressed chunks are copied to HBM as they are read and one kernel does both stages (include/xpg_decode.h), writing
the frame store the integrator reads (raw uint32 counts, :class:`xpg_amd.frames.DeviceFrames`) or value-coded
float32 frames identical to ``ImageSet.get`` / ``unshuffle_frame_torch``.

Errors: a malformed chunk raises ``ValueError`` naming the frames and the reason (``lz4_stage`` raises
``ValueError`` for a chunk of the wrong size and python-lz4's error for a corrupt block); frames are never read or
written out of bounds.
"""
from __future__ import annotations

import ctypes as C
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib as L

BSLZ4_FILTER_ID = 32008
STATUS_TEXT = {1: "chunk shorter than its header", 2: "uncompressed size is not the frame's",
               4: "block size differs from the batch's", 8: "chunk truncated", 16: "compressed size beyond the LZ4 bound",
               32: "corrupt LZ4 block"}


def available() -> bool:
    """The device decoder is built (the HIP library loads)."""
    try:
        L.lib()
        return True
    except L.XpgError:
        return False


def read_header(raw) -> tuple[int, int]:
    """(uncompressed bytes, block size field in bytes) of a chunk (xpg/io/bslz4.py:39-40)."""
    mv = memoryview(raw)
    if len(mv) < 12:
        raise ValueError("not a bitshuffle-LZ4 chunk: shorter than its header")
    return int.from_bytes(mv[0:8], "big"), int.from_bytes(mv[8:12], "big")


def block_elems(block_bytes: int, esize: int) -> int:
    """Elements per bitshuffle block for a header's block size field (0 = 8192 bytes), as lz4_stage computes it."""
    return int(L.lib().xpg_bslz4_block_elems(int(block_bytes), int(esize)))


def _status_error(st: np.ndarray, frames) -> ValueError:
    bad = np.flatnonzero(st)
    what = []
    for i in bad[:8]:
        bits = [t for b, t in STATUS_TEXT.items() if st[i] & b]
        what.append(f"frame {frames[i]}: {', '.join(bits)}")
    more = f" (+{bad.size - 8} more)" if bad.size > 8 else ""
    return ValueError("not a decodable bitshuffle-LZ4 chunk: " + "; ".join(what) + more)


def output_frames(shape, nel: int, n_chunks: int, slotted: bool):
    """(elements between output frames, output frames) of a decode output of ``shape``: a 2-D or higher output holds
    shape[0] frames of prod(shape[1:]) >= nel elements; a 1-D output one frame, of at least nel elements.  Raises
    ValueError for an output that cannot hold the frames (checked before any launch: the kernel writes nel elements per
    frame at slot * stride)."""
    numel = int(np.prod(shape)) if len(shape) else 1
    if len(shape) > 1:
        stride, nslot = int(np.prod(shape[1:])), int(shape[0])
    else:
        stride, nslot = nel, 1
    if stride < nel:
        raise ValueError("output frames smaller than the chunks' frames")
    if numel < (nslot - 1) * stride + nel:
        raise ValueError(f"output of {numel} elements cannot hold a frame of {nel}")
    if not slotted and n_chunks > nslot:
        raise ValueError("more chunks than output frames")
    return stride, nslot


class ChunkBatch:
    """Compressed chunks of frames of ``nel`` elements of ``esize`` bytes, resident on the device with their block
    table.  ``index="host"`` walks each chunk's block chain on the host while it is in host memory (one pass over 4
    bytes per 8 KB block); ``index="device"`` walks them on the device, one lane per frame."""

    def __init__(self, chunks, nel: int, esize: int, device, index: str = "host", frames=None):
        if esize not in (2, 4):
            raise ValueError("esize must be 2 or 4")
        if not chunks:
            raise ValueError("no chunks")
        self.nel, self.esize, self.device = int(nel), int(esize), torch.device(device)
        self.n = len(chunks)
        self.frames = list(frames) if frames is not None else list(range(self.n))
        _, bb = read_header(chunks[0])
        self.block_elems = block_elems(bb, esize)
        if self.block_elems < 8:
            raise ValueError("not a bitshuffle-LZ4 chunk: block of fewer than 8 elements")
        self.n_blocks = int(L.lib().xpg_bslz4_n_blocks(self.nel, self.block_elems))
        lens = np.array([len(c) for c in chunks], np.int64)
        offs = np.zeros(self.n, np.int64)
        offs[1:] = np.cumsum((lens[:-1] + 15) // 16 * 16)          # chunk starts on 16-byte boundaries
        total = int(offs[-1] + lens[-1])
        host = torch.empty(max(total, 1), dtype=torch.uint8, pin_memory=True)
        hv = host.numpy()
        for c, o, n in zip(chunks, offs, lens):
            hv[o:o + n] = np.frombuffer(memoryview(c), np.uint8)
        self.src = host.to(self.device, non_blocking=True)
        self.chunk_off = torch.as_tensor(offs, device=self.device)
        self.chunk_len = torch.as_tensor(lens, device=self.device)
        self.table = torch.zeros((self.n, self.n_blocks + 1, 2), dtype=torch.int32, device=self.device)
        if index == "host":
            tab = np.zeros((self.n, self.n_blocks + 1, 2), np.uint32)
            st = np.zeros(self.n, np.int32)
            h = L.lib()
            for i in range(self.n):
                st[i] = h.xpg_bslz4_index_host(hv[offs[i]:].ctypes.data_as(C.c_void_p), int(lens[i]), self.nel,
                                               self.esize, self.block_elems, tab[i].ctypes.data_as(C.c_void_p))
            self.table.copy_(torch.from_numpy(tab.view(np.int32)))
            self.status = torch.as_tensor(st, device=self.device)
        elif index == "device":
            self.status = torch.zeros(self.n, dtype=torch.int32, device=self.device)
            L.check(L.lib().xpg_bslz4_index(C.byref(self._batch()), L.stream_ptr(self.device)))
        else:
            raise ValueError("index must be 'host' or 'device'")
        self._host = host          # the pinned staging buffer lives until the upload has completed
        # decodes may run on another stream than the uploads above: they wait for this event first
        self._uploaded = torch.cuda.Event()
        self._uploaded.record(torch.cuda.current_stream(self.device))

    @property
    def compressed_bytes(self) -> int:
        return int(self.src.numel())

    def _batch(self, out=None, kind=0, stride=0, slot=None, codes=None) -> L.Bslz4Batch:
        b = L.Bslz4Batch()
        b.src, b.chunk_off, b.chunk_len = L.ptr(self.src), L.ptr(self.chunk_off), L.ptr(self.chunk_len)
        b.n_frames, b.esize, b.nel = self.n, self.esize, self.nel
        b.block_elems, b.n_blocks = self.block_elems, self.n_blocks
        b.table, b.status = L.ptr(self.table), L.ptr(self.status)
        b.out, b.out_kind, b.out_stride, b.out_slot = L.ptr(out), kind, stride, L.ptr(slot)
        if codes is not None:
            (b.use_min_valid, b.min_valid_ceil, b.use_overload, b.overload_ceil, sb) = codes
            b.static_bad = L.ptr(sb)
        return b

    def decode_into(self, out: torch.Tensor, kind: str = "raw", slots=None, overload=None, min_valid=0.0,
                    static_bad=None):
        """Decode every frame into ``out`` ((frames, ...) with >= nel elements per frame): ``kind="raw"`` uint32
        counts, ``kind="coded"`` float32 with unshuffle_frame_torch's codes.  ``slots`` (device int64, one per frame)
        picks the output frame of each chunk.  Asynchronous on the current stream; :meth:`check` raises for malformed
        chunks."""
        L.require_device(out)
        k = {"raw": 0, "coded": 1}[kind]
        if out.dtype != (torch.float32 if k else torch.int32) or not out.is_contiguous():
            raise ValueError("output: contiguous int32 (raw uint32 counts) or float32 (coded) frames")
        stride, nslot = output_frames(tuple(out.shape), self.nel, self.n, slots is not None)
        codes = None
        if k == 1:
            sb = None
            if static_bad is not None:
                sb = torch.as_tensor(np.asarray(static_bad, np.uint8).reshape(-1) if not torch.is_tensor(static_bad)
                                     else static_bad.reshape(-1).to(torch.uint8), device=self.device).contiguous()
                if sb.numel() != self.nel:
                    raise ValueError("static_bad must hold one flag per pixel")
            codes = (int(min_valid > 0), math.ceil(min_valid) if min_valid > 0 else 0, int(overload is not None),
                     math.ceil(overload) if overload is not None else 0, sb)
            self._sb = sb
        if slots is not None:
            slots = torch.as_tensor(slots, dtype=torch.int64, device=self.device).contiguous()
            if slots.numel() != self.n:
                raise ValueError("one slot per chunk")
            lo, hi = int(slots.min()), int(slots.max())
            if lo < 0 or hi >= nslot:
                raise ValueError("slot outside the output")
        self._slots = slots
        b = self._batch(out, k, stride, slots, codes)
        torch.cuda.current_stream(self.device).wait_event(self._uploaded)
        L.check(L.lib().xpg_bslz4_decode(C.byref(b), L.stream_ptr(self.device)))
        return out

    def check(self):
        """Raise ValueError for frames whose chunk was malformed (synchronises the stream)."""
        st = self.status.cpu().numpy()
        if st.any():
            raise _status_error(st, self.frames)


def decode_frame(chunk, nel: int, esize: int, overload=None, min_valid=0.0, static_bad=None, device="cuda"):
    """One chunk to a float32 frame (nel,) on the device with ImageSet.get's values for unsigned counts: the LZ4 stage
    of lz4_stage and the unshuffle and codes of unshuffle_frame_torch (xpg/io/bslz4.py:34-108) in one kernel."""
    cb = ChunkBatch([chunk], nel, esize, device)
    out = torch.empty((1, nel), dtype=torch.float32, device=cb.device)
    cb.decode_into(out, "coded", overload=overload, min_valid=min_valid, static_bad=static_bad)
    cb.check()
    return out[0]


def frames_from_chunks(chunks, ny: int, nx: int, esize: int, device, first: int = 1, overload=None, min_valid=0,
                       static_mask=None, index: str = "host"):
    """A device frame store (raw uint32 counts, held in an int32 tensor as the synthetic frames are) decoded from one chunk per frame, for the integrator: the counts'
    codes (all ones = flagged, min_valid, overload, static mask) stay with the store, as for raw Eiger frames."""
    from .frames import DeviceFrames
    cb = ChunkBatch(chunks, ny * nx, esize, device, index=index, frames=range(first, first + len(chunks)))
    data = torch.empty((len(chunks), ny, nx), dtype=torch.int32, device=cb.device)
    cb.decode_into(data, "raw")
    cb.check()
    return DeviceFrames(data, first=first, overload=overload, min_valid=math.ceil(min_valid),
                        flag_value=(1 << (8 * esize)) - 1, static_mask=static_mask)


# ---- writer (workloads and tests): bit shuffle + LZ4 per block, through the system's liblz4 ------------------------

_lz4 = None


def _lz4_compress_ptr():
    global _lz4
    if _lz4 is None:
        try:
            h = C.CDLL("liblz4.so.1")
        except OSError as e:
            raise RuntimeError("writing bitshuffle-LZ4 chunks needs the LZ4 library (liblz4.so.1)") from e
        _lz4 = (h, C.cast(h.LZ4_compress_default, C.c_void_p))
    return _lz4[1]


def encode_frame(arr: np.ndarray, block_bytes: int = 0) -> bytes:
    """One frame (uint16 or uint32 array) as a bitshuffle-LZ4 chunk (block size field block_bytes, 0 = 8192)."""
    a = np.ascontiguousarray(arr)
    if a.dtype not in (np.uint16, np.uint32, np.int32, np.int16):
        raise ValueError("frames are 16- or 32-bit counts")
    esize, nel = a.dtype.itemsize, a.size
    bb = block_bytes or 8192
    be = max(block_elems(bb, esize), 8)
    nb = -(-nel // be)
    cap = 12 + nb * (4 + bb + bb // 255 + 16) + 8 * esize
    out = np.empty(cap, np.uint8)
    n = L.lib().xpg_bslz4_encode_host(a.ctypes.data_as(C.c_void_p), nel, esize, int(block_bytes), _lz4_compress_ptr(),
                                      out.ctypes.data_as(C.c_void_p), cap)
    if n < 0:
        raise RuntimeError("bitshuffle-LZ4 encode failed")
    return out[:n].tobytes()


def encode_frames(frames, block_bytes: int = 0, threads: int = 16) -> list[bytes]:
    """encode_frame over frames (an iterable of arrays) on a thread pool (the C encoder releases the GIL)."""
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(lambda f: encode_frame(f, block_bytes), frames))
