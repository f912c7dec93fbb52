"""CBF byte-offset frames (x-CBF_BYTE_OFFSET, PILATUS / miniCBF) decoded on the device.

The reference decodes every CBF frame on the host (``decode_byte_offset``, xpg/io/cbf.py:12-44, from ``read_cbf``,
:74-98, for ``CBFImageSet``, xpg/io/images.py:213-235).  Here the frames' binary sections are copied to HBM as read and
one batch of five kernels (include/xpg_cbf.h) rebuilds the int32 pixel values of every frame: ``read_cbf``'s ``data``
(``decode_byte_offset(buf, n).astype(np.int32)``), written straight into device memory.

Errors: a stream with fewer than n elements raises ``ValueError`` naming the frames.  The reference raises
``IndexError`` when a stream runs out before its n-th element or inside an int16 escape; when the stream ends inside
the int32 or int64 payload of its n-th element, the reference reads the bytes that are there as a shorter signed
integer and returns a value.  Here that case is an error too (deliberately stricter: such a stream is cut, and the
value it would give is not the pixel's).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

CHUNK, GROUP = 256, 64


class StreamBatch:
    """The byte-offset streams of several frames (``n_elem[i]`` pixels each), resident on the device with the
    decode's workspace."""

    def __init__(self, streams, n_elem, device, frames=None):
        if len(streams) == 0:
            raise ValueError("no streams")
        self.n = len(streams)
        self.device = torch.device(device)
        self.frames = list(frames) if frames is not None else list(range(self.n))
        self.n_elem = np.asarray(n_elem, np.int64).reshape(-1)
        if self.n_elem.size == 1 and self.n > 1:
            self.n_elem = np.full(self.n, int(self.n_elem[0]), np.int64)
        if self.n_elem.size != self.n or (self.n_elem < 0).any():
            raise ValueError("one non-negative element count per stream")
        lens = np.array([len(s) for s in streams], np.int64)
        offs = np.zeros(self.n, np.int64)
        offs[1:] = np.cumsum((lens[:-1] + 15) // 16 * 16)
        total = int(offs[-1] + lens[-1])
        host = torch.empty(max(total, 1), dtype=torch.uint8, pin_memory=True)
        hv = host.numpy()
        for s, o, n in zip(streams, offs, lens):
            hv[o:o + n] = np.frombuffer(memoryview(s), np.uint8)
        chunks = (lens + CHUNK - 1) // CHUNK
        groups = (chunks + GROUP - 1) // GROUP
        chunk0 = np.concatenate([[0], np.cumsum(chunks)]).astype(np.int64)
        group0 = np.concatenate([[0], np.cumsum(groups)]).astype(np.int64)
        self.n_chunks, self.n_groups = int(chunk0[-1]), int(group0[-1])
        dev = self.device
        self.src = host.to(dev, non_blocking=True)
        self._host = host          # the pinned staging buffer lives until the upload has completed
        self.stream_off = torch.as_tensor(offs, device=dev)
        self.stream_len = torch.as_tensor(lens, device=dev)
        self.n_elem_d = torch.as_tensor(self.n_elem, device=dev)
        self.chunk0 = torch.as_tensor(chunk0, device=dev)
        self.group0 = torch.as_tensor(group0, device=dev)
        self.work_bytes = int(L.lib().xpg_cbf_workspace_bytes(self.n_chunks, self.n_groups))
        self.work = None               # allocated by the first decode that uses a caller workspace
        self.status = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.stream_bytes = total
        # decodes may run on another stream than the uploads above: they wait for this event first
        self._uploaded = torch.cuda.Event()
        self._uploaded.record(torch.cuda.current_stream(dev))

    def decode_into(self, out: torch.Tensor, out_off=None, static_scratch: bool = False):
        """Decode every frame into ``out`` (contiguous int32): frame i at element ``out_off[i]`` (default: frames back
        to back).  Asynchronous on the current stream; :meth:`check` raises for short streams.  ``static_scratch``:
        through xpg_cbf_decode_static (the library's per-GPU scratch array instead of a workspace allocated here;
        batches up to XPG_CBF_STATIC_SCRATCH_BYTES of workspace, calls on one GPU must not overlap)."""
        L.require_device(out)
        if out.dtype != torch.int32 or not out.is_contiguous():
            raise ValueError("output: contiguous int32")
        if out_off is None:
            out_off = np.concatenate([[0], np.cumsum(self.n_elem)[:-1]]).astype(np.int64)
        out_off = np.asarray(out_off, np.int64)
        if out_off.size != self.n or (out_off < 0).any() or (out_off + self.n_elem > out.numel()).any():
            raise ValueError("output offsets outside the output")
        self._out_off = torch.as_tensor(out_off, device=self.device)
        b = L.CbfBatch()
        b.src, b.stream_off, b.stream_len = L.ptr(self.src), L.ptr(self.stream_off), L.ptr(self.stream_len)
        b.n_elem, b.chunk0, b.group0 = L.ptr(self.n_elem_d), L.ptr(self.chunk0), L.ptr(self.group0)
        b.n_frames, b.n_chunks, b.n_groups = self.n, self.n_chunks, self.n_groups
        if not static_scratch and self.work is None:
            self.work = torch.empty(max(self.work_bytes, 256), dtype=torch.uint8, device=self.device)
        b.work = None if static_scratch else L.ptr(self.work)
        b.out, b.out_off, b.status = L.ptr(out), L.ptr(self._out_off), L.ptr(self.status)
        torch.cuda.current_stream(self.device).wait_event(self._uploaded)
        decode = L.lib().xpg_cbf_decode_static if static_scratch else L.lib().xpg_cbf_decode
        L.check(decode(C.byref(b), L.stream_ptr(self.device)))
        return out

    def check(self):
        st = self.status.cpu().numpy()
        if st.any():
            bad = [self.frames[i] for i in np.flatnonzero(st)[:8]]
            raise ValueError(f"byte-offset stream shorter than its frame (frames {bad})")


def decode_byte_offset(buf, n: int, device="cuda") -> torch.Tensor:
    """``read_cbf``'s pixel values from one binary section: ``decode_byte_offset(buf, n).astype(np.int32)``
    (xpg/io/cbf.py:12-44, :97) as an int32 device tensor (n,)."""
    sb = StreamBatch([buf], [n], device)
    out = torch.empty(max(n, 1), dtype=torch.int32, device=sb.device)
    sb.decode_into(out)
    sb.check()
    return out[:n]


def encode_byte_offset(data) -> bytes:
    """encode_byte_offset (xpg/io/cbf.py:47-70) of int32 values, in C (the writer side; makes workloads)."""
    a = np.ascontiguousarray(np.asarray(data).reshape(-1), np.int32)
    out = np.empty(max(7 * a.size, 1), np.uint8)
    n = L.lib().xpg_cbf_encode_host(a.ctypes.data_as(C.c_void_p), a.size, out.ctypes.data_as(C.c_void_p), out.size)
    if n < 0:
        raise ValueError("values too large for 32-bit byte offset encoding")
    return out[:n].tobytes()
