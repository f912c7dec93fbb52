"""Device frame store: rotation frames resident in HBM, read by the extraction inside the integration kernels.

Replaces the reference's host frame loop and shoebox buffers (xpg/integrate.py:648-803, xpg/shoebox.py:202-256):
instead of copying every frame into a padded host buffer and gathering windows on CPU threads, frames stay on the
device as raw uint32 counts (or value-coded float32 when they come from an ImageSet-like source) and each kernel reads
its shoebox rows directly.  A 4148 x 4362 uint32 frame is 72.4 MB, so a 1800-frame sweep (130 GB) fits in one
MI355X's 288 GB.

The value codes follow xpg/io/images.py:80-96 (min_valid / flagged / static mask -> -1, overload -> -2, masked wins)
and the integrator's mask ceiling of xpg/integrate.py:661-664 (masked pixels min(value, -1), a masked overload keeps
-2).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L


class DeviceFrames:
    """Frames ``first .. last`` (1-based, inclusive) on the device.

    data: torch tensor (n, ny, nx) uint32 (raw counts, dtype 0) or float32 (value-coded, dtype 1).
    overload / min_valid / flag_value: raw-count codes (uint32 data only).  static_mask: bool (ny, nx), True =
    trusted, as ImageSet.static_mask (untrusted -> -1 even when overloaded).
    """

    def __init__(self, data, first=1, overload=None, min_valid=0, flag_value=None, static_mask=None, n_frames=None):
        L.require_device(data)
        if data.dim() != 3:
            raise ValueError("frames must be (n, ny, nx)")
        self.data = data.contiguous()
        self.dtype_code = 1 if self.data.dtype == torch.float32 else 0
        if self.dtype_code == 0 and self.data.element_size() != 4:
            raise ValueError("raw frames must be 32-bit")
        self.device = self.data.device
        self.first = int(first)
        # n_frames > len(data): a ring of len(data) slots holding frames first .. first + n_frames - 1 as they are
        # streamed in (frame first + t in slot t % len(data)); the caller keeps the frames a launch reads resident
        self.ring = int(self.data.shape[0]) if n_frames is not None and n_frames > self.data.shape[0] else 0
        self.n = int(n_frames) if self.ring else int(self.data.shape[0])
        self.last = self.first + self.n - 1
        self.ny, self.nx = int(self.data.shape[1]), int(self.data.shape[2])
        self.overload = overload
        self.min_valid = int(min_valid)
        self.flag_value = flag_value
        self.static_mask = None
        if static_mask is not None:
            self.static_mask = torch.as_tensor(np.asarray(static_mask, np.uint8), device=self.device)

    # --- the frame protocol of xpg/io/images.py (first, last, nx, ny) plus the device store ----------------------

    def store(self, mask=None, exclude_frames=(), frames=None):
        """ctypes view for the kernels.  mask: the integrator's trusted-pixel mask (bool (ny, nx) or None);
        exclude_frames / frames: frames never read (their shoebox slices are unrecorded)."""
        s = L.Frames()
        s.data = L.ptr(self.data)
        s.frame_stride = self.ny * self.nx
        s.nx, s.ny = self.nx, self.ny
        s.frame0, s.n_frames = self.first - 1, self.n
        avail = np.ones(self.n, np.uint8)
        for f in exclude_frames:
            if self.first <= f <= self.last:
                avail[f - self.first] = 0
        if frames is not None:
            lo, hi = frames
            fr = np.arange(self.first, self.last + 1)
            avail[(fr < lo) | (fr > hi)] = 0
        keep = []
        if not avail.all():
            t = torch.as_tensor(avail, device=self.device)
            keep.append(t)
            s.available = L.ptr(t)
        if self.static_mask is not None:
            s.src_mask = L.ptr(self.static_mask)
        if mask is not None and not np.asarray(mask).all():
            t = torch.as_tensor(np.asarray(mask, np.uint8), device=self.device)
            keep.append(t)
            s.ceil_mask = L.ptr(t)
        s.ring = self.ring
        s.min_valid = self.min_valid
        s.use_flag_value = int(self.flag_value is not None)
        s.flag_value = int(self.flag_value or 0)
        s.overload = int(self.overload or 0)
        s.dtype = self.dtype_code
        self._keep = keep                      # device arrays referenced by the struct
        return s

    def nbytes(self):
        return self.data.numel() * self.data.element_size()


def from_imageset(images, frames, device, batch=10):
    """Upload the value-coded float32 frames of an ImageSet-like source (``iter_frames``) into a device store, through
    two pinned host buffers of ``batch`` frames: the host fills one while the other's copy runs."""
    f0, f1 = frames
    n = f1 - f0 + 1
    data = torch.empty((n, images.ny, images.nx), dtype=torch.float32, device=device)
    up = FrameUploader(images, device, batch)
    for k0 in range(0, n, batch):
        up.upload(f0 + k0, min(batch, n - k0), data[k0:k0 + batch])
    torch.cuda.current_stream(device).synchronize()
    return DeviceFrames(data, first=f0)


class FrameUploader:
    """Host frames to the device in batches, through two pinned buffers: filling one waits only for the copy that
    last read it (an event), so reading the next batch on the host overlaps the previous batch's copy."""

    def __init__(self, images, device, batch):
        self.images, self.device, self.batch = images, device, int(batch)
        self.buf = [torch.empty((self.batch, images.ny, images.nx), dtype=torch.float32, pin_memory=True)
                    for _ in range(2)]
        self.done = [None, None]
        self.k = 0

    def upload(self, f0, nf, dst):
        """Frames f0 .. f0 + nf - 1 into ``dst`` (nf, ny, nx) float32 on the device, asynchronously on the current
        stream."""
        i = self.k % 2
        self.k += 1
        if self.done[i] is not None:
            self.done[i].synchronize()
        hb = self.buf[i]
        for j, (f, img) in enumerate(self.images.iter_frames(range(f0, f0 + nf), copy=False)):
            hb[j].copy_(torch.from_numpy(np.ascontiguousarray(img, np.float32)))
        dst.copy_(hb[:nf], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self.done[i] = ev
