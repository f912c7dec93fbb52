"""CPU oracle of shoebox extraction -- TEST INFRASTRUCTURE ONLY.

Restates, with the reference's own mechanics, how value-coded shoeboxes are gathered from frames:

  frame_codes      xpg/io/images.py:80-96     raw counts -> float32 codes (min_valid / flagged / static mask -> -1,
                                              overload -> -2, -1 written last so a masked overload is -1)
  prepared_frame   xpg/integrate.py:648-702   padded frame with a -1 border, the integrator's mask as a ceiling
                                              (min(value, -1): a masked overload keeps -2), s x s super-pixel frames
  gather           xpg/integrate.py:704-803   per frame: rows with iz0 <= f < iz0 + B*fb, slot (f - iz0) // fb,
                                              window V[iy, ix] of the padded frame; unbinned rows store value + 3,
                                              binned rows add IEEE-encoded values (NaN overload, +inf masked) and count
  decode_bins      xpg/shoebox.py:264-268 with xpg/integrate.py:484-488 (bins missing a frame -> unrecorded)

Pinning: decode_bins against the reference's golden vector (tests/test_robustness.py:44-54); gather against a
brute-force per-voxel definition on small frames (tests/test_extract_oracle.py).
"""
from __future__ import annotations

import numpy as np

SB_OFF = np.float32(3.0)


def frame_codes(raw, min_valid=0, flag_value=None, overload=None, static_mask=None):
    raw = np.asarray(raw)
    img = raw.astype(np.float32)
    bad = raw < min_valid
    if flag_value is not None:
        bad |= raw == flag_value
    if static_mask is not None:
        bad |= ~np.asarray(static_mask, bool)
    if overload is not None:
        np.copyto(img, np.float32(-2.0), where=raw >= overload)
    np.copyto(img, np.float32(-1.0), where=bad)
    return img


def prepared_frame(img, pad, mask=None):
    ny, nx = img.shape
    P = np.full((ny + 2 * pad, nx + 2 * pad), -1.0, np.float32)
    inner = P[pad:pad + ny, pad:pad + nx]
    if mask is None or np.asarray(mask).all():
        np.copyto(inner, img)
    else:
        mval = np.where(np.asarray(mask), np.float32(np.inf), np.float32(-1.0)).astype(np.float32)
        np.minimum(img, mval, out=inner, dtype=np.float32)
    return P


def superpixel_frame(inner, sp, pk):
    ny, nx = inner.shape
    Pk = np.full((ny // sp + 2 * pk, nx // sp + 2 * pk), -1.0, np.float32)
    dst = Pk[pk:pk + ny // sp, pk:pk + nx // sp]
    a_ = inner[:dst.shape[0] * sp, :dst.shape[1] * sp]
    np.copyto(dst, a_[0::sp, 0::sp])
    mn = a_[0::sp, 0::sp].copy()
    for i_ in range(sp):
        for j_ in range(sp):
            if i_ or j_:
                v_ = a_[i_::sp, j_::sp]
                np.add(dst, v_, out=dst)
                np.minimum(mn, v_, out=mn)
    neg = mn < 0
    if neg.any():
        dst[neg] = np.where(mn[neg] <= -1.5, np.float32(-2.0), np.float32(-1.0))
    return Pk


def decode_bins(arr, cnt):
    out = np.nan_to_num(arr, nan=-2.0, posinf=-1.0, neginf=-1.0).astype(np.float32) + SB_OFF
    out[cnt == 0] = 0.0
    return out


def gather(frames, read, ix0, iy0, iz0, fb, pxs, B, bx, mask=None):
    """Shoeboxes (n, B, bx, bx) of one group, as value codes (-3 unrecorded ... counts).

    frames: dict {0-based frame index: coded float32 frame (ny, nx)}; read: iterable of 0-based frames read in order.
    ix0, iy0 in the reflections' super-pixel units (pxs, one value for the group), iz0, fb per reflection."""
    n = len(ix0)
    sp = int(pxs)
    half = bx // 2
    pad = half + 1
    binned = bool((np.asarray(fb) > 1).any())
    arr = np.zeros((n, B, bx, bx), np.float32)
    cnt = np.zeros((n, B), np.int16) if binned else None
    for fi in read:
        img = frames[fi]
        P = prepared_frame(img, pad, mask)
        if sp > 1:
            ny, nx = img.shape
            P = superpixel_frame(P[pad:pad + ny, pad:pad + nx], sp, pad)
        V = np.lib.stride_tricks.sliding_window_view(P, (bx, bx))
        for r in range(n):
            kk = fi - iz0[r]
            if kk < 0 or kk >= B * fb[r]:
                continue
            iy = int(np.clip(iy0[r] + pad, 0, P.shape[0] - bx))
            ix = int(np.clip(ix0[r] + pad, 0, P.shape[1] - bx))
            v = V[iy, ix].copy()
            if not binned:
                arr[r, kk] = v + SB_OFF
            else:
                kk //= fb[r]
                if kk >= B:
                    continue
                neg = v < 0
                v[neg] = np.where(v[neg] == -2.0, np.float32(np.nan), np.float32(np.inf))
                arr[r, kk] += v
                cnt[r, kk] += 1
    if binned:
        arr = decode_bins(arr, cnt)
        arr[cnt < np.asarray(fb)[:, None]] = 0.0
    return arr - SB_OFF
