"""ctypes binding of libxpg_integrate.so (include/xpg_integrate.h, include/xpg_decode.h, include/xpg_cbf.h,
include/xpg_synth.h, include/xpg_forward.h, include/xpg_finalise.h, include/xpg_refine.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc, gfx950).  There is no fallback: when the library
is missing or a HIP call fails, the calls raise.  Device memory and streams come from torch tensors; the structs
below carry their raw pointers.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "libxpg_integrate.so"
RC_K = 21


class XpgError(RuntimeError):
    pass


class Detector(C.Structure):
    _fields_ = [("n_seg", C.c_int32), ("P0", C.c_void_p), ("Bx", C.c_void_p), ("By", C.c_void_p),
                ("U1", C.c_void_p), ("U2", C.c_void_p), ("UN", C.c_void_p),
                ("orgx", C.c_double), ("orgy", C.c_double), ("qx", C.c_double), ("qy", C.c_double),
                ("wl", C.c_double), ("tsens", C.c_double), ("mu", C.c_double), ("plx", C.c_double),
                ("segmap", C.c_void_p), ("seg_nx", C.c_int32), ("seg_ny", C.c_int32),
                ("gx", C.c_void_p), ("gy", C.c_void_p), ("geo_nx", C.c_int32), ("geo_ny", C.c_int32),
                ("bkg", C.c_void_p), ("bkg_nx", C.c_int32), ("bkg_ny", C.c_int32)]


class Call(C.Structure):
    _fields_ = [("axis", C.c_double * 3), ("s0", C.c_double * 3), ("frames_phi0", C.c_double), ("osc", C.c_double),
                ("sd", C.c_double), ("sm", C.c_double), ("gain", C.c_double), ("rcw", C.c_float * RC_K),
                ("use_rcw", C.c_int32), ("widths", C.c_void_p)]


class Widths(C.Structure):
    _fields_ = [("sd", C.c_double), ("sm", C.c_double), ("rcw", C.c_float * RC_K), ("n_strong", C.c_int32),
                ("active", C.c_uint32), ("sd_fit", C.c_double), ("sm_fit", C.c_double)]


class Refl(C.Structure):
    _fields_ = [("n", C.c_int32), ("Dz", C.c_int32), ("Dy", C.c_int32), ("Dx", C.c_int32),
                ("S", C.c_void_p), ("phi", C.c_void_p), ("ix0", C.c_void_p), ("iy0", C.c_void_p),
                ("iz0", C.c_void_p), ("fb", C.c_void_p), ("pxs", C.c_void_p), ("region", C.c_void_p),
                ("has_p", C.c_void_p), ("pS", C.c_void_p), ("pphi", C.c_void_p), ("K", C.c_int32),
                ("nbx", C.c_void_p), ("nby", C.c_void_p), ("nbz", C.c_void_p), ("nbsx", C.c_void_p),
                ("nbsz", C.c_void_p), ("rdz", C.c_void_p), ("rdy", C.c_void_p), ("rdx", C.c_void_p),
                ("voff", C.c_void_p), ("poff", C.c_void_p), ("wave_lds", C.c_int32), ("wg_waves", C.c_int32),
                ("lds_rows", C.c_int32)]


class Pass1Out(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("bg", "nbg", "isum", "vsum", "vsum0", "xo", "yo", "zo", "m_d", "m_m",
                                          "frac_rec", "nfg", "zeta", "seff", "n_over", "n_fg_masked", "n_fg_nb",
                                          "fz")] + [("fz_stride", C.c_int32)] + \
               [(k, C.c_void_p) for k in ("img2", "J", "u1", "u2", "all_valid2", "vmask")]


class Pass2Out(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("iprf", "vprf", "vprf0", "corr", "p_valid", "p_total", "ipartner", "rho",
                                          "p_fg", "vmask")]


class ProfileAcc(C.Structure):
    _fields_ = [("num", C.c_void_p), ("den", C.c_void_p), ("n_regions", C.c_int32), ("copies", C.c_int32)]


class Frames(C.Structure):
    _fields_ = [("data", C.c_void_p), ("frame_stride", C.c_int64), ("nx", C.c_int32), ("ny", C.c_int32),
                ("frame0", C.c_int32), ("n_frames", C.c_int32), ("available", C.c_void_p),
                ("src_mask", C.c_void_p), ("ceil_mask", C.c_void_p), ("min_valid", C.c_uint32),
                ("flag_value", C.c_uint32), ("overload", C.c_uint32), ("use_flag_value", C.c_int32),
                ("dtype", C.c_int32), ("ring", C.c_int32)]


class Prepared(C.Structure):
    _fields_ = [("sb", C.c_void_p), ("geo", C.c_void_p), ("zeta", C.c_void_p), ("nbm", C.c_void_p),
                ("geo_p", C.c_void_p), ("zeta_p", C.c_void_p), ("seff", C.c_void_p), ("R", C.c_void_p),
                ("r_stride", C.c_int32)]



class Bslz4Batch(C.Structure):
    _fields_ = [("src", C.c_void_p), ("chunk_off", C.c_void_p), ("chunk_len", C.c_void_p), ("n_frames", C.c_int32),
                ("esize", C.c_int32), ("nel", C.c_int64), ("block_elems", C.c_int32), ("n_blocks", C.c_int32),
                ("table", C.c_void_p), ("status", C.c_void_p), ("out", C.c_void_p), ("out_kind", C.c_int32),
                ("out_stride", C.c_int64), ("out_slot", C.c_void_p), ("use_min_valid", C.c_int32),
                ("use_overload", C.c_int32), ("min_valid_ceil", C.c_int64), ("overload_ceil", C.c_int64),
                ("static_bad", C.c_void_p)]



class CbfBatch(C.Structure):
    _fields_ = [("src", C.c_void_p), ("stream_off", C.c_void_p), ("stream_len", C.c_void_p), ("n_elem", C.c_void_p),
                ("chunk0", C.c_void_p), ("group0", C.c_void_p), ("n_frames", C.c_int32), ("n_chunks", C.c_int64),
                ("n_groups", C.c_int64), ("work", C.c_void_p), ("out", C.c_void_p), ("out_off", C.c_void_p),
                ("status", C.c_void_p)]


class FmSupports(C.Structure):
    _fields_ = [("R", C.c_void_p), ("P", C.c_void_p), ("iz0", C.c_void_p), ("iy0", C.c_void_p), ("ix0", C.c_void_p),
                ("T", C.c_int32), ("n", C.c_int32), ("Dz", C.c_int32), ("b", C.c_int32)]


class FmWindow(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("f0", "nf", "y0", "ny", "x0", "nx")]


class FmRows(C.Structure):
    _fields_ = [("S", C.c_void_p), ("phi", C.c_void_p), ("ix0", C.c_void_p), ("iy0", C.c_void_p),
                ("iz0", C.c_void_p), ("region", C.c_void_p), ("sd", C.c_void_p), ("sm", C.c_void_p),
                ("n", C.c_int32), ("Dz", C.c_int32), ("b", C.c_int32), ("pxs", C.c_int32)]


class FmModel(C.Structure):
    _fields_ = [("axis", C.c_double * 3), ("s0", C.c_double * 3), ("frames_phi0", C.c_double), ("osc", C.c_double),
                ("sd", C.c_double), ("sm", C.c_double), ("prof", C.c_void_p), ("n_regions", C.c_int32),
                ("G", C.c_int32), ("umax", C.c_double), ("halo", C.c_int32), ("halo_w", C.c_double),
                ("halo_kappa", C.c_double), ("rcw", C.c_float * RC_K), ("use_rcw", C.c_int32)]


class LpGeom(C.Structure):
    _fields_ = [("e0", C.c_double * 3), ("axis", C.c_double * 3), ("pn", C.c_double * 3), ("pm", C.c_double * 3),
                ("pfrac", C.c_double), ("ed3", C.c_double * 3), ("distance", C.c_double), ("tsens", C.c_double),
                ("mu", C.c_double), ("mu_air", C.c_double), ("segmented", C.c_int32)]


class FinalIn(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("S", "x", "y", "iprf", "vprf", "isum", "vsum", "p_valid", "p_total", "corr",
                                          "coverage")]


class FinalOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("rlp", "I", "sig", "Isum", "sigsum", "peak", "corr", "good")]


class RfSpots(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("h", "phi", "branch", "xo", "yo", "seg")] + [("n", C.c_int64)]


class RfSetup(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("qx", "qy", "sxy", "sphi", "tsens", "mu", "plx", "robust_c")] + \
        [(k, C.c_int32) for k in ("n_seg", "n_models")]


class RfJob(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("spot0", "model0", "r0", "t0", "w0")] + [("n_spots", C.c_int32),
                                                                                 ("n_models", C.c_int32)]


RF_HEAD, RF_SEG = 17, 9

P = C.POINTER
_SIGS = {
    "xpg_version": ([], C.c_char_p),
    "xpg_last_error": ([], C.c_char_p),
    "xpg_extract": ([P(Frames), P(Refl), C.c_void_p, C.c_void_p], C.c_int),
    "xpg_decode_bins": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "xpg_pass1": ([P(Detector), P(Call), P(Refl), C.c_void_p, P(Pass1Out), P(ProfileAcc), C.c_void_p], C.c_int),
    "xpg_pass1_frames": ([P(Detector), P(Call), P(Refl), P(Frames), P(Pass1Out), P(ProfileAcc), C.c_void_p],
                         C.c_int),
    "xpg_accumulate_profiles": ([C.c_void_p] * 9 + [C.c_int32] * 3 + [P(ProfileAcc), C.c_void_p], C.c_int),
    "xpg_normalise_profiles": ([P(ProfileAcc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "xpg_pass2": ([P(Detector), P(Call), P(Refl), C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                   C.c_int32, P(Pass2Out), C.c_void_p], C.c_int),
    "xpg_pass2_frames": ([P(Detector), P(Call), P(Refl), P(Frames), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                          C.c_int32, P(Pass2Out), C.c_void_p], C.c_int),
    "xpg_prepare": ([P(Detector), P(Call), P(Refl), P(Frames), P(Prepared), C.c_void_p], C.c_int),
    "xpg_prepare_sb": ([P(Detector), P(Call), P(Refl), C.c_void_p, P(Prepared), C.c_void_p], C.c_int),
    "xpg_rocking": ([P(Call), P(Refl), P(Prepared), C.c_void_p], C.c_int),
    "xpg_pass1_prepared": ([P(Detector), P(Call), P(Refl), P(Prepared), P(Pass1Out), P(ProfileAcc), C.c_void_p],
                           C.c_int),
    "xpg_pass2_prepared": ([P(Detector), P(Call), P(Refl), P(Prepared), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                            C.c_int32, P(Pass2Out), C.c_void_p], C.c_int),
    "xpg_lds_bytes": ([C.c_int32] * 3, C.c_int),
    "xpg_lds_limit": ([], C.c_int),
    "xpg_copy_rows": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p], C.c_int),
    "xpg_adapt_block": ([C.c_int32] + [C.c_void_p] * 12 + [C.c_double] * 5 + [C.c_void_p] * 4 + [C.c_int64,
                                                                                         C.c_void_p], C.c_int),
    "xpg_adapt_workspace_bytes": ([C.c_int32], C.c_int64),
    "xpg_synth_background": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "xpg_synth_spots_add": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "xpg_bslz4_block_elems": ([C.c_uint32, C.c_int32], C.c_int32),
    "xpg_bslz4_n_blocks": ([C.c_int64, C.c_int32], C.c_int32),
    "xpg_bslz4_index_host": ([C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_void_p], C.c_int32),
    "xpg_bslz4_index": ([P(Bslz4Batch), C.c_void_p], C.c_int),
    "xpg_bslz4_decode": ([P(Bslz4Batch), C.c_void_p], C.c_int),
    "xpg_fm_scatter": ([C.c_void_p, P(FmSupports), C.c_void_p, C.c_void_p, P(FmWindow), C.c_void_p], C.c_int),
    "xpg_fm_gather": ([C.c_void_p, C.c_void_p, P(FmSupports), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_int32, C.c_int32, P(FmWindow), C.c_void_p], C.c_int),
    "xpg_cbf_workspace_bytes": ([C.c_int64, C.c_int64], C.c_int64),
    "xpg_cbf_decode": ([P(CbfBatch), C.c_void_p], C.c_int),
    "xpg_cbf_decode_static": ([P(CbfBatch), C.c_void_p], C.c_int),
    "xpg_cbf_encode_host": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_int64], C.c_int64),
    "xpg_fm_terms": ([P(Detector), P(FmModel), P(FmRows), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "xpg_finalise": ([P(LpGeom), P(Detector), C.c_int64, P(FinalIn), P(FinalOut), C.c_void_p], C.c_int),
    "xpg_format_integrate_rows": ([C.c_void_p, C.c_int64, C.c_int64] + [C.c_void_p] * 12 + [C.c_double] * 2
                                  + [C.c_void_p] * 2, C.c_int64),
    "xpg_rf_residuals": ([P(RfSpots), P(RfSetup), C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                          C.c_void_p], C.c_int),
    "xpg_rf_normal": ([P(RfSpots), P(RfSetup), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p], C.c_int),
    "xpg_bslz4_encode_host": ([C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64],
                              C.c_int64),
}
EXPORTS = tuple(_SIGS)

_lib = None


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise XpgError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
        h = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def check(rc):
    if rc != 0:
        raise XpgError(f"xpg kernel call failed ({rc}): {lib().xpg_last_error().decode()}")


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise XpgError("kernel arguments must be contiguous")
    return C.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(t):
    if not (torch.is_tensor(t) and t.is_cuda):
        raise XpgError("the integration kernels run on the GPU: device tensors expected")
