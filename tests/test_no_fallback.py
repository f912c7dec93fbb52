"""The product path has no CPU fallback: host tensors and a missing library raise (CPU only)."""
import numpy as np
import pytest
import torch


def test_kernels_refuse_host_tensors():
    from xpg_amd._lib import XpgError
    from xpg_amd.integrate_kernels import make_kernels
    p1, acc, p2 = make_kernels(7, 11, 11, 9)
    sb = torch.zeros((2, 7, 11, 11))
    with pytest.raises(XpgError):
        p1(sb, {}, {}, np.zeros(3), np.zeros(3), 0.0, 0.001, 0.001, 0.001, 1.0, None, None, None)
    with pytest.raises(XpgError):
        acc(torch.zeros((2, 11, 11)), None, None, None, None, None, None, None, None)


def test_missing_library_raises(monkeypatch, tmp_path):
    from xpg_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(_lib.XpgError):
        _lib.lib()


def test_integrate_needs_a_gpu(monkeypatch):
    from types import SimpleNamespace
    from xpg_amd import integrate as I, synth
    from xpg_amd._lib import XpgError
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    spec = synth.spec_small()
    exp = synth.make_experiment(spec)
    img = SimpleNamespace(first=1, last=100, nx=1024, ny=1024)
    model = I.IntegrationModel(0.02, 0.1, box_xy=9, box_z=5)
    with pytest.raises(XpgError):
        I.integrate(img, [exp, exp], synth.block_edges(spec), model, 3.8, verbose=False)


def test_frame_decoders_have_no_cpu_path():
    """The bitshuffle-LZ4 and CBF decoders run on the device only: without a GPU (or with host tensors) they raise."""
    from xpg_amd import bslz4, cbf
    from xpg_amd._lib import XpgError
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    chunk = bslz4.encode_frame(np.arange(64, dtype=np.uint32))
    with pytest.raises((XpgError, RuntimeError)):
        bslz4.decode_frame(chunk, 64, 4, device="cpu")
    with pytest.raises((XpgError, RuntimeError)):
        cbf.decode_byte_offset(cbf.encode_byte_offset(np.arange(10, dtype=np.int32)), 10, device="cpu")
