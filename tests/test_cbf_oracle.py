"""CBF byte-offset oracle (oracle/codec.c: cbf_decode) and the product's encoder pinned without a GPU: hand-built
streams of every escape class (int8, int16 after 0x80, int32 after 0x80 0x00 0x80, int64 after a further int32
escape) with the values the format defines, wrap to int32 of the int64 running sum (read_cbf's astype), the
reference test's vector (tests/test_core.py:15-22 of the reference: Poisson(3) 37 x 53 with 1e6, -1, 40000, -2e6),
short and cut streams, and trailing bytes after the n-th element (a CBF binary section's padding)."""
import numpy as np
import pytest

from oracle import codec as oc


def le(v, n):
    return int(v).to_bytes(n, "little", signed=True)


HAND = [
    (bytes([5, 0xFB, 0x7F, 0x81]), [5, 0, 127, 0]),                                  # int8 differences
    (b"\x80" + le(200, 2) + b"\x80" + le(-32767, 2), [200, 200 - 32767]),           # int16
    (b"\x80\x00\x80" + le(100000, 4) + bytes([1]), [100000, 100001]),                 # int32
    (b"\x80\x00\x80" + le(-2 ** 31 + 1, 4), [-2 ** 31 + 1]),
    (b"\x80\x00\x80\x00\x00\x00\x80" + le(2 ** 40 + 7, 8), [7]),                      # int64, wrapped to int32
    (b"\x80\x00\x80\x00\x00\x00\x80" + le(-(2 ** 33), 8) + b"\x80" + le(-300, 2), [0, -300]),
    (bytes([0x80, 0x80, 0xFF]), [-128]),                                             # int16 escape of -128
]


@pytest.mark.parametrize("i", range(len(HAND)))
def test_hand_built_streams(i):
    buf, want = HAND[i]
    st, got = oc.cbf_decode(buf, len(want))
    assert st == 0
    np.testing.assert_array_equal(got, np.array(want, np.int64).astype(np.int32))
    # the stream cut inside its last element, or asked for one element more, is short
    assert oc.cbf_decode(buf[:-1], len(want))[0] == 1
    assert oc.cbf_decode(buf, len(want) + 1)[0] == 1
    # trailing padding after the n-th element is ignored
    assert np.array_equal(oc.cbf_decode(buf + b"\x00" * 4095, len(want))[1], got)


def test_encoder_bytes_and_roundtrip():
    from xpg_amd import cbf
    assert cbf.encode_byte_offset([5, 5, 132, 4]) == bytes([5, 0, 127, 0x80]) + le(-128, 2)
    assert cbf.encode_byte_offset([40000]) == b"\x80\x00\x80" + le(40000, 4)
    assert cbf.encode_byte_offset([-1, -32768]) == bytes([0xFF, 0x80]) + le(-32767, 2)
    assert cbf.encode_byte_offset([-1, -32769]) == bytes([0xFF]) + b"\x80\x00\x80" + le(-32768, 4)
    rng = np.random.default_rng(0)
    a = rng.poisson(3, (37, 53)).astype(np.int32)
    a[3, 4] = 1_000_000
    a[5, 6] = -1
    a[7, 8] = 40_000
    a[9, 10] = -2_000_000
    buf = cbf.encode_byte_offset(a)
    st, got = oc.cbf_decode(buf, a.size)
    assert st == 0 and np.array_equal(got.reshape(a.shape), a)
    big = np.array([2 ** 31 - 1, -2 ** 31 + 1], np.int32)          # a difference beyond int32 is refused
    with pytest.raises(ValueError):
        cbf.encode_byte_offset(big)


def test_stream_cut_inside_the_last_wide_escape_is_short():
    # The n-th element's int32 (7-byte) or int64 (15-byte) escape cut short at the stream's end: the reference reads
    # the bytes that are there as a shorter integer (xpg/io/cbf.py:35-41 slice without a length check); the oracle and
    # the device decode flag the frame as short instead (a cut stream's value is not the pixel's).
    int32_tail = bytes([3]) + b"\x80\x00\x80" + le(100000, 4)
    int64_tail = bytes([3]) + b"\x80\x00\x80\x00\x00\x00\x80" + le(2 ** 40, 8)
    for full in (int32_tail, int64_tail):
        assert oc.cbf_decode(full, 2)[0] == 0
        for cut in range(1, 4):
            assert oc.cbf_decode(full[:-cut], 2)[0] == 1
