"""Pins the oracle's unpacker (the checker for the GPU unpack).

Checks: the lookup tables against their closed forms (mark5access
format_vdif.c:74-126, mark5_format_mark5b.c:59-96); decoding against an
independent bit-level reading of the frame layouts on random frames; the
blanking rules (invalid bit, zero frame length, fill pattern, Mark5B header
flag and partial fill); the weight with Mk5Mode's granularity erasure; the
2..4-channel 8-bit decoders' no-advance behaviour; and the reference's own
VDIF sample file (libraries/vdifio/python/examples/vdifio_1.vdif).
"""
import ctypes as C
import os

import numpy as np
import pytest

from fxcorr import FXG_BLANKER_NONE, FXG_FORMAT_MARK5B, FXG_FORMAT_VDIF, Datastream, check

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILL = np.frombuffer(np.array([0x11223344, 0x11223344], dtype=np.uint32).tobytes(), dtype=np.uint8)


def lut(orc, fmt, nbits):
    out = np.zeros(256, dtype=np.float32)
    n = orc.orc_lut(fmt, nbits, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert n > 0
    return out[:n]


def test_lut_closed_forms(oracle):
    h = np.float32(3.3359)
    assert np.array_equal(lut(oracle, FXG_FORMAT_VDIF, 2), np.array([-h, -1, 1, h], dtype=np.float32))
    assert np.array_equal(lut(oracle, FXG_FORMAT_MARK5B, 2), np.array([-h, 1, -1, h], dtype=np.float32))
    b = np.arange(256)
    assert np.array_equal(lut(oracle, FXG_FORMAT_VDIF, 8), ((b - 128) / 3.3).astype(np.float32))
    # 8-bit code 128 is exactly zero, and the table is odd about it
    l8 = lut(oracle, FXG_FORMAT_VDIF, 8)
    assert l8[128] == 0.0 and np.array_equal(l8[129:], -l8[127:0:-1])


def desc(fmt, nchan, nbits, framebytes, headerbytes, blanker=0):
    d = Datastream()
    d.format, d.nbits, d.framebytes, d.headerbytes, d.blanker = fmt, nbits, framebytes, headerbytes, blanker
    d.nrecfreqs, d.nrecbands, d.nzoombands = 1, nchan, 0
    return d


def frames(rng, fmt, nchan, nbits, nframes, fb, hb):
    data = rng.integers(0, 256, size=(nframes, fb), dtype=np.uint8)
    for f in range(nframes):
        if fmt == FXG_FORMAT_VDIF:
            w = np.zeros(8, dtype=np.uint32)
            w[2] = (fb // 8) | (int(np.log2(nchan)) << 24)
            w[3] = (nbits - 1) << 26
            data[f, :32] = w.view(np.uint8)
        else:
            w = np.array([0xABADDEED, f, 0, 0], dtype=np.uint32)
            data[f, :16] = w.view(np.uint8)
    return data


def brute(fmt, data, nchan, nbits, hb, t0, n, table):
    """Bit-level decode: sample (t, c) is field t*nchan + c of the payload
    bitstream, fields LSB first; returns out[nchan][n] for valid frames."""
    fb = data.shape[1]
    fs = (fb - hb) * 8 // (nchan * nbits)
    out = np.zeros((nchan, n), dtype=np.float32)
    for k in range(n):
        t = t0 + k
        f, tin = divmod(t, fs)
        for c in range(nchan):
            bit = (tin * nchan + c) * nbits
            byte = data[f, hb + bit // 8]
            code = (int(byte) >> (bit % 8)) & ((1 << nbits) - 1)
            out[c, k] = table[code]
    return out


def run(orc, d, data, off, nsamp):
    out = np.zeros(d.nrecbands * nsamp, dtype=np.float32)
    w = C.c_float()
    flat = np.ascontiguousarray(data).reshape(-1)
    check(orc, orc.orc_unpack(C.byref(d), flat.ctypes.data, flat.size, off, nsamp, out.ctypes.data, C.byref(w)),
          "orc_unpack")
    return out.reshape(d.nrecbands, nsamp), w.value


CASES = [(FXG_FORMAT_VDIF, 1, 2, 1056, 32), (FXG_FORMAT_VDIF, 2, 2, 1056, 32), (FXG_FORMAT_VDIF, 4, 2, 1056, 32),
         (FXG_FORMAT_VDIF, 16, 2, 1056, 32), (FXG_FORMAT_VDIF, 1, 8, 1056, 32), (FXG_FORMAT_VDIF, 4, 8, 1056, 32),
         (FXG_FORMAT_MARK5B, 4, 2, 10016, 16), (FXG_FORMAT_MARK5B, 16, 2, 10016, 16)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"fmt{c[0]}-{c[1]}ch-{c[2]}bit")
def test_decode_matches_bit_layout(oracle, case):
    fmt, nchan, nbits, fb, hb = case
    rng = np.random.default_rng(nchan + 10 * nbits + fmt)
    data = frames(rng, fmt, nchan, nbits, 6, fb, hb)
    fs = (fb - hb) * 8 // (nchan * nbits)
    table = lut(oracle, fmt, nbits)
    gran = max(1, 8 // (nchan * nbits))
    for off in (0, gran, fs - 2 * gran, fs + 3 * gran, 2 * fs):
        n = min(600, 5 * fs - off)
        n -= n % gran
        got, w = run(oracle, desc(fmt, nchan, nbits, fb, hb), data, off, n)
        assert np.array_equal(got, brute(fmt, data, nchan, nbits, hb, off, n, table))
        assert w == 1.0


def test_vdif_blanking_rules(oracle):
    rng = np.random.default_rng(3)
    fb, hb, nchan = 1056, 32, 4
    fs = (fb - hb) * 8 // (nchan * 2)
    data = frames(rng, FXG_FORMAT_VDIF, nchan, 2, 6, fb, hb)
    data[1, 3] |= 0x80                  # invalid bit (word 0 bit 31)
    data[2, 8:12] = 0                   # zero frame length (word 2)
    data[3, hb:hb + 8] = FILL           # fill pattern at the payload start
    data[4, fb - 8:fb] = FILL           # fill pattern at the payload end
    got, w = run(oracle, desc(FXG_FORMAT_VDIF, nchan, 2, fb, hb), data, 0, 6 * fs)
    for f in range(6):
        seg = got[:, f * fs:(f + 1) * fs]
        assert (np.all(seg == 0) if f in (1, 2, 3, 4) else np.all(seg != 0)), f
    assert w == np.float32(2.0 / 6.0)
    # blanker_none keeps fill-pattern frames but still drops invalid ones
    got, w = run(oracle, desc(FXG_FORMAT_VDIF, nchan, 2, fb, hb, FXG_BLANKER_NONE), data, 0, 6 * fs)
    for f in range(6):
        seg = got[:, f * fs:(f + 1) * fs]
        assert (np.all(seg == 0) if f in (1, 2) else np.all(seg != 0)), f


def test_granularity_weight(oracle):
    """2-channel 2-bit: two sample-times per byte; an odd offset unpacks from
    the even sample before it and erases it, so the weight stays exact."""
    rng = np.random.default_rng(4)
    fb, hb, nchan = 1056, 32, 2
    fs = (fb - hb) * 8 // (nchan * 2)
    data = frames(rng, FXG_FORMAT_VDIF, nchan, 2, 4, fb, hb)
    data[2, 3] |= 0x80
    table = lut(oracle, FXG_FORMAT_VDIF, 2)
    off, n = fs + 101, 2 * fs
    got, w = run(oracle, desc(FXG_FORMAT_VDIF, nchan, 2, fb, hb), data, off, n)
    exp = brute(FXG_FORMAT_VDIF, data, nchan, 2, hb, off, n, table)
    t = np.arange(off, off + n)
    exp[:, (t // fs) == 2] = 0.0
    assert np.array_equal(got, exp)
    assert w == np.float32(np.count_nonzero((t // fs) != 2) / n)


@pytest.mark.parametrize("nchan", [1, 2, 4])
def test_8bit_blank_no_advance(oracle, nchan):
    """format_vdif.c:1879-2020: the 2..4-channel 8-bit decoders do not move on
    after a blank frame, so the rest of the unpack window stays zero."""
    rng = np.random.default_rng(5 + nchan)
    fb, hb = 1056, 32
    fs = (fb - hb) // nchan
    data = frames(rng, FXG_FORMAT_VDIF, nchan, 8, 5, fb, hb)
    data[1, 3] |= 0x80
    table = lut(oracle, FXG_FORMAT_VDIF, 8)
    got, w = run(oracle, desc(FXG_FORMAT_VDIF, nchan, 8, fb, hb), data, fs // 2, 3 * fs)
    exp = brute(FXG_FORMAT_VDIF, data, nchan, 8, hb, fs // 2, 3 * fs, table)
    t = np.arange(fs // 2, fs // 2 + 3 * fs)
    if nchan == 1:
        exp[:, (t // fs) == 1] = 0.0
        good = np.count_nonzero((t // fs) != 1)
    else:
        exp[:, t >= fs] = 0.0
        good = np.count_nonzero(t < fs)
    assert np.array_equal(got, exp)
    assert w == np.float32(good / (3 * fs))


def test_mark5b_zones_and_header_flag(oracle):
    rng = np.random.default_rng(6)
    fb, hb, nchan = 10016, 16, 4
    fs = 10000 * 8 // (nchan * 2)
    data = frames(rng, FXG_FORMAT_MARK5B, nchan, 2, 4, fb, hb)
    data[1, 5] |= 0x80                              # header invalid flag
    k = 300                                         # fill pattern over the payload tail
    data[2, fb - 8 * k:fb] = np.tile(FILL, k)
    data[2, fb - 8 * k - 4:fb - 8 * k] = FILL[:4]   # a 32-bit aligned straggler before it
    table = lut(oracle, FXG_FORMAT_MARK5B, 2)
    got, w = run(oracle, desc(FXG_FORMAT_MARK5B, nchan, 2, fb, hb), data, 0, 4 * fs)
    exp = brute(FXG_FORMAT_MARK5B, data, nchan, 2, hb, 0, 4 * fs, table)
    t = np.arange(4 * fs)
    exp[:, (t // fs) == 1] = 0.0
    # frame 2: bytes from the first fill word (less the straggler) onwards are blank
    first_blank_byte = 10000 - 8 * k - 4
    tin = t - 2 * fs
    exp[:, ((t // fs) == 2) & (tin >= first_blank_byte)] = 0.0   # one byte per sample-time here
    assert np.array_equal(got, exp)
    # blanker_none: only the header flag blanks
    got, w = run(oracle, desc(FXG_FORMAT_MARK5B, nchan, 2, fb, hb, FXG_BLANKER_NONE), data, 0, 4 * fs)
    exp = brute(FXG_FORMAT_MARK5B, data, nchan, 2, hb, 0, 4 * fs, table)
    exp[:, (t // fs) == 1] = 0.0
    assert np.array_equal(got, exp)


def test_reference_vdif_sample_file(oracle):
    raw = np.fromfile(os.path.join(GOLDEN, "vdifio_1.vdif"), dtype=np.uint8)
    ref = np.load(os.path.join(GOLDEN, "vdifio_1_decoded.npz"))
    nchan = int(ref["nchan"])
    fb = 8224
    d = desc(FXG_FORMAT_VDIF, nchan, 2, fb, 32)
    n = raw.size // fb * ((fb - 32) * 8 // (nchan * 2))
    got, w = run(oracle, d, raw.reshape(-1, fb), 0, n)
    assert np.array_equal(got, ref["samples"])
    assert w == 1.0
