"""The oracle's VDIF multiplexer (vdifmux.c restated), pinned by the
reference's own corner-turn layouts (cornerturners.c comments), by decoding
the multiplexed frames with the separately pinned VDIF unpacker and
comparing each channel with the thread it came from, and by the counting
rules of vdifmux's statistics on streams with known defects."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_BLANKER_NONE, FXG_FORMAT_VDIF, Datastream
from muxlib import header, orc_mux, thread_stream


@pytest.fixture(scope="module")
def orc():
    return fxcorr.load_oracle()


def test_three_thread_two_bit_layout(orc):
    """cornerturners.c:1417-1431: three 2-bit threads a, b, c become four
    channels, byte n = (0, c_n, b_n, a_n) from the top, the fourth channel zero."""
    ids = [0, 1, 2]
    fb = 32 + 8
    frames = []
    vals = {0: 0b11100100, 1: 0b00011011, 2: 0b10011100}   # samples 0..3 of each thread, LSB first
    for k in ids:
        frames.append(np.concatenate([header(5, 0, k, fb, 1, 2), np.full(8, vals[k], np.uint8)]))
    for k in ids:                                              # a second, complete output frame
        frames.append(np.concatenate([header(5, 1, k, fb, 1, 2), np.full(8, vals[k], np.uint8)]))
    r, dest, st = orc_mux(orc, np.concatenate(frames), ids, 2, framebytes=fb, nsort=1, ngap=1,
                          flags=0x01)
    assert st[8] >= 1
    out = dest[32:32 + 4 * 8]
    for n in range(4):                                         # sample-time n -> output byte n of each word
        a, b, c = [(vals[k] >> (2 * n)) & 3 for k in ids]
        assert out[n] == a | (b << 2) | (c << 4)


def decode(orc, frames, nchan, nbits, framebytes, nsamp):
    ds = Datastream()
    ds.format, ds.nbits, ds.framebytes, ds.headerbytes = FXG_FORMAT_VDIF, nbits, framebytes, 32
    ds.blanker, ds.nrecfreqs, ds.nrecbands = FXG_BLANKER_NONE, 1, nchan
    out = np.zeros(nchan * nsamp, np.float32)
    w = C.c_float()
    fxcorr.check(orc, orc.orc_unpack(C.byref(ds), frames.ctypes.data, frames.size, 0, nsamp, out.ctypes.data,
                                     C.byref(w)), "unpack")
    return out.reshape(nchan, nsamp), w.value


@pytest.mark.parametrize("nthr,nbits,nchan", [(2, 2, 1), (4, 2, 1), (8, 2, 1), (16, 2, 1), (3, 2, 1),
                                              (5, 2, 1), (2, 8, 1), (4, 2, 2), (2, 2, 4)])
def test_channels_decode_to_their_threads(orc, nthr, nbits, nchan):
    """Every channel of the multiplexed stream, decoded by the VDIF unpacker,
    equals the samples of the thread (and thread channel) it came from."""
    ids = list(range(10, 10 + nthr))
    fb = 32 + 256
    src, pay = thread_stream(ids, 6, framebytes=fb, nbits=nbits, nchan=nchan, seed=nthr + nbits)
    r, dest, st = orc_mux(orc, src, ids, 8, framebytes=fb, nbits=nbits, nchan=nchan, nsort=2, ngap=2,
                          flags=0x01)
    units = 1
    while units < nthr:
        units *= 2
    outframe = 256 * units + 32
    nout = st[14]
    assert nout == 6 and st[8] == 6
    spt = 256 * 8 // (nbits * nchan)                          # samples per thread frame
    mux, w = decode(orc, dest[:outframe * nout], units * nchan, nbits, outframe, spt * nout)
    assert w == 1.0
    for k in range(nthr):
        thr = np.concatenate([np.concatenate([header(0, 0, 0, fb, nchan, nbits), pay[(f, k)]]) for f in range(6)])
        ref, _ = decode(orc, thr, nchan, nbits, fb, spt * 6)
        for c in range(nchan):
            assert np.array_equal(mux[k * nchan + c], ref[c]), (k, c)
    for k in range(nthr * nchan, units * nchan):                # padding channels decode as the zero code
        assert np.all(mux[k] == mux[k][0])


def test_statistics_of_a_defective_stream(orc):
    """Duplicates, a fill-pattern frame, junk bytes, an unwanted thread and a
    flagged-invalid frame are counted as vdifmux counts them (vdifmux.c:594-755);
    the frame with a dropped thread comes out invalid."""
    ids = [0, 1, 2, 3]
    fb = 1032
    src, pay = thread_stream(ids, 12, framebytes=fb, drop={(4, 2)}, dup={(6, 1)}, invalid={(8, 3)},
                             fill_after={2}, junk_after={5}, wrong_thread_after={7})
    r, dest, st = orc_mux(orc, src, ids, 16, framebytes=fb, nsort=4, ngap=4, flags=0x05)
    n_valid, n_invalid, n_discard, n_wrong, n_skip, n_fill, n_dup = st[:7]
    assert n_invalid == 1 and n_wrong == 1 and n_dup == 1 and n_fill == fb and n_skip == 12
    outframe = 1000 * 4 + 32
    hdr_words = dest[:outframe * st[14]].reshape(st[14], outframe)[:, :16].copy().view(np.uint32)
    invalid = (hdr_words[:, 0] >> 31).astype(bool)
    assert invalid[4] and invalid[8]                  # thread 2 dropped, thread 3 flagged invalid
    assert not invalid[[0, 1, 2, 3, 5, 6, 7]].any()
    assert (hdr_words[:, 1] & 0xFFFFFF).tolist() == list(range(st[14]))
    assert ((hdr_words[:, 3] >> 16) & 0x3FF).tolist() == [0] * st[14]


def test_out_of_order_within_nsort(orc):
    """Frames permuted within windows shorter than nsort land in order."""
    ids = [0, 1, 2, 3]
    fb = 1032
    a, _ = thread_stream(ids, 10, framebytes=fb, seed=3)
    b, _ = thread_stream(ids, 10, framebytes=fb, seed=3, shuffle=6)
    ra, da, sa = orc_mux(orc, a, ids, 12, framebytes=fb, nsort=8, ngap=8, flags=0x01)
    rb, db, sb = orc_mux(orc, b, ids, 12, framebytes=fb, nsort=8, ngap=8, flags=0x01)
    assert sa[14] == sb[14] == 10
    n = sa[12]
    assert np.array_equal(da[:n], db[:n])
