"""VDIF multiplexing on the GPU (fxg_vdifmux) against the oracle's vdifmux:
the same bytes consumed, the same statistics and byte-identical output
frames; and demultiplexing a station's frames into single-channel threads
and multiplexing them back on the GPU gives the original frames."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check
from fxsynth import make_subint
from muxlib import header, mux_desc, orc_mux, thread_stream

pytestmark = pytest.mark.gpu


def gpu_mux(lib, ctx, src, ids, dest_frames, start=-1, device_src=False, **kw):
    d = mux_desc(ids, **kw)
    fb, nch = C.c_int32(), C.c_int32()
    check(lib, lib.fxg_vdifmux_output_shape(C.byref(d), C.byref(fb), C.byref(nch)), "shape")
    nbytes = fb.value * dest_frames
    dptr = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, max(nbytes, 8), C.byref(dptr)), "alloc")
    sdev = C.c_void_p()
    src = np.ascontiguousarray(src)
    if device_src:
        check(lib, lib.fxg_device_alloc(ctx, src.size, C.byref(sdev)), "alloc")
        check(lib, lib.fxg_memcpy_h2d(ctx, sdev, src.ctypes.data, src.size), "h2d")
    st = fxcorr.VdifMuxStats()
    r = lib.fxg_vdifmux(ctx, C.byref(d), src.ctypes.data, sdev if device_src else None, src.size, dptr, nbytes,
                        start, C.byref(st))
    out = np.zeros(nbytes, np.uint8)
    if r >= 0:
        check(lib, lib.fxg_memcpy_d2h(ctx, out.ctypes.data, dptr, nbytes), "d2h")
    lib.fxg_device_free(ctx, dptr)
    if device_src:
        lib.fxg_device_free(ctx, sdev)
    return r, out, st.as_list()


CASES = {
    "clean_4x2bit": dict(ids=[0, 1, 2, 3], stream=dict(nframes=12), mux=dict(nsort=4, ngap=4)),
    "16_threads": dict(ids=list(range(16)), stream=dict(nframes=6), mux=dict(nsort=16, ngap=16)),
    "3_threads_padded": dict(ids=[5, 9, 2], stream=dict(nframes=10), mux=dict(nsort=6, ngap=6)),
    "8bit": dict(ids=[0, 1], stream=dict(nframes=10, nbits=8), mux=dict(nbits=8, nsort=4, ngap=4)),
    "2chan_threads": dict(ids=[0, 1, 2, 3], stream=dict(nframes=8, nchan=2), mux=dict(nchan=2, nsort=4, ngap=4)),
    "defects": dict(ids=[0, 1, 2, 3],
                    stream=dict(nframes=14, drop={(4, 2)}, dup={(6, 1)}, invalid={(8, 3)}, fill_after={2},
                                junk_after={5}, wrong_thread_after={7}),
                    mux=dict(nsort=4, ngap=4, flags=0x04)),
    "shuffled": dict(ids=[0, 1, 2, 3], stream=dict(nframes=16, shuffle=6), mux=dict(nsort=8, ngap=8)),
    "dest_fills_up": dict(ids=[0, 1, 2, 3], stream=dict(nframes=20), mux=dict(nsort=4, ngap=4), dest=9),
    "gap_ends_call": dict(ids=[0, 1], stream=dict(nframes=30, drop={(f, k) for f in range(10, 20) for k in (0, 1)}),
                          mux=dict(nsort=2, ngap=3), dest=40),
    "start_frame_granularity": dict(ids=[0, 1, 2, 3], stream=dict(nframes=12, first=3, fps=1024),
                                    mux=dict(nsort=4, ngap=4, fps=1024, flags=0x03), start=5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_vdifmux_matches_oracle(gpu, oracle, name):
    lib, ctx = gpu
    c = CASES[name]
    src, _ = thread_stream(c["ids"], framebytes=1032, seed=len(name), **c["stream"])
    mkw = dict(framebytes=1032, **c["mux"])
    ndest = c.get("dest", 24)
    start = c.get("start", -1)
    r0, d0, s0 = orc_mux(oracle, src, c["ids"], ndest, start=start, **mkw)
    for dev in (False, True):
        r1, d1, s1 = gpu_mux(lib, ctx, src, c["ids"], ndest, start=start, device_src=dev, **mkw)
        assert r1 == r0, (name, r1, r0)
        assert s1 == s0, (name, s1, s0)
        used = s0[12]
        assert np.array_equal(d1[:used], d0[:used]), name


def test_mux_of_demuxed_station_is_the_original(gpu):
    """One station's 8-channel 2-bit frames split into 8 one-channel threads
    (thread k = channel k) and multiplexed back on the GPU are the original
    frames, byte for byte."""
    lib, ctx = gpu
    spec = JobSpec(nstations=1, nsubbands=4, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=8)
    sd = make_subint(spec, seed=4)
    raw = sd.stations[0].raw
    fb, nch = spec.framebytes, spec.nrecbands
    frames = raw[: raw.size // fb * fb].reshape(-1, fb)
    nfr = frames.shape[0]
    spf = (fb - 32) * 8 // (2 * nch)                         # samples per frame
    # demultiplex: channel k of every sample, 2 bits each, into thread k's payload
    codes = np.unpackbits(frames[:, 32:], axis=1, bitorder="little").reshape(nfr, spf, nch, 2)
    hdr = frames[:, :32].copy().view(np.uint32)
    fps = int(round(1e6 / spec.sampletime_us)) // spf
    threads = []
    for f in range(nfr):
        sec = int(hdr[f, 0] & 0x3FFFFFFF)
        fn = int(hdr[f, 1] & 0xFFFFFF)
        for k in range(nch):
            pay = np.packbits(codes[f, :, k, :].reshape(-1), bitorder="little")
            threads.append(np.concatenate([header(sec, fn, k, 32 + pay.size, 1, 2, station=0), pay]))
    src = np.concatenate(threads)
    r, out, st = gpu_mux(lib, ctx, src, list(range(nch)), nfr, framebytes=32 + spf * 2 // 8, fps=fps, nbits=2,
                         nsort=nch, ngap=nch, flags=0x01)
    assert st[14] == nfr and st[8] == nfr
    assert np.array_equal(out[: nfr * fb].reshape(nfr, fb), frames)
