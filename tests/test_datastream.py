"""The station data window (include/fxcorr_datastream.h) against the
oracle's restatement of DataStream / Mk5DataStream::calculateControlParams
(oracle/datastream.py), and the window's contract: every block flagged
valid finds all its samples inside the window, at the Mode's own
nearest-sample arithmetic (mode.cpp:686-705)."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_FORMAT_MARK5B, JobSpec, check

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import datastream as ods  # noqa: E402

GUARD_NS = 20000.0

SPECS = {
    "vdif_2bit_16ch": JobSpec(nstations=1, nsubbands=8, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=250),
    "vdif_8bit_4ch": JobSpec(nstations=1, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, nbits=8, blockspersend=400),
    "vdif_2bit_2ch": JobSpec(nstations=1, nsubbands=1, npol=2, bw_mhz=64.0, nchan=4096, blockspersend=40),
    "mark5b_2bit_8ch": JobSpec(nstations=1, nsubbands=4, npol=2, bw_mhz=16.0, nchan=512, blockspersend=100,
                               fmt=FXG_FORMAT_MARK5B, framebytes=10016, headerbytes=16),
}


@pytest.fixture(scope="module")
def host():
    return fxcorr.load_host_lib()


def make(host, spec):
    job, keep = spec.build()
    f = fxcorr.DsFormat()
    check(host, host.fxds_format_of(C.byref(job), 0, GUARD_NS, C.byref(f)), "format")
    h = C.c_void_p()
    check(host, host.fxds_create(C.byref(f), C.byref(h)), "create")
    o = ods.Stream(f.framebytes, f.payloadbytes, f.framespersecond, f.framegranularity, f.bytespersamplenum,
                   f.bytespersampledenom, f.numchannels, f.blockspersend, f.bw_mhz, f.databytes)
    return h, o, f, (job, keep)


def control(host, h, scan, sec, ns, d1, d2, ok=1):
    st = fxcorr.StationInput()
    flags = (C.c_int32 * 64)()
    bi = C.c_int64()
    check(host, host.fxds_control(h, scan, sec, ns, d1, d2, ok, C.byref(st), flags, 64, C.byref(bi)), "control")
    nw = st.nvalidflags
    return st.sec, st.ns, [flags[i] for i in range(nw)], bi.value, st.nbytes


@pytest.mark.parametrize("name", sorted(SPECS))
def test_window_matches_oracle(host, name):
    spec = SPECS[name]
    h, o, f, keep = make(host, spec)
    rng = np.random.default_rng(len(name))
    nframes = 3000
    base = 0x10000000
    host.fxds_set_buffer(h, C.c_void_p(base), nframes * f.framebytes, 0, 100, 0)
    o.set_buffer(nframes * f.framebytes, 0, 100, 0)
    span_s = nframes / f.framespersecond
    subint_s = spec.blockspersend * spec.fftchannels * spec.sampletime_us * 1e-6
    n_ok = n_bad = 0
    for trial in range(400):
        t = rng.uniform(-0.2 * span_s, 1.1 * span_s)
        sec = 100 + int(np.floor(t))
        ns = int((t - np.floor(t)) * 1e9)
        d1 = rng.uniform(-10000.0, 10000.0)
        d2 = d1 + rng.uniform(-3.0, 3.0) * subint_s
        scan = 0 if trial % 37 else 1
        got = control(host, h, scan, sec, ns, d1, d2)
        want = o.control(scan, sec, ns, d1, d2)
        assert got[:4] == (want[0], want[1] if want[0] != ods.INVALID_SUBINT else 0,
                           want[2], want[3]), (trial, t, d1, d2, got, want)
        if got[0] == ods.INVALID_SUBINT:
            n_bad += 1
        else:
            n_ok += 1
            assert got[4] == min(f.databytes, nframes * f.framebytes - got[3])
    assert n_ok > 50 and n_bad > 20
    host.fxds_destroy(h)


@pytest.mark.parametrize("name", sorted(SPECS))
def test_valid_blocks_lie_inside_the_window(host, name):
    """Mode::process reads samples [nearestsample, nearestsample + L) of the
    window for a valid block; they must exist (datastream.cpp:588-621 with
    the guard of getDataBytes)."""
    spec = SPECS[name]
    h, o, f, keep = make(host, spec)
    nframes = 4000
    host.fxds_set_buffer(h, C.c_void_p(0x10000000), nframes * f.framebytes, 0, 100, 0)
    ts = spec.sampletime_us
    L = spec.fftchannels
    spf = f.payloadbytes * f.bytespersampledenom // f.bytespersamplenum     # samples per frame
    subint_us = spec.blockspersend * L * ts
    rng = np.random.default_rng(7)
    checked = 0
    for trial in range(200):
        t_us = rng.uniform(0.05, 0.8) * nframes / f.framespersecond * 1e6
        sec, ns = 100 + int(t_us // 1e6), int(round((t_us % 1e6) * 1000))
        d1 = rng.uniform(-8000.0, 8000.0)
        rate = rng.uniform(-3e-6, 3e-6)                      # delay rate, us per us
        d2 = d1 + rate * subint_us
        dsec, dns, flags, bi, nbytes = control(host, h, 0, sec, ns, d1, d2)
        if dsec == ods.INVALID_SUBINT:
            continue
        nsamp = (nbytes // f.framebytes) * spf
        for i in range(spec.blockspersend):
            if not (flags[i // 30] >> (i % 30)) & 1:
                continue
            centre_delay = d1 + rate * (i + 0.5) * L * ts     # the block's model delay
            start = (sec - dsec) * 1e6 + (ns - dns) / 1000.0 + i * L * ts - centre_delay
            nearest = int(start / ts + 0.5)
            assert nearest >= -1 and nearest + L <= nsamp, (trial, i, nearest, nsamp)
            checked += 1
    assert checked > 1000
    host.fxds_destroy(h)


def test_aligned_subints_start_on_their_frame(host):
    """No delay, subint times on frame boundaries: the window starts at the
    subint's own frame, its time is the subint time, all blocks valid."""
    spec = SPECS["vdif_2bit_16ch"]
    h, o, f, keep = make(host, spec)
    nframes = 5000
    host.fxds_set_buffer(h, C.c_void_p(0x10000000), nframes * f.framebytes, 0, 100, 0)
    fps = int(round(f.framespersecond))
    for k in (0, 7, 100, 1000):
        t_ns = int(round(k * 1e9 / fps))
        sec, ns, flags, bi, nbytes = control(host, h, 0, 100 + t_ns // 10**9, t_ns % 10**9, 0.0, 0.0)
        assert (sec, ns, bi) == (100 + t_ns // 10**9, t_ns % 10**9, k * f.framebytes)
        assert flags == [(1 << 30) - 1] * (spec.blockspersend // 30) + [(1 << (spec.blockspersend % 30)) - 1]
    host.fxds_destroy(h)


def test_data_bytes_formula(host):
    """getDataBytes worked by hand: 16 channels x 2 bits x 2048 samples x 250
    blocks = 2 048 000 B; the 20 us guard is int(20000 / 15.625 + 1) = 1281
    samples = 5124 B; 8000-byte payloads: 2 053 124 // 8000 + 2 = 258 frames
    of 8032 bytes."""
    spec = SPECS["vdif_2bit_16ch"]
    job, keep = spec.build()
    assert host.fxds_data_bytes(C.byref(job), 0, GUARD_NS) == 258 * 8032
