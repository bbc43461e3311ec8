"""INTERLACEDVDIF on the GPU against the oracle: EDV4 per-band blanking and
weights (mk5mode.cpp:78-90, 145-168; format_vdif.c:4699-4775) and VDIF
multiplexing with validity propagation (vdifmux.c:664-675, 799-805,
899-940).

  - jobs whose datastreams carry EDV4 masks with absent threads (2- and
    8-bit, one or two bands per mask bit, an invalid frame): visibilities
    within 1e-5 of the FP32 oracle, block geometry and every weight of the
    result (baseline and autocorrelation weights) bit for bit;
  - fxg_vdifmux with validity propagation: output bytes and statistics as
    the oracle's multiplexer;
  - a station's frames split into one thread per band, some thread frames
    dropped, multiplexed back on the GPU with validity propagation and
    correlated as INTERLACEDVDIF: the same results as the oracle from the
    same bytes.
"""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_FORMAT_INTERLACEDVDIF, FXG_MUX_PROPAGATEVALIDITY, JobSpec, check
from fxsynth import make_subint, set_edv4
from fxtestlib import GpuJob, oracle_block_info, oracle_results, oracle_spectra, plan_info, region_slices, rel_err
from muxlib import header, orc_mux, thread_stream
from test_gpu_vdifmux import gpu_mux

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def edv4_spec(nbits, **kw):
    base = dict(nstations=4, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=48, nbits=nbits,
                fmt=FXG_FORMAT_INTERLACEDVDIF)
    base.update(kw)
    return JobSpec(**base)


def compare(lib, ctx, oracle, spec, sd, tag):
    job, keep = spec.build()
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd)
    info = plan_info(oracle, "orc_", job)
    assert g.info.result_length == info.result_length
    xend, a0, a1 = region_slices(oracle, job, info)
    ex, ea = rel_err(got[:xend], ref[:xend]), rel_err(got[a0:a1], ref[a0:a1])
    print(f"{tag}: cross {ex:.3e} auto {ea:.3e}", flush=True)
    assert ex < FP32_TOL and ea < FP32_TOL
    # baseline and autocorrelation weights: the same FP32 terms in block order
    assert np.array_equal(got[xend:a0].view(np.float32), ref[xend:a0].view(np.float32))
    assert np.array_equal(got[a1:].view(np.float32), ref[a1:].view(np.float32))
    for ds in range(spec.nstations):
        for blk in range(spec.blockspersend):
            a, b = g.block_info(ds, blk), oracle_block_info(oracle, job, sd, ds, blk)
            assert (a.status, a.nearestsample) == (b.status, b.nearestsample), (ds, blk)
            if b.status == 0:
                assert a.weight == b.weight, (ds, blk, a.weight, b.weight)
    return g, job, keep, got


@pytest.mark.parametrize("nbits,masklength", [(2, 4), (2, 2), (8, 4)])
def test_edv4_job_matches_oracle(gpu, oracle, nbits, masklength):
    lib, ctx = gpu
    spec = edv4_spec(nbits)
    sd = make_subint(spec, seed=17 + nbits + masklength, rho=0.1, delays_us=[0.0, 3.0, -7.5, 11.0],
                     rates_us_per_s=[0.0, 0.5, -0.3, 0.8], invalid_frames={2: [3]})
    full = (1 << masklength) - 1
    set_edv4(sd.stations[0].raw, spec.framebytes, masklength, {1: full & ~1, 2: full & ~0b10, 4: 0})
    set_edv4(sd.stations[1].raw, spec.framebytes, masklength, {2: full & ~(1 << (masklength - 1))})
    set_edv4(sd.stations[2].raw, spec.framebytes, masklength, {})
    set_edv4(sd.stations[3].raw, spec.framebytes, masklength, {3: 1})
    g, job, keep, got = compare(lib, ctx, oracle, spec, sd, f"edv4 {nbits}-bit mask {masklength}")
    # frame 2 of station 0 lacks mask bit 1: over a block inside it, the
    # bands of that bit have zero spectra on both sides, the others match
    fs, L = spec.framesamples, spec.fftchannels
    d = spec.nrecbands // masklength
    checked = 0
    for blk in range(spec.blockspersend):
        bi = oracle_block_info(oracle, job, sd, 0, blk)
        if not (bi.status == 0 and bi.nearestsample >= 2 * fs and bi.nearestsample + L <= 3 * fs):
            continue
        ref = oracle_spectra(oracle, job, sd, 0, blk, spec.nrecbands, spec.nchan)
        for band in range(spec.nrecbands):
            gs = g.spectrum(0, band, blk, spec.nchan)
            if band // d == 1:
                assert not np.any(ref[band]) and not np.any(gs), (blk, band)
            else:
                assert rel_err(gs, ref[band]) < FP32_TOL, (blk, band)
        checked += 1
    assert checked > 0
    g.close()


MUXCASES = {
    "4_threads_gaps": dict(ids=[0, 1, 2, 3], stream=dict(nframes=14, drop={(3, 1), (5, 0), (5, 2), (8, 3)}),
                           mux=dict(nsort=4, ngap=4)),
    "3_threads_padded": dict(ids=[5, 9, 2], stream=dict(nframes=12, drop={(4, 2), (6, 0), (6, 1), (6, 2)}),
                             mux=dict(nsort=6, ngap=6)),
    "8bit_invalid": dict(ids=[0, 1], stream=dict(nframes=10, nbits=8, drop={(2, 1)}, invalid={(5, 0)}),
                         mux=dict(nbits=8, nsort=4, ngap=4)),
}


@pytest.mark.parametrize("name", sorted(MUXCASES))
def test_mux_propagating_validity_matches_oracle(gpu, oracle, name):
    lib, ctx = gpu
    c = MUXCASES[name]
    src, _ = thread_stream(c["ids"], framebytes=1032, seed=len(name) + 40, **c["stream"])
    flags = FXG_MUX_PROPAGATEVALIDITY | 0x01 | (0x04 if "invalid" in name else 0)
    mkw = dict(framebytes=1032, flags=flags, **c["mux"])
    r0, d0, s0 = orc_mux(oracle, src, c["ids"], 24, **mkw)
    r1, d1, s1 = gpu_mux(lib, ctx, src, c["ids"], 24, **mkw)
    assert r1 == r0 and s1 == s0, (name, r1, r0, s1, s0)
    assert s0[9] >= 1                                           # partial frames kept
    used = s0[12]
    assert np.array_equal(d1[:used], d0[:used]), name


def test_demuxed_station_with_missing_threads_correlates_like_oracle(gpu, oracle):
    """Station 0's 4-band frames split into 4 one-band threads, a few thread
    frames dropped, multiplexed back on the GPU with validity propagation
    into an INTERLACEDVDIF stream; the job (its other stations plain VDIF
    frames with complete EDV4 masks) on the GPU equals the oracle."""
    lib, ctx = gpu
    spec = edv4_spec(2, nstations=3)
    sd = make_subint(spec, seed=23, rho=0.1, delays_us=[0.0, 2.0, -4.0], rates_us_per_s=[0.0, 0.4, -0.6])
    for k in (1, 2):
        set_edv4(sd.stations[k].raw, spec.framebytes, spec.nrecbands, {})
    fb, nch = spec.framebytes, spec.nrecbands
    raw = sd.stations[0].raw
    frames = raw[: raw.size // fb * fb].reshape(-1, fb)
    nfr = frames.shape[0]
    spf = (fb - 32) * 8 // (2 * nch)
    codes = np.unpackbits(frames[:, 32:], axis=1, bitorder="little").reshape(nfr, spf, nch, 2)
    hdr = frames[:, :32].copy().view(np.uint32)
    fps = int(round(1e6 / spec.sampletime_us)) // spf
    drop = {(2, 1), (3, 3), (5, 0), (5, 2)}
    threads = []
    for f in range(nfr):
        sec = int(hdr[f, 0] & 0x3FFFFFFF)
        fn = int(hdr[f, 1] & 0xFFFFFF)
        for k in range(nch):
            if (f, k) in drop:
                continue
            pay = np.packbits(codes[f, :, k, :].reshape(-1), bitorder="little")
            threads.append(np.concatenate([header(sec, fn, k, 32 + pay.size, 1, 2, station=0), pay]))
    src = np.concatenate(threads)
    mkw = dict(framebytes=32 + spf * 2 // 8, fps=fps, nbits=2, nsort=nch, ngap=nch,
               flags=FXG_MUX_PROPAGATEVALIDITY | 0x01)
    r, out, st = gpu_mux(lib, ctx, src, list(range(nch)), nfr, **mkw)
    r0, out0, st0 = orc_mux(oracle, src, list(range(nch)), nfr, **mkw)
    assert st == st0 and st[14] == nfr and st[9] == 3             # frames 2, 3 and 5 partial
    assert np.array_equal(out, out0)
    sd.stations[0].raw = out[: nfr * fb].copy()
    g, job, keep, got = compare(lib, ctx, oracle, spec, sd, "demuxed station")
    g.close()
