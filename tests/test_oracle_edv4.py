"""The oracle's INTERLACEDVDIF path (Mk5Mode with per-band weights,
mpifxcorr/src/mk5mode.cpp:78-90, 145-168, blank_vdif_EDV4 of
libraries/mark5access/mark5access/format_vdif.c:4699-4775) and its
validity-propagating VDIF multiplexer (vdifmux.c:664-675, 799-805,
899-940), pinned independently of the restatement:

  - block and band weights recomputed in numpy from the frame headers: a
    band loses the samples of frames whose EDV4 mask (bit band / (nchan /
    mask length)) is clear, the block keeps good - sum(invalid) / nchan;
  - the autocorrelation weights are the band weights summed over blocks,
    the baseline weights the products of the two record bands' weights;
  - a band blanked over a whole block has a zero spectrum, the other bands
    are the plain VDIF spectra of the same bytes, and a stream whose first
    frame is not EDV4 correlates exactly as plain VDIF;
  - the multiplexer writes EDV4 headers (version 4, mask length = threads,
    sync word, the threads present as the mask), keeps a frame with some
    threads valid with the missing threads' units zero, and marks a frame
    with none invalid.
"""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import FXG_FORMAT_INTERLACEDVDIF, FXG_FORMAT_VDIF, JobSpec, check
from fxsynth import make_subint, set_edv4
from fxtestlib import oracle_block_info, oracle_results, oracle_spectra, plan_info
from muxlib import header, orc_mux, thread_stream


@pytest.fixture(scope="module")
def orc():
    return fxcorr.load_oracle()


def spec_of(fmt, nbits=2):
    # 2 stations x 2 subbands x 2 pol: four bands per datastream
    return JobSpec(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=40, nbits=nbits, fmt=fmt)


def edv4_case(nbits, masklength, masks0, masks1=None, seed=3):
    spec = spec_of(FXG_FORMAT_INTERLACEDVDIF, nbits)
    sd = make_subint(spec, seed=seed, rho=0.2, delays_us=[0.0, 1.0], rates_us_per_s=[0.0, 0.2],
                     invalid_frames={1: [2]})
    set_edv4(sd.stations[0].raw, spec.framebytes, masklength, masks0)
    set_edv4(sd.stations[1].raw, spec.framebytes, masklength, masks1 or {})
    return spec, sd


def expected_weights(orc, job, spec, sd, ds, masklength, nbits):
    """(block weights, band weights[block][band]) from the headers and bytes."""
    raw = sd.stations[ds].raw.reshape(-1, spec.framebytes)
    nfr = raw.shape[0]
    words = raw[:, :32].view(np.uint32)
    nch = spec.nrecbands
    fs = spec.framesamples
    d = nch // masklength
    good_mask = (1 << nch) - 1
    L = spec.fftchannels
    bw, bandw = [], []
    for blk in range(spec.blockspersend):
        bi = oracle_block_info(orc, job, sd, ds, blk)
        if bi.status in (1, 2):
            bw.append(np.float32(0))
            bandw.append(np.zeros(nch, np.float32))
            continue
        good = 0
        inv = np.zeros(nch, np.int64)
        for t in range(bi.nearestsample, bi.nearestsample + L):
            f = t // fs
            if f >= nfr or (words[f, 0] >> 31) & 1:
                continue                                  # decodes as zero: neither good nor counted
            good += 1
            vmask = int(words[f, 6]) | (int(words[f, 7]) << 32)
            if vmask == good_mask:
                continue
            tin = t - f * fs
            for b in range(nch):
                if (vmask >> (b // d)) & 1:
                    continue
                if nbits == 8:                            # (byte - 128) / 3.3 is zero for 128
                    if raw[f, 32 + tin * nch + b] == 128:
                        continue
                inv[b] += 1
        g = np.float32(good)
        bandw.append(np.array([(g - np.float32(inv[b])) / np.float32(L) for b in range(nch)], np.float32))
        g2 = g - np.float32(inv.sum()) / np.float32(nch)
        bw.append(np.float32(max(g2, np.float32(0)) / np.float32(L)))
    return bw, bandw


@pytest.mark.parametrize("nbits,masklength", [(2, 4), (2, 2), (8, 4)])
def test_block_and_band_weights(orc, nbits, masklength):
    masks0 = {1: 0b0101 if masklength == 4 else 0b01, 3: 0, 4: 0b1110 if masklength == 4 else 0b10}
    spec, sd = edv4_case(nbits, masklength, masks0)
    job, keep = spec.build()
    bw, bandw = expected_weights(orc, job, spec, sd, 0, masklength, nbits)
    assert min(bw) < 1.0                                # some blocks lose samples
    for blk in range(spec.blockspersend):
        bi = oracle_block_info(orc, job, sd, 0, blk)
        got = np.float32(bi.weight) if bi.status == 0 else np.float32(0)
        want = bw[blk] if bw[blk] > 0 else np.float32(0)
        assert got == want, (blk, got, want)
    # autocorrelation weights: the band weights summed in block order
    res = oracle_results(orc, job, sd, nthreads=1)
    off = C.c_int64()
    check(orc, orc.orc_plan_acweight_offset(C.byref(job), 0, C.byref(off)), "acw")
    acw = res.view(np.float32)[2 * off.value: 2 * off.value + spec.nrecbands]
    for b in range(spec.nrecbands):
        acc = np.float32(0)
        for blk in range(spec.blockspersend):
            bi = oracle_block_info(orc, job, sd, 0, blk)
            if bi.status == 0:
                acc = np.float32(acc + bandw[blk][b])
        assert acw[b] == acc, (b, acw[b], acc)
    # baseline weights of each product: the two record bands' weights multiplied
    bw1, bandw1 = expected_weights(orc, job, spec, sd, 1, masklength, nbits)
    bl = job.baselines[0]
    for f in range(bl.nfreqs):
        fi = bl.freq_index[f]
        check(orc, orc.orc_plan_bweight_offset(C.byref(job), fi, 0, C.byref(off)), "bw")
        for p in range(bl.npolproducts[f]):
            acc = np.float32(0)
            for blk in range(spec.blockspersend):
                b0 = oracle_block_info(orc, job, sd, 0, blk)
                b1 = oracle_block_info(orc, job, sd, 1, blk)
                w0 = bandw[blk][bl.band1[f][p]] if b0.status == 0 else np.float32(0)
                w1 = bandw1[blk][bl.band2[f][p]] if b1.status == 0 else np.float32(0)
                acc = np.float32(acc + np.float32(w0 * w1))
            assert res.view(np.float32)[2 * off.value + p] == acc, (f, p)


def test_blanked_band_is_zero_and_others_are_plain_vdif(orc):
    spec, sd = edv4_case(2, 4, {2: 0b1011})               # frame 2: band 2 absent
    job, keep = spec.build()
    plain = spec_of(FXG_FORMAT_VDIF)
    pjob, pkeep = plain.build()
    fs, L = spec.framesamples, spec.fftchannels
    inside = [blk for blk in range(spec.blockspersend)
              if (lambda bi: bi.status == 0 and bi.nearestsample >= 2 * fs and bi.nearestsample + L <= 3 * fs)(
                  oracle_block_info(orc, job, sd, 0, blk))]
    assert inside
    for blk in inside[:3]:
        got = oracle_spectra(orc, job, sd, 0, blk, spec.nrecbands, spec.nchan)
        ref = oracle_spectra(orc, pjob, sd, 0, blk, spec.nrecbands, spec.nchan)
        assert not np.any(got[2])
        for b in (0, 1, 3):
            assert np.array_equal(got[b], ref[b])


def test_first_frame_not_edv4_means_plain_vdif(orc):
    spec, sd = edv4_case(2, 4, {1: 0b0101, 4: 0})
    w = sd.stations[0].raw.reshape(-1, spec.framebytes)[0, :32].view(np.uint32)
    w[4] = np.uint32((3 << 24) | (4 << 16))              # extended version 3
    job, keep = spec.build()
    plain = spec_of(FXG_FORMAT_VDIF)
    pjob, pkeep = plain.build()
    a = oracle_results(orc, job, sd, nthreads=1)
    b = oracle_results(orc, pjob, sd, nthreads=1)
    # station 1 is EDV4 with complete masks: the same as plain VDIF too
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_mux_propagates_validity(orc):
    ids = [4, 5, 6]
    fb = 32 + 64
    src, pay = thread_stream(ids, 10, framebytes=fb, seed=9, drop=[(2, 1), (4, 0), (4, 1), (4, 2)])
    r, dest, st = orc_mux(orc, src, ids, 12, framebytes=fb, nsort=4, ngap=4, flags=0x41)
    out = dest.reshape(-1, (fb - 32) * 4 + 32)
    assert st[9] == 1                                    # one partial output frame
    nout = st[14]
    assert nout >= 6 and st[10] == 1                     # frame 4 (no thread) is the one invalid frame
    for f in range(nout):
        w = out[f, :32].view(np.uint32)
        assert w[4] >> 24 == 4 and (w[4] >> 16) & 0xFF == 3 and w[5] == 0xACABFEED
        present = [k for k in range(3) if (f, k) not in [(2, 1), (4, 0), (4, 1), (4, 2)]]
        mask = int(w[6]) | (int(w[7]) << 32)
        if not present:
            assert w[0] >> 31 == 1 and mask == 0
            continue
        assert w[0] >> 31 == 0 and mask == sum(1 << k for k in present)
        # 2-bit units: byte n of the payload holds sample n of threads 0, 1, 2 (and a zero unit)
        payload = out[f, 32:]
        for n in range(0, 4 * (fb - 32), 37):
            byte = payload[n]
            for k in range(3):
                unit = (byte >> (2 * k)) & 3
                if k in present:
                    src_byte = pay[(f, k)][n // 4]
                    assert unit == (src_byte >> (2 * (n % 4))) & 3
                else:
                    assert unit == 0
            assert byte >> 6 == 0
