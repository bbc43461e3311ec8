"""Complex sampling in the oracle (Mode with usecomplex / usedouble,
mpifxcorr/src/mode.cpp:51-105, 886-908, 1184-1241; mark5access complex VDIF
decoders format_vdif.c:2325-3054), pinned without the reference binary:

  - decoding: each unit's real part is the low nbits and its imaginary part
    the high nbits, read here bit by bit from the frames with numpy; complex
    8-bit decoders stop advancing at a blank frame (everything after it in
    the window is zero);
  - one station block in FP64: the FFT (L = nchan) of the complex samples
    times exp(-2 pi i (lo (a n/L + b + intdelay) - looffset t_n)) evaluated
    directly, with lo the band edge, the band centre for double sideband
    and negated for lower sideband; channels reordered (DSB halves swapped,
    LSB conjugated in reverse order with DC kept); times the fractional-
    sample rotation exp(2 pi i f_c (fse - clock)), f_c = c BW/C, (c - C)
    BW/C for LSB, (c - C/2) BW/C for DSB;
  - physics: identical voltages at two stations whose delay models differ by
    delta give a cross spectrum whose phase is 2 pi (sky frequency) delta,
    channel 0 being the band's lowest sky frequency, for USB, LSB and both
    double-sideband senses;
  - FP32 flavour against FP64.
"""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import oracle_block_info, oracle_results, oracle_spectra, rel_err

LUT2 = np.array([-3.3359, -1.0, 1.0, 3.3359], dtype=np.float32)
LUT8 = ((np.arange(256) - 128) / 3.3).astype(np.float32)


def unpack_complex(orc, spec, job, sd, ds, first, nsamp):
    d = job.datastreams[ds]
    out = np.zeros(spec.nrecbands * nsamp * 2, dtype=np.float32)
    w = C.c_float()
    raw = sd.stations[ds].raw
    check(orc, orc.orc_unpack(C.byref(d), raw.ctypes.data, raw.size, first, nsamp, out.ctypes.data,
                              C.byref(w)), "unpack")
    z = out.reshape(spec.nrecbands, nsamp, 2).astype(np.float64)
    return z[..., 0] + 1j * z[..., 1], w.value


def numpy_decode(spec, raw, first, nsamp):
    """Units of the frames read bit by bit: [band][sample] complex."""
    fb, hb = spec.framebytes, spec.headerbytes
    frames = raw.reshape(-1, fb)[:, hb:]
    bits = np.unpackbits(frames, axis=1, bitorder="little").reshape(-1)   # payload bits in order
    ub, nb, nbits = spec.unitbits, spec.nrecbands, spec.nbits
    t = first + np.arange(nsamp)
    out = np.zeros((nb, nsamp), dtype=np.complex128)
    lut = LUT2 if nbits == 2 else LUT8
    w = 1 << np.arange(nbits)
    for c in range(nb):
        base = (t[:, None] * nb + c) * ub + np.arange(nbits)[None, :]
        re = (bits[base] * w).sum(axis=1)
        im = (bits[base + nbits] * w).sum(axis=1)
        out[c] = lut[re] + 1j * lut[im].astype(np.float64)
    return out


@pytest.mark.parametrize("nbits,nsub,npol", [(2, 1, 2), (2, 2, 2), (2, 8, 2), (8, 1, 1), (8, 1, 2), (8, 2, 2)])
def test_complex_decode_bits(oracle, nbits, nsub, npol):
    spec = JobSpec(nstations=1, nsubbands=nsub, npol=npol, bw_mhz=16.0, nchan=128, nbits=nbits,
                   blockspersend=4, sampling="complex")
    job, keep = spec.build()
    sd = make_subint(spec, seed=3, delays_us=[0.0])
    first, n = 1000, 3000
    got, w = unpack_complex(oracle, spec, job, sd, 0, first, n)
    want = numpy_decode(spec, sd.stations[0].raw, first, n)
    np.testing.assert_array_equal(got, want)
    assert w == 1.0


def test_complex_8bit_stops_at_blank_frame(oracle):
    spec = JobSpec(nstations=1, nsubbands=1, npol=2, bw_mhz=16.0, nchan=128, nbits=8, blockspersend=4,
                   sampling="complex")
    job, keep = spec.build()
    sd = make_subint(spec, seed=4, delays_us=[0.0], invalid_frames={0: [2]})
    fs = spec.framesamples
    first, n = fs + 100, 2 * fs
    got, w = unpack_complex(oracle, spec, job, sd, 0, first, n)
    cut = 2 * fs - first                     # the invalid frame starts here
    want = numpy_decode(spec, sd.stations[0].raw, first, n)
    np.testing.assert_array_equal(got[:, :cut], want[:, :cut])
    assert np.all(got[:, cut:] == 0)         # no advance: the rest of the window is blank
    assert abs(w - cut / n) < 1e-6


@pytest.mark.parametrize("first", [1000, 1001, 2 * 16000 - 37, 2 * 16000 - 36, 2 * 16000 - 65, 2 * 16000 - 64])
def test_complex_single_channel_2bit_erasure(oracle, first):
    """One band of 2-bit complex samples: 4 bits per sample-time, so the
    decoder works on byte pairs (sample granularity 2, format_vdif.c:3845)
    and Mk5Mode::unpack unpacks nchan + 2 samples from the even sample at or
    before the block start, then erases floats [0, munged) and [nchan +
    munged, nchan + 2) of the band array (mk5mode.cpp:121-140): indices of
    floats, not of complex samples.  Viewed as floats, real then imaginary,
    that zeroes the whole sample nchan / 2 (even start) or the real part of
    the sample before the window and the imaginary part of sample nchan / 2
    (odd start), and the weight is (decoded good samples - nonzero floats
    erased) / nchan.  Checked against the frames' bits read with numpy, with
    a frame marked invalid so that the window runs into it."""
    spec = JobSpec(nstations=1, nsubbands=1, npol=1, bw_mhz=16.0, nchan=128, nbits=2, blockspersend=4,
                   sampling="complex")
    job, keep = spec.build()
    fs = spec.framesamples
    assert fs == 16000 and spec.unitbits == 4
    sd = make_subint(spec, seed=6, delays_us=[0.0], invalid_frames={0: [2]})
    n = spec.nchan
    munged = first % 2
    start = first - munged
    got, w = unpack_complex(oracle, spec, job, sd, 0, first, n)
    dec = numpy_decode(spec, sd.stations[0].raw, start, n + 2)[0]
    # the decoder's blanking: a frame marked invalid decodes as zeros
    f_of = (start + np.arange(n + 2)) // fs
    dec[f_of == 2] = 0
    good = int(np.count_nonzero(f_of != 2))
    fl = np.empty(2 * (n + 2), np.float64)
    fl[0::2], fl[1::2] = dec.real, dec.imag
    erase = list(range(munged)) + list(range(n + munged, n + 2))
    erased = int(np.count_nonzero(fl[erase]))
    fl[erase] = 0.0
    want = (fl[0::2] + 1j * fl[1::2])[munged:munged + n]
    np.testing.assert_array_equal(got[0], want)
    assert w == np.float32(np.float32(good) - np.float32(erased)) / np.float32(n)
    # the erased sample sits inside the window
    k = n // 2 - munged
    assert want[k].imag == 0 and (munged or want[k].real == 0)


def lo_of(spec, sb, lsb, dsb):
    lo = spec.base_freq_mhz + sb * spec.bw_mhz
    if dsb:
        lo += -spec.bw_mhz / 2 if lsb else spec.bw_mhz / 2
    return -lo if lsb else lo


def reorder(X, lsb, dsb):
    Cn = X.size
    c = np.arange(Cn)
    if not lsb:
        return X[(c + Cn // 2) % Cn] if dsb else X
    if dsb:
        return np.conj(X[(Cn // 2 - c) % Cn])
    out = np.conj(X[(Cn - c) % Cn])
    out[0] = X[0]
    return out


MODES = [(0, 0), (1, 0), (0, 1), (1, 1)]   # (lsb, dsb)


@pytest.mark.parametrize("lsb,dsb", MODES)
@pytest.mark.parametrize("looff", [0.0, 1234.567])
def test_complex_spectrum_closed_form(oracle, lsb, dsb, looff):
    spec = JobSpec(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16,
                   base_freq_mhz=8400.123, lsb=[lsb, lsb], clockoffset_us=[0.013, -0.021],
                   looffset_hz=[looff, -looff], sampling="complex_dsb" if dsb else "complex")
    job, keep = spec.build()
    sd = make_subint(spec, seed=7, delays_us=[3.2111, -7.777], rates_us_per_s=[0.8, -0.3])
    L, Cn, ts, bw = spec.fftchannels, spec.nchan, spec.sampletime_us, spec.bw_mhz
    assert L == Cn and ts == 1.0 / bw
    for ds in range(2):
        for blk in (0, 9):
            bi = oracle_block_info(oracle, job, sd, ds, blk)
            assert bi.status == 0
            x, _ = unpack_complex(oracle, spec, job, sd, ds, bi.nearestsample, L)
            got = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, Cn, fp64=True)
            st = sd.stations[ds]
            ip = st.interpolator
            fc = blk + 0.5
            t0 = (sd.sub_sec - st.datasec) * 1e6 + (sd.sub_ns - st.datans) / 1e3 + blk * L * ts - (
                ip[0] * fc * fc + ip[1] * fc + ip[2])
            fse = t0 - bi.nearestsample * ts
            walltime = bi.intwalltime + bi.fracwalltime
            n = np.arange(L)
            c = np.arange(Cn)
            lo_off = looff if ds == 0 else -looff
            clk = 0.013 if ds == 0 else -0.021
            fchan = (c - Cn / 2) if dsb else (c - Cn * lsb)
            for band in range(spec.nrecbands):
                lo = lo_of(spec, band // spec.npol, lsb, dsb)
                phase = lo * (bi.a * n / L + bi.b + bi.integerdelay) - lo_off * (n * ts * 1e-6 + walltime)
                X = reorder(np.fft.fft(x[band] * np.exp(-2j * np.pi * phase)), lsb, dsb)
                exp = X * np.exp(2j * np.pi * fchan * bw / Cn * (fse - clk))
                assert rel_err(got[band], exp) < 1e-9, (ds, blk, band, rel_err(got[band], exp))


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("lsb,dsb", MODES)
def test_complex_cross_phase_equals_sky_frequency_times_delay(oracle, lsb, dsb, order):
    spec = JobSpec(nstations=2, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=8,
                   base_freq_mhz=8400.0, lsb=[lsb], fringe_order=order,
                   sampling="complex_dsb" if dsb else "complex")
    job, keep = spec.build()
    ts = spec.sampletime_us
    delta = 0.3 * ts
    sd = make_subint(spec, seed=10, delays_us=[0.0, 0.0])
    sd.stations[1].raw = sd.stations[0].raw.copy()
    sd.stations[1].interpolator = np.array([0.0, 0.0, delta])
    res = oracle_results(oracle, job, sd, fp64=True)
    Cn, bw = spec.nchan, spec.bw_mhz
    v = res[0:Cn]
    c = np.arange(Cn)
    lowest = spec.base_freq_mhz - (bw if lsb else 0.0)
    sky = lowest + c * bw / Cn
    want = np.exp(2j * np.pi * sky * delta)
    got = v / np.abs(v)
    assert np.max(np.abs(got - want)) < 1e-9


@pytest.mark.parametrize("lsb,dsb", MODES)
def test_complex_fp32_matches_fp64(oracle, lsb, dsb):
    spec = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=32.0, nchan=512, blockspersend=10, lsb=[lsb, lsb],
                   sampling="complex_dsb" if dsb else "complex")
    job, keep = spec.build()
    sd = make_subint(spec, seed=12, rho=0.05, delays_us=[1.1, -3.3, 5.5], rates_us_per_s=[0.1, 0.2, -0.3])
    r32 = oracle_results(oracle, job, sd, fp64=False)
    r64 = oracle_results(oracle, job, sd, fp64=True)
    assert rel_err(r32, r64) < 1e-6
