"""Pins the oracle's station processing (Mode::process, mode.cpp:649-1397)
and cross-multiplication (Core::processdata, core.cpp:826-975) against
closed forms and physics, independently of the sub/step rotator structure:

  - block geometry: nearest sample rounds the model start time to the sample
    grid and the fractional-sample error stays within half a sample;
  - one station block: FFT of x[n] exp(-2 pi i (lofreq (a n/L + b) -
    looffset t_n)) evaluated directly, times the fractional-sample rotation
    exp(2 pi i (c - C lsb) BW/C (fse - clock)), in FP64;
  - autocorrelation Parseval identity with a zero delay model;
  - two stations with identical data whose models differ by delta: the
    cross spectrum's phase is 2 pi (sky frequency) delta, channel by channel
    (sign and scale of fringe and fractional-sample corrections and the
    conj(second station) order of the product);
  - fringe order 0 (post-F): FFT of the unrotated samples times the block's
    fringe phase at its centre delay and the fractional-sample rotation;
  - fringe order 2 (quadratic): the piecewise phase of the sub/step/quad
    rotators written out directly, and its distance from the exact quadratic
    phase bounded by the chord error lofreq a (s/L)^2 / 4;
  - FP32 flavour against the FP64 flavour.
"""
import ctypes as C

import numpy as np
import pytest

from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import (mixed_sideband_job, oracle_block_info, oracle_results, oracle_spectra, plan_info, rel_err,
                       usb_to_lsb_raw)


def unpacked(orc, spec, job, sd, ds, ns):
    d = job.datastreams[ds]
    out = np.zeros(spec.nrecbands * spec.fftchannels, dtype=np.float32)
    w = C.c_float()
    raw = sd.stations[ds].raw
    check(orc, orc.orc_unpack(C.byref(d), raw.ctypes.data, raw.size, ns, spec.fftchannels, out.ctypes.data,
                              C.byref(w)), "unpack")
    return out.reshape(spec.nrecbands, spec.fftchannels).astype(np.float64)


def model_starttime(spec, sd, ds, index):
    st = sd.stations[ds]
    ip = st.interpolator
    fc = index + 0.5
    delay = ip[0] * fc * fc + ip[1] * fc + ip[2]
    ts = spec.sampletime_us
    return (sd.sub_sec - st.datasec) * 1e6 + (sd.sub_ns - st.datans) / 1e3 + index * spec.fftchannels * ts - delay


def test_block_geometry_properties(oracle):
    spec = JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=64)
    job, keep = spec.build()
    sd = make_subint(spec, seed=2, delays_us=[0.37, -12.81, 44.449], rates_us_per_s=[0.9, -0.7, 0.33])
    ts = spec.sampletime_us
    for ds in range(3):
        for blk in range(0, 64, 7):
            bi = oracle_block_info(oracle, job, sd, ds, blk)
            assert bi.status == 0
            t = model_starttime(spec, sd, ds, blk)
            assert bi.nearestsample == int(np.floor(t / ts + 0.5))
            assert abs(bi.fracsampleerror - np.float32(t - bi.nearestsample * ts)) == 0.0
            assert abs(bi.fracsampleerror) <= ts / 2 + 1e-6
            ip = sd.stations[ds].interpolator
            d0 = ip[0] * blk * blk + ip[1] * blk + ip[2]
            d2 = ip[0] * (blk + 1) ** 2 + ip[1] * (blk + 1) + ip[2]
            assert abs(bi.a - (d2 - d0)) < 1e-12
            assert bi.integerdelay == int(np.trunc(bi.b + bi.integerdelay))


@pytest.mark.parametrize("lsb", [0, 1])
@pytest.mark.parametrize("looff", [0.0, 1234.567])
def test_station_spectrum_closed_form(oracle, lsb, looff):
    spec = JobSpec(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16,
                   base_freq_mhz=8400.123, lsb=[lsb, lsb], clockoffset_us=[0.013, -0.021],
                   looffset_hz=[looff, -looff])
    job, keep = spec.build()
    sd = make_subint(spec, seed=7, delays_us=[3.2111, -7.777], rates_us_per_s=[0.8, -0.3])
    L, Cn, ts, bw = spec.fftchannels, spec.nchan, spec.sampletime_us, spec.bw_mhz
    for ds in range(2):
        for blk in (0, 9):
            bi = oracle_block_info(oracle, job, sd, ds, blk)
            assert bi.status == 0
            x = unpacked(oracle, spec, job, sd, ds, bi.nearestsample)
            got = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, Cn, fp64=True)
            fse = model_starttime(spec, sd, ds, blk) - bi.nearestsample * ts
            walltime = bi.intwalltime + bi.fracwalltime
            n = np.arange(L)
            c = np.arange(Cn)
            lo_off = looff if ds == 0 else -looff
            clk = 0.013 if ds == 0 else -0.021
            for band in range(spec.nrecbands):
                sb = band // spec.npol
                lofreq = spec.base_freq_mhz + sb * bw
                phase = lofreq * (bi.a * n / L + bi.b + bi.integerdelay) - lo_off * (n * ts * 1e-6 + walltime)
                X = np.fft.fft(x[band] * np.exp(-2j * np.pi * phase))[c + Cn * lsb]
                theta = 2 * np.pi * (c - Cn * lsb) * bw / Cn * (fse - clk)
                exp = X * np.exp(1j * theta)
                assert rel_err(got[band], exp) < 1e-9, (ds, blk, band, rel_err(got[band], exp))


@pytest.mark.parametrize("lsb", [0, 1])
@pytest.mark.parametrize("looff", [0.0, 1234.567])
def test_station_spectrum_order0_closed_form(oracle, lsb, looff):
    spec = JobSpec(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16,
                   base_freq_mhz=8400.123, lsb=[lsb, lsb], clockoffset_us=[0.013, -0.021],
                   looffset_hz=[looff, -looff], fringe_order=0)
    job, keep = spec.build()
    sd = make_subint(spec, seed=7, delays_us=[3.2111, -7.777], rates_us_per_s=[0.8, -0.3])
    L, Cn, ts, bw = spec.fftchannels, spec.nchan, spec.sampletime_us, spec.bw_mhz
    for ds in range(2):
        ip = sd.stations[ds].interpolator
        for blk in (0, 9):
            bi = oracle_block_info(oracle, job, sd, ds, blk)
            assert bi.status == 0
            fc = blk + 0.5
            avgdelay = ip[0] * fc * fc + ip[1] * fc + ip[2]
            assert bi.integerdelay == int(np.trunc(avgdelay)) and abs(bi.c - avgdelay) < 1e-12
            x = unpacked(oracle, spec, job, sd, ds, bi.nearestsample)
            got = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, Cn, fp64=True)
            fse = model_starttime(spec, sd, ds, blk) - bi.nearestsample * ts
            walltime = bi.intwalltime + bi.fracwalltime
            c = np.arange(Cn)
            lo_off = looff if ds == 0 else -looff
            clk = 0.013 if ds == 0 else -0.021
            for band in range(spec.nrecbands):
                lofreq = spec.base_freq_mhz + (band // spec.npol) * bw
                X = np.fft.fft(x[band])[c + Cn * lsb]
                fringe = np.exp(-2j * np.pi * (avgdelay * lofreq - walltime * lo_off))
                theta = 2 * np.pi * (c - Cn * lsb) * bw / Cn * (fse - clk)
                exp = X * fringe * np.exp(1j * theta)
                assert rel_err(got[band], exp) < 1e-9, (ds, blk, band, rel_err(got[band], exp))


@pytest.mark.parametrize("lsb", [0, 1])
def test_station_spectrum_order2_closed_form(oracle, lsb):
    spec = JobSpec(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16,
                   base_freq_mhz=8400.123, lsb=[lsb, lsb], clockoffset_us=[0.013, -0.021],
                   looffset_hz=[777.25, -31.5], fringe_order=2)
    job, keep = spec.build()
    sd = make_subint(spec, seed=7, delays_us=[3.2111, -7.777], rates_us_per_s=[0.8, -0.3])
    # a strongly curved model so the quadratic term matters
    for st, q in zip(sd.stations, (2.5e-5, -4.0e-5)):
        st.interpolator = st.interpolator + np.array([q, 0.0, 0.0])
    L, Cn, ts, bw = spec.fftchannels, spec.nchan, spec.sampletime_us, spec.bw_mhz
    s = oracle.orc_calcstridelength(Cn)
    n = np.arange(L)
    k, j = n % s, n // s
    for ds in range(2):
        ip = sd.stations[ds].interpolator
        for blk in (0, 9):
            bi = oracle_block_info(oracle, job, sd, ds, blk)
            assert bi.status == 0
            cc = ip[2] + blk * ip[1] + blk * blk * ip[0]
            assert bi.integerdelay == int(np.trunc(cc)) and abs(bi.a - ip[0]) < 1e-15
            x = unpacked(oracle, spec, job, sd, ds, bi.nearestsample)
            got = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, Cn, fp64=True)
            fse = model_starttime(spec, sd, ds, blk) - bi.nearestsample * ts
            walltime = bi.intwalltime + bi.fracwalltime
            c = np.arange(Cn)
            lo_off = 777.25 if ds == 0 else -31.5
            clk = 0.013 if ds == 0 else -0.021
            # piecewise quadratic delay in block units: exact at the stride
            # points j s, linear in between (the sub/quad/step rotators)
            u = (bi.b * n + bi.a * (k * s + 2 * j * k * s + (j * s) ** 2) / L) / L + bi.c
            exact = (bi.b * n + bi.a * n * n / L) / L + bi.c
            assert np.max(np.abs(u - exact)) <= abs(bi.a) * (s / L) ** 2 / 4 * 1.000001
            for band in range(spec.nrecbands):
                lofreq = spec.base_freq_mhz + (band // spec.npol) * bw
                phase = lofreq * (u + bi.integerdelay) - lo_off * (n * ts * 1e-6 + walltime)
                X = np.fft.fft(x[band] * np.exp(-2j * np.pi * phase))[c + Cn * lsb]
                theta = 2 * np.pi * (c - Cn * lsb) * bw / Cn * (fse - clk)
                exp = X * np.exp(1j * theta)
                assert rel_err(got[band], exp) < 1e-9, (ds, blk, band, rel_err(got[band], exp))


def test_autocorrelation_parseval(oracle):
    spec = JobSpec(nstations=2, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=8)
    job, keep = spec.build()
    sd = make_subint(spec, seed=9, delays_us=[0.0, 0.0])
    res = oracle_results(oracle, job, sd, fp64=True)
    off = C.c_int64()
    check(oracle, oracle.orc_plan_autocorr_offset(C.byref(job), 0, C.byref(off)), "ac")
    L, Cn = spec.fftchannels, spec.nchan
    for band in range(spec.nrecbands):
        ac = res[off.value + band * Cn: off.value + (band + 1) * Cn]
        tot = 0.0
        for blk in range(spec.blockspersend):
            bi = oracle_block_info(oracle, job, sd, 0, blk)
            x = unpacked(oracle, spec, job, sd, 0, bi.nearestsample)[band]
            x0, xc = x.sum(), (x * (-1.0) ** np.arange(L)).sum()
            tot += (L * np.sum(x * x) + x0 * x0 - xc * xc) / 2
        assert abs(ac.real.sum() - tot) < 1e-10 * tot
        assert np.max(np.abs(ac.imag)) == 0.0


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("lsb", [0, 1])
def test_cross_phase_equals_sky_frequency_times_delay(oracle, lsb, order):
    spec = JobSpec(nstations=2, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=8,
                   base_freq_mhz=8400.0, lsb=[lsb], fringe_order=order)
    job, keep = spec.build()
    ts = spec.sampletime_us
    delta = 0.3 * ts
    sd = make_subint(spec, seed=10, delays_us=[0.0, 0.0])
    sd.stations[1].raw = sd.stations[0].raw.copy()            # identical voltages
    sd.stations[1].interpolator = np.array([0.0, 0.0, delta])  # model differs by delta
    res = oracle_results(oracle, job, sd, fp64=True)
    Cn, bw = spec.nchan, spec.bw_mhz
    v = res[0:Cn]                                              # baseline 0, RR
    c = np.arange(Cn)
    lowest = spec.base_freq_mhz - (bw if lsb else 0.0)          # sky frequency of channel 0
    sky = lowest + c * bw / Cn
    want = np.exp(2j * np.pi * sky * delta)
    got = v / np.abs(v)
    assert np.max(np.abs(got - want)) < 1e-9


def test_fp32_matches_fp64(oracle):
    spec = JobSpec(nstations=4, nsubbands=2, npol=2, bw_mhz=32.0, nchan=512, blockspersend=12)
    job, keep = spec.build()
    sd = make_subint(spec, seed=12, rho=0.05, delays_us=[1.1, -3.3, 5.5, 0.0], rates_us_per_s=[0.1, 0.2, -0.3, 0])
    r32 = oracle_results(oracle, job, sd, fp64=False)
    r64 = oracle_results(oracle, job, sd, fp64=True)
    assert rel_err(r32, r64) < 1e-6


@pytest.mark.parametrize("order", [1, 2])
def test_usb_against_lsb_phase_equals_sky_frequency_times_delay(oracle, order):
    """A USB station against an LSB station recording the same sky band (the
    USB x LSB baseline products of configuration.cpp:2926-2937): the LSB
    voltages are the USB ones mirrored about the band centre, x[n] (-1)^n,
    and the delay models differ by delta; the cross spectrum's phase is
    2 pi (sky frequency) delta in every channel, channel c of both bands
    being the same sky frequency."""
    spec = JobSpec(nstations=2, nsubbands=1, npol=1, bw_mhz=16.0, nchan=256, blockspersend=8,
                   base_freq_mhz=8400.0, fringe_order=order, pols="R")
    job, keep = mixed_sideband_job(spec, [1])
    ts = spec.sampletime_us
    delta = 0.3 * ts
    sd = make_subint(spec, seed=10, delays_us=[0.0, 0.0])
    sd.stations[1].raw = usb_to_lsb_raw(spec, sd.stations[0].raw)
    sd.stations[1].interpolator = np.array([0.0, 0.0, delta])
    res = oracle_results(oracle, job, sd, fp64=True)
    Cn, bw = spec.nchan, spec.bw_mhz
    v = res[0:Cn]
    sky = spec.base_freq_mhz + np.arange(Cn) * bw / Cn
    got = v / np.abs(v)
    want = np.exp(2j * np.pi * sky * delta)
    assert np.max(np.abs(got - want)) < 1e-9, np.max(np.abs(got - want))


def _polturn(spec, s, band):
    """e^{i theta/2} of the band turned by station s's phase offset theta
    (the second of a frequency's first two bands if the first is R, else the
    first), 1 for the other band."""
    npol = spec.npol
    sb, p = divmod(band, npol)
    turned = 1 if spec.pols[0] == "R" else 0
    po = spec.phaseoffset_deg[s]
    th = po[sb] if isinstance(po, (list, tuple)) else po
    return np.exp(0.5j * np.deg2rad(th)) if p == turned else 1.0


@pytest.mark.parametrize("pols", ["RL", "LR", "XY"])
def test_pol_phase_offset_turns_one_band_per_frequency(oracle, pols):
    """Recorded-frequency polarisation phase offsets (mode.cpp:337-346,
    435-460, 1341-1353): with any offset set, each frequency's LCP band (the
    second of its first two bands when the first is R, otherwise the first)
    is turned by e^{i theta/2} after the fractional-sample rotation.  Against
    the same data without offsets, product p x q* of stations a, b changes by
    turn(a, p) conj(turn(b, q)); self autocorrelations do not change and the
    cross-pol ones change by turn(R) conj(turn(L)) and its conjugate.  The
    offsets differ per station and subband, one of them zero."""
    base = dict(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=6,
                base_freq_mhz=8400.0, pols=pols, clockoffset_us=[0.01, -0.02, 0.0])
    spec0 = JobSpec(**base)
    spec1 = JobSpec(**base, phaseoffset_deg=[[30.0, -70.0], [110.0, 0.0], 45.0])
    job0, k0 = spec0.build()
    job1, k1 = spec1.build()
    sd = make_subint(spec0, seed=21, rho=0.2, delays_us=[0.3, -1.1, 0.7])
    r0 = oracle_results(oracle, job0, sd, fp64=True)
    r1 = oracle_results(oracle, job1, sd, fp64=True)
    Cn = spec0.nchan
    off = C.c_int64()
    pairs = [(i, j) for i in range(3) for j in range(i + 1, 3)]
    prods = [(0, 0), (0, 1), (1, 0), (1, 1)]
    checked = 0
    for b, (sa, sb_) in enumerate(pairs):
        for f in range(spec0.nsubbands):
            check(oracle, oracle.orc_plan_baseline_offset(C.byref(job1), f, b, C.byref(off)), "bl")
            for k, (p, q) in enumerate(prods):
                s0 = slice(off.value + k * Cn, off.value + (k + 1) * Cn)
                want = r0[s0] * _polturn(spec1, sa, 2 * f + p) * np.conj(_polturn(spec1, sb_, 2 * f + q))
                assert np.max(np.abs(r1[s0] - want)) <= 1e-9 * np.max(np.abs(r0[s0])), (b, f, p, q)
                checked += 1
    assert checked == 3 * 2 * 4
    for s in range(3):
        check(oracle, oracle.orc_plan_autocorr_offset(C.byref(job1), s, C.byref(off)), "ac")
        a = off.value
        nb = spec0.nrecbands
        self0, self1 = r0[a:a + nb * Cn], r1[a:a + nb * Cn]
        assert np.max(np.abs(self1 - self0)) <= 1e-9 * np.max(np.abs(self0))
        for band in range(nb):
            other = band ^ 1
            x = slice(a + (nb + band) * Cn, a + (nb + band + 1) * Cn)
            want = r0[x] * _polturn(spec1, s, band) * np.conj(_polturn(spec1, s, other))
            assert np.max(np.abs(r1[x] - want)) <= 1e-9 * np.max(np.abs(r0[x])), (s, band)
    # the turn is not the identity anywhere it should act
    assert abs(_polturn(spec1, 2, 1 if pols[0] == "R" else 0) - 1.0) > 0.3
