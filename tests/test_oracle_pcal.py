"""Phase-cal extraction in the oracle, pinned by the reference's own test
pattern (mpifxcorr/src/pcal.cpp:1532-1640: tones with a 5 degree phase slope,
expected phase -90 deg + 5 deg per tone), by the reference's slow direct
method (PCal::extractAndIntegrate_reference, pcal.cpp:253-331), by the
agreement of the shifting and implicit-shift extractors on whole periods,
and by the tone tables of configuration.cpp:1771-1845 worked by hand."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check

P = C.POINTER
PC_DUMMY, PC_TRIVIAL, PC_SHIFT, PC_IMPLICIT = 0, 1, 2, 3


@pytest.fixture(scope="module")
def orc():
    lib = fxcorr.load_oracle()
    lib.orc_pcal_extract.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                     P(C.c_int64), P(C.c_int64), P(C.c_int64), P(C.c_float), P(C.c_float),
                                     P(C.c_int32), P(C.c_int64)]
    lib.orc_pcal_direct.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int64, P(C.c_float), C.c_int64,
                                    P(C.c_double), P(C.c_int32)]
    return lib


def tone_pattern(bw, offset, spacing, n, so, slope_deg=5.0):
    """The reference test signal: sum over tones t of sin(w_t (n + so) + t slope)."""
    fs = 2 * bw
    x = np.zeros(n)
    f, t = float(offset), 0
    while spacing > 0 and f <= bw:
        x += np.sin(2 * np.pi * f / fs * (np.arange(n) + so) + np.radians(t * slope_deg))
        f += spacing
        t += 1
    return x.astype(np.float32)


def extract(orc, bw, spacing, offset, kind, minres, segs, data):
    """segs: (sampleoffset, start, length) per extractAndIntegrate call."""
    n = len(segs)
    so = (C.c_int64 * n)(*[s[0] for s in segs])
    st = (C.c_int64 * n)(*[s[1] for s in segs])
    ln = (C.c_int64 * n)(*[s[2] for s in segs])
    out = np.zeros(2 * 4096, np.float32)
    nt, ns = C.c_int32(), C.c_int64()
    k = orc.orc_pcal_extract(bw, spacing, int(offset), kind, minres, n, so, st, ln,
                             data.ctypes.data_as(P(C.c_float)), out.ctypes.data_as(P(C.c_float)),
                             C.byref(nt), C.byref(ns))
    z = out[:2 * nt.value].reshape(-1, 2).astype(np.float64)
    return k, z[:, 0] + 1j * z[:, 1], ns.value


# (bandwidth, offset, spacing, mode) from pcal.cpp:1532-1547; the two modes the
# reference marks "fails" are left out
REF_CASES = [
    (16e6, 0, 1e6, -1), (3e6, 0, 2e6, -1), (16e6, 10e3, 1e6, -1), (16e6, 10e3, 3e6, -1),
    (16e6, 10e3, 5e6, -1), (16e6, 0, 5e6, -1), (1e6, 10e3, 5e6, -1),
    (1e6, 10e3, 0, -1), (1e6, 0, 0, -1), (1e6, 2e6, 0, -1), (32e3, 2e6, 100e3, -1),
    (16e6, 10e3, 1e6, PC_SHIFT), (16e6, 10e3, 3e6, PC_SHIFT), (16e6, 0, 1e6, PC_TRIVIAL),
]


@pytest.mark.parametrize("bw,offset,spacing,mode", REF_CASES)
def test_reference_unit_pattern(orc, bw, offset, spacing, mode):
    """The reference's test_pcal_case: 32000 samples at sample offset 11, then
    the same data again from sample 11 with the offset adjusted; the tones'
    phases must slope by 5 degrees from -90 (-85 when tone 0 sits at DC and is
    dropped), amplitudes = half the samples integrated."""
    so, n, skip = 11, 32000, 11
    data = tone_pattern(bw, offset, spacing, n, so)
    k, tones, count = extract(orc, bw, spacing, offset, mode, 10e3,
                              [(so, 0, n), (so + skip, skip, n - skip)], data)
    if spacing == 0 or offset > bw:
        assert k == PC_DUMMY and len(tones) == 0
        return
    assert count == 2 * n - skip
    start = -90.0 + (5.0 if offset == 0 else 0.0)
    for t, v in enumerate(tones):
        ftone = (t + (1 if offset == 0 else 0)) * spacing + offset
        if ftone >= bw:                    # a tone on the Nyquist edge has no phase
            continue
        ph = np.degrees(np.angle(v))
        assert abs((ph - (start + 5.0 * t) + 180) % 360 - 180) < 0.05, (t, ph)
        assert abs(abs(v) - count / 2) < 2e-3 * count, (t, abs(v))


def test_getnew_choices(orc):
    """PCal::getNew: offset 0 -> first tone at DC, offset period a multiple of
    the comb period -> implicit shift, otherwise shifting; minimum resolution
    doubling changes the implicit test (pcal.cpp:150-200)."""
    x = np.zeros(64, np.float32)
    kinds = {}
    for bw, off, sp in [(16e6, 0, 1e6), (16e6, 10e3, 1e6), (16e6, 10e3, 3e6), (16e6, 333e3, 1e6),
                        (16e6, -1, 1e6)]:
        kinds[(bw, off, sp)] = extract(orc, bw, sp, off, -1, 1e6, [(0, 0, 64)], x)[0]
    assert kinds[(16e6, 0, 1e6)] == PC_TRIVIAL
    assert kinds[(16e6, 10e3, 1e6)] == PC_IMPLICIT
    # No = 3200, Np = 32e6 / gcd(32e6, 3e6) = 32: 3200 % 32 == 0 -> implicit
    assert kinds[(16e6, 10e3, 3e6)] == PC_IMPLICIT
    # 333 kHz: No = 32e6 / gcd(32e6, 333e3) = 32000, Np = 32 -> implicit
    assert extract(orc, 16e6, 3e6, 333e3, -1, 1e6, [(0, 0, 64)], x)[0] == PC_IMPLICIT
    assert extract(orc, 16e6, 7e6, 10e3, -1, 1e6, [(0, 0, 64)], x)[0] == PC_IMPLICIT
    # 1.5 MHz comb from 1 MHz: Np = 64, No = 32 -> shifting
    assert extract(orc, 16e6, 1.5e6, 1e6, -1, 1e6, [(0, 0, 64)], x)[0] == PC_SHIFT
    # with 10 kHz resolution No doubles to 2048 -> implicit (pcal.cpp:158-160)
    assert extract(orc, 16e6, 1.5e6, 1e6, -1, 10e3, [(0, 0, 64)], x)[0] == PC_IMPLICIT
    assert kinds[(16e6, -1, 1e6)] == PC_DUMMY


@pytest.mark.parametrize("bw,offset,spacing", [(16e6, 10e3, 1e6), (8e6, 250e3, 1e6), (16e6, 10e3, 4e6)])
def test_shifting_equals_implicit_on_whole_periods(orc, bw, offset, spacing):
    """Both extractors are exact for a comb over whole periods of the mixer and
    the bins: the tones they return agree (pcal.cpp:590-650 vs 1037-1104)."""
    fs = 2 * bw
    per = int(fs // np.gcd(int(fs), int(offset)))
    n = per * 8
    data = tone_pattern(bw, offset, spacing, n, 7, slope_deg=11.0)
    _, a, _ = extract(orc, bw, spacing, offset, PC_SHIFT, 1e6, [(7, 0, n)], data)
    _, b, _ = extract(orc, bw, spacing, offset, PC_IMPLICIT, 1e6, [(7, 0, n)], data)
    m = min(len(a), len(b))
    assert m > 0
    assert np.max(np.abs(a[:m] - b[:m])) < 1e-4 * n


def test_direct_reference_method(orc):
    """The reference's slow direct method (mix by the first tone's offset,
    2 x ntones bins, DFT) and the implicit-shift extractor agree for a comb
    that fills the band (16 MHz, 1 MHz tones from 10 kHz)."""
    bw, off, sp = 16e6, 10e3, 1e6
    n = 3200 * 10
    data = tone_pattern(bw, off, sp, n, 0, slope_deg=7.0)
    _, fast, _ = extract(orc, bw, sp, off, PC_IMPLICIT, 1e6, [(0, 0, n)], data)
    out = np.zeros(2 * 64)
    nt = C.c_int32()
    orc.orc_pcal_direct(bw, sp, int(off), 0, data.ctypes.data_as(P(C.c_float)), n,
                        out.ctypes.data_as(P(C.c_double)), C.byref(nt))
    slow = out[:2 * nt.value:2] + 1j * out[1:2 * nt.value:2]
    assert nt.value == len(fast) == 16
    assert np.max(np.abs(fast - slow)) < 1e-4 * n
    assert np.allclose(np.degrees(np.angle(fast[:15])), -90 + 7.0 * np.arange(15), atol=0.05)


def test_bins_follow_absolute_sample_number(orc):
    """Linearity in the data and the absolute-sample binning: a block given
    at sample offset so contributes x[n] to bin (so + n) mod N, so splitting a
    stream into blocks at their own offsets changes nothing (the per-block
    adjustSampleOffset of Mode::process, mode.cpp:777)."""
    bw, off, sp = 8e6, 250e3, 1e6
    rng = np.random.default_rng(9)
    n = 50000
    x = rng.standard_normal(n).astype(np.float32)
    for kind in (PC_IMPLICIT, PC_SHIFT):
        _, whole, _ = extract(orc, bw, sp, off, kind, 1e6, [(123, 0, n)], x)
        cuts = [0, 4096, 9000, 20000, 33333, n]
        segs = [(123 + a, a, b - a) for a, b in zip(cuts[:-1], cuts[1:])]
        _, parts, _ = extract(orc, bw, sp, off, kind, 1e6, segs, x)
        assert np.max(np.abs(whole - parts)) < 1e-3 * np.sqrt(n)


def _job(**kw):
    base = dict(nstations=2, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=4,
                pcal_interval_mhz=1.0)
    base.update(kw)
    return JobSpec(**base)


@pytest.mark.parametrize("lib_name", ["oracle", "gpu"])
def test_config_tone_tables(lib_name):
    """configuration.cpp:1771-1845 worked by hand: USB band 8400.01-8416.01
    MHz has tones 8401..8416 MHz (16, first 990 kHz above the edge); an LSB
    band edge at 8400.01 MHz reaches down to 8384.01: tones 8400..8385 (16,
    first 10 kHz below the edge); a band edge on a tone moves the first tone
    a full interval away."""
    lib = fxcorr.load_oracle() if lib_name == "oracle" else fxcorr.load_gpu_lib()
    prefix = "orc_" if lib_name == "oracle" else "fxg_"
    cases = [(8400.01, 0, 16, 990000), (8400.01, 1, 16, 10000), (8400.0, 0, 16, 1000000),
             (8400.0, 1, 16, 1000000), (8400.5, 0, 16, 500000)]
    for lo, lsb, nt_exp, off_exp in cases:
        spec = _job(base_freq_mhz=lo, lsb=(lsb, lsb), nsubbands=1)
        job, keep = spec.build()
        nt, off = C.c_int32(), C.c_int32()
        check(lib, getattr(lib, prefix + "plan_pcal_tones")(C.byref(job), 0, 0, C.byref(nt), C.byref(off)), "tones")
        assert (nt.value, off.value) == (nt_exp, off_exp), (lo, lsb)


def test_pcal_layout_matches_oracle():
    """Result length and phase-cal offsets of the HIP library's planner equal
    the oracle's restatement of populateResultLengths."""
    lib = fxcorr.load_gpu_lib()
    orc = fxcorr.load_oracle()
    for spec in [_job(), _job(nstations=3, nsubbands=4, pcal_interval_mhz=5.0, base_freq_mhz=8400.25),
                 _job(lsb=(1, 0), zoom=((0, 16, 32),)), _job(pcal_interval_mhz=0.0)]:
        job, keep = spec.build()
        a, b = fxcorr.PlanInfo(), fxcorr.PlanInfo()
        check(lib, lib.fxg_plan(C.byref(job), C.byref(a)), "plan")
        check(orc, orc.orc_plan(C.byref(job), C.byref(b)), "plan")
        assert a.result_length == b.result_length
        for d in range(spec.nstations):
            x, y = C.c_int64(), C.c_int64()
            check(lib, lib.fxg_plan_pcal_offset(C.byref(job), d, C.byref(x)), "off")
            check(orc, orc.orc_plan_pcal_offset(C.byref(job), d, C.byref(y)), "off")
            assert x.value == y.value
        if spec.pcal_interval_mhz > 0:
            x = C.c_int64()
            check(lib, lib.fxg_plan_pcal_offset(C.byref(job), 0, C.byref(x)), "off")
            # every recorded band carries its frequency's tones after the AC weights
            ntot = 0
            for d in range(spec.nstations):
                for sb in range(spec.nsubbands):
                    nt, off = C.c_int32(), C.c_int32()
                    check(lib, lib.fxg_plan_pcal_tones(C.byref(job), d, sb, C.byref(nt), C.byref(off)), "t")
                    ntot += spec.npol * nt.value
            assert a.result_length - x.value == ntot
