"""Phase-cal extraction in the oracle, pinned by the reference's own test
pattern (mpifxcorr/src/pcal.cpp:1532-1640: tones with a 5 degree phase slope,
expected phase -90 deg + 5 deg per tone), by the reference's slow direct
method (PCal::extractAndIntegrate_reference, pcal.cpp:253-331), by the
agreement of the shifting and implicit-shift extractors on whole periods,
and by the tone tables of configuration.cpp:1771-1845 worked by hand.  The
complex-sampling extractors (pcal.cpp:707-925, 1107-1285) are pinned by
closed-form complex combs (single and double sideband) and by the direct sum
over samples that their bin folding computes."""
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


# ---- complex samples (PCalExtractorComplex, PCalExtractorComplexImplicitShift)
PC_CSHIFT, PC_CIMPLICIT = 4, 5


@pytest.fixture(scope="module")
def orc_c(orc):
    orc.orc_pcal_extract_c.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                       P(C.c_int64), P(C.c_int64), P(C.c_int64), P(C.c_float), P(C.c_float),
                                       P(C.c_int32), P(C.c_int64)]
    return orc


def extract_c(orc, bw, spacing, offset, kind, minres, dsb, segs, x):
    """x: complex samples; segs: (sampleoffset, start, length) in complex samples."""
    n = len(segs)
    so = (C.c_int64 * n)(*[s[0] for s in segs])
    st = (C.c_int64 * n)(*[s[1] for s in segs])
    ln = (C.c_int64 * n)(*[s[2] for s in segs])
    data = np.empty(2 * len(x), np.float32)
    data[0::2], data[1::2] = x.real, x.imag
    out = np.zeros(2 * 4096, np.float32)
    nt, ns = C.c_int32(), C.c_int64()
    k = orc.orc_pcal_extract_c(bw, spacing, int(offset), kind, minres, int(dsb), n, so, st, ln,
                               data.ctypes.data_as(P(C.c_float)), out.ctypes.data_as(P(C.c_float)),
                               C.byref(nt), C.byref(ns))
    assert k >= 0
    z = out[:2 * nt.value].reshape(-1, 2).astype(np.float64)
    return k, z[:, 0] + 1j * z[:, 1], ns.value


def comb_c(freqs_hz, fs, n, so, phases):
    """sum_k exp(i (2 pi f_k (so + m) / fs + phase_k)), m < n, at absolute sample so + m."""
    m = so + np.arange(n, dtype=np.float64)
    x = np.zeros(n, np.complex128)
    for f, ph in zip(freqs_hz, phases):
        x += np.exp(1j * (2 * np.pi * np.mod(f * m, fs) / fs + ph))
    return x.astype(np.complex64)


def test_complex_getnew_choices(orc_c):
    """PCal::getNew for complex data (pcal.cpp:186-191): fs = bandwidth; the
    offset period (doubled to the minimum resolution) a multiple of the comb
    period -> implicit shift, else the shifting extractor; no tone -> dummy."""
    x = np.zeros(64, np.complex64)
    k = lambda bw, off, sp, res=1e6: extract_c(orc_c, bw, sp, off, -1, res, 0, [(0, 0, 64)], x)[0]
    # No = 16e6 / 0.5e6 = 32, Np = 16 -> implicit
    assert k(16e6, 0.5e6, 1e6) == PC_CIMPLICIT
    # No = 8 (16e6 / (2 x 8) = 1 MHz stops the doubling), Np = 16 -> shifting
    assert k(16e6, 2e6, 3e6) == PC_CSHIFT
    # a 10 kHz resolution doubles No to 1024 -> implicit
    assert k(16e6, 2e6, 3e6, 10e3) == PC_CIMPLICIT
    assert k(16e6, -1, 1e6) == PC_DUMMY
    # tones past the band: none fit -> dummy
    assert k(1e6, 2e6, 1e6) == PC_DUMMY


@pytest.mark.parametrize("dsb", [0, 1])
@pytest.mark.parametrize("bw,offset,spacing,kind,bins", [
    # tones 0.5, 1.5, ... 15.5 MHz; N = 32 bins of 0.5 MHz: tone k in bin 1 + 2k
    (16e6, 0.5e6, 1e6, PC_CIMPLICIT, [1 + 2 * k for k in range(16)]),
    # tones 2, 5, 8, 11, 14 MHz mixed down by 2 MHz; N = 16 bins of 1 MHz: bin 3k
    (16e6, 2e6, 3e6, PC_CSHIFT, [3 * k for k in range(5)]),
    # the shifting extractor on the implicit case: N = 16, mixer period 32, bin k
    (16e6, 0.5e6, 1e6, PC_CSHIFT, list(range(16))),
])
def test_complex_tones_closed_form(orc_c, bw, offset, spacing, kind, bins, dsb):
    """A complex comb exp(i (2 pi f_k t + k slope)) locked to the absolute
    sample number comes back as count x exp(i k slope) in tone k: the
    extractor mixes the first tone to DC (shifting), bins by absolute sample
    number mod N and reads DFT bin [offset bin +] k x step.  Double sideband
    reads the DFT reversed (pcal.cpp:906-913, 1269-1276): a comb at bins
    N - 1 - bin_k then lands in tone k."""
    fs = bw
    so, n = 11, 32 * 200
    N = {PC_CIMPLICIT: 32, PC_CSHIFT: 16}[kind]
    slope = np.radians(7.0)
    if kind == PC_CSHIFT:
        f = [offset + (N - 1 - b if dsb else b) * fs / N for b in bins]
    else:
        f = [(N - 1 - b if dsb else b) * fs / N for b in bins]
    x = comb_c(f, fs, n + so, 0, slope * np.arange(len(bins)))
    k, tones, count = extract_c(orc_c, bw, spacing, offset, kind, 1e6, dsb,
                                [(so, so, n // 2), (so + n // 2, so + n // 2, n // 2)], x)
    assert k == kind and count == n and len(tones) == len(bins)
    expect = count * np.exp(1j * slope * np.arange(len(bins)))
    assert np.max(np.abs(tones - expect)) < 2e-4 * count


@pytest.mark.parametrize("kind,bw,offset,spacing", [(PC_CIMPLICIT, 8e6, 250e3, 1e6),
                                                     (PC_CSHIFT, 16e6, 2e6, 3e6)])
@pytest.mark.parametrize("dsb", [0, 1])
def test_complex_equals_direct_sum(orc_c, kind, bw, offset, spacing, dsb):
    """On random complex samples in blocks at their own sample offsets (the
    per-block adjustSampleOffset of mode.cpp:777), tone k equals the direct
    sum over samples of x[m] exp(-2 pi i (offset (shifting) / fs + b_k / N)
    (so + m)), b_k the tone's DFT bin (N - 1 - b_k for double sideband): the
    bin folding, the mixer table and the tail handling change nothing."""
    fs = bw
    rng = np.random.default_rng(31)
    n = 40000
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    cuts = [0, 4096, 9001, 20000, 33333, n]
    so0 = 123
    segs = [(so0 + a, a, b - a) for a, b in zip(cuts[:-1], cuts[1:])]
    k, tones, _ = extract_c(orc_c, bw, spacing, offset, kind, 1e6, dsb, segs, x)
    assert k == kind
    if kind == PC_CIMPLICIT:
        N = int(fs // np.gcd(int(fs), int(offset)))
        while fs / (2 * N) > 1e6:
            N *= 2
        first, mix = int(np.floor(N * offset / fs)), 0.0
    else:
        N = int(fs // np.gcd(int(fs), int(spacing)))
        while fs / (2 * N) > 1e6:
            N *= 2
        first, mix = 0, offset
    step = int(np.floor(N * spacing / fs))
    t = so0 + np.arange(n, dtype=np.float64)
    xd = x.astype(np.complex128)
    for j, v in enumerate(tones):
        b = first + j * step
        if b >= N:
            assert v == 0
            continue
        b = N - 1 - b if dsb else b
        ph = np.mod(mix * t, fs) / fs + np.mod(b * t, N) / N
        ref = np.sum(xd * np.exp(-2j * np.pi * ph))
        assert abs(v - ref) < 2e-3 * np.sqrt(n), (j, v, ref)
