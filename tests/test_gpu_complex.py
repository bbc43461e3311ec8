"""Complex-sampled VDIF (single and double sideband, USB and LSB, 2- and
8-bit) through the GPU against the oracle: the FP32 path within 1e-5 of the
FP32 oracle (spectra and visibilities), reduced precision against the FP64
oracle, with a bad frame (a complex 8-bit decoder stops at it) and a
flagged block; one band of 2-bit complex samples, whose unpacker erases a
sample inside each block's window."""
import ctypes as C

import numpy as np
import pytest

from fxcorr import JobSpec
from fxsynth import make_subint
from fxtestlib import GpuJob, oracle_block_info, oracle_results, oracle_spectra, rel_err

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def cspec(lsb, dsb, nbits=2, **kw):
    base = dict(nstations=4, nsubbands=2, npol=2, bw_mhz=16.0, nchan=512, blockspersend=12, nbits=nbits,
                lsb=[lsb, lsb], sampling="complex_dsb" if dsb else "complex", base_freq_mhz=8400.25,
                clockoffset_us=[0.011, -0.02, 0.0, 0.004])
    base.update(kw)
    return JobSpec(**base)


def cdata(spec, seed=5):
    rng = np.random.default_rng(seed)
    return make_subint(spec, seed=seed, rho=0.05, delays_us=rng.uniform(-20, 20, spec.nstations),
                       rates_us_per_s=rng.uniform(-1, 1, spec.nstations),
                       invalid_frames={1: [2]}, invalid_blocks={2: [3]})


MODES = [(0, 0), (1, 0), (0, 1), (1, 1)]


@pytest.mark.parametrize("nbits", [2, 8])
@pytest.mark.parametrize("lsb,dsb", MODES)
def test_complex_visibilities_fp32(gpu, oracle, lsb, dsb, nbits):
    lib, ctx = gpu
    spec = cspec(lsb, dsb, nbits)
    job, keep = spec.build()
    sd = cdata(spec)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd, fp64=False, nthreads=8)
    assert rel_err(got, ref) < FP32_TOL, rel_err(got, ref)
    # spectra of one block of station 1 (whose frame 2 is bad) as Mode::getFreqs
    for blk in (0, 5):
        sp = oracle_spectra(oracle, job, sd, 1, blk, spec.nrecbands, spec.nchan)
        for b in range(spec.nrecbands):
            gs = g.spectrum(1, b, blk, spec.nchan)
            assert rel_err(gs, sp[b]) < FP32_TOL, (blk, b, rel_err(gs, sp[b]))
    g.close()


@pytest.mark.parametrize("nchan", [128, 1024, 4096, 8192, 16384])
def test_complex_fft_lengths_and_order2(gpu, oracle, nchan):
    lib, ctx = gpu
    spec = cspec(0, 1, 2, nchan=nchan, fringe_order=2, nstations=3, blockspersend=6)
    job, keep = spec.build()
    sd = cdata(spec, seed=8)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd, fp64=False, nthreads=8)
    assert rel_err(got, ref) < FP32_TOL, rel_err(got, ref)
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp16", 1e-3), ("bf16", 8e-3)])
@pytest.mark.parametrize("lsb,dsb", [(0, 0), (1, 1)])
def test_complex_mixed_vs_fp64(gpu, oracle, prec, tol, lsb, dsb):
    lib, ctx = gpu
    spec = cspec(lsb, dsb, 2, xmac=prec)
    job, keep = spec.build()
    sd = cdata(spec, seed=9)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd, fp64=True, nthreads=8)
    assert rel_err(got, ref) < tol, rel_err(got, ref)
    g.close()


@pytest.mark.parametrize("nchan", [8192, 16384])
def test_complex_8bit_largest(gpu, oracle, nchan):
    """8-bit complex samples at the largest complex transforms; at 16384
    channels the staged stream words share the tail of the FFT buffer."""
    lib, ctx = gpu
    spec = cspec(1, 0, 8, nchan=nchan, nstations=3, blockspersend=5)
    job, keep = spec.build()
    sd = cdata(spec, seed=12)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd, fp64=False, nthreads=8)
    assert rel_err(got, ref) < FP32_TOL, rel_err(got, ref)
    sp = oracle_spectra(oracle, job, sd, 2, 4, spec.nrecbands, spec.nchan)
    for b in range(spec.nrecbands):
        assert rel_err(g.spectrum(2, b, 4, spec.nchan), sp[b]) < FP32_TOL, b
    g.close()


@pytest.mark.parametrize("dsb,order", [(0, 1), (1, 2), (0, 2)])
def test_complex_single_band_2bit(gpu, oracle, dsb, order):
    """One band of 2-bit complex samples: 4 bits per sample-time, sample
    granularity 2, so Mk5Mode::unpack erases by float index inside the
    window (sample L/2 for an even block start, the imaginary part of sample
    L/2 - 1 for an odd one, mk5mode.cpp:121-140) and moves the weight.
    Block geometry and weights exactly, spectra and visibilities (phase cal
    included) within 1e-5 of the FP32 oracle, blocks of both start parities,
    with a bad frame and a flagged block."""
    lib, ctx = gpu
    spec = cspec(0, dsb, 2, nsubbands=1, npol=1, nstations=3, blockspersend=40, fringe_order=order,
                 pcal_interval_mhz=1.0, base_freq_mhz=8400.01)
    assert spec.nrecbands * spec.unitbits == 4
    job, keep = spec.build()
    # whole-sample delays of 48, -81 and 113 samples: even and odd block
    # starts; frame 3 of stations 0 (even starts) and 1 (odd starts) is bad,
    # from inside block 31, whose erased sample falls in it
    sd = make_subint(spec, seed=14, rho=0.05, delays_us=[3.0, -5.0625, 7.0625], rates_us_per_s=[0.3, -0.5, 0.2],
                     invalid_frames={0: [3], 1: [3]}, invalid_blocks={2: [3]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    parity = set()
    for ds in range(spec.nstations):
        for blk in range(spec.blockspersend):
            a = g.block_info(ds, blk)
            b = oracle_block_info(oracle, job, sd, ds, blk)
            for f in ("status", "nearestsample", "integerdelay"):
                assert getattr(a, f) == getattr(b, f), (ds, blk, f)
            if b.status == 0:
                parity.add(b.nearestsample % 2)
                for f in ("weight", "fracsampleerror"):
                    assert getattr(a, f) == getattr(b, f), (ds, blk, f, getattr(a, f), getattr(b, f))
    assert parity == {0, 1}
    ref = oracle_results(oracle, job, sd, fp64=False, nthreads=8)
    assert rel_err(got, ref) < FP32_TOL, rel_err(got, ref)
    for ds, blk in ((0, 0), (1, 5), (1, 31), (2, 7)):
        sp = oracle_spectra(oracle, job, sd, ds, blk, 1, spec.nchan)
        assert rel_err(g.spectrum(ds, 0, blk, spec.nchan), sp[0]) < FP32_TOL, (ds, blk)
    g.close()
