"""GPU parity: the HIP library against the CPU oracle on the same seeded
inputs.  Tolerances (||dV|| / ||V||):

  - unpack, block geometry, weights: bit for bit;
  - FP32 spectra and visibilities vs the FP32 oracle: 1e-5 (FP32_TOL);
  - fp16 XMAC vs the FP64 oracle: 1e-3; bf16: 8e-3 (MIXED_TOL: an 11- or
    8-bit significand per spectrum value);
  - fp8 e4m3 XMAC vs the FP64 oracle: 0.08 on the cross products, 0.04 on
    the autocorrelations (FP8_TOL: 3 mantissa bits).

Only fp16 meets a 1e-3 visibility error; bf16 and fp8 are checked against
their own significands."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import (FXG_BLANKER_NONE, FXG_FORMAT_MARK5B, FXG_FORMAT_VDIF, Datastream, JobSpec, check)
from fxsynth import make_subint
from fxtestlib import (GpuJob, oracle_block_info, oracle_results, oracle_spectra, plan_info, rel_err,
                       region_slices)

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5      # plain FP32 path vs the FP32 oracle, ||dV||/||V||
# reduced-precision XMAC vs the FP64 oracle, ||dV||/||V||: the spectra are
# rounded to an 11-bit (fp16) or 8-bit (bf16) significand, so the expected
# relative visibility error is ~2^-11 (fp16) and ~2^-8 (bf16)
MIXED_TOL = {"fp16": 1e-3, "bf16": 8e-3}


def random_frames(rng, fmt, nchan, nbits, nframes, framebytes, headerbytes, corrupt=()):
    data = rng.integers(0, 256, size=(nframes, framebytes), dtype=np.uint8)
    fill = np.frombuffer(np.array([0x11223344, 0x11223344], dtype=np.uint32).tobytes(), dtype=np.uint8)
    for f in range(nframes):
        if fmt == FXG_FORMAT_VDIF:
            w = np.zeros(8, dtype=np.uint32)
            w[0] = 1000 + f // 100
            w[2] = (framebytes // 8) | (int(np.log2(nchan)) << 24)
            w[3] = (nbits - 1) << 26
            data[f, :32] = w.view(np.uint8)
        else:
            w = np.zeros(4, dtype=np.uint32)
            w[0] = 0xABADDEED
            w[1] = f & 0x7FFF
            data[f, :16] = w.view(np.uint8)
    for f, kind in corrupt:
        pl = headerbytes
        if kind == "invalid":
            if fmt == FXG_FORMAT_VDIF:
                data[f, 3] |= 0x80
            else:
                data[f, 5] |= 0x80
        elif kind == "zerolen":
            data[f, 8:12] = 0
        elif kind == "fillstart":
            data[f, pl:pl + 8] = fill
        elif kind == "fillend":
            data[f, framebytes - 8:framebytes] = fill
        elif kind == "filltail":       # fill pattern over the last part of the payload
            n = (framebytes - pl) // 8
            k = n // 3
            data[f, framebytes - 8 * k:framebytes] = np.tile(fill, k)
        elif kind == "fillhead":
            n = (framebytes - pl) // 8
            k = n // 4
            data[f, pl:pl + 8 * k] = np.tile(fill, k)
    return data.reshape(-1)


def ds_desc(fmt, nchan, nbits, framebytes, headerbytes, blanker=0):
    d = Datastream()
    d.format = fmt
    d.nbits = nbits
    d.framebytes = framebytes
    d.headerbytes = headerbytes
    d.blanker = blanker
    d.nrecfreqs = 1
    d.nrecbands = nchan
    d.nzoombands = 0
    d.recfreq_index[0] = 0
    return d


UNPACK_CASES = [
    # fmt, nchan, nbits, framebytes, headerbytes
    (FXG_FORMAT_VDIF, 1, 2, 8032, 32),
    (FXG_FORMAT_VDIF, 2, 2, 8032, 32),
    (FXG_FORMAT_VDIF, 4, 2, 1056, 32),
    (FXG_FORMAT_VDIF, 16, 2, 8032, 32),
    (FXG_FORMAT_VDIF, 32, 2, 8032, 32),
    (FXG_FORMAT_VDIF, 1, 8, 1056, 32),
    (FXG_FORMAT_VDIF, 2, 8, 1056, 32),
    (FXG_FORMAT_VDIF, 4, 8, 8032, 32),
    (FXG_FORMAT_VDIF, 16, 8, 8032, 32),
    (FXG_FORMAT_MARK5B, 4, 2, 10016, 16),
    (FXG_FORMAT_MARK5B, 8, 2, 10016, 16),
    (FXG_FORMAT_MARK5B, 16, 2, 10016, 16),
]


@pytest.mark.parametrize("case", UNPACK_CASES, ids=lambda c: f"fmt{c[0]}-{c[1]}ch-{c[2]}bit")
@pytest.mark.parametrize("blanker", [0, FXG_BLANKER_NONE])
def test_unpack_bitexact(gpu, oracle, case, blanker):
    lib, ctx = gpu
    fmt, nchan, nbits, fb, hb = case
    rng = np.random.default_rng(nchan * 100 + nbits + 7 * blanker + fmt)
    nframes = 24
    corrupt = [(3, "invalid"), (7, "fillstart"), (11, "fillend"), (15, "filltail"), (18, "fillhead")]
    if fmt == FXG_FORMAT_VDIF:
        corrupt.append((21, "zerolen"))
    data = random_frames(rng, fmt, nchan, nbits, nframes, fb, hb, corrupt)
    d = ds_desc(fmt, nchan, nbits, fb, hb, blanker)
    framesamples = (fb - hb) * 8 // (nchan * nbits)
    dptr = C.c_void_p()
    check(lib, lib.fxg_device_alloc(ctx, data.size, C.byref(dptr)), "alloc")
    check(lib, lib.fxg_memcpy_h2d(ctx, dptr, data.ctypes.data, data.size), "h2d")
    total = nframes * framesamples
    offsets = [0, 1, 5, framesamples - 3, 2 * framesamples + 17, 3 * framesamples - 100,
               7 * framesamples + 33, 15 * framesamples - 5, 20 * framesamples + 1, total - 600]
    for nsamp in (512, 4096):
        outdev = C.c_void_p()
        check(lib, lib.fxg_device_alloc(ctx, 4 * nchan * nsamp, C.byref(outdev)), "alloc")
        try:
            for off in offsets:
                if off < 0:
                    continue
                ref = np.zeros(nchan * nsamp, dtype=np.float32)
                wref = C.c_float()
                check(oracle, oracle.orc_unpack(C.byref(d), data.ctypes.data, data.size, off, nsamp,
                                                ref.ctypes.data, C.byref(wref)), "orc_unpack")
                w = C.c_float()
                check(lib, lib.fxg_unpack(ctx, C.byref(d), dptr, data.size, off, nsamp, outdev, C.byref(w)),
                      "fxg_unpack")
                got = np.zeros_like(ref)
                check(lib, lib.fxg_memcpy_d2h(ctx, got.ctypes.data, outdev, got.nbytes), "d2h")
                assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"unpack mismatch at offset {off}"
                assert w.value == wref.value, f"weight {w.value} != {wref.value} at offset {off}"
        finally:
            lib.fxg_device_free(ctx, outdev)
    lib.fxg_device_free(ctx, dptr)


def c1_spec(**kw):
    base = dict(nstations=4, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=48)
    base.update(kw)
    return JobSpec(**base)


def c2_spec(**kw):
    base = dict(nstations=10, nsubbands=8, npol=2, bw_mhz=32.0, nchan=1024, blockspersend=6,
                base_freq_mhz=6000.0)
    base.update(kw)
    return JobSpec(**base)


def bands_spec(nbands, **kw):
    """Frames of 32 or 64 bands: 2 or 4 words per sample-time (register corner turn)."""
    base = dict(nstations=3, nsubbands=nbands // 2, npol=2, bw_mhz=16.0, nchan=128, blockspersend=10)
    base.update(kw)
    return JobSpec(**base)


SPECS = {"c1": c1_spec, "c2": c2_spec, "b32": lambda **kw: bands_spec(32, **kw),
         "b64": lambda **kw: bands_spec(64, **kw)}


def make_data(spec, seed=5, rho=0.05, **kw):
    rng = np.random.default_rng(seed + 1000)
    delays = rng.uniform(-20.0, 20.0, spec.nstations)
    rates = rng.uniform(-1.0, 1.0, spec.nstations)
    return make_subint(spec, seed=seed, rho=rho, delays_us=delays, rates_us_per_s=rates, **kw)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_block_geometry_exact(gpu, oracle, order):
    lib, ctx = gpu
    spec = c1_spec(fringe_order=order)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={1: [2, 3]}, invalid_blocks={2: [5, 6, 40]})
    g = GpuJob(lib, ctx, job)
    g.process(sd)
    for ds in range(spec.nstations):
        for blk in range(spec.blockspersend):
            a = g.block_info(ds, blk)
            b = oracle_block_info(oracle, job, sd, ds, blk)
            for f in ("status", "nearestsample", "integerdelay", "intwalltime"):
                assert getattr(a, f) == getattr(b, f), (ds, blk, f)
            if b.status == 0:
                for f in ("weight", "fracsampleerror", "a", "b", "c", "fracwalltime"):
                    assert getattr(a, f) == getattr(b, f), (ds, blk, f, getattr(a, f), getattr(b, f))
    g.close()


@pytest.mark.parametrize("mk", ["c1", "c2"])
def test_spectra_fp32(gpu, oracle, mk):
    lib, ctx = gpu
    spec = c1_spec() if mk == "c1" else c2_spec()
    job, keep = spec.build()
    sd = make_data(spec)
    g = GpuJob(lib, ctx, job)
    g.process(sd, keep_taps=True)
    chunk = g.info.chunk_blocks
    last0 = ((spec.blockspersend - 1) // chunk) * chunk
    for ds in range(spec.nstations):
        for blk in (last0, spec.blockspersend - 1):
            ref = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, spec.nchan)
            for band in range(spec.nrecbands):
                got = g.spectrum(ds, band, blk, spec.nchan)
                assert rel_err(got, ref[band]) < FP32_TOL, (ds, band, blk, rel_err(got, ref[band]))
    g.close()


@pytest.mark.parametrize("mk", ["c1", "c2", "b32", "b64"])
def test_visibilities_fp32(gpu, oracle, mk):
    lib, ctx = gpu
    spec = SPECS[mk]()
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]}, invalid_blocks={1: [2]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd)
    info = plan_info(oracle, "orc_", job)
    assert g.info.result_length == info.result_length
    xend, a0, a1 = region_slices(oracle, job, info)
    assert rel_err(got[:xend], ref[:xend]) < FP32_TOL
    assert rel_err(got[a0:a1], ref[a0:a1]) < FP32_TOL
    # weights are sums of the same FP32 terms in block order
    assert np.array_equal(got[xend:a0].view(np.float32), ref[xend:a0].view(np.float32))
    assert np.array_equal(got[a1:].view(np.float32), ref[a1:].view(np.float32))
    g.close()


def wide_spec(**kw):
    """40 inputs (20 stations x 2 pol): two 32-row tiles on the matrix cores."""
    base = dict(nstations=20, nsubbands=1, npol=2, bw_mhz=16.0, nchan=256, blockspersend=40)
    base.update(kw)
    return JobSpec(**base)


@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("mk", ["c1", "c2"])
def test_visibilities_fp32_fringe_orders(gpu, oracle, mk, order):
    """Post-F (order 0) and quadratic (order 2) fringe rotation, FP32 path,
    with a curved delay model so the quadratic term matters."""
    lib, ctx = gpu
    spec = SPECS[mk](fringe_order=order, looffset_hz=[12.5] * SPECS[mk]().nsubbands)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]}, invalid_blocks={1: [2]})
    for i, st in enumerate(sd.stations):
        st.interpolator = st.interpolator + np.array([3e-5 * (i - 1.5), 0.0, 0.0])
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    assert rel_err(got[:xend], ref[:xend]) < FP32_TOL
    assert rel_err(got[a0:a1], ref[a0:a1]) < FP32_TOL
    assert np.array_equal(got[xend:a0].view(np.float32), ref[xend:a0].view(np.float32))
    for ds in range(spec.nstations):
        blk = spec.blockspersend - 1
        refs = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, spec.nchan)
        for band in range(spec.nrecbands):
            assert rel_err(g.spectrum(ds, band, blk, spec.nchan), refs[band]) < FP32_TOL
    g.close()


@pytest.mark.parametrize("nchan", [8192, 16384, 32768])
@pytest.mark.parametrize("nbits", [2, 8])
@pytest.mark.parametrize("xmac", ["fp32", "fp16"])
def test_visibilities_largest_fft(gpu, oracle, nchan, nbits, xmac):
    """8192 channels: the 16384-point F stage (1024-thread workgroups; 8-bit
    stream words sharing the FFT buffer); 16384 and 32768 channels: the
    two-kernel 32768- and 65536-point transforms through the scratch."""
    lib, ctx = gpu
    spec = JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=64.0, nchan=nchan, blockspersend=5, nbits=nbits,
                   xmac=xmac)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [3]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd, fp64=(xmac != "fp32"))
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    tol = FP32_TOL if xmac == "fp32" else MIXED_TOL[xmac]
    assert rel_err(got[:xend], ref[:xend]) < tol
    assert rel_err(got[a0:a1], ref[a0:a1]) < tol
    if xmac == "fp32":
        for ds in range(spec.nstations):
            refs = oracle_spectra(oracle, job, sd, ds, 4, spec.nrecbands, spec.nchan)
            for band in range(spec.nrecbands):
                assert rel_err(g.spectrum(ds, band, 4, spec.nchan), refs[band]) < FP32_TOL
    g.close()


@pytest.mark.parametrize("avg,zoom", [(4, [(0, 37, 64)]), (256, [])])
def test_visibilities_fp32_channel_averaging(gpu, oracle, avg, zoom):
    """Channel averaging of cross products (core.cpp:1788-1802) and of
    autocorrelations (mode.cpp:1399-1424), with a zoom band whose offset is
    not a multiple of the averaging."""
    lib, ctx = gpu
    spec = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=12, zoom=zoom,
                   chans_to_avg=avg)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd)
    info = plan_info(oracle, "orc_", job)
    assert g.info.result_length == info.result_length
    xend, a0, a1 = region_slices(oracle, job, info)
    assert rel_err(got[:xend], ref[:xend]) < FP32_TOL
    assert rel_err(got[a0:a1], ref[a0:a1]) < FP32_TOL
    assert np.array_equal(got[xend:a0].view(np.float32), ref[xend:a0].view(np.float32))
    g.close()


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("mk", ["c1", "c2", "wide"])
def test_xmac_streaming_equals_register_staging(gpu, prec, mk):
    """The direct-to-LDS XMAC accumulates the same products in the same order
    as the register-staged one: identical visibilities (block counts that
    leave a partial stage, input counts that leave zero rows)."""
    lib, ctx = gpu
    spec = {"c1": c1_spec, "c2": c2_spec, "wide": wide_spec}[mk](xmac=prec, blockspersend=37)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]})
    outs = []
    for on in (0, 1):
        g = GpuJob(lib, ctx, job)
        check(lib, lib.fxg_job_set_xmac_streaming(g.h, on), "streaming")
        outs.append(g.process(sd))
        g.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("mk", ["c1", "c2", "wide", "x80"])
def test_fp32_xmac_matrix_cores_and_valu_tiles(gpu, oracle, mk):
    """FP32 spectra cross-multiplied on the FP32 matrix cores (the streamed
    schedule, the default) and on the VALU register tiles (streaming off):
    both within 1e-5 of the FP32 oracle, cross products and
    autocorrelations, for 8, 20, 40 and 80 inputs (the last: a 64-input
    chunk, a partial chunk and their cross pair) over 37 blocks."""
    lib, ctx = gpu
    mkspec = {"c1": c1_spec, "c2": c2_spec, "wide": wide_spec,
              "x80": lambda **kw: JobSpec(nstations=40, nsubbands=1, npol=2, bw_mhz=16.0, nchan=128, **kw)}[mk]
    spec = mkspec(blockspersend=37)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]})
    ref = oracle_results(oracle, job, sd, nthreads=8)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    for on in (1, 0):
        g = GpuJob(lib, ctx, job)
        check(lib, lib.fxg_job_set_xmac_streaming(g.h, on), "streaming")
        got = g.process(sd)
        g.close()
        assert rel_err(got[:xend], ref[:xend]) < FP32_TOL, (on, rel_err(got[:xend], ref[:xend]))
        assert rel_err(got[a0:a1], ref[a0:a1]) < FP32_TOL, on


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("nst", [16, 32])
def test_xmac_streaming_many_stages(gpu, prec, nst):
    """One grid row over 117 blocks of 4096 channels: the streamed kernel's
    ring of 8-block stages (8 buffers for 32 inputs, 4 for 64) turns 15
    times, the last stage partial, and its visibilities equal the
    register-staged kernel's bit for bit."""
    lib, ctx = gpu
    spec = JobSpec(nstations=nst, nsubbands=1, npol=2, bw_mhz=64.0, nchan=4096, blockspersend=117, xmac=prec)
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={1: [2]})
    outs = []
    for on in (0, 1):
        g = GpuJob(lib, ctx, job)
        check(lib, lib.fxg_job_set_xmac_streaming(g.h, on), "streaming")
        outs.append(g.process(sd))
        g.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("nst", [40, 75])
def test_xmac_above_64_inputs(gpu, oracle, prec, nst):
    """More than 64 inputs on the matrix cores: 40 stations x 2 pol (80
    inputs: a full and a partial chunk of 64, one chunk pair) and 75 x 2
    (150 inputs: three chunks, three chunk pairs), every baseline and
    autocorrelation against the FP64 oracle within the significand rounding;
    the result does not depend on the streaming switch (no register-staged
    kernel above 64 inputs)."""
    lib, ctx = gpu
    spec = JobSpec(nstations=nst, nsubbands=1, npol=2, bw_mhz=16.0, nchan=128, blockspersend=100, xmac=prec)
    job, keep = spec.build()
    sd = make_data(spec, rho=0.1)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    g.close()
    ref = oracle_results(oracle, job, sd, fp64=True, nthreads=8)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    assert rel_err(got[:xend], ref[:xend]) < MIXED_TOL[prec], rel_err(got[:xend], ref[:xend])
    assert rel_err(got[a0:a1], ref[a0:a1]) < MIXED_TOL[prec]
    g = GpuJob(lib, ctx, job)
    check(lib, lib.fxg_job_set_xmac_streaming(g.h, 0), "streaming")
    again = g.process(sd)
    g.close()
    assert np.array_equal(again.view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
@pytest.mark.parametrize("mk", ["c1", "c2", "wide"])
def test_visibilities_mixed_vs_fp64(gpu, oracle, prec, mk):
    lib, ctx = gpu
    spec = {"c1": c1_spec, "c2": c2_spec, "wide": wide_spec}[mk](xmac=prec)
    job, keep = spec.build()
    sd = make_data(spec)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd, fp64=True)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    e = rel_err(got[:xend], ref[:xend])
    assert e < MIXED_TOL[prec], e
    assert rel_err(got[a0:a1], ref[a0:a1]) < MIXED_TOL[prec]
    g.close()


def _uneven_cases():
    from fxsynth import uneven_workload_specs
    out = []
    for spec, label in uneven_workload_specs(xmac="fp32", max_blocks=3):
        if spec.nchan >= 4096:          # the oracle's cost grows with baselines x channels
            spec.nstations = min(spec.nstations, 6)
        out.append(pytest.param(spec, id=label))
    return out


@pytest.mark.parametrize("spec", _uneven_cases())
def test_uneven_workload_jobs(gpu, oracle, spec):
    """The six uneven jobs of the batched workload (64..32768 channels, 2-
    and 8-bit, 4..16 subbands, two zoom bands each), FP32 path vs the oracle
    and fp16 vs the FP64 oracle."""
    import dataclasses
    lib, ctx = gpu
    sd = make_data(spec, invalid_frames={1: [2]})
    for xmac in ("fp32", "fp16"):
        sp = dataclasses.replace(spec, xmac=xmac)
        job, keep = sp.build()
        g = GpuJob(lib, ctx, job)
        got = g.process(sd)
        ref = oracle_results(oracle, job, sd, fp64=(xmac != "fp32"), nthreads=8)
        info = plan_info(oracle, "orc_", job)
        assert g.info.result_length == info.result_length
        xend, a0, a1 = region_slices(oracle, job, info)
        tol = FP32_TOL if xmac == "fp32" else MIXED_TOL[xmac]
        assert rel_err(got[:xend], ref[:xend]) < tol, (xmac, rel_err(got[:xend], ref[:xend]))
        assert rel_err(got[a0:a1], ref[a0:a1]) < tol
        g.close()


# phase-cal configurations: the implicit-shift extractor (fractional-MHz band
# edge), the shifting extractor (No % Np != 0), lower sideband, and a comb
# shifted onto DC (the first-tone-at-DC extractor); complex samples (sample
# rate = bandwidth): the complex implicit-shift extractor (1600 bins), the
# complex shifting extractor (16 bins, 2 MHz mixer), lower sideband, 8-bit,
# and double sideband (the DFT read reversed)
PCAL_CASES = {
    "implicit": dict(base_freq_mhz=8400.01, pcal_interval_mhz=1.0),
    "shifting": dict(base_freq_mhz=8401.0, pcal_interval_mhz=3.0),
    "lsb": dict(base_freq_mhz=8400.01, pcal_interval_mhz=1.0, lsb=(1, 1)),
    "dc": dict(base_freq_mhz=8400.0, pcal_interval_mhz=3.0, pcal_base_mhz=-3.0, nsubbands=1),
    "c_implicit": dict(base_freq_mhz=8400.01, pcal_interval_mhz=1.0, sampling="complex"),
    "c_shifting": dict(base_freq_mhz=8401.0, pcal_interval_mhz=3.0, sampling="complex"),
    "c_lsb": dict(base_freq_mhz=8400.01, pcal_interval_mhz=1.0, lsb=(1, 1), sampling="complex"),
    "c_8bit": dict(base_freq_mhz=8401.0, pcal_interval_mhz=3.0, sampling="complex", nbits=8),
    "c_dsb": dict(base_freq_mhz=8400.01, pcal_interval_mhz=1.0, sampling="complex_dsb"),
}


@pytest.mark.parametrize("case", sorted(PCAL_CASES))
def test_phase_cal(gpu, oracle, case):
    """Phase-cal tones of every recorded band (Mode::process extraction,
    copyPCalTones) against the oracle's restated extractors, FP32, with a
    flagged block, a bad frame and a comb injected at known phases p_k
    (Schroeder's 180 k^2 / K degrees, whose low crest factor keeps the 2-bit
    quantiser from distorting the comb): tone k reads p_k - 90 degrees for
    sines (real samples), p_k for complex exponentials."""
    lib, ctx = gpu
    kw = dict(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=24)
    kw.update(PCAL_CASES[case])
    spec = JobSpec(**kw)
    job, keep = spec.build()
    ntones, offs = [], []
    for sb in range(spec.nsubbands):
        nt, off = C.c_int32(), C.c_int32()
        check(lib, lib.fxg_plan_pcal_tones(C.byref(job), 0, sb, C.byref(nt), C.byref(off)), "pcal tones")
        ntones.append(nt.value)
        offs.append(max(off.value, 0))
    K = max(ntones)
    phases = [180.0 * k * k / K for k in range(K)]
    comb = dict(amp=0.5, spacing_hz=1e6 * spec.pcal_interval_mhz, offsets_hz=offs, ntones=ntones, phases_deg=phases)
    sd = make_data(spec, invalid_frames={1: [2]}, invalid_blocks={2: [3, 4]}, pcal=comb)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd, fp64=False, nthreads=8)
    p0 = C.c_int64()
    check(lib, lib.fxg_plan_pcal_offset(C.byref(job), 0, C.byref(p0)), "pcal offset")
    p0 = p0.value
    assert g.info.result_length - p0 == spec.nstations * spec.npol * sum(ntones)
    assert rel_err(got[p0:], ref[p0:]) < FP32_TOL, rel_err(got[p0:], ref[p0:])
    assert np.max(np.abs(got[p0:] - ref[p0:])) < 1e-4 * np.max(np.abs(ref[p0:]))
    # the rest of the result is unchanged by the extraction
    assert rel_err(got[:p0], ref[:p0]) < FP32_TOL
    if case in ("implicit", "shifting", "lsb", "c_implicit", "c_shifting", "c_lsb", "c_8bit"):
        # station 0, band 0
        ph = np.degrees(np.angle(got[p0:p0 + ntones[0]]))
        exp = (0.0 if spec.is_complex else -90.0) + np.array(phases[:ntones[0]])
        fs_edge = offs[0] + np.arange(ntones[0]) * comb["spacing_hz"] < 1e6 * spec.bw_mhz
        d = (ph - exp + 180) % 360 - 180
        # 12288 real or 6144 complex samples of 2-bit data: a few degrees of
        # noise per tone; a sign, bin or sample-offset mistake would move the
        # phases by tens of degrees
        assert np.all(np.abs(d[fs_edge]) < 12.0), d
        assert np.mean(np.abs(d[fs_edge])) < 4.0, d
    g.close()


FP8_TOL = 0.08   # e4m3 keeps 3 mantissa bits: ~4 % per product, averaged over the blocks


@pytest.mark.parametrize("mk", ["c2", "wide", "zoom"])
def test_visibilities_fp8_vs_fp64(gpu, oracle, mk):
    """fp8 e4m3 spectra with per-(band, channel) power-of-two scales on the
    microscaled matrix cores, against the FP64 oracle: with the default
    scales and again after calibrating them on the subintegration's own
    power; a wrong scale would be off by powers of two."""
    lib, ctx = gpu
    if mk == "zoom":
        spec = c2_spec(xmac="fp8", zoom=((0, 128, 256), (1, 40, 64)))
    else:
        spec = {"c2": c2_spec, "wide": wide_spec}[mk](xmac="fp8")
    job, keep = spec.build()
    sd = make_data(spec, invalid_frames={0: [1]})
    g = GpuJob(lib, ctx, job)
    ref = oracle_results(oracle, job, sd, fp64=True, nthreads=8)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    first = g.process(sd)
    check(lib, lib.fxg_job_calibrate_scales(g.h), "calibrate")
    second = g.process(sd)
    for got in (first, second):
        assert rel_err(got[:xend], ref[:xend]) < FP8_TOL, rel_err(got[:xend], ref[:xend])
        assert rel_err(got[a0:a1], ref[a0:a1]) < FP8_TOL / 2, rel_err(got[a0:a1], ref[a0:a1])
    g.close()


@pytest.mark.parametrize("xmac,tol,fp64", [("fp32", FP32_TOL, False), ("fp16", 1e-3, True)])
def test_usb_against_lsb(gpu, oracle, xmac, tol, fp64):
    """USB x LSB baseline products (configuration.cpp:2926-2937): stations 1
    and 3 record the sky band as lower sideband, the others as upper; the
    cross-sideband products ride in the USB frequency's cross-multiply."""
    from fxtestlib import mixed_sideband_job, usb_to_lsb_raw
    lib, ctx = gpu
    spec = JobSpec(nstations=4, nsubbands=1, npol=1, bw_mhz=16.0, nchan=512, blockspersend=10, pols="R",
                   xmac=xmac)
    job, keep = mixed_sideband_job(spec, [1, 3])
    sd = make_data(spec, rho=0.1)
    for s in (1, 3):
        sd.stations[s].raw = usb_to_lsb_raw(spec, sd.stations[s].raw)
    g = GpuJob(lib, ctx, job)
    got = g.process(sd)
    ref = oracle_results(oracle, job, sd, fp64=fp64, nthreads=8)
    assert rel_err(got, ref) < tol, rel_err(got, ref)
    g.close()


POL_OFFSETS = [[30.0, -70.0], [110.0, 0.0], 45.0, [-15.0, 200.0]]


@pytest.mark.parametrize("case", ["real-fp32", "real-big-fp32", "lsb-order0-fp32", "complex-fp32", "zoom-fp16",
                                  "real-bf16"])
def test_pol_phase_offsets(gpu, oracle, case):
    """Recorded-frequency polarisation phase offsets (mode.cpp:435-460,
    1341-1353): each frequency's LCP band turned by e^{i theta/2}, the turn
    riding on the fractional-sample rotators; FP32 against the FP32 oracle
    (visibilities, autocorrelations and one block's spectra), reduced
    precision against the FP64 oracle; a zoom band inherits its parent's
    turn."""
    lib, ctx = gpu
    kw = dict(nstations=4, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=10,
              phaseoffset_deg=POL_OFFSETS)
    if case == "real-big-fp32":
        kw.update(nchan=32768, blockspersend=3, nstations=3, phaseoffset_deg=POL_OFFSETS[:3])
    if case == "lsb-order0-fp32":
        kw.update(lsb=[1, 0], fringe_order=0, pols="LR")
    if case == "complex-fp32":
        kw.update(sampling="complex_dsb", lsb=[0, 1], pols="XY")
    if case == "zoom-fp16":
        kw.update(xmac="fp16", zoom=[(1, 32, 64)])
    if case == "real-bf16":
        kw.update(xmac="bf16")
    spec = JobSpec(**kw)
    job, keep = spec.build()
    sd = make_data(spec, rho=0.1, invalid_blocks={1: [2]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    fp64 = not case.endswith("fp32")
    ref = oracle_results(oracle, job, sd, fp64=fp64, nthreads=8)
    tol = FP32_TOL if not fp64 else MIXED_TOL[spec.xmac]
    assert rel_err(got, ref) < tol, rel_err(got, ref)
    if not fp64:
        blk = spec.blockspersend - 1
        for ds in range(spec.nstations):
            refs = oracle_spectra(oracle, job, sd, ds, blk, spec.nrecbands, spec.nchan)
            for band in range(spec.nrecbands):
                assert rel_err(g.spectrum(ds, band, blk, spec.nchan), refs[band]) < FP32_TOL, (ds, band)
    g.close()


@pytest.mark.parametrize("order", [0, 2])
@pytest.mark.parametrize("nchan,nbits", [(32768, 2), (16384, 8)])
def test_largest_fft_fringe_orders(gpu, oracle, order, nchan, nbits):
    """The two-kernel 65536- and 32768-point transforms with post-F (order 0)
    and quadratic (order 2) fringe rotation, FP32 against the FP32 oracle,
    with a curved delay model.  Blocks wholly inside valid frames take the
    first kernel's word-aligned path (station 1, block 3: spectra compared
    on their own); the blocks of station 2 that overlap its invalid frame
    take the general per-sample decoder, and their spectra are compared on
    their own too."""
    lib, ctx = gpu
    spec = JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=64.0, nchan=nchan, blockspersend=4, nbits=nbits,
                   fringe_order=order, looffset_hz=[7.5])
    job, keep = spec.build()
    fs, L = spec.framesamples, spec.fftchannels

    def data(invalid):
        sd = make_data(spec, invalid_frames=invalid)
        for i, st in enumerate(sd.stations):
            st.interpolator = st.interpolator + np.array([2e-5 * (i - 1), 0.0, 0.0])
        return sd

    # station 2's invalid frame: the one under the middle of its block 1
    # (the block geometry does not depend on frame validity)
    mid = oracle_block_info(oracle, job, data({}), 2, 1).nearestsample + L // 2
    sd = data({2: [mid // fs]})
    g = GpuJob(lib, ctx, job)
    got = g.process(sd, keep_taps=True)
    ref = oracle_results(oracle, job, sd)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    assert rel_err(got[:xend], ref[:xend]) < FP32_TOL
    assert rel_err(got[a0:a1], ref[a0:a1]) < FP32_TOL
    refs = oracle_spectra(oracle, job, sd, 1, 3, spec.nrecbands, spec.nchan)
    for band in range(spec.nrecbands):
        assert rel_err(g.spectrum(1, band, 3, spec.nchan), refs[band]) < FP32_TOL
    general = []
    for blk in range(spec.blockspersend):
        bi = oracle_block_info(oracle, job, sd, 2, blk)
        if bi.status == 0 and bi.nearestsample < (mid // fs + 1) * fs and bi.nearestsample + L > (mid // fs) * fs:
            assert bi.weight < 1.0
            general.append(blk)
    assert 1 in general
    for blk in general:
        refs = oracle_spectra(oracle, job, sd, 2, blk, spec.nrecbands, spec.nchan)
        for band in range(spec.nrecbands):
            assert rel_err(g.spectrum(2, band, blk, spec.nchan), refs[band]) < FP32_TOL, (blk, band)
    g.close()


@pytest.mark.parametrize("nchan,nbits", [(32768, 2), (16384, 8)])
def test_largest_fft_fp8(gpu, oracle, nchan, nbits):
    """fp8 e4m3 spectra written by the two-kernel transforms (per-(band,
    channel) exponents applied in the second kernel), against the FP64
    oracle, before and after calibrating the scales."""
    lib, ctx = gpu
    spec = JobSpec(nstations=3, nsubbands=1, npol=2, bw_mhz=64.0, nchan=nchan, blockspersend=6, nbits=nbits,
                   xmac="fp8")
    job, keep = spec.build()
    sd = make_data(spec, rho=0.1)
    g = GpuJob(lib, ctx, job)
    ref = oracle_results(oracle, job, sd, fp64=True, nthreads=8)
    info = plan_info(oracle, "orc_", job)
    xend, a0, a1 = region_slices(oracle, job, info)
    first = g.process(sd)
    check(lib, lib.fxg_job_calibrate_scales(g.h), "calibrate")
    second = g.process(sd)
    for got in (first, second):
        assert rel_err(got[:xend], ref[:xend]) < FP8_TOL, rel_err(got[:xend], ref[:xend])
        assert rel_err(got[a0:a1], ref[a0:a1]) < FP8_TOL / 2, rel_err(got[a0:a1], ref[a0:a1])
    g.close()
