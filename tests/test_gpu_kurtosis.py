"""Spectral kurtosis records (Mode's dumpkurtosis path, mode.cpp:1260-1277,
1426-1478; Core::averageAndSendKurtosis core.cpp:1371-1420) on the GPU
against the oracle: FP32 spectra to FP32 rounding, reduced-precision spectra
to their rounding, with thread block ranges that cut across the F/X passes,
graph replays, and the validity of a station without data."""
import ctypes as C

import numpy as np
import pytest

from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import GpuJob, oracle_kurtosis

pytestmark = pytest.mark.gpu

# FP32 spectra: the GPU reads |X| after the unit-modulus fractional rotation
# and its own FFT, the oracle before the rotation; both FP32
TOL = {"fp32": 5e-4, "fp16": 5e-3, "bf16": 5e-2}


def spec_of(xmac, **kw):
    base = dict(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=10, xmac=xmac,
                max_chunk_blocks=4)
    base.update(kw)
    return JobSpec(**base)


@pytest.mark.parametrize("xmac", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("nthreads,maxch", [(1, 32), (3, 256), (3, 16)])
def test_kurtosis_vs_oracle(gpu, oracle, xmac, nthreads, maxch):
    lib, ctx = gpu
    spec = spec_of(xmac)
    job, keep = spec.build()
    sd = make_subint(spec, seed=9, rho=0.05, invalid_blocks={2: list(range(spec.blockspersend))},
                     invalid_frames={0: [1]})
    g = GpuJob(lib, ctx, job)
    check(lib, lib.fxg_job_set_kurtosis(g.h, 1, maxch, nthreads), "set_kurtosis")
    g.process(sd)
    nout = min(maxch, spec.nchan)
    for ds in range(spec.nstations):
        ref, rvalid = oracle_kurtosis(oracle, job, sd, ds, nthreads, maxch, spec.nrecbands, spec.nchan)
        for t in range(nthreads):
            for b in range(spec.nrecbands):
                got, valid = g.kurtosis(t, ds, b, nout)
                assert valid == rvalid[t], (ds, t)
                np.testing.assert_allclose(got, ref[t, b], rtol=0, atol=TOL[xmac], err_msg=f"ds {ds} t {t} b {b}")
    # station 2 has no usable block: no valid record
    assert all(g.kurtosis(t, 2, 0, nout)[1] == 0 for t in range(nthreads))
    g.close()


def test_kurtosis_graph_replays_and_off(gpu, oracle):
    """Graph replays give the same records as launches one by one; turning
    kurtosis off leaves the visibilities unchanged and the getter refuses."""
    lib, ctx = gpu
    spec = spec_of("fp32", nstations=2)
    job, keep = spec.build()
    sd = make_subint(spec, seed=4)
    g = GpuJob(lib, ctx, job)
    check(lib, lib.fxg_job_set_kurtosis(g.h, 1, 32, 2), "set_kurtosis")
    vis0 = g.process(sd)
    rec0 = [g.kurtosis(t, d, b, 32)[0] for t in range(2) for d in range(2) for b in range(spec.nrecbands)]
    check(lib, lib.fxg_job_set_graph(g.h, 1), "set_graph")
    for _ in range(3):
        vis = g.process(sd)
        rec = [g.kurtosis(t, d, b, 32)[0] for t in range(2) for d in range(2) for b in range(spec.nrecbands)]
        np.testing.assert_array_equal(vis, vis0)
        for a, b in zip(rec, rec0):
            np.testing.assert_array_equal(a, b)
    check(lib, lib.fxg_job_set_kurtosis(g.h, 0, 32, 1), "kurtosis off")
    np.testing.assert_array_equal(g.process(sd), vis0)
    out = np.zeros(32, dtype=np.float32)
    n, v = C.c_int32(), C.c_int32()
    assert lib.fxg_job_get_kurtosis(g.h, 0, 0, 0, out.ctypes.data_as(C.POINTER(C.c_float)), 32,
                                    C.byref(n), C.byref(v)) != 0
    g.close()


@pytest.mark.parametrize("xmac", ["fp32", "fp16"])
def test_kurtosis_largest_fft(gpu, oracle, xmac):
    """Kurtosis of the spectra the two-kernel 65536-point transform writes
    (row-major layout, scratch sub-chunks), two records over six blocks."""
    lib, ctx = gpu
    spec = spec_of(xmac, nstations=2, nsubbands=1, bw_mhz=64.0, nchan=32768, blockspersend=6, max_chunk_blocks=0)
    job, keep = spec.build()
    sd = make_subint(spec, seed=6, rho=0.05)
    g = GpuJob(lib, ctx, job)
    check(lib, lib.fxg_job_set_kurtosis(g.h, 1, 64, 2), "set_kurtosis")
    g.process(sd)
    for ds in range(spec.nstations):
        ref, rvalid = oracle_kurtosis(oracle, job, sd, ds, 2, 64, spec.nrecbands, spec.nchan)
        for t in range(2):
            for b in range(spec.nrecbands):
                got, valid = g.kurtosis(t, ds, b, 64)
                assert valid == rvalid[t], (ds, t)
                np.testing.assert_allclose(got, ref[t, b], rtol=0, atol=TOL[xmac], err_msg=f"ds {ds} t {t} b {b}")
    g.close()
