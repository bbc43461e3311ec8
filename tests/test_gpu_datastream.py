"""A station stream resident on the device, cut into subintegration windows
by the datastream boundary (fxds_control) and correlated straight from
device memory, gives the visibilities of the same subintegrations read from
the whole stream."""
import ctypes as C

import numpy as np
import pytest

import fxcorr
from fxcorr import JobSpec, check
from fxsynth import make_subint
from fxtestlib import GpuJob, rel_err

pytestmark = pytest.mark.gpu


def test_windows_from_a_resident_stream(gpu):
    lib, ctx = gpu
    host = fxcorr.load_host_lib()
    nsub = 4
    spec = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16)
    longspec = JobSpec(nstations=3, nsubbands=2, npol=2, bw_mhz=16.0, nchan=256, blockspersend=16 * (nsub + 1))
    delays = [3.7, -12.25, 40.0]
    sd = make_subint(longspec, seed=12, delays_us=delays, rho=0.1)
    job, keep = spec.build()
    g = GpuJob(lib, ctx, job)
    sub_us = spec.blockspersend * spec.fftchannels * spec.sampletime_us
    dev, streams = [], []
    for s, st in enumerate(sd.stations):
        p = C.c_void_p()
        check(lib, lib.fxg_device_alloc(ctx, st.raw.size, C.byref(p)), "alloc")
        check(lib, lib.fxg_memcpy_h2d(ctx, p, st.raw.ctypes.data, st.raw.size), "h2d")
        dev.append(p)
        f = fxcorr.DsFormat()
        check(host, host.fxds_format_of(C.byref(job), s, 20000.0, C.byref(f)), "format")
        h = C.c_void_p()
        check(host, host.fxds_create(C.byref(f), C.byref(h)), "create")
        check(host, host.fxds_set_buffer(h, p, st.raw.size, 0, st.datasec, st.datans), "set_buffer")
        streams.append(h)
    try:
        for k in range(nsub):
            t_ns = int(round(k * sub_us * 1000))
            sub = fxcorr.SubintDesc()
            sub.scan, sub.sec, sub.ns, sub.scan_start_sec, sub.model_ok = 0, sd.sub_sec + t_ns // 10**9, t_ns % 10**9, 0, 1
            # the whole stream, host data
            whole = (fxcorr.StationInput * 3)()
            keepf = []
            for s, st in enumerate(sd.stations):
                a = whole[s]
                a.data, a.nbytes, a.scan, a.sec, a.ns = st.raw.ctypes.data, st.raw.size, 0, st.datasec, st.datans
                vf = np.full(1, (1 << 16) - 1, np.int32)
                keepf.append(vf)
                a.nvalidflags, a.validflags = 1, vf.ctypes.data_as(C.POINTER(C.c_int32))
                a.interpolator[:] = [0.0, 0.0, delays[s]]
            ref = np.zeros(2 * g.info.result_length, np.float32)
            check(lib, lib.fxg_subint_process(g.h, C.byref(sub), whole, 0, ref.ctypes.data, 0), "whole")
            # the windows of the resident stream
            win = (fxcorr.StationInput * 3)()
            flags = [(C.c_int32 * 4)() for _ in range(3)]
            for s in range(3):
                bi = C.c_int64()
                check(host, host.fxds_control(streams[s], 0, sub.sec, sub.ns, delays[s], delays[s], 1, C.byref(win[s]),
                                              flags[s], 4, C.byref(bi)), "control")
                assert win[s].sec != fxcorr.FXG_INVALID_SUBINT
                assert flags[s][0] == (1 << 16) - 1
                assert bi.value > 0 or k == 0
                win[s].interpolator[:] = [0.0, 0.0, delays[s]]
            got = np.zeros_like(ref)
            check(lib, lib.fxg_subint_process(g.h, C.byref(sub), win, 1, got.ctypes.data, 0), "windows")
            assert rel_err(got.view(np.complex64), ref.view(np.complex64)) < 1e-6, k
    finally:
        for h in streams:
            host.fxds_destroy(h)
        for p in dev:
            lib.fxg_device_free(ctx, p)
        g.close()
