"""Shared helpers for the parity tests: run one subintegration through the
CPU oracle and through the HIP library, and compare result buffers."""
from __future__ import annotations

import ctypes as C

import numpy as np

import fxcorr
from fxcorr import BlockInfo, PlanInfo, check


def plan_info(lib, prefix, job) -> PlanInfo:
    info = PlanInfo()
    check(lib, getattr(lib, prefix + "plan")(C.byref(job), C.byref(info)), "plan")
    return info


def oracle_results(orc, job, sd, fp64=False, nthreads=4, block0=0, nblocks=None):
    info = plan_info(orc, "orc_", job)
    nb = job.blockspersend if nblocks is None else nblocks
    dt = np.float64 if fp64 else np.float32
    out = np.zeros(2 * info.result_length, dtype=dt)
    sub = sd.subint_desc()
    st, keep = sd.station_inputs()
    check(orc, orc.orc_process(C.byref(job), C.byref(sub), st, 1 if fp64 else 0, nthreads, block0, nb,
                               out.ctypes.data), "orc_process")
    return out.view(np.complex128 if fp64 else np.complex64)


def oracle_block_info(orc, job, sd, ds, block) -> BlockInfo:
    sub = sd.subint_desc()
    st, keep = sd.station_inputs()
    bi = BlockInfo()
    check(orc, orc.orc_block_info(C.byref(job), C.byref(sub), st, ds, block, C.byref(bi)), "orc_block_info")
    return bi


def oracle_spectra(orc, job, sd, ds, block, nbands, nchan, fp64=False):
    sub = sd.subint_desc()
    st, keep = sd.station_inputs()
    dt = np.complex128 if fp64 else np.complex64
    out = np.zeros(nbands * nchan, dtype=dt)
    check(orc, orc.orc_spectra(C.byref(job), C.byref(sub), st, ds, block, 1 if fp64 else 0,
                               out.ctypes.data), "orc_spectra")
    return out.reshape(nbands, nchan)


def oracle_kurtosis(orc, job, sd, ds, nthreads, maxchannels, nbands, nchan):
    """(sk[nthreads][nbands][nout], valid[nthreads]) of datastream ds."""
    sub = sd.subint_desc()
    st, keep = sd.station_inputs()
    nout = min(maxchannels, nchan)
    out = np.zeros(nthreads * nbands * nout, dtype=np.float32)
    valid = np.zeros(nthreads, dtype=np.int32)
    check(orc, orc.orc_kurtosis(C.byref(job), C.byref(sub), st, ds, nthreads, maxchannels,
                                out.ctypes.data_as(C.POINTER(C.c_float)),
                                valid.ctypes.data_as(C.POINTER(C.c_int))), "orc_kurtosis")
    return out.reshape(nthreads, nbands, nout), valid


def mixed_sideband_job(spec, lsb_stations):
    """spec's job (one subband, real sampling) with the stations in
    lsb_stations recording the same sky band as lower sideband: a second
    freq table entry, band edge + bandwidth, LSB, that their recorded band
    carries; the baselines keep the USB entry (Configuration's rewrite of
    LSB freqs correlated against USB, configuration.cpp:1143-1186).
    Returns (job, keepalive)."""
    job, keep = spec.build()
    f0 = job.freqs[0]
    freqs = (fxcorr.Freq * 2)()
    C.memmove(C.addressof(freqs[0]), C.addressof(f0), C.sizeof(fxcorr.Freq))
    freqs[1].freq_mhz = f0.freq_mhz + f0.bw_mhz
    freqs[1].bw_mhz = f0.bw_mhz
    freqs[1].lsb = 1
    freqs[1].nchan = f0.nchan
    freqs[1].chans_to_avg = f0.chans_to_avg
    job.nfreqs = 2
    job.freqs = C.cast(freqs, C.POINTER(fxcorr.Freq))
    for s in lsb_stations:
        job.datastreams[s].recfreq_index[0] = 1
    return job, (keep, freqs)


def usb_to_lsb_raw(spec, raw):
    """The frames of a 1-band 2-bit USB stream as the same sky band seen in
    LSB from an LO at the upper band edge: x_lsb[n] = (-1)^n x_usb[n], i.e.
    the odd samples' codes k -> 3 - k (bits 2-3 and 6-7 of every byte)."""
    assert spec.nrecbands == 1 and spec.nbits == 2 and spec.sampling == "real"
    fr = raw.reshape(-1, spec.framebytes).copy()
    fr[:, spec.headerbytes:] ^= 0xCC
    return fr.reshape(-1)


class GpuJob:
    """A job on the HIP library; results come back to the host."""

    def __init__(self, lib, ctx, job):
        self.lib, self.ctx, self.job = lib, ctx, job
        self.h = C.c_void_p()
        check(lib, lib.fxg_job_create(ctx, C.byref(job), C.byref(self.h)), "fxg_job_create")
        self.info = PlanInfo()
        check(lib, lib.fxg_job_plan_info(self.h, C.byref(self.info)), "plan_info")

    def process(self, sd, keep_taps=False):
        lib = self.lib
        check(lib, lib.fxg_job_set_keep_taps(self.h, 1 if keep_taps else 0), "keep_taps")
        sub = sd.subint_desc()
        st, keep = sd.station_inputs()
        out = np.zeros(2 * self.info.result_length, dtype=np.float32)
        check(lib, lib.fxg_subint_process(self.h, C.byref(sub), st, 0, out.ctypes.data, 0), "fxg_subint_process")
        return out.view(np.complex64)

    def block_info(self, ds, block) -> BlockInfo:
        bi = BlockInfo()
        check(self.lib, self.lib.fxg_tap_block_info(self.h, ds, block, C.byref(bi)), "tap_block_info")
        return bi

    def spectrum(self, ds, band, block, nchan):
        out = np.zeros(nchan, dtype=np.complex64)
        check(self.lib, self.lib.fxg_tap_spectrum(self.h, ds, band, block, out.ctypes.data), "tap_spectrum")
        return out

    def kurtosis(self, thread, ds, band, nout):
        """(sk[nout], valid) of one record of the last subintegration."""
        out = np.zeros(nout, dtype=np.float32)
        n, v = C.c_int32(), C.c_int32()
        check(self.lib, self.lib.fxg_job_get_kurtosis(self.h, thread, ds, band,
                                                      out.ctypes.data_as(C.POINTER(C.c_float)), nout,
                                                      C.byref(n), C.byref(v)), "get_kurtosis")
        assert n.value == nout
        return out, v.value

    def close(self):
        if self.h:
            self.lib.fxg_job_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_err(a, b):
    """||a - b|| / ||b||."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def region_slices(orc, job, info):
    """(xcorr_end, autocorr_start, autocorr_end) cf32 indices of the layout."""
    off = C.c_int64()
    check(orc, orc.orc_plan_bweight_offset(C.byref(job), job.baselines[0].freq_index[0], 0, C.byref(off)), "bw")
    xend = off.value
    check(orc, orc.orc_plan_autocorr_offset(C.byref(job), 0, C.byref(off)), "ac")
    a0 = off.value
    check(orc, orc.orc_plan_acweight_offset(C.byref(job), 0, C.byref(off)), "acw")
    a1 = off.value
    return xend, a0, a1
