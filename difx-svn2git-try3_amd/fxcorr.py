"""ctypes binding of the libfxcorr_gpu C ABI (include/fxcorr_gpu.h) and a
builder for correlator job descriptors.

The descriptor structs mirror the reference's configuration tables:
frequency table (mpifxcorr/src/configuration.h freqdata), datastream table
(datastreamdata), baseline table (baselinedata), as read by
Configuration::processDatastreamTable / processBaselineTable
(mpifxcorr/src/configuration.cpp:1051-1135).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

FXG_MAX_FREQS = 64
FXG_MAX_RECFREQS = 32
FXG_MAX_BANDS = 64
FXG_MAX_ZOOMBANDS = 128
FXG_MAX_BLFREQS = 48
FXG_MAX_POLPRODUCTS = 4

FXG_FORMAT_VDIF = 0
FXG_FORMAT_MARK5B = 1
FXG_FORMAT_INTERLACEDVDIF = 2
FXG_BLANKER_FORMAT = 0
FXG_BLANKER_NONE = 1
FXG_SAMPLING_REAL = 0
FXG_SAMPLING_COMPLEX = 1
FXG_XMAC_FP32 = 0
FXG_XMAC_FP16 = 1
FXG_XMAC_BF16 = 2
FXG_XMAC_FP8 = 3
FXG_OK, FXG_EINVAL, FXG_ENOMEM, FXG_EHIP, FXG_EUNSUPPORTED, FXG_ESTATE, FXG_ENODEV = 0, -1, -2, -3, -4, -5, -6
FXG_INVALID_SUBINT = -99999999
FXG_FLAGS_PER_INT = 30
STAGE_NAMES = ["frames", "prologue", "fstage", "xmac", "weights", "scatter", "h2d", "d2h", "pcal", "kurtosis"]

XMAC_NAMES = {"fp32": FXG_XMAC_FP32, "fp16": FXG_XMAC_FP16, "bf16": FXG_XMAC_BF16, "fp8": FXG_XMAC_FP8}


class Freq(C.Structure):
    _fields_ = [("freq_mhz", C.c_double), ("bw_mhz", C.c_double), ("lsb", C.c_int32),
                ("nchan", C.c_int32), ("chans_to_avg", C.c_int32), ("pad", C.c_int32)]


class Datastream(C.Structure):
    _fields_ = [
        ("format", C.c_int32), ("nbits", C.c_int32), ("framebytes", C.c_int32),
        ("headerbytes", C.c_int32), ("blanker", C.c_int32), ("nrecfreqs", C.c_int32),
        ("nrecbands", C.c_int32), ("nzoombands", C.c_int32),
        ("recfreq_index", C.c_int32 * FXG_MAX_RECFREQS),
        ("recfreq_clockoffset_us", C.c_double * FXG_MAX_RECFREQS),
        ("recfreq_clockoffsetdelta_us", C.c_double * FXG_MAX_RECFREQS),
        ("recfreq_phaseoffset_deg", C.c_double * FXG_MAX_RECFREQS),
        ("recfreq_looffset_hz", C.c_double * FXG_MAX_RECFREQS),
        ("recband_freq", C.c_int32 * FXG_MAX_BANDS),
        ("recband_pol", C.c_char * FXG_MAX_BANDS),
        ("zoomband_freqindex", C.c_int32 * FXG_MAX_ZOOMBANDS),
        ("zoomband_pol", C.c_char * FXG_MAX_ZOOMBANDS),
        ("zoomband_parent", C.c_int32 * FXG_MAX_ZOOMBANDS),
        ("zoomband_choffset", C.c_int32 * FXG_MAX_ZOOMBANDS),
        ("pcal_interval_mhz", C.c_double),
        ("pcal_base_mhz", C.c_double),
        ("sampling", C.c_int32), ("complex_dsb", C.c_int32),
    ]


class Baseline(C.Structure):
    _fields_ = [
        ("ds1", C.c_int32), ("ds2", C.c_int32), ("nfreqs", C.c_int32),
        ("freq_index", C.c_int32 * FXG_MAX_BLFREQS),
        ("npolproducts", C.c_int32 * FXG_MAX_BLFREQS),
        ("band1", (C.c_int32 * FXG_MAX_POLPRODUCTS) * FXG_MAX_BLFREQS),
        ("band2", (C.c_int32 * FXG_MAX_POLPRODUCTS) * FXG_MAX_BLFREQS),
    ]


class JobDesc(C.Structure):
    _fields_ = [
        ("nfreqs", C.c_int32), ("ndatastreams", C.c_int32), ("nbaselines", C.c_int32),
        ("blockspersend", C.c_int32), ("fringe_order", C.c_int32),
        ("arraystridelen", C.c_int32), ("xmacstridelen", C.c_int32),
        ("crosspol_autocorrs", C.c_int32), ("xmac_precision", C.c_int32),
        ("max_chunk_blocks", C.c_int32),
        ("freqs", C.POINTER(Freq)), ("datastreams", C.POINTER(Datastream)),
        ("baselines", C.POINTER(Baseline)),
    ]


class StationInput(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("nbytes", C.c_int64),
        ("scan", C.c_int32), ("sec", C.c_int32), ("ns", C.c_int32),
        ("nvalidflags", C.c_int32), ("validflags", C.POINTER(C.c_int32)),
        ("interpolator", C.c_double * 3),
    ]


class SubintDesc(C.Structure):
    _fields_ = [("scan", C.c_int32), ("sec", C.c_int32), ("ns", C.c_int32),
                ("scan_start_sec", C.c_int32), ("model_ok", C.c_int32), ("pad", C.c_int32)]


FXG_MUX_MAX_THREADS = 31
FXG_MUX_GOTOEND, FXG_MUX_RESPECTGRANULARITY, FXG_MUX_ENABLEVALIDITY = 0x01, 0x02, 0x04
FXG_MUX_PROPAGATEVALIDITY = 0x40


class VdifMuxDesc(C.Structure):
    _fields_ = [("input_frame_bytes", C.c_int32), ("frames_per_second", C.c_int32),
                ("bits_per_sample", C.c_int32), ("nthreads", C.c_int32),
                ("thread_ids", C.c_int32 * FXG_MUX_MAX_THREADS), ("channels_per_thread", C.c_int32),
                ("nsort", C.c_int32), ("ngap", C.c_int32), ("flags", C.c_int32)]


class VdifMuxStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "valid_frames", "invalid_frames", "discarded_frames", "wrong_thread_frames", "skipped_bytes",
        "fill_bytes", "duplicate_frames", "bytes_processed", "good_frames", "partial_frames", "filler_frames",
        "src_used", "dest_used", "output_frame_bytes", "output_frames", "start_frame_number",
        "output_frame_granularity", "epoch")]

    def as_list(self):
        return [getattr(self, n) for n, _ in self._fields_]


class DsFormat(C.Structure):
    _fields_ = [("framebytes", C.c_int32), ("payloadbytes", C.c_int32), ("framespersecond", C.c_double),
                ("framegranularity", C.c_int32), ("bytespersamplenum", C.c_int32),
                ("bytespersampledenom", C.c_int32), ("numchannels", C.c_int32), ("blockspersend", C.c_int32),
                ("pad", C.c_int32), ("bw_mhz", C.c_double), ("databytes", C.c_int64)]


class Schedule(C.Structure):
    _fields_ = [("xmac_streaming", C.c_int32), ("xmac_ct", C.c_int32), ("xmac_tb", C.c_int32),
                ("chunk_blocks", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [("schedule", Schedule), ("ms", C.c_double), ("rel_err", C.c_double),
                ("accepted", C.c_int32), ("pad", C.c_int32)]


class BlockInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("nearestsample", C.c_int32), ("integerdelay", C.c_int32),
                ("weight", C.c_float), ("fracsampleerror", C.c_float), ("intwalltime", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("fracwalltime", C.c_double), ("c", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [("result_length", C.c_int64), ("device_bytes", C.c_int64),
                ("arraystridelen", C.c_int32), ("xmacstridelen", C.c_int32),
                ("rotatestridelen", C.c_int32), ("bandsperautocorr", C.c_int32),
                ("fftchannels", C.c_int32), ("chunk_blocks", C.c_int32)]


# --------------------------------------------------------------------------
# Job specification
# --------------------------------------------------------------------------
@dataclass
class JobSpec:
    """A correlator configuration in the reference's table terms.

    stations x subbands x pols recorded bands per datastream, band index
    = subband*npol + pol (as in the reference's .input REC BAND tables),
    baselines for every station pair i<j with pol products RR, RL, LR, LL
    (cross-pol) or RR, LL (parallel hands only).
    """
    nstations: int
    nsubbands: int
    npol: int
    bw_mhz: float
    nchan: int
    nbits: int = 2
    blockspersend: int = 16
    crosspol: bool = True
    xmac: str = "fp32"
    framebytes: int = 8032
    headerbytes: int = 32
    fmt: int = FXG_FORMAT_VDIF
    blanker: int = FXG_BLANKER_FORMAT
    base_freq_mhz: float = 8400.0
    lsb: Sequence[int] = ()
    clockoffset_us: Sequence[float] = ()
    looffset_hz: Sequence[float] = ()
    phaseoffset_deg: Sequence = ()      # per station: one value, or one per subband
    zoom: Sequence[tuple] = ()          # (subband, channel offset, nchan) per zoom freq
    arraystridelen: int = 0
    xmacstridelen: int = 0
    max_chunk_blocks: int = 0
    pols: str = "RL"
    fringe_order: int = 1
    chans_to_avg: int = 1               # channel averaging of every frequency
    pcal_interval_mhz: float = 0.0      # phase-cal tone interval (0 = no extraction)
    pcal_base_mhz: float = 0.0
    sampling: str = "real"              # "real", "complex" (single sideband) or "complex_dsb"

    @property
    def is_complex(self) -> bool:
        return self.sampling != "real"

    @property
    def nrecbands(self) -> int:
        return self.nsubbands * self.npol

    @property
    def fftchannels(self) -> int:
        return self.nchan if self.is_complex else 2 * self.nchan

    @property
    def sampletime_us(self) -> float:
        return (1.0 if self.is_complex else 0.5) / self.bw_mhz

    @property
    def unitbits(self) -> int:
        """Bits of one sample of one band (both components if complex)."""
        return 2 * self.nbits if self.is_complex else self.nbits

    @property
    def framesamples(self) -> int:
        return (self.framebytes - self.headerbytes) * 8 // (self.nrecbands * self.unitbits)

    def build(self):
        """Return (JobDesc, keepalive) with ctypes arrays owned by keepalive."""
        nsb, npol = self.nsubbands, self.npol
        nz = len(self.zoom)
        nfreqs = nsb + nz
        freqs = (Freq * nfreqs)()
        for sb in range(nsb):
            f = freqs[sb]
            f.freq_mhz = self.base_freq_mhz + sb * self.bw_mhz
            f.bw_mhz = self.bw_mhz
            f.lsb = int(self.lsb[sb]) if sb < len(self.lsb) else 0
            f.nchan = self.nchan
            f.chans_to_avg = self.chans_to_avg
        for z, (sb, choff, znchan) in enumerate(self.zoom):
            f = freqs[nsb + z]
            parent = freqs[sb]
            zbw = self.bw_mhz * znchan / self.nchan
            f.freq_mhz = parent.freq_mhz + choff * self.bw_mhz / self.nchan
            f.bw_mhz = zbw
            f.lsb = parent.lsb
            f.nchan = znchan
            f.chans_to_avg = self.chans_to_avg
        dss = (Datastream * self.nstations)()
        for s in range(self.nstations):
            d = dss[s]
            d.format = self.fmt
            d.nbits = self.nbits
            d.framebytes = self.framebytes
            d.headerbytes = self.headerbytes
            d.blanker = self.blanker
            d.nrecfreqs = nsb
            d.nrecbands = nsb * npol
            d.nzoombands = nz * npol
            d.pcal_interval_mhz = self.pcal_interval_mhz
            d.pcal_base_mhz = self.pcal_base_mhz
            d.sampling = FXG_SAMPLING_COMPLEX if self.is_complex else FXG_SAMPLING_REAL
            d.complex_dsb = 1 if self.sampling == "complex_dsb" else 0
            recpol = bytearray(FXG_MAX_BANDS)
            zpol = bytearray(FXG_MAX_ZOOMBANDS)
            for sb in range(nsb):
                d.recfreq_index[sb] = sb
                d.recfreq_clockoffset_us[sb] = self.clockoffset_us[s] if s < len(self.clockoffset_us) else 0.0
                d.recfreq_looffset_hz[sb] = self.looffset_hz[s] if s < len(self.looffset_hz) else 0.0
                if s < len(self.phaseoffset_deg):
                    po = self.phaseoffset_deg[s]
                    d.recfreq_phaseoffset_deg[sb] = po[sb] if isinstance(po, (list, tuple)) else po
                for p in range(npol):
                    d.recband_freq[sb * npol + p] = sb
                    recpol[sb * npol + p] = ord(self.pols[p])
            for z, (sb, choff, znchan) in enumerate(self.zoom):
                for p in range(npol):
                    k = z * npol + p
                    d.zoomband_freqindex[k] = nsb + z
                    zpol[k] = ord(self.pols[p])
                    d.zoomband_parent[k] = sb * npol + p
                    d.zoomband_choffset[k] = choff
            C.memmove(C.addressof(d) + Datastream.recband_pol.offset, bytes(recpol), FXG_MAX_BANDS)
            C.memmove(C.addressof(d) + Datastream.zoomband_pol.offset, bytes(zpol), FXG_MAX_ZOOMBANDS)
        pairs = [(i, j) for i in range(self.nstations) for j in range(i + 1, self.nstations)]
        bls = (Baseline * max(1, len(pairs)))()
        for b, (i, j) in enumerate(pairs):
            bl = bls[b]
            bl.ds1, bl.ds2 = i, j
            bl.nfreqs = nfreqs
            for f in range(nfreqs):
                bl.freq_index[f] = f
                if f < nsb:
                    base = f * npol
                else:
                    base = nsb * npol + (f - nsb) * npol
                if npol == 2 and self.crosspol:
                    prods = [(0, 0), (0, 1), (1, 0), (1, 1)]
                elif npol == 2:
                    prods = [(0, 0), (1, 1)]
                else:
                    prods = [(0, 0)]
                bl.npolproducts[f] = len(prods)
                for p, (a, c) in enumerate(prods):
                    bl.band1[f][p] = base + a
                    bl.band2[f][p] = base + c
        job = JobDesc()
        job.nfreqs = nfreqs
        job.ndatastreams = self.nstations
        job.nbaselines = len(pairs)
        job.blockspersend = self.blockspersend
        job.fringe_order = self.fringe_order
        job.arraystridelen = self.arraystridelen
        job.xmacstridelen = self.xmacstridelen
        job.crosspol_autocorrs = 1 if (self.crosspol and npol == 2) else 0
        job.xmac_precision = XMAC_NAMES[self.xmac]
        job.max_chunk_blocks = self.max_chunk_blocks
        job.freqs = C.cast(freqs, C.POINTER(Freq))
        job.datastreams = C.cast(dss, C.POINTER(Datastream))
        job.baselines = C.cast(bls, C.POINTER(Baseline))
        return job, (freqs, dss, bls)


# --------------------------------------------------------------------------
# Library loading
# --------------------------------------------------------------------------
_GPU_LIB_PATH = os.path.join(HERE, "_build", "libfxcorr_gpu.so")
_ORACLE_PATH = os.path.join(REPO, "oracle", "_build", "libfxoracle.so")

_gpu_lib = None
_oracle = None


def _declare_common(lib, prefix):
    P = C.POINTER
    lib.__getattr__(prefix + "plan").argtypes = [P(JobDesc), P(PlanInfo)]
    for name in ("plan_baseline_offset", "plan_bweight_offset"):
        lib.__getattr__(prefix + name).argtypes = [P(JobDesc), C.c_int, C.c_int, P(C.c_int64)]
    for name in ("plan_autocorr_offset", "plan_acweight_offset", "plan_pcal_offset"):
        lib.__getattr__(prefix + name).argtypes = [P(JobDesc), C.c_int, P(C.c_int64)]
    lib.__getattr__(prefix + "plan_pcal_tones").argtypes = [P(JobDesc), C.c_int, C.c_int, P(C.c_int32),
                                                            P(C.c_int32)]


def load_gpu_lib():
    """Load libfxcorr_gpu.so; raises if it has not been built."""
    global _gpu_lib
    if _gpu_lib is not None:
        return _gpu_lib
    if not os.path.exists(_GPU_LIB_PATH):
        raise RuntimeError(f"HIP library not built: {_GPU_LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(_GPU_LIB_PATH)
    P = C.POINTER
    lib.fxg_last_error.restype = C.c_char_p
    _declare_common(lib, "fxg_")
    lib.fxg_ctx_create.argtypes = [C.c_int, P(C.c_void_p)]
    lib.fxg_ctx_destroy.argtypes = [C.c_void_p]
    lib.fxg_device_alloc.argtypes = [C.c_void_p, C.c_int64, P(C.c_void_p)]
    lib.fxg_device_free.argtypes = [C.c_void_p, C.c_void_p]
    lib.fxg_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    lib.fxg_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    lib.fxg_unpack.argtypes = [C.c_void_p, P(Datastream), C.c_void_p, C.c_int64, C.c_int64,
                               C.c_int32, C.c_void_p, P(C.c_float)]
    lib.fxg_unpack_noalloc.argtypes = [C.c_void_p, P(Datastream), C.c_void_p, C.c_int64, C.c_int64,
                               C.c_int32, C.c_void_p, P(C.c_float)]
    lib.fxg_job_create.argtypes = [C.c_void_p, P(JobDesc), P(C.c_void_p)]
    lib.fxg_job_destroy.argtypes = [C.c_void_p]
    lib.fxg_job_plan_info.argtypes = [C.c_void_p, P(PlanInfo)]
    lib.fxg_subint_process.argtypes = [C.c_void_p, P(SubintDesc), P(StationInput), C.c_int,
                                       C.c_void_p, C.c_int]
    lib.fxg_subint_submit.argtypes = [C.c_void_p, P(SubintDesc), P(StationInput), C.c_int,
                                      C.c_void_p, C.c_int, P(C.c_int64)]
    lib.fxg_subint_query.argtypes = [C.c_void_p, C.c_int64, P(C.c_int)]
    lib.fxg_subint_wait.argtypes = [C.c_void_p, C.c_int64]
    lib.fxg_station_process_block.argtypes = [C.c_void_p, P(SubintDesc), P(StationInput), C.c_int,
                                              C.c_int, C.c_int, C.c_void_p, P(BlockInfo)]
    lib.fxg_job_set_keep_taps.argtypes = [C.c_void_p, C.c_int]
    lib.fxg_job_set_xmac_streaming.argtypes = [C.c_void_p, C.c_int]
    lib.fxg_job_set_profiling.argtypes = [C.c_void_p, C.c_int]
    lib.fxg_job_stage_times.argtypes = [C.c_void_p, P(C.c_double), P(C.c_int64), C.c_int]
    lib.fxg_tap_block_info.argtypes = [C.c_void_p, C.c_int, C.c_int, P(BlockInfo)]
    lib.fxg_tap_band_weights.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.fxg_tap_spectrum.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.fxg_vdifmux_output_shape.argtypes = [P(VdifMuxDesc), P(C.c_int32), P(C.c_int32)]
    lib.fxg_vdifmux.restype = C.c_int64
    lib.fxg_vdifmux.argtypes = [C.c_void_p, P(VdifMuxDesc), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int64, C.c_int64, P(VdifMuxStats)]
    lib.fxg_job_set_schedule.argtypes = [C.c_void_p, P(Schedule)]
    lib.fxg_job_get_schedule.argtypes = [C.c_void_p, P(Schedule)]
    lib.fxg_job_autotune.argtypes = [C.c_void_p, P(SubintDesc), P(StationInput), C.c_int, C.c_int,
                                     P(TuneResult), C.c_int, P(C.c_int)]
    lib.fxg_job_set_graph.argtypes = [C.c_void_p, C.c_int]
    lib.fxg_job_calibrate_scales.argtypes = [C.c_void_p]
    lib.fxg_job_set_kurtosis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.fxg_job_get_kurtosis.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, P(C.c_float), C.c_int,
                                         P(C.c_int32), P(C.c_int32)]
    lib.fxg_job_graph_captures.argtypes = [C.c_void_p, P(C.c_int64)]
    lib.fxg_batch_create.argtypes = [C.c_void_p, P(C.c_void_p), C.c_int, P(C.c_void_p)]
    lib.fxg_batch_destroy.argtypes = [C.c_void_p]
    lib.fxg_batch_process.argtypes = [C.c_void_p, P(SubintDesc), P(P(StationInput)), C.c_int,
                                      P(C.c_void_p), C.c_int]
    lib.fxg_batch_graph_captures.argtypes = [C.c_void_p, P(C.c_int64)]
    _gpu_lib = lib
    return lib


class ServeStats(C.Structure):
    _fields_ = [("budget_bytes", C.c_int64), ("used_bytes", C.c_int64), ("nconfigs", C.c_int32),
                ("nslots", C.c_int32), ("submitted", C.c_int64), ("completed", C.c_int64),
                ("queued", C.c_int64), ("refused", C.c_int64)]


_HOST_LIB_PATH = os.path.join(os.path.dirname(_GPU_LIB_PATH), "libfxcorr_host.so")
_host_lib = None


def load_host_lib():
    """Load libfxcorr_host.so (the C++ Mode/Core/Visibility layer and the
    request loop, include/fxcorr_serve.h); raises if it has not been built."""
    global _host_lib
    if _host_lib is not None:
        return _host_lib
    load_gpu_lib()
    if not os.path.exists(_HOST_LIB_PATH):
        raise RuntimeError(f"host library not built: {_HOST_LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(_HOST_LIB_PATH)
    P = C.POINTER
    lib.fxs_create.argtypes = [C.c_int, C.c_int64, P(C.c_void_p)]
    lib.fxs_destroy.argtypes = [C.c_void_p]
    lib.fxs_add_config.argtypes = [C.c_void_p, P(JobDesc), C.c_int, P(C.c_int)]
    lib.fxs_submit.argtypes = [C.c_void_p, C.c_int, P(SubintDesc), P(StationInput), C.c_void_p, P(C.c_int64)]
    lib.fxs_wait.argtypes = [C.c_void_p, C.c_int64, P(C.c_int64)]
    lib.fxs_stats.argtypes = [C.c_void_p, P(ServeStats)]
    lib.fxs_set_max_inflight.argtypes = [C.c_void_p, C.c_int]
    lib.fxds_data_bytes.restype = C.c_int64
    lib.fxds_data_bytes.argtypes = [P(JobDesc), C.c_int, C.c_double]
    lib.fxds_format_of.argtypes = [P(JobDesc), C.c_int, C.c_double, P(DsFormat)]
    lib.fxds_create.argtypes = [P(DsFormat), P(C.c_void_p)]
    lib.fxds_destroy.argtypes = [C.c_void_p]
    lib.fxds_set_buffer.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32]
    lib.fxds_control.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                 P(StationInput), P(C.c_int32), C.c_int32, P(C.c_int64)]
    _host_lib = lib
    return lib


def load_oracle():
    """Load the CPU oracle (test infrastructure only)."""
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(_ORACLE_PATH):
        raise RuntimeError(f"oracle not built: {_ORACLE_PATH}")
    lib = C.CDLL(_ORACLE_PATH)
    P = C.POINTER
    lib.orc_last_error.restype = C.c_char_p
    _declare_common(lib, "orc_")
    lib.orc_lut.argtypes = [C.c_int, C.c_int, P(C.c_float)]
    lib.orc_calcstridelength.argtypes = [C.c_int]
    lib.orc_vdifmux.restype = C.c_int
    lib.orc_vdifmux.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                P(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, P(C.c_int64)]
    lib.orc_unpack.argtypes = [P(Datastream), C.c_void_p, C.c_int64, C.c_int64, C.c_int,
                               C.c_void_p, P(C.c_float)]
    lib.orc_block_info.argtypes = [P(JobDesc), P(SubintDesc), P(StationInput), C.c_int, C.c_int,
                                   P(BlockInfo)]
    lib.orc_spectra.argtypes = [P(JobDesc), P(SubintDesc), P(StationInput), C.c_int, C.c_int,
                                C.c_int, C.c_void_p]
    lib.orc_process.argtypes = [P(JobDesc), P(SubintDesc), P(StationInput), C.c_int, C.c_int,
                                C.c_int, C.c_int, C.c_void_p]
    lib.orc_fft.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.orc_kurtosis.argtypes = [P(JobDesc), P(SubintDesc), P(StationInput), C.c_int, C.c_int, C.c_int,
                                 P(C.c_float), P(C.c_int)]
    lib.orc_delay_interpolator.argtypes = [P(C.c_double), C.c_int, C.c_double, C.c_double,
                                           C.c_int, P(C.c_double)]
    _oracle = lib
    return lib


class FxgError(RuntimeError):
    pass


def check(lib, rc, what=""):
    if rc != 0:
        fn = getattr(lib, "fxg_last_error", None) or getattr(lib, "orc_last_error")
        msg = fn().decode() if fn() else ""
        raise FxgError(f"{what} failed ({rc}): {msg}")
    return rc
