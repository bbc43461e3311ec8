"""Seeded synthetic VDIF / Mark5B data for correlator jobs.

Signal model (after the reference's simulator, applications/datasim/src/
subband.cpp:580-662 and util.cpp:272-273): for every subband a common
white Gaussian sky signal g, and per station and polarisation
    x(t) = sqrt(rho) * g(t + tau_station) + sqrt(1 - rho) * n(t)
with unit variance and independent station noise n.  Station delays are
whole samples here, so the correlator's delay model can equal the data's
delay exactly.  2-bit samples are quantised at +-0.98 sigma into codes
0..3 (x < -0.98, [-0.98, 0), [0, 0.98), >= 0.98: the VDIF code order of
format_vdif.c:78), 8-bit samples as clip(round(8 x) + 128).

Frames are VDIF (32-byte header, word layout of libraries/vdifio/src/
vdifio.h:42-74) or Mark5B (16-byte header + 10000-byte payload); samples are
packed sample-time major, channel minor, LSB first.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from fxcorr import (FXG_FORMAT_MARK5B, FXG_FORMAT_VDIF, JobSpec, StationInput,
                    SubintDesc, FXG_FLAGS_PER_INT)
import ctypes as C


@dataclass
class StationData:
    raw: np.ndarray          # uint8 frames
    datasec: int
    datans: int
    validflags: np.ndarray   # int32
    interpolator: np.ndarray  # float64[3]
    delay_samples: int


@dataclass
class SubintData:
    spec: JobSpec
    sub_sec: int
    sub_ns: int
    scan_start_sec: int
    stations: List[StationData]

    def subint_desc(self) -> SubintDesc:
        d = SubintDesc()
        d.scan = 0
        d.sec = self.sub_sec
        d.ns = self.sub_ns
        d.scan_start_sec = self.scan_start_sec
        d.model_ok = 1
        return d

    def station_inputs(self, device_ptrs: Optional[Sequence[int]] = None):
        """ctypes StationInput array (host data unless device_ptrs given)."""
        arr = (StationInput * len(self.stations))()
        keep = []
        for i, st in enumerate(self.stations):
            a = arr[i]
            if device_ptrs is None:
                a.data = st.raw.ctypes.data
            else:
                a.data = device_ptrs[i]
            a.nbytes = st.raw.size
            a.scan = 0
            a.sec = st.datasec
            a.ns = st.datans
            vf = np.ascontiguousarray(st.validflags, dtype=np.int32)
            keep.append(vf)
            a.nvalidflags = vf.size
            a.validflags = vf.ctypes.data_as(C.POINTER(C.c_int32))
            for k in range(3):
                a.interpolator[k] = float(st.interpolator[k])
        return arr, keep


def quantise2(x: np.ndarray) -> np.ndarray:
    t = 0.98
    return ((x >= -t).astype(np.uint8) + (x >= 0).astype(np.uint8) + (x >= t).astype(np.uint8))


def quantise8(x: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(x * 8.0) + 128, 0, 255).astype(np.uint8)


def pack_codes(codes: np.ndarray, nbits: int) -> np.ndarray:
    """codes[T, nchan] -> bytes, sample-time major, channel minor, LSB first
    (nbits 2 or 8; 4 and 16 for complex 2- and 8-bit units)."""
    if nbits == 16:
        return np.ascontiguousarray(codes, dtype=np.uint16).reshape(-1).astype("<u2").view(np.uint8).copy()
    flat = np.ascontiguousarray(codes, dtype=np.uint8).reshape(-1)
    if nbits == 8:
        return flat.copy()
    if nbits == 4:
        flat = flat.reshape(-1, 2)
        return (flat[:, 0] | (flat[:, 1] << 4)).astype(np.uint8)
    if nbits == 2:
        flat = flat.reshape(-1, 4)
        return (flat[:, 0] | (flat[:, 1] << 2) | (flat[:, 2] << 4) | (flat[:, 3] << 6)).astype(np.uint8)
    raise ValueError("nbits")


def vdif_header(sec: int, frameno: int, framebytes: int, nchan: int, nbits: int,
                station: int, invalid: bool = False, complex_data: bool = False) -> np.ndarray:
    w = np.zeros(8, dtype=np.uint32)
    w[0] = (sec & 0x3FFFFFFF) | ((1 << 31) if invalid else 0)
    w[1] = (frameno & 0xFFFFFF) | (20 << 24)
    w[2] = ((framebytes // 8) & 0xFFFFFF) | ((int(math.log2(nchan)) & 0x1F) << 24)
    w[3] = (station & 0xFFFF) | (((nbits - 1) & 0x1F) << 26) | ((1 << 31) if complex_data else 0)
    return w.view(np.uint8)


def vdif_headers(datasec: int, first_frameno: int, nframes: int, frames_per_sec: int, framebytes: int,
                 nchan: int, nbits: int, station: int) -> np.ndarray:
    """vdif_header for nframes consecutive frames: [nframes, 32] bytes."""
    fn = first_frameno + np.arange(nframes, dtype=np.int64)
    w = np.zeros((nframes, 8), dtype=np.uint32)
    w[:, 0] = ((datasec + fn // frames_per_sec) & 0x3FFFFFFF).astype(np.uint32)
    w[:, 1] = ((fn % frames_per_sec) & 0xFFFFFF).astype(np.uint32) | np.uint32(20 << 24)
    w[:, 2] = ((framebytes // 8) & 0xFFFFFF) | ((int(math.log2(nchan)) & 0x1F) << 24)
    w[:, 3] = (station & 0xFFFF) | (((nbits - 1) & 0x1F) << 26)
    return w.view(np.uint8).reshape(nframes, 32)


def mark5b_header(frameno: int, invalid: bool = False) -> np.ndarray:
    w = np.zeros(4, dtype=np.uint32)
    w[0] = 0xABADDEED
    w[1] = (frameno & 0x7FFF) | ((1 << 15) if invalid else 0)
    return w.view(np.uint8)


def frame_stream(spec: JobSpec, codes: np.ndarray, first_frame_sec: int, first_frameno: int,
                 frames_per_sec: int, station: int, invalid_frames: Sequence[int] = ()) -> np.ndarray:
    """Pack codes[T, nchan] (T a multiple of framesamples) into frames."""
    payload = pack_codes(codes, spec.unitbits)
    databytes = spec.framebytes - spec.headerbytes
    nframes = payload.size // databytes
    payload = payload[: nframes * databytes].reshape(nframes, databytes)
    out = np.empty((nframes, spec.framebytes), dtype=np.uint8)
    out[:, spec.headerbytes:] = payload
    inval = set(int(i) for i in invalid_frames)
    for f in range(nframes):
        fn = first_frameno + f
        sec = first_frame_sec + fn // frames_per_sec
        fno = fn % frames_per_sec
        if spec.fmt != FXG_FORMAT_MARK5B:
            out[f, : spec.headerbytes] = vdif_header(sec, fno, spec.framebytes, spec.nrecbands,
                                                     spec.nbits, station, f in inval,
                                                     spec.is_complex)[: spec.headerbytes]
        else:
            out[f, :16] = mark5b_header(fno, f in inval)
    return out.reshape(-1)


def set_edv4(raw: np.ndarray, framebytes: int, masklength: int, masks: dict) -> np.ndarray:
    """The frames of raw (VDIF, 32-byte headers) as EDV4 frames: extended
    version 4 with the given validity mask length, the sync word 0xACABFEED
    and a 64-bit validity mask per frame (masks: frame -> mask; other frames
    get every one of the masklength bits set), in place; returns raw."""
    fr = raw.reshape(-1, framebytes)
    words = fr[:, :32].view(np.uint32)
    words[:, 4] = np.uint32((4 << 24) | ((masklength & 0xFF) << 16))
    words[:, 5] = np.uint32(0xACABFEED)
    full = (1 << masklength) - 1
    for f in range(fr.shape[0]):
        m = int(masks.get(f, full))
        words[f, 6] = np.uint32(m & 0xFFFFFFFF)
        words[f, 7] = np.uint32(m >> 32)
    return raw


def make_subint(spec: JobSpec, seed: int = 1, rho: float | Sequence[float] = 0.05,
                delays_us: Optional[Sequence[float]] = None,
                rates_us_per_s: Optional[Sequence[float]] = None,
                sub_sec: int = 10, guard_frames: Optional[int] = None,
                invalid_frames: Optional[dict] = None,
                invalid_blocks: Optional[dict] = None,
                extra_frames: int = 1,
                pcal: Optional[dict] = None) -> SubintData:
    """Generate one subintegration of raw frames for every station.

    delays_us: model delay per station (the data carry the whole-sample part
    of it; any fractional part is then a model-vs-data offset).  rates add a
    linear delay rate to the model only (exercises the fringe rotator).
    Complex sampling: the common and station signals are complex white
    noise, each component quantised like a real sample (real part in the
    low bits of the unit).
    pcal: phase-cal comb added to every band before quantisation, locked to
    the second boundary: {"amp": a, "spacing_hz": s, "offsets_hz": [per
    subband], "ntones": [per subband], "slope_deg": d}; tone k of a band is
    a * sin(2 pi (offset + k s) n / fs + p_k) at absolute sample n, p_k =
    k d, or "phases_deg"[k] when the dict has that list (real
    sampling, fs = 2 x bandwidth), a * exp(i (2 pi (offset + k s) n / fs +
    k d)) (complex sampling, fs = bandwidth; double sideband: the comb shifted
    down by half the band, the LO sitting at the band centre).
    """
    rng = np.random.default_rng(seed)
    ts = spec.sampletime_us
    nst = spec.nstations
    nsb, npol = spec.nsubbands, spec.npol
    fs = spec.framesamples
    frames_per_sec = int(round(1e6 / ts)) // fs
    if delays_us is None:
        delays_us = rng.uniform(-50.0, 50.0, nst) * min(1.0, spec.blockspersend * spec.fftchannels * ts / 1e4)
    if rates_us_per_s is None:
        rates_us_per_s = np.zeros(nst)
    dsamp = [int(round(d / ts)) for d in delays_us]
    maxd = max(abs(d) for d in dsamp) + 2 * spec.fftchannels
    if guard_frames is None:
        guard_frames = (maxd + fs - 1) // fs + 1
        while abs(guard_frames * fs * ts * 1000.0 - round(guard_frames * fs * ts * 1000.0)) > 1e-6:
            guard_frames += 1  # the data start must fall on a whole nanosecond
    G = guard_frames * fs
    nsamp_sub = spec.blockspersend * spec.fftchannels
    nframes = (nsamp_sub + 2 * G + fs - 1) // fs + extra_frames
    T = nframes * fs
    rhos = np.broadcast_to(np.asarray(rho, dtype=np.float64), (nsb,))
    # data start: G samples before the subint start
    sub_ns = 0
    start_ns_total = sub_sec * 1_000_000_000 - G * ts * 1000.0
    if abs(start_ns_total - round(start_ns_total)) > 1e-6:
        raise ValueError("guard does not land on an integer ns")
    start_ns_total = int(round(start_ns_total))
    datasec, datans = divmod(start_ns_total, 1_000_000_000)
    first_frameno = int(round((datans / 1000.0) / ts)) // fs
    lo = min(dsamp)
    hi = max(dsamp)
    stations = []
    cplx = spec.is_complex
    codes_all = [np.empty((T, spec.nrecbands), dtype=np.uint16 if cplx and spec.nbits == 8 else np.uint8)
                 for _ in range(nst)]
    quant = quantise2 if spec.nbits == 2 else quantise8
    datasamples = int(datans / (ts * 1e3) + 0.5)
    for sb in range(nsb):
        g = rng.standard_normal(T + hi - lo)
        if cplx:
            g = g + 1j * rng.standard_normal(T + hi - lo)
        r = float(rhos[sb])
        comb = 0.0
        if pcal:
            fs_hz = (1e6 if cplx else 2e6) * spec.bw_mhz
            nabs = datasamples + np.arange(T, dtype=np.float64)
            comb = np.zeros(T, np.complex128 if cplx else np.float64)
            for k in range(pcal["ntones"][sb]):
                f = pcal["offsets_hz"][sb] + k * pcal["spacing_hz"]
                if spec.sampling == "complex_dsb":
                    f -= fs_hz / 2
                pk = pcal["phases_deg"][k] if "phases_deg" in pcal else k * pcal.get("slope_deg", 0.0)
                ph = 2 * np.pi * np.mod(f * nabs, fs_hz) / fs_hz + np.radians(pk)
                comb += pcal["amp"] * (np.exp(1j * ph) if cplx else np.sin(ph))
        for s in range(nst):
            gs = g[dsamp[s] - lo: dsamp[s] - lo + T]
            for p in range(npol):
                if cplx:
                    n = rng.standard_normal(T) + 1j * rng.standard_normal(T)
                    x = math.sqrt(r) * gs + math.sqrt(1.0 - r) * n + comb
                    q = quant(x.real).astype(np.uint16) | (quant(x.imag).astype(np.uint16) << spec.nbits)
                else:
                    x = math.sqrt(r) * gs + math.sqrt(1.0 - r) * rng.standard_normal(T) + comb
                    q = quant(x)
                codes_all[s][:, sb * npol + p] = q
    nflags = (spec.blockspersend + FXG_FLAGS_PER_INT - 1) // FXG_FLAGS_PER_INT
    for s in range(nst):
        inv = (invalid_frames or {}).get(s, ())
        raw = frame_stream(spec, codes_all[s], datasec, first_frameno, frames_per_sec, s, inv)
        vf = np.zeros(nflags, dtype=np.int64)
        for b in range(spec.blockspersend):
            if b in set((invalid_blocks or {}).get(s, ())):
                continue
            vf[b // FXG_FLAGS_PER_INT] |= 1 << (b % FXG_FLAGS_PER_INT)
        blk_s = spec.fftchannels * ts * 1e-6
        interp = np.array([0.0, rates_us_per_s[s] * blk_s, delays_us[s]], dtype=np.float64)
        stations.append(StationData(raw=raw, datasec=int(datasec), datans=int(datans),
                                    validflags=vf.astype(np.int32), interpolator=interp,
                                    delay_samples=dsamp[s]))
    return SubintData(spec=spec, sub_sec=sub_sec, sub_ns=sub_ns, scan_start_sec=0, stations=stations)


# ---------------------------------------------------------------------------
# Device-side generation for large jobs (same model, generated with torch on
# the GPU so that multi-GB subintegrations start resident in HBM)
# ---------------------------------------------------------------------------
def make_subint_device(spec: JobSpec, seed: int = 1, rho=0.05, delay_samples=None,
                       rates_us_per_s=None, sub_sec: int = 10, device: str = "cuda:0"):
    """Returns (SubintData with raw=None, list of torch uint8 tensors on device).

    Delays are whole samples (delay_samples per station); the model carries
    the same delay.  Only VDIF 2-bit/8-bit is generated here.
    """
    import torch
    if spec.fmt != FXG_FORMAT_VDIF:
        raise ValueError("device generator writes VDIF")
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ts = spec.sampletime_us
    nst, nsb, npol = spec.nstations, spec.nsubbands, spec.npol
    fs = spec.framesamples
    frames_per_sec = int(round(1e6 / ts)) // fs
    if delay_samples is None:
        rng = np.random.default_rng(seed)
        delay_samples = [int(v) for v in rng.integers(-2000, 2000, nst)]
    if rates_us_per_s is None:
        rates_us_per_s = np.zeros(nst)
    maxd = max(abs(d) for d in delay_samples) + 2 * spec.fftchannels
    guard_frames = (maxd + fs - 1) // fs + 1
    while abs(guard_frames * fs * ts * 1000.0 - round(guard_frames * fs * ts * 1000.0)) > 1e-6:
        guard_frames += 1      # the data start must fall on a whole nanosecond
    G = guard_frames * fs
    nsamp_sub = spec.blockspersend * spec.fftchannels
    nframes = (nsamp_sub + 2 * G + fs - 1) // fs + 1
    T = nframes * fs
    rhos = np.broadcast_to(np.asarray(rho, dtype=np.float64), (nsb,))
    start_ns_total = sub_sec * 1_000_000_000 - G * ts * 1000.0
    if abs(start_ns_total - round(start_ns_total)) > 1e-6:
        raise ValueError("guard does not land on an integer ns")
    datasec, datans = divmod(int(round(start_ns_total)), 1_000_000_000)
    first_frameno = int(round((datans / 1000.0) / ts)) // fs
    lo, hi = min(delay_samples), max(delay_samples)
    nchan = spec.nrecbands
    databytes = spec.framebytes - spec.headerbytes
    # common signal per subband, kept on device
    common = [torch.randn(T + hi - lo, generator=g, device=device) for _ in range(nsb)]
    raws = []
    stations = []
    nflags = (spec.blockspersend + FXG_FLAGS_PER_INT - 1) // FXG_FLAGS_PER_INT
    for s in range(nst):
        codes = torch.empty((T, nchan), dtype=torch.uint8, device=device)
        for sb in range(nsb):
            r = float(rhos[sb])
            gs = common[sb][delay_samples[s] - lo: delay_samples[s] - lo + T]
            for p in range(npol):
                x = math.sqrt(r) * gs + math.sqrt(1.0 - r) * torch.randn(T, generator=g, device=device)
                if spec.nbits == 2:
                    q = (x >= -0.98).to(torch.uint8) + (x >= 0).to(torch.uint8) + (x >= 0.98).to(torch.uint8)
                else:
                    q = torch.clamp(torch.round(x * 8.0) + 128, 0, 255).to(torch.uint8)
                codes[:, sb * npol + p] = q
                del x, q
        flat = codes.reshape(-1)
        if spec.nbits == 2:
            f4 = flat.view(-1, 4)
            payload = f4[:, 0] | (f4[:, 1] << 2) | (f4[:, 2] << 4) | (f4[:, 3] << 6)
        else:
            payload = flat
        del codes, flat
        out = torch.empty((nframes, spec.framebytes), dtype=torch.uint8, device=device)
        out[:, spec.headerbytes:] = payload.view(nframes, databytes)
        hdr = vdif_headers(datasec, first_frameno, nframes, frames_per_sec, spec.framebytes, nchan,
                           spec.nbits, s)[:, : spec.headerbytes]
        out[:, : spec.headerbytes] = torch.from_numpy(np.ascontiguousarray(hdr)).to(device)
        raws.append(out.view(-1))
        vf = np.zeros(nflags, dtype=np.int64)
        for b in range(spec.blockspersend):
            vf[b // FXG_FLAGS_PER_INT] |= 1 << (b % FXG_FLAGS_PER_INT)
        blk_s = spec.fftchannels * ts * 1e-6
        interp = np.array([0.0, rates_us_per_s[s] * blk_s, delay_samples[s] * ts], dtype=np.float64)
        stations.append(StationData(raw=np.empty(0, np.uint8), datasec=int(datasec), datans=int(datans),
                                    validflags=vf.astype(np.int32), interpolator=interp,
                                    delay_samples=delay_samples[s]))
    del common
    sd = SubintData(spec=spec, sub_sec=sub_sec, sub_ns=0, scan_start_sec=0, stations=stations)
    return sd, raws


def station_inputs_device(sd: SubintData, raws):
    """StationInput array pointing at device tensors."""
    arr = (StationInput * len(sd.stations))()
    keep = []
    for i, st in enumerate(sd.stations):
        a = arr[i]
        a.data = raws[i].data_ptr()
        a.nbytes = raws[i].numel()
        a.scan = 0
        a.sec = st.datasec
        a.ns = st.datans
        vf = np.ascontiguousarray(st.validflags, dtype=np.int32)
        keep.append(vf)
        a.nvalidflags = vf.size
        a.validflags = vf.ctypes.data_as(C.POINTER(C.c_int32))
        for k in range(3):
            a.interpolator[k] = float(st.interpolator[k])
    return arr, keep


def uneven_workload_specs(subint_s: float = 0.128, xmac: str = "fp16", max_blocks: Optional[int] = None):
    """Six jobs of uneven size (stations 4..30, channels 64..32768, 4..16
    subbands of 16..128 MHz, 2- and 8-bit, two zoom bands of 1/16 of a
    subband each): (spec, label) pairs, blockspersend covering subint_s."""
    rows = [  # stations, channels, subbands, MHz, bits
        (4, 64, 4, 16.0, 2),
        (7, 256, 8, 32.0, 8),
        (12, 1024, 16, 64.0, 2),
        (18, 4096, 8, 128.0, 2),
        (25, 16384, 4, 64.0, 8),
        (30, 32768, 4, 128.0, 2),
    ]
    out = []
    for nst, nchan, nsb, bw, bits in rows:
        L = 2 * nchan
        blocks = max(1, int(round(subint_s * 2e6 * bw / L)))
        if max_blocks:
            blocks = min(blocks, max_blocks)
        zw = nchan // 16
        zoom = [(0, nchan // 4, zw), (nsb - 1, nchan // 2 + zw // 2, zw)]
        spec = JobSpec(nstations=nst, nsubbands=nsb, npol=2, bw_mhz=bw, nchan=nchan, nbits=bits,
                       blockspersend=blocks, xmac=xmac, zoom=zoom)
        out.append((spec, f"{nst}st_{nchan}ch_{nsb}x{int(bw)}MHz_{bits}bit"))
    return out
